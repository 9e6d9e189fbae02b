"""GPU parity of the component-level C ABI beyond the force field: Mass::addMDx / accFromF /
addForce, the projective constraints' projectResponse / projectVelocity / projectPosition,
vDot in both summation orders and CGLinearSolver::solve on caller device vectors -- each call
against the oracle's restatement of the same component on the same inputs.

Bars: bit-exact on the FP64 engine (same operations in the same order, contraction off); the
oscillator's sin / cos come from the device's libm (the oracle's from the host's), so the
oscillator entries are held to 1 ulp-level (4e-16 relative); the CG solve on the FP64 tiled
engine in the reference summation order is bit-exact as well.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _capi():
    from sofa_amd import capi
    return capi


def _scene(kind):
    """cfg0 cantilever (5x5x21, z = 0 fixed) with the requested mass and two oscillating nodes."""
    from sofa_amd import workloads
    s = workloads.cfg0()
    if kind == "diagonal":
        s["mass"] = ("diagonal", 1000.0)
    elif kind == "meshmatrix":
        s["mass"] = ("meshmatrix", 1000.0)
    elif kind == "lumped":
        s["mass"] = ("meshmatrix_lumped", 1000.0)
    top = np.nonzero(s["pos"][:, 2] == s["pos"][:, 2].max())[0][:2]
    amp = np.zeros((len(top), 3))
    amp[:, 0] = 0.003
    amp[:, 1] = -0.002
    s["oscillators"] = dict(index=top.astype(np.int32), mean=s["pos"][top].copy(), amplitude=amp,
                            pulsation=np.full(len(top), 2 * np.pi * 1.7), phase=np.array([0.0, 0.4]))
    return s


def _oracle(spec, dot_mode=0):
    from specs import oracle_from_spec
    return oracle_from_spec(spec, dot_mode)


def _dev(a, dtype=np.float64):
    from sofa_amd import capi
    return capi.DeviceArray(np.asarray(a, dtype=dtype))


@pytest.mark.parametrize("kind", ["uniform", "diagonal", "meshmatrix", "lumped"])
def test_mass_and_projection_calls(kind):
    """addMDx (factor 1 and not), accFromF (refused for the consistent mass, as the reference
    does), addForce, projectResponse, projectVelocity (with activate_projectVelocity) and
    projectPosition at a given time, on the FP64 engine."""
    from sofa_amd import workloads
    capi = _capi()
    print(f"case: mass and projection calls, {kind}", flush=True)
    spec = _scene(kind)
    spec["project_velocity"] = True
    n = len(spec["pos"])
    rng = np.random.default_rng(7)
    dx, f0 = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    o = _oracle(spec)
    o.set_project_velocity(True)
    eng = capi.Engine(0, capi.FP64)
    sid = workloads.add_to_engine(eng, spec)
    eng.commit()
    for factor in (1.0, -0.37):
        dxd, fd = _dev(dx), _dev(f0)
        eng.mass_add_mdx(sid, dxd.ptr, fd.ptr, factor)
        eng.sync()
        assert np.array_equal(fd.numpy(), o.mass_add_mdx(dx, f0, factor))
    fd, ad = _dev(f0), _dev(np.zeros((n, 3)))
    if kind == "meshmatrix":
        with pytest.raises(capi.SofaAmdError) as ei:
            eng.mass_acc_from_f(sid, fd.ptr, ad.ptr)
        assert ei.value.code == capi.EINVAL
        with pytest.raises(ValueError):
            o.mass_acc_from_f(f0)
    else:
        eng.mass_acc_from_f(sid, fd.ptr, ad.ptr)
        eng.sync()
        a_ref = o.mass_acc_from_f(f0)
        assert np.array_equal(ad.numpy(), a_ref)
        # closed form: M (M^-1 f) = f up to the division's rounding
        back = o.mass_add_mdx(a_ref, np.zeros((n, 3)), 1.0)
        assert np.allclose(back, f0, rtol=1e-14, atol=1e-14)
    fd = _dev(f0)
    eng.mass_add_force(sid, fd.ptr)
    eng.sync()
    assert np.array_equal(fd.numpy(), o.mass_add_force(f0))
    # projective constraints
    vd = _dev(dx)
    eng.project_response(sid, vd.ptr)
    eng.sync()
    pr = o.project(0, dx)
    assert np.array_equal(vd.numpy(), pr)
    assert np.all(pr[spec["fixed"]] == 0) and np.all(pr[spec["oscillators"]["index"]] == 0)
    t = 0.37
    for what, call in ((1, eng.project_velocity), (2, eng.project_position)):
        vd = _dev(dx)
        call(sid, vd.ptr, t)
        eng.sync()
        got, ref = vd.numpy(), o.project(what, dx, t)
        osc = spec["oscillators"]["index"]
        rest = np.setdiff1d(np.arange(n), osc)
        assert np.array_equal(got[rest], ref[rest])
        assert np.allclose(got[osc], ref[osc], rtol=4e-16, atol=1e-18)
    eng.close()


@pytest.mark.parametrize("prec", ["FP64", "FP32"])
def test_vdot_orders(prec):
    """vDot of two caller vectors: order 1 equals the oracle's sequential sum bit for bit (FP64;
    FP32 vectors: the same sum of the float values' exact products), order 0 the oracle's blocked
    order; the scene (21,525 nodes) spans 84 blocks."""
    from sofa_amd import meshes
    capi = _capi()
    print(f"case: vdot {prec}", flush=True)
    pos, tets, fixed = meshes.beam((15, 15, 97), (1.0, 1.0, 4.0))
    n = len(pos)
    eng = capi.Engine(0, getattr(capi, prec))
    sid = eng.add_scene(pos, tets, [1e5], [0.45], fixed, mass_value=1.0)
    eng.commit()
    rng = np.random.default_rng(11)
    a, b = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    dt = np.float64 if prec == "FP64" else np.float32
    if prec == "FP32":
        a, b = a.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64)
    ad, bd = _dev(a, dt), _dev(b, dt)
    out = _dev(np.zeros(2))
    eng.vdot(sid, ad.ptr, bd.ptr, 1, out.ptr)
    eng.vdot(sid, ad.ptr, bd.ptr, 0, out.ptr + 8)
    eng.sync()
    from oracle import OracleScene
    o = OracleScene(pos)
    seq, blk = o.vdot(a, b, 0), o.vdot(a, b, 1)
    got = out.numpy()
    assert got[0] == seq, (got[0], seq)
    assert got[1] == blk, (got[1], blk)
    assert abs(seq - float(np.sum(a * b))) <= 1e-12 * np.sum(np.abs(a * b))
    eng.close()


@pytest.mark.parametrize("prec,order", [("FP64", 0), ("FP64", 1), ("FP64_TILED", 1)])
def test_cg_solve_component(prec, order):
    """CGLinearSolver::solve on a caller right-hand side, after addForce at a deformed state:
    x, the iteration count, the end reason and timeStepCount equal the oracle's solve of the same
    system (the Euler step's factors m = 1 + h rm, b = -h, k = -h (h + rs)) bit for bit; a second
    solve (timeStepCount 1: no first-solve exemption) as well."""
    from sofa_amd import workloads
    capi = _capi()
    print(f"case: cg solve {prec} order {order}", flush=True)
    spec = workloads.cfg0()
    spec["iters"], spec["tol"] = 40, 1e-7
    n = len(spec["pos"])
    rng = np.random.default_rng(5)
    x = spec["pos"] + 0.03 * rng.standard_normal((n, 3))
    o = _oracle(spec, dot_mode=1 if order == 0 else 0)
    o.fem_add_force(x, np.zeros((n, 3)))
    eng = capi.Engine(0, getattr(capi, prec))
    if order == 1:
        eng.set_dot_order(True)
    sid = workloads.add_to_engine(eng, spec)
    eng.commit()
    eng.write(sid, x=x)                      # the tiled operator recomputes rotations from these
    xd, fd = _dev(x), _dev(np.zeros((n, 3)))
    eng.fem_add_force(sid, xd.ptr, fd.ptr)
    h, rs, rm = spec["dt"], spec["rs"], spec["rm"]
    mf, bf, kf = 1 + h * rm, -h, -h * (h + rs)
    for k in range(2):
        b = o.project(0, rng.standard_normal((n, 3)))
        bd, sol = _dev(b), _dev(np.zeros((n, 3)))
        st = eng.cg_solve(sid, mf, bf, kf, bd.ptr, sol.ptr)
        ref = o.cg_solve(b, mf, bf, kf)
        so = o.stats()
        got = sol.numpy()
        print(f"  solve {k}: {st['iterations']} iterations, end {st['end_reason']}, err {st['err']:.3e}", flush=True)
        assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))
        assert st["iterations"] == so["iterations"] and st["end_reason"] == so["endcond"]
        assert st["step_count"] == so["step_count"] == k + 1
    eng.close()


def test_potential_energy_large_is_zero():
    """TetrahedronFEMForceField::getPotentialEnergy sums the small method only: 0 for large."""
    from sofa_amd import workloads
    capi = _capi()
    spec = workloads.cfg0()
    eng = capi.Engine(0, capi.FP64)
    sid = workloads.add_to_engine(eng, spec)
    eng.commit()
    xd = _dev(spec["pos"] * 1.01)
    assert eng.fem_potential_energy(sid, xd.ptr) == 0.0
    eng.close()


def _lm_spec(relative):
    """cfg0 cantilever with a LinearMovementProjectiveConstraint on its tip face: free for the
    first 10 steps, driven through two segments, free again after the last key time."""
    from sofa_amd import workloads
    s = workloads.cfg0()
    top = np.nonzero(s["pos"][:, 2] == s["pos"][:, 2].max())[0].astype(np.int32)
    moves = np.array([[0.0, 0.0, 0.0], [0.03, -0.02, 0.01], [0.0, 0.04, -0.02]])
    if not relative:
        moves = moves + np.array([0.5, 0.5, 4.0])
    s["linear_movement"] = dict(indices=top, key_times=[0.1, 0.25, 0.45], movements=moves, relative=relative)
    return s


@pytest.mark.parametrize("relative", [True, False])
@pytest.mark.parametrize("prec", ["FP64", "FP64_TILED", "FP32"])
def test_linear_movement(prec, relative):
    """LinearMovementProjectiveConstraint through the whole step, 60 steps (before, inside and after
    the movement windows): the FP64 engine and the FP64 tiled engine in the reference summation
    order equal the oracle's reference-order run bit for bit; FP32 within 2e-5 of it."""
    from sofa_amd import workloads
    capi = _capi()
    print(f"case: linear movement {prec} relative {relative}", flush=True)
    spec = _lm_spec(relative)
    eng = capi.Engine(0, getattr(capi, prec))
    if prec != "FP32":
        eng.set_dot_order(True)
    sid = workloads.add_to_engine(eng, spec)
    eng.commit()
    eng.set_graphs(True)
    eng.step([sid], 60)
    x, v, _ = eng.read(sid)
    eng.close()
    o = _oracle(spec, dot_mode=0)
    o.step(spec["dt"], 60)
    xo, vo = o.get("x"), o.get("v")
    top = spec["linear_movement"]["indices"]
    assert np.max(np.abs(xo[top] - spec["pos"][top])) > 1e-3        # the tip was driven
    if prec == "FP32":
        e = float(np.max(np.linalg.norm(x - xo, axis=1)) / np.max(np.linalg.norm(xo, axis=1)))
        assert e <= 2e-5, e
    else:
        assert np.array_equal(x, xo), float(np.max(np.abs(x - xo)))
        assert np.array_equal(v, vo)


def test_linear_movement_component_calls():
    """projectResponse / projectVelocity / projectPosition of the movement on caller vectors at a
    time inside a window and at one outside, against the oracle's (FP64, bit for bit)."""
    from sofa_amd import workloads
    capi = _capi()
    print("case: linear movement component calls", flush=True)
    spec = _lm_spec(True)
    n = len(spec["pos"])
    eng = capi.Engine(0, capi.FP64)
    sid = workloads.add_to_engine(eng, spec)
    eng.commit()
    o = _oracle(spec)
    rng = np.random.default_rng(3)
    # projectPosition first at t = 0.3 (inside: x0 captured from this vector), then later times
    for what, t in ((2, 0.3), (1, 0.3), (1, 0.05), (2, 0.33), (2, 0.5)):
        a = rng.standard_normal((n, 3))
        ad = _dev(a)
        eng.project_position(sid, ad.ptr, t) if what == 2 else eng.project_velocity(sid, ad.ptr, t)
        eng.sync()
        ref = o.project(what, a, t)
        assert np.array_equal(ad.numpy(), ref), (what, t)
        ad.close()
    a = rng.standard_normal((n, 3))
    ad = _dev(a)
    eng.project_response(sid, ad.ptr)           # at the scene's time 0: before the window
    eng.sync()
    assert np.array_equal(ad.numpy(), o.project(0, a, 0.0))
    eng.close()


def _pcg_spec(which, kind):
    from sofa_amd import workloads
    if which == "beam":
        s = workloads.cfg0()
        rng = np.random.default_rng(17)
        s["E"] = 10 ** rng.uniform(4.0, 6.0, len(s["tets"]))     # stiffness spread over two decades
        s["iters"], s["tol"] = 60, 1e-12
        n = 30
    else:
        # the ellipsoid recipe at 12 mm (consistent MeshMatrixMass, per-element materials), the
        # driven face at rest so that no sin / cos enters
        s = workloads.cfg2(pitch=0.012, amplitude=0.0)
        s["iters"], s["tol"] = 200, 1e-16
        n = 12
    s["preconditioner"] = kind
    return s, n


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("which", ["beam", "ellipsoid"])
def test_pcg_bitexact(which, kind, order):
    """PCGLinearSolver with a Jacobi (1) / block-Jacobi (2) preconditioner on the FP64 engine:
    positions, velocities and the last solve's iteration count equal the oracle's bit for bit, in
    the blocked (0) and the reference's sequential (1) dot-product order, over steps whose solves
    start from the previous step's solution."""
    from sofa_amd import workloads
    capi = _capi()
    print(f"case: pcg {which} kind {kind} order {order}", flush=True)
    spec, nsteps = _pcg_spec(which, kind)
    eng = capi.Engine(0, capi.FP64)
    if order == 1:
        eng.set_dot_order(True)
    sid = workloads.add_to_engine(eng, spec)
    eng.commit()
    eng.set_graphs(True)
    eng.step([sid], nsteps)
    x, v, _ = eng.read(sid)
    st = eng.stats(sid)
    eng.close()
    o = _oracle(spec, dot_mode=1 if order == 0 else 0)
    o.step(spec["dt"], nsteps)
    so = o.stats()
    print(f"  last solve: {st['iterations']} iterations (oracle {so['iterations']}), r.z/b.b {st['err']:.3e}", flush=True)
    assert st["iterations"] == so["iterations"]
    assert np.array_equal(x, o.get("x")), float(np.max(np.abs(x - o.get("x"))))
    assert np.array_equal(v, o.get("v"))
    assert np.max(np.abs(x - spec["pos"])) > 1e-5


@pytest.mark.parametrize("which", ["beam", "ellipsoid"])
def test_mixed_jacobi_inner_cg(which):
    """MIXED with PCGLinearSolver's Jacobi (1) or block-Jacobi (2) preconditioner in its inner
    FP32 CG (sweeps of 10 iterations continuing one search direction, until the FP64 |r|/|b| <=
    1e-8) stays within 1e-6 relative of the oracle's FP64 Jacobi PCG run to convergence, and needs
    no more inner iterations than the plain inner CG on the same scene (element stiffness spread
    over two decades; on the ellipsoid, whose systems take the plain CG thousands of iterations,
    strictly fewer).  The
    inner tolerance is 0: the inner CG measures its residual in the preconditioned norm, which
    can be below the outer test's bound while |r| is not, and would then end every sweep at once.
    MIXED has no reference counterpart: the comparison point is the converged FP64 solution."""
    from sofa_amd import workloads
    capi = _capi()
    print(f"case: mixed jacobi {which}", flush=True)
    spec, nsteps = _pcg_spec(which, 1)
    o = _oracle(dict(spec, iters=5000, tol=1e-26))
    o.step(spec["dt"], nsteps)
    xo = o.get("x")
    totals = {}
    for kind in (2, 1, 0):
        s = dict(spec, tol=1e-8, preconditioner=kind,
                 refinement=dict(inner_iterations=10, inner_tolerance=0.0, max_sweeps=1000, keep_direction=True))
        eng = capi.Engine(0, capi.MIXED)
        sid = workloads.add_to_engine(eng, s)
        eng.commit()
        its = 0
        for _ in range(nsteps):
            eng.step([sid], 1)
            st = eng.stats(sid)
            assert st["end_reason"] == 1, (kind, st)
            its += st["iterations"]
        x, _, _ = eng.read(sid)
        eng.close()
        e = float(np.max(np.linalg.norm(x - xo, axis=1)) / np.max(np.linalg.norm(xo, axis=1)))
        print(f"  preconditioner {kind}: {its} inner FP32 iterations over {nsteps} steps, eps_rel {e:.3e} "
              "against the converged FP64 PCG", flush=True)
        assert e <= 1e-6
        totals[kind] = its
    assert totals[1] <= totals[0] and totals[2] <= totals[0]
    if which == "ellipsoid":
        assert totals[1] < totals[0] and totals[2] < totals[0]


def test_cg_warm_start_tiled_reference_order():
    """CGLinearSolver warmStart on FP64_TILED in the reference's summation order, batched with
    cold-start scenes in one launch set: the cfg0 beam with spread stiffness (CG to 1e-4, the
    solves end on the tolerance) and a 9,104-node beam of the 256-beam mix (segmented dot
    products) warm-start, a second copy of the mix beam does not; every scene's positions,
    velocities and last iteration count equal the oracle's reference-order run bit for bit."""
    from sofa_amd import workloads
    capi = _capi()
    print("case: cg warm start, FP64_TILED reference order, 3 scenes", flush=True)
    beam, _ = _pcg_spec("beam", 0)
    beam["preconditioner"] = 0
    beam["tol"] = 1e-4
    big = dict(workloads.cfg3()[10])
    big["tol"] = 1e-6
    specs = [dict(beam, warm_start=True), dict(big, warm_start=True), dict(big)]
    nsteps = 12
    eng = capi.Engine(0, capi.FP64_TILED)
    eng.set_dot_order(True)
    ids = [workloads.add_to_engine(eng, sp) for sp in specs]
    eng.commit()
    eng.set_graphs(True)
    eng.step(ids, nsteps)
    got = [eng.read(i)[:2] + (eng.stats(i),) for i in ids]
    eng.close()
    for sp, (x, v, st) in zip(specs, got):
        o = _oracle(sp, dot_mode=0)
        o.step(sp["dt"], nsteps)
        so = o.stats()
        print(f"  warm {sp.get('warm_start', False)}: last solve {st['iterations']} iterations "
              f"(oracle {so['iterations']}), |r|/|b| {st['err']:.3e}", flush=True)
        assert st["iterations"] == so["iterations"]
        assert np.array_equal(x, o.get("x")), float(np.max(np.abs(x - o.get("x"))))
        assert np.array_equal(v, o.get("v"))
    # the warm start changes the big beam's trajectory (its solves start elsewhere)
    assert not np.array_equal(got[1][0], got[2][0])


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("which", ["beam", "ellipsoid"])
def test_cg_warm_start_bitexact(which, order):
    """CGLinearSolver warmStart on the FP64 engine: each solve starts from the previous step's
    solution (r = b - A x); positions, velocities and the last solve's iteration count equal the
    oracle's bit for bit in both dot-product orders."""
    from sofa_amd import workloads
    capi = _capi()
    print(f"case: cg warm start {which} order {order}", flush=True)
    spec, nsteps = _pcg_spec(which, 0)
    spec["preconditioner"] = 0
    spec["warm_start"] = True
    spec["tol"] = 1e-4 if which == "beam" else 1e-6      # the beam's solves end on the tolerance
    eng = capi.Engine(0, capi.FP64)
    if order == 1:
        eng.set_dot_order(True)
    sid = workloads.add_to_engine(eng, spec)
    eng.commit()
    eng.set_graphs(True)
    eng.step([sid], nsteps)
    x, v, _ = eng.read(sid)
    st = eng.stats(sid)
    eng.close()
    o = _oracle(spec, dot_mode=1 if order == 0 else 0)
    o.step(spec["dt"], nsteps)
    so = o.stats()
    print(f"  last solve: {st['iterations']} iterations (oracle {so['iterations']}), |r|/|b| {st['err']:.3e}", flush=True)
    assert st["iterations"] == so["iterations"]
    assert np.array_equal(x, o.get("x")), float(np.max(np.abs(x - o.get("x"))))
    assert np.array_equal(v, o.get("v"))
