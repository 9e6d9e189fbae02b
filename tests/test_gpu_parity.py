"""GPU parity: libsofa_amd.so (HIP, through the C ABI) against the FP64 CPU oracle.

Bars (stated per test):
* FP64 component calls (addForce / addDForce): bit-exact against the oracle -- same operation
  order, contraction off on both sides.
* FP64 trajectories: bit-exact against the oracle replaying the GPU's dot-product order, and
  <= 1e-10 relative (max node error / max node norm) against the reference's sequential order.
* FP32: <= 1e-5 relative position error after the stated number of steps.
"""
import numpy as np
import pytest

from sofa_amd import meshes

pytestmark = pytest.mark.gpu


def _capi():
    from sofa_amd import capi
    return capi


def beam_cfg0():
    # 5x5x21 grid over [0,1]x[0,1]x[0,4], 5-tet split, E=1e5, nu=0.45,
    # UniformMass totalMass = 1000 kg/m^3 * 4 m^3, z=0 face fixed, g=(0,-9.81,0),
    # rs = rm = 0.1, dt = 0.01, CG 25 / 1e-9 / 1e-9
    pos, tets, fixed = meshes.beam((5, 5, 21), (1.0, 1.0, 4.0))
    return dict(pos=pos, tets=tets, fixed=fixed, E=1e5, nu=0.45, total_mass=4000.0, g=(0.0, -9.81, 0.0),
                rs=0.1, rm=0.1, dt=0.01, iters=25, tol=1e-9, thr=1e-9)


def make_oracle(c, dot_mode=0, diagonal=None):
    from oracle import OracleScene
    s = OracleScene(c["pos"])
    s.set_gravity(c["g"])
    s.set_fem(c["tets"], [c["E"]], [c["nu"]], "large")
    if diagonal is not None:
        s.set_diagonal_mass(diagonal)
    else:
        s.set_uniform_mass(total_mass=c["total_mass"])
    s.set_fixed(c["fixed"])
    s.set_solver(c["rs"], c["rm"], c["iters"], c["tol"], c["thr"])
    s.set_dot_mode(dot_mode)
    return s


def add_gpu_scene(eng, c, diagonal=None):
    capi = _capi()
    n = len(c["pos"])
    if diagonal is not None:
        kind, val = capi.MASS_DIAGONAL, diagonal
    else:
        kind, val = capi.MASS_UNIFORM, c["total_mass"] / float(n)
    return eng.add_scene(c["pos"], c["tets"], [c["E"]], [c["nu"]], c["fixed"], mass_kind=kind, mass_value=val,
                         gravity=c["g"], dt=c["dt"], rayleigh_stiffness=c["rs"], rayleigh_mass=c["rm"],
                         cg_iterations=c["iters"], cg_tolerance=c["tol"], cg_threshold=c["thr"])


def rel_err(a, ref):
    return float(np.max(np.linalg.norm(a - ref, axis=1)) / np.max(np.linalg.norm(ref, axis=1)))


def test_fp64_component_calls_bitexact():
    """addForce then addDForce on a deformed state: bit-exact vs the oracle."""
    import torch
    capi = _capi()
    print("case: fp64 component calls, cfg0 mesh", flush=True)
    c = beam_cfg0()
    rng = np.random.default_rng(1234)
    n = len(c["pos"])
    x = c["pos"] + 0.02 * rng.standard_normal((n, 3))
    dx = rng.standard_normal((n, 3))
    f0 = rng.standard_normal((n, 3))
    k = -0.01 * (0.01 + 0.1)
    orc = make_oracle(c)
    f_ref = orc.fem_add_force(x, f0)
    df_ref = orc.fem_add_dforce(dx, f0, k)

    eng = capi.Engine(0, capi.FP64)
    sid = add_gpu_scene(eng, c)
    eng.commit()
    xd = torch.tensor(x, dtype=torch.float64, device="cuda")
    fd = torch.tensor(f0, dtype=torch.float64, device="cuda")
    dxd = torch.tensor(dx, dtype=torch.float64, device="cuda")
    dfd = torch.tensor(f0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    eng.fem_add_force(sid, xd.data_ptr(), fd.data_ptr())
    eng.fem_add_dforce(sid, dxd.data_ptr(), dfd.data_ptr(), k)
    eng.sync()
    f_gpu = fd.cpu().numpy()
    df_gpu = dfd.cpu().numpy()
    assert np.array_equal(f_gpu, f_ref), np.max(np.abs(f_gpu - f_ref))
    assert np.array_equal(df_gpu, df_ref), np.max(np.abs(df_gpu - df_ref))
    # the element rotations updated by addForce match as well
    for e in (0, 17, len(c["tets"]) - 1):
        rot, kk, rest = eng.element(sid, e)
        el = orc.fem_element(e)
        assert np.array_equal(rot, el["R"])
        assert np.array_equal(rest[1:], el["rest"][1:])
        assert kk[0] == el["K"][0, 0] and kk[1] == el["K"][0, 1] and kk[2] == el["K"][3, 3]
    eng.close()


@pytest.mark.parametrize("nsteps", [1, 100])
def test_fp64_trajectory_cfg0(nsteps):
    """cfg0 beam, FP64: bit-exact vs the oracle in GPU dot order; <= 1e-10 vs reference order."""
    capi = _capi()
    print(f"case: fp64 trajectory cfg0, {nsteps} steps", flush=True)
    c = beam_cfg0()
    eng = capi.Engine(0, capi.FP64)
    sid = add_gpu_scene(eng, c)
    eng.commit()
    eng.step([sid], nsteps)
    x, v, f = eng.read(sid)
    st = eng.stats(sid)
    orc_gpu_order = make_oracle(c, dot_mode=1)
    orc_gpu_order.step(c["dt"], nsteps)
    xo, vo = orc_gpu_order.get("x"), orc_gpu_order.get("v")
    so = orc_gpu_order.stats()
    assert st["iterations"] == so["iterations"] and st["end_reason"] == so["endcond"], (st, so)
    assert np.array_equal(x, xo), rel_err(x, xo)
    assert np.array_equal(v, vo), rel_err(v, vo)
    orc_ref = make_oracle(c, dot_mode=0)
    orc_ref.step(c["dt"], nsteps)
    assert rel_err(x, orc_ref.get("x")) <= 1e-10
    vr = orc_ref.get("v")
    assert float(np.max(np.linalg.norm(v - vr, axis=1)) / np.max(np.linalg.norm(vr, axis=1))) <= 1e-10
    eng.close()


def test_fp64_golden_checkFEMValues():
    """The reference's checkFEMValues scene on the GPU: position[159] after 100 steps
    (BaseTetrahedronFEMForceField_test.h:379-431) within its 1e-4 tolerance."""
    capi = _capi()
    print("case: golden checkFEMValues scene on GPU", flush=True)
    pos, hexa = meshes.regular_grid((4, 10, 4), (0, 0, 20), (10, 40, 30))
    tets = meshes.split_hexa_6(hexa)
    fixed = meshes.box_roi(pos, (-1, -1, 0, 10, 1, 50))
    c = dict(pos=pos, tets=tets, fixed=fixed, E=600.0, nu=0.3, total_mass=None, g=(0.0, 10.0, 0.0),
             rs=0.0, rm=0.0, dt=0.01, iters=20, tol=1e-5, thr=1e-6)
    eng = capi.Engine(0, capi.FP64)
    sid = add_gpu_scene(eng, c, diagonal=1.0)
    eng.commit()
    eng.step([sid], 100)
    x, v, f = eng.read(sid)
    np.testing.assert_allclose(x[159], [9.99985, 45.0487, 30.0011], atol=1e-4)
    orc = make_oracle(c, dot_mode=1, diagonal=1.0)
    orc.step(c["dt"], 100)
    assert np.array_equal(x, orc.get("x"))
    rot, kk, rest = eng.element(sid, 100)
    exp_rot = np.array([[-1, 8.01488e-06, 0.000541687], [-0.000320764, -0.814541, -0.580106],
                        [0.000436576, -0.580106, 0.814541]])
    np.testing.assert_allclose(rot, exp_rot, atol=1e-4)
    eng.close()


def test_fp32_trajectory_cfg0():
    """cfg0 beam in FP32 (compact element record): <= 1e-5 relative after 100 steps."""
    capi = _capi()
    print("case: fp32 trajectory cfg0, 100 steps", flush=True)
    c = beam_cfg0()
    eng = capi.Engine(0, capi.FP32)
    sid = add_gpu_scene(eng, c)
    eng.commit()
    eng.step([sid], 100)
    x, v, f = eng.read(sid)
    orc = make_oracle(c)
    orc.step(c["dt"], 100)
    err = rel_err(x, orc.get("x"))
    print("fp32 rel err", err, flush=True)
    assert err <= 1e-5
    eng.close()


def test_batched_equals_alone():
    """Three different beams stepped in one merged launch give the same bits as alone."""
    capi = _capi()
    print("case: batched vs alone, 3 beams, 20 steps", flush=True)
    cfgs = []
    for n, E in (((3, 4, 9), 2e4), ((5, 5, 21), 1e5), ((4, 3, 13), 5e5)):
        pos, tets, fixed = meshes.beam(n, (1.0, 1.0, 4.0))
        cfgs.append(dict(pos=pos, tets=tets, fixed=fixed, E=E, nu=0.4, total_mass=4000.0, g=(0.0, -9.81, 0.0),
                         rs=0.1, rm=0.1, dt=0.01, iters=25, tol=1e-9, thr=1e-9))
    eng = capi.Engine(0, capi.FP64)
    ids = [add_gpu_scene(eng, c) for c in cfgs]
    eng.commit()
    eng.step(ids, 20)
    merged = [eng.read(i)[0] for i in ids]
    eng.close()
    for c, xm in zip(cfgs, merged):
        e1 = capi.Engine(0, capi.FP64)
        s1 = add_gpu_scene(e1, c)
        e1.commit()
        e1.step([s1], 20)
        assert np.array_equal(e1.read(s1)[0], xm)
        e1.close()


# ------------------------------------------------------------------------------------------
# tiled LDS-gather path
# ------------------------------------------------------------------------------------------
def beam_medium(E=1e5):
    pos, tets, fixed = meshes.beam((9, 9, 41), (1.0, 1.0, 4.0))
    return dict(pos=pos, tets=tets, fixed=fixed, E=E, nu=0.45, total_mass=4000.0, g=(0.0, -9.81, 0.0),
                rs=0.1, rm=0.1, dt=0.01, iters=25, tol=1e-9, thr=1e-9)


@pytest.mark.parametrize("cfgname,nsteps", [("cfg0", 100), ("medium", 20)])
def test_fp64_tiled_trajectory(cfgname, nsteps):
    """FP64 on the tiled kernels (shared nodes summed tile by tile): <= 1e-10 relative
    position and velocity error against the oracle in the reference's order."""
    capi = _capi()
    print(f"case: fp64 tiled trajectory {cfgname}, {nsteps} steps", flush=True)
    c = beam_cfg0() if cfgname == "cfg0" else beam_medium()
    eng = capi.Engine(0, capi.FP64_TILED)
    sid = add_gpu_scene(eng, c)
    eng.commit()
    eng.step([sid], nsteps)
    x, v, f = eng.read(sid)
    st = eng.stats(sid)
    orc = make_oracle(c)
    orc.step(c["dt"], nsteps)
    so = orc.stats()
    assert st["iterations"] == so["iterations"], (st, so)
    assert rel_err(x, orc.get("x")) <= 1e-10
    vr = orc.get("v")
    assert float(np.max(np.linalg.norm(v - vr, axis=1)) / np.max(np.linalg.norm(vr, axis=1))) <= 1e-10
    eng.close()


def test_fp32_tiled_medium():
    """FP32 tiled path on a 12,800-tet beam (25 tiles): <= 2e-5 relative after 20 steps."""
    capi = _capi()
    print("case: fp32 tiled medium, 20 steps", flush=True)
    c = beam_medium()
    eng = capi.Engine(0, capi.FP32)
    sid = add_gpu_scene(eng, c)
    eng.commit()
    eng.step([sid], 20)
    x = eng.read(sid)[0]
    orc = make_oracle(c)
    orc.step(c["dt"], 20)
    err = rel_err(x, orc.get("x"))
    print("fp32 tiled rel err", err, flush=True)
    assert err <= 2e-5
    eng.close()


def test_tiled_golden_checkFEMValues():
    capi = _capi()
    print("case: golden checkFEMValues scene, fp64 tiled", flush=True)
    pos, hexa = meshes.regular_grid((4, 10, 4), (0, 0, 20), (10, 40, 30))
    tets = meshes.split_hexa_6(hexa)
    fixed = meshes.box_roi(pos, (-1, -1, 0, 10, 1, 50))
    c = dict(pos=pos, tets=tets, fixed=fixed, E=600.0, nu=0.3, total_mass=None, g=(0.0, 10.0, 0.0),
             rs=0.0, rm=0.0, dt=0.01, iters=20, tol=1e-5, thr=1e-6)
    eng = capi.Engine(0, capi.FP64_TILED)
    sid = add_gpu_scene(eng, c, diagonal=1.0)
    eng.commit()
    eng.step([sid], 100)
    x = eng.read(sid)[0]
    np.testing.assert_allclose(x[159], [9.99985, 45.0487, 30.0011], atol=1e-4)
    eng.close()


@pytest.mark.parametrize("prec", ["FP32", "FP64_TILED"])
def test_tiled_batched_equals_alone(prec):
    """Beams of different sizes stepped in one merged launch give the same bits as alone."""
    capi = _capi()
    print(f"case: tiled batched vs alone ({prec}), 4 beams, 10 steps", flush=True)
    cfgs = []
    for n, E in (((3, 4, 9), 2e4), ((5, 5, 21), 1e5), ((9, 9, 41), 3e5), ((4, 3, 13), 5e5)):
        pos, tets, fixed = meshes.beam(n, (1.0, 1.0, 4.0))
        cfgs.append(dict(pos=pos, tets=tets, fixed=fixed, E=E, nu=0.4, total_mass=4000.0, g=(0.0, -9.81, 0.0),
                         rs=0.1, rm=0.1, dt=0.01, iters=25, tol=1e-9, thr=1e-9))
    P = getattr(capi, prec)
    eng = capi.Engine(0, P)
    ids = [add_gpu_scene(eng, c) for c in cfgs]
    eng.commit()
    eng.step(ids, 10)
    merged = [eng.read(i)[0] for i in ids]
    eng.close()
    for c, xm in zip(cfgs, merged):
        e1 = capi.Engine(0, P)
        s1 = add_gpu_scene(e1, c)
        e1.commit()
        e1.step([s1], 10)
        assert np.array_equal(e1.read(s1)[0], xm)
        e1.close()


def test_tiled_component_calls_and_element_map():
    """On a tiled engine the component calls still sum in the reference's element order
    (bit-exact in FP64) and element ids map back to the caller's order."""
    import torch
    capi = _capi()
    print("case: fp64 tiled component calls", flush=True)
    c = beam_medium()
    rng = np.random.default_rng(99)
    n = len(c["pos"])
    x = c["pos"] + 0.02 * rng.standard_normal((n, 3))
    dx = rng.standard_normal((n, 3))
    k = -0.0011
    orc = make_oracle(c)
    f_ref = orc.fem_add_force(x, np.zeros((n, 3)))
    df_ref = orc.fem_add_dforce(dx, np.zeros((n, 3)), k)
    eng = capi.Engine(0, capi.FP64_TILED)
    sid = add_gpu_scene(eng, c)
    eng.commit()
    xd = torch.tensor(x, dtype=torch.float64, device="cuda")
    fd = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    dxd = torch.tensor(dx, dtype=torch.float64, device="cuda")
    dfd = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    eng.fem_add_force(sid, xd.data_ptr(), fd.data_ptr())
    eng.fem_add_dforce(sid, dxd.data_ptr(), dfd.data_ptr(), k)
    eng.sync()
    assert np.array_equal(fd.cpu().numpy(), f_ref)
    assert np.array_equal(dfd.cpu().numpy(), df_ref)
    for e in (0, 1234, len(c["tets"]) - 1):
        rot, kk, rest = eng.element(sid, e)
        assert np.array_equal(rot, orc.fem_element(e)["R"])
    eng.close()


@pytest.mark.parametrize("prec", ["FP64", "FP32", "FP64_TILED"])
def test_graph_replay_equals_eager(prec):
    """A step captured as a hipGraph and replayed gives, replay after replay, the same bits as
    the eager launches (eager reads and synchronisations between the replays)."""
    capi = _capi()
    print(f"case: graph replay vs eager ({prec}), 2 beams, 6 x 3 steps", flush=True)
    cfgs = [beam_cfg0(), beam_medium(3e5)]
    P = getattr(capi, prec)
    engines = []
    for graphs in (False, True):
        eng = capi.Engine(0, P)
        ids = [add_gpu_scene(eng, c) for c in cfgs]
        eng.commit()
        eng.set_graphs(graphs)
        engines.append((eng, ids))
    (eager, ids_e), (graph, ids_g) = engines
    for rep in range(6):
        eager.step(ids_e, 3)
        graph.step(ids_g, 3)
        for se, sg in zip(ids_e, ids_g):
            xe, ve, fe = eager.read(se)
            xg, vg, fg = graph.read(sg)
            assert np.array_equal(xe, xg) and np.array_equal(ve, vg), (rep, se)
            assert eager.stats(se) == graph.stats(sg)
    # a different launch set drops and recaptures the graph
    eager.step(ids_e[:1], 2)
    graph.step(ids_g[:1], 2)
    assert np.array_equal(eager.read(ids_e[0])[0], graph.read(ids_g[0])[0])
    eager.close()
    graph.close()
