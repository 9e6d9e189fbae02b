"""GPU parity of the tiled kernels in the reference's summation order, and of the precision split.

* FP64_TILED with the reference summation order (sofa_amd_engine_set_dot_order(eng, 1) before the
  scenes are placed): positions, velocities and CG statistics equal the oracle's reference-order
  trajectory (dot_mode 0: MechanicalObject::vDot's sequential sums, element contributions in
  ascending element id) bit for bit -- alone and batched with other beams of the 256-beam mix,
  with graphs, with DiagonalMass, lumped MeshMatrixMass, oscillators and a PlaneForceField.
* SPLIT (FP32 element arithmetic on FP64 node vectors, CG and state): error against the same
  reference-order trajectory, stated per beam.
"""
import numpy as np
import pytest

from specs import oracle_from_spec

pytestmark = pytest.mark.gpu

# beams of the 256-beam mix (workloads.cfg3(), seed 3) small enough for the oracle to run 100
# steps in seconds; 244 is a beam whose trajectory moves by ~1e-5 when only the association of
# its dot-product sums changes
MIX_BEAMS = [53, 153, 176, 244, 0, 39, 165, 68]


def _capi():
    from sofa_amd import capi
    return capi


def rel_err(a, ref):
    return float(np.max(np.linalg.norm(a - ref, axis=1)) / np.max(np.linalg.norm(ref, axis=1)))


def _specs(which):
    from sofa_amd import workloads
    if which == "cfg0":
        return [workloads.cfg0()], 100
    if which == "sensitive_beam":
        return [workloads.cfg3()[244]], 100
    if which == "diagonal_plane":
        s = workloads.cfg0()
        s["mass"] = ("diagonal", 1000.0)
        s["plane"] = dict(normal=(0.0, 1.0, 0.0), d=-0.3, stiffness=2000.0, damping=5.0)
        return [s], 60
    if which == "lumped_oscillators":
        s = workloads.build_job(dict(kind="ellipsoid", n_target=1500, E=3e4, nu=0.45, amp=0.004, freq=2.0))
        return [s], 40
    mix = workloads.cfg3()
    return [mix[i] for i in MIX_BEAMS], 100


def _run(prec, specs, nsteps, reference_order=False, graphs=True, branches=0, turns=False, segments=None):
    from sofa_amd import workloads
    capi = _capi()
    eng = capi.Engine(0, getattr(capi, prec))
    if reference_order:
        eng.set_dot_order(True)
    if segments is not None:
        eng.set_dot_segments(*segments)
    ids = [workloads.add_to_engine(eng, s) for s in specs]
    eng.commit()
    eng.set_graphs(graphs)
    eng.set_branches(branches)
    eng.set_branch_turns(turns)
    eng.step(ids, nsteps)
    out = [eng.read(i)[:2] + (eng.stats(i),) for i in ids]
    eng.close()
    return out


@pytest.mark.parametrize("which", ["cfg0", "sensitive_beam", "diagonal_plane", "lumped_oscillators", "mix8"])
def test_exact_tiled_bitexact(which):
    """FP64 tiled kernels in the reference's summation order: every scene's positions,
    velocities, CG iteration count and end reason equal the oracle's reference-order run bit
    for bit (mix8: eight beams of the mix batched in one engine, each against its own oracle)."""
    specs, nsteps = _specs(which)
    print(f"case: exact tiled {which}, {len(specs)} scene(s), {nsteps} steps", flush=True)
    got = _run("FP64_TILED", specs, nsteps, reference_order=True)
    # the oscillator's sin / cos come from the device's libm on the GPU and from the host's in the
    # oracle (not bit-identical): there the reference is the FP64 parity engine in the sequential
    # dot order (bit-exact against the oracle wherever no oscillator runs), the oracle within 1e-10
    osc = which == "lumped_oscillators"
    ref_gpu = _run("FP64", specs, nsteps, reference_order=True) if osc else None
    for k, (spec, (x, v, st)) in enumerate(zip(specs, got)):
        o = oracle_from_spec(spec, dot_mode=0)
        o.step(spec["dt"], nsteps)
        xo, vo = o.get("x"), o.get("v")
        so = o.stats()
        if osc:
            xg, vg, sg = ref_gpu[k]
            assert np.array_equal(x, xg), rel_err(x, xg)
            assert np.array_equal(v, vg), rel_err(v, vg)
            assert st["iterations"] == sg["iterations"] and st["end_reason"] == sg["end_reason"]
            assert rel_err(x, xo) <= 1e-10, rel_err(x, xo)
        else:
            assert np.array_equal(x, xo), rel_err(x, xo)
            assert np.array_equal(v, vo), rel_err(v, vo)
            assert st["iterations"] == so["iterations"] and st["end_reason"] == so["endcond"]
        assert rel_err(xo, spec["pos"]) > 1e-6          # the scene moved


def test_exact_tiled_eager_equals_graph():
    """The reference-order step replayed from a graph equals the eager launches bit for bit."""
    specs, _ = _specs("mix8")
    a = _run("FP64_TILED", specs[:3], 20, reference_order=True, graphs=False)
    b = _run("FP64_TILED", specs[:3], 20, reference_order=True, graphs=True)
    for (xa, va, _), (xb, vb, _) in zip(a, b):
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)


@pytest.mark.parametrize("mode", ["turns", "free_eager", "free_graphs"])
def test_exact_tiled_branches(mode):
    """A merged step in concurrent branches (scenes split over streams) equals the one-branch step
    bit for bit, with the branches' operator phases taking turns and running freely (eager and
    replayed from graphs): scenes are independent.  The scenes' CG budgets differ (10 to 25
    iterations), so that groups end their iterations at different turns."""
    specs, _ = _specs("mix8")
    specs = [dict(s, iters=10 + 5 * (k % 4)) for k, s in enumerate(specs)]
    graphs, turns = mode == "free_graphs", mode == "turns"
    print(f"case: exact tiled, 8 beams, branches 1 / 2 / 3 / 4, {mode}", flush=True)
    one = _run("FP64_TILED", specs, 15, reference_order=True, graphs=graphs, branches=1, turns=turns)
    for br in (2, 3, 4):
        got = _run("FP64_TILED", specs, 15, reference_order=True, graphs=graphs, branches=br, turns=turns)
        for (xa, va, sa), (xb, vb, sb) in zip(one, got):
            assert np.array_equal(xa, xb) and np.array_equal(va, vb)
            assert sa["iterations"] == sb["iterations"]


# beams of the 256-beam mix above the 8192 nodes from which the segmented dot products apply
# (8528, 9088 and 9104 nodes), and one below
SEG_BEAMS = [67, 151, 10, 53]


def test_exact_tiled_dot_segments():
    """Reference-order dot products with the chain's segments run at once from candidate starts
    give the single chain's bits: 2, 4, 8 and 16 segments, windows of 32 to 128 candidates per side,
    and window 0, where every boundary falls back to summing its segment from the known start.
    The largest beam's trajectory also equals the oracle's reference-order run bit for bit."""
    from sofa_amd import workloads
    mix = workloads.cfg3()
    specs = [mix[i] for i in SEG_BEAMS]
    nsteps = 20
    print(f"case: exact tiled dot segments, beams {SEG_BEAMS}, {nsteps} steps", flush=True)
    one = _run("FP64_TILED", specs, nsteps, reference_order=True, segments=(1, 128))
    for seg in [(4, 128), (8, 0), (2, 32), (8, 96), (8, 32), (16, 32), (16, 0)]:
        got = _run("FP64_TILED", specs, nsteps, reference_order=True, segments=seg)
        for k, ((xa, va, sa), (xb, vb, sb)) in enumerate(zip(one, got)):
            assert np.array_equal(xa, xb) and np.array_equal(va, vb), (seg, k, rel_err(xa, xb))
            assert sa["iterations"] == sb["iterations"] and sa["end_reason"] == sb["end_reason"]
            small = len(specs[k]["pos"]) < 8192
            # per step |b|^2 and per CG iteration p.q and r.r, K - 1 boundaries each
            bounds = (seg[0] - 1) * nsteps * (1 + 2 * sb["iterations"])
            print(f"  segments {seg}: scene {k} fallbacks {sb['dot_fallbacks']} of about {bounds} boundaries",
                  flush=True)
            if small:
                assert sb["dot_fallbacks"] == 0
            elif seg[1] == 0:
                assert sb["dot_fallbacks"] >= bounds // 2
            else:   # the modelled prediction lands within the candidates (0 misses measured)
                assert sb["dot_fallbacks"] <= bounds // 10
    o = oracle_from_spec(specs[2], dot_mode=0)
    o.step(specs[2]["dt"], nsteps)
    x, v, st = one[2]
    assert np.array_equal(got[2][0], o.get("x")), rel_err(got[2][0], o.get("x"))
    assert np.array_equal(got[2][1], o.get("v"))


def test_exact_tiled_refuses_consistent_mass():
    """A consistent MeshMatrixMass is not summed in the reference's order by the tiled kernels:
    the engine refuses it at commit instead of running another order."""
    from sofa_amd import workloads
    capi = _capi()
    s = workloads.cfg0()
    s["mass"] = ("meshmatrix", 1000.0)
    eng = capi.Engine(0, capi.FP64_TILED)
    eng.set_dot_order(True)
    workloads.add_to_engine(eng, s)
    with pytest.raises(capi.SofaAmdError) as ei:
        eng.commit()
    assert ei.value.code == capi.EINVAL
    eng.close()


def test_split_precision_mix():
    """SPLIT (FP32 element arithmetic, FP64 vectors / CG / state) on eight beams of the mix and
    the cfg0 cantilever, batched, against the reference-order trajectory after 100 steps.  Bars:
    cfg0 (CG converged to 1e-9 each step) <= 1e-6; each beam no worse than 1e-2 (the mix's CG 25
    stops far from convergence, where the trajectory amplifies any rounding: the oracle in
    another dot-product association alone moves some beams by 1e-5 to 1e-3).  The per-beam
    errors are printed beside the blocked-order FP64 oracle's."""
    from sofa_amd import workloads
    specs = [workloads.cfg0()] + _specs("mix8")[0]
    got = _run("SPLIT", specs, 100)
    for k, (spec, (x, _, _)) in enumerate(zip(specs, got)):
        o = oracle_from_spec(spec, dot_mode=0)
        o.step(spec["dt"], 100)
        ob = oracle_from_spec(spec, dot_mode=1)
        ob.step(spec["dt"], 100)
        e = rel_err(x, o.get("x"))
        eb = rel_err(ob.get("x"), o.get("x"))
        print(f"  scene {k}: split {e:.3e}  fp64 blocked order {eb:.3e}", flush=True)
        assert e <= (1e-6 if k == 0 else 1e-2), e
