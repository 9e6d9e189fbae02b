"""The allocation-free vDot (sofa_amd_vdot_noalloc) against sofa_amd_vdot, through the C ABI.

Both orders, two calls in a row on one stream: the first on a 21,525-node scene (85 block
partials), the second on a smaller 2,205-node scene of another engine (9 partials), so that the
second call runs over partials the first left in the shared static array.  Bar: bit-exact (the
same kernels in the same order).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _engine(capi, shape):
    from sofa_amd import meshes
    pos, tets, fixed = meshes.beam(shape, (1.0, 1.0, 4.0))
    eng = capi.Engine(0, capi.FP64)
    sid = eng.add_scene(pos, tets, [1e5], [0.45], fixed, mass_value=1.0)
    eng.commit()
    return eng, sid, len(pos)


def test_vdot_noalloc_matches_vdot():
    from sofa_amd import capi
    print("case: vdot_noalloc vs vdot, scenes of 21,525 then 2,205 nodes", flush=True)
    big, sb, nb = _engine(capi, (15, 15, 97))
    small, ss, ns = _engine(capi, (7, 7, 45))
    assert nb > ns
    rng = np.random.default_rng(23)
    stream = capi.Stream()
    for order in (0, 1):
        out = capi.DeviceArray(np.zeros(4))
        vecs = []
        for eng, sid, n, k in ((big, sb, nb, 0), (small, ss, ns, 2)):
            a = capi.DeviceArray(rng.standard_normal((n, 3)))
            b = capi.DeviceArray(rng.standard_normal((n, 3)))
            vecs += [a, b]
            eng.vdot(sid, a.ptr, b.ptr, order, out.ptr + 8 * k, stream=stream)
            eng.vdot_noalloc(sid, a.ptr, b.ptr, order, out.ptr + 8 * (k + 1), stream=stream)
        stream.synchronize()
        got = out.numpy()
        print(f"  order {order}: {got}", flush=True)
        assert got[0] == got[1] and got[2] == got[3], got
        assert got[0] != 0.0 and got[2] != 0.0
        for v in vecs:
            v.close()
        out.close()
    stream.close()
    big.close()
    small.close()


def test_vdot_noalloc_refuses_invalid_arguments():
    """Null vectors and an unknown order are refused with SOFA_AMD_EINVAL before any launch (the
    node-count bound of the static buffer is checked in the same place)."""
    from sofa_amd import capi
    eng, sid, n = _engine(capi, (5, 5, 21))
    out = capi.DeviceArray(np.zeros(1))
    a = capi.DeviceArray(np.ones((n, 3)))
    for args in ((0, 0, 0), (a.ptr, a.ptr, 2)):
        with pytest.raises(capi.SofaAmdError) as ei:
            eng.vdot_noalloc(sid, args[0], args[1], args[2], out.ptr)
        assert ei.value.code == capi.EINVAL
    a.close()
    out.close()
    eng.close()
