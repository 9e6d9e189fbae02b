"""Host-side workload and serving logic (CPU only): job streams, meshes, footprints."""
import numpy as np

from sofa_amd import workloads


def test_cfg4_stream_statistics():
    jobs = workloads.cfg4_jobs(rate=200.0, duration=20.0, seed=4)
    arr = np.array([j["arrival"] for j in jobs])
    # Poisson process: count ~ 4000 +- 4 sigma, increasing arrivals, exponential gaps (mean 5 ms)
    assert abs(len(jobs) - 4000) < 4 * np.sqrt(4000)
    assert np.all(np.diff(arr) > 0) and arr[-1] < 20.0
    assert abs(np.diff(arr).mean() - 0.005) < 0.0005
    steps = np.array([j["steps"] for j in jobs])
    assert steps.min() >= 50 and steps.max() <= 500
    kinds = [j["kind"] for j in jobs]
    assert 0.45 < kinds.count("beam") / len(kinds) < 0.55
    assert all(0.0 <= j["amp"] <= 0.01 and 0.5 <= j["freq"] <= 5.0 for j in jobs)
    # the same seed gives the same stream
    again = workloads.cfg4_jobs(rate=200.0, duration=20.0, seed=4)
    assert [j["arrival"] for j in again] == list(arr)


def test_cfg4_job_meshes():
    jobs = workloads.cfg4_jobs(rate=200.0, duration=0.2, seed=11)
    for job in jobs[:12]:
        spec = workloads.build_job(job)
        pos, tets = spec["pos"], spec["tets"]
        assert tets.min() >= 0 and tets.max() < len(pos)
        assert np.unique(tets.ravel()).size == len(pos)          # no unused node
        A = pos[tets]
        vol = np.einsum("ij,ij->i", np.cross(A[:, 1] - A[:, 0], A[:, 2] - A[:, 0]), A[:, 3] - A[:, 0])
        assert np.all(np.abs(vol) > 0)
        osc = spec["oscillators"]
        assert len(osc["index"]) > 0 and len(spec["fixed"]) > 0
        assert not set(osc["index"]) & set(spec["fixed"].tolist())
        if job["kind"] == "beam":
            assert spec["mass"][0] == "uniform"
        else:
            assert spec["mass"][0] == "meshmatrix_lumped"
            assert 0.3 * job["n_target"] < len(pos) < 3.0 * job["n_target"]


def test_ellipsoid_delaunay_mesh():
    pos, tets = workloads.ellipsoid(pitch=0.012, seed=2)
    a = np.array([0.15, 0.10, 0.08])
    assert np.all(((pos / a) ** 2).sum(1) <= 1.0 + 1e-12)
    A = pos[tets]
    vol6 = np.einsum("ij,ij->i", np.cross(A[:, 1] - A[:, 0], A[:, 2] - A[:, 0]), A[:, 3] - A[:, 0])
    assert np.all(vol6 >= 0.012 ** 3 / 20.0)                    # positively oriented, no slivers
    assert np.unique(tets.ravel()).size == len(pos)
    # the kept tetrahedra fill most of the ellipsoid's volume
    assert vol6.sum() / 6.0 > 0.85 * (4.0 / 3.0 * np.pi * a.prod())


def test_footprint_monotonic():
    from sofa_amd import capi, serving
    for prec in (capi.FP32, capi.FP64_TILED, capi.MIXED):
        a = serving.footprint_bytes(prec, 1000, 5000)
        assert serving.footprint_bytes(prec, 2000, 5000) > a and serving.footprint_bytes(prec, 1000, 6000) > a
    assert serving.footprint_bytes(capi.FP64_TILED, 1000, 5000) > serving.footprint_bytes(capi.FP32, 1000, 5000)
    assert serving.footprint_bytes(capi.MIXED, 1000, 5000) > serving.footprint_bytes(capi.FP64_TILED, 1000, 5000)
