"""Host-side workload and serving logic (CPU only): job streams, meshes, footprints."""
import numpy as np
import pytest

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
    # the reference summation order adds 4 x 2 slots of 3 doubles and 4 x 2 rank bytes per
    # tetrahedron, a dot term per node, and per scene the segmented dot products' 2 x 64 prediction
    # chunks and 1 + 7 (64 x 4 + 1) candidate sums (doubles)
    d = serving.footprint_bytes(capi.FP64_TILED, 1000, 5000, True) - serving.footprint_bytes(capi.FP64_TILED, 1000, 5000)
    assert d == 5000 * (4 * 2 * 3 * 8 + 8) + 1000 * 8 + 8 * (2 * 64 + 1 + 7 * (64 * 4 + 1))
    # the arenas are sized per tetrahedron, without the per-scene part
    per_tet = serving.footprint_bytes(capi.FP64_TILED, 0, 1, True, scenes=0)
    assert per_tet == serving.footprint_bytes(capi.FP64_TILED, 0, 2, True) - serving.footprint_bytes(capi.FP64_TILED, 0, 1, True)
    assert per_tet < 1024


class _FakeEngine:
    """Host stand-in for a live engine: arenas as node / tetrahedron counters, a step log."""

    def __init__(self, max_nodes, max_tets):
        self.cap = (max_nodes, max_tets)
        self.used = [0, 0]
        self.scenes = []          # (nodes, tets)
        self.placed = set()
        self.steps = {}
        self.log = []

    def add_scene(self, positions, tets, *a, **k):
        self.scenes.append((len(positions), len(tets)))
        return len(self.scenes) - 1

    def set_refinement(self, *a, **k):
        pass

    def prepare_scene(self, sid):
        self.prepared = getattr(self, "prepared", set()) | {sid}

    def commit_scene(self, sid):
        n, t = self.scenes[sid]
        if self.used[0] + n > self.cap[0] or self.used[1] + t > self.cap[1]:
            return False
        self.used[0] += n
        self.used[1] += t
        self.placed.add(sid)
        self.steps[sid] = 0
        return True

    def step(self, ids, n_steps=1, stream=None):
        import time
        time.sleep(0.004)                                          # the device's time per tick
        assert all(i in self.placed for i in ids)
        self.log.append(tuple(ids))
        for i in ids:
            self.steps[i] += n_steps

    def release_scene(self, sid):
        n, t = self.scenes[sid]
        self.placed.remove(sid)
        self.used[0] -= n
        self.used[1] -= t

    def close(self):
        pass


class _FakeEvent:
    def record(self, stream=None):
        import time
        self.t = time.perf_counter()

    def elapsed_time(self, other):
        return (other.t - self.t) * 1e3

    def synchronize(self):
        pass


def _tiny_jobs(n, seed=1):
    rng = np.random.default_rng(seed)
    return [dict(arrival=0.002 * k, steps=int(rng.integers(2, 6)), kind="beam",
                 dims=(3, 3, int(rng.integers(3, 8))), E=1e5, nu=0.3, amp=0.0, freq=1.0) for k in range(n)]


def test_serving_scheduler_batched():
    """Every job runs exactly its number of steps, merged ticks hold several jobs, the budget is
    never exceeded, arrivals are admitted first come first served."""
    from sofa_amd import capi, serving
    jobs = _tiny_jobs(24)
    budget = 3 * serving.footprint_bytes(capi.FP32, 3 * 3 * 8, 5 * 2 * 2 * 7)
    eng = _FakeEngine(10 ** 6, 10 ** 7)
    seen = []
    srv = serving.Server(capi.FP32, budget_bytes=budget, engine=eng, batched=True, build_workers=2)
    res = srv.run(jobs, wall_limit=30.0, event=_FakeEvent, on_finish=lambda e, s, j: seen.append((s, j)))
    srv.close()
    assert res["jobs_done"] == 24 and res["jobs_rejected"] == 0 and res["queue_at_end"] == 0
    assert sorted(j for _, j in seen) == list(range(24))
    for s, j in seen:
        assert eng.steps[s] == jobs[j]["steps"]
    assert max(len(t) for t in eng.log) >= 2                       # merged ticks
    for t in eng.log:                                              # the budget is never exceeded
        assert sum(serving.footprint_bytes(capi.FP32, *eng.scenes[s]) for s in t) <= budget
    first_tick = {}
    for k, t in enumerate(eng.log):
        for s in t:
            first_tick.setdefault(s, k)
    ticks_by_job = [first_tick[s] for s, j in sorted(seen, key=lambda p: p[1])]
    assert ticks_by_job == sorted(ticks_by_job)                    # FIFO admission
    assert res["element_steps"] == sum(eng.scenes[s][1] * eng.steps[s] for s, _ in seen)


def test_serving_scheduler_plain_and_refusal():
    """Plain mode steps one scene per launch; a job larger than the whole reservation is refused
    instead of blocking the queue forever."""
    from sofa_amd import capi, serving
    jobs = _tiny_jobs(6, seed=3)
    jobs[2] = dict(jobs[2], dims=(9, 9, 40))
    eng = _FakeEngine(2000, 10 ** 7)
    srv = serving.Server(capi.FP32, budget_bytes=1 << 40, engine=eng, batched=False, build_workers=2)
    res = srv.run(jobs, wall_limit=30.0, event=_FakeEvent)
    srv.close()
    assert res["jobs_rejected"] == 1 and res["jobs_done"] == 5
    assert all(len(t) == 1 for t in eng.log)


def test_serving_close_waits_for_preparations():
    """Server.close() while host preparations (tiling, element set-up) are still running on the
    worker threads: the engine they read is closed only after every running preparation has
    returned, and queued ones are dropped (they never start after the close)."""
    import threading
    import time
    from sofa_amd import capi, serving

    class SlowPrepareEngine(_FakeEngine):
        def __init__(self):
            super().__init__(10 ** 6, 10 ** 7)
            self.running = 0
            self.started = 0
            self.closed = False
            self.prepared_after_close = 0
            self.lock = threading.Lock()
            self.gate = threading.Event()

        def prepare_scene(self, sid):
            with self.lock:
                if self.closed:
                    self.prepared_after_close += 1
                self.running += 1
                self.started += 1
            self.gate.wait(2.0)
            time.sleep(0.05)
            with self.lock:
                self.running -= 1

        def close(self):
            with self.lock:
                assert self.running == 0, "engine closed while a preparation reads it"
                self.closed = True

    eng = SlowPrepareEngine()
    srv = serving.Server(capi.FP32, budget_bytes=1 << 40, engine=eng, build_workers=2)
    futs = [srv.pool.submit(eng.prepare_scene, k) for k in range(8)]
    while eng.started < 2:                       # two preparations in flight, six queued
        time.sleep(0.001)
    threading.Timer(0.05, eng.gate.set).start()
    srv.close()
    assert eng.closed and eng.running == 0
    assert eng.prepared_after_close == 0
    assert sum(f.cancelled() for f in futs) >= 1


def test_scene_reader_linear_movement():
    """LinearMovementProjectiveConstraint Data (indices through a BoxROI link, keyTimes,
    movements, relativeMovements) and its defaults."""
    from sofa_amd import scene_xml
    text = """<Node name="root" dt="0.01" gravity="0 -9.81 0">
      <Node name="beam">
        <EulerImplicitSolver rayleighStiffness="0.1" rayleighMass="0.1"/>
        <CGLinearSolver iterations="25" tolerance="1e-9" threshold="1e-9"/>
        <RegularGridTopology name="grid" n="3 3 5" min="0 0 0" max="1 1 2"/>
        <MechanicalObject name="mo"/>
        <TetrahedronFEMForceField youngModulus="1e5" poissonRatio="0.4" method="large"/>
        <UniformMass totalMass="10"/>
        <BoxROI name="top" box="-0.1 -0.1 1.9 1.1 1.1 2.1"/>
        <LinearMovementProjectiveConstraint indices="@top.indices" keyTimes="0 0.5 1"
            movements="0 0 0  0.1 0 0  0.1 0.2 0" relativeMovements="false"/>
        <FixedProjectiveConstraint indices="0 1 2"/>
      </Node>
    </Node>"""
    (spec,), info = scene_xml.load_scene(text)
    lm = spec["linear_movement"]
    top = np.nonzero(spec["pos"][:, 2] >= 1.9)[0]
    assert np.array_equal(np.sort(lm["indices"]), top) and len(top) == 9
    assert list(lm["key_times"]) == [0.0, 0.5, 1.0] and lm["relative"] is False
    assert np.array_equal(lm["movements"], [[0, 0, 0], [0.1, 0, 0], [0.1, 0.2, 0]])
    (spec2,), _ = scene_xml.load_scene(text.replace(
        '<LinearMovementProjectiveConstraint indices="@top.indices" keyTimes="0 0.5 1"\n'
        '            movements="0 0 0  0.1 0 0  0.1 0.2 0" relativeMovements="false"/>',
        "<LinearMovementProjectiveConstraint/>"))
    d = spec2["linear_movement"]
    assert list(d["indices"]) == [0] and list(d["key_times"]) == [0.0] and d["relative"] is True
    with pytest.raises(ValueError):
        scene_xml.load_scene(text.replace('keyTimes="0 0.5 1"', 'keyTimes="0 0.5"'))


def test_fem_hexa_tessellation():
    """TetrahedronFEMForceField::init's 6-tetrahedra tessellation of a grid (cube corners mirrored
    by cube parity): conforming (every interior triangle in exactly two tetrahedra, every boundary
    triangle in one and on the box), filling the box exactly, and neighbouring cubes cut as
    mirror images of each other across their shared face, along x, y and z alike -- which a wrong
    parity, edge pair or pattern breaks."""
    from collections import Counter
    from sofa_amd import meshes
    n = (5, 4, 4)
    L = (4.0, 3.0, 3.0)
    pos, hexa = meshes.regular_grid(n, (0, 0, 0), L)
    nx, ny, nz = n[0] - 1, n[1] - 1, n[2] - 1
    tets = meshes.split_hexa_fem(hexa, nx, ny)
    assert tets.shape == (6 * len(hexa), 4)
    A = pos[tets]
    vol = np.einsum("ij,ij->i", np.cross(A[:, 1] - A[:, 0], A[:, 2] - A[:, 0]), A[:, 3] - A[:, 0]) / 6.0
    assert np.all(np.abs(vol) > 1e-12) and abs(np.abs(vol).sum() - np.prod(L)) < 1e-9
    cnt = Counter()
    for t in tets:
        for f in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
            cnt[tuple(sorted(t[list(f)]))] += 1
    for f, c in cnt.items():
        P = pos[list(f)]
        on_box = any(np.all(P[:, d] == v) for d in range(3) for v in (0.0, L[d]))
        assert (c == 2 and not on_box) or (c == 1 and on_box), (f, c)

    def gidx(ix, iy, iz):
        return ix + n[0] * (iy + n[1] * iz)

    coords = np.stack(np.unravel_index(np.arange(len(pos)), (n[2], n[1], n[0]))[::-1], axis=1)   # (ix, iy, iz)
    cube_tets = [set(frozenset(t) for t in tets[6 * c:6 * c + 6]) for c in range(len(hexa))]
    checked = 0
    for c in range(len(hexa)):
        cx, cy, cz = c % nx, (c // nx) % ny, c // (nx * ny)
        for d, step in ((0, 1), (1, nx), (2, nx * ny)):
            if (cx, cy, cz)[d] + 1 >= (nx, ny, nz)[d]:
                continue
            plane = (cx, cy, cz)[d] + 1                     # shared face at grid index `plane`
            mapped = set()
            for t in cube_tets[c + step]:
                m = []
                for v in t:
                    q = list(coords[v])
                    q[d] = 2 * plane - q[d]
                    m.append(gidx(*q))
                mapped.add(frozenset(m))
            assert mapped == cube_tets[c], (c, d)
            checked += 1
    assert checked == 75
    # absolute parity: cube 1 (odd x, even y and z) is cut by the plain 6-tetrahedra pattern,
    # cube 0 (even x) by that pattern reflected across its own mid-plane x = 1/2
    plain = meshes.split_hexa_6(hexa)
    assert cube_tets[1] == set(frozenset(t) for t in plain[6:12])
    refl = set()
    for t in plain[0:6]:
        refl.add(frozenset(gidx(1 - coords[v][0], coords[v][1], coords[v][2]) for v in t))
    assert cube_tets[0] == refl


def test_scene_reader_benchmark_scenes():
    """The SOFA scene reader on the SofaCUDA TetrahedronFEMForceField benchmark configurations:
    grid sizes, the 6-tetrahedra split, the BoxROI-clamped face, DiagonalMass totalMass, the solver
    Data and the PlaneForceField floor come out as the scenes state them. Where the reference tree
    is present, its own scene files read into the same bodies (their visual and collision
    components listed as ignored)."""
    import os
    from sofa_amd import scene_xml
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sofa_amd", "scenes")
    expect = {"beam10x10x40": (4000, 18954, 100, 20, 0.01, 2000.0, None),
              "beam16x16x76": (19456, 101250, 256, 10, 0.04, 1000.0, 2.0)}
    ours = {}
    for name, (n, e, nf, iters, dt, E, plane_d) in expect.items():
        specs, info = scene_xml.load_scene(os.path.join(here, name + ".scn"))
        assert len(specs) == 1 and info["ignored"] == []
        s = specs[0]
        assert (len(s["pos"]), len(s["tets"]), len(s["fixed"])) == (n, e, nf)
        assert s["iters"] == iters and s["dt"] == dt and s["E"][0] == E and s["nu"][0] == 0.3
        assert s["tol"] == 1e-6 and s["thr"] == 1e-6 and s["rs"] == 0.1 and s["rm"] == 0.1
        assert s["mass"] == ("diagonal_total", 50.0) and s["g"] == (0.0, -9.0, 0.0)
        assert np.all(s["pos"][s["fixed"], 0] == 0.0)                        # the clamped face x = 0
        assert np.array_equal(s["tets"], meshes_split6(s["pos"], name))
        assert (s.get("plane") or {}).get("d") == plane_d
        ours[name] = s
    ref = "/root/reference/applications/plugins/SofaCUDA/scenes/benchmarks"
    if os.path.isdir(ref):
        for name, s in ours.items():
            for dev in ("cpu", "gpu"):
                specs, info = scene_xml.load_scene(os.path.join(ref, f"TetrahedronFEMForceField_{name}_{dev}.scn"))
                assert len(specs) == 1 and "OglModel" in info["ignored"]
                r = specs[0]
                for k in ("pos", "tets", "fixed"):
                    assert np.array_equal(r[k], s[k]), (name, dev, k)
                for k in ("iters", "tol", "thr", "rs", "rm", "dt", "mass", "g", "plane"):
                    assert r.get(k) == s.get(k), (name, dev, k)
                assert np.array_equal(r["E"], s["E"]) and np.array_equal(r["nu"], s["nu"])


def meshes_split6(pos, name):
    from sofa_amd import meshes
    n = {"beam10x10x40": (40, 10, 10), "beam16x16x76": (76, 16, 16)}[name]
    lo, hi = (0.0, 6.0, -2.0), ((16.0, 10.0, 2.0) if name == "beam10x10x40" else (19.0, 10.0, 2.0))
    p, hexa = meshes.regular_grid(n, lo, hi)
    assert np.array_equal(p, pos)
    return meshes.split_hexa_6(hexa)


def _write_mesh_files(tmp, pos, tets):
    """The same mesh in Gmsh 1, 2.2 and 4.1 and legacy VTK, node numbers offset in the Gmsh files
    (1-based, with a gap) and one triangle element mixed in, as real files have."""
    import os
    num = [i + 1 + (7 if i >= 3 else 0) for i in range(len(pos))]
    f1 = os.path.join(tmp, "m1.msh")
    with open(f1, "w") as fh:
        fh.write(f"$NOD\n{len(pos)}\n")
        for i, p in enumerate(pos):
            fh.write(f"{num[i]} {float(p[0])!r} {float(p[1])!r} {float(p[2])!r}\n")
        fh.write(f"$ENDNOD\n$ELM\n{len(tets) + 1}\n1 2 0 0 3 {num[0]} {num[1]} {num[2]}\n")
        for k, t in enumerate(tets):
            fh.write(f"{k + 2} 4 1 1 4 " + " ".join(str(num[v]) for v in t) + "\n")
        fh.write("$ENDELM\n")
    f2 = os.path.join(tmp, "m2.msh")
    with open(f2, "w") as fh:
        fh.write(f"$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n{len(pos)}\n")
        for i, p in enumerate(pos):
            fh.write(f"{num[i]} {float(p[0])!r} {float(p[1])!r} {float(p[2])!r}\n")
        fh.write(f"$EndNodes\n$Elements\n{len(tets) + 1}\n1 2 2 0 1 {num[0]} {num[1]} {num[2]}\n")
        for k, t in enumerate(tets):
            fh.write(f"{k + 2} 4 2 0 1 " + " ".join(str(num[v]) for v in t) + "\n")
        fh.write("$EndElements\n")
    f4 = os.path.join(tmp, "m4.msh")
    with open(f4, "w") as fh:
        fh.write(f"$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n1 {len(pos)} 1 {max(num)}\n3 1 0 {len(pos)}\n")
        fh.write("".join(f"{v}\n" for v in num))
        fh.write("".join(f"{float(p[0])!r} {float(p[1])!r} {float(p[2])!r}\n" for p in pos))
        fh.write(f"$EndNodes\n$Elements\n2 {len(tets) + 1} 1 {len(tets) + 1}\n2 1 2 1\n1 {num[0]} {num[1]} {num[2]}\n")
        fh.write(f"3 1 4 {len(tets)}\n")
        for k, t in enumerate(tets):
            fh.write(f"{k + 2} " + " ".join(str(num[v]) for v in t) + "\n")
        fh.write("$EndElements\n")
    fv = os.path.join(tmp, "m.vtk")
    with open(fv, "w") as fh:
        fh.write("# vtk DataFile Version 3.0\nmesh\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        fh.write(f"POINTS {len(pos)} double\n" + "".join(f"{float(p[0])!r} {float(p[1])!r} {float(p[2])!r}\n" for p in pos))
        fh.write(f"CELLS {len(tets) + 1} {5 * len(tets) + 4}\n3 0 1 2\n" + "".join("4 " + " ".join(map(str, t)) + "\n" for t in tets))
        fh.write(f"CELL_TYPES {len(tets) + 1}\n5\n" + "10\n" * len(tets))
    return f1, f2, f4, fv


def test_mesh_files(tmp_path):
    """Gmsh 1 / 2.2 / 4.1 and legacy VTK: points in file order (Gmsh numbers mapped), the
    tetrahedra in file order, other element types skipped; a scene file naming the mesh
    (MeshGmshLoader with scale3d and translation) reads into the same body. Where the reference
    tree is present, its liver scene (SofaCUDA examples) with its Gmsh 1 mesh reads too."""
    import os
    from sofa_amd import mesh_io, meshes, scene_xml
    pos, tets, fixed = meshes.beam((3, 3, 4), (1.0, 1.0, 1.5), split=6)
    for f in _write_mesh_files(str(tmp_path), pos, tets):
        p, t = mesh_io.read_mesh(f)
        assert np.array_equal(p, pos) and np.array_equal(t, tets), f
    scn = f"""<Node name="root" gravity="0 -9.81 0" dt="0.01"><Node name="B">
      <EulerImplicitSolver rayleighStiffness="0.1" rayleighMass="0.1"/><CGLinearSolver iterations="30"/>
      <MeshGmshLoader name="loader" filename="m2.msh" scale3d="2 2 2" translation="1 0 -1"/>
      <MechanicalObject src="@loader"/><TetrahedronSetTopologyContainer src="@loader"/>
      <UniformMass totalMass="12"/><FixedProjectiveConstraint indices="0 1 2"/>
      <TetrahedronFEMForceField youngModulus="3000" poissonRatio="0.3" method="large"/></Node></Node>"""
    path = os.path.join(str(tmp_path), "s.scn")
    with open(path, "w") as fh:
        fh.write(scn)
    (s,), info = scene_xml.load_scene(path)
    assert np.array_equal(s["pos"], pos * 2.0 + np.array([1.0, 0.0, -1.0])) and np.array_equal(s["tets"], tets)
    assert s["mass"] == ("uniform", 12.0 / len(pos)) and list(s["fixed"]) == [0, 1, 2] and s["iters"] == 30
    ref = "/root/reference"
    if os.path.isdir(ref):
        p, t = mesh_io.read_gmsh(os.path.join(ref, "share/mesh/liver.msh"))
        assert p.shape == (181, 3) and t.shape == (596, 4) and t.min() == 0 and t.max() == 180
        (s,), info = scene_xml.load_scene(os.path.join(ref, "applications/plugins/SofaCUDA/examples/liver-CUDA.scn"),
                                          search_paths=[os.path.join(ref, "share")])
        assert np.array_equal(s["pos"], p) and np.array_equal(s["tets"], t) and list(s["fixed"]) == [3, 39, 64]
        assert s["mass"] == ("diagonal", 1.0) and s["E"][0] == 3000.0 and s["iters"] == 25
        assert "SphereCollisionModel" in info["ignored"] and "BarycentricMapping" in info["ignored"]


def test_scene_reader_von_mises_attribute():
    """TetrahedronFEMForceField computeVonMisesStress (TetrahedronFEMForceField.inl:57: 0 none,
    1 corotational, 2 Green strain) is kept as the spec's von_mises mode."""
    from sofa_amd import scene_xml
    text = """<Node name="root" dt="0.01" gravity="0 -9.81 0">
      <EulerImplicitSolver/><CGLinearSolver iterations="20" tolerance="1e-9" threshold="1e-9"/>
      <Node name="beam">
        <RegularGridTopology name="grid" n="3 3 5" min="0 0 0" max="1 1 2"/>
        <MechanicalObject/>
        <TetrahedronFEMForceField youngModulus="1e5" poissonRatio="0.3" method="large" computeVonMisesStress="%s"/>
        <UniformMass totalMass="10"/>
        <BoxROI name="roi" box="-0.1 -0.1 -0.1 1.1 1.1 0.1"/>
        <FixedProjectiveConstraint indices="@roi.indices"/>
      </Node></Node>"""
    for mode, want in (("0", None), ("1", 1), ("2", 2)):
        specs, info = scene_xml.load_scene(text % mode)
        assert len(specs) == 1 and specs[0].get("von_mises") == want, (mode, specs[0].get("von_mises"))


_STRICT_SCENE = """<Node name="root" dt="0.01" gravity="0 -9.81 0">
  <EulerImplicitSolver %(solver)s/><CGLinearSolver iterations="20" tolerance="1e-9" threshold="1e-9"/>
  <Node name="beam">
    <RegularGridTopology name="grid" n="3 3 5" min="0 0 0" max="1 1 2"/>
    <MechanicalObject/>
    <TetrahedronFEMForceField youngModulus="1e5 2e5" poissonRatio="0.3" method="large" %(fem)s/>
    <%(mass)s/>
    <BoxROI name="roi" box="-0.1 -0.1 -0.1 1.1 1.1 0.1"/>
    <FixedProjectiveConstraint indices="@roi.indices"/>
    %(extra)s
  </Node>%(second)s</Node>"""


def _strict(fem="", mass='UniformMass totalMass="10"', solver="", extra="", second=""):
    return _STRICT_SCENE % dict(fem=fem, mass=mass, solver=solver, extra=extra, second=second)


@pytest.mark.parametrize("fem,mass,extra", [
    ('plasticMaxThreshold="0.1"', None, ""),
    ('updateStiffnessMatrix="true"', None, ""),
    ('updateStiffness="1"', None, ""),
    ('computeGlobalMatrix="true"', None, ""),
    ('rayleighStiffness="0.2"', None, ""),
    ("", 'UniformMass totalMass="10" rayleighMass="0.1"', ""),
    ("", "UniformMass", ""),
    ("", 'DiagonalMass vertexMass="1 1 1"', ""),
    ("", None, '<PlaneForceField normal="0 1 0" d="-1" rayleighStiffness="0.5"/>'),
    ("", None, '<PlaneForceField normal="0 1 0" d="-1" localRange="0 10"/>'),
])
def test_scene_reader_refuses_unrun_data(fem, mass, extra):
    """Data values this path does not run are refused, not silently dropped: plasticity,
    updateStiffnessMatrix, updateStiffness, the assembled global matrix, a component's own
    Rayleigh coefficient, a UniformMass with no mass Data (invalid in the reference), a
    DiagonalMass of per-vertex masses only, PlaneForceField localRange."""
    from sofa_amd import scene_xml
    text = _strict(fem=fem, mass=mass or 'UniformMass totalMass="10"', extra=extra)
    with pytest.raises(ValueError):
        scene_xml.load_scene(text)
    # the same scene without the offending value reads
    assert len(scene_xml.load_scene(_strict())[0]) == 1


def test_scene_reader_warm_start():
    """CGLinearSolver warmStart is read into the spec (the FP64 engine runs it)."""
    from sofa_amd import scene_xml
    text = _strict().replace('<CGLinearSolver iterations="20"', '<CGLinearSolver warmStart="true" iterations="20"')
    (spec,), _ = scene_xml.load_scene(text)
    assert spec.get("warm_start") is True
    (spec,), _ = scene_xml.load_scene(_strict())
    assert not spec.get("warm_start")


def test_scene_reader_mass_precedence():
    """UniformMass: totalMass over vertexMass (UniformMass.inl:233-266); DiagonalMass: totalMass,
    then massDensity, then the default totalMass 1 (DiagonalMass.inl:880-935)."""
    from sofa_amd import scene_xml
    cases = [('UniformMass vertexMass="2" totalMass="45"', ("uniform", 1.0)),   # 45 kg over 45 nodes
             ('UniformMass vertexMass="2"', ("uniform", 2.0)),
             ('DiagonalMass massDensity="3" totalMass="7"', ("diagonal_total", 7.0)),
             ('DiagonalMass massDensity="3"', ("diagonal", 3.0)),
             ("DiagonalMass", ("diagonal_total", 1.0))]
    for mass, want in cases:
        (spec,), _ = scene_xml.load_scene(_strict(mass=mass))
        assert spec["mass"] == want, (mass, spec["mass"])


def test_scene_reader_local_stiffness_factor():
    """localStiffnessFactor: element i's Young modulus is factor[i M / N] times its own
    (youngModulus entry i, or entry 0 past the list), as computeMaterialStiffness computes it
    (TetrahedronFEMForceField.inl:256-261); checked element by element."""
    from sofa_amd import scene_xml
    (spec,), _ = scene_xml.load_scene(_strict(fem='localStiffnessFactor="1 0.5 4"'))
    n = len(spec["tets"])
    E = np.asarray(spec["E"])
    assert E.shape == (n,)
    base = [1e5, 2e5]
    for i in range(n):
        f = [1.0, 0.5, 4.0][i * 3 // n]
        assert E[i] == f * (base[i] if i < 2 else base[0])
    (plain,), _ = scene_xml.load_scene(_strict())
    assert list(plain["E"]) == [1e5, 2e5]


def test_scene_reader_shared_solver_refused():
    """One EulerImplicitSolver over two bodies is one linear system in SOFA (one CG over both
    mechanical objects); the reader refuses it rather than solving the bodies apart."""
    from sofa_amd import scene_xml
    second = """<Node name="beam2">
    <RegularGridTopology name="grid2" n="3 3 5" min="2 0 0" max="3 1 2"/>
    <MechanicalObject/>
    <TetrahedronFEMForceField youngModulus="1e5" poissonRatio="0.3"/>
    <UniformMass totalMass="10"/>
    <FixedProjectiveConstraint indices="0 1 2"/>
  </Node>"""
    with pytest.raises(ValueError, match="governs bodies"):
        scene_xml.load_scene(_strict(second=second))
    # each body with its own solver reads as two scenes
    text = """<Node name="root" dt="0.01" gravity="0 -9.81 0">
  <Node name="beam"><EulerImplicitSolver/><CGLinearSolver/>
    <RegularGridTopology name="grid" n="3 3 5" min="0 0 0" max="1 1 2"/>
    <MechanicalObject/><TetrahedronFEMForceField youngModulus="1e5" poissonRatio="0.3"/>
    <UniformMass totalMass="10"/><FixedProjectiveConstraint indices="0 1 2"/></Node>
  <Node name="beam2"><EulerImplicitSolver/><CGLinearSolver/>
    <RegularGridTopology name="grid2" n="3 3 5" min="2 0 0" max="3 1 2"/>
    <MechanicalObject/><TetrahedronFEMForceField youngModulus="1e5" poissonRatio="0.3"/>
    <UniformMass totalMass="10"/><FixedProjectiveConstraint indices="0 1 2"/></Node></Node>"""
    specs, info = scene_xml.load_scene(text)
    assert info["bodies"] == ["beam", "beam2"] and len(specs) == 2
