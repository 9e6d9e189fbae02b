"""Synthetic workloads of the benchmark configurations (seeded, reproducible).

Each scene spec is a dict: pos (N,3) float64, tets (E,4) int32, fixed (F,) int32, E and nu
(scalar or per element), mass ("uniform", vertexMass) or ("diagonal", density), g, dt, rs, rm,
iters, tol, thr -- the Data fields of MechanicalObject, TetrahedronFEMForceField, UniformMass /
DiagonalMass, FixedProjectiveConstraint, EulerImplicitSolver and CGLinearSolver.

* cfg0: cantilever beam 5x5x21 nodes over 1x1x4 m (1,600 tets), E 1e5, nu 0.45, UniformMass
  totalMass = 1000 kg/m^3 x 4 m^3, z = 0 face fixed, g = -9.81 y, rs = rm = 0.1, dt 0.01,
  CG 25 / 1e-9 / 1e-9.
* cfg1: the same beam at 33x33x257 nodes over 1x1x8 m (1,310,720 tets), dt 0.005, CG 50.
* cfg2: soft ellipsoid (semi-axes 0.15 x 0.10 x 0.08 m): a grid of pitch h = 2.93 mm jittered by
  U(-h/4, h/4) per axis, points inside the ellipsoid, Delaunay tetrahedra (scipy) without the
  slivers (volume below h^3/120, about 0.1 % of them, mostly on the hull) and without nodes left
  unused; E_e = 10^U(3,5), nu_e = U(0.40, 0.49) per element, MeshMatrixMass massDensity 1050
  (consistent, lumping false), nodes with x < -0.13 fixed, nodes with x > 0.13 driven by
  OscillatorProjectiveConstraint (amplitude 5 mm along x, 1 Hz, phase 0, mean = rest
  position), g = -9.81 y, rs = rm = 0.1, dt 0.01; solver: CG tolerance 1e-8 on |r|/|b| with
  threshold 1e-30; MIXED runs it as FP32 inner CG of 30 iterations inside at most 3 FP64
  refinement sweeps.  About 200k nodes and 1.3M tetrahedra.
* cfg3: 256 independent beams, node counts log-uniform in [500, 50,000] realised as
  nx = ny drawn from {3..9} and nz = max(2, round(N / nx^2)), 1 x 1 x 4 m, E log-uniform in
  [1e4, 1e6], nu uniform in [0.3, 0.49], z = 0 face fixed, CG 25, dt 0.01; the beam's mass is
  1000 kg/m^3 x 4 m^3 spread uniformly (UniformMass totalMass).
"""
from __future__ import annotations

import numpy as np

from . import meshes


def _beam_spec(n, length, E, nu, dt, iters, density=1000.0, tol=1e-9, thr=1e-9):
    pos, tets, fixed = meshes.beam(n, length)
    volume = float(np.prod(length))
    return dict(pos=pos, tets=tets, fixed=fixed, E=E, nu=nu, mass=("uniform", density * volume / len(pos)),
                total_mass=density * volume, g=(0.0, -9.81, 0.0), dt=dt, rs=0.1, rm=0.1, iters=iters, tol=tol,
                thr=thr)


def cfg0():
    return _beam_spec((5, 5, 21), (1.0, 1.0, 4.0), 1e5, 0.45, 0.01, 25)


def cfg1():
    return _beam_spec((33, 33, 257), (1.0, 1.0, 8.0), 1e5, 0.45, 0.005, 50)


def ellipsoid(pitch=0.00293, axes=(0.15, 0.10, 0.08), seed=2):
    """Jittered-grid Delaunay mesh of an ellipsoid: positions (N,3) float64, tets (E,4) int32,
    every tetrahedron positively oriented."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    a = np.asarray(axes, dtype=np.float64)
    g = [np.arange(-ax, ax + pitch / 2, pitch) for ax in a]
    X, Y, Z = np.meshgrid(*g, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    P = P + rng.uniform(-0.25 * pitch, 0.25 * pitch, P.shape)
    P = P[((P / a) ** 2).sum(1) <= 1.0]
    T = Delaunay(P).simplices.astype(np.int64)
    A = P[T]
    vol6 = np.einsum("ij,ij->i", np.cross(A[:, 1] - A[:, 0], A[:, 2] - A[:, 0]), A[:, 3] - A[:, 0])
    keep = np.abs(vol6) >= pitch ** 3 / 20.0          # |V| >= h^3 / 120
    T, vol6 = T[keep], vol6[keep]
    neg = vol6 < 0
    T[neg, 2], T[neg, 3] = T[neg, 3].copy(), T[neg, 2].copy()
    used = np.zeros(len(P), dtype=bool)
    used[T.ravel()] = True
    remap = np.cumsum(used) - 1
    return np.ascontiguousarray(P[used]), np.ascontiguousarray(remap[T].astype(np.int32))


def cfg2(pitch=0.00293, seed=2, frequency=1.0, amplitude=0.005):
    pos, tets = ellipsoid(pitch, seed=seed)
    rng = np.random.default_rng(seed + 1000)
    E = 10 ** rng.uniform(3.0, 5.0, len(tets))
    nu = rng.uniform(0.40, 0.49, len(tets))
    fixed = np.nonzero(pos[:, 0] < -0.13)[0].astype(np.int32)
    drv = np.nonzero(pos[:, 0] > 0.13)[0].astype(np.int32)
    amp = np.zeros((len(drv), 3))
    amp[:, 0] = amplitude
    osc = dict(index=drv, mean=pos[drv].copy(), amplitude=amp, pulsation=np.full(len(drv), 2 * np.pi * frequency),
               phase=np.zeros(len(drv)))
    return dict(pos=pos, tets=tets, fixed=fixed, E=E, nu=nu, mass=("meshmatrix", 1050.0), g=(0.0, -9.81, 0.0),
                dt=0.01, rs=0.1, rm=0.1, iters=90, tol=1e-8, thr=1e-30, oscillators=osc,
                refinement=dict(inner_iterations=30, inner_tolerance=1e-6, max_sweeps=3))


def voxel_ellipsoid(n_target, axes=(0.15, 0.10, 0.08)):
    """Ellipsoid from a regular grid: the hexahedra whose centre lies inside, split 6 ways
    (Hexa2TetraTopologicalMapping, conforming), unused nodes dropped; the pitch is chosen so
    that the node count is close to `n_target`."""
    a = np.asarray(axes, dtype=np.float64)
    h = (4.0 / 3.0 * np.pi * np.prod(a) / n_target) ** (1.0 / 3.0)
    n = [max(2, int(np.ceil(2 * ax / h)) + 1) for ax in a]
    pos, hexa = meshes.regular_grid(n, -a, a)
    centre = pos[hexa].mean(axis=1)
    hexa = hexa[((centre / a) ** 2).sum(1) <= 1.0]
    tets = meshes.split_hexa_6(hexa)
    used = np.zeros(len(pos), dtype=bool)
    used[tets.ravel()] = True
    remap = np.cumsum(used) - 1
    return np.ascontiguousarray(pos[used]), np.ascontiguousarray(remap[tets].astype(np.int32))


def _driven(pos, nodes, amplitude, frequency):
    amp = np.zeros((len(nodes), 3))
    amp[:, 0] = amplitude
    return dict(index=nodes.astype(np.int32), mean=pos[nodes].copy(), amplitude=amp,
                pulsation=np.full(len(nodes), 2 * np.pi * frequency), phase=np.zeros(len(nodes)))


def cfg4_jobs(rate=200.0, duration=60.0, seed=4):
    """The served job stream: Poisson arrivals at `rate` per second over `duration` seconds.
    Each job is a beam (1 x 1 x 4 m, cubic cells, 5-tet split, z = 0 face fixed, z = 4 face
    oscillating) or an ellipsoid (voxel ellipsoid, x < -0.13 fixed, x > 0.13 oscillating), by a
    coin; node count log-uniform in [1e3, 2e5], steps uniform in {50..500}, oscillation
    amplitude U[0, 10 mm] along x at U[0.5, 5] Hz; E 10^U(4,6) (beams) / 10^U(3,5)
    (ellipsoids), nu U[0.3, 0.49]; UniformMass 1000 kg/m^3 (beams) / MeshMatrixMass 1050
    lumped (ellipsoids); CG 25 / 1e-9, dt 0.01, rs = rm = 0.1.  Returns a list of dicts with
    `arrival` (s), `steps` and a lazily built scene spec (`build()`)."""
    rng = np.random.default_rng(seed)
    jobs, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        n_target = 10 ** rng.uniform(3.0, np.log10(2e5))
        steps = int(rng.integers(50, 501))
        amp = rng.uniform(0.0, 0.01)
        freq = rng.uniform(0.5, 5.0)
        nu = rng.uniform(0.3, 0.49)
        if rng.random() < 0.5:
            nxy = max(2, int(round((n_target / 4.0) ** (1.0 / 3.0))))
            dims = (nxy, nxy, max(2, int(round(n_target / (nxy * nxy)))))
            E = 10 ** rng.uniform(4.0, 6.0)
            jobs.append(dict(arrival=t, steps=steps, kind="beam", dims=dims, E=E, nu=nu, amp=amp, freq=freq))
        else:
            E = 10 ** rng.uniform(3.0, 5.0)
            jobs.append(dict(arrival=t, steps=steps, kind="ellipsoid", n_target=n_target, E=E, nu=nu, amp=amp,
                             freq=freq))
    return jobs


def build_job(job):
    """Scene spec of one cfg4 job."""
    common = dict(g=(0.0, -9.81, 0.0), dt=0.01, rs=0.1, rm=0.1, iters=25, tol=1e-9, thr=1e-9)
    if job["kind"] == "beam":
        spec = _beam_spec(job["dims"], (1.0, 1.0, 4.0), job["E"], job["nu"], 0.01, 25)
        pos = spec["pos"]
        top = np.nonzero(pos[:, 2] == pos[:, 2].max())[0]
        spec["oscillators"] = _driven(pos, top, job["amp"], job["freq"])
        return spec
    pos, tets = voxel_ellipsoid(job["n_target"])
    fixed = np.nonzero(pos[:, 0] < -0.13)[0].astype(np.int32)
    drv = np.nonzero(pos[:, 0] > 0.13)[0]
    return dict(pos=pos, tets=tets, fixed=fixed, E=job["E"], nu=job["nu"], mass=("meshmatrix_lumped", 1050.0),
                oscillators=_driven(pos, drv, job["amp"], job["freq"]), **common)


def cfg3(count=256, seed=3):
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(count):
        n_target = 10 ** rng.uniform(np.log10(500.0), np.log10(50000.0))
        nxy = int(rng.integers(3, 10))
        nz = max(2, int(round(n_target / (nxy * nxy))))
        E = 10 ** rng.uniform(4.0, 6.0)
        nu = rng.uniform(0.3, 0.49)
        specs.append(_beam_spec((nxy, nxy, nz), (1.0, 1.0, 4.0), E, nu, 0.01, 25))
    return specs


def add_to_engine(eng, spec):
    """Registers one spec as a scene of a capi.Engine; returns the scene id."""
    from . import capi
    kind, val = spec["mass"]
    mk = {"uniform": capi.MASS_UNIFORM, "diagonal": capi.MASS_DIAGONAL, "meshmatrix": capi.MASS_MESHMATRIX,
          "meshmatrix_lumped": capi.MASS_MESHMATRIX_LUMPED, "diagonal_total": capi.MASS_DIAGONAL_TOTAL}[kind]
    s = eng.add_scene(spec["pos"], spec["tets"], np.atleast_1d(spec["E"]), np.atleast_1d(spec["nu"]),
                      spec["fixed"], mass_kind=mk, mass_value=val, gravity=spec["g"], dt=spec["dt"],
                      rayleigh_stiffness=spec["rs"], rayleigh_mass=spec["rm"], cg_iterations=spec["iters"],
                      cg_tolerance=spec["tol"], cg_threshold=spec["thr"], oscillators=spec.get("oscillators"))
    if "refinement" in spec:
        eng.set_refinement(s, **spec["refinement"])
    if spec.get("fix_all"):
        eng.set_fix_all(s, True)
    if spec.get("project_velocity"):
        eng.set_project_velocity(s, True)
    if "integrator" in spec:
        eng.set_integrator(s, **spec["integrator"])
    if "plane" in spec:
        eng.add_plane(s, **spec["plane"])
    if "linear_movement" in spec:
        eng.add_linear_movement(s, **spec["linear_movement"])
    if spec.get("preconditioner"):
        eng.set_preconditioner(s, spec["preconditioner"])
    if spec.get("warm_start"):
        eng.set_warm_start(s, True)
    return s
