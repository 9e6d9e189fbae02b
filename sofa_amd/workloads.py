"""Synthetic workloads of the benchmark configurations (seeded, reproducible).

Each scene spec is a dict: pos (N,3) float64, tets (E,4) int32, fixed (F,) int32, E and nu
(scalar or per element), mass ("uniform", vertexMass) or ("diagonal", density), g, dt, rs, rm,
iters, tol, thr -- the Data fields of MechanicalObject, TetrahedronFEMForceField, UniformMass /
DiagonalMass, FixedProjectiveConstraint, EulerImplicitSolver and CGLinearSolver.

* cfg0: cantilever beam 5x5x21 nodes over 1x1x4 m (1,600 tets), E 1e5, nu 0.45, UniformMass
  totalMass = 1000 kg/m^3 x 4 m^3, z = 0 face fixed, g = -9.81 y, rs = rm = 0.1, dt 0.01,
  CG 25 / 1e-9 / 1e-9.
* cfg1: the same beam at 33x33x257 nodes over 1x1x8 m (1,310,720 tets), dt 0.005, CG 50.
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
    mk = capi.MASS_UNIFORM if kind == "uniform" else capi.MASS_DIAGONAL
    return eng.add_scene(spec["pos"], spec["tets"], np.atleast_1d(spec["E"]), np.atleast_1d(spec["nu"]),
                         spec["fixed"], mass_kind=mk, mass_value=val, gravity=spec["g"], dt=spec["dt"],
                         rayleigh_stiffness=spec["rs"], rayleigh_mass=spec["rm"], cg_iterations=spec["iters"],
                         cg_tolerance=spec["tol"], cg_threshold=spec["thr"])
