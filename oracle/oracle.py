"""ctypes loader for the FP64 CPU oracle (fem_oracle.cpp) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, by __graft_entry__.smoke() (as the checker) and by bench.py's
cpu_baseline leg.  The product package (sofa_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def use_native_build():
    """Load the -O3 -march=native build (made here, for this host) instead of the parity build:
    the CPU-baseline timing leg calls this before anything else touches the oracle."""
    global _LIB_PATH
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle_native.so"])
    _LIB_PATH = os.path.join(_HERE, "liboracle_native.so")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int, _dp]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_gravity.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_set_fem.restype = ctypes.c_int
        L.orc_set_fem.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_set_uniform_mass.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_set_uniform_total_mass.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_set_diagonal_mass.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_set_diagonal_total_mass.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_add_plane.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_add_axis_spring.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_set_fixed.argtypes = [ctypes.c_void_p, _ip, ctypes.c_int, ctypes.c_int]
        L.orc_add_oscillator.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_double]
        L.orc_set_solver.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_set_dot_mode.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_integrator.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.orc_set_project_velocity.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_preconditioner.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_warm_start.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_precond_blocks.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_add_linear_movement.argtypes = [ctypes.c_void_p, _ip, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
        L.orc_get_vec.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_set_vec.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_get_stats.argtypes = [ctypes.c_void_p, _ip, _dp]
        L.orc_fem_add_force.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_fem_add_dforce.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double]
        L.orc_fem_get_element.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_stiffness_dense.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp]
        L.orc_von_mises.restype = ctypes.c_int
        L.orc_von_mises.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.orc_fem_rest_volume.restype = ctypes.c_double
        L.orc_fem_rest_volume.argtypes = [ctypes.c_void_p]
        L.orc_get_masses.argtypes = [ctypes.c_void_p, _dp]
        L.orc_set_meshmatrix_mass.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_get_edge_count.restype = ctypes.c_int
        L.orc_get_edge_count.argtypes = [ctypes.c_void_p]
        L.orc_get_edges.argtypes = [ctypes.c_void_p, _ip, _dp]
        L.orc_mass_add_mdx.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double]
        L.orc_mass_acc_from_f.restype = ctypes.c_int
        L.orc_mass_acc_from_f.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_mass_add_force.argtypes = [ctypes.c_void_p, _dp]
        L.orc_project.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_double]
        L.orc_vdot.restype = ctypes.c_double
        L.orc_vdot.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.orc_cg_solve.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


class OracleScene:
    """One SOFA scene: MechanicalObject + mass + TetrahedronFEMForceField + constraints,
    stepped by EulerImplicitSolver + CGLinearSolver in FP64 on the CPU."""

    def __init__(self, positions):
        L = lib()
        self.n = int(len(positions))
        pos, pp = _d(positions)
        self._h = ctypes.c_void_p(L.orc_create(self.n, pp))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_destroy(self._h)
            self._h = None

    def set_gravity(self, g):
        lib().orc_set_gravity(self._h, float(g[0]), float(g[1]), float(g[2]))

    def set_fem(self, tets, young, poisson, method="large"):
        t, tp = _i(np.asarray(tets).reshape(-1))
        y, yp = _d(np.atleast_1d(young))
        p, pp = _d(np.atleast_1d(poisson))
        m = {"large": 0, "small": 1}[method]
        rc = lib().orc_set_fem(self._h, len(t) // 4, tp, yp, len(y), pp, len(p), m)
        if rc != 0:
            raise ValueError("invalid tetrahedron index")
        self.n_tets = len(t) // 4

    def set_uniform_mass(self, vertex_mass=None, total_mass=None):
        if total_mass is not None:
            lib().orc_set_uniform_total_mass(self._h, float(total_mass))
        else:
            lib().orc_set_uniform_mass(self._h, float(vertex_mass))

    def set_diagonal_mass(self, density):
        lib().orc_set_diagonal_mass(self._h, float(density))

    def add_plane(self, normal=(0.0, 1.0, 0.0), d=0.0, stiffness=500.0, damping=5.0, max_force=0.0, bilateral=False):
        """PlaneForceField (after the FEM in scene order)."""
        n, np_ = _d(np.asarray(normal, dtype=np.float64))
        lib().orc_add_plane(self._h, np_, float(d), float(stiffness), float(damping), float(max_force),
                            1 if bilateral else 0)

    def set_diagonal_total_mass(self, total_mass):
        """DiagonalMass given totalMass (masses for density 1, scaled to the total)."""
        lib().orc_set_diagonal_total_mass(self._h, float(total_mass))

    def set_meshmatrix_mass(self, density, lumping=False):
        d, dp = _d(np.atleast_1d(density))
        lib().orc_set_meshmatrix_mass(self._h, dp, len(d), 1 if lumping else 0)

    def edges(self):
        ne = lib().orc_get_edge_count(self._h)
        e = np.zeros((ne, 2), dtype=np.int32)
        m = np.zeros(ne, dtype=np.float64)
        lib().orc_get_edges(self._h, e.ctypes.data_as(_ip), m.ctypes.data_as(_dp))
        return e, m

    def add_axis_spring(self, node, axis, stiffness, anchor):
        lib().orc_add_axis_spring(self._h, int(node), int(axis), float(stiffness), float(anchor))

    def set_fixed(self, idx, fix_all=False):
        i, ip = _i(idx)
        lib().orc_set_fixed(self._h, ip, len(i), 1 if fix_all else 0)

    def add_oscillator(self, index, mean, amplitude, pulsation, phase=0.0):
        m, mp = _d(mean)
        a, ap = _d(amplitude)
        lib().orc_add_oscillator(self._h, int(index), mp, ap, float(pulsation), float(phase))

    def set_solver(self, rayleigh_stiffness=0.0, rayleigh_mass=0.0, iterations=25, tolerance=1e-5, threshold=1e-5):
        lib().orc_set_solver(self._h, float(rayleigh_stiffness), float(rayleigh_mass), int(iterations),
                             float(tolerance), float(threshold))

    def set_integrator(self, first_order=False, trapezoidal=False, vdamping=0.0):
        """EulerImplicitSolver firstOrder, trapezoidalScheme and vdamping."""
        lib().orc_set_integrator(self._h, 1 if first_order else 0, 1 if trapezoidal else 0, float(vdamping))

    def set_project_velocity(self, on=True):
        """FixedProjectiveConstraint activate_projectVelocity."""
        lib().orc_set_project_velocity(self._h, 1 if on else 0)

    def set_preconditioner(self, kind):
        """PCGLinearSolver in place of CGLinearSolver: 0 none, 1 Jacobi, 2 block Jacobi."""
        lib().orc_set_preconditioner(self._h, int(kind))

    def set_warm_start(self, on=True):
        """CGLinearSolver warmStart: each solve starts from the previous step's solution."""
        lib().orc_set_warm_start(self._h, 1 if on else 0)

    def precond_blocks(self, m_factor, k_factor):
        """The preconditioner's inverted 3x3 blocks (n, 3, 3) with the current rotations."""
        out = np.zeros((self.n, 3, 3))
        lib().orc_precond_blocks(self._h, float(m_factor), float(k_factor), out.ctypes.data_as(_dp))
        return out

    def add_linear_movement(self, indices, key_times, movements, relative=True):
        """LinearMovementProjectiveConstraint: indices, keyTimes, movements (one 3-vector per key),
        relativeMovements."""
        ia, ip = _i(indices)
        ta, tp = _d(key_times)
        ma, mp = _d(np.asarray(movements, dtype=np.float64).reshape(-1, 3))
        assert len(ma) == len(ta)
        lib().orc_add_linear_movement(self._h, ip, len(ia), tp, mp, len(ta), 1 if relative else 0)

    def set_dot_mode(self, mode):
        lib().orc_set_dot_mode(self._h, int(mode))

    def set_threads(self, n):
        """Threads of the FEM element loops (results identical for any count)."""
        lib().orc_set_threads(self._h, int(n))

    def step(self, dt, nsteps=1):
        lib().orc_step(self._h, float(dt), int(nsteps))

    def get(self, which="x"):
        out = np.zeros((self.n, 3), dtype=np.float64)
        lib().orc_get_vec(self._h, {"x": 0, "v": 1, "f": 2, "x0": 3}[which], out.ctypes.data_as(_dp))
        return out

    def set(self, which, arr):
        a, ap = _d(arr)
        lib().orc_set_vec(self._h, {"x": 0, "v": 1}[which], ap)

    def stats(self):
        ist = np.zeros(3, dtype=np.int32)
        dst = np.zeros(3, dtype=np.float64)
        lib().orc_get_stats(self._h, ist.ctypes.data_as(_ip), dst.ctypes.data_as(_dp))
        return {"iterations": int(ist[0]), "endcond": int(ist[1]), "step_count": int(ist[2]),
                "err": float(dst[0]), "den": float(dst[1]), "time": float(dst[2])}

    def fem_add_force(self, x, f):
        x, xp = _d(x)
        f = np.array(f, dtype=np.float64, copy=True)
        lib().orc_fem_add_force(self._h, xp, f.ctypes.data_as(_dp))
        return f

    def fem_add_dforce(self, dx, df, kfactor):
        dx, dxp = _d(dx)
        df = np.array(df, dtype=np.float64, copy=True)
        lib().orc_fem_add_dforce(self._h, dxp, df.ctypes.data_as(_dp), float(kfactor))
        return df

    def fem_element(self, e):
        R0 = np.zeros(9); R = np.zeros(9); K = np.zeros(36); J = np.zeros(72); rest = np.zeros(12)
        lib().orc_fem_get_element(self._h, int(e), R0.ctypes.data_as(_dp), R.ctypes.data_as(_dp),
                                  K.ctypes.data_as(_dp), J.ctypes.data_as(_dp), rest.ctypes.data_as(_dp))
        return {"R0": R0.reshape(3, 3), "R": R.reshape(3, 3), "K": K.reshape(6, 6), "J": J.reshape(12, 6),
                "rest": rest.reshape(4, 3)}

    def stiffness_dense(self, k=1.0):
        """addKToMatrix (TetrahedronFEMForceField.inl:1921-1957) with the current rotations: the
        dense 3N x 3N matrix -k sum_e R J K J^T R^T."""
        m = np.zeros((3 * self.n, 3 * self.n))
        lib().orc_stiffness_dense(self._h, float(k), m.ctypes.data_as(_dp))
        return m

    def von_mises(self, mode=1):
        """computeVonMisesStress (TetrahedronFEMForceField.inl:2197-2347) at the current positions:
        (per element, per node)."""
        pe = np.zeros(self.n_tets); pn = np.zeros(self.n)
        if lib().orc_von_mises(self._h, int(mode), pe.ctypes.data_as(_dp), pn.ctypes.data_as(_dp)) != 0:
            raise ValueError("von Mises stress not available (mode, method or a singular rest element)")
        return pe, pn

    # ---- component-level calls on (n, 3) arrays (the inputs are not modified) ----
    def mass_add_mdx(self, dx, res, factor=1.0):
        """Mass::addMDx: res + factor M dx."""
        dx, dxp = _d(dx)
        out = np.array(res, dtype=np.float64, copy=True)
        lib().orc_mass_add_mdx(self._h, dxp, out.ctypes.data_as(_dp), float(factor))
        return out

    def mass_acc_from_f(self, f):
        """Mass::accFromF: M^-1 f; ValueError where the reference refuses (consistent mass)."""
        f, fp = _d(f)
        a = np.zeros((self.n, 3))
        if lib().orc_mass_acc_from_f(self._h, fp, a.ctypes.data_as(_dp)) != 0:
            raise ValueError("accFromF needs a lumped mass")
        return a

    def mass_add_force(self, f):
        """Mass::addForce: f + the weight."""
        out = np.array(f, dtype=np.float64, copy=True)
        lib().orc_mass_add_force(self._h, out.ctypes.data_as(_dp))
        return out

    def project(self, what, vec, t=0.0):
        """0 projectResponse, 1 projectVelocity at time t, 2 projectPosition at time t."""
        out = np.array(vec, dtype=np.float64, copy=True)
        lib().orc_project(self._h, int(what), out.ctypes.data_as(_dp), float(t))
        return out

    def vdot(self, a, b, mode=0):
        """vDot: mode 0 the reference's sequential order, 1 the GPU's blocked order."""
        a, ap = _d(a)
        b, bp = _d(b)
        return lib().orc_vdot(self._h, ap, bp, int(mode))

    def cg_solve(self, b, m_factor, b_factor, k_factor):
        """CGLinearSolver::solve of (m M + b B + k K) x = b with the scene's CG Data and the
        rotations of the last addForce; returns x (stats() gives the solve's statistics)."""
        b, bp = _d(b)
        x = np.zeros((self.n, 3))
        lib().orc_cg_solve(self._h, bp, x.ctypes.data_as(_dp), float(m_factor), float(b_factor), float(k_factor))
        return x

    def rest_volume(self):
        return lib().orc_fem_rest_volume(self._h)

    def masses(self):
        out = np.zeros(self.n, dtype=np.float64)
        lib().orc_get_masses(self._h, out.ctypes.data_as(_dp))
        return out
