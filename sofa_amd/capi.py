"""ctypes binding of libsofa_amd.so (include/sofa_amd.h).

The library is the product: this module only marshals arguments.  It raises if the library
has not been built -- there is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsofa_amd.so")

OK, EINVAL, EDEVICE, ENOMEM, EDEGENERATE, ESTATE, ENAN = 0, -1, -2, -3, -4, -5, -6
FP64, FP32, MIXED, FP64_TILED, SPLIT = 0, 1, 2, 3, 4
MASS_UNIFORM, MASS_DIAGONAL, MASS_MESHMATRIX, MASS_MESHMATRIX_LUMPED, MASS_DIAGONAL_TOTAL = 0, 1, 2, 3, 4
END_REASONS = {0: "iterations", 1: "tolerance", 2: "threshold", 3: "den==0", 4: "null norm of vector b"}

# every symbol include/sofa_amd.h declares
EXPORTS = [
    "sofa_amd_module_name", "sofa_amd_module_version", "sofa_amd_last_error", "sofa_amd_device_check", "sofa_amd_device_memory",
    "sofa_amd_engine_create", "sofa_amd_engine_destroy", "sofa_amd_scene_create", "sofa_amd_scene_set_fem",
    "sofa_amd_scene_set_mass", "sofa_amd_scene_set_gravity", "sofa_amd_scene_set_fixed",
    "sofa_amd_scene_set_solver", "sofa_amd_engine_commit", "sofa_amd_engine_step", "sofa_amd_engine_sync",
    "sofa_amd_scene_read", "sofa_amd_scene_write", "sofa_amd_scene_stats", "sofa_amd_scene_element",
    "sofa_amd_scene_von_mises", "sofa_amd_scene_stiffness",
    "sofa_amd_fem_add_force", "sofa_amd_fem_add_dforce", "sofa_amd_engine_bench_apply", "sofa_amd_scene_tiling",
    "sofa_amd_engine_set_graphs", "sofa_amd_engine_set_branches", "sofa_amd_engine_set_branch_turns", "sofa_amd_engine_set_dot_segments", "sofa_amd_scene_set_mass_density", "sofa_amd_scene_add_oscillators", "sofa_amd_scene_add_linear_movement", "sofa_amd_scene_set_preconditioner", "sofa_amd_scene_set_warm_start",
    "sofa_amd_scene_set_refinement", "sofa_amd_engine_reserve", "sofa_amd_scene_commit", "sofa_amd_scene_release",
    "sofa_amd_engine_usage", "sofa_amd_scene_set_fix_all", "sofa_amd_scene_set_integrator", "sofa_amd_scene_set_fem_hexa", "sofa_amd_scene_add_plane", "sofa_amd_scene_set_project_velocity", "sofa_amd_engine_set_dot_order",
    "sofa_amd_scene_prepare", "sofa_amd_engine_autotune", "sofa_amd_engine_set_apply_variant",
    "sofa_amd_mass_add_mdx", "sofa_amd_mass_acc_from_f", "sofa_amd_mass_add_force", "sofa_amd_project_response",
    "sofa_amd_project_velocity", "sofa_amd_project_position", "sofa_amd_vdot", "sofa_amd_vdot_noalloc", "sofa_amd_cg_solve",
    "sofa_amd_fem_potential_energy", "sofa_amd_stream_create", "sofa_amd_stream_destroy", "sofa_amd_stream_sync",
    "sofa_amd_device_sync", "sofa_amd_event_create", "sofa_amd_event_destroy", "sofa_amd_event_record",
    "sofa_amd_event_sync", "sofa_amd_event_elapsed", "sofa_amd_malloc", "sofa_amd_free", "sofa_amd_memcpy",
    "sofa_amd_copy_bench",
]


class SofaAmdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sofa_amd error {code}: {msg}")
        self.code = code


class CGStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("end_reason", ctypes.c_int32), ("step_count", ctypes.c_int32),
                ("ir_sweeps", ctypes.c_int32), ("err", ctypes.c_double), ("den", ctypes.c_double),
                ("time", ctypes.c_double), ("dot_fallbacks", ctypes.c_int32), ("reserved", ctypes.c_int32)]


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p


def lib():
    """Load libsofa_amd.so; raises if it was not built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m sofa_amd.build` "
                          "(or __graft_entry__.build()); there is no CPU fallback")
    # The library binds the HIP runtime it was built with (RUNPATH /opt/rocm), and streams, events
    # and device buffers come from it (Stream, Event, DeviceArray below): a process that uses this
    # package maps one HIP / HSA runtime.  torch bundles another (its libamdhip64 needs an
    # unversioned libhsa-runtime64.so, so a runtime already mapped under the versioned name -- by
    # the library, or by rocprofv3's preloaded tool -- is not reused and a second HSA runtime
    # comes up beside it), so nothing here imports torch.
    L = ctypes.CDLL(LIB_PATH)
    c_int, c_double = ctypes.c_int, ctypes.c_double
    sig = {
        "sofa_amd_module_name": (ctypes.c_char_p, []),
        "sofa_amd_module_version": (ctypes.c_char_p, []),
        "sofa_amd_last_error": (ctypes.c_char_p, []),
        "sofa_amd_device_check": (c_int, [c_int]),
        "sofa_amd_device_memory": (c_int, [c_int, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
        "sofa_amd_engine_create": (c_int, [c_int, c_int, ctypes.POINTER(_vp)]),
        "sofa_amd_engine_destroy": (c_int, [_vp]),
        "sofa_amd_scene_create": (c_int, [_vp, c_int, _dp, _ip]),
        "sofa_amd_scene_set_fem": (c_int, [_vp, c_int, c_int, _ip, _dp, c_int, _dp, c_int]),
        "sofa_amd_scene_set_mass": (c_int, [_vp, c_int, c_int, c_double]),
        "sofa_amd_scene_set_gravity": (c_int, [_vp, c_int, c_double, c_double, c_double]),
        "sofa_amd_scene_set_fixed": (c_int, [_vp, c_int, _ip, c_int]),
        "sofa_amd_scene_set_solver": (c_int, [_vp, c_int, c_double, c_double, c_double, c_int, c_double, c_double]),
        "sofa_amd_engine_commit": (c_int, [_vp]),
        "sofa_amd_engine_step": (c_int, [_vp, _ip, c_int, c_int, _vp]),
        "sofa_amd_engine_sync": (c_int, [_vp]),
        "sofa_amd_scene_read": (c_int, [_vp, c_int, _dp, _dp, _dp]),
        "sofa_amd_scene_write": (c_int, [_vp, c_int, _dp, _dp]),
        "sofa_amd_scene_stats": (c_int, [_vp, c_int, ctypes.POINTER(CGStats)]),
        "sofa_amd_scene_element": (c_int, [_vp, c_int, c_int, _dp, _dp, _dp]),
        "sofa_amd_scene_von_mises": (c_int, [_vp, c_int, c_int, _dp, _dp]),
        "sofa_amd_scene_stiffness": (c_int, [_vp, c_int, ctypes.c_double, _ip, _ip, _dp, ctypes.POINTER(ctypes.c_int64)]),
        "sofa_amd_fem_add_force": (c_int, [_vp, c_int, _vp, _vp, _vp]),
        "sofa_amd_fem_add_dforce": (c_int, [_vp, c_int, _vp, _vp, c_double, _vp]),
        "sofa_amd_engine_bench_apply": (c_int, [_vp, c_int, c_int, _vp]),
        "sofa_amd_scene_tiling": (c_int, [_vp, c_int, _ip]),
        "sofa_amd_engine_set_graphs": (c_int, [_vp, c_int]),
        "sofa_amd_engine_set_branches": (c_int, [_vp, c_int]),
        "sofa_amd_engine_set_branch_turns": (c_int, [_vp, c_int]),
        "sofa_amd_engine_set_dot_segments": (c_int, [_vp, c_int, c_int]),
        "sofa_amd_scene_set_mass_density": (c_int, [_vp, c_int, c_int, _dp, c_int]),
        "sofa_amd_scene_add_oscillators": (c_int, [_vp, c_int, c_int, _ip, _dp, _dp, _dp, _dp]),
        "sofa_amd_scene_add_linear_movement": (c_int, [_vp, c_int, c_int, _ip, c_int, _dp, _dp, c_int]),
        "sofa_amd_scene_set_preconditioner": (c_int, [_vp, c_int, c_int]),
        "sofa_amd_scene_set_warm_start": (c_int, [_vp, c_int, c_int]),
        "sofa_amd_scene_set_refinement": (c_int, [_vp, c_int, c_int, c_double, c_int, c_int]),
        "sofa_amd_engine_reserve": (c_int, [_vp, ctypes.c_int64, ctypes.c_int64, c_int]),
        "sofa_amd_scene_commit": (c_int, [_vp, c_int]),
        "sofa_amd_scene_release": (c_int, [_vp, c_int]),
        "sofa_amd_engine_usage": (c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
        "sofa_amd_scene_set_fix_all": (c_int, [_vp, c_int, c_int]),
        "sofa_amd_scene_set_integrator": (c_int, [_vp, c_int, c_int, c_int, c_double]),
        "sofa_amd_scene_set_fem_hexa": (c_int, [_vp, c_int, c_int, _ip, c_int, c_int, _dp, c_int, _dp, c_int]),
        "sofa_amd_scene_add_plane": (c_int, [_vp, c_int, _dp, c_double, c_double, c_double, c_double, c_int]),
        "sofa_amd_scene_set_project_velocity": (c_int, [_vp, c_int, c_int]),
        "sofa_amd_engine_set_dot_order": (c_int, [_vp, c_int]),
        "sofa_amd_scene_prepare": (c_int, [_vp, c_int]),
        "sofa_amd_engine_autotune": (c_int, [_vp, c_int, ctypes.POINTER(c_int), ctypes.POINTER(ctypes.c_float)]),
        "sofa_amd_engine_set_apply_variant": (c_int, [_vp, c_int]),
        "sofa_amd_mass_add_mdx": (c_int, [_vp, c_int, _vp, _vp, c_double, _vp]),
        "sofa_amd_mass_acc_from_f": (c_int, [_vp, c_int, _vp, _vp, _vp]),
        "sofa_amd_mass_add_force": (c_int, [_vp, c_int, _vp, _vp]),
        "sofa_amd_project_response": (c_int, [_vp, c_int, _vp, _vp]),
        "sofa_amd_project_velocity": (c_int, [_vp, c_int, _vp, c_double, _vp]),
        "sofa_amd_project_position": (c_int, [_vp, c_int, _vp, c_double, _vp]),
        "sofa_amd_vdot": (c_int, [_vp, c_int, _vp, _vp, c_int, _vp, _vp]),
        "sofa_amd_vdot_noalloc": (c_int, [_vp, c_int, _vp, _vp, c_int, _vp, _vp]),
        "sofa_amd_cg_solve": (c_int, [_vp, c_int, c_double, c_double, c_double, _vp, _vp,
                                      ctypes.POINTER(CGStats), _vp]),
        "sofa_amd_fem_potential_energy": (c_int, [_vp, c_int, _vp, _dp]),
        "sofa_amd_stream_create": (c_int, [c_int, ctypes.POINTER(_vp)]),
        "sofa_amd_stream_destroy": (c_int, [_vp]),
        "sofa_amd_stream_sync": (c_int, [_vp]),
        "sofa_amd_device_sync": (c_int, [c_int]),
        "sofa_amd_event_create": (c_int, [ctypes.POINTER(_vp)]),
        "sofa_amd_event_destroy": (c_int, [_vp]),
        "sofa_amd_event_record": (c_int, [_vp, _vp]),
        "sofa_amd_event_sync": (c_int, [_vp]),
        "sofa_amd_event_elapsed": (c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_float)]),
        "sofa_amd_malloc": (c_int, [c_int, ctypes.c_int64, ctypes.POINTER(_vp)]),
        "sofa_amd_free": (c_int, [_vp]),
        "sofa_amd_memcpy": (c_int, [_vp, _vp, ctypes.c_int64, c_int]),
        "sofa_amd_copy_bench": (c_int, [c_int, ctypes.c_int64, c_int, _vp, ctypes.POINTER(ctypes.c_float)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != OK:
        msg = lib().sofa_amd_last_error().decode(errors="replace")
        raise SofaAmdError(rc, msg)
    return rc


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def _sh(stream):
    """A stream argument: a Stream, a raw handle, or None (the engine's own stream)."""
    return stream.handle if isinstance(stream, Stream) else stream


def device_available(device=0):
    """True when `device` is a usable gfx950 GPU (the library's own runtime; no exception)."""
    return lib().sofa_amd_device_check(int(device)) == OK


def device_memory(device=0):
    """(free, total) bytes of device memory."""
    fr, tot = ctypes.c_int64(), ctypes.c_int64()
    check(lib().sofa_amd_device_memory(int(device), ctypes.byref(fr), ctypes.byref(tot)))
    return fr.value, tot.value


def device_sync(device=0):
    check(lib().sofa_amd_device_sync(int(device)))


class Stream:
    """A HIP stream of the library's runtime (default priority).  `handle` is what the engine
    calls take as `stream`."""

    def __init__(self, device=0):
        h = _vp()
        check(lib().sofa_amd_stream_create(int(device), ctypes.byref(h)))
        self.handle = h.value

    def synchronize(self):
        check(lib().sofa_amd_stream_sync(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            lib().sofa_amd_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    """A timing event: record(stream), synchronize(), elapsed_time(end) in milliseconds."""

    def __init__(self):
        h = _vp()
        check(lib().sofa_amd_event_create(ctypes.byref(h)))
        self.handle = h.value

    def record(self, stream=None):
        check(lib().sofa_amd_event_record(self.handle, stream.handle if isinstance(stream, Stream) else stream))

    def synchronize(self):
        check(lib().sofa_amd_event_sync(self.handle))

    def elapsed_time(self, end):
        ms = ctypes.c_float()
        check(lib().sofa_amd_event_elapsed(self.handle, end.handle, ctypes.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().sofa_amd_event_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class DeviceArray:
    """A device buffer of the library's runtime holding a numpy array's dtype and shape:
    DeviceArray(np_array) uploads, .numpy() downloads, .ptr is the device address."""

    def __init__(self, array=None, shape=None, dtype=np.float64, device=0):
        if array is not None:
            host = np.ascontiguousarray(array)
        else:
            host = np.zeros(shape, dtype=dtype)
        self.shape, self.dtype = host.shape, host.dtype
        self.nbytes = int(host.nbytes)
        h = _vp()
        check(lib().sofa_amd_malloc(int(device), self.nbytes, ctypes.byref(h)))
        self.ptr = h.value
        check(lib().sofa_amd_memcpy(self.ptr, host.ctypes.data, self.nbytes, 0))

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().sofa_amd_memcpy(out.ctypes.data, self.ptr, self.nbytes, 1))
        return out

    def close(self):
        if getattr(self, "ptr", None):
            lib().sofa_amd_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_bench(nbytes, reps=5, stream=None, device=0):
    """A plain device copy kernel of `nbytes` timed over `reps` launches: microseconds per copy."""
    us = ctypes.c_float()
    check(lib().sofa_amd_copy_bench(int(device), int(nbytes), int(reps),
                                    stream.handle if isinstance(stream, Stream) else stream, ctypes.byref(us)))
    return float(us.value)


class Engine:
    """A set of device-resident scenes advanced together (one launch per phase)."""

    def __init__(self, device=0, precision=FP64):
        L = lib()
        h = _vp()
        check(L.sofa_amd_engine_create(int(device), int(precision), ctypes.byref(h)))
        self._h = h
        self.precision = precision
        self.n_nodes = []
        self.n_tets = []

    def close(self):
        if getattr(self, "_h", None):
            lib().sofa_amd_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- scene construction ----
    def add_scene(self, positions, tets, young, poisson, fixed, mass_kind=MASS_UNIFORM, mass_value=1.0,
                  gravity=(0.0, 0.0, 0.0), dt=0.01, rayleigh_stiffness=0.0, rayleigh_mass=0.0,
                  cg_iterations=25, cg_tolerance=1e-5, cg_threshold=1e-5, oscillators=None, hexa_grid=None):
        """oscillators: None or dict(index, mean (n,3), amplitude (n,3), pulsation (n,), phase (n,));
        mass_value may be an array of per-element densities for the MeshMatrixMass kinds;
        hexa_grid=(nx, ny): `tets` holds hexahedra (H,8), tessellated by the library as
        TetrahedronFEMForceField::init does (6 per cube; nx, ny cubes along x and y)."""
        L = lib()
        pos, pp = _d(np.asarray(positions).reshape(-1, 3))
        sid = ctypes.c_int32()
        check(L.sofa_amd_scene_create(self._h, len(pos), pp, ctypes.byref(sid)))
        s = sid.value
        t, tp = _i(np.asarray(tets).reshape(-1))
        per = 4 if hexa_grid is None else 8
        # recorded at creation, so that the lists stay indexed by scene id even when a later
        # call of this set-up is refused (the engine keeps the half-made scene's id)
        assert s == len(self.n_nodes), (s, len(self.n_nodes))
        self.n_nodes.append(len(pos))
        self.n_tets.append(len(t) // per * (1 if hexa_grid is None else 6))
        y, yp = _d(np.atleast_1d(young))
        nu, nup = _d(np.atleast_1d(poisson))
        if hexa_grid is None:
            check(L.sofa_amd_scene_set_fem(self._h, s, len(t) // 4, tp, yp, len(y), nup, len(nu)))
        else:
            check(L.sofa_amd_scene_set_fem_hexa(self._h, s, len(t) // 8, tp, int(hexa_grid[0]), int(hexa_grid[1]),
                                                yp, len(y), nup, len(nu)))
        mv = np.atleast_1d(np.asarray(mass_value, dtype=np.float64))
        if mv.size > 1:
            d, dp = _d(mv)
            check(L.sofa_amd_scene_set_mass_density(self._h, s, int(mass_kind), dp, len(d)))
        else:
            check(L.sofa_amd_scene_set_mass(self._h, s, int(mass_kind), float(mv[0])))
        if oscillators is not None and len(oscillators["index"]):
            oi, oip = _i(oscillators["index"])
            om, omp = _d(np.asarray(oscillators["mean"]).reshape(-1))
            oa, oap = _d(np.asarray(oscillators["amplitude"]).reshape(-1))
            ow, owp = _d(oscillators["pulsation"])
            op, opp = _d(oscillators["phase"])
            check(L.sofa_amd_scene_add_oscillators(self._h, s, len(oi), oip, omp, oap, owp, opp))
        check(L.sofa_amd_scene_set_gravity(self._h, s, *[float(g) for g in gravity]))
        fx, fxp = _i(np.atleast_1d(fixed) if len(np.atleast_1d(fixed)) else np.zeros(0, np.int32))
        check(L.sofa_amd_scene_set_fixed(self._h, s, fxp, len(fx)))
        check(L.sofa_amd_scene_set_solver(self._h, s, float(dt), float(rayleigh_stiffness), float(rayleigh_mass),
                                          int(cg_iterations), float(cg_tolerance), float(cg_threshold)))
        return s

    def commit(self):
        check(lib().sofa_amd_engine_commit(self._h))

    # ---- live scenes (tiled engines) ----
    def reserve(self, max_nodes, max_tets, max_scenes):
        """Arenas for scenes committed and released one at a time (the serving loop)."""
        check(lib().sofa_amd_engine_reserve(self._h, int(max_nodes), int(max_tets), int(max_scenes)))

    def commit_scene(self, scene):
        """Places one scene; False if the reservation has no room for it now."""
        rc = lib().sofa_amd_scene_commit(self._h, int(scene))
        if rc == ENOMEM:
            return False
        check(rc)
        return True

    def prepare_scene(self, scene):
        """Host tables of a created scene, ahead of commit_scene; callable from a worker thread
        (the call releases the GIL)."""
        check(lib().sofa_amd_scene_prepare(self._h, int(scene)))

    def release_scene(self, scene):
        check(lib().sofa_amd_scene_release(self._h, int(scene)))

    def usage(self):
        out = (ctypes.c_int64 * 22)()
        check(lib().sofa_amd_engine_usage(self._h, out))
        names = ("nodes", "elements", "incidences", "tiles", "tile_local_nodes", "shared_nodes", "shared_slots",
                 "shared_blocks", "node_blocks", "oscillators", "scenes")
        return {k: (int(out[2 * i]), int(out[2 * i + 1])) for i, k in enumerate(names)}

    # ---- stepping ----
    def step(self, scenes=None, n_steps=1, stream=None):
        ids = np.arange(len(self.n_nodes), dtype=np.int32) if scenes is None else np.asarray(scenes, dtype=np.int32)
        a, ap = _i(ids)
        check(lib().sofa_amd_engine_step(self._h, ap, len(a), int(n_steps), _sh(stream)))

    def sync(self):
        check(lib().sofa_amd_engine_sync(self._h))

    def add_plane(self, scene, normal=(0.0, 1.0, 0.0), d=0.0, stiffness=500.0, damping=5.0, max_force=0.0,
                  bilateral=False):
        """PlaneForceField (one per scene)."""
        n, np_ = _d(np.asarray(normal, dtype=np.float64))
        check(lib().sofa_amd_scene_add_plane(self._h, int(scene), np_, float(d), float(stiffness), float(damping),
                                             float(max_force), 1 if bilateral else 0))

    def add_linear_movement(self, scene, indices, key_times, movements, relative=True):
        """LinearMovementProjectiveConstraint: indices, keyTimes, movements (one 3-vector per key),
        relativeMovements."""
        ia, ip = _i(np.asarray(indices).reshape(-1))
        ta, tp = _d(np.asarray(key_times, dtype=np.float64).reshape(-1))
        ma, mp = _d(np.asarray(movements, dtype=np.float64).reshape(-1))
        if len(ma) != 3 * len(ta):
            raise ValueError("one 3-vector movement per key time")
        check(lib().sofa_amd_scene_add_linear_movement(self._h, int(scene), len(ia), ip, len(ta), tp, mp,
                                                        1 if relative else 0))

    def set_preconditioner(self, scene, kind):
        """PCGLinearSolver in place of CGLinearSolver: 0 none, 1 Jacobi, 2 block Jacobi."""
        check(lib().sofa_amd_scene_set_preconditioner(self._h, int(scene), int(kind)))

    def set_warm_start(self, scene, on=True):
        """CGLinearSolver warmStart: each solve starts from the previous step's solution (FP64)."""
        check(lib().sofa_amd_scene_set_warm_start(self._h, int(scene), 1 if on else 0))

    def set_project_velocity(self, scene, on=True):
        """FixedProjectiveConstraint activate_projectVelocity."""
        check(lib().sofa_amd_scene_set_project_velocity(self._h, int(scene), 1 if on else 0))

    def set_integrator(self, scene, first_order=False, trapezoidal=False, vdamping=0.0):
        """EulerImplicitSolver firstOrder, trapezoidalScheme, vdamping."""
        check(lib().sofa_amd_scene_set_integrator(self._h, int(scene), 1 if first_order else 0,
                                                  1 if trapezoidal else 0, float(vdamping)))

    def set_fix_all(self, scene, fix_all=True):
        """FixedProjectiveConstraint fixAll."""
        check(lib().sofa_amd_scene_set_fix_all(self._h, int(scene), 1 if fix_all else 0))

    def set_refinement(self, scene, inner_iterations=30, inner_tolerance=1e-6, max_sweeps=3, keep_direction=False):
        """MIXED engines: FP32 inner CG limits and FP64 refinement sweeps per step; with
        keep_direction the sweeps continue one CG (residual replacement) instead of restarting."""
        check(lib().sofa_amd_scene_set_refinement(self._h, int(scene), int(inner_iterations),
                                                  float(inner_tolerance), int(max_sweeps),
                                                  1 if keep_direction else 0))

    def set_dot_order(self, sequential=True):
        """FP64 engines: dot products in the reference's sequential node order (bit-exact
        trajectories) instead of the parallel blocked order."""
        check(lib().sofa_amd_engine_set_dot_order(self._h, 1 if sequential else 0))

    def autotune(self, reps=5):
        """Tiled engines, before the first step: time the operator variants on the committed
        scenes and keep the fastest.  Returns (chosen variant, microseconds per launch of each)."""
        ch = ctypes.c_int()
        us = (ctypes.c_float * 4)()
        check(lib().sofa_amd_engine_autotune(self._h, int(reps), ctypes.byref(ch), us))
        return ch.value, [float(u) for u in us]

    def set_apply_variant(self, variant):
        check(lib().sofa_amd_engine_set_apply_variant(self._h, int(variant)))

    def set_branches(self, branches):
        """Concurrent branches of a merged step (0: the engine's default)."""
        check(lib().sofa_amd_engine_set_branches(self._h, int(branches)))

    def set_dot_segments(self, segments, window=32):
        """Reference-order dot products in `segments` segments run at once (1: the single chain),
        from 2 `window` candidate starts per boundary."""
        check(lib().sofa_amd_engine_set_dot_segments(self._h, int(segments), int(window)))

    def set_branch_turns(self, enable):
        """Reference-order steps in branches: operator phases take turns (True) or the branches
        run freely (False, default)."""
        check(lib().sofa_amd_engine_set_branch_turns(self._h, 1 if enable else 0))

    def set_graphs(self, enable=True):
        check(lib().sofa_amd_engine_set_graphs(self._h, 1 if enable else 0))

    # ---- state ----
    def read(self, scene):
        n = self.n_nodes[scene]
        x = np.zeros((n, 3)); v = np.zeros((n, 3)); f = np.zeros((n, 3))
        check(lib().sofa_amd_scene_read(self._h, int(scene), x.ctypes.data_as(_dp), v.ctypes.data_as(_dp),
                                        f.ctypes.data_as(_dp)))
        return x, v, f

    def write(self, scene, x=None, v=None):
        xa = None if x is None else _d(np.asarray(x).reshape(-1))
        va = None if v is None else _d(np.asarray(v).reshape(-1))
        check(lib().sofa_amd_scene_write(self._h, int(scene), None if xa is None else xa[1],
                                         None if va is None else va[1]))

    def stats(self, scene):
        st = CGStats()
        check(lib().sofa_amd_scene_stats(self._h, int(scene), ctypes.byref(st)))
        return {"iterations": st.iterations, "end_reason": st.end_reason, "step_count": st.step_count,
                "ir_sweeps": st.ir_sweeps, "err": st.err, "den": st.den, "time": st.time,
                "dot_fallbacks": st.dot_fallbacks}

    def element(self, scene, e):
        rot = np.zeros(9); k = np.zeros(3); rest = np.zeros(12)
        check(lib().sofa_amd_scene_element(self._h, int(scene), int(e), rot.ctypes.data_as(_dp),
                                           k.ctypes.data_as(_dp), rest.ctypes.data_as(_dp)))
        return rot.reshape(3, 3), k, rest.reshape(4, 3)

    def von_mises(self, scene, mode=1):
        """TetrahedronFEMForceField computeVonMisesStress (1: corotational strain, 2: Green strain)
        from the scene's current positions: (per element, per node)."""
        pe = np.zeros(self.n_tets[scene]); pn = np.zeros(self.n_nodes[scene])
        check(lib().sofa_amd_scene_von_mises(self._h, int(scene), int(mode), pe.ctypes.data_as(_dp),
                                             pn.ctypes.data_as(_dp)))
        return pe, pn

    def stiffness(self, scene, k_factor=1.0):
        """addKToMatrix / buildStiffnessMatrix: the scene's assembled stiffness (-k sum R J K J^T R^T
        with the last step's or addForce's rotations) as (row_ptr, cols, blocks (nnzb, 3, 3))."""
        nnzb = ctypes.c_int64()
        check(lib().sofa_amd_scene_stiffness(self._h, int(scene), float(k_factor), None, None, None, ctypes.byref(nnzb)))
        rp = np.zeros(self.n_nodes[scene] + 1, dtype=np.int32)
        cols = np.zeros(nnzb.value, dtype=np.int32)
        vals = np.zeros((nnzb.value, 3, 3))
        check(lib().sofa_amd_scene_stiffness(self._h, int(scene), float(k_factor), rp.ctypes.data_as(_ip),
                                             cols.ctypes.data_as(_ip), vals.ctypes.data_as(_dp), ctypes.byref(nnzb)))
        return rp, cols, vals

    # ---- component-level calls on device pointers ----
    def fem_add_force(self, scene, x_dev_ptr, f_dev_ptr, stream=None):
        check(lib().sofa_amd_fem_add_force(self._h, int(scene), x_dev_ptr, f_dev_ptr, _sh(stream)))

    def fem_add_dforce(self, scene, dx_dev_ptr, df_dev_ptr, k_factor, stream=None):
        check(lib().sofa_amd_fem_add_dforce(self._h, int(scene), dx_dev_ptr, df_dev_ptr, float(k_factor), _sh(stream)))

    # ---- component-level calls of the masses, constraints, state and CG on device pointers ----
    def mass_add_mdx(self, scene, dx_dev_ptr, f_dev_ptr, factor=1.0, stream=None):
        check(lib().sofa_amd_mass_add_mdx(self._h, int(scene), dx_dev_ptr, f_dev_ptr, float(factor), _sh(stream)))

    def mass_acc_from_f(self, scene, f_dev_ptr, a_dev_ptr, stream=None):
        check(lib().sofa_amd_mass_acc_from_f(self._h, int(scene), f_dev_ptr, a_dev_ptr, _sh(stream)))

    def mass_add_force(self, scene, f_dev_ptr, stream=None):
        check(lib().sofa_amd_mass_add_force(self._h, int(scene), f_dev_ptr, _sh(stream)))

    def project_response(self, scene, dx_dev_ptr, stream=None):
        check(lib().sofa_amd_project_response(self._h, int(scene), dx_dev_ptr, _sh(stream)))

    def project_velocity(self, scene, v_dev_ptr, time, stream=None):
        check(lib().sofa_amd_project_velocity(self._h, int(scene), v_dev_ptr, float(time), _sh(stream)))

    def project_position(self, scene, x_dev_ptr, time, stream=None):
        check(lib().sofa_amd_project_position(self._h, int(scene), x_dev_ptr, float(time), _sh(stream)))

    def vdot(self, scene, a_dev_ptr, b_dev_ptr, order, result_dev_ptr, stream=None):
        """vDot into one double at result_dev_ptr (order 1: the reference's sequential sum)."""
        check(lib().sofa_amd_vdot(self._h, int(scene), a_dev_ptr, b_dev_ptr, int(order), result_dev_ptr, _sh(stream)))

    def vdot_noalloc(self, scene, a_dev_ptr, b_dev_ptr, order, result_dev_ptr, stream=None):
        """vDot with the block partials in the library's static device array (no allocation)."""
        check(lib().sofa_amd_vdot_noalloc(self._h, int(scene), a_dev_ptr, b_dev_ptr, int(order), result_dev_ptr,
                                          _sh(stream)))

    def cg_solve(self, scene, m_factor, b_factor, k_factor, b_dev_ptr, x_dev_ptr, stream=None):
        """CGLinearSolver::solve of (m M + b B + k K) x = b; returns the statistics."""
        st = CGStats()
        check(lib().sofa_amd_cg_solve(self._h, int(scene), float(m_factor), float(b_factor), float(k_factor),
                                      b_dev_ptr, x_dev_ptr, ctypes.byref(st), _sh(stream)))
        return {"iterations": st.iterations, "end_reason": st.end_reason, "step_count": st.step_count,
                "err": st.err, "den": st.den}

    def fem_potential_energy(self, scene, x_dev_ptr):
        e = ctypes.c_double()
        check(lib().sofa_amd_fem_potential_energy(self._h, int(scene), x_dev_ptr, ctypes.byref(e)))
        return e.value

    # ---- measurement ----
    def bench_apply(self, scene, reps, stream=None):
        check(lib().sofa_amd_engine_bench_apply(self._h, int(scene), int(reps), _sh(stream)))

    def tiling(self, scene):
        out = np.zeros(7, dtype=np.int32)
        check(lib().sofa_amd_scene_tiling(self._h, int(scene), out.ctypes.data_as(_ip)))
        keys = ("tiles", "tile_local_nodes", "finalize_nodes", "shared_slots", "incidences", "elements", "nodes")
        return dict(zip(keys, (int(v) for v in out)))
