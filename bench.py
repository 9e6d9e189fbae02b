#!/usr/bin/env python3
"""Benchmark of the implicit corotational-FEM step on one MI355X.

Workload (one "step" = one implicit Euler step = RHS + CG + update, i.e. one node::animate):
the cantilever beam of the plain-at-scale configuration: RegularGridTopology 33x33x257 nodes
over 1x1x8 m, 5 tetrahedra per hexahedron (1,310,720 tets, 279,873 nodes), E = 1e5 Pa,
nu = 0.45, UniformMass totalMass = 1000 kg/m^3 * 8 m^3, z = 0 face fixed, gravity -9.81 y,
EulerImplicitSolver rs = rm = 0.1, dt = 0.005, CGLinearSolver 50 iterations / 1e-9 / 1e-9.

Prints ONE JSON line: element-steps/s of the FP32 path (whole job, inputs resident in HBM),
p50/p99 per-step latency, the position error against the FP64 path after 100 steps, the
roofline of the addDForce launch (HIP events on its stream) and the CPU oracle timed on this
host (a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, 8 TB/s


def cfg1():
    from sofa_amd import meshes
    pos, tets, fixed = meshes.beam((33, 33, 257), (1.0, 1.0, 8.0))
    return dict(pos=pos, tets=tets, fixed=fixed, E=1e5, nu=0.45, total_mass=8000.0, g=(0.0, -9.81, 0.0),
                rs=0.1, rm=0.1, dt=0.005, iters=50, tol=1e-9, thr=1e-9)


def make_engine(c, precision):
    from sofa_amd import capi
    eng = capi.Engine(0, precision)
    n = len(c["pos"])
    sid = eng.add_scene(c["pos"], c["tets"], [c["E"]], [c["nu"]], c["fixed"], mass_kind=capi.MASS_UNIFORM,
                        mass_value=c["total_mass"] / float(n), gravity=c["g"], dt=c["dt"],
                        rayleigh_stiffness=c["rs"], rayleigh_mass=c["rm"], cg_iterations=c["iters"],
                        cg_tolerance=c["tol"], cg_threshold=c["thr"])
    eng.commit()
    return eng, sid


def bench_stream():
    """A dedicated stream (default priority): the library's calls and the HIP events that time
    them must share a non-null stream handle (NULL selects the engine's own stream)."""
    import torch
    global _STREAM
    if _STREAM is None:
        _STREAM = torch.cuda.Stream()
    return _STREAM


_STREAM = None


def time_steps(eng, sid, steps, warmup):
    import torch
    stream = bench_stream()
    sp = stream.cuda_stream
    eng.step([sid], warmup, stream=sp)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(steps):
        eng.step([sid], 1, stream=sp)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_ms = ev[0].elapsed_time(ev[steps])
    lat = np.array([ev[i].elapsed_time(ev[i + 1]) for i in range(steps)])
    return total_ms, wall, lat


def apply_roofline(c, precision, reps, warm_steps=3):
    """The dominant kernel: the tiled CG operator apply (k_tile_apply: p = r + beta p staged in
    LDS, element K p, per-node sums in LDS, q of tile-owned nodes, partials of shared nodes),
    launched `reps` times on a state advanced by `warm_steps` steps and timed with HIP events on
    the stream it is launched on.
    Algorithmic bytes per launch: E (16 + 18 p_e) + N 12 p_v
      per element: 16 B of indices (4 x u16 local node ids + 4 x u16 incidence slots),
                   rotation 9, rest coordinates 6, stiffness 3 reals;
      per node: r and p read, p and q written (3 reals each)."""
    import torch
    from sofa_amd import capi
    eng, sid = make_engine(c, precision)
    n, E = len(c["pos"]), len(c["tets"])
    stream = bench_stream()
    sp = stream.cuda_stream
    eng.step([sid], warm_steps, stream=sp)
    eng.bench_apply(sid, 3, stream=sp)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    eng.bench_apply(sid, reps, stream=sp)
    e1.record(stream)
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / reps
    tiling = eng.tiling(sid)
    eng.close()
    pe = pv = 8 if precision == capi.FP64_TILED else 4
    bytes_alg = E * (16 + 18 * pe) + n * 12 * pv
    if precision == capi.FP64_TILED:
        bytes_alg = E * (16 + 21 * pe) + n * 12 * pv     # exact record: 9 rest coordinates
    return us, bytes_alg, tiling


def copy_probe(mb=512, reps=5):
    """A plain device copy of known size (calibrates the HBM byte counters)."""
    import torch
    nbytes = mb * (1 << 20)
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    stream = bench_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        e0.record(stream)
        for _ in range(reps):
            b.copy_(a)
        e1.record(stream)
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / reps
    return {"bytes": 2 * nbytes, "us": us, "GBs": 2 * nbytes / us / 1e3}


def position_error(c, steps):
    """FP32 vs FP64 positions after `steps` steps: eps_rel = max|x - x64| / max|x64| and the
    stricter eps_disp = max|x - x64| / max|x64 - x0|."""
    from sofa_amd import capi
    out = {}
    res = {}
    for name, prec in (("fp64", capi.FP64), ("fp32", capi.FP32)):
        eng, sid = make_engine(c, prec)
        eng.step([sid], steps)
        res[name] = eng.read(sid)[0]
        st = eng.stats(sid)
        out[f"{name}_cg_iterations_last"] = st["iterations"]
        eng.close()
    d = np.linalg.norm(res["fp32"] - res["fp64"], axis=1)
    out["eps_rel"] = float(d.max() / np.linalg.norm(res["fp64"], axis=1).max())
    disp = np.linalg.norm(res["fp64"] - c["pos"], axis=1).max()
    out["eps_disp"] = float(d.max() / disp) if disp > 0 else None
    out["steps"] = steps
    out["max_displacement_m"] = float(disp)
    return out


def cpu_baseline(c, steps):
    """The FP64 oracle (a C++ restatement of the reference path, one thread) on the host."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import OracleScene
    s = OracleScene(c["pos"])
    s.set_gravity(c["g"])
    s.set_fem(c["tets"], [c["E"]], [c["nu"]], "large")
    s.set_uniform_mass(total_mass=c["total_mass"])
    s.set_fixed(c["fixed"])
    s.set_solver(c["rs"], c["rm"], c["iters"], c["tol"], c["thr"])
    t0 = time.perf_counter()
    s.step(c["dt"], steps)
    dt = time.perf_counter() - t0
    return {"value": len(c["tets"]) * steps / dt, "unit": "element-steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} step(s) of the same 1.31M-tet beam, FP64, CG 50, 1 thread "
                      f"({dt:.1f} s); oracle/fem_oracle.cpp built -O2 -ffp-contract=off"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--dforce-reps", type=int, default=100)
    ap.add_argument("--error-steps", type=int, default=100)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(independent jobs shard across GPUs with no communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    from sofa_amd import capi
    torch.cuda.set_device(0)
    c = cfg1()
    n, E = len(c["pos"]), len(c["tets"])

    # main timed leg: FP32 plain path
    eng, sid = make_engine(c, capi.FP32)
    total_ms, wall, lat = time_steps(eng, sid, args.steps, args.warmup)
    st = eng.stats(sid)
    eng.close()
    ms_per_step = total_ms / args.steps
    value = E * args.steps / (total_ms / 1e3)

    # FP64 legs (same workload): the parity mode (every sum in the reference's order) and the
    # tiled FP64 kernels
    k64 = max(5, args.steps // 5)
    eng64, sid64 = make_engine(c, capi.FP64)
    total64, _, lat64 = time_steps(eng64, sid64, k64, 2)
    eng64.close()
    engt, sidt = make_engine(c, capi.FP64_TILED)
    total64t, _, lat64t = time_steps(engt, sidt, k64, 2)
    engt.close()

    us, bytes_alg, tiling = apply_roofline(c, capi.FP32, args.dforce_reps)
    achieved = bytes_alg / (us * 1e-6) / 1e9
    us64, bytes64, _ = apply_roofline(c, capi.FP64_TILED, max(10, args.dforce_reps // 4))
    probe = copy_probe()

    err = None if args.no_error else position_error(c, args.error_steps)
    cpu = None if args.no_cpu_baseline else cpu_baseline(c, args.cpu_steps)

    out = {
        "metric": "FEM element-steps/s (implicit corotational FEM step: TetrahedronFEMForceField large + CG + implicit Euler)",
        "value": value,
        "unit": "element-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "cantilever beam 33x33x257 nodes, 1x1x8 m, 5-tet split: 1,310,720 tets, "
                               "279,873 nodes; CG 50 iterations, dt 0.005, E 1e5, nu 0.45; plain one-instance path",
                   "tets": E, "nodes": n, "cg_iterations": c["iters"], "path": "plain", "precision": "fp32"},
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "wall_s_timed": wall,
        "cg_last": {"iterations": st["iterations"], "end_reason": capi.END_REASONS.get(st["end_reason"])},
        "fp64_parity_mode": {"value": E * k64 / (total64 / 1e3), "ms_per_step": total64 / k64,
                             "p50_ms": float(np.percentile(lat64, 50)), "p99_ms": float(np.percentile(lat64, 99))},
        "fp64_tiled": {"value": E * k64 / (total64t / 1e3), "ms_per_step": total64t / k64,
                       "p50_ms": float(np.percentile(lat64t, 50)), "p99_ms": float(np.percentile(lat64t, 99)),
                       "apply_us": us64, "apply_bytes": bytes64,
                       "apply_frac": bytes64 / (us64 * 1e-6) / 1e9 / HBM_PEAK_GBS},
        "tiling": tiling,
        "error_vs_fp64": err,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_tile_apply<float> (CG operator: addDForce + mass + projection + p update)",
                     "us_per_launch": us,
                     "algorithmic_bytes": bytes_alg},
        "copy_probe": probe,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
