"""Serving loop: a stream of independent FEM jobs on one GPU with continuous batching.

Jobs arrive over time (wall clock).  A job is admitted when its device footprint fits the HBM
budget next to the running jobs (first come, first served: a job that does not fit holds the
queue), is placed in a live engine (`Engine.reserve` / `commit_scene`), and from then on is
advanced one implicit Euler step per tick together with every other running job, in ONE merged
launch sequence per tick; a job leaves (`release_scene`) after its last step and its room is
reused.  The alternative "plain" mode advances the same running jobs one at a time, each with
its own launches (the one-instance-at-a-time path the batched mode is compared with).

Reported: element-steps/s over the run; per-step latency = device time of the merged step (HIP
events around it on its stream) attributed to every job it advances, p50 / p99 over all
(job, step) pairs; admission wait = wall time from arrival to the start of the job's first
step, p50 / p99; in plain mode also each job step's own device time, summarised per job size
(element-steps/s of one job alone on the GPU, by tetrahedron count).  The scheduler is host code
around the C ABI; all per-step work is on the GPU.  With graphs on, every launch set's step is
captured once and replayed (one graph per set of running jobs in batched mode, one per job in
plain mode).  Without an explicit budget the admission budget is 80 % of the device's free memory
when the server starts.
"""
from __future__ import annotations

import collections
import concurrent.futures as cf
import time

import numpy as np

from . import capi, workloads

# device bytes per node / per tetrahedron of a tiled engine (the arenas of
# sofa_amd_engine_reserve): node vectors x v f b r p q x_cg (3 reals each), mass (8), mask (1),
# incidence pointer (4), tile-local node tables (~1.3 x 12 B), shared tables; per element: index
# (16 + 8), rotation 9, rest 6 or 9, stiffness 3, edge mass 1 reals, element force 12 reals,
# incidences 4 x (4 + 2); in the reference summation order also two per-incidence slot arrays
# (4 x 2 x 3 reals) and two per-incidence rank tables (4 x 2 bytes)
def footprint_bytes(precision, nodes, tets, reference_order=False, scenes=1):
    pv = 4 if precision == capi.FP32 else 8
    pe = 4 if precision == capi.FP32 else 8
    nr = 6 if precision == capi.FP32 else 9
    per_node = 8 * 3 * pv + 8 + 1 + 4 + 16 + 12 + 6 * pv
    per_tet = 24 + (9 + nr + 3 + 1) * pe + 12 * pv + 24
    if precision == capi.MIXED:
        # float r, p, q, d, x, x0 and shared partials; float K entries and edge masses
        per_node += 6 * 3 * 4 + 3 * 4
        per_tet += (3 + 1) * 4
    per_scene = 0
    if reference_order:
        per_tet += 4 * 2 * 3 * pv + 4 * 2
        per_node += 8
        # the segmented dot products' prediction chunks (plain and modelled) and candidate sums
        per_scene += 8 * (2 * 64 + 1800)
    return per_node * nodes + per_tet * tets + per_scene * scenes


class Server:
    """One GPU, one live engine, a FIFO admission queue bounded by an HBM budget."""

    def __init__(self, precision=capi.FP32, budget_bytes=None, max_scenes=2048, batched=True,
                 build_workers=4, engine=None, graphs=False, reference_order=False):
        """reference_order (FP64_TILED): every sum in the reference's order, the served jobs'
        trajectories those of the reference bit for bit."""
        self.precision = precision
        self.reference_order = reference_order
        if budget_bytes is None:
            free, _ = capi.device_memory(0)
            budget_bytes = int(0.8 * free)
        self.budget = int(budget_bytes)
        self.batched = batched
        if engine is None:
            # arena capacity: the budget expressed in nodes and tetrahedra (a tetrahedral mesh
            # has about 5-6 tetrahedra per node; the node arena is sized for the sparser side)
            per_tet = footprint_bytes(precision, 0, 1, reference_order, scenes=0)
            # the engine's incidence space (4 per tetrahedron) is addressed with 32-bit indices
            max_tets = min(self.budget // per_tet, 500_000_000)
            max_nodes = max_tets // 3
            engine = capi.Engine(0, precision)
            if reference_order:
                engine.set_dot_order(True)         # before the arenas: the slot space depends on it
            engine.reserve(max_nodes, max_tets, max_scenes)
            engine.set_graphs(graphs)
        self.eng = engine
        self.pool = cf.ThreadPoolExecutor(build_workers)

    def close(self):
        # queued preparations are dropped, running ones finish before the engine they read goes
        self.pool.shutdown(wait=True, cancel_futures=True)
        self.eng.close()

    def run(self, jobs, wall_limit=30.0, stream=None, build=workloads.build_job, log=None, on_finish=None,
            event=None):
        """Serve `jobs` (dicts with `arrival` seconds and `steps`, built by `build`) until every
        job is done or `wall_limit` seconds have passed.  `on_finish(engine, scene, job_index)` is
        called for each finished job before its scene is released.  `event()` makes the timing
        events (HIP events by default).  Returns the measurements."""
        if event is None:
            own = stream is None
            if own:
                stream = capi.Stream()
            sp = stream

            def event():
                return capi.Event()

            def sync_all():
                stream.synchronize()
        else:
            sp = None

            def sync_all():
                pass
        jobs = sorted(jobs, key=lambda j: j["arrival"])
        specs = {}                                    # job index -> Future of its scene spec
        staged = {}                                   # job index -> (scene, Future of its host tables, spec)
        nxt = 0                                       # next job not yet arrived
        queue = collections.deque()                   # arrived, not admitted (job indices)
        active = {}                                   # scene id -> [job index, steps left, tets]
        used = 0
        admitted, done, rejected = {}, 0, 0
        waits, step_records = [], []                  # (start event, end event, [tets per job])
        job_records = []                              # plain mode: (start event, end event, tets) per job step
        elem_steps = 0
        t0 = time.perf_counter()
        inflight = collections.deque()

        prepare = getattr(self.eng, "prepare_scene", None)

        def stage(j, wait=False):
            """Job j's mesh is built, then its scene created and its host tables (tiling,
            element set-up) built on a worker thread, while the GPU runs the current tick."""
            if j in staged:
                return True
            if j not in specs:
                specs[j] = self.pool.submit(build, jobs[j])
            if not wait and not specs[j].done():
                return False
            spec = specs.pop(j).result()
            sid = workloads.add_to_engine(self.eng, spec)
            fut = self.pool.submit(prepare, sid) if prepare is not None else None
            staged[j] = (sid, fut, spec)
            return True

        while True:
            now = time.perf_counter() - t0
            if now >= wall_limit:
                break
            while nxt < len(jobs) and jobs[nxt]["arrival"] <= now:
                queue.append(nxt)
                nxt += 1
            for k, j in enumerate(queue):               # the queue's head is staged ahead
                if k >= 32:
                    break
                stage(j)
            # admission: FIFO while the head fits the budget
            while queue:
                j = queue[0]
                if not stage(j, wait=not active):
                    break                              # its mesh is still being built
                sid, fut, spec = staged[j]
                need = footprint_bytes(self.precision, len(spec["pos"]), len(spec["tets"]), self.reference_order)
                if used + need > self.budget and active:
                    break
                if fut is not None:
                    if active and not fut.done():
                        break                          # host tables not ready: next tick
                    fut.result()
                if not self.eng.commit_scene(sid):
                    if active:
                        break                          # arenas full (fragmentation): wait
                    queue.popleft()                    # larger than the whole reservation
                    del staged[j]
                    rejected += 1
                    continue
                queue.popleft()
                del staged[j]
                used += need
                active[sid] = [j, jobs[j]["steps"], len(spec["tets"]), need]
                admitted[j] = time.perf_counter() - t0
                waits.append(admitted[j] - jobs[j]["arrival"])
            if not active:
                if nxt >= len(jobs) and not queue:
                    break
                time.sleep(0.001)
                continue
            ids = list(active.keys())
            e0 = event()
            e1 = event()
            e0.record(stream)
            if self.batched:
                self.eng.step(ids, 1, stream=sp)
            else:
                for s in ids:
                    j0, j1 = event(), event()
                    j0.record(stream)
                    self.eng.step([s], 1, stream=sp)
                    j1.record(stream)
                    job_records.append((j0, j1, active[s][2]))
            e1.record(stream)
            tets = [active[s][2] for s in ids]
            step_records.append((e0, e1, tets))
            elem_steps += sum(tets)
            inflight.append(e1)
            finished = []
            for s in ids:
                active[s][1] -= 1
                if active[s][1] == 0:
                    finished.append(s)
            # at most two ticks in flight: the host prepares the next one meanwhile
            while len(inflight) > 1:
                inflight.popleft().synchronize()
            for s in finished:
                j, _, _, need = active.pop(s)
                if on_finish is not None:
                    on_finish(self.eng, s, j)
                self.eng.release_scene(s)
                used -= need
                done += 1
            if log is not None and len(step_records) % 50 == 0:
                log(f"t={now:.1f}s ticks={len(step_records)} active={len(active)} queue={len(queue)} done={done}")
        sync_all()
        wall = time.perf_counter() - t0
        lat = []
        for e0, e1, tets in step_records:
            ms = e0.elapsed_time(e1)
            lat.extend([ms] * len(tets))
        lat = np.asarray(lat) if lat else np.zeros(1)
        waits_a = np.asarray(waits) if waits else np.zeros(1)
        by_size = None
        if job_records:
            tets = np.array([t for _, _, t in job_records], dtype=np.float64)
            ms = np.array([a.elapsed_time(b) for a, b, _ in job_records])
            edges = [0, 1e4, 3e4, 1e5, 3e5, 1e6, np.inf]
            by_size = []
            for lo, hi in zip(edges[:-1], edges[1:]):
                sel = (tets >= lo) & (tets < hi)
                if sel.any():
                    by_size.append({"tets": [lo, None if hi == np.inf else hi], "job_steps": int(sel.sum()),
                                    "element_steps": int(tets[sel].sum()),
                                    "median_ms_per_step": float(np.median(ms[sel])),
                                    "element_steps_per_s_alone": float(tets[sel].sum() / (ms[sel].sum() / 1e3))})
        return {"mode": "batched" if self.batched else "plain", "wall_s": wall, "element_steps": int(elem_steps),
                "value": elem_steps / wall, "ticks": len(step_records), "jobs_arrived": nxt,
                "jobs_admitted": len(admitted), "jobs_done": done, "jobs_rejected": rejected, "jobs_running_at_end": len(active),
                "queue_at_end": len(queue),
                "step_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
                "admission_wait_s": {"p50": float(np.percentile(waits_a, 50)), "p99": float(np.percentile(waits_a, 99))},
                "budget_bytes": self.budget, "plain_by_job_size": by_size}
