/* sofa_amd.h -- C ABI of the MI355X (gfx950) implicit corotational-FEM step.
 *
 * One shared library, libsofa_amd.so, exports everything below.  Plain C types only: device
 * vectors are pointers to AoS xyz triples (SOFA's Vec3d layout, VecTypes.h:200-234), streams
 * are hipStream_t passed as void*.  Every function returns an int status (SOFA_AMD_OK = 0,
 * negative on error) and never throws; sofa_amd_last_error() gives the message.  This maps
 * onto SOFA's msg_error + ComponentState::Invalid convention (TetrahedronFEMForceField.inl
 * :1262-1311) in the plugin adapter.
 *
 * Reference interfaces each entry point stands in for are cited per function (paths relative
 * to the sofa-framework/sofa source tree).
 */
#ifndef SOFA_AMD_H
#define SOFA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SOFA_AMD_OK            0
#define SOFA_AMD_EINVAL       -1   /* invalid argument / mesh index out of range            */
#define SOFA_AMD_EDEVICE      -2   /* HIP runtime error or no gfx950 device                  */
#define SOFA_AMD_ENOMEM       -3   /* device allocation failed / admission refused           */
#define SOFA_AMD_EDEGENERATE  -4   /* zero-volume tetrahedron (K / (36 V) would divide by 0)  */
#define SOFA_AMD_ESTATE       -5   /* call out of order (e.g. step before commit)            */
#define SOFA_AMD_ENAN         -6   /* non-finite state detected                              */

/* precision modes */
#define SOFA_AMD_FP64        0   /* double, every sum in the reference's order: the parity mode */
#define SOFA_AMD_FP32        1   /* float element arithmetic, state and CG vectors; FP64 dots;
                                    tiled LDS-gather kernels                                     */
#define SOFA_AMD_MIXED       2   /* FP32 element arithmetic + FP32 inner CG, FP64 state and
                                    iterative refinement (r = b - A x in FP64)                   */
#define SOFA_AMD_FP64_TILED  3   /* double on the tiled kernels: nodes shared by several tiles
                                    are summed tile by tile (deterministic, not reference order);
                                    with sofa_amd_engine_set_dot_order(eng, 1), set before the
                                    scenes are placed, every sum in the reference's order: the
                                    FP64 parity trajectories bit for bit, on the tiled kernels   */
#define SOFA_AMD_SPLIT       4   /* FP32 element record and element arithmetic (tiled kernels);
                                    FP64 node vectors (state, CG vectors), per-node sums, CG dot
                                    products and scalars                                         */

/* mass kinds */
#define SOFA_AMD_MASS_UNIFORM            0  /* UniformMass (UniformMass.inl:403-496)                  */
#define SOFA_AMD_MASS_DIAGONAL           1  /* DiagonalMass on tetrahedra (DiagonalMass.inl:1060-1080) */
#define SOFA_AMD_MASS_MESHMATRIX         2  /* MeshMatrixMass, lumping = false: consistent mass with
                                               vertex and edge terms (MeshMatrixMass.inl:547-657,
                                               1987-2092)                                          */
#define SOFA_AMD_MASS_MESHMATRIX_LUMPED  3  /* MeshMatrixMass, lumping = true (coefficient 2.5)      */
#define SOFA_AMD_MASS_DIAGONAL_TOTAL     4  /* DiagonalMass given its totalMass: vertex masses for
                                               density 1, then scaled by totalMass / their sum
                                               (DiagonalMass.inl:1235-1257); the mass value is the
                                               total mass                                          */

typedef struct sofa_amd_engine_s* sofa_amd_engine;

/* CG statistics of the last solve of one scene: CGLinearSolver's "graph" Data and the
 * AdvancedTimer "CG iterations" value (CGLinearSolver.inl:109-114,306). */
typedef struct sofa_amd_cg_stats {
    int32_t iterations;     /* nb_iter at loop exit                                          */
    int32_t end_reason;     /* 0 iterations, 1 tolerance, 2 threshold, 3 den == 0, 4 |b| == 0 */
    int32_t step_count;     /* CGLinearSolver::timeStepCount                                 */
    int32_t ir_sweeps;      /* iterative-refinement sweeps (MIXED), else 0                   */
    double  err;            /* last sqrt(r.r)/|b|                                            */
    double  den;            /* last p.A.p                                                    */
    double  time;           /* context time of the scene                                     */
    int32_t dot_fallbacks;  /* FP64_TILED in the reference order with segmented dot products:
                               segment boundaries where no candidate start matched (the segment
                               was summed from the known start instead) since the scene was
                               placed; else 0                                                    */
    int32_t reserved;       /* 0                                                                 */
} sofa_amd_cg_stats;

/* ---- library ---------------------------------------------------------------------------- */

/* Plugin-module identity, as PluginManager resolves it (PluginManager.cpp:89-94). */
const char* sofa_amd_module_name(void);
const char* sofa_amd_module_version(void);
/* Message of the last failing call on this thread. */
const char* sofa_amd_last_error(void);
/* 0 if `device` is a usable gfx950 GPU; the plugin's moduleIsInitialized() reports it
 * (PluginManager.cpp:258-266 unloads a plugin that answers false). */
int sofa_amd_device_check(int device);

/* Free and total device memory of `device` (hipMemGetInfo): a serving loop sizes its admission
 * budget from it.  No reference counterpart. */
int sofa_amd_device_memory(int device, int64_t* free_bytes, int64_t* total_bytes);

/* ---- engine: a set of device-resident scenes advanced in merged launches ----------------- */

/* No reference counterpart (SOFA runs one scene per process); holds device buffers. */
int sofa_amd_engine_create(int device, int precision, sofa_amd_engine* out);
int sofa_amd_engine_destroy(sofa_amd_engine eng);

/* A scene = MechanicalObject<Vec3d> (MechanicalObject.h:85-93) with rest positions
 * `positions` (n_nodes x 3, host).  Returns the scene id. */
int sofa_amd_scene_create(sofa_amd_engine eng, int n_nodes, const double* positions, int* out_scene);

/* TetrahedronFEMForceField<Vec3d> method="large": init + reinit
 * (TetrahedronFEMForceField.inl:1257-1545).  `tets` n_tets x 4 node indices (host);
 * youngModulus / poissonRatio per element or a single value
 * (BaseLinearElasticityFEMForceField.inl:113-144). */
int sofa_amd_scene_set_fem(sofa_amd_engine eng, int scene, int n_tets, const int32_t* tets,
                           const double* young, int n_young, const double* poisson, int n_poisson);

/* The same on hexahedra (n_hexa x 8 corners, RegularGridTopology order), tessellated as
 * TetrahedronFEMForceField::init does when its topology has no tetrahedra
 * (TetrahedronFEMForceField.inl:1318-1370): 6 tetrahedra per cube, corners mirrored by cube
 * parity along x / y / z; nx, ny = cubes along x and y of the grid (the reference's defaults for
 * a topology that is not a grid: 2 and 1). Element e of the scene is tetrahedron e % 6 of cube
 * e / 6. */
int sofa_amd_scene_set_fem_hexa(sofa_amd_engine eng, int scene, int n_hexa, const int32_t* hexa, int nx, int ny,
                                const double* young, int n_young, const double* poisson, int n_poisson);

/* UniformMass vertexMass (UniformMass.inl:403-496), DiagonalMass massDensity on the FEM's
 * tetrahedra (DiagonalMass.inl:1060-1080, 536-559, 1393-1413) or a uniform MeshMatrixMass
 * massDensity. */
int sofa_amd_scene_set_mass(sofa_amd_engine eng, int scene, int kind, double value);

/* MeshMatrixMass massDensity, one value or one per tetrahedron (MeshMatrixMass.h:88). */
int sofa_amd_scene_set_mass_density(sofa_amd_engine eng, int scene, int kind, const double* density, int n);

/* OscillatorProjectiveConstraint entries (OscillatorProjectiveConstraint.inl:65-106): node index,
 * mean (3), amplitude (3), pulsation, phase.  The node's response is cleared in the solve and,
 * after it, x = mean + amplitude sin(w t + phase), v = amplitude w cos(w t + phase) with t the
 * scene time at the start of the step. */
int sofa_amd_scene_add_oscillators(sofa_amd_engine eng, int scene, int n, const int32_t* indices,
                                   const double* mean, const double* amplitude, const double* pulsation,
                                   const double* phase);

/* LinearMovementProjectiveConstraint<Vec3d> (LinearMovementProjectiveConstraint.inl:40-330): the
 * `n` nodes `indices` follow the piecewise-linear movements given at `n_keys` key times (in
 * chronological order; movements: 3 per key) while the scene time t (the step's start) lies in
 * [keyTimes[k], keyTimes[k+1]) for some k: their responses are cleared in the solve, and after
 * the position update x = x0 + m(t) (relative = 1; x0 = their positions at the first projection)
 * or x = m(t), v = (m[k+1] - m[k]) / (keyTimes[k+1] - keyTimes[k]), with
 * m(t) = m[k] + (m[k+1] - m[k]) (t - keyTimes[k]) / (keyTimes[k+1] - keyTimes[k]).  Outside the
 * window (and at the last key time itself) the nodes are free.  Applied after the fixed and
 * oscillator constraints; one per scene. */
int sofa_amd_scene_add_linear_movement(sofa_amd_engine eng, int scene, int n, const int32_t* indices, int n_keys,
                                       const double* key_times, const double* movements, int relative);

/* PCGLinearSolver (PCGLinearSolver.inl:178-262) in place of the scene's CGLinearSolver, with a
 * JacobiPreconditioner (kind 1) or a BlockJacobiPreconditioner (kind 2) linked; 0 restores the
 * CGLinearSolver.  The CG Data of sofa_amd_scene_set_solver are the PCG's: `iterations` and
 * `tolerance` (on r.z relative to b.b; no threshold).  The preconditioner is assembled every
 * step (the PreconditionedMatrixFreeSystem's assemblingRate 1): the 3x3 diagonal blocks of
 * (1 + h rm) M + h (h + rs) K with the step's rotations (Jacobi keeps their diagonals), the
 * identity on fixed and LinearMovement nodes (Dirichlet rows), inverted.  The solve starts from
 * the previous step's solution, as the reference's does (the solver's solution vector persists);
 * the statistics report PCG's iteration count (maxIter + 1 when the budget runs out) and
 * err = r.z / b.b.  FP64 engine.  On a MIXED engine (kind 1 or 2): the inner FP32 CG runs
 * preconditioned by the same inverted diagonal or 3x3 blocks (in float), z = P^-1 r, its
 * residual measured as r.z; the FP64 refinement and its |r| / |b| test are unchanged (no
 * reference counterpart: MIXED is compared with converged FP64). */
int sofa_amd_scene_set_preconditioner(sofa_amd_engine eng, int scene, int kind);

/* CGLinearSolver Data warmStart (CGLinearSolver.inl:39,94-105): each solve starts from the
 * solver's solution vector, which keeps the previous step's solution (0 before the first step):
 * r = A x, then r.eq(b, r, -1.0); |b| and the loop as without it.  The first step is the cold
 * solve bit for bit.  FP64 engine, or FP64_TILED in the reference's summation order (set with
 * sofa_amd_engine_set_dot_order before this call): there the operator of an iteration 0 gives
 * A x.  A PCGLinearSolver scene warm-starts anyway. */
int sofa_amd_scene_set_warm_start(sofa_amd_engine eng, int scene, int on);

/* Context gravity (Node::setGravity), used by the masses' addForce. */
int sofa_amd_scene_set_gravity(sofa_amd_engine eng, int scene, double gx, double gy, double gz);

/* FixedProjectiveConstraint indices (FixedProjectiveConstraint.inl:184-207). */
int sofa_amd_scene_set_fixed(sofa_amd_engine eng, int scene, const int32_t* indices, int n);

/* FixedProjectiveConstraint fixAll (FixedProjectiveConstraint.inl:190-195): every node's response
 * is cleared. */
int sofa_amd_scene_set_fix_all(sofa_amd_engine eng, int scene, int fix_all);

/* FixedProjectiveConstraint activate_projectVelocity (FixedProjectiveConstraint.inl:237-258;
 * default off): after each step's position update the fixed nodes' velocities (every node's
 * with fixAll) are cleared -- DefaultAnimationLoop's projectPositionAndVelocity
 * (DefaultAnimationLoop.cpp:228-234), applied before the oscillator's projection. */
int sofa_amd_scene_set_project_velocity(sofa_amd_engine eng, int scene, int on);

/* PlaneForceField (PlaneForceField.inl:86-224), applied after the FEM in scene order: every node
 * below the plane x.normal = d (either side if bilateral) gets the force
 * normal (-stiffness dist) + v (damping dist), dist = x.normal - d, clamped to max_force when
 * max_force > 0; its stiffness -stiffness normal normal^T enters the implicit system for the
 * nodes in contact at the step's force evaluation (the damping does not). One plane per scene;
 * the reference's defaults: normal (0, 1, 0), d 0, stiffness 500, damping 5, max_force 0. */
int sofa_amd_scene_add_plane(sofa_amd_engine eng, int scene, const double* normal, double d, double stiffness,
                             double damping, double max_force, int bilateral);

/* EulerImplicitSolver rayleighStiffness / rayleighMass (EulerImplicitSolver.cpp:40-51) and
 * CGLinearSolver iterations / tolerance / threshold (CGLinearSolver.inl:36-40), dt. */
int sofa_amd_scene_set_solver(sofa_amd_engine eng, int scene, double dt, double rayleigh_stiffness,
                              double rayleigh_mass, int cg_iterations, double cg_tolerance,
                              double cg_threshold);

/* EulerImplicitSolver firstOrder, trapezoidalScheme and vdamping (EulerImplicitSolver.cpp:40-51;
 * used at :118-123, :136-173 and :246-299): firstOrder solves (M - h K) v = f and sets v = dv;
 * trapezoidal uses h/2 in the stiffness and Rayleigh terms; vdamping multiplies the new
 * velocity by exp(-h vdamping) after the position update.  Defaults: 0, 0, 0. */
int sofa_amd_scene_set_integrator(sofa_amd_engine eng, int scene, int first_order, int trapezoidal,
                                  double vdamping);

/* SOFA_AMD_MIXED engines: the linear solve of each step is iterative refinement -- r = b - A x
 * in FP64, stop when |r| / |b| <= cg_tolerance (the CGLinearSolver tolerance of
 * sofa_amd_scene_set_solver) or after `max_sweeps` corrections, each correction an FP32 CG
 * solve of A d = r of at most `inner_iterations` iterations stopping at relative residual
 * `inner_tolerance`.  With keep_direction = 0 each correction starts CG afresh (classic
 * refinement); with 1 it continues the previous correction's search direction from the
 * replaced residual (FP32 CG with FP64 residual replacement: the convergence of one unbroken CG,
 * the inner tolerance then being relative to the first sweep's residual).  Defaults 30, 1e-6, 3,
 * 0.  No reference counterpart (SOFA's CG runs in the state's precision); it replaces
 * CGLinearSolver::solve (CGLinearSolver.inl:71-315) in this mode.  sofa_amd_cg_stats reports
 * the inner iterations summed over the sweeps and the sweep count. */
int sofa_amd_scene_set_refinement(sofa_amd_engine eng, int scene, int inner_iterations,
                                  double inner_tolerance, int max_sweeps, int keep_direction);

/* Uploads every scene created so far; after this the scenes can be stepped. */
int sofa_amd_engine_commit(sofa_amd_engine eng);

/* node::animate x n_steps (Simulation.cpp:207-227 -> DefaultAnimationLoop::step ->
 * EulerImplicitSolver::solve -> CGLinearSolver::solve) for the listed scenes, merged into one
 * launch per phase.  Asynchronous on `stream` (NULL = the engine's stream). */
int sofa_amd_engine_step(sofa_amd_engine eng, const int32_t* scenes, int n_scenes, int n_steps,
                         void* stream);

/* ---- live scenes (tiled engines: FP32, FP64_TILED, MIXED) -------------------------------- */

/* Instead of sofa_amd_engine_commit: reserve device arenas for up to `max_nodes` nodes,
 * `max_tets` tetrahedra and `max_scenes` scenes in total, after which scenes are created,
 * committed and released one at a time while the others keep running (the serving loop's
 * continuous batching; admission is by the arenas' room).  No reference counterpart: SOFA
 * builds one scene graph per process. */
int sofa_amd_engine_reserve(sofa_amd_engine eng, int64_t max_nodes, int64_t max_tets, int max_scenes);
/* Places one created scene in the reserved arenas; SOFA_AMD_ENOMEM (and nothing kept) if they
 * have no room for it now. */
int sofa_amd_scene_commit(sofa_amd_engine eng, int scene);
/* Builds a created scene's host tables (tiling, element set-up) ahead of its commit, which then
 * only places and uploads them.  May run on another host thread than the one stepping the
 * engine (several scenes in parallel); the scene must not be changed or committed meanwhile. */
int sofa_amd_scene_prepare(sofa_amd_engine eng, int scene);
/* Waits for the device, then returns the scene's ranges to the arenas; its id stays invalid. */
int sofa_amd_scene_release(sofa_amd_engine eng, int scene);
/* Arena use: out22[2k] used, out22[2k+1] capacity, k = nodes, elements, incidences, tiles,
 * tile-local nodes, shared nodes, shared partial slots, shared-node blocks, node blocks,
 * oscillator entries, scene records. */
int sofa_amd_engine_usage(sofa_amd_engine eng, int64_t* out22);

/* Replay each step from a captured hipGraph (one graph launch per step; the graph of a launch
 * set is captured on its first step and kept until the set changes).  Off by default.  No
 * reference counterpart: SOFA has no device graphs. */
int sofa_amd_engine_set_graphs(sofa_amd_engine eng, int enable);

/* A merged step in concurrent branches: the stepped scenes are split into `branches` groups
 * (largest node counts dealt round-robin), each advanced by its own launch sequence on its own
 * stream, forked from and joined back into the step's stream, so that one group's latency-bound
 * phases overlap the others' kernels.  0 (default): on FP64_TILED engines in the reference's
 * summation order, whose node-after-node dot products are single-wave chains, one branch per
 * hardware queue the runtime leaves free (HIP's GPU_MAX_HW_QUEUES, 4 unless set, minus the two
 * its own streams hold), at most 4; else 1.  At most 8, and at most one branch per two scenes.
 * Branches only run concurrently on distinct hardware queues.  Results do not depend on it (scenes are
 * independent).  No reference counterpart. */
int sofa_amd_engine_set_branches(sofa_amd_engine eng, int branches);

/* FP64_TILED engines in the reference's summation order, stepped in more than one branch: 1
 * makes the branches' operator phases take turns: in every CG iteration each group's
 * operator and finalize kernels wait for the previous group's (round-robin), so that one group's
 * two node-after-node dot products (single-wave chains) run while the GPU runs another group's
 * operator alone, instead of every group's operator at once and then every group's chains on an
 * idle GPU.  Such steps are launched directly, without graphs (a graph per branch cannot wait
 * for another branch's kernels).  0 (default): the branches run freely, each replayed from its own
 * graph when graphs are on (measured faster on the 256-beam mix: 21.6 against 22.3 ms/step with
 * 4 branches).  Results do not depend on it.  No reference counterpart. */
int sofa_amd_engine_set_branch_turns(sofa_amd_engine eng, int enable);

/* FP64_TILED engines in the reference's summation order (set before the scenes are placed): how
 * each scene's node-after-node dot product (|b|^2, p.q, r.r) is computed.  segments = 1: one
 * single-wave chain of dependent adds per scene.  segments = 2, 4, 8 (default) or 16 (with a
 * window of 32 at most): the chain's
 * segments run at once, each segment after the first from 2 `window` candidate start values
 * (consecutive doubles around a prediction of the sum before it: the terms each rounded as the
 * chain would round them; window 0, 32, 64, 96 or 128, default 32), and the sum is then resolved
 * boundary by boundary: the candidate equal to the
 * sum so far, bit for bit, continues it; where none is, the segment is summed from the sum so
 * far (counted in sofa_amd_cg_stats::dot_fallbacks).  Every add of the chain happens in the
 * reference's order on its exact operands: the result has the same bits either way.  Scenes
 * under 8192 nodes always use the single chain.  No reference counterpart. */
int sofa_amd_engine_set_dot_segments(sofa_amd_engine eng, int segments, int window);

/* SOFA_AMD_FP64 and SOFA_AMD_FP64_TILED engines: order of the sums.  0 (default): CG dot
 * products (|b|^2, p.q, r.r) as per-block sums combined in a fixed order (parallel,
 * deterministic); on the tiled engine, nodes shared by several tiles summed tile by tile.  1:
 * the reference's order -- dot products node after node from 0 (MechanicalObject::vDot,
 * MechanicalObject.inl:2334-2356), and on the tiled engine every node's element contributions
 * one by one in ascending element id (TetrahedronFEMForceField.inl:1630-1634) -- so that
 * trajectories equal the reference's bit for bit however sensitive the scene; costs one dependent
 * add per node and dot.  A tiled engine takes it before its scenes are placed (commit / reserve)
 * and then refuses a consistent MeshMatrixMass (SOFA_AMD_EINVAL at commit). */
int sofa_amd_engine_set_dot_order(sofa_amd_engine eng, int order);

/* Tiled engines: the CG operator kernel's variant -- threads per element and how much element
 * geometry it recomputes from positions: 0: 1 thread per element (512 per tile), rotation and
 * rest coordinates recomputed (the default); 1: 512 threads, rotation only; 2: 256 threads (2
 * elements each), rotation and rest coordinates; 3: 256 threads, rotation only.  Variants differ
 * in rounding only (the order of block sums, geometry recomputed in the engine's precision).
 * sofa_amd_engine_autotune times all of them on the committed scenes (`reps` launches each, one
 * launch over every scene) and keeps the fastest, a later variant replacing an earlier one only
 * when faster by 1 % or more; it must run before the first step.  `chosen`
 * and `us4` (the 4 timings in microseconds per launch) may be NULL.  No reference counterpart. */
int sofa_amd_engine_autotune(sofa_amd_engine eng, int reps, int* chosen, float* us4);
int sofa_amd_engine_set_apply_variant(sofa_amd_engine eng, int variant);

/* Blocks until the engine's work is done. */
int sofa_amd_engine_sync(sofa_amd_engine eng);

/* Host copies of a scene's position / velocity / force (n_nodes x 3 doubles each; NULL skips).
 * The copies are made in any case; SOFA_AMD_ENAN when one of them holds a NaN or an Inf. */
int sofa_amd_scene_read(sofa_amd_engine eng, int scene, double* x, double* v, double* f);
/* Overwrites position / velocity (NULL skips); SOFA_AMD_ENAN (nothing written) for a NaN or Inf. */
int sofa_amd_scene_write(sofa_amd_engine eng, int scene, const double* x, const double* v);
int sofa_amd_scene_stats(sofa_amd_engine eng, int scene, sofa_amd_cg_stats* out);
/* TetrahedronFEMForceField::computeVonMisesStress (TetrahedronFEMForceField.inl:2197-2347, the
 * Data vonMisesPerElement / vonMisesPerNode it fills at each AnimateEndEvent) from the scene's
 * current positions: mode 1 corotational strain, mode 2 full Green strain (the reference's
 * computeVonMisesStress values; 0 is "no computation" there and SOFA_AMD_EINVAL here).
 * per_element: n_tets values in element order, per_node: n_nodes values (the mean over the
 * node's tetrahedra; NULL skips either).  The element shape functions are built on the first
 * call for a scene and kept on the device (120 B per element + 32 B per node).  Mode 1 also
 * stores the element rotations of the current positions, as the reference does.  Blocks until
 * done; SOFA_AMD_EDEGENERATE when a rest tetrahedron's matrix cannot be inverted. */
int sofa_amd_scene_von_mises(sofa_amd_engine eng, int scene, int mode, double* per_element, double* per_node);
/* TetrahedronFEMForceField::addKToMatrix / buildStiffnessMatrix (TetrahedronFEMForceField.inl:
 * 1921-1999, computeStiffnessMatrix :212-232): the scene's assembled stiffness
 * -k_factor * sum_e R_e J_e K_e J_e^T R_e^T with the rotations of the last step or addForce, as a
 * block-sparse matrix of 3x3 blocks over the scene's nodes: block (a, b) exists when nodes a and
 * b share an element; row_ptr (n_nodes + 1) and cols (*nnzb, ascending in a row) give the
 * pattern, values (9 * *nnzb, each block row-major) the entries, every block summed over its
 * elements in ascending element id.  NULL arrays are skipped: a first call with cols and values
 * NULL returns *nnzb.  The pattern and each block's element list are built on the first call
 * for a scene and kept on the device (about 64 B per element + 76 B per block).  Blocks until
 * done. */
int sofa_amd_scene_stiffness(sofa_amd_engine eng, int scene, double k_factor, int32_t* row_ptr, int32_t* cols,
                             double* values, int64_t* nnzb);
/* Device-side element data of one element as the reference stores it (for parity checks):
 * rotations[e] (3x3), K diagonal/off-diagonal/shear (3), rotated rest coordinates (4x3). */
int sofa_amd_scene_element(sofa_amd_engine eng, int scene, int elem, double* rot9, double* k3,
                           double* rest12);

/* ---- component-level calls, one TetrahedronFEMForceField call each ----------------------- */

/* TetrahedronFEMForceField::addForce (TetrahedronFEMForceField.inl:1548-1604):
 * f += F(x), and the element rotations are updated from x.  x_dev, f_dev: device, n_nodes x 3
 * AoS in the engine's state type: double for SOFA_AMD_FP64 (Vec3d), float for SOFA_AMD_FP32
 * (Vec3f), as a MechanicalObject<Vec3d> / <Vec3f> would hold them. */
int sofa_amd_fem_add_force(sofa_amd_engine eng, int scene, const void* x_dev, void* f_dev,
                           void* stream);

/* TetrahedronFEMForceField::addDForce (TetrahedronFEMForceField.inl:1606-1636):
 * df -= kFactor * K(x) * dx with the rotations of the last addForce; same vector types as
 * sofa_amd_fem_add_force. */
int sofa_amd_fem_add_dforce(sofa_amd_engine eng, int scene, const void* dx_dev, void* df_dev,
                            double k_factor, void* stream);

/* ---- component-level calls of the other components on the path ---------------------------
 * Caller vectors are device pointers to the scene's n_nodes x 3 AoS values in the engine's state
 * type (double for SOFA_AMD_FP64, FP64_TILED, MIXED and SPLIT; float for SOFA_AMD_FP32), enqueued on
 * `stream` (NULL = the engine's stream) unless stated otherwise. */

/* Mass::addMDx (Mass.h:69): f += factor M dx -- UniformMass.inl:403-419, DiagonalMass.inl:536-559,
 * MeshMatrixMass.inl:1987-2047 (the consistent mass's edge terms after the vertex term, in edge
 * order; on the FP64 engine only: SOFA_AMD_EINVAL on the tiled engines). */
int sofa_amd_mass_add_mdx(sofa_amd_engine eng, int scene, const void* dx_dev, void* f_dev, double factor,
                          void* stream);
/* Mass::accFromF (Mass.h:77): a = M^-1 f -- UniformMass.inl:423-436, DiagonalMass.inl:564-576,
 * MeshMatrixMass.inl:2051-2069 (lumped only; SOFA_AMD_EINVAL for the consistent mass, which the
 * reference refuses with msg_error). */
int sofa_amd_mass_acc_from_f(sofa_amd_engine eng, int scene, const void* f_dev, void* a_dev, void* stream);
/* Mass::addForce (Mass.h:66): f += the weight under the scene's gravity -- UniformMass.inl:471-496,
 * DiagonalMass.inl:1393-1413, MeshMatrixMass.inl:2073-2092. */
int sofa_amd_mass_add_force(sofa_amd_engine eng, int scene, void* f_dev, void* stream);
/* ProjectiveConstraintSet::projectResponse (ProjectiveConstraintSet.h:112-116) of the scene's
 * constraints: FixedProjectiveConstraint (FixedProjectiveConstraint.inl:184-207, fixAll) and
 * OscillatorProjectiveConstraint (OscillatorProjectiveConstraint.inl:65-70) clear their nodes. */
int sofa_amd_project_response(sofa_amd_engine eng, int scene, void* dx_dev, void* stream);
/* ProjectiveConstraintSet::projectVelocity (ProjectiveConstraintSet.h:118-120) at context time
 * `time`: FixedProjectiveConstraint clears its nodes when activate_projectVelocity is set
 * (FixedProjectiveConstraint.inl:237-258), then OscillatorProjectiveConstraint sets
 * v = a w cos(w t + phi) (OscillatorProjectiveConstraint.inl:72-87). */
int sofa_amd_project_velocity(sofa_amd_engine eng, int scene, void* v_dev, double time, void* stream);
/* ProjectiveConstraintSet::projectPosition (ProjectiveConstraintSet.h:122-123): a no-op for
 * FixedProjectiveConstraint (.inl:262-265); OscillatorProjectiveConstraint sets
 * x = mean + a sin(w t + phi) (OscillatorProjectiveConstraint.inl:89-106). */
int sofa_amd_project_position(sofa_amd_engine eng, int scene, void* x_dev, double time, void* stream);
/* BaseMechanicalState::vDot (BaseMechanicalState.h:147; MechanicalObject::vDot,
 * MechanicalObject.inl:2334-2356) of two caller vectors over the scene's nodes, into
 * *result_dev (one double in device memory).  order 1: the reference's sequential sum, node after
 * node from 0, bit for bit; order 0: the FP64 engine's parallel blocked order (256-node block
 * trees, then the block partials), whose per-block partials live in a buffer the library
 * allocates on the engine and grows to the largest scene asked for. */
int sofa_amd_vdot(sofa_amd_engine eng, int scene, const void* a_dev, const void* b_dev, int order,
                  double* result_dev, void* stream);
/* sofa_amd_vdot without any device allocation: no reference counterpart.  For callers that
 * cannot allocate device memory (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same kernels and result as sofa_amd_vdot, bit for bit, but the order-0
 * block partials go to a file-scope device array of the library (512 KiB per GPU, held until the
 * process ends) instead of a buffer the library allocates on the engine: scenes of up to
 * 16,777,216 nodes (SOFA_AMD_EINVAL beyond).  Every call on one GPU shares that GPU's array, so
 * calls that overlap on one GPU (other streams, other host threads, other engines) race; calls
 * ordered on one stream do not.  Enqueued on `stream` (NULL = the engine's stream). */
int sofa_amd_vdot_noalloc(sofa_amd_engine eng, int scene, const void* a_dev, const void* b_dev, int order,
                          double* result_dev, void* stream);
/* CGLinearSolver::solve (CGLinearSolver.h:68, CGLinearSolver.inl:71-315) of
 * (m_factor M + b_factor B + k_factor K) x = b on the scene, through MatrixLinearSolver::solveSystem
 * (MatrixLinearSolver.inl:218-233) and GraphScatteredMatrix::apply (GraphScatteredTypes.cpp:33-45):
 * mass addMDx, FEM addDForce with the rotations of the last step or addForce, PlaneForceField
 * contacts, projectResponse.  The CG Data are the scene's (sofa_amd_scene_set_solver; x starts at
 * 0: warmStart false), its dot products in the engine's order (sofa_amd_engine_set_dot_order).  B
 * is the components' own Rayleigh damping, which this path's components do not have (their
 * rayleighMass / rayleighStiffness Data are 0), so b_factor adds nothing.  Counts as one solve of the
 * scene's CGLinearSolver (timeStepCount).  Not on MIXED engines.  Blocks until x_dev holds the
 * solution; stats may be NULL. */
int sofa_amd_cg_solve(sofa_amd_engine eng, int scene, double m_factor, double b_factor, double k_factor,
                      const void* b_dev, void* x_dev, sofa_amd_cg_stats* stats, void* stream);
/* ForceField::getPotentialEnergy (ForceField.h:121) of TetrahedronFEMForceField: the reference
 * sums it for method "small" only and returns 0 for "large" (TetrahedronFEMForceField.inl:625-717),
 * which is what this gives (synchronous). */
int sofa_amd_fem_potential_energy(sofa_amd_engine eng, int scene, const void* x_dev, double* energy);

/* ---- runtime plumbing (no reference counterpart) -------------------------------------------
 * Streams (default priority), timing events and device memory of the library's own HIP runtime,
 * for callers without a device-array library: the Python driver, the tests and the benchmark use
 * these, so that a process maps one HIP / HSA runtime only.  All synchronous but event_record. */
int sofa_amd_stream_create(int device, void** stream_out);
int sofa_amd_stream_destroy(void* stream);
int sofa_amd_stream_sync(void* stream);
int sofa_amd_device_sync(int device);
int sofa_amd_event_create(void** event_out);
int sofa_amd_event_destroy(void* event);
int sofa_amd_event_record(void* event, void* stream);
int sofa_amd_event_sync(void* event);
int sofa_amd_event_elapsed(void* start, void* end, float* ms);
int sofa_amd_malloc(int device, int64_t bytes, void** ptr_out);
int sofa_amd_free(void* ptr);
/* kind 0 host to device, 1 device to host, 2 device to device */
int sofa_amd_memcpy(void* dst, const void* src, int64_t bytes, int kind);
/* A plain copy kernel of `bytes` (a multiple of 16) timed over `reps` launches on `stream`: the
 * device-copy rate and the calibration of the HBM byte counters. */
int sofa_amd_copy_bench(int device, int64_t bytes, int reps, void* stream, float* us_per_copy);

/* ---- measurement and introspection (no reference counterpart) ---------------------------- */

/* `reps` launches of the CG operator kernel of one scene (scene >= 0) or of every scene of the
 * engine in one merged launch (scene = -1) -- tiled engines: the tile apply kernel with the
 * mid-loop data flow p = r + beta p, q = A p -- on the current state, for roofline timing
 * with events on `stream`.  Leaves p, q and the CG scalars of the scene undefined: step the
 * scene afresh (or discard the engine) afterwards. */
int sofa_amd_engine_bench_apply(sofa_amd_engine eng, int scene, int reps, void* stream);

/* Tiling of a scene: out7 = {tiles, tile-local nodes, nodes finished by the shared-node
 * kernel, shared partial slots, incidences, elements, nodes}; zeros but the last two on the
 * exact (untiled) path. */
int sofa_amd_scene_tiling(sofa_amd_engine eng, int scene, int32_t* out7);

#ifdef __cplusplus
}
#endif

#endif /* SOFA_AMD_H */
