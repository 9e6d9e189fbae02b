// engine.hip -- host side of libsofa_amd.so: scene tables, element initialisation, device
// buffers, launch schedule and the C ABI declared in include/sofa_amd.h.
//
// Element initialisation restates TetrahedronFEMForceField::reinit for method "large"
// (TetrahedronFEMForceField.inl:1391-1545): computeMaterialStiffness (:256-291), initLarge
// (:835-868) with computeRotationLarge (:755-778).  It runs once per scene on the host in
// double, in the reference's operation order (contraction off in these functions whatever the
// translation unit's flags), and the result is rounded to the element precision of the engine.
//
// The file is compiled twice: once with contraction off (the FP64 engines and the C ABI) and
// once with SOFA_AMD_TU_FP32 and contraction on (the FP32 engine, whose kernels may fuse
// multiply-adds), so that each precision gets its own device code.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <limits>
#include <mutex>
#include <array>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "sofa_amd.h"
#include "fem_kernels.h"
#include "fem_tiled.h"
#include "fem_mixed.h"
#include "tiling.h"

namespace sofa_amd {

static thread_local std::string g_last_error;

static int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

struct HipError : std::runtime_error {
    explicit HipError(const std::string& m) : std::runtime_error(m) {}
};

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess)                                                                     \
            throw HipError(std::string(#expr) + ": " + hipGetErrorString(_e));                   \
    } while (0)

// ---------------------------------------------------------------------------------------
// host restatement of the element set-up (double, reference order)
// ---------------------------------------------------------------------------------------
namespace host {
struct V3 { double v[3]; };
static inline V3 sub(const V3& a, const V3& b) {
    return {{a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]}};
}
static inline V3 cross(const V3& a, const V3& b) {
#pragma clang fp contract(off)
    return {{a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2], a.v[0] * b.v[1] - a.v[1] * b.v[0]}};
}
static inline double dot(const V3& a, const V3& b) {
#pragma clang fp contract(off)
    double r = a.v[0] * b.v[0];
    r += a.v[1] * b.v[1];
    r += a.v[2] * b.v[2];
    return r;
}
static inline V3 normalized(V3 a) {
#pragma clang fp contract(off)
    double n2 = a.v[0] * a.v[0];
    n2 += a.v[1] * a.v[1];
    n2 += a.v[2] * a.v[2];
    const double n = std::sqrt(n2);
    if (n > std::numeric_limits<double>::epsilon())
        for (int i = 0; i < 3; i++) a.v[i] /= n;
    return a;
}
// Tetrahedron.h:55-83
static inline double volume(const V3& n0, const V3& n1, const V3& n2, const V3& n3) {
#pragma clang fp contract(off)
    return std::abs(dot(cross(sub(n1, n0), sub(n2, n0)), sub(n3, n0)) / 6.0);
}

struct ElemInit {
    double rot[9];     // rotations[e] = _initialRotations[e] (transpose of R_0_1)
    double rest[9];    // rotated rest coordinates b, c, d (a = 0)
    double k[3];       // K00, K01, K33 after / (36 V)
    double vol;
};

// computeMaterialStiffness before the division by 36 V: K00, K01, K33 times 36 V, in
// init_element's operation order
static void material_factors(double E, double nu, double ks[3]) {
#pragma clang fp contract(off)
    const double young = 1.0f * E;
    const double k00 = 1, k01 = nu / (1 - nu), k33 = (1 - 2 * nu) / (2 * (1 - nu));
    const double s = (young * (1 - nu)) / ((1 + nu) * (1 - 2 * nu));
    ks[0] = k00 * s;
    ks[1] = k01 * s;
    ks[2] = k33 * s;
}

static bool init_element(const std::vector<V3>& P, const int32_t* t, double E, double nu, ElemInit& out) {
#pragma clang fp contract(off)
    const V3 &pa = P[t[0]], &pb = P[t[1]], &pc = P[t[2]], &pd = P[t[3]];
    // computeMaterialStiffness (localStiffnessFactor empty: youngModulus = 1.0f * E)
    const double young = 1.0f * E;
    const double k00 = 1, k01 = nu / (1 - nu), k33 = (1 - 2 * nu) / (2 * (1 - nu));
    const double s = (young * (1 - nu)) / ((1 + nu) * (1 - 2 * nu));
    const double vol = volume(pa, pb, pc, pd);
    const double div = vol * 36;
    out.k[0] = (k00 * s) / div;
    out.k[1] = (k01 * s) / div;
    out.k[2] = (k33 * s) / div;
    out.vol = vol;
    // initLarge: R_0_1 from (a, b, c); rotated rest shape relative to a
    const V3 ex = normalized(sub(pb, pa));
    V3 ey = sub(pc, pa);
    const V3 ez = normalized(cross(ex, ey));
    ey = cross(ez, ex);
    const V3 rows[3] = {ex, ey, ez};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out.rot[3 * i + j] = rows[j].v[i];
    V3 ri[4];
    const V3* pts[4] = {&pa, &pb, &pc, &pd};
    for (int k = 0; k < 4; k++)
        for (int i = 0; i < 3; i++) {
            double r = rows[i].v[0] * pts[k]->v[0];
            r += rows[i].v[1] * pts[k]->v[1];
            r += rows[i].v[2] * pts[k]->v[2];
            ri[k].v[i] = r;
        }
    for (int k = 1; k < 4; k++)
        for (int i = 0; i < 3; i++) ri[k].v[i] -= ri[0].v[i];
    for (int k = 1; k < 4; k++)
        for (int i = 0; i < 3; i++) out.rest[3 * (k - 1) + i] = ri[k].v[i];
    return vol > 0.0 && std::isfinite(out.k[0]);
}

// the reference's general matrix inverse (sofa/type/Mat.h:1104-1168), used for the shape
// functions of computeVonMisesStress: Gauss-Jordan elimination with full pivoting, the pivot the
// first strictly largest |entry| in row-major order over unused rows and columns, the pivot row
// divided, every other row reduced, rows permuted back at the end; false when a pivot is
// within the type's epsilon of zero
static bool invert4(const double a[4][4], double out[4][4]) {
#pragma clang fp contract(off)
    double m1[4][4], m2[4][4];
    int r[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0}, row[4] = {0, 0, 0, 0}, col[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            m1[i][j] = a[i][j];
            m2[i][j] = i == j ? 1.0 : 0.0;
        }
    for (int k = 0; k < 4; k++) {
        double pivot = 0;
        for (int i = 0; i < 4; i++) {
            if (row[i]) continue;
            for (int j = 0; j < 4; j++) {
                if (col[j]) continue;
                double t = m1[i][j];
                if (t < 0) t = -t;
                if (t > pivot) { pivot = t; r[k] = i; c[k] = j; }
            }
        }
        if (std::abs(pivot) <= std::numeric_limits<double>::epsilon()) return false;
        const int rk = r[k], ck = c[k];
        row[rk] = col[ck] = 1;
        pivot = m1[rk][ck];
        for (int j = 0; j < 4; j++) m1[rk][j] /= pivot;
        m1[rk][ck] = 1;
        for (int j = 0; j < 4; j++) m2[rk][j] /= pivot;
        for (int i = 0; i < 4; i++) {
            if (i == rk) continue;
            const double f = m1[i][ck];
            for (int j = 0; j < 4; j++) m1[i][j] -= m1[rk][j] * f;
            m1[i][ck] = 0;
            for (int j = 0; j < 4; j++) m2[i][j] -= m2[rk][j] * f;
        }
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            if (c[j] == i) row[i] = r[j];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) out[i][j] = m2[row[i]][j];
    return true;
}

}  // namespace host

// ---------------------------------------------------------------------------------------
// scene description (host)
// ---------------------------------------------------------------------------------------
struct SceneHost {
    int n_nodes = 0;
    std::vector<double> rest;               // 3N
    std::vector<int32_t> tets;              // 4E
    std::vector<double> young{5000.0}, poisson{0.45};
    int mass_kind = SOFA_AMD_MASS_UNIFORM;
    bool mass_total = false;                // DiagonalMass: mass_value is the total mass
    double mass_value = 1.0;                // vertexMass or massDensity
    double g[3] = {0, 0, 0};
    std::vector<int32_t> fixed{0};          // FixedProjectiveConstraint default indices = {0}
    bool fix_all = false;                   // FixedProjectiveConstraint fixAll
    bool project_velocity = false;          // FixedProjectiveConstraint activate_projectVelocity
    bool plane_on = false, plane_bilateral = false;   // PlaneForceField
    double plane_n[3] = {0, 1, 0}, plane_d = 0, plane_k = 500, plane_damp = 5, plane_max = 0;
    std::vector<double> density{1.0};       // MeshMatrixMass massDensity (one or per element)
    // OscillatorProjectiveConstraint entries
    std::vector<int32_t> osc_idx;
    std::vector<double> osc_mean, osc_amp, osc_w, osc_phase;
    // LinearMovementProjectiveConstraint: indices, keyTimes, movements (3 per key), relativeMovements
    bool lm_on = false, lm_relative = true;
    int precond = 0;                        // PCGLinearSolver: 0 none (CGLinearSolver), 1 Jacobi, 2 block Jacobi
    bool warm_start = false;                // CGLinearSolver warmStart (FP64 engine)
    std::vector<int32_t> lm_idx;
    std::vector<double> lm_times, lm_moves;
    double dt = 0.01, rs = 0, rm = 0;
    bool first_order = false, trapezoidal = false;   // EulerImplicitSolver firstOrder, trapezoidalScheme
    double vdamping = 0;                    // EulerImplicitSolver vdamping
    int cg_iter = 25;
    double tol = 1e-5, thr = 1e-5;
    // MIXED: inner FP32 CG iterations / relative tolerance per sweep, sweeps per step at most
    int inner_iter = 30;
    double inner_tol = 1e-6;
    int max_sweeps = 3;
    int keep_dir = 0;                       // sweeps continue the previous search direction
    bool released = false;                  // live engines: the scene's data has been released
    // placement after commit
    int node_begin = 0, elem_begin = 0;
    int nblk_begin = 0, nblk_count = 0, eblk_begin = 0, eblk_count = 0;
};

// ---------------------------------------------------------------------------------------
// typed engine
// ---------------------------------------------------------------------------------------
// first-fit allocator of index ranges in [0, cap): the index spaces of a tiled engine (nodes,
// elements, tiles, ...) are arenas, so that scenes can be placed and released while the others
// keep their device data where it is
struct RangeAlloc {
    int64_t cap = 0, used = 0;
    std::map<int64_t, int64_t> free_;   // begin -> length
    void reset(int64_t c) {
        cap = c; used = 0; free_.clear();
        if (c > 0) free_[0] = c;
    }
    int64_t alloc(int64_t n) {          // -1: no room
        if (n <= 0) return 0;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const int64_t b = it->first, len = it->second;
            free_.erase(it);
            if (len > n) free_[b + n] = len - n;
            used += n;
            return b;
        }
        return -1;
    }
    void release(int64_t b, int64_t n) {
        if (n <= 0) return;
        used -= n;
        auto it = free_.emplace(b, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) { it->second += nx->second; free_.erase(nx); }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); }
        }
    }
};

// sizes in the index spaces of a tiled engine (one scene's needs, or an engine's capacity)
struct Extent {
    int64_t node = 0;    // nodes (+1 per scene: its incidence CSR ends inside its own range)
    int64_t elem = 0;    // elements
    int64_t inc = 0;     // incidences (4 per element): CSR entries and tile-local entries
    int64_t tile = 0;    // tiles (and their partial slots)
    int64_t lnode = 0;   // tile-local nodes (+1 per tile)
    int64_t shn = 0;     // nodes finished by the shared-node kernel (+1 per scene)
    int64_t slot = 0;    // shared-node partial slots
    int64_t fin = 0;     // shared-node blocks (and their partial slots)
    int64_t nblk = 0;    // node blocks (and their partial slots)
    int64_t osc = 0;     // oscillator entries
    int64_t sid = 0;     // scene records
    void add(const Extent& o) {
        node += o.node; elem += o.elem; inc += o.inc; tile += o.tile; lnode += o.lnode; shn += o.shn;
        slot += o.slot; fin += o.fin; nblk += o.nblk; osc += o.osc; sid += o.sid;
    }
};

struct EngineBase {
    int device = 0;
    int precision = SOFA_AMD_FP64;
    hipStream_t stream = nullptr;
    std::deque<SceneHost> scenes;     // a deque: references stay valid while scenes are added
    std::mutex scenes_mu;             // guards the container against a concurrent scene_prepare
    bool committed = false;
    bool use_graphs = false;
    virtual ~EngineBase() {
        if (vdot_part) (void)hipFree(vdot_part);
    }
    virtual void commit() = 0;
    virtual void step(const int32_t* ids, int n, int nsteps, hipStream_t s) = 0;
    virtual void read(int scene, double* x, double* v, double* f) = 0;
    virtual void write(int scene, const double* x, const double* v) = 0;
    virtual void stats(int scene, sofa_amd_cg_stats* out) = 0;
    virtual void element(int scene, int e, double* rot9, double* k3, double* rest12) = 0;
    virtual void add_force(int scene, const void* x, void* f, hipStream_t s) = 0;
    virtual void add_dforce(int scene, const void* dx, void* df, double k, hipStream_t s) = 0;
    virtual void bench_apply(int scene, int reps, hipStream_t s) = 0;
    virtual void tiling(int scene, int32_t* out) = 0;
    virtual void reset_graph() = 0;
    // live scenes (tiled engines): capacity, one scene placed / released at a time
    virtual bool live() const = 0;
    virtual void reserve(const Extent& c) = 0;
    virtual void prepare_scene(int scene) = 0;
    virtual bool commit_scene(int scene) = 0;
    virtual void release_scene(int scene) = 0;
    virtual bool placed(int scene) const = 0;
    virtual void usage(int64_t* out22) const = 0;
    bool live_mode = false;           // reserved: scenes are committed and released one by one
    bool seq_dots = false;            // exact path: dot products in the reference's sequential order
    bool stepped = false;             // a step has been enqueued
    int branches = 0;                 // concurrent branches of a merged step (0: the engine's default)
    bool branch_turns = false;        // the branches' operator phases take turns (reference order)
    int dot_segments = 8;             // reference-order dots on FP64_TILED: segments run at once
    int dot_window = 32;              // ... candidate start values per side of each boundary
    double* vdot_part = nullptr;      // vDot's per-block partials (grown on demand)
    int64_t vdot_len = 0;
    virtual int autotune(int reps, hipStream_t s, float* us_out) = 0;
    virtual void set_apply_variant(int v) = 0;
    virtual void von_mises(int scene, int mode, double* per_elem, double* per_node) = 0;
    virtual int64_t stiffness(int scene, double kf, int32_t* row_ptr, int32_t* cols, double* values) = 0;
    // component-level calls of the masses (0 addMDx, 1 accFromF, 2 addForce), the projective
    // constraints (0 projectResponse, 1 projectVelocity, 2 projectPosition at time t), vDot and
    // CGLinearSolver::solve on caller vectors of one scene
    virtual void mass_call(int scene, int what, const void* in, void* out, double factor, hipStream_t s) = 0;
    virtual void project(int scene, int what, void* vec, double t, hipStream_t s) = 0;
    virtual void vdot(int scene, const void* a, const void* b, int order, double* out, double* scratch,
                      int64_t scratch_len, hipStream_t s) = 0;
    virtual void cg_solve(int scene, double mf, double bf, double kf, const void* bvec, void* xvec,
                          sofa_amd_cg_stats* st, hipStream_t s) = 0;
    virtual int scene_nodes(int scene) const = 0;
};

template <typename T>
static T* dalloc(size_t n) {
    T* p = nullptr;
    if (n == 0) n = 1;
    HIPCHK(hipMalloc((void**)&p, n * sizeof(T)));
    return p;
}

// the inner FP32 CG of a MIXED engine: its kernels are the FP32 engine's (compiled in the
// translation unit that allows fused multiply-adds), on float copies of the element record
struct InnerCg {
    TileArgs<float, 6> ta;          // inst = the inner scene records
    const FinBlk* fb;
    const NodeBlk* nb;
    int na, nfb, nnb, maxit;
    bool wide;                      // some scene has more than kWavePartials partial sums
    bool plane;                     // some scene has a PlaneForceField
    SharedNodes sn;
    float *r, *p, *q, *d, *sp;
    float* z;                       // with a preconditioner: z = P^-1 r, staged by the operator
    const float *dinv, *binv;       // ... the inverted diagonal (3 per node) / 3x3 blocks (9)
    const float *x, *x0;            // float positions and rest positions (geometry recomputed)
    double *pden, *prho;            // prho: |r|^2 partials per node block (iteration 1 included)
};
void launch_inner_cg(const InnerCg& c, hipStream_t s);

template <typename R, int NR, typename VT>
struct EngineT : EngineBase {
    bool tiled = false;
    bool mixed = false;             // FP64 engine whose CG is FP32 inside FP64 refinement
    int64_t Ntot = 0, Etot = 0;     // node / element capacity (the SoA stride of element arrays)
    // node vectors (AoS xyz)
    VT *x = nullptr, *v = nullptr, *f = nullptr, *b = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *xcg = nullptr;
    VT* x0 = nullptr;                              // rest positions (tiled path)
    double* mass = nullptr;
    int8_t* fixed = nullptr;
    int8_t* contact = nullptr;                     // PlaneForceField contacts per node
    int *inc_ptr = nullptr, *inc = nullptr;        // node -> (position*4 + vertex), ascending element id
    // element record (SoA, in tile order on the tiled path, else in element order)
    int4* tet = nullptr;
    R *rot = nullptr, *rest = nullptr, *kmat = nullptr;
    VT *ef = nullptr, *ef2 = nullptr;
    std::vector<std::vector<int>> iperm;           // per scene: original element -> position
    VT* p2 = nullptr;                              // previous CG direction (exact path)
    double* nd = nullptr;                          // per-node dot terms (exact path, reference order)
    double* spec_part = nullptr;                   // segmented reference-order dots: prediction chunks
    double* spec_model = nullptr;                  // ... with the chain's roundings modelled
    double* spec_seg = nullptr;                    // ... and the segments' candidate sums
    double* pinv = nullptr;                        // PCG: inverted preconditioner block per node (9)
    VT* zv = nullptr;                              // PCG: preconditioned residual z = P^-1 r
    int *em_ptr = nullptr, *em_nbr = nullptr;      // MeshMatrixMass edges per node (exact path)
    double* em_val = nullptr;
    R* emass = nullptr;                            // per element rho V / 20 (tiled path)
    int* osc_node = nullptr;
    double *osc_mean = nullptr, *osc_amp = nullptr, *osc_w = nullptr, *osc_phase = nullptr;
    // scene records
    InstDev* inst = nullptr;
    // exact path: block lists of every scene, reduction partials
    std::vector<Blk> eblk_h, nblk_h;
    double* partial = nullptr;
    // tiled path: device tables, partial buffers
    TileDesc* tiles_d = nullptr;
    Ranges* ranges_d = nullptr;
    uint2* ltet = nullptr;
    int *ln_gid = nullptr, *ln_inc_ptr = nullptr, *ln_pslot = nullptr, *sh_gid = nullptr, *sh_ptr = nullptr;
    int* sh_of = nullptr;                          // per node: index into sh_gid / sh_ptr, or -1
    uint8_t* inc_rank = nullptr;                   // reference order: incidence ranks of shared nodes
    uint8_t* inc_arank = nullptr;                  // reference order, CG operator: slot offsets (0: folded)
    uint8_t* sh_acnt = nullptr;                    // reference order, CG operator: slots after the first
    uint8_t* ln_lead = nullptr;                    // reference order, CG operator: folded incidences per local node
    float* dinv32 = nullptr;                       // MIXED: the inner CG's Jacobi preconditioner (3 per node)
    float* binv32 = nullptr;                       // ... its block-Jacobi preconditioner (9 per node)
    float* z32 = nullptr;                          // ... z = P^-1 r of its preconditioned CG
    uint16_t* ln_inc = nullptr;
    VT *sp0 = nullptr, *sp1 = nullptr;
    double *pbb = nullptr, *pden = nullptr, *prho = nullptr;
    // MIXED: float element record and inner CG vectors, inner scene records, |r|^2 partials
    float *kmat32 = nullptr, *emass32 = nullptr;   // MIXED: float K entries and edge masses
    float *x32 = nullptr, *x0_32 = nullptr;   // MIXED: float positions (per step) and rest positions
    float *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *d32 = nullptr, *sp32 = nullptr;
    double *pden32 = nullptr, *prho32 = nullptr, *prr = nullptr;
    InstDev* inst_in = nullptr;
    // placement of each scene: where its ranges begin in every index space, its block lists
    struct Place {
        Extent at, size;
        bool placed = false;
        std::vector<TileDesc> tiles;
        std::vector<FinBlk> fins;
        std::vector<NodeBlk> nblks;
        std::vector<Blk> oblks;
        std::vector<Blk> lblks;          // LinearMovementProjectiveConstraint entries
    };
    std::vector<Place> place;
    // per scene (by scene index): the LinearMovementProjectiveConstraint's device data, one
    // allocation: node ids (int), then x0 (3 doubles each) and the keys (4 doubles each), then the
    // base masks (int8)
    std::map<int, void*> lmd;
    Extent cap;
    RangeAlloc a_node, a_elem, a_inc, a_tile, a_lnode, a_shn, a_slot, a_fin, a_nblk, a_osc, a_sid;
    // launch sets: the block lists (a: element blocks / tiles, b: node blocks / finalize blocks,
    // c: tiled node blocks) and the captured step graph of each set of scenes stepped together
    struct LaunchSet {
        void *a = nullptr, *b = nullptr, *c = nullptr, *o = nullptr, *l = nullptr;
        int* scenes = nullptr;          // scene record indices
        int na = 0, nb = 0, nc = 0, no = 0, nl = 0, maxit = 0, maxsw = 0, n = 0;
        int maxpart = 0;                // most partial sums the operator reduces for one scene
        bool plane = false;             // some scene has a PlaneForceField
        bool pcg = false;               // some scene solves with PCGLinearSolver
        bool warm = false;              // some scene's CGLinearSolver has warmStart
        hipGraphExec_t gexec = nullptr;
        ~LaunchSet() {
            if (gexec) (void)hipGraphExecDestroy(gexec);
            for (void* ptr : {a, b, c, o, l, (void*)scenes})
                if (ptr) (void)hipFree(ptr);
        }
    };
    std::map<std::vector<int32_t>, std::unique_ptr<LaunchSet>> sets;
    LaunchSet* cur = nullptr;
    uint64_t sets_dropped = 0;          // times prepare_set emptied the cache

    ~EngineT() override {   // the device is synchronized before an engine is destroyed
        sets.clear();
        for (hipStream_t a : aux)
            if (a) (void)hipStreamDestroy(a);
        for (hipEvent_t e : join_ev)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : turn_ev)
            if (e) (void)hipEventDestroy(e);
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        while (!vmd.empty()) free_vm(vmd.begin()->first);
        while (!kd.empty()) free_k(kd.begin()->first);
        while (!lmd.empty()) free_lm(lmd.begin()->first);
        free_device();
    }
    void free_device() {
        void* ptrs[] = {x, v, f, b, r, p, q, xcg, mass, fixed, inc_ptr, inc, tet, rot, rest, kmat, ef, ef2,
                        inst, partial, tiles_d, ranges_d, ltet, ln_gid, ln_inc_ptr, ln_pslot, sh_gid, sh_ptr,
                        ln_inc, sp0, sp1, pbb, pden, prho, p2, em_ptr, em_nbr, em_val, emass, osc_node,
                        osc_mean, osc_amp, osc_w, osc_phase, kmat32, emass32, r32, p32, q32,
                        d32, sp32, pden32, prho32, prr, inst_in, sh_of, nd, x0, x32, x0_32, contact, inc_rank, pinv, zv,
                        inc_arank, sh_acnt, ln_lead, dinv32, binv32, z32, spec_part, spec_model, spec_seg};
        for (void* ptr : ptrs)
            if (ptr) (void)hipFree(ptr);
    }

    bool live() const override { return tiled; }

    ElemView<R, NR> view() const { return ElemView<R, NR>{tet, rot, rest, kmat, Etot}; }
    NodeArrays nodes() const { return NodeArrays{inc_ptr, inc, mass, fixed, em_ptr, em_nbr, em_val, contact}; }
    Oscillators oscillators() const { return Oscillators{osc_node, osc_mean, osc_amp, osc_w, osc_phase}; }
    int node_begin(int scene) const { return (int)place[scene].at.node; }
    int elem_begin(int scene) const { return (int)place[scene].at.elem; }
    int sid_of(int scene) const { return (int)place[scene].at.sid; }

    // MeshMatrixMass edges (consistent mass): TetrahedronSetTopologyContainer::createEdgeSetArray
    // order (local edges {0,1},{0,2},{0,3},{1,2},{1,3},{2,3}, first appearance, endpoints
    // sorted), edge mass (rho V) / 20 summed in tetrahedron order (MeshMatrixMass.inl:603-650);
    // each node lists its edges in ascending edge id with the other end (global ids)
    void build_edges(const SceneHost& s, int nE, const std::vector<host::V3>& P, std::vector<int>& emp,
                     std::vector<int>& emn, std::vector<double>& emv) {
        std::vector<std::vector<std::pair<int, double>>> per(s.n_nodes);
        if (s.mass_kind == SOFA_AMD_MASS_MESHMATRIX) {
            static const int le[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
            std::map<std::pair<int, int>, int> emap;
            std::vector<std::array<int, 2>> edges;
            std::vector<double> emass_e;
            for (int e = 0; e < nE; e++) {
                const int32_t* t = &s.tets[4 * (size_t)e];
                const double rho = s.density.size() > (size_t)e ? s.density[e] : s.density[0];
                const double m = (rho * host::volume(P[t[0]], P[t[1]], P[t[2]], P[t[3]])) / 20.0;
                int ids[6];
                for (int j = 0; j < 6; j++) {
                    int v1 = t[le[j][0]], v2 = t[le[j][1]];
                    std::pair<int, int> key = v1 < v2 ? std::make_pair(v1, v2) : std::make_pair(v2, v1);
                    auto it = emap.find(key);
                    if (it == emap.end()) {
                        ids[j] = (int)edges.size();
                        emap[key] = ids[j];
                        edges.push_back({key.first, key.second});
                        emass_e.push_back(0.0);
                    } else {
                        ids[j] = it->second;
                    }
                }
                for (int j = 0; j < 6; j++) emass_e[ids[j]] += m;
            }
            for (size_t j = 0; j < edges.size(); j++) {
                per[edges[j][0]].push_back({edges[j][1], emass_e[j]});
                per[edges[j][1]].push_back({edges[j][0], emass_e[j]});
            }
        }
        for (int i = 0; i < s.n_nodes; i++) {
            for (auto& pr : per[i]) {
                emn.push_back(pr.first + s.node_begin);
                emv.push_back(pr.second);
            }
            emp.push_back((int)emn.size());
        }
    }

    template <typename T>
    void upload(T*& dst, const std::vector<T>& src) {
        dst = dalloc<T>(src.size());
        if (!src.empty()) HIPCHK(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    template <typename T>
    static void put_at(T* dst, int64_t at, const T* src, size_t n) {
        if (n) HIPCHK(hipMemcpy(dst + at, src, n * sizeof(T), hipMemcpyHostToDevice));
    }
    template <typename T>
    static void put_at(T* dst, int64_t at, const std::vector<T>& src) { put_at(dst, at, src.data(), src.size()); }

    // ---- one scene's element record, masses and incidence (host, scene-local ids) ----------
    struct ElemHost {
        int nE = 0;
        std::vector<int4> tet;                       // placement order, scene-local node ids
        std::vector<R> rot, rest, kmat, emass;       // SoA with stride nE
        std::vector<float> k_lo, em_lo;     // MIXED
        std::vector<double> mass;                    // per node (DiagonalMass / MeshMatrixMass)
        std::vector<int> inc_ptr, inc;               // n + 1, 4 nE: position * 4 + vertex
        std::vector<int8_t> fixed;
        std::vector<host::V3> P;
    };
    // TetrahedronFEMForceField::reinit per element (rotation, rest frame, K), the masses' vertex
    // terms, FixedProjectiveConstraint / oscillator masks, the incidence in ascending element id
    void setup_scene(const SceneHost& s, int si, const std::vector<int>& ip, ElemHost& h) const {
        const int nE = (int)(s.tets.size() / 4), n = s.n_nodes;
        h.nE = nE;
        h.P.resize(n);
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) h.P[i].v[k] = s.rest[3 * i + k];
        h.tet.resize(nE);
        h.rot.assign(9 * (size_t)nE, (R)0); h.rest.assign(NR * (size_t)nE, (R)0); h.kmat.assign(3 * (size_t)nE, (R)0);
        h.emass.assign(nE, (R)0);
        if (mixed) { h.k_lo.assign(3 * (size_t)nE, 0.f); h.em_lo.assign(nE, 0.f); }
        h.mass.assign(n, 0.0);
        std::vector<int> cnt(n + 1, 0);
        double mass_sum = 0.0;                       // DiagonalMass computeVertexMass total
        for (int e = 0; e < nE; e++) {
            const int32_t* t = &s.tets[4 * (size_t)e];
            for (int k = 0; k < 4; k++)
                if (t[k] < 0 || t[k] >= n) throw std::invalid_argument("tetrahedron index out of range");
            const double E = s.young.size() > (size_t)e ? s.young[e] : s.young[0];
            const double nu = s.poisson.size() > (size_t)e ? s.poisson[e] : s.poisson[0];
            host::ElemInit ei;
            if (!host::init_element(h.P, t, E, nu, ei))
                throw std::domain_error("degenerate tetrahedron " + std::to_string(e) + " in scene " + std::to_string(si));
            const int64_t pe = ip[e];
            h.tet[pe] = make_int4(t[0], t[1], t[2], t[3]);
            for (int k = 0; k < 9; k++) h.rot[k * (size_t)nE + pe] = (R)ei.rot[k];
            const int pick[6] = {0, 3, 4, 6, 7, 8};   // b0, c0, c1, d0, d1, d2
            if (NR == 9) {
                for (int k = 0; k < 9; k++) h.rest[k * (size_t)nE + pe] = (R)ei.rest[k];
            } else {
                for (int k = 0; k < 6; k++) h.rest[k * (size_t)nE + pe] = (R)ei.rest[pick[k]];
            }
            for (int k = 0; k < 3; k++) h.kmat[k * (size_t)nE + pe] = (R)ei.k[k];
            if (mixed)
                for (int k = 0; k < 3; k++) h.k_lo[k * (size_t)nE + pe] = (float)ei.k[k];
            for (int k = 0; k < 4; k++) cnt[t[k] + 1]++;
            if (s.mass_kind == SOFA_AMD_MASS_DIAGONAL) {
                // DiagonalMass.inl:1075-1080; given the total mass, first with density 1
                const double m = ((s.mass_total ? 1.0 : s.mass_value) * ei.vol) / 4.0;
                for (int k = 0; k < 4; k++) {
                    h.mass[t[k]] += m;
                    mass_sum += m;
                }
            } else if (s.mass_kind == SOFA_AMD_MASS_MESHMATRIX || s.mass_kind == SOFA_AMD_MASS_MESHMATRIX_LUMPED) {
                const double rho = s.density.size() > (size_t)e ? s.density[e] : s.density[0];
                const double m = (rho * ei.vol) / 10.0;              // MeshMatrixMass.inl:584
                for (int k = 0; k < 4; k++) h.mass[t[k]] += m;
                if (s.mass_kind == SOFA_AMD_MASS_MESHMATRIX) {
                    h.emass[pe] = (R)((rho * ei.vol) / 20.0);
                    if (mixed) h.em_lo[pe] = (float)((rho * ei.vol) / 20.0);
                }
            }
        }
        // DiagonalMass::initFromTotalMass (DiagonalMass.inl:1235-1257): density = total / sum
        // (1 when the sum is below epsilon), every vertex mass times it
        if (s.mass_kind == SOFA_AMD_MASS_DIAGONAL && s.mass_total) {
            const double density = mass_sum < std::numeric_limits<double>::epsilon() ? 1.0 : s.mass_value / mass_sum;
            for (int i = 0; i < n; i++) h.mass[i] *= density;
        }
        for (int i = 0; i < n; i++) cnt[i + 1] += cnt[i];
        h.inc_ptr = cnt;
        h.inc.assign(4 * (size_t)nE, 0);
        std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
        for (int e = 0; e < nE; e++)
            for (int k = 0; k < 4; k++) h.inc[fillp[s.tets[4 * (size_t)e + k]]++] = 4 * ip[e] + k;
        h.fixed.assign(n, s.fix_all ? 1 : 0);
        for (int32_t i : s.fixed) {
            if (i < 0 || i >= n) throw std::invalid_argument("fixed index out of range");
            h.fixed[i] = 1;
        }
        // oscillators: projectResponse clears their nodes like a fixed constraint (mask 2: not
        // a fixed node for projectVelocity)
        for (int32_t i : s.osc_idx) {
            if (i < 0 || i >= n) throw std::invalid_argument("oscillator index out of range");
            if (!h.fixed[i]) h.fixed[i] = 2;
        }
    }

    // one E and one nu for every element (BaseLinearElasticityFEMForceField.inl:113-144 reads
    // entry 0 for elements beyond the vectors' length)
    static bool one_material(const SceneHost& s) {
        for (double y : s.young) if (y != s.young[0]) return false;
        for (double p : s.poisson) if (p != s.poisson[0]) return false;
        return true;
    }

    // the scene record: EulerImplicitSolver / CGLinearSolver parameters and the factors of the
    // mechanical operations (EulerImplicitSolver.cpp:147-171)
    InstDev make_inst(int si) const {
        const SceneHost& s = scenes[si];
        InstDev in;
        std::memset(&in, 0, sizeof(in));
        in.node_count = s.n_nodes;
        in.elem_count = (int)(s.tets.size() / 4);
        in.mass_kind = s.mass_kind;
        in.max_iter = s.cg_iter;
        in.dt = s.dt; in.rm = s.rm; in.rs = s.rs;
        in.vertex_mass = s.mass_kind == SOFA_AMD_MASS_UNIFORM ? s.mass_value : 0.0;
        for (int k = 0; k < 3; k++) {
            in.gravity[k] = s.g[k];
            // UniformMass::addForce: mg = theGravity * m, in the state's precision
            in.mg[k] = (double)((VT)s.g[k] * (VT)in.vertex_mass);
        }
        // EulerImplicitSolver.cpp:118-123,136-173: trapezoidal tr = 1/2; firstOrder: b = f, A = M - h tr K
        const double h = s.dt, tr = s.trapezoidal ? 0.5 : 1.0;
        if (s.first_order) {
            in.k_rhs = 0.0;
            in.m_rhs = 0.0;
            in.m_cg = 1.0;
            in.k_cg = -h * tr;
            in.b_scale = 1.0;
        } else {
            in.k_rhs = h * tr + s.rs;
            in.m_rhs = -s.rm;
            in.m_cg = 1 + tr * h * s.rm;
            in.k_cg = -tr * h * (tr * h + s.rs);
            in.b_scale = h;
        }
        in.first_order = s.first_order ? 1 : 0;
        in.project_velocity = s.project_velocity ? 1 : 0;
        in.plane_on = s.plane_on ? 1 : 0;
        in.plane_bilateral = s.plane_bilateral ? 1 : 0;
        for (int k = 0; k < 3; k++) in.plane_n[k] = s.plane_n[k];
        in.plane_d = s.plane_d; in.plane_k = s.plane_k; in.plane_damp = s.plane_damp; in.plane_max = s.plane_max;
        in.vdamp = s.vdamping != 0.0 ? std::exp(-h * s.vdamping) : 1.0;
        in.tol = s.tol < 0 ? 1e-5 : s.tol;
        in.thr = s.thr < 0 ? 1e-5 : s.thr;
        in.first_solve = 1;
        in.stopped = tiled ? 0 : 1;
        in.max_sweeps = s.max_sweeps;
        in.pcg = s.precond;
        in.warm_start = s.warm_start ? 1 : 0;
        if (one_material(s)) host::material_factors(s.young[0], s.poisson[0], in.ks);
        return in;
    }
    // the inner FP32 CG's record: the same operator, FP32 CG limits, no threshold test (the
    // inner right-hand side shrinks sweep by sweep), closed until a sweep opens it
    InstDev make_inner(int si, const InstDev& o) const {
        InstDev in = o;
        in.max_iter = scenes[si].inner_iter;
        in.tol = scenes[si].inner_tol;
        in.thr = 0.0;
        in.first_solve = 0;
        in.stopped = 1;
        in.mg[0] = in.mg[1] = in.mg[2] = 0.0;
        in.keep_dir = scenes[si].keep_dir;
        in.sweep = 0;
        in.rho_last = 0.0;
        return in;
    }

    // the scene's LinearMovementProjectiveConstraint on the device: its entries, x0 and keys in
    // one allocation, the record's pointers and counts set in `in`, the entry blocks returned;
    // `mask` is the scene's response mask without the movement (0 free, 1 fixed, 2 oscillator)
    std::vector<Blk> place_lm(int si, int rec, int nb0, const std::vector<int8_t>& mask, InstDev& in) {
        std::vector<Blk> blocks;
        const SceneHost& s = scenes[si];
        if (!s.lm_on) return blocks;
        free_lm(si);
        const int n = (int)s.lm_idx.size(), nk = (int)s.lm_times.size();
        const size_t b_node = 0, b_x0 = ((size_t)n * 4 + 15) / 16 * 16, b_keys = b_x0 + (size_t)n * 24;
        const size_t b_base = b_keys + (size_t)nk * 32, bytes = b_base + (size_t)n + 16;
        char* buf = dalloc<char>(bytes);
        lmd[si] = buf;
        std::vector<int32_t> node(n);
        std::vector<int8_t> base(n);
        for (int j = 0; j < n; j++) {
            node[j] = nb0 + s.lm_idx[j];
            base[j] = mask[s.lm_idx[j]];
        }
        std::vector<double> keys(s.lm_times);
        keys.insert(keys.end(), s.lm_moves.begin(), s.lm_moves.end());
        if (n) HIPCHK(hipMemcpy(buf + b_node, node.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(buf + b_x0, 0, (size_t)n * 24));
        HIPCHK(hipMemcpy(buf + b_keys, keys.data(), keys.size() * sizeof(double), hipMemcpyHostToDevice));
        if (n) HIPCHK(hipMemcpy(buf + b_base, base.data(), n, hipMemcpyHostToDevice));
        in.lm_node = (const int*)(buf + b_node);
        in.lm_x0 = (double*)(buf + b_x0);
        in.lm_keys = (const double*)(buf + b_keys);
        in.lm_base = (const int8_t*)(buf + b_base);
        in.lm_n = n;
        in.lm_nk = nk;
        in.lm_relative = s.lm_relative ? 1 : 0;
        in.lm_x0_set = 0;
        for (int j0 = 0; j0 < n; j0 += kBlock) blocks.push_back(Blk{rec, j0, std::min(n, j0 + kBlock), 0});
        return blocks;
    }
    // PCG buffers (per node: the inverted preconditioner block and z), once some scene asks for
    // PCGLinearSolver; the solution vector starts at 0 for a scene's first solve
    void alloc_pcg() {
        if (pinv) return;
        bool any = false;
        for (const SceneHost& sc : scenes) any = any || sc.precond != 0;
        if (!any) return;
        const int64_t n = tiled ? cap.node : Ntot;
        pinv = dalloc<double>(9 * (size_t)n);
        zv = dalloc<VT>(3 * (size_t)n);
        HIPCHK(hipMemset(zv, 0, 3 * (size_t)n * sizeof(VT)));
    }
    void free_lm(int si) {
        auto it = lmd.find(si);
        if (it == lmd.end()) return;
        (void)hipFree(it->second);
        lmd.erase(it);
    }

    void commit() override {
        HIPCHK(hipSetDevice(device));
        if (tiled) commit_tiled();
        else commit_exact();
        committed = true;
    }

    // ---- exact path: every scene placed once, in order, into exactly sized buffers ---------
    void commit_exact() {
        const int S = (int)scenes.size();
        Ntot = 0; Etot = 0;
        place.assign(S, Place());
        for (int si = 0; si < S; si++) {
            place[si].at.node = Ntot;
            place[si].at.elem = Etot;
            place[si].at.sid = si;
            place[si].placed = true;
            scenes[si].node_begin = (int)Ntot;
            scenes[si].elem_begin = (int)Etot;
            Ntot += scenes[si].n_nodes;
            Etot += (int64_t)scenes[si].tets.size() / 4;
        }
        if (Ntot >= (int64_t)1 << 30 || Etot * 4 >= (int64_t)1 << 31)
            throw std::invalid_argument("mesh too large for 32-bit indices");
        std::vector<InstDev> inst_h(S);
        eblk_h.clear(); nblk_h.clear();
        iperm.assign(S, {});
        std::vector<double> xh(3 * Ntot), massh(Ntot, 0.0);
        std::vector<int8_t> fixh(Ntot, 0);
        std::vector<int4> teth(Etot);
        std::vector<R> roth(9 * Etot), resth(NR * Etot), kh(3 * Etot);
        std::vector<int> cnt{0}, inch(4 * Etot);
        std::vector<int> emp{0}, emn;
        std::vector<double> emv;
        std::vector<int> oscn;
        std::vector<double> oscm, osca, oscw, oscp;
        for (int si = 0; si < S; si++) {
            SceneHost& s = scenes[si];
            const int nE = (int)(s.tets.size() / 4);
            const int nb0 = s.node_begin;
            const int64_t eb0 = s.elem_begin;
            std::vector<int>& ip = iperm[si];
            ip.resize(nE);
            for (int e = 0; e < nE; e++) ip[e] = e;
            ElemHost h;
            setup_scene(s, si, ip, h);
            for (int i = 0; i < s.n_nodes; i++) {
                for (int k = 0; k < 3; k++) xh[3 * (nb0 + (int64_t)i) + k] = s.rest[3 * i + k];
                massh[nb0 + i] = h.mass[i];
                fixh[nb0 + i] = h.fixed[i];
            }
            for (int e = 0; e < nE; e++) {
                const int4 t = h.tet[e];
                teth[eb0 + e] = make_int4(t.x + nb0, t.y + nb0, t.z + nb0, t.w + nb0);
                for (int k = 0; k < 9; k++) roth[k * Etot + eb0 + e] = h.rot[k * (size_t)nE + e];
                for (int k = 0; k < NR; k++) resth[k * Etot + eb0 + e] = h.rest[k * (size_t)nE + e];
                for (int k = 0; k < 3; k++) kh[k * Etot + eb0 + e] = h.kmat[k * (size_t)nE + e];
            }
            const int ibase = cnt.back();
            for (int i = 0; i < s.n_nodes; i++) cnt.push_back(ibase + h.inc_ptr[i + 1]);
            for (int j = 0; j < 4 * nE; j++) inch[ibase + j] = h.inc[j] + (int)(4 * eb0);
            build_edges(s, nE, h.P, emp, emn, emv);
            const int ob = (int)oscn.size();
            for (size_t j = 0; j < s.osc_idx.size(); j++) {
                oscn.push_back(nb0 + s.osc_idx[j]);
                for (int k = 0; k < 3; k++) { oscm.push_back(s.osc_mean[3 * j + k]); osca.push_back(s.osc_amp[3 * j + k]); }
                oscw.push_back(s.osc_w[j]);
                oscp.push_back(s.osc_phase[j]);
            }
            for (int j0 = 0; j0 < (int)s.osc_idx.size(); j0 += kBlock)
                place[si].oblks.push_back(Blk{si, ob + j0, ob + std::min((int)s.osc_idx.size(), j0 + kBlock), 0});
            // blocks: 256 elements / 256 nodes
            s.eblk_begin = (int)eblk_h.size();
            for (int e0 = 0; e0 < nE; e0 += kBlock)
                eblk_h.push_back(Blk{si, (int)eb0 + e0, (int)eb0 + std::min(nE, e0 + kBlock), 0});
            s.eblk_count = (int)eblk_h.size() - s.eblk_begin;
            s.nblk_begin = (int)nblk_h.size();
            for (int n0 = 0; n0 < s.n_nodes; n0 += kBlock)
                nblk_h.push_back(Blk{si, nb0 + n0, nb0 + std::min(s.n_nodes, n0 + kBlock), (int)nblk_h.size()});
            s.nblk_count = (int)nblk_h.size() - s.nblk_begin;
            InstDev& in = inst_h[si];
            in = make_inst(si);
            in.node_begin = nb0; in.elem_begin = (int)eb0;
            in.nblk_begin = s.nblk_begin; in.nblk_count = s.nblk_count;
            place[si].lblks = place_lm(si, si, nb0, h.fixed, in);
        }
        x = dalloc<VT>(3 * Ntot); v = dalloc<VT>(3 * Ntot); f = dalloc<VT>(3 * Ntot); b = dalloc<VT>(3 * Ntot);
        r = dalloc<VT>(3 * Ntot); p = dalloc<VT>(3 * Ntot); q = dalloc<VT>(3 * Ntot); xcg = dalloc<VT>(3 * Ntot);
        std::vector<VT> xv(3 * Ntot);
        for (int64_t i = 0; i < 3 * Ntot; i++) xv[i] = (VT)xh[i];
        HIPCHK(hipMemcpy(x, xv.data(), xv.size() * sizeof(VT), hipMemcpyHostToDevice));
        for (VT* vec : {v, f, b, r, p, q, xcg}) HIPCHK(hipMemset(vec, 0, 3 * Ntot * sizeof(VT)));
        contact = dalloc<int8_t>(Ntot);
        HIPCHK(hipMemset(contact, 0, std::max<int64_t>(1, Ntot)));
        upload(mass, massh); upload(fixed, fixh);
        upload(inc_ptr, cnt); upload(inc, inch);
        upload(tet, teth); upload(rot, roth); upload(rest, resth); upload(kmat, kh);
        upload(em_ptr, emp); upload(em_nbr, emn); upload(em_val, emv);
        upload(osc_node, oscn); upload(osc_mean, oscm); upload(osc_amp, osca); upload(osc_w, oscw); upload(osc_phase, oscp);
        ef = dalloc<VT>(12 * Etot);
        ef2 = dalloc<VT>(12 * Etot);
        p2 = dalloc<VT>(3 * Ntot);
        HIPCHK(hipMemset(p2, 0, 3 * Ntot * sizeof(VT)));
        nd = dalloc<double>(Ntot);
        alloc_pcg();
        upload(inst, inst_h);
        partial = dalloc<double>(nblk_h.size());
        HIPCHK(hipMemset(partial, 0, std::max<size_t>(1, nblk_h.size()) * sizeof(double)));
        HIPCHK(hipDeviceSynchronize());
    }

    // ---- tiled path: arenas, scenes placed one at a time ------------------------------------
    struct TiledHost {
        TilePlan plan;
        ElemHost eh;
        std::vector<int> iperm;            // original element -> position
        std::vector<int> sh_gid, sh_ptr;   // nodes finished by the shared-node kernel
        Extent need;
    };
    // host tables of one scene (tiling, element set-up); touches no engine state, so it may run
    // on another host thread than the one stepping the engine
    void build_tiled(const SceneHost& s, int si, TiledHost& th) const {
        const int nE = (int)(s.tets.size() / 4), n = s.n_nodes;
        if (exact_tiled()) {
            // the consistent mass's edge terms ride on the element contributions in the tiles,
            // which is not the reference's order (its edge loop runs before the FEM)
            if (s.mass_kind == SOFA_AMD_MASS_MESHMATRIX)
                throw std::invalid_argument("the reference summation order on a tiled engine needs a lumped mass "
                                            "(UniformMass, DiagonalMass, lumped MeshMatrixMass)");
        }
        th.plan = plan_best(n, s.rest.data(), s.tets.data(), nE, kTileElems, kTileNodes, exact_tiled(), (int)sizeof(VT));
        th.iperm = th.plan.iperm;
        setup_scene(s, si, th.iperm, th.eh);
        th.sh_gid = th.plan.sh_gid;
        th.sh_ptr = th.plan.sh_ptr;
        if (th.sh_ptr.empty()) th.sh_ptr.push_back(0);
        Extent& e = th.need;
        const int64_t T = (int64_t)th.plan.t_ebegin.size();
        const int64_t Sn = (int64_t)th.sh_gid.size();
        e.node = n + 1;
        e.elem = nE;
        e.inc = 4 * (int64_t)nE;
        e.tile = T;
        e.lnode = (int64_t)th.plan.ln_gid.size() + T;
        e.shn = Sn + 1;
        e.slot = th.sh_ptr.back();
        e.fin = (Sn + kBlock - 1) / kBlock;
        e.nblk = (n + kBlock * kNodesPerThread - 1) / (kBlock * kNodesPerThread);
        e.osc = (int64_t)s.osc_idx.size();
        e.sid = 1;
    }

    RangeAlloc* arenas(int k) {
        RangeAlloc* a[] = {&a_node, &a_elem, &a_inc, &a_tile, &a_lnode, &a_shn, &a_slot, &a_fin, &a_nblk, &a_osc, &a_sid};
        return a[k];
    }
    static int64_t& field(Extent& e, int k) {
        int64_t* f[] = {&e.node, &e.elem, &e.inc, &e.tile, &e.lnode, &e.shn, &e.slot, &e.fin, &e.nblk, &e.osc, &e.sid};
        return *f[k];
    }

    // device buffers of capacity `c` in every index space (all scenes' data is dropped)
    void reserve(const Extent& c) override {
        if (!tiled) throw std::invalid_argument("live scenes need a tiled engine (FP32, FP64_TILED or MIXED)");
        {   // every index space is addressed with 32-bit indices on the device
            Extent t = c;
            for (int k = 0; k < 11; k++)
                if (field(t, k) >= INT32_MAX) throw std::invalid_argument("capacity beyond the 32-bit index spaces");
        }
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipDeviceSynchronize());
        sets.clear();
        cur = nullptr;
        free_device();
        x = v = f = b = r = p = q = xcg = x0 = nullptr; mass = nullptr; fixed = nullptr; contact = nullptr; inc_ptr = inc = nullptr;
        tet = nullptr; rot = rest = kmat = nullptr; ef = ef2 = nullptr; p2 = nullptr; em_ptr = em_nbr = nullptr;
        em_val = nullptr; emass = nullptr; osc_node = nullptr; osc_mean = osc_amp = osc_w = osc_phase = nullptr;
        inst = nullptr; partial = nullptr; tiles_d = nullptr; ranges_d = nullptr; ltet = nullptr;
        ln_gid = ln_inc_ptr = ln_pslot = sh_gid = sh_ptr = nullptr; ln_inc = nullptr; sp0 = sp1 = nullptr;
        pbb = pden = prho = nullptr; kmat32 = emass32 = nullptr; x32 = x0_32 = nullptr;
        r32 = p32 = q32 = d32 = sp32 = nullptr; pden32 = prho32 = prr = nullptr; inst_in = nullptr; sh_of = nullptr; inc_rank = nullptr;
        pinv = nullptr; zv = nullptr; inc_arank = nullptr; sh_acnt = nullptr; ln_lead = nullptr; dinv32 = nullptr;
        binv32 = z32 = nullptr;
        spec_part = spec_model = spec_seg = nullptr;
        cap = c;
        if (cap.node >= (int64_t)1 << 30 || cap.inc >= (int64_t)1 << 31 || cap.lnode >= (int64_t)1 << 31 ||
            cap.slot >= (int64_t)kSlotOwner)
            throw std::invalid_argument("capacity too large for 32-bit indices");
        for (int k = 0; k < 11; k++) arenas(k)->reset(field(cap, k));
        Ntot = cap.node; Etot = cap.elem;
        const size_t N = std::max<int64_t>(1, cap.node), E = std::max<int64_t>(1, cap.elem);
        for (VT** vec : {&x, &v, &f, &b, &r, &p, &q, &xcg, &x0}) {
            *vec = dalloc<VT>(3 * N);
            HIPCHK(hipMemset(*vec, 0, 3 * N * sizeof(VT)));
        }
        mass = dalloc<double>(N); fixed = dalloc<int8_t>(N); inc_ptr = dalloc<int>(N); sh_of = dalloc<int>(N);
        nd = dalloc<double>(N);
        HIPCHK(hipMemset(nd, 0, N * sizeof(double)));
        inc_rank = dalloc<uint8_t>(std::max<int64_t>(1, cap.inc));
        inc_arank = dalloc<uint8_t>(std::max<int64_t>(1, cap.inc));
        sh_acnt = dalloc<uint8_t>(std::max<int64_t>(1, cap.shn));
        ln_lead = dalloc<uint8_t>(std::max<int64_t>(1, cap.lnode));
        if (mixed) {
            dinv32 = dalloc<float>(3 * N);
            binv32 = dalloc<float>(9 * N);
            z32 = dalloc<float>(3 * N);
        }
        contact = dalloc<int8_t>(N);
        HIPCHK(hipMemset(contact, 0, std::max<int64_t>(1, N)));
        tet = dalloc<int4>(E); rot = dalloc<R>(9 * E); rest = dalloc<R>(NR * E); kmat = dalloc<R>(3 * E);
        emass = dalloc<R>(E); ltet = dalloc<uint2>(E); ef = dalloc<VT>(12 * E);
        inc = dalloc<int>(cap.inc); ln_inc = dalloc<uint16_t>(cap.inc);
        tiles_d = dalloc<TileDesc>(cap.tile);
        const size_t np = cap.tile + cap.fin;
        pbb = dalloc<double>(np); pden = dalloc<double>(np);
        ln_gid = dalloc<int>(cap.lnode); ln_inc_ptr = dalloc<int>(cap.lnode); ln_pslot = dalloc<int>(cap.lnode);
        sh_gid = dalloc<int>(cap.shn); sh_ptr = dalloc<int>(cap.shn);
        sp0 = dalloc<VT>(3 * (size_t)std::max<int64_t>(1, cap.slot)); sp1 = dalloc<VT>(3 * (size_t)std::max<int64_t>(1, cap.slot));
        prho = dalloc<double>(cap.nblk);
        osc_node = dalloc<int>(cap.osc); osc_mean = dalloc<double>(3 * cap.osc); osc_amp = dalloc<double>(3 * cap.osc);
        osc_w = dalloc<double>(cap.osc); osc_phase = dalloc<double>(cap.osc);
        inst = dalloc<InstDev>(cap.sid); ranges_d = dalloc<Ranges>(cap.sid);
        if (seq_dots && !mixed && std::is_same<R, double>::value) {   // segmented reference-order dots
            spec_part = dalloc<double>((size_t)std::max<int64_t>(1, cap.sid) * kSpecChunks);
            spec_model = dalloc<double>((size_t)std::max<int64_t>(1, cap.sid) * kSpecChunks);
            spec_seg = dalloc<double>((size_t)std::max<int64_t>(1, cap.sid) * kSpecStride);
        }
        for (double* pb : {pbb, pden}) HIPCHK(hipMemset(pb, 0, std::max<size_t>(1, np) * sizeof(double)));
        HIPCHK(hipMemset(prho, 0, std::max<int64_t>(1, cap.nblk) * sizeof(double)));
        if (mixed) {
            kmat32 = dalloc<float>(3 * E);
            emass32 = dalloc<float>(E);
            for (float** vec : {&r32, &p32, &q32, &d32, &x32, &x0_32}) {
                *vec = dalloc<float>(3 * N);
                HIPCHK(hipMemset(*vec, 0, 3 * N * sizeof(float)));
            }
            sp32 = dalloc<float>(3 * (size_t)std::max<int64_t>(1, cap.slot));
            pden32 = dalloc<double>(np); prho32 = dalloc<double>(cap.nblk); prr = dalloc<double>(cap.nblk);
            HIPCHK(hipMemset(pden32, 0, std::max<size_t>(1, np) * sizeof(double)));
            HIPCHK(hipMemset(prho32, 0, std::max<int64_t>(1, cap.nblk) * sizeof(double)));
            HIPCHK(hipMemset(prr, 0, std::max<int64_t>(1, cap.nblk) * sizeof(double)));
            inst_in = dalloc<InstDev>(cap.sid);
        }
        place.assign(scenes.size(), Place());
        iperm.resize(scenes.size());
        HIPCHK(hipDeviceSynchronize());
    }

    void release_ranges(Place& pl) {
        for (int k = 0; k < 11; k++) arenas(k)->release(field(pl.at, k), field(pl.size, k));
        pl.placed = false;
        pl.tiles.clear(); pl.fins.clear(); pl.nblks.clear(); pl.oblks.clear();
    }

    // allocate the scene's ranges and upload its tables; false (nothing kept) if it does not fit
    bool place_scene(int si, TiledHost& th) {
        if ((int)place.size() <= si) place.resize(si + 1);
        if ((int)iperm.size() <= si) iperm.resize(si + 1);
        Place& pl = place[si];
        pl.size = th.need;
        for (int k = 0; k < 11; k++) {
            const int64_t at = arenas(k)->alloc(field(pl.size, k));
            if (at < 0) {
                for (int j = 0; j < k; j++) arenas(j)->release(field(pl.at, j), field(pl.size, j));
                return false;
            }
            field(pl.at, k) = at;
        }
        pl.placed = true;
        iperm[si] = std::move(th.iperm);
        const SceneHost& s = scenes[si];
        const int n = s.n_nodes, nE = th.eh.nE;
        const int nb0 = (int)pl.at.node, sid = (int)pl.at.sid;
        const int64_t eb0 = pl.at.elem, ib0 = pl.at.inc;
        const TilePlan& P = th.plan;
        scenes[si].node_begin = nb0;
        scenes[si].elem_begin = (int)eb0;
        // nodes: state, masses, masks, incidence CSR (n + 1 entries: the end stays in range)
        std::vector<VT> xv(3 * (size_t)(n + 1), (VT)0), zero(3 * (size_t)(n + 1), (VT)0);
        for (int i = 0; i < 3 * n; i++) xv[i] = (VT)s.rest[i];
        put_at(x, 3 * (int64_t)nb0, xv);
        put_at(x0, 3 * (int64_t)nb0, xv);
        for (VT* vec : {v, f, b, r, p, q, xcg}) put_at(vec, 3 * (int64_t)nb0, zero);
        std::vector<double> mh(th.eh.mass);
        mh.push_back(0.0);
        put_at(mass, nb0, mh);
        std::vector<int8_t> fx(th.eh.fixed);
        fx.push_back(1);
        put_at(fixed, nb0, fx);
        std::vector<int> ptr(n + 1);
        for (int i = 0; i <= n; i++) ptr[i] = (int)ib0 + th.eh.inc_ptr[i];
        put_at(inc_ptr, nb0, ptr);
        std::vector<int> incv(th.eh.inc);
        for (int& j : incv) j += (int)(4 * eb0);
        put_at(inc, ib0, incv);
        // elements
        std::vector<int4> tv(th.eh.tet);
        for (int4& t : tv) { t.x += nb0; t.y += nb0; t.z += nb0; t.w += nb0; }
        put_at(tet, eb0, tv);
        for (int k = 0; k < 9; k++) put_at(rot, k * Etot + eb0, th.eh.rot.data() + k * (size_t)nE, nE);
        for (int k = 0; k < NR; k++) put_at(rest, k * Etot + eb0, th.eh.rest.data() + k * (size_t)nE, nE);
        for (int k = 0; k < 3; k++) put_at(kmat, k * Etot + eb0, th.eh.kmat.data() + k * (size_t)nE, nE);
        put_at(emass, eb0, th.eh.emass);
        put_at(ltet, eb0, P.ltet);
        put_at(ln_inc, ib0, P.ln_inc);
        if (mixed) {
            for (int k = 0; k < 3; k++) put_at(kmat32, k * Etot + eb0, th.eh.k_lo.data() + k * (size_t)nE, nE);
            put_at(emass32, eb0, th.eh.em_lo);
            std::vector<float> xf(3 * (size_t)(n + 1), 0.0f);
            for (int i = 0; i < 3 * n; i++) xf[i] = (float)s.rest[i];
            put_at(x32, 3 * (int64_t)nb0, xf);
            put_at(x0_32, 3 * (int64_t)nb0, xf);
        }
        // tiles and their local nodes (each tile's list keeps one extra entry for its end)
        const int T = (int)P.t_ebegin.size();
        // the first slot of each shared node belongs to its first tile (the node's owner tile)
        std::vector<int> first_slot(n, -1);
        for (size_t k = 0; k < P.sh_gid.size(); k++) first_slot[P.sh_gid[k]] = P.sh_ptr[k];
        std::vector<TileDesc> tds(T);
        std::vector<int> lg, lp, lps;
        std::vector<uint8_t> lld;
        lg.reserve(pl.size.lnode); lp.reserve(pl.size.lnode); lps.reserve(pl.size.lnode);
        for (int t = 0; t < T; t++) {
            TileDesc& td = tds[t];
            td.scene = sid;
            td.e_begin = (int)eb0 + P.t_ebegin[t];
            td.e_count = P.t_ecount[t];
            td.ln_begin = (int)pl.at.lnode + (int)lg.size();
            td.ln_count = P.t_lncount[t];
            td.flags = (t == 0 ? kTileFirst : 0) | (s.mass_kind == SOFA_AMD_MASS_MESHMATRIX ? kTileConsistent : 0) |
                       (one_material(s) ? kTileOneMaterial : 0);
            td.pslot = (int)pl.at.tile + t;
            td.inc_begin = (int)ib0 + P.ln_inc_ptr[P.t_lnbegin[t]];
            for (int j = 0; j <= P.t_lncount[t]; j++) {
                const int li = P.t_lnbegin[t] + j;
                const bool end = j == P.t_lncount[t];
                lg.push_back(end ? nb0 : P.ln_gid[li] + nb0);
                lp.push_back((int)ib0 + P.ln_inc_ptr[li]);
                if (exact_tiled()) lld.push_back(end ? 0 : P.ln_lead[li]);
                if (end || P.ln_pslot[li] < 0) {
                    lps.push_back(-1);
                } else {
                    const int owner = !exact_tiled() && P.ln_pslot[li] == first_slot[P.ln_gid[li]] ? kSlotOwner : 0;
                    lps.push_back((P.ln_pslot[li] + (int)pl.at.slot) | owner);
                }
            }
        }
        put_at(tiles_d, pl.at.tile, tds);
        put_at(ln_gid, pl.at.lnode, lg); put_at(ln_inc_ptr, pl.at.lnode, lp); put_at(ln_pslot, pl.at.lnode, lps);
        if (exact_tiled()) put_at(ln_lead, pl.at.lnode, lld);
        pl.tiles = tds;
        // shared nodes and their blocks
        const int Sn = (int)th.sh_gid.size();
        std::vector<int> sg(Sn + 1), spp(Sn + 1);
        for (int j = 0; j < Sn; j++) sg[j] = th.sh_gid[j] + nb0;
        sg[Sn] = nb0;
        for (int j = 0; j <= Sn; j++) spp[j] = (int)pl.at.slot + th.sh_ptr[j];
        put_at(sh_gid, pl.at.shn, sg); put_at(sh_ptr, pl.at.shn, spp);
        if (exact_tiled()) {   // slot offsets of shared nodes' incidences (force pass, CG operator)
            put_at(inc_rank, ib0, P.inc_rank);
            put_at(inc_arank, ib0, P.inc_arank);
            put_at(sh_acnt, pl.at.shn, P.sh_acnt);
        }
        std::vector<int> of(n + 1, -1);
        for (int j = 0; j < Sn; j++) of[th.sh_gid[j]] = (int)pl.at.shn + j;
        put_at(sh_of, nb0, of);
        for (int j0 = 0, bi = 0; j0 < Sn; j0 += kBlock, bi++)
            pl.fins.push_back(FinBlk{sid, (int)pl.at.shn + j0, (int)pl.at.shn + std::min(Sn, j0 + kBlock),
                                     (int)(cap.tile + pl.at.fin) + bi, bi, 0, 0, 0});
        const int per = kBlock * kNodesPerThread;
        for (int n0 = 0, bi = 0; n0 < n; n0 += per, bi++)
            pl.nblks.push_back(NodeBlk{sid, nb0 + n0, nb0 + std::min(n, n0 + per), (int)pl.at.nblk + bi, n0 == 0, 0, 0, 0});
        // oscillators
        const int no = (int)s.osc_idx.size();
        if (no > 0) {
            std::vector<int> on(no);
            for (int j = 0; j < no; j++) on[j] = s.osc_idx[j] + nb0;
            put_at(osc_node, pl.at.osc, on);
            put_at(osc_mean, 3 * pl.at.osc, s.osc_mean); put_at(osc_amp, 3 * pl.at.osc, s.osc_amp);
            put_at(osc_w, pl.at.osc, s.osc_w); put_at(osc_phase, pl.at.osc, s.osc_phase);
            for (int j0 = 0; j0 < no; j0 += kBlock)
                pl.oblks.push_back(Blk{sid, (int)pl.at.osc + j0, (int)pl.at.osc + std::min(no, j0 + kBlock), 0});
        }
        // scene records
        InstDev in = make_inst(si);
        in.node_begin = nb0; in.elem_begin = (int)eb0;
        pl.lblks = place_lm(si, sid, nb0, th.eh.fixed, in);
        put_at(inst, sid, &in, 1);
        const Ranges rg{(int)pl.at.tile, T, (int)(cap.tile + pl.at.fin), (int)pl.size.fin, (int)pl.at.nblk,
                        (int)pl.size.nblk};
        put_at(ranges_d, sid, &rg, 1);
        if (mixed) {
            const InstDev ii = make_inner(si, in);
            put_at(inst_in, sid, &ii, 1);
        }
        return true;
    }

    void commit_tiled() {
        const int S = (int)scenes.size();
        iperm.assign(S, {});
        std::vector<TiledHost> th(S);
        Extent tot;
        for (int si = 0; si < S; si++) {
            build_tiled(scenes[si], si, th[si]);
            tot.add(th[si].need);
        }
        reserve(tot);
        for (int si = 0; si < S; si++) {
            if (!place_scene(si, th[si])) throw std::runtime_error("internal: scene does not fit its exact reservation");
            th[si] = TiledHost();
        }
        HIPCHK(hipDeviceSynchronize());
    }

    // live engines: the host tables of a scene, built ahead of its commit (any host thread)
    std::mutex prep_mu;
    std::map<int, std::unique_ptr<TiledHost>> prepared;
    void prepare_scene(int si) override {
        const SceneHost* s;
        {
            std::lock_guard<std::mutex> g(scenes_mu);
            s = &scenes[si];
        }
        std::unique_ptr<TiledHost> th(new TiledHost());
        build_tiled(*s, si, *th);
        std::lock_guard<std::mutex> g(prep_mu);
        prepared[si] = std::move(th);
    }
    // live engines: place one more scene; false if the arenas have no room for it now (its
    // prepared tables are kept for the next attempt)
    bool commit_scene(int si) override {
        if (si >= (int)place.size()) place.resize(si + 1);
        if (place[si].placed) throw std::logic_error("scene already committed");
        std::unique_ptr<TiledHost> th;
        {
            std::lock_guard<std::mutex> g(prep_mu);
            auto it = prepared.find(si);
            if (it != prepared.end()) th = std::move(it->second);
        }
        if (!th) {
            th.reset(new TiledHost());
            build_tiled(scenes[si], si, *th);
        }
        // the ranges it gets are free: no work in flight reads them, so the (synchronous)
        // uploads need no device-wide wait
        if (place_scene(si, *th)) {
            std::lock_guard<std::mutex> g(prep_mu);
            prepared.erase(si);
            return true;
        }
        std::lock_guard<std::mutex> g(prep_mu);
        prepared[si] = std::move(th);
        return false;
    }
    void release_scene(int si) override {
        HIPCHK(hipDeviceSynchronize());   // nothing in flight reads the ranges any more
        sets.clear();                     // launch sets holding the scene's blocks
        branch_plans.clear();
        cur = nullptr;
        release_ranges(place[si]);
        free_vm(si);
        free_k(si);
        free_lm(si);
        iperm[si].clear();
        iperm[si].shrink_to_fit();
        std::lock_guard<std::mutex> g(prep_mu);
        prepared.erase(si);
    }
    bool placed(int si) const override { return si < (int)place.size() && place[si].placed; }
    void usage(int64_t* out) const override {
        const RangeAlloc* a[] = {&a_node, &a_elem, &a_inc, &a_tile, &a_lnode, &a_shn, &a_slot, &a_fin, &a_nblk, &a_osc, &a_sid};
        for (int k = 0; k < 11; k++) { out[2 * k] = a[k]->used; out[2 * k + 1] = a[k]->cap; }
    }

    // block lists of a launch set (a: element blocks / tiles, b: node blocks / finalize blocks,
    // c: tiled node blocks)
    template <typename T>
    static void push_range(std::vector<T>& dst, const std::vector<T>& src, int b0, int n) {
        dst.insert(dst.end(), src.begin() + b0, src.begin() + b0 + n);
    }
    template <typename T>
    static void put(void*& buf, const std::vector<T>& v) {
        buf = dalloc<T>(v.size());
        if (!v.empty()) HIPCHK(hipMemcpy(buf, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    void prepare_set(const int32_t* ids, int n, hipStream_t s) {
        std::vector<int32_t> want(ids, ids + n);
        auto it = sets.find(want);
        if (it != sets.end()) { cur = it->second.get(); return; }
        if (sets.size() >= 4096) {       // bounded cache: drop every set (nothing is in flight after the sync)
            HIPCHK(hipDeviceSynchronize());
            sets.clear();
            branch_plans.clear();
            sets_dropped++;
        }
        std::unique_ptr<LaunchSet> ls(new LaunchSet());
        int maxit = 0;
        std::vector<int> sids;
        if (!tiled) {
            std::vector<Blk> eb, nb;
            for (int32_t id : want) {
                const SceneHost& sc = scenes[id];
                push_range(eb, eblk_h, sc.eblk_begin, sc.eblk_count);
                push_range(nb, nblk_h, sc.nblk_begin, sc.nblk_count);
                ls->pcg = ls->pcg || sc.precond != 0;
                ls->warm = ls->warm || (sc.warm_start && sc.precond == 0);
                maxit = std::max(maxit, sc.cg_iter);
            }
            put(ls->a, eb); put(ls->b, nb);
            ls->na = (int)eb.size(); ls->nb = (int)nb.size(); ls->nc = 0;
        } else {
            std::vector<TileDesc> ta;
            std::vector<FinBlk> fa;
            std::vector<NodeBlk> na;
            for (int32_t id : want) {
                const Place& pl = place[id];
                ta.insert(ta.end(), pl.tiles.begin(), pl.tiles.end());
                fa.insert(fa.end(), pl.fins.begin(), pl.fins.end());
                na.insert(na.end(), pl.nblks.begin(), pl.nblks.end());
                maxit = std::max(maxit, mixed ? scenes[id].inner_iter : scenes[id].cg_iter);
                ls->maxsw = std::max(ls->maxsw, scenes[id].max_sweeps);
                // the partial sums the operator kernel reduces: node blocks (r.r), and for |b|^2 the
                // finalize blocks, or the tiles of a scene without shared nodes
                const int bparts = pl.fins.empty() ? (int)pl.tiles.size() : (int)pl.fins.size();
                ls->maxpart = std::max({ls->maxpart, bparts, (int)pl.nblks.size()});
                ls->plane = ls->plane || scenes[id].plane_on;
                ls->pcg = ls->pcg || scenes[id].precond != 0;
                ls->warm = ls->warm || (scenes[id].warm_start && scenes[id].precond == 0);
            }
            put(ls->a, ta); put(ls->b, fa); put(ls->c, na);
            ls->na = (int)ta.size(); ls->nb = (int)fa.size(); ls->nc = (int)na.size();
        }
        std::vector<Blk> ob, lb;
        for (int32_t id : want) {
            ob.insert(ob.end(), place[id].oblks.begin(), place[id].oblks.end());
            lb.insert(lb.end(), place[id].lblks.begin(), place[id].lblks.end());
            sids.push_back(sid_of(id));
        }
        put(ls->o, ob);
        ls->no = (int)ob.size();
        put(ls->l, lb);
        ls->nl = (int)lb.size();
        void* sc = nullptr;
        put(sc, sids);
        ls->scenes = (int*)sc;
        ls->maxit = maxit;
        ls->n = n;
        cur = ls.get();
        sets[want] = std::move(ls);
        (void)s;
    }

    // one implicit Euler step of the current launch set, enqueued on `s`
    void launch_step(int n, hipStream_t s) {
        if constexpr (std::is_same<VT, double>::value) {
            if (mixed) { launch_step_mixed(n, s); return; }
        }
        if (tiled) { launch_step_tiled(n, s); return; }
        const ElemView<R, NR> ev = view();
        const NodeArrays na = nodes();
        const Blk* eb = (const Blk*)cur->a;
        const Blk* nb = (const Blk*)cur->b;
        if (cur->nl > 0) k_lm_mask<false><<<cur->nl, kBlock, 0, s>>>((const Blk*)cur->l, inst, 0, fixed);
        if (cur->na > 0) k_rhs_elem<R, NR, VT><<<cur->na, kBlock, 0, s>>>(ev, eb, inst, x, v, ef, ef2);
        double* ndp = seq_dots ? nd : nullptr;
        if (cur->nb > 0) k_rhs_node<VT><<<cur->nb, kBlock, 0, s>>>(nb, inst, na, ef, ef2, v, f, b, r, xcg, partial, ndp, x);
        if (cur->pcg) {
            // PCGLinearSolver: the preconditioner from this step's rotations, then r = b - A x from the
            // previous solution, z = P^-1 r (CG scenes of the set pass through untouched)
            if (cur->nb > 0)
                k_pcg_diag<R, NR, Blk><<<cur->nb, kBlock, 0, s>>>(ev, nb, inst, na.inc_ptr, na.inc, mass, fixed, pinv);
            if (cur->nl > 0) k_pcg_lm_identity<<<cur->nl, kBlock, 0, s>>>((const Blk*)cur->l, inst, pinv);
            k_pcg_open<<<(n + 63) / 64, 64, 0, s>>>(cur->scenes, n, inst);
            if (cur->na > 0) k_cg_elem<R, NR, VT><<<cur->na, kBlock, 0, s>>>(ev, eb, inst, xcg, p2, ef);
            if (cur->nb > 0) k_pcg_begin<VT><<<cur->nb, kBlock, 0, s>>>(nb, inst, na, ef, xcg, b, r, zv, pinv, partial, ndp);
        }
        if (cur->warm) {
            // CGLinearSolver warmStart: r = b - A x from the previous solution (other scenes of the
            // set pass through untouched)
            k_warm_open<<<(n + 63) / 64, 64, 0, s>>>(cur->scenes, n, inst);
            if (cur->na > 0) k_cg_elem<R, NR, VT><<<cur->na, kBlock, 0, s>>>(ev, eb, inst, xcg, p2, ef);
            if (cur->nb > 0) k_cg_warm_begin<VT><<<cur->nb, kBlock, 0, s>>>(nb, inst, na, ef, xcg, b, r, partial, ndp);
        }
        for (int it = 1; it <= cur->maxit; it++) {
            // direction of iteration it in buffer it & 1, the previous one in the other
            VT* p_new = (it & 1) ? p : p2;
            const VT* p_old = (it & 1) ? p2 : p;
            if (cur->na > 0) k_cg_elem<R, NR, VT><<<cur->na, kBlock, 0, s>>>(ev, eb, inst, r, p_old, ef, zv);
            if (cur->nb > 0) {
                k_cg_node<VT><<<cur->nb, kBlock, 0, s>>>(nb, inst, na, ef, r, p_old, p_new, q, partial, ndp, zv);
                k_cg_update<VT><<<cur->nb, kBlock, 0, s>>>(nb, inst, p_new, q, xcg, r, partial, ndp, pinv, zv);
            }
        }
        if (cur->nb > 0) k_euler_update<VT><<<cur->nb, kBlock, 0, s>>>(nb, inst, fixed, xcg, v, x);
        if (cur->no > 0) k_oscillate<VT><<<cur->no, kBlock, 0, s>>>((const Blk*)cur->o, inst, oscillators(), x, v);
        if (cur->nl > 0) k_lm_project<VT><<<cur->nl, kBlock, 0, s>>>((const Blk*)cur->l, inst, x, v);
        k_finalize<<<(n + 63) / 64, 64, 0, s>>>(cur->scenes, n, inst);
    }

    // the reference's summation order on the tiled kernels (FP64 tiled engines, chosen before
    // the scenes are placed): shared nodes summed incidence by incidence in element order, every
    // CG dot product node after node (k_seqdot); the trajectories are then the FP64 parity
    // path's in its sequential dot order, bit for bit
    bool exact_tiled() const { return tiled && !mixed && seq_dots && std::is_same<R, double>::value; }

    // one step in the reference's summation order (FP64 tiled engines): per CG iteration the tile
    // apply (interior nodes finished, shared nodes' contributions per incidence), the shared
    // nodes' finalize, p.q node after node, the update, r.r node after node
    void launch_step_exact(int n, hipStream_t s) {
        exact_open(n, s);
        for (int it = 1; it <= cur->maxit; it++) {
            exact_operator(it, s);
            exact_dots(it, n, s);
        }
        exact_close(n, s);
    }
    // its phases, on the current launch set: the right-hand side and |b|^2
    void exact_open(int n, hipStream_t s) {
        const TileArgs<R, NR> ta = tile_args();
        const FinBlk* fb = (const FinBlk*)cur->b;
        const SharedNodes sn{sh_gid, sh_ptr, sh_of, sh_acnt};
        if (cur->nl > 0) k_lm_mask<false><<<cur->nl, kBlock, 0, s>>>((const Blk*)cur->l, inst, 0, fixed);
        if (cur->na > 0) k_tile_rhs<R, NR, VT, 512, true><<<cur->na, 512, 0, s>>>(ta, x, v, f, b, r, xcg, sp0, sp1, pbb);
        if (cur->nb > 0)
            k_fin_rhs<VT, true><<<cur->nb, kBlock, 0, s>>>(fb, inst, ranges_d, sn, mass, fixed, v, sp0, sp1, f, b, r, xcg,
                                                             pbb, x, contact, nd);
        launch_seqdot(n, s, kDotBB);
        if (cur->warm) {
            // CGLinearSolver warmStart: q = A x on the previous solution (the operator of
            // iteration 0), r = b - q and r.r for the tolerance test of iteration 1 (scenes of
            // the set without it pass through untouched: their x is 0 and their r, p, q and
            // terms are rewritten before they are read)
            const NodeBlk* nb = (const NodeBlk*)cur->c;
            if (cur->na > 0) launch_apply(cur->na, s, tile_args(), 0, xcg, nullptr, q, sp0, nullptr, pden, 0);
            if (cur->nb > 0)
                k_fin_exact<VT><<<cur->nb, kBlock, 0, s>>>(fb, inst, sn, mass, fixed, 0, xcg, p, q, sp0, nd, contact);
            if (cur->nc > 0) k_warm_begin_exact<VT><<<cur->nc, kBlock, 0, s>>>(nb, inst, b, q, r, nd);
            launch_seqdot(n, s, kDotRR);
        }
    }
    // CG iteration `it`'s operator (q = A p, the p.q terms): the throughput-bound phase
    void exact_operator(int it, hipStream_t s) {
        const FinBlk* fb = (const FinBlk*)cur->b;
        const SharedNodes sn{sh_gid, sh_ptr, sh_of, sh_acnt};
        if (cur->na > 0) launch_apply(cur->na, s, tile_args(), it, r, p, q, sp0, nullptr, pden, 0);
        if (cur->nb > 0)
            k_fin_exact<VT><<<cur->nb, kBlock, 0, s>>>(fb, inst, sn, mass, fixed, it, r, p, q, sp0, nd, contact);
    }
    // its dot products and update: p.q node after node, x and r updated, r.r node after node (one
    // wave per scene: the latency-bound phase)
    void exact_dots(int it, int n, hipStream_t s) {
        const NodeBlk* nb = (const NodeBlk*)cur->c;
        const SharedNodes sn{sh_gid, sh_ptr, sh_of, sh_acnt};
        launch_seqdot(n, s, kDotPQ);
        if (cur->nc > 0)
            k_update<VT, true><<<cur->nc, kBlock, 0, s>>>(nb, inst, ranges_d, it, sn, sp0, mass, fixed, p, q, xcg, r,
                                                          pden, prho, contact, nd);
        launch_seqdot(n, s, kDotRR);
    }
    // a reference-order dot product of the launch set's n scenes into InstDev::dots[which]: the
    // single chain per scene (k_seqdot), or the chain in dot_segments segments run at once from
    // 2 dot_window candidate starts per boundary (k_seqdot_part / _model / _spec / _resolve, the
    // same bits)
    static constexpr int kSpecMinNodes = 8192;   // smaller scenes: the single chain (under 40 us)
    void launch_seqdot(int n, hipStream_t s, int which) {
        if (spec_seg && dot_segments > 1) {
            const SpecDots sd{spec_part, spec_model, spec_seg, dot_segments, dot_window / 32, kSpecMinNodes};
            k_seqdot_part<<<n * kSpecChunks, 64, 0, s>>>(cur->scenes, inst, nd, sd);
            k_seqdot_model<<<n * kSpecChunks, 64, 0, s>>>(cur->scenes, inst, nd, sd);
            k_seqdot_spec<<<n * (1 + (sd.K - 1) * sd.W), 64, 0, s>>>(cur->scenes, inst, nd, sd, which);
            k_seqdot_resolve<<<n, 64, 0, s>>>(cur->scenes, inst, nd, sd, which);
        } else {
            k_seqdot<<<n, 64, 0, s>>>(cur->scenes, inst, nd, which);
        }
    }
    // the Euler update and the step's end
    void exact_close(int n, hipStream_t s) {
        const NodeBlk* nb = (const NodeBlk*)cur->c;
        if (cur->nc > 0) k_integrate<VT><<<cur->nc, kBlock, 0, s>>>(nb, inst, fixed, xcg, v, x);
        if (cur->no > 0) k_oscillate<VT><<<cur->no, kBlock, 0, s>>>((const Blk*)cur->o, inst, oscillators(), x, v);
        if (cur->nl > 0) k_lm_project<VT><<<cur->nl, kBlock, 0, s>>>((const Blk*)cur->l, inst, x, v);
        k_step_end_tiled<<<(n + 63) / 64, 64, 0, s>>>(cur->scenes, n, inst);
    }

    // the tile operator's launch shape and how much geometry it recomputes (results of the
    // variants differ in rounding only: block sums and geometry recomputed in R instead of
    // rounded from the host's double); picked by sofa_amd_engine_autotune or set explicitly
    static constexpr int kApplyVariants = 4;
    static int variant_threads(int v) { return v < 2 ? 512 : 256; }
    static int variant_geometry(int v) { return (v & 1) ? 1 : 2; }
    int apply_variant = 0;
    void launch_apply(int na, hipStream_t s, const TileArgs<R, NR>& ta, int it, const VT* rr, VT* pp, VT* qq, VT* sp,
                      const double* pin, double* pd, int pin_nodes) {
        if constexpr (std::is_same<R, double>::value) {
            if (exact_tiled()) {   // scalars from k_seqdot: no partial sums reduced
                if (cur->plane) launch_apply_t<false, true, true>(na, s, ta, it, rr, pp, qq, sp, pin, pd, pin_nodes);
                else launch_apply_t<false, false, true>(na, s, ta, it, rr, pp, qq, sp, pin, pd, pin_nodes);
                return;
            }
        }
        const bool wide = cur->maxpart > kWavePartials;
        if (cur->plane) {
            if (wide) launch_apply_t<true, true>(na, s, ta, it, rr, pp, qq, sp, pin, pd, pin_nodes);
            else launch_apply_t<false, true>(na, s, ta, it, rr, pp, qq, sp, pin, pd, pin_nodes);
        } else {
            if (wide) launch_apply_t<true, false>(na, s, ta, it, rr, pp, qq, sp, pin, pd, pin_nodes);
            else launch_apply_t<false, false>(na, s, ta, it, rr, pp, qq, sp, pin, pd, pin_nodes);
        }
    }
    template <bool WIDE, bool PLANE, bool EXACT = false>
    void launch_apply_t(int na, hipStream_t s, const TileArgs<R, NR>& ta, int it, const VT* rr, VT* pp, VT* qq, VT* sp,
                        const double* pin, double* pd, int pin_nodes) {
        switch (apply_variant) {
        case 1: k_tile_apply<R, NR, VT, 512, 1, WIDE, PLANE, EXACT><<<na, 512, 0, s>>>(ta, it, rr, pp, qq, sp, pin, pd, pin_nodes, x, x0); break;
        case 2: k_tile_apply<R, NR, VT, 256, 2, WIDE, PLANE, EXACT><<<na, 256, 0, s>>>(ta, it, rr, pp, qq, sp, pin, pd, pin_nodes, x, x0); break;
        case 3: k_tile_apply<R, NR, VT, 256, 1, WIDE, PLANE, EXACT><<<na, 256, 0, s>>>(ta, it, rr, pp, qq, sp, pin, pd, pin_nodes, x, x0); break;
        default: k_tile_apply<R, NR, VT, 512, 2, WIDE, PLANE, EXACT><<<na, 512, 0, s>>>(ta, it, rr, pp, qq, sp, pin, pd, pin_nodes, x, x0); break;
        }
    }
    // times every variant on the engine's scenes (all of them in one launch, `reps` launches
    // each, events on `s`) and keeps the fastest (by at least 1 %); before the first step only
    // (the timing runs the operator on the CG buffers)
    int autotune(int reps, hipStream_t s, float* us_out) override {
        if (!tiled) throw std::invalid_argument("autotune needs a tiled engine");
        if (stepped) throw std::logic_error("autotune before the first step");
        std::vector<int32_t> ids;
        for (int i = 0; i < (int)scenes.size(); i++) if (placed(i)) ids.push_back(i);
        if (ids.empty()) throw std::invalid_argument("no committed scene");
        prepare_set(ids.data(), (int)ids.size(), s);
        const TileArgs<R, NR> ta = tile_args();
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        int best = 0;
        float best_us = 0;
        for (int v = 0; v < kApplyVariants; v++) {
            apply_variant = v;
            for (int i = 0; i < 2; i++) launch_apply(cur->na, s, ta, 2, r, p, q, sp0, prho, pden, 0);
            HIPCHK(hipEventRecord(e0, s));
            for (int i = 0; i < reps; i++) launch_apply(cur->na, s, ta, 2, r, p, q, sp0, prho, pden, 0);
            HIPCHK(hipEventRecord(e1, s));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            const float us = 1e3f * ms / (float)std::max(1, reps);
            if (us_out) us_out[v] = us;
            // a variant displaces the kept one only when measurably faster (1 %): within timing
            // noise the earlier one stays, variant 0 (everything recomputed) reading the fewest bytes
            if (v == 0 || us < 0.99f * best_us) { best = v; best_us = us; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        apply_variant = best;
        reset_graph();
        return best;
    }
    void set_apply_variant(int v) override {
        if (!tiled || v < 0 || v >= kApplyVariants) throw std::invalid_argument("invalid operator variant");
        HIPCHK(hipDeviceSynchronize());
        apply_variant = v;
        reset_graph();
    }

    TileArgs<R, NR> tile_args() const {
        TileArgs<R, NR> ta;
        ta.tiles = (const TileDesc*)cur->a;
        ta.inst = inst;
        ta.ranges = ranges_d;
        ta.ev = view();
        ta.ltet = ltet;
        ta.tn = TileNodes{ln_gid, ln_inc_ptr, ln_inc, ln_pslot, inc_rank, inc_arank, ln_lead};
        ta.mass = mass;
        ta.fixed = fixed;
        ta.contact = contact;
        ta.emass = emass;
        ta.x_lo = nullptr;
        ta.nd = nd;
        ta.dinv = nullptr;
        return ta;
    }

    void launch_step_tiled(int n, hipStream_t s) {
        const TileArgs<R, NR> ta = tile_args();
        const FinBlk* fb = (const FinBlk*)cur->b;
        const NodeBlk* nb = (const NodeBlk*)cur->c;
        const SharedNodes sn{sh_gid, sh_ptr, sh_of, sh_acnt};
        if constexpr (std::is_same<R, double>::value) {
            if (exact_tiled()) { launch_step_exact(n, s); return; }
        }
        if (cur->nl > 0) k_lm_mask<false><<<cur->nl, kBlock, 0, s>>>((const Blk*)cur->l, inst, 0, fixed);
        if (cur->na > 0) k_tile_rhs<R, NR, VT, 512><<<cur->na, 512, 0, s>>>(ta, x, v, f, b, r, xcg, sp0, sp1, pbb);
        if (cur->nb > 0) k_fin_rhs<VT><<<cur->nb, kBlock, 0, s>>>(fb, inst, ranges_d, sn, mass, fixed, v, sp0, sp1, f, b, r, xcg, pbb, x, contact);
        for (int it = 1; it <= cur->maxit; it++) {
            if (cur->na > 0)
                launch_apply(cur->na, s, ta, it, r, p, q, sp0, it == 1 ? pbb : prho, pden, 0);
            if (cur->nc > 0)
                k_update<VT><<<cur->nc, kBlock, 0, s>>>(nb, inst, ranges_d, it, sn, sp0, mass, fixed, p, q, xcg, r, pden, prho, contact);
        }
        if (cur->nc > 0) k_integrate<VT><<<cur->nc, kBlock, 0, s>>>(nb, inst, fixed, xcg, v, x);
        if (cur->no > 0) k_oscillate<VT><<<cur->no, kBlock, 0, s>>>((const Blk*)cur->o, inst, oscillators(), x, v);
        if (cur->nl > 0) k_lm_project<VT><<<cur->nl, kBlock, 0, s>>>((const Blk*)cur->l, inst, x, v);
        k_step_end_tiled<<<(n + 63) / 64, 64, 0, s>>>(cur->scenes, n, inst);
    }

    // MIXED step: FP64 RHS (which also writes the float rotations), refinement sweeps with the
    // FP32 inner CG, FP64 update
    void launch_step_mixed(int n, hipStream_t s) {
        TileArgs<R, NR> ta = tile_args();
        ta.x_lo = x32;                 // the inner operator recomputes the rotations from these
        const FinBlk* fb = (const FinBlk*)cur->b;
        const NodeBlk* nb = (const NodeBlk*)cur->c;
        const SharedNodes sn{sh_gid, sh_ptr, sh_of, sh_acnt};
        InnerCg ic;
        ic.ta.tiles = ta.tiles;
        ic.ta.inst = inst_in;
        ic.ta.ranges = ranges_d;
        // rotations and rest coordinates are recomputed from x32 / x0_32: only K is read
        ic.ta.ev = ElemView<float, 6>{tet, nullptr, nullptr, kmat32, Etot};
        ic.ta.ltet = ltet;
        ic.ta.tn = ta.tn;
        ic.ta.mass = mass;
        ic.ta.fixed = fixed;
        ic.ta.contact = contact;
        ic.ta.emass = emass32;
        ic.ta.x_lo = nullptr;
        ic.ta.nd = nullptr;
        // PCGLinearSolver with a Jacobi preconditioner: the inner CG runs preconditioned
        ic.ta.dinv = nullptr;   // the preconditioner acts in the node kernels (z), not in the operator
        ic.z = cur->pcg ? z32 : nullptr;
        ic.dinv = dinv32;
        ic.binv = binv32;
        ic.x = x32; ic.x0 = x0_32;
        ic.fb = fb; ic.nb = nb;
        ic.na = cur->na; ic.nfb = cur->nb; ic.nnb = cur->nc; ic.maxit = cur->maxit;
        ic.wide = cur->maxpart > kWavePartials;
        ic.plane = cur->plane;
        ic.sn = sn;
        ic.r = r32; ic.p = p32; ic.q = q32; ic.d = d32; ic.sp = sp32;
        ic.pden = pden32; ic.prho = prho32;
        if (cur->nl > 0) k_lm_mask<false><<<cur->nl, kBlock, 0, s>>>((const Blk*)cur->l, inst, 0, fixed);
        if (cur->na > 0) k_tile_rhs<R, NR, VT, 512><<<cur->na, 512, 0, s>>>(ta, x, v, f, b, r, xcg, sp0, sp1, pbb);
        if (cur->nb > 0) k_fin_rhs<VT><<<cur->nb, kBlock, 0, s>>>(fb, inst, ranges_d, sn, mass, fixed, v, sp0, sp1, f, b, r, xcg, pbb, x, contact);
        if (cur->pcg) {   // the inner CG's Jacobi preconditioner from this step's rotations
            if (cur->nc > 0)
                k_jacobi32<R, NR><<<cur->nc, kBlock, 0, s>>>(view(), nb, inst, inc_ptr, inc, mass, fixed, dinv32, binv32);
            if (cur->nl > 0) k_jacobi32_lm<<<cur->nl, kBlock, 0, s>>>((const Blk*)cur->l, inst, dinv32, binv32);
        }
        for (int sw = 1; sw <= cur->maxsw + 1; sw++) {
            if (sw > 1) {   // q = A x, FP64
                if (cur->na > 0)
                    launch_apply(cur->na, s, ta, 0, xcg, nullptr, q, sp0, nullptr, pden, 0);
                if (cur->nb > 0)
                    k_fin_apply<VT><<<cur->nb, kBlock, 0, s>>>(fb, inst, ranges_d, sn, mass, fixed, 0, xcg, nullptr, q, sp0, pden, contact);
            }
            if (cur->nc > 0)
                k_ir_residual<<<cur->nc, kBlock, 0, s>>>(nb, inst, sw, b, q, r32, d32, prr, prho32, dinv32, binv32, ic.z);
            k_ir_control<<<n, kBlock, 0, s>>>(cur->scenes, inst, inst_in, ranges_d, sw, prr);
            if (sw <= cur->maxsw) {
                launch_inner_cg(ic, s);
                if (cur->nc > 0) k_ir_accumulate<<<cur->nc, kBlock, 0, s>>>(nb, inst, inst_in, d32, xcg);
            }
        }
        if (cur->nc > 0) k_integrate<VT><<<cur->nc, kBlock, 0, s>>>(nb, inst, fixed, xcg, v, x);
        if (cur->no > 0) k_oscillate<VT><<<cur->no, kBlock, 0, s>>>((const Blk*)cur->o, inst, oscillators(), x, v);
        if (cur->nl > 0) k_lm_project<VT><<<cur->nl, kBlock, 0, s>>>((const Blk*)cur->l, inst, x, v);
        k_step_end_mixed<<<(n + 63) / 64, 64, 0, s>>>(cur->scenes, n, inst, inst_in);
    }

    void reset_graph() override {
        for (auto& kv : sets)
            if (kv.second->gexec) {
                (void)hipGraphExecDestroy(kv.second->gexec);
                kv.second->gexec = nullptr;
            }
    }

    // a merged step in concurrent branches: the scenes are split into G groups, each stepped by its
    // own launch sequence on its own stream (forked from and joined back into `s` by events), so
    // that one group's latency-bound phases (the reference order's node-after-node dot products:
    // one wave per scene, one dependent add per node) overlap the other groups' element and node
    // kernels.  Scenes are dealt to the groups by node count, largest first, round-robin: each
    // group holds one of the G longest dot-product chains and about 1/G of the work.
    static constexpr int kMaxBranches = 8;
    std::map<std::vector<int32_t>, std::vector<std::vector<int32_t>>> branch_plans;
    hipStream_t aux[kMaxBranches - 1] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kMaxBranches - 1] = {};
    // the default: one branch per hardware queue the runtime's own streams leave free (HIP gives a
    // process GPU_MAX_HW_QUEUES of them, 4 unless set; two are taken), at most 4 (measured on the
    // 256-beam mix with 8 queues: 1 branch 30.6 ms/step, 2 26.0, 3 24.0, 4 23.7); branches sharing
    // a queue run one after the other
    static int default_branches() {
        const char* e = std::getenv("GPU_MAX_HW_QUEUES");
        const int queues = e ? std::atoi(e) : 4;
        return std::max(1, std::min(4, queues - 2));
    }
    int branch_count(int n) const {
        const int g = branches > 0 ? branches : (exact_tiled() ? default_branches() : 1);
        return std::max(1, std::min({g, kMaxBranches, n / 2}));
    }
    void step_branches(const int32_t* ids, int n, int nsteps, hipStream_t s, int G) {
        std::vector<int32_t> key(ids, ids + n);
        key.push_back(-G);
        auto it = branch_plans.find(key);
        if (it == branch_plans.end()) {
            std::vector<int32_t> order(ids, ids + n);
            std::stable_sort(order.begin(), order.end(),
                             [&](int32_t a, int32_t b) { return scenes[a].n_nodes > scenes[b].n_nodes; });
            std::vector<std::vector<int32_t>> gr(G);
            for (int k = 0; k < n; k++) gr[k % G].push_back(order[k]);
            for (auto& g : gr) std::sort(g.begin(), g.end());
            it = branch_plans.emplace(key, std::move(gr)).first;
        }
        if (!fork_ev) HIPCHK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
        for (int g = 1; g < G; g++) {
            if (!aux[g - 1]) HIPCHK(hipStreamCreateWithFlags(&aux[g - 1], hipStreamNonBlocking));
            if (!join_ev[g - 1]) HIPCHK(hipEventCreateWithFlags(&join_ev[g - 1], hipEventDisableTiming));
        }
        // a copy: step_set may empty the plan cache (prepare_set's bounded cache of launch sets)
        const std::vector<std::vector<int32_t>> groups = it->second;
        if (branch_turns && exact_tiled()) { step_turns(groups, nsteps, s, G); return; }
        for (int st = 0; st < nsteps; st++) {
            HIPCHK(hipEventRecord(fork_ev, s));
            for (int g = 1; g < G; g++) HIPCHK(hipStreamWaitEvent(aux[g - 1], fork_ev, 0));
            for (int g = 0; g < G; g++) {
                const std::vector<int32_t>& grp = groups[g];
                step_set(grp.data(), (int)grp.size(), 1, g == 0 ? s : aux[g - 1]);
            }
            for (int g = 1; g < G; g++) {
                HIPCHK(hipEventRecord(join_ev[g - 1], aux[g - 1]));
                HIPCHK(hipStreamWaitEvent(s, join_ev[g - 1], 0));
            }
        }
    }

    // the branches' operator phases taking turns (the reference order's step): per CG iteration
    // each group's operator and finalize wait for the previous group's (round-robin over the
    // groups), so that while one group runs its two node-after-node dot products (one wave per
    // scene) the GPU runs the other groups' operators one at a time, instead of every group's
    // operator at once and then every group's chains on an idle GPU.  Launched directly: a graph
    // per group cannot wait for another group's kernels.  Off by default: measured on the 256-beam
    // mix (4 branches, 8 queues) 22.3 ms/step against 21.6 running freely -- a group's operator
    // alone takes no less time than beside the others' (185-211 us for a quarter of the beams),
    // its turn's wait adds about 200 us per iteration, and each group's own chain of dependent
    // phases (two dot chains of about 190 us, its operator, finalize and update) sets the step.
    hipEvent_t turn_ev[kMaxBranches] = {};
    void step_turns(const std::vector<std::vector<int32_t>>& groups, int nsteps, hipStream_t s, int G) {
        LaunchSet* ls[kMaxBranches] = {};
        for (;;) {   // again if preparing a group's set emptied the cache under an earlier group's
            const uint64_t d0 = sets_dropped;
            for (int g = 0; g < G; g++) {
                prepare_set(groups[g].data(), (int)groups[g].size(), s);
                ls[g] = cur;
            }
            if (sets_dropped == d0) break;
        }
        int maxit = 0;
        for (int g = 0; g < G; g++) {
            maxit = std::max(maxit, ls[g]->maxit);
            if (!turn_ev[g]) HIPCHK(hipEventCreateWithFlags(&turn_ev[g], hipEventDisableTiming));
        }
        auto stream = [&](int g) { return g == 0 ? s : aux[g - 1]; };
        for (int st = 0; st < nsteps; st++) {
            HIPCHK(hipEventRecord(fork_ev, s));
            for (int g = 1; g < G; g++) HIPCHK(hipStreamWaitEvent(aux[g - 1], fork_ev, 0));
            for (int g = 0; g < G; g++) {
                cur = ls[g];
                exact_open(cur->n, stream(g));
            }
            int last = -1;   // the group whose operator phase was enqueued last
            for (int it = 1; it <= maxit; it++)
                for (int g = 0; g < G; g++) {
                    if (it > ls[g]->maxit) continue;
                    cur = ls[g];
                    const hipStream_t sg = stream(g);
                    if (last >= 0 && last != g) HIPCHK(hipStreamWaitEvent(sg, turn_ev[last], 0));
                    exact_operator(it, sg);
                    HIPCHK(hipEventRecord(turn_ev[g], sg));
                    last = g;
                    exact_dots(it, cur->n, sg);
                }
            for (int g = 0; g < G; g++) {
                cur = ls[g];
                exact_close(cur->n, stream(g));
            }
            for (int g = 1; g < G; g++) {
                HIPCHK(hipEventRecord(join_ev[g - 1], aux[g - 1]));
                HIPCHK(hipStreamWaitEvent(s, join_ev[g - 1], 0));
            }
        }
        HIPCHK(hipGetLastError());
    }

    void step(const int32_t* ids, int n, int nsteps, hipStream_t s) override {
        if (nsteps > 0) stepped = true;
        const int G = branch_count(n);
        if (G > 1) step_branches(ids, n, nsteps, s, G);
        else step_set(ids, n, nsteps, s);
    }

    // `nsteps` steps of one launch set (the listed scenes in one launch per phase) on `s`
    void step_set(const int32_t* ids, int n, int nsteps, hipStream_t s) {
        prepare_set(ids, n, s);
        if (!use_graphs) {
            for (int st = 0; st < nsteps; st++) launch_step(n, s);
            HIPCHK(hipGetLastError());
            return;
        }
        // the whole step (RHS, every CG iteration, update) is captured once per launch set and
        // replayed: one graph launch per step instead of 3 n_cg + 5 kernel launches
        hipGraphExec_t& gexec = cur->gexec;
        if (!gexec && nsteps > 0) {
            hipGraph_t graph = nullptr;
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            launch_step(n, s);
            const hipError_t le = hipGetLastError();
            const hipError_t ce = hipStreamEndCapture(s, &graph);
            if (le != hipSuccess) throw HipError(std::string("launch during capture: ") + hipGetErrorString(le));
            if (ce != hipSuccess) throw HipError(std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
            const hipError_t ie = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) { gexec = nullptr; throw HipError(std::string("hipGraphInstantiate: ") + hipGetErrorString(ie)); }
        }
        for (int st = 0; st < nsteps; st++) HIPCHK(hipGraphLaunch(gexec, s));
    }

    // measurement hook: `reps` launches of the scene's CG operator kernel (the tiled apply, with
    // the iteration-2 data flow: p = r + beta p, q = A p) on the current state
    void bench_apply(int scene, int reps, hipStream_t s) override {
        if (!tiled) throw std::invalid_argument("bench_apply needs a tiled engine");
        std::vector<int32_t> ids;
        if (scene >= 0) ids.push_back(scene);
        else for (int i = 0; i < (int)scenes.size(); i++) if (placed(i)) ids.push_back(i);
        prepare_set(ids.data(), (int)ids.size(), s);
        const TileArgs<R, NR> ta = tile_args();
        for (int i = 0; i < reps; i++)
            if (cur->na > 0) launch_apply(cur->na, s, ta, 2, r, p, q, sp0, prho, pden, 0);
        HIPCHK(hipGetLastError());
    }
    // tiles, tile-local nodes, nodes finished by the finalize kernel, shared partial slots,
    // incidences, elements, nodes
    void tiling(int scene, int32_t* out) override {
        const SceneHost& sc = scenes[scene];
        for (int k = 0; k < 7; k++) out[k] = 0;
        out[5] = (int32_t)(sc.tets.size() / 4);
        out[6] = sc.n_nodes;
        if (!tiled) return;
        const Place& pl = place[scene];
        out[0] = (int32_t)pl.tiles.size();
        for (const TileDesc& td : pl.tiles) out[1] += td.ln_count;
        for (const FinBlk& fb : pl.fins) out[2] += fb.end - fb.begin;
        out[3] = (int32_t)pl.size.slot;
        out[4] = 4 * out[5];
    }

    void read_vec(const VT* src, int scene, double* out) {
        const SceneHost& sc = scenes[scene];
        std::vector<VT> tmp(3 * (size_t)sc.n_nodes);
        HIPCHK(hipMemcpy(tmp.data(), src + 3 * (int64_t)node_begin(scene), tmp.size() * sizeof(VT), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < tmp.size(); i++) out[i] = (double)tmp[i];
    }
    void read(int scene, double* xo, double* vo, double* fo) override {
        HIPCHK(hipDeviceSynchronize());
        if (xo) read_vec(x, scene, xo);
        if (vo) read_vec(v, scene, vo);
        if (fo) read_vec(f, scene, fo);
    }
    void write(int scene, const double* xi, const double* vi) override {
        HIPCHK(hipDeviceSynchronize());
        const SceneHost& sc = scenes[scene];
        std::vector<VT> tmp(3 * (size_t)sc.n_nodes);
        if (xi) {
            for (size_t i = 0; i < tmp.size(); i++) tmp[i] = (VT)xi[i];
            HIPCHK(hipMemcpy(x + 3 * (int64_t)node_begin(scene), tmp.data(), tmp.size() * sizeof(VT), hipMemcpyHostToDevice));
        }
        if (vi) {
            for (size_t i = 0; i < tmp.size(); i++) tmp[i] = (VT)vi[i];
            HIPCHK(hipMemcpy(v + 3 * (int64_t)node_begin(scene), tmp.data(), tmp.size() * sizeof(VT), hipMemcpyHostToDevice));
        }
    }
    void stats(int scene, sofa_amd_cg_stats* out) override {
        InstDev d;
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(&d, inst + sid_of(scene), sizeof(d), hipMemcpyDeviceToHost));
        out->iterations = d.last_iters;
        out->end_reason = d.end_reason;
        out->step_count = d.step_count;
        out->ir_sweeps = mixed ? d.ir_sweeps : 0;
        out->err = d.err;
        out->den = d.den;
        out->time = d.time;
        out->dot_fallbacks = d.dot_fallbacks;
        out->reserved = 0;
    }
    void element(int scene, int e, double* rot9, double* k3, double* rest12) override {
        HIPCHK(hipDeviceSynchronize());
        const int64_t ge = elem_begin(scene) + (int64_t)iperm[scene][e];
        R tmp;
        for (int k = 0; k < 9; k++) {
            HIPCHK(hipMemcpy(&tmp, rot + k * Etot + ge, sizeof(R), hipMemcpyDeviceToHost));
            rot9[k] = (double)tmp;
        }
        for (int k = 0; k < 3; k++) {
            HIPCHK(hipMemcpy(&tmp, kmat + k * Etot + ge, sizeof(R), hipMemcpyDeviceToHost));
            k3[k] = (double)tmp;
        }
        double rc[9] = {0};
        if (NR == 9) {
            for (int k = 0; k < 9; k++) {
                HIPCHK(hipMemcpy(&tmp, rest + k * Etot + ge, sizeof(R), hipMemcpyDeviceToHost));
                rc[k] = (double)tmp;
            }
        } else {
            const int pick[6] = {0, 3, 4, 6, 7, 8};
            for (int k = 0; k < 6; k++) {
                HIPCHK(hipMemcpy(&tmp, rest + k * Etot + ge, sizeof(R), hipMemcpyDeviceToHost));
                rc[pick[k]] = (double)tmp;
            }
        }
        for (int k = 0; k < 3; k++) rest12[k] = 0.0;
        for (int k = 0; k < 9; k++) rest12[3 + k] = rc[k];
    }

    // component-level calls: vectors in the engine's state precision (double for FP64,
    // float for FP32), scene-local AoS xyz
    void add_force(int scene, const void* xd, void* fd, hipStream_t s) override {
        const SceneHost& sc = scenes[scene];
        const int64_t nE = (int64_t)sc.tets.size() / 4;
        const int nb0 = node_begin(scene);
        const int64_t eb0 = elem_begin(scene);
        if (nE > 0) {
            const int grid = (int)((nE + kBlock - 1) / kBlock);
            k_elem_add_force<R, NR, VT><<<grid, kBlock, 0, s>>>(view(), eb0, eb0 + nE, nb0, (const VT*)xd, ef);
        }
        const int gn = (sc.n_nodes + kBlock - 1) / kBlock;
        if (gn > 0) k_gather<VT><<<gn, kBlock, 0, s>>>(inc_ptr, inc, nb0, nb0 + sc.n_nodes, ef, (VT*)fd, 0);
        HIPCHK(hipGetLastError());
    }
    void add_dforce(int scene, const void* dxd, void* dfd, double kf, hipStream_t s) override {
        const SceneHost& sc = scenes[scene];
        const int64_t nE = (int64_t)sc.tets.size() / 4;
        const int nb0 = node_begin(scene);
        const int64_t eb0 = elem_begin(scene);
        if (nE > 0) {
            const int grid = (int)((nE + kBlock - 1) / kBlock);
            // addDForce: kFactor = (Real) mparams->kFactor(); applied as `KJtD *= fact`
            k_elem_add_dforce<R, NR, VT><<<grid, kBlock, 0, s>>>(view(), eb0, eb0 + nE, nb0, (const VT*)dxd, (R)kf, ef);
        }
        const int gn = (sc.n_nodes + kBlock - 1) / kBlock;
        if (gn > 0) k_gather<VT><<<gn, kBlock, 0, s>>>(inc_ptr, inc, nb0, nb0 + sc.n_nodes, ef, (VT*)dfd, 1);
        HIPCHK(hipGetLastError());
    }

    int scene_nodes(int scene) const override { return scenes[scene].n_nodes; }

    // Mass::addMDx / accFromF / addForce on the caller's vectors (state type VT, scene-local)
    void mass_call(int scene, int what, const void* in, void* out, double factor, hipStream_t s) override {
        const SceneHost& sc = scenes[scene];
        if (sc.mass_kind == SOFA_AMD_MASS_MESHMATRIX) {
            if (what == 1)   // MeshMatrixMass.inl:2055-2060: the sparse mass is not inverted
                throw std::invalid_argument("accFromF needs a lumped mass (MeshMatrixMass lumping = false)");
            if (what == 0 && tiled)   // the tiled engines keep the edge masses per element only
                throw std::invalid_argument("addMDx of a consistent MeshMatrixMass runs on the FP64 engine");
        }
        const int n = sc.n_nodes, nb0 = node_begin(scene), sid = sid_of(scene);
        const int grid = (n + kBlock - 1) / kBlock;
        if (grid == 0) return;
        if (what == 0)
            k_mass_mdx<VT><<<grid, kBlock, 0, s>>>(inst, sid, nodes(), nb0, n, (const VT*)in, (VT*)out, factor);
        else if (what == 1)
            k_mass_acc<VT><<<grid, kBlock, 0, s>>>(inst, sid, mass, nb0, n, (const VT*)in, (VT*)out);
        else
            k_mass_add_force<VT><<<grid, kBlock, 0, s>>>(inst, sid, mass, nb0, n, (VT*)out);
        HIPCHK(hipGetLastError());
    }

    // the scene's oscillator entries [o0, o1)
    void osc_range(int scene, int& o0, int& o1) const {
        const auto& ob = place[scene].oblks;
        o0 = ob.empty() ? 0 : ob.front().begin;
        o1 = ob.empty() ? 0 : ob.back().end;
    }
    // FixedProjectiveConstraint then OscillatorProjectiveConstraint (scene order) on a caller vector
    void project(int scene, int what, void* vec, double t, hipStream_t s) override {
        const int n = scenes[scene].n_nodes, nb0 = node_begin(scene), sid = sid_of(scene);
        const int grid = (n + kBlock - 1) / kBlock;
        const int nlm = scenes[scene].lm_on ? (int)scenes[scene].lm_idx.size() : 0;
        // the movement's response mask at the scene's current time
        if (what == 0 && nlm > 0)
            k_lm_mask<true><<<(nlm + kBlock - 1) / kBlock, kBlock, 0, s>>>(nullptr, inst, sid, fixed);
        if (what < 2 && grid > 0) k_project<VT><<<grid, kBlock, 0, s>>>(inst, sid, fixed, nb0, n, (VT*)vec, what);
        int o0, o1;
        osc_range(scene, o0, o1);
        if (what > 0 && o1 > o0)
            k_oscillate_vec<VT><<<(o1 - o0 + kBlock - 1) / kBlock, kBlock, 0, s>>>(oscillators(), o0, o1, nb0, t, (VT*)vec,
                                                                                  what == 2 ? 1 : 0);
        if (what > 0 && nlm > 0) {
            k_lm_vec<VT><<<(nlm + kBlock - 1) / kBlock, kBlock, 0, s>>>(inst, sid, t, nb0, (VT*)vec, what == 2 ? 1 : 0);
            if (what == 2) k_lm_mark_x0<<<1, 64, 0, s>>>(inst, sid);
        }
        HIPCHK(hipGetLastError());
    }

    // vDot into out (device), order 1 the reference's sequential sum, 0 the blocked order with
    // its per-block partials in `scratch` (scratch_len doubles)
    void vdot(int scene, const void* a, const void* b, int order, double* out, double* scratch, int64_t scratch_len,
              hipStream_t s) override {
        const int n = scenes[scene].n_nodes;
        if (order == 1) {
            k_vdot_seq<VT><<<1, 64, 0, s>>>((const VT*)a, (const VT*)b, n, out);
        } else {
            const int nb = (n + kBlock - 1) / kBlock;
            if (nb > scratch_len) throw std::length_error("vDot: scene larger than the scratch buffer");
            if (nb > 0) k_vdot_blocks<VT><<<nb, kBlock, 0, s>>>((const VT*)a, (const VT*)b, n, scratch);
            k_vdot_final<<<1, kBlock, 0, s>>>(scratch, nb, out);
        }
        HIPCHK(hipGetLastError());
    }

    // CGLinearSolver::solve (CGLinearSolver.inl:71-315) on the caller's right-hand side with the
    // scene's CG Data: x = (m M + b B + k K)^-1 b through GraphScatteredMatrix::apply's operator
    // (mass addMDx, FEM addDForce with the last addForce's rotations, plane, projectResponse).
    // The components carry no Rayleigh damping of their own here (rayleighMass / rayleighStiffness
    // Data 0), so B = 0 and bFactor adds nothing (Mass.inl:94-104, BaseForceField.cpp:38-47).
    // Blocks until done; stats as sofa_amd_scene_stats reports them.
    void cg_solve(int scene, double mf, double bf, double kf, const void* bvec, void* xvec, sofa_amd_cg_stats* st,
                  hipStream_t s) override {
        if (mixed) throw std::invalid_argument("the MIXED engine solves by refinement; CG solve runs on the others");
        (void)bf;
        const int sid = sid_of(scene), n = scenes[scene].n_nodes, nb0 = node_begin(scene);
        InstDev saved;
        HIPCHK(hipMemcpy(&saved, inst + sid, sizeof(InstDev), hipMemcpyDeviceToHost));
        InstDev tmp = saved;
        tmp.m_cg = mf;
        tmp.k_cg = kf;
        HIPCHK(hipMemcpy(inst + sid, &tmp, sizeof(InstDev), hipMemcpyHostToDevice));
        const int32_t one = scene;
        prepare_set(&one, 1, s);
        if (!tiled) {
            const ElemView<R, NR> ev = view();
            const NodeArrays na = nodes();
            const Blk* eb = (const Blk*)cur->a;
            const Blk* nbk = (const Blk*)cur->b;
            double* ndp = seq_dots ? nd : nullptr;
            if (cur->nb > 0) k_cg_begin<VT><<<cur->nb, kBlock, 0, s>>>(nbk, inst, (const VT*)bvec, nb0, b, r, xcg, partial, ndp);
            for (int it = 1; it <= cur->maxit; it++) {
                VT* p_new = (it & 1) ? p : p2;
                const VT* p_old = (it & 1) ? p2 : p;
                if (cur->na > 0) k_cg_elem<R, NR, VT><<<cur->na, kBlock, 0, s>>>(ev, eb, inst, r, p_old, ef);
                if (cur->nb > 0) {
                    k_cg_node<VT><<<cur->nb, kBlock, 0, s>>>(nbk, inst, na, ef, r, p_old, p_new, q, partial, ndp);
                    k_cg_update<VT><<<cur->nb, kBlock, 0, s>>>(nbk, inst, p_new, q, xcg, r, partial, ndp);
                }
            }
            k_solve_end<<<1, 64, 0, s>>>(cur->scenes, 1, inst, 0);
        } else {
            const TileArgs<R, NR> ta = tile_args();
            const FinBlk* fb = (const FinBlk*)cur->b;
            const NodeBlk* nbk = (const NodeBlk*)cur->c;
            const SharedNodes sn{sh_gid, sh_ptr, sh_of, sh_acnt};
            bool exact = false;
            if constexpr (std::is_same<R, double>::value) exact = exact_tiled();
            if (exact) {
                if constexpr (std::is_same<R, double>::value) {
                    if (cur->nc > 0)
                        k_cg_begin_tiled<VT, true><<<cur->nc, kBlock, 0, s>>>(nbk, (const VT*)bvec, nb0, b, r, xcg, prho, nd);
                    k_seqdot<<<1, 64, 0, s>>>(cur->scenes, inst, nd, kDotBB);
                    for (int it = 1; it <= cur->maxit; it++) {
                        if (cur->na > 0) launch_apply(cur->na, s, ta, it, r, p, q, sp0, nullptr, pden, 0);
                        if (cur->nb > 0)
                            k_fin_exact<VT><<<cur->nb, kBlock, 0, s>>>(fb, inst, sn, mass, fixed, it, r, p, q, sp0, nd, contact);
                        k_seqdot<<<1, 64, 0, s>>>(cur->scenes, inst, nd, kDotPQ);
                        if (cur->nc > 0)
                            k_update<VT, true><<<cur->nc, kBlock, 0, s>>>(nbk, inst, ranges_d, it, sn, sp0, mass, fixed, p, q,
                                                                          xcg, r, pden, prho, contact, nd);
                        k_seqdot<<<1, 64, 0, s>>>(cur->scenes, inst, nd, kDotRR);
                    }
                }
            } else {
                if (cur->nc > 0)
                    k_cg_begin_tiled<VT, false><<<cur->nc, kBlock, 0, s>>>(nbk, (const VT*)bvec, nb0, b, r, xcg, prho, nd);
                for (int it = 1; it <= cur->maxit; it++) {
                    if (cur->na > 0) launch_apply(cur->na, s, ta, it, r, p, q, sp0, prho, pden, it == 1 ? 1 : 0);
                    if (cur->nc > 0)
                        k_update<VT><<<cur->nc, kBlock, 0, s>>>(nbk, inst, ranges_d, it, sn, sp0, mass, fixed, p, q, xcg, r,
                                                                pden, prho, contact);
                }
            }
            k_solve_end<<<1, 64, 0, s>>>(cur->scenes, 1, inst, 1);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(xvec, xcg + 3 * (int64_t)nb0, 3 * (size_t)n * sizeof(VT), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        InstDev after;
        HIPCHK(hipMemcpy(&after, inst + sid, sizeof(InstDev), hipMemcpyDeviceToHost));
        if (st) {
            st->iterations = after.last_iters;
            st->end_reason = after.end_reason;
            st->step_count = after.step_count;
            st->ir_sweeps = 0;
            st->err = after.err;
            st->den = after.den;
            st->time = after.time;
            st->dot_fallbacks = 0;
            st->reserved = 0;
        }
        after.m_cg = saved.m_cg;
        after.k_cg = saved.k_cg;
        HIPCHK(hipMemcpy(inst + sid, &after, sizeof(InstDev), hipMemcpyHostToDevice));
    }

    // computeVonMisesStress (TetrahedronFEMForceField.inl:2197-2347) from the scene's current
    // positions: per scene, built on first use, rows 1..3 of the inverse of each element's rest
    // matrix [1 x y z] (init, :1521-1542) and its Lame coefficients (computeMaterialStiffness:
    // K(0,1) and K(3,3) before the division by 36 V, :278-282), in the engine's element order;
    // the rest positions (mode 2); the output buffers
    struct VmData {
        double *shf = nullptr, *lm = nullptr, *x0 = nullptr, *vme = nullptr, *vmn = nullptr;
    };
    std::map<int, VmData> vmd;
    void free_vm(int si) {
        auto it = vmd.find(si);
        if (it == vmd.end()) return;
        for (double* ptr : {it->second.shf, it->second.lm, it->second.x0, it->second.vme, it->second.vmn})
            if (ptr) (void)hipFree(ptr);
        vmd.erase(it);
    }
    // addKToMatrix / buildStiffnessMatrix: per scene, built on first use, the block pattern (node
    // pairs that share an element; rows ascending, columns ascending in a row) and, per block,
    // its (element, n1, n2) contributions in ascending element id
    struct KData {
        std::vector<int32_t> row_ptr, cols;
        int* slot_ptr = nullptr;
        int* slots = nullptr;
        double* vals = nullptr;
    };
    std::map<int, KData> kd;
    void free_k(int si) {
        auto it = kd.find(si);
        if (it == kd.end()) return;
        for (void* ptr : {(void*)it->second.slot_ptr, (void*)it->second.slots, (void*)it->second.vals})
            if (ptr) (void)hipFree(ptr);
        kd.erase(it);
    }
    int64_t stiffness(int scene, double kf, int32_t* row_ptr, int32_t* cols, double* values) override {
        HIPCHK(hipSetDevice(device));
        const SceneHost& sc = scenes[scene];
        const int64_t nE = (int64_t)sc.tets.size() / 4;
        const int nN = sc.n_nodes;
        auto it = kd.find(scene);
        if (it == kd.end()) {
            if (elem_begin(scene) + nE > ((int64_t)1 << 27)) throw std::invalid_argument("too many elements for the block slots");
            KData d;
            std::vector<std::vector<int32_t>> nbr((size_t)nN);
            for (int64_t e = 0; e < nE; e++) {
                const int32_t* t = &sc.tets[4 * (size_t)e];
                for (int a = 0; a < 4; a++)
                    for (int b = 0; b < 4; b++) nbr[t[a]].push_back(t[b]);
            }
            d.row_ptr.assign((size_t)nN + 1, 0);
            for (int n = 0; n < nN; n++) {
                auto& v = nbr[n];
                std::sort(v.begin(), v.end());
                v.erase(std::unique(v.begin(), v.end()), v.end());
                d.row_ptr[n + 1] = d.row_ptr[n] + (int32_t)v.size();
            }
            d.cols.reserve((size_t)d.row_ptr[nN]);
            for (int n = 0; n < nN; n++) d.cols.insert(d.cols.end(), nbr[n].begin(), nbr[n].end());
            nbr.clear();
            const int64_t nb = d.row_ptr[nN];
            std::vector<int> bid(16 * (size_t)nE), sp((size_t)nb + 1, 0), sl(16 * (size_t)nE);
            for (int64_t e = 0; e < nE; e++) {
                const int32_t* t = &sc.tets[4 * (size_t)e];
                for (int a = 0; a < 4; a++)
                    for (int b = 0; b < 4; b++) {
                        const int32_t* c0 = d.cols.data() + d.row_ptr[t[a]];
                        const int32_t* c1 = d.cols.data() + d.row_ptr[t[a] + 1];
                        const int k = (int)(std::lower_bound(c0, c1, t[b]) - d.cols.data());
                        bid[16 * (size_t)e + 4 * a + b] = k;
                        sp[k + 1]++;
                    }
            }
            for (int64_t k = 0; k < nb; k++) sp[k + 1] += sp[k];
            std::vector<int> fill(sp.begin(), sp.end() - 1);
            const std::vector<int>& ip = iperm[scene];
            const int64_t eb0 = elem_begin(scene);
            for (int64_t e = 0; e < nE; e++)
                for (int j = 0; j < 16; j++)
                    sl[fill[bid[16 * (size_t)e + j]]++] = (int)((eb0 + ip[e]) * 16 + j);
            d.slot_ptr = dalloc<int>(sp.size());
            d.slots = dalloc<int>(sl.size());
            d.vals = dalloc<double>(9 * (size_t)nb);
            it = kd.emplace(scene, std::move(d)).first;
            HIPCHK(hipMemcpy(it->second.slot_ptr, sp.data(), sp.size() * sizeof(int), hipMemcpyHostToDevice));
            if (!sl.empty()) HIPCHK(hipMemcpy(it->second.slots, sl.data(), sl.size() * sizeof(int), hipMemcpyHostToDevice));
        }
        KData& d = it->second;
        const int nb = d.row_ptr[nN];
        if (row_ptr) std::copy(d.row_ptr.begin(), d.row_ptr.end(), row_ptr);
        if (cols) std::copy(d.cols.begin(), d.cols.end(), cols);
        if (values && nb > 0) {
            HIPCHK(hipDeviceSynchronize());     // the rotations of the last step / addForce
            // the element blocks (16 per element, 72 B each), for this call only
            const int64_t eb0 = elem_begin(scene), n16 = 16 * nE;
            double* eblk = dalloc<double>(9 * (size_t)n16);
            k_stiffness_elem<R, NR><<<(unsigned)((n16 + kBlock - 1) / kBlock), kBlock, 0, stream>>>(view(), eb0, n16, eblk);
            const int grid = (nb + kBlock - 1) / kBlock;
            k_stiffness_blocks<R, NR><<<grid, kBlock, 0, stream>>>(d.slot_ptr, d.slots, nb, 16 * eb0, kf, eblk, d.vals);
            const hipError_t le = hipGetLastError();
            const hipError_t se = hipStreamSynchronize(stream);
            (void)hipFree(eblk);
            HIPCHK(le);
            HIPCHK(se);
            HIPCHK(hipMemcpy(values, d.vals, 9 * (size_t)nb * sizeof(double), hipMemcpyDeviceToHost));
        }
        return nb;
    }
    void von_mises(int scene, int mode, double* pe, double* pn) override {
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipDeviceSynchronize());
        const SceneHost& sc = scenes[scene];
        const int64_t nE = (int64_t)sc.tets.size() / 4;
        const int nN = sc.n_nodes;
        const std::vector<int>& ip = iperm[scene];
        auto it = vmd.find(scene);
        if (it == vmd.end()) {
            std::vector<double> shf(12 * (size_t)nE), lm(2 * (size_t)nE);
            for (int64_t e = 0; e < nE; e++) {
                const int32_t* t = &sc.tets[4 * (size_t)e];
                double M[4][4], inv[4][4];
                for (int k = 0; k < 4; k++) {
                    M[k][0] = 1.0;
                    for (int l = 1; l < 4; l++) M[k][l] = sc.rest[3 * (size_t)t[k] + l - 1];
                }
                if (!host::invert4(M, inv))
                    throw std::domain_error("rest tetrahedron " + std::to_string(e) + " has no shape functions");
                const size_t pos = (size_t)ip[e];
                for (int l = 0; l < 3; l++)
                    for (int m = 0; m < 4; m++) shf[12 * pos + 4 * l + m] = inv[l + 1][m];
                const double E = sc.young.size() > (size_t)e ? sc.young[e] : sc.young[0];
                const double nu = sc.poisson.size() > (size_t)e ? sc.poisson[e] : sc.poisson[0];
                double ks[3];
                host::material_factors(E, nu, ks);
                lm[2 * pos] = ks[1];
                lm[2 * pos + 1] = ks[2];
            }
            VmData d;
            d.shf = dalloc<double>(shf.size());
            d.lm = dalloc<double>(lm.size());
            d.x0 = dalloc<double>(3 * (size_t)nN);
            d.vme = dalloc<double>((size_t)nE);
            d.vmn = dalloc<double>((size_t)nN);
            it = vmd.emplace(scene, d).first;
            if (nE > 0) {
                HIPCHK(hipMemcpy(d.shf, shf.data(), shf.size() * sizeof(double), hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(d.lm, lm.data(), lm.size() * sizeof(double), hipMemcpyHostToDevice));
            }
            if (nN > 0) HIPCHK(hipMemcpy(d.x0, sc.rest.data(), 3 * (size_t)nN * sizeof(double), hipMemcpyHostToDevice));
        }
        const VmData& d = it->second;
        const int nb0 = node_begin(scene);
        const int64_t eb0 = elem_begin(scene);
        if (nE > 0) {
            const int grid = (int)((nE + kBlock - 1) / kBlock);
            // x of the scene (node ids of the element record minus nb0 index it), as add_force's
            k_von_mises_elem<R, NR, VT><<<grid, kBlock, 0, stream>>>(view(), eb0, eb0 + nE, nb0, x + 3 * (int64_t)nb0,
                                                                     d.x0, d.shf, d.lm, mode, d.vme);
        }
        const int gn = (nN + kBlock - 1) / kBlock;
        if (gn > 0) k_von_mises_node<<<gn, kBlock, 0, stream>>>(inc_ptr, inc, nb0, nb0 + nN, eb0, d.vme, d.vmn);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
        if (pe && nE > 0) {
            std::vector<double> tmp((size_t)nE);
            HIPCHK(hipMemcpy(tmp.data(), d.vme, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (int64_t e = 0; e < nE; e++) pe[e] = tmp[(size_t)ip[e]];
        }
        if (pn && nN > 0) HIPCHK(hipMemcpy(pn, d.vmn, (size_t)nN * sizeof(double), hipMemcpyDeviceToHost));
    }
};

#if defined(SOFA_AMD_TU_FP32)
EngineBase* make_engine_fp32() {
    auto* e = new EngineT<float, 6, float>();
    e->tiled = true;
    return e;
}

// FP32 element record and element arithmetic on FP64 node vectors (state, CG vectors, per-node
// sums) with FP64 CG scalars: the tiled kernels instantiated with R = float, VT = double
EngineBase* make_engine_split() {
    auto* e = new EngineT<float, 6, double>();
    e->tiled = true;
    return e;
}

void launch_inner_cg(const InnerCg& c, hipStream_t s) {
    const float* rs = c.z ? c.z : c.r;   // the vector p is formed from (p = rs + beta p)
    for (int it = 1; it <= c.maxit; it++) {
        if (c.na > 0)
        {
            if (c.wide && c.plane)
                k_tile_apply<float, 6, float, kTileThreads, 2, true, true><<<c.na, kTileThreads, 0, s>>>(
                    c.ta, it, rs, c.p, c.q, c.sp, c.prho, c.pden, 1, c.x, c.x0);
            else if (c.wide)
                k_tile_apply<float, 6, float, kTileThreads, 2, true, false><<<c.na, kTileThreads, 0, s>>>(
                    c.ta, it, rs, c.p, c.q, c.sp, c.prho, c.pden, 1, c.x, c.x0);
            else if (c.plane)
                k_tile_apply<float, 6, float, kTileThreads, 2, false, true><<<c.na, kTileThreads, 0, s>>>(
                    c.ta, it, rs, c.p, c.q, c.sp, c.prho, c.pden, 1, c.x, c.x0);
            else
                k_tile_apply<float, 6, float, kTileThreads, 2, false, false><<<c.na, kTileThreads, 0, s>>>(
                    c.ta, it, rs, c.p, c.q, c.sp, c.prho, c.pden, 1, c.x, c.x0);
        }
        if (c.nnb > 0)
            k_update<float><<<c.nnb, kBlock, 0, s>>>(c.nb, c.ta.inst, c.ta.ranges, it, c.sn, c.sp, c.ta.mass, c.ta.fixed,
                                                      c.p, c.q, c.d, c.r, c.pden, c.prho, c.ta.contact, nullptr,
                                                      c.dinv, c.binv, c.z);
    }
}
#else
EngineBase* make_engine_fp32();
EngineBase* make_engine_split();
#endif

}  // namespace sofa_amd

#if !defined(SOFA_AMD_TU_FP32)
// =========================================================================================
// C ABI
// =========================================================================================
using namespace sofa_amd;

struct sofa_amd_engine_s {
    std::unique_ptr<EngineBase> impl;
};

#define GUARD_BEGIN try {
#define GUARD_END                                                                                 \
    }                                                                                             \
    catch (const HipError& e) { return fail(SOFA_AMD_EDEVICE, e.what()); }                        \
    catch (const std::domain_error& e) { return fail(SOFA_AMD_EDEGENERATE, e.what()); }           \
    catch (const std::invalid_argument& e) { return fail(SOFA_AMD_EINVAL, e.what()); }            \
    catch (const std::bad_alloc& e) { return fail(SOFA_AMD_ENOMEM, "host allocation failed"); }   \
    catch (const std::exception& e) { return fail(SOFA_AMD_EINVAL, e.what()); }

static SceneHost* scene_of(sofa_amd_engine eng, int scene) {
    if (!eng || !eng->impl) return nullptr;
    if (scene < 0 || scene >= (int)eng->impl->scenes.size()) return nullptr;
    return &eng->impl->scenes[scene];
}

// a scene's parameters can change until it is placed on the device (engine commit, or its own
// commit on a live engine)
static bool editable(sofa_amd_engine eng, int scene) {
    EngineBase* e = eng->impl.get();
    if (e->live_mode) return !e->placed(scene);
    return !e->committed;
}

// a scene can be stepped / read once it is placed
static bool usable(sofa_amd_engine eng, int scene) {
    EngineBase* e = eng->impl.get();
    return e->committed && e->placed(scene);
}

extern "C" {

const char* sofa_amd_module_name(void) { return "SofaMI355X"; }
const char* sofa_amd_module_version(void) { return "0.1.0"; }
const char* sofa_amd_last_error(void) { return g_last_error.c_str(); }

int sofa_amd_device_check(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0)
        return fail(SOFA_AMD_EDEVICE, "no HIP device " + std::to_string(device));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(SOFA_AMD_EDEVICE, "hipGetDeviceProperties failed");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SOFA_AMD_EDEVICE, std::string("device is ") + prop.gcnArchName + ", not gfx950");
    return SOFA_AMD_OK;
}

int sofa_amd_device_memory(int device, int64_t* free_bytes, int64_t* total_bytes) {
    GUARD_BEGIN
    if (!free_bytes || !total_bytes) return fail(SOFA_AMD_EINVAL, "null output");
    const int rc = sofa_amd_device_check(device);
    if (rc != SOFA_AMD_OK) return rc;
    HIPCHK(hipSetDevice(device));
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    *free_bytes = (int64_t)fr;
    *total_bytes = (int64_t)tot;
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_engine_create(int device, int precision, sofa_amd_engine* out) {
    GUARD_BEGIN
    if (!out) return fail(SOFA_AMD_EINVAL, "null output");
    int rc = sofa_amd_device_check(device);
    if (rc != SOFA_AMD_OK) return rc;
    std::unique_ptr<EngineBase> impl;
    if (precision == SOFA_AMD_FP64) {
        impl.reset(new EngineT<double, 9, double>());
    } else if (precision == SOFA_AMD_FP64_TILED) {
        auto* e = new EngineT<double, 9, double>();
        e->tiled = true;
        impl.reset(e);
    } else if (precision == SOFA_AMD_FP32) {
        impl.reset(make_engine_fp32());
    } else if (precision == SOFA_AMD_SPLIT) {
        impl.reset(make_engine_split());
    } else if (precision == SOFA_AMD_MIXED) {
        auto* e = new EngineT<double, 9, double>();
        e->tiled = true;
        e->mixed = true;
        impl.reset(e);
    } else {
        return fail(SOFA_AMD_EINVAL, "unsupported precision mode");
    }
    impl->device = device;
    impl->precision = precision;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&impl->stream, hipStreamNonBlocking));
    auto* e = new sofa_amd_engine_s();
    e->impl = std::move(impl);
    *out = e;
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_engine_destroy(sofa_amd_engine eng) {
    GUARD_BEGIN
    if (!eng) return SOFA_AMD_OK;
    if (eng->impl) {
        (void)hipSetDevice(eng->impl->device);
        (void)hipDeviceSynchronize();
        if (eng->impl->stream) (void)hipStreamDestroy(eng->impl->stream);
    }
    delete eng;
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_scene_create(sofa_amd_engine eng, int n_nodes, const double* positions, int* out_scene) {
    GUARD_BEGIN
    if (!eng || !eng->impl || !positions || !out_scene || n_nodes <= 0) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (eng->impl->committed && !eng->impl->live_mode) return fail(SOFA_AMD_ESTATE, "engine already committed");
    SceneHost s;
    s.n_nodes = n_nodes;
    s.rest.assign(positions, positions + 3 * (size_t)n_nodes);
    for (double d : s.rest)
        if (!std::isfinite(d)) return fail(SOFA_AMD_EINVAL, "non-finite position");
    {
        std::lock_guard<std::mutex> g(eng->impl->scenes_mu);
        eng->impl->scenes.push_back(std::move(s));
        *out_scene = (int)eng->impl->scenes.size() - 1;
    }
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_scene_set_fem(sofa_amd_engine eng, int scene, int n_tets, const int32_t* tets, const double* young,
                           int n_young, const double* poisson, int n_poisson) {
    GUARD_BEGIN
    SceneHost* s = scene_of(eng, scene);
    if (!s || n_tets < 0 || (n_tets > 0 && !tets)) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!editable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene already committed");
    for (int i = 0; i < 4 * n_tets; i++)
        if (tets[i] < 0 || tets[i] >= s->n_nodes) return fail(SOFA_AMD_EINVAL, "tetrahedron index out of range");
    s->tets.assign(tets, tets + 4 * (size_t)n_tets);
    if (young && n_young > 0) s->young.assign(young, young + n_young);
    if (poisson && n_poisson > 0) s->poisson.assign(poisson, poisson + n_poisson);
    // BaseLinearElasticityFEMForceField::checkYoungModulus / checkPoissonRatio (:49-84)
    for (double& y : s->young) if (y < 0) y = 5000;
    for (double& p : s->poisson) if (p < 0 || p >= 0.5) p = 0.45;
    return SOFA_AMD_OK;
    GUARD_END
}

// TetrahedronFEMForceField::init on a topology with hexahedra and no tetrahedra
// (TetrahedronFEMForceField.inl:1318-1370): every cube becomes 6 tetrahedra, its corners first
// mirrored across x where (i mod nx) is even, across y where floor(i / nx) mod ny is odd and
// across z where floor(i / (nx ny)) is odd (the Hexahedron x/y/z edge pairs,
// sofa/geometry/Hexahedron.h:67-88), so that neighbouring cubes share their face diagonals.
int sofa_amd_scene_set_fem_hexa(sofa_amd_engine eng, int scene, int n_hexa, const int32_t* hexa, int nx, int ny,
                                const double* young, int n_young, const double* poisson, int n_poisson) {
    GUARD_BEGIN
    SceneHost* s = scene_of(eng, scene);
    if (!s || n_hexa < 0 || (n_hexa > 0 && !hexa) || nx < 1 || ny < 1) return fail(SOFA_AMD_EINVAL, "invalid argument");
    static const int xe[4][2] = {{0, 1}, {4, 5}, {3, 2}, {7, 6}};
    static const int ye[4][2] = {{4, 7}, {5, 6}, {1, 2}, {0, 3}};
    static const int ze[4][2] = {{4, 0}, {5, 1}, {6, 2}, {7, 3}};
    static const int six[6][4] = {{0, 5, 1, 6}, {0, 1, 3, 6}, {1, 3, 6, 2}, {6, 3, 0, 7}, {6, 7, 0, 5}, {7, 5, 4, 0}};
    std::vector<int32_t> t(24 * (size_t)n_hexa);
    for (int i = 0; i < n_hexa; i++) {
        int32_t c[8];
        for (int k = 0; k < 8; k++) c[k] = hexa[8 * (size_t)i + k];
        if (!((i % nx) & 1))
            for (const auto& e : xe) std::swap(c[e[0]], c[e[1]]);
        if (((i / nx) % ny) & 1)
            for (const auto& e : ye) std::swap(c[e[0]], c[e[1]]);
        if ((i / (nx * ny)) & 1)
            for (const auto& e : ze) std::swap(c[e[0]], c[e[1]]);
        for (int k = 0; k < 6; k++)
            for (int v = 0; v < 4; v++) t[24 * (size_t)i + 4 * k + v] = c[six[k][v]];
    }
    return sofa_amd_scene_set_fem(eng, scene, 6 * n_hexa, t.data(), young, n_young, poisson, n_poisson);
    GUARD_END
}

int sofa_amd_scene_set_mass(sofa_amd_engine eng, int scene, int kind, double value) {
    SceneHost* s = scene_of(eng, scene);
    if (!s || kind < SOFA_AMD_MASS_UNIFORM || kind > SOFA_AMD_MASS_DIAGONAL_TOTAL || !(value > 0))
        return fail(SOFA_AMD_EINVAL, "invalid mass");
    if (!editable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene already committed");
    s->mass_total = kind == SOFA_AMD_MASS_DIAGONAL_TOTAL;
    if (s->mass_total) kind = SOFA_AMD_MASS_DIAGONAL;
    s->mass_kind = kind;
    s->mass_value = value;
    s->density.assign(1, value);
    return SOFA_AMD_OK;
}

int sofa_amd_scene_set_mass_density(sofa_amd_engine eng, int scene, int kind, const double* density, int n) {
    SceneHost* s = scene_of(eng, scene);
    if (!s || (kind != SOFA_AMD_MASS_MESHMATRIX && kind != SOFA_AMD_MASS_MESHMATRIX_LUMPED) || !density || n <= 0)
        return fail(SOFA_AMD_EINVAL, "invalid mass density");
    if (!editable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene already committed");
    for (int i = 0; i < n; i++)
        if (!(density[i] > 0)) return fail(SOFA_AMD_EINVAL, "non-positive mass density");
    s->mass_kind = kind;
    s->mass_value = density[0];
    s->density.assign(density, density + n);
    return SOFA_AMD_OK;
}

int sofa_amd_scene_add_oscillators(sofa_amd_engine eng, int scene, int n, const int32_t* indices, const double* mean,
                                   const double* amplitude, const double* pulsation, const double* phase) {
    SceneHost* s = scene_of(eng, scene);
    if (!s || n < 0 || (n > 0 && (!indices || !mean || !amplitude || !pulsation || !phase)))
        return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!editable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene already committed");
    for (int i = 0; i < n; i++) {
        if (indices[i] < 0 || indices[i] >= s->n_nodes) return fail(SOFA_AMD_EINVAL, "oscillator index out of range");
        s->osc_idx.push_back(indices[i]);
        for (int k = 0; k < 3; k++) { s->osc_mean.push_back(mean[3 * i + k]); s->osc_amp.push_back(amplitude[3 * i + k]); }
        s->osc_w.push_back(pulsation[i]);
        s->osc_phase.push_back(phase[i]);
    }
    return SOFA_AMD_OK;
}

int sofa_amd_scene_add_linear_movement(sofa_amd_engine eng, int scene, int n, const int32_t* indices, int n_keys,
                                       const double* key_times, const double* movements, int relative) {
    SceneHost* s = scene_of(eng, scene);
    if (!s || n < 0 || n_keys < 1 || (n > 0 && !indices) || !key_times || !movements)
        return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!editable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene already committed");
    if (s->lm_on) return fail(SOFA_AMD_EINVAL, "one LinearMovementProjectiveConstraint per scene");
    for (int i = 0; i < n; i++)
        if (indices[i] < 0 || indices[i] >= s->n_nodes) return fail(SOFA_AMD_EINVAL, "movement index out of range");
    for (int k = 0; k < n_keys; k++) {
        if (!std::isfinite(key_times[k])) return fail(SOFA_AMD_EINVAL, "non-finite key time");
        if (k > 0 && key_times[k] < key_times[k - 1])
            return fail(SOFA_AMD_EINVAL, "key times must be in chronological order");
    }
    s->lm_on = true;
    s->lm_relative = relative != 0;
    s->lm_idx.assign(indices, indices + n);
    s->lm_times.assign(key_times, key_times + n_keys);
    s->lm_moves.assign(movements, movements + 3 * (size_t)n_keys);
    return SOFA_AMD_OK;
}

int sofa_amd_scene_set_preconditioner(sofa_amd_engine eng, int scene, int kind) {
    SceneHost* s = scene_of(eng, scene);
    if (!s || kind < 0 || kind > 2) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!editable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene already committed");
    if (kind != 0 && eng->impl->precision != SOFA_AMD_FP64 && eng->impl->precision != SOFA_AMD_MIXED)
        return fail(SOFA_AMD_EINVAL, "PCGLinearSolver runs on the FP64 engine (and in the MIXED inner CG)");
    s->precond = kind;
    return SOFA_AMD_OK;
}

int sofa_amd_scene_set_warm_start(sofa_amd_engine eng, int scene, int on) {
    SceneHost* s = scene_of(eng, scene);
    if (!s) return fail(SOFA_AMD_EINVAL, "invalid scene");
    if (!editable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene already committed");
    if (on && eng->impl->precision != SOFA_AMD_FP64 &&
        !(eng->impl->precision == SOFA_AMD_FP64_TILED && eng->impl->seq_dots))
        return fail(SOFA_AMD_EINVAL, "CGLinearSolver warmStart runs on the FP64 engine, or on FP64_TILED in the "
                                     "reference order (sofa_amd_engine_set_dot_order first)");
    s->warm_start = on != 0;
    return SOFA_AMD_OK;
}

int sofa_amd_scene_set_gravity(sofa_amd_engine eng, int scene, double gx, double gy, double gz) {
    SceneHost* s = scene_of(eng, scene);
    if (!s) return fail(SOFA_AMD_EINVAL, "invalid scene");
    if (!editable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene already committed");
    s->g[0] = gx; s->g[1] = gy; s->g[2] = gz;
    return SOFA_AMD_OK;
}

int sofa_amd_scene_set_fixed(sofa_amd_engine eng, int scene, const int32_t* indices, int n) {
    SceneHost* s = scene_of(eng, scene);
    if (!s || n < 0 || (n > 0 && !indices)) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!editable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene already committed");
    for (int i = 0; i < n; i++)
        if (indices[i] < 0 || indices[i] >= s->n_nodes) return fail(SOFA_AMD_EINVAL, "fixed index out of range");
    s->fixed.assign(indices, indices + n);
    return SOFA_AMD_OK;
}

int sofa_amd_scene_set_fix_all(sofa_amd_engine eng, int scene, int fix_all) {
    SceneHost* s = scene_of(eng, scene);
    if (!s) return fail(SOFA_AMD_EINVAL, "invalid scene");
    if (!editable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene already committed");
    s->fix_all = fix_all != 0;
    return SOFA_AMD_OK;
}

int sofa_amd_scene_set_solver(sofa_amd_engine eng, int scene, double dt, double rs, double rm, int iters,
                              double tol, double thr) {
    SceneHost* s = scene_of(eng, scene);
    if (!s || !(dt > 0) || iters < 0) return fail(SOFA_AMD_EINVAL, "invalid solver parameters");
    if (!editable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene already committed");
    s->dt = dt; s->rs = rs; s->rm = rm; s->cg_iter = iters; s->tol = tol; s->thr = thr;
    return SOFA_AMD_OK;
}

int sofa_amd_scene_add_plane(sofa_amd_engine eng, int scene, const double* normal, double d, double stiffness,
                             double damping, double max_force, int bilateral) {
    SceneHost* s = scene_of(eng, scene);
    if (!s || !normal) return fail(SOFA_AMD_EINVAL, "invalid argument");
    for (double v : {normal[0], normal[1], normal[2], d, stiffness, damping, max_force})
        if (!std::isfinite(v)) return fail(SOFA_AMD_EINVAL, "invalid plane parameters");
    if (!editable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene already committed");
    if (s->plane_on) return fail(SOFA_AMD_EINVAL, "one PlaneForceField per scene");
    s->plane_on = true;
    s->plane_bilateral = bilateral != 0;
    for (int k = 0; k < 3; k++) s->plane_n[k] = normal[k];
    s->plane_d = d; s->plane_k = stiffness; s->plane_damp = damping; s->plane_max = max_force;
    return SOFA_AMD_OK;
}

int sofa_amd_scene_set_project_velocity(sofa_amd_engine eng, int scene, int on) {
    SceneHost* s = scene_of(eng, scene);
    if (!s) return fail(SOFA_AMD_EINVAL, "invalid scene");
    if (!editable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene already committed");
    s->project_velocity = on != 0;
    return SOFA_AMD_OK;
}

int sofa_amd_scene_set_integrator(sofa_amd_engine eng, int scene, int first_order, int trapezoidal, double vdamping) {
    SceneHost* s = scene_of(eng, scene);
    if (!s || !std::isfinite(vdamping)) return fail(SOFA_AMD_EINVAL, "invalid integrator parameters");
    if (!editable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene already committed");
    s->first_order = first_order != 0;
    s->trapezoidal = trapezoidal != 0;
    s->vdamping = vdamping;
    return SOFA_AMD_OK;
}

int sofa_amd_scene_set_refinement(sofa_amd_engine eng, int scene, int inner_iterations, double inner_tolerance,
                                  int max_sweeps, int keep_direction) {
    SceneHost* s = scene_of(eng, scene);
    if (!s || inner_iterations < 1 || max_sweeps < 1 || !(inner_tolerance >= 0))
        return fail(SOFA_AMD_EINVAL, "invalid refinement parameters");
    if (!editable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene already committed");
    s->inner_iter = inner_iterations;
    s->inner_tol = inner_tolerance;
    s->max_sweeps = max_sweeps;
    s->keep_dir = keep_direction != 0;
    return SOFA_AMD_OK;
}

int sofa_amd_engine_commit(sofa_amd_engine eng) {
    GUARD_BEGIN
    if (!eng || !eng->impl) return fail(SOFA_AMD_EINVAL, "null engine");
    if (eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine already committed");
    if (eng->impl->scenes.empty()) return fail(SOFA_AMD_EINVAL, "no scene");
    if (eng->impl->live_mode) return fail(SOFA_AMD_ESTATE, "live engine: commit scenes one by one");
    eng->impl->commit();
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_engine_step(sofa_amd_engine eng, const int32_t* scenes, int n_scenes, int n_steps, void* stream) {
    GUARD_BEGIN
    if (!eng || !eng->impl || !scenes || n_scenes <= 0 || n_steps < 0) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine not committed");
    for (int i = 0; i < n_scenes; i++) {
        if (scenes[i] < 0 || scenes[i] >= (int)eng->impl->scenes.size()) return fail(SOFA_AMD_EINVAL, "invalid scene id");
        if (!eng->impl->placed(scenes[i])) return fail(SOFA_AMD_ESTATE, "scene not committed");
        for (int j = 0; j < i; j++)
            if (scenes[j] == scenes[i]) return fail(SOFA_AMD_EINVAL, "duplicate scene id");
    }
    HIPCHK(hipSetDevice(eng->impl->device));
    hipStream_t s = stream ? (hipStream_t)stream : eng->impl->stream;
    eng->impl->step(scenes, n_scenes, n_steps, s);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_engine_sync(sofa_amd_engine eng) {
    GUARD_BEGIN
    if (!eng || !eng->impl) return fail(SOFA_AMD_EINVAL, "null engine");
    HIPCHK(hipSetDevice(eng->impl->device));
    HIPCHK(hipDeviceSynchronize());
    return SOFA_AMD_OK;
    GUARD_END
}

static bool all_finite(const double* a, size_t n) {
    for (size_t i = 0; i < n; i++)
        if (!std::isfinite(a[i])) return false;
    return true;
}

int sofa_amd_scene_read(sofa_amd_engine eng, int scene, double* x, double* v, double* f) {
    GUARD_BEGIN
    if (!scene_of(eng, scene)) return fail(SOFA_AMD_EINVAL, "invalid scene");
    if (!usable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene not committed");
    eng->impl->read(scene, x, v, f);
    // the buffers are filled either way; a state that has left the finite numbers is reported
    const size_t n3 = 3 * (size_t)scene_of(eng, scene)->n_nodes;
    for (const double* a : {(const double*)x, (const double*)v, (const double*)f})
        if (a && !all_finite(a, n3)) return fail(SOFA_AMD_ENAN, "non-finite state (NaN or Inf) in the scene read back");
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_scene_write(sofa_amd_engine eng, int scene, const double* x, const double* v) {
    GUARD_BEGIN
    if (!scene_of(eng, scene)) return fail(SOFA_AMD_EINVAL, "invalid scene");
    if (!usable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene not committed");
    const size_t n3 = 3 * (size_t)scene_of(eng, scene)->n_nodes;
    if ((x && !all_finite(x, n3)) || (v && !all_finite(v, n3)))
        return fail(SOFA_AMD_ENAN, "non-finite position or velocity written (NaN or Inf)");
    eng->impl->write(scene, x, v);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_scene_stats(sofa_amd_engine eng, int scene, sofa_amd_cg_stats* out) {
    GUARD_BEGIN
    if (!scene_of(eng, scene) || !out) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!usable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene not committed");
    eng->impl->stats(scene, out);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_scene_element(sofa_amd_engine eng, int scene, int elem, double* rot9, double* k3, double* rest12) {
    GUARD_BEGIN
    SceneHost* s = scene_of(eng, scene);
    if (!s || !rot9 || !k3 || !rest12 || elem < 0 || elem >= (int)(s->tets.size() / 4))
        return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!usable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene not committed");
    eng->impl->element(scene, elem, rot9, k3, rest12);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_scene_von_mises(sofa_amd_engine eng, int scene, int mode, double* per_element, double* per_node) {
    GUARD_BEGIN
    SceneHost* s = scene_of(eng, scene);
    if (!s || (mode != 1 && mode != 2)) return fail(SOFA_AMD_EINVAL, "invalid argument (mode must be 1 or 2)");
    if (!usable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene not committed");
    eng->impl->von_mises(scene, mode, per_element, per_node);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_scene_stiffness(sofa_amd_engine eng, int scene, double k_factor, int32_t* row_ptr, int32_t* cols,
                             double* values, int64_t* nnzb) {
    GUARD_BEGIN
    SceneHost* s = scene_of(eng, scene);
    if (!s || !nnzb) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!usable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene not committed");
    if (!std::isfinite(k_factor)) return fail(SOFA_AMD_ENAN, "k_factor is not finite");
    *nnzb = eng->impl->stiffness(scene, k_factor, row_ptr, cols, values);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_fem_add_force(sofa_amd_engine eng, int scene, const void* x_dev, void* f_dev, void* stream) {
    GUARD_BEGIN
    if (!scene_of(eng, scene) || !x_dev || !f_dev) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!usable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene not committed");
    HIPCHK(hipSetDevice(eng->impl->device));
    eng->impl->add_force(scene, x_dev, f_dev, stream ? (hipStream_t)stream : eng->impl->stream);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_fem_add_dforce(sofa_amd_engine eng, int scene, const void* dx_dev, void* df_dev, double k_factor,
                            void* stream) {
    GUARD_BEGIN
    if (!scene_of(eng, scene) || !dx_dev || !df_dev) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!usable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene not committed");
    HIPCHK(hipSetDevice(eng->impl->device));
    eng->impl->add_dforce(scene, dx_dev, df_dev, k_factor, stream ? (hipStream_t)stream : eng->impl->stream);
    return SOFA_AMD_OK;
    GUARD_END
}

// ---- component-level calls of the masses, the projective constraints, the state and the CG ----
static int mass_call(sofa_amd_engine eng, int scene, int what, const void* in, void* out, double factor, void* stream) {
    GUARD_BEGIN
    if (!scene_of(eng, scene) || !out || (what < 2 && !in)) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!usable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene not committed");
    HIPCHK(hipSetDevice(eng->impl->device));
    eng->impl->mass_call(scene, what, in, out, factor, stream ? (hipStream_t)stream : eng->impl->stream);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_mass_add_mdx(sofa_amd_engine eng, int scene, const void* dx_dev, void* f_dev, double factor, void* stream) {
    return mass_call(eng, scene, 0, dx_dev, f_dev, factor, stream);
}

int sofa_amd_mass_acc_from_f(sofa_amd_engine eng, int scene, const void* f_dev, void* a_dev, void* stream) {
    return mass_call(eng, scene, 1, f_dev, a_dev, 1.0, stream);
}

int sofa_amd_mass_add_force(sofa_amd_engine eng, int scene, void* f_dev, void* stream) {
    return mass_call(eng, scene, 2, nullptr, f_dev, 1.0, stream);
}

static int project_call(sofa_amd_engine eng, int scene, int what, void* vec, double t, void* stream) {
    GUARD_BEGIN
    if (!scene_of(eng, scene) || !vec) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!usable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene not committed");
    HIPCHK(hipSetDevice(eng->impl->device));
    eng->impl->project(scene, what, vec, t, stream ? (hipStream_t)stream : eng->impl->stream);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_project_response(sofa_amd_engine eng, int scene, void* dx_dev, void* stream) {
    return project_call(eng, scene, 0, dx_dev, 0.0, stream);
}

int sofa_amd_project_velocity(sofa_amd_engine eng, int scene, void* v_dev, double time, void* stream) {
    return project_call(eng, scene, 1, v_dev, time, stream);
}

int sofa_amd_project_position(sofa_amd_engine eng, int scene, void* x_dev, double time, void* stream) {
    return project_call(eng, scene, 2, x_dev, time, stream);
}

int sofa_amd_vdot(sofa_amd_engine eng, int scene, const void* a_dev, const void* b_dev, int order, double* result_dev,
                  void* stream) {
    GUARD_BEGIN
    if (!scene_of(eng, scene) || !a_dev || !b_dev || !result_dev || order < 0 || order > 1)
        return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!usable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene not committed");
    EngineBase* e = eng->impl.get();
    HIPCHK(hipSetDevice(e->device));
    const hipStream_t s = stream ? (hipStream_t)stream : e->stream;
    const int64_t need = (e->scene_nodes(scene) + kBlock - 1) / kBlock;
    if (order == 0 && need > e->vdot_len) {
        // the partials buffer grows to the largest scene asked for (no call in flight uses it
        // after the device-wide wait)
        HIPCHK(hipDeviceSynchronize());
        if (e->vdot_part) HIPCHK(hipFree(e->vdot_part));
        e->vdot_part = nullptr;
        e->vdot_len = 0;
        e->vdot_part = dalloc<double>((size_t)need);
        e->vdot_len = need;
    }
    e->vdot(scene, a_dev, b_dev, order, result_dev, e->vdot_part, e->vdot_len, s);
    return SOFA_AMD_OK;
    GUARD_END
}

}  // extern "C"

// ---- the allocation-free vDot: partial sums in a file-scope device array ---------------------
// kVdotScratchMax doubles (512 KiB per GPU, held until the process ends): the block partials of
// a scene of up to kVdotScratchMax * kBlock = 16,777,216 nodes.  In host code the array's name is
// not a device address: each device's copy is reached through the address hipGetSymbolAddress
// returns with that device current, taken once per device on its first call and kept.
constexpr int64_t kVdotScratchMax = 65536;
static __device__ __align__(256) double scratch[kVdotScratchMax];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// the device's copy of `scratch` (the device must be current)
static double* scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device index out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = (double*)ptr;
    }
    return g_scratch_ptr[device];
}

extern "C" {

int sofa_amd_vdot_noalloc(sofa_amd_engine eng, int scene, const void* a_dev, const void* b_dev, int order,
                          double* result_dev, void* stream) {
    GUARD_BEGIN
    if (!scene_of(eng, scene) || !a_dev || !b_dev || !result_dev || order < 0 || order > 1)
        return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!usable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene not committed");
    EngineBase* e = eng->impl.get();
    HIPCHK(hipSetDevice(e->device));
    const hipStream_t s = stream ? (hipStream_t)stream : e->stream;
    const int64_t n = e->scene_nodes(scene);
    const int64_t need = (n + kBlock - 1) / kBlock;
    if (need > kVdotScratchMax)
        return fail(SOFA_AMD_EINVAL, "vDot without allocation: the scene has " + std::to_string(n) +
                                         " nodes, more than the " + std::to_string(kVdotScratchMax * kBlock) +
                                         " its static partials buffer covers");
    double* part = scratch_on(e->device);
    e->vdot(scene, a_dev, b_dev, order, result_dev, part, kVdotScratchMax, s);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_cg_solve(sofa_amd_engine eng, int scene, double m_factor, double b_factor, double k_factor,
                      const void* b_dev, void* x_dev, sofa_amd_cg_stats* stats, void* stream) {
    GUARD_BEGIN
    if (!scene_of(eng, scene) || !b_dev || !x_dev) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!usable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene not committed");
    HIPCHK(hipSetDevice(eng->impl->device));
    eng->impl->cg_solve(scene, m_factor, b_factor, k_factor, b_dev, x_dev, stats,
                        stream ? (hipStream_t)stream : eng->impl->stream);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_fem_potential_energy(sofa_amd_engine eng, int scene, const void* x_dev, double* energy) {
    GUARD_BEGIN
    if (!scene_of(eng, scene) || !x_dev || !energy) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!usable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene not committed");
    *energy = 0.0;   // method "large": TetrahedronFEMForceField.inl:625-717 sums the small method only
    return SOFA_AMD_OK;
    GUARD_END
}

// ---- runtime plumbing: streams, events and device memory of the library's HIP runtime, for
// callers that have no device-array library of their own (the Python driver, tests, benchmark) ----
int sofa_amd_stream_create(int device, void** out) {
    GUARD_BEGIN
    if (!out) return fail(SOFA_AMD_EINVAL, "null output");
    HIPCHK(hipSetDevice(device));
    hipStream_t s = nullptr;
    HIPCHK(hipStreamCreate(&s));          // default priority
    *out = (void*)s;
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_stream_destroy(void* stream) {
    GUARD_BEGIN
    if (!stream) return fail(SOFA_AMD_EINVAL, "null stream");
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    HIPCHK(hipStreamDestroy((hipStream_t)stream));
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_stream_sync(void* stream) {
    GUARD_BEGIN
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_device_sync(int device) {
    GUARD_BEGIN
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipDeviceSynchronize());
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_event_create(void** out) {
    GUARD_BEGIN
    if (!out) return fail(SOFA_AMD_EINVAL, "null output");
    hipEvent_t e = nullptr;
    HIPCHK(hipEventCreate(&e));
    *out = (void*)e;
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_event_destroy(void* event) {
    GUARD_BEGIN
    if (!event) return fail(SOFA_AMD_EINVAL, "null event");
    HIPCHK(hipEventDestroy((hipEvent_t)event));
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_event_record(void* event, void* stream) {
    GUARD_BEGIN
    if (!event) return fail(SOFA_AMD_EINVAL, "null event");
    HIPCHK(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_event_sync(void* event) {
    GUARD_BEGIN
    if (!event) return fail(SOFA_AMD_EINVAL, "null event");
    HIPCHK(hipEventSynchronize((hipEvent_t)event));
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_event_elapsed(void* start, void* end, float* ms) {
    GUARD_BEGIN
    if (!start || !end || !ms) return fail(SOFA_AMD_EINVAL, "invalid argument");
    HIPCHK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_malloc(int device, int64_t bytes, void** out) {
    GUARD_BEGIN
    if (!out || bytes < 0) return fail(SOFA_AMD_EINVAL, "invalid argument");
    HIPCHK(hipSetDevice(device));
    void* p = nullptr;
    if (hipMalloc(&p, (size_t)std::max<int64_t>(bytes, 1)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(SOFA_AMD_ENOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    *out = p;
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_free(void* ptr) {
    GUARD_BEGIN
    if (ptr) HIPCHK(hipFree(ptr));
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_memcpy(void* dst, const void* src, int64_t bytes, int kind) {
    GUARD_BEGIN
    if ((!dst || !src) && bytes > 0) return fail(SOFA_AMD_EINVAL, "null pointer");
    if (bytes < 0 || kind < 0 || kind > 2) return fail(SOFA_AMD_EINVAL, "invalid argument");
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    if (bytes > 0) HIPCHK(hipMemcpy(dst, src, (size_t)bytes, k));
    return SOFA_AMD_OK;
    GUARD_END
}

// a plain device copy kernel: each thread moves 16-byte words, grid-strided
static __global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ a, uint4* __restrict__ b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) b[i] = a[i];
}

int sofa_amd_copy_bench(int device, int64_t bytes, int reps, void* stream, float* us_per_copy) {
    GUARD_BEGIN
    if (bytes <= 0 || bytes % 16 != 0 || reps < 1 || !us_per_copy) return fail(SOFA_AMD_EINVAL, "invalid argument");
    HIPCHK(hipSetDevice(device));
    const hipStream_t s = (hipStream_t)stream;
    void *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, (size_t)bytes));
    HIPCHK(hipMalloc(&b, (size_t)bytes));
    HIPCHK(hipMemset(a, 1, (size_t)bytes));
    const int64_t n = bytes / 16;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 256 * 64);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    k_copy<<<grid, 256, 0, s>>>((const uint4*)a, (uint4*)b, n);
    HIPCHK(hipEventRecord(e0, s));
    for (int i = 0; i < reps; i++) k_copy<<<grid, 256, 0, s>>>((const uint4*)a, (uint4*)b, n);
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *us_per_copy = 1e3f * ms / (float)reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipFree(a));
    HIPCHK(hipFree(b));
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_engine_bench_apply(sofa_amd_engine eng, int scene, int reps, void* stream) {
    GUARD_BEGIN
    if (!eng || !eng->impl || (scene >= 0 && !scene_of(eng, scene)) || reps < 0) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!eng->impl->committed || (scene >= 0 && !eng->impl->placed(scene)))
        return fail(SOFA_AMD_ESTATE, "scene not committed");
    HIPCHK(hipSetDevice(eng->impl->device));
    eng->impl->bench_apply(scene, reps, stream ? (hipStream_t)stream : eng->impl->stream);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_scene_tiling(sofa_amd_engine eng, int scene, int32_t* out7) {
    GUARD_BEGIN
    if (!scene_of(eng, scene) || !out7) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!usable(eng, scene)) return fail(SOFA_AMD_ESTATE, "scene not committed");
    eng->impl->tiling(scene, out7);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_engine_reserve(sofa_amd_engine eng, int64_t max_nodes, int64_t max_tets, int max_scenes) {
    GUARD_BEGIN
    if (!eng || !eng->impl || max_nodes <= 0 || max_tets < 0 || max_scenes <= 0)
        return fail(SOFA_AMD_EINVAL, "invalid capacity");
    EngineBase* e = eng->impl.get();
    if (!e->live()) return fail(SOFA_AMD_EINVAL, "live scenes need a tiled engine (FP32, FP64_TILED or MIXED)");
    if (e->committed) return fail(SOFA_AMD_ESTATE, "engine already committed");
    Extent c;
    c.node = max_nodes + max_scenes;
    c.elem = max_tets;
    c.inc = 4 * max_tets;
    c.tile = max_tets / 128 + max_scenes;
    c.lnode = 2 * max_nodes + c.tile;
    c.shn = max_nodes + max_scenes;
    // shared-partial slots: one per (shared node, tile), or one per incidence of a shared node in
    // the reference summation order
    c.slot = e->seq_dots ? std::min<int64_t>(4 * max_tets, INT32_MAX - 1) : 2 * max_nodes;
    c.fin = max_nodes / kBlock + max_scenes;
    c.nblk = max_nodes / (kBlock * kNodesPerThread) + max_scenes;
    c.osc = max_nodes;
    c.sid = max_scenes;
    e->reserve(c);
    e->committed = true;
    e->live_mode = true;
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_scene_commit(sofa_amd_engine eng, int scene) {
    GUARD_BEGIN
    if (!scene_of(eng, scene)) return fail(SOFA_AMD_EINVAL, "invalid scene");
    EngineBase* e = eng->impl.get();
    if (!e->live_mode) return fail(SOFA_AMD_ESTATE, "engine not reserved (sofa_amd_engine_reserve)");
    if (e->placed(scene)) return fail(SOFA_AMD_ESTATE, "scene already committed");
    HIPCHK(hipSetDevice(e->device));
    if (!e->commit_scene(scene)) return fail(SOFA_AMD_ENOMEM, "no room for the scene in the engine's reservation");
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_scene_prepare(sofa_amd_engine eng, int scene) {
    GUARD_BEGIN
    if (!eng || !eng->impl) return fail(SOFA_AMD_EINVAL, "null engine");
    EngineBase* e = eng->impl.get();
    {
        std::lock_guard<std::mutex> g(e->scenes_mu);   // another thread may be adding scenes
        if (scene < 0 || scene >= (int)e->scenes.size()) return fail(SOFA_AMD_EINVAL, "invalid scene");
    }
    if (!e->live_mode) return fail(SOFA_AMD_ESTATE, "engine not reserved (sofa_amd_engine_reserve)");
    e->prepare_scene(scene);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_scene_release(sofa_amd_engine eng, int scene) {
    GUARD_BEGIN
    if (!scene_of(eng, scene)) return fail(SOFA_AMD_EINVAL, "invalid scene");
    EngineBase* e = eng->impl.get();
    if (!e->live_mode) return fail(SOFA_AMD_ESTATE, "engine not reserved (sofa_amd_engine_reserve)");
    if (!e->placed(scene)) return fail(SOFA_AMD_ESTATE, "scene not committed");
    HIPCHK(hipSetDevice(e->device));
    e->release_scene(scene);
    SceneHost& s = e->scenes[scene];
    s.rest.clear(); s.rest.shrink_to_fit();
    s.tets.clear(); s.tets.shrink_to_fit();
    s.young.assign(1, 5000.0); s.poisson.assign(1, 0.45); s.density.assign(1, 1.0);
    s.osc_idx.clear(); s.osc_mean.clear(); s.osc_amp.clear(); s.osc_w.clear(); s.osc_phase.clear();
    s.lm_idx.clear(); s.lm_times.clear(); s.lm_moves.clear();
    s.fixed.clear();
    s.released = true;
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_engine_usage(sofa_amd_engine eng, int64_t* out22) {
    GUARD_BEGIN
    if (!eng || !eng->impl || !out22) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!eng->impl->live_mode) return fail(SOFA_AMD_ESTATE, "engine not reserved (sofa_amd_engine_reserve)");
    eng->impl->usage(out22);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_engine_set_dot_order(sofa_amd_engine eng, int order) {
    GUARD_BEGIN
    if (!eng || !eng->impl || order < 0 || order > 1) return fail(SOFA_AMD_EINVAL, "invalid argument");
    const int prec = eng->impl->precision;
    if (prec != SOFA_AMD_FP64 && prec != SOFA_AMD_FP64_TILED)
        return fail(SOFA_AMD_EINVAL, "the summation order applies to the FP64 and FP64_TILED engines");
    // the tiled engine's tables depend on it (per-incidence slots of the shared nodes)
    if (prec == SOFA_AMD_FP64_TILED && (eng->impl->committed || eng->impl->live_mode))
        return fail(SOFA_AMD_ESTATE, "set the summation order of a tiled engine before its scenes are placed");
    HIPCHK(hipSetDevice(eng->impl->device));
    HIPCHK(hipDeviceSynchronize());
    eng->impl->seq_dots = order == 1;
    eng->impl->reset_graph();
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_engine_autotune(sofa_amd_engine eng, int reps, int* chosen, float* us4) {
    GUARD_BEGIN
    if (!eng || !eng->impl || reps < 1) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine not committed");
    if (eng->impl->stepped) return fail(SOFA_AMD_ESTATE, "autotune before the first step");
    HIPCHK(hipSetDevice(eng->impl->device));
    const int v = eng->impl->autotune(reps, eng->impl->stream, us4);
    if (chosen) *chosen = v;
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_engine_set_apply_variant(sofa_amd_engine eng, int variant) {
    GUARD_BEGIN
    if (!eng || !eng->impl) return fail(SOFA_AMD_EINVAL, "null engine");
    HIPCHK(hipSetDevice(eng->impl->device));
    eng->impl->set_apply_variant(variant);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_engine_set_branches(sofa_amd_engine eng, int branches) {
    GUARD_BEGIN
    if (!eng || !eng->impl) return fail(SOFA_AMD_EINVAL, "null engine");
    if (branches < 0 || branches > 8) return fail(SOFA_AMD_EINVAL, "branches must be 0 (default) to 8");
    HIPCHK(hipSetDevice(eng->impl->device));
    HIPCHK(hipDeviceSynchronize());
    eng->impl->branches = branches;
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_engine_set_dot_segments(sofa_amd_engine eng, int segments, int window) {
    GUARD_BEGIN
    if (!eng || !eng->impl) return fail(SOFA_AMD_EINVAL, "null engine");
    if (segments != 1 && segments != 2 && segments != 4 && segments != 8 && segments != 16)
        return fail(SOFA_AMD_EINVAL, "segments must be 1, 2, 4, 8 or 16");
    if (window < 0 || window > 32 * kSpecMaxWaves || window % 32 != 0)
        return fail(SOFA_AMD_EINVAL, "window must be 0, 32, 64, 96 or 128");
    if (1 + (segments - 1) * (2 * window + 1) > kSpecStride)
        return fail(SOFA_AMD_EINVAL, "16 segments take a window of 32 at most");
    HIPCHK(hipSetDevice(eng->impl->device));
    HIPCHK(hipDeviceSynchronize());
    eng->impl->dot_segments = segments;
    eng->impl->dot_window = window;
    eng->impl->reset_graph();
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_engine_set_branch_turns(sofa_amd_engine eng, int enable) {
    GUARD_BEGIN
    if (!eng || !eng->impl) return fail(SOFA_AMD_EINVAL, "null engine");
    HIPCHK(hipSetDevice(eng->impl->device));
    HIPCHK(hipDeviceSynchronize());
    eng->impl->branch_turns = enable != 0;
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_engine_set_graphs(sofa_amd_engine eng, int enable) {
    GUARD_BEGIN
    if (!eng || !eng->impl) return fail(SOFA_AMD_EINVAL, "null engine");
    HIPCHK(hipSetDevice(eng->impl->device));
    HIPCHK(hipDeviceSynchronize());
    eng->impl->use_graphs = enable != 0;
    eng->impl->reset_graph();
    return SOFA_AMD_OK;
    GUARD_END
}

}  // extern "C"

#endif  // !SOFA_AMD_TU_FP32
