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
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "sofa_amd.h"
#include "fem_kernels.h"
#include "fem_tiled.h"
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
    double mass_value = 1.0;                // vertexMass or massDensity
    double g[3] = {0, 0, 0};
    std::vector<int32_t> fixed{0};          // FixedProjectiveConstraint default indices = {0}
    double dt = 0.01, rs = 0, rm = 0;
    int cg_iter = 25;
    double tol = 1e-5, thr = 1e-5;
    // placement after commit
    int node_begin = 0, elem_begin = 0;
    int nblk_begin = 0, nblk_count = 0, eblk_begin = 0, eblk_count = 0;
};

// ---------------------------------------------------------------------------------------
// typed engine
// ---------------------------------------------------------------------------------------
struct EngineBase {
    int device = 0;
    int precision = SOFA_AMD_FP64;
    hipStream_t stream = nullptr;
    std::vector<SceneHost> scenes;
    bool committed = false;
    bool use_graphs = false;
    virtual ~EngineBase() {}
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
};

template <typename T>
static T* dalloc(size_t n) {
    T* p = nullptr;
    if (n == 0) n = 1;
    HIPCHK(hipMalloc((void**)&p, n * sizeof(T)));
    return p;
}

template <typename R, int NR, typename VT>
struct EngineT : EngineBase {
    bool tiled = false;
    int64_t Ntot = 0, Etot = 0;
    // node vectors (AoS xyz)
    VT *x = nullptr, *v = nullptr, *f = nullptr, *b = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *xcg = nullptr;
    double* mass = nullptr;
    int8_t* fixed = nullptr;
    int *inc_ptr = nullptr, *inc = nullptr;        // node -> (position*4 + vertex), ascending element id
    // element record (SoA, in tile order on the tiled path, else in element order)
    int4* tet = nullptr;
    R *rot = nullptr, *rest = nullptr, *kmat = nullptr;
    VT *ef = nullptr, *ef2 = nullptr;
    std::vector<std::vector<int>> iperm;           // per scene: original element -> position
    // scene records and block lists
    InstDev* inst = nullptr;
    std::vector<InstDev> inst_h;
    std::vector<Blk> eblk_h, nblk_h;
    double* partial = nullptr;
    // tiled path
    std::vector<TileDesc> tiles_h;
    std::vector<FinBlk> fin_h;
    std::vector<NodeBlk> nb_h;
    std::vector<int> tile_begin, tile_count, fin_begin, fin_count, nbk_begin, nbk_count, shared_slots;
    TileDesc* tiles_d = nullptr;
    Ranges* ranges_d = nullptr;
    uint2* ltet = nullptr;
    int *ln_gid = nullptr, *ln_inc_ptr = nullptr, *ln_pslot = nullptr, *sh_gid = nullptr, *sh_ptr = nullptr;
    uint16_t* ln_inc = nullptr;
    VT *sp0 = nullptr, *sp1 = nullptr;
    double *pbb = nullptr, *pden = nullptr, *prho = nullptr;
    // launch-set cache
    // launch sets: the block lists (a: element blocks / tiles, b: node blocks / finalize blocks,
    // c: tiled node blocks) and the captured step graph of each set of scenes stepped together
    struct LaunchSet {
        void *a = nullptr, *b = nullptr, *c = nullptr;
        int* scenes = nullptr;
        int na = 0, nb = 0, nc = 0, maxit = 0, n = 0;
        hipGraphExec_t gexec = nullptr;
        ~LaunchSet() {
            if (gexec) (void)hipGraphExecDestroy(gexec);
            for (void* ptr : {a, b, c, (void*)scenes})
                if (ptr) (void)hipFree(ptr);
        }
    };
    std::map<std::vector<int32_t>, std::unique_ptr<LaunchSet>> sets;
    LaunchSet* cur = nullptr;

    ~EngineT() override {
        sets.clear();
        void* ptrs[] = {x, v, f, b, r, p, q, xcg, mass, fixed, inc_ptr, inc, tet, rot, rest, kmat, ef, ef2,
                        inst, partial, tiles_d, ranges_d, ltet, ln_gid, ln_inc_ptr, ln_pslot, sh_gid, sh_ptr,
                        ln_inc, sp0, sp1, pbb, pden, prho};
        for (void* ptr : ptrs)
            if (ptr) (void)hipFree(ptr);
    }

    ElemView<R, NR> view() const { return ElemView<R, NR>{tet, rot, rest, kmat, Etot}; }
    NodeArrays nodes() const { return NodeArrays{inc_ptr, inc, mass, fixed}; }

    template <typename T>
    void upload(T*& dst, const std::vector<T>& src) {
        dst = dalloc<T>(src.size());
        if (!src.empty()) HIPCHK(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    }

    void commit() override {
        HIPCHK(hipSetDevice(device));
        Ntot = 0; Etot = 0;
        for (auto& s : scenes) {
            s.node_begin = (int)Ntot;
            s.elem_begin = (int)Etot;
            Ntot += s.n_nodes;
            Etot += (int64_t)s.tets.size() / 4;
        }
        if (Ntot >= (int64_t)1 << 30 || Etot * 4 >= (int64_t)1 << 31)
            throw std::invalid_argument("mesh too large for 32-bit indices");
        const int S = (int)scenes.size();
        inst_h.assign(S, InstDev());
        eblk_h.clear(); nblk_h.clear();
        tiles_h.clear(); fin_h.clear(); nb_h.clear();
        tile_begin.assign(S, 0); tile_count.assign(S, 0); fin_begin.assign(S, 0); fin_count.assign(S, 0);
        nbk_begin.assign(S, 0); nbk_count.assign(S, 0); shared_slots.assign(S, 0);
        iperm.assign(S, {});
        std::vector<double> xh(3 * Ntot), massh(Ntot, 0.0);
        std::vector<int8_t> fixh(Ntot, 0);
        std::vector<int4> teth(Etot);
        std::vector<R> roth(9 * Etot), resth(NR * Etot), kh(3 * Etot);
        std::vector<int> cnt(Ntot + 1, 0);
        // tiled-path host tables (global ids)
        std::vector<uint2> ltet_h(tiled ? Etot : 0);
        std::vector<int> lgid, lptr{0}, lps, shg, shp{0};
        std::vector<uint16_t> linc;
        int slot_base = 0;
        for (int si = 0; si < S; si++) {
            SceneHost& s = scenes[si];
            const int nE = (int)(s.tets.size() / 4);
            std::vector<host::V3> P(s.n_nodes);
            for (int i = 0; i < s.n_nodes; i++) {
                for (int k = 0; k < 3; k++) P[i].v[k] = s.rest[3 * i + k];
                for (int k = 0; k < 3; k++) xh[3 * (s.node_begin + (int64_t)i) + k] = s.rest[3 * i + k];
            }
            TilePlan plan;
            std::vector<int>& ip = iperm[si];
            if (tiled && nE > 0) {
                plan = plan_tiles(s.n_nodes, s.rest.data(), s.tets.data(), nE, kTileElems, kTileNodes);
                ip = plan.iperm;
            } else {
                ip.resize(nE);
                for (int e = 0; e < nE; e++) ip[e] = e;
            }
            for (int e = 0; e < nE; e++) {
                const int32_t* t = &s.tets[4 * (size_t)e];
                for (int k = 0; k < 4; k++)
                    if (t[k] < 0 || t[k] >= s.n_nodes) throw std::invalid_argument("tetrahedron index out of range");
                const double E = s.young.size() > (size_t)e ? s.young[e] : s.young[0];
                const double nu = s.poisson.size() > (size_t)e ? s.poisson[e] : s.poisson[0];
                host::ElemInit ei;
                if (!host::init_element(P, t, E, nu, ei))
                    throw std::domain_error("degenerate tetrahedron " + std::to_string(e) + " in scene " + std::to_string(si));
                const int64_t ge = s.elem_begin + (int64_t)ip[e];
                teth[ge] = make_int4(t[0] + s.node_begin, t[1] + s.node_begin, t[2] + s.node_begin, t[3] + s.node_begin);
                for (int k = 0; k < 9; k++) roth[k * Etot + ge] = (R)ei.rot[k];
                if (NR == 9) {
                    for (int k = 0; k < 9; k++) resth[k * Etot + ge] = (R)ei.rest[k];
                } else {
                    const int pick[6] = {0, 3, 4, 6, 7, 8};   // b0, c0, c1, d0, d1, d2
                    for (int k = 0; k < 6; k++) resth[k * Etot + ge] = (R)ei.rest[pick[k]];
                }
                for (int k = 0; k < 3; k++) kh[k * Etot + ge] = (R)ei.k[k];
                for (int k = 0; k < 4; k++) cnt[s.node_begin + t[k] + 1]++;
                if (s.mass_kind == SOFA_AMD_MASS_DIAGONAL) {
                    const double m = (s.mass_value * ei.vol) / 4.0;     // DiagonalMass.inl:1075-1080
                    for (int k = 0; k < 4; k++) massh[s.node_begin + t[k]] += m;
                }
            }
            for (int32_t i : s.fixed) {
                if (i < 0 || i >= s.n_nodes) throw std::invalid_argument("fixed index out of range");
                fixh[s.node_begin + i] = 1;
            }
            // exact-path blocks: 256 elements / 256 nodes
            s.eblk_begin = (int)eblk_h.size();
            for (int e0 = 0; e0 < nE; e0 += kBlock)
                eblk_h.push_back(Blk{si, s.elem_begin + e0, s.elem_begin + std::min(nE, e0 + kBlock), 0});
            s.eblk_count = (int)eblk_h.size() - s.eblk_begin;
            s.nblk_begin = (int)nblk_h.size();
            for (int n0 = 0; n0 < s.n_nodes; n0 += kBlock)
                nblk_h.push_back(Blk{si, s.node_begin + n0, s.node_begin + std::min(s.n_nodes, n0 + kBlock),
                                     (int)nblk_h.size()});
            s.nblk_count = (int)nblk_h.size() - s.nblk_begin;
            if (tiled) {
                // tiles
                tile_begin[si] = (int)tiles_h.size();
                const int T = (int)plan.t_ebegin.size();
                const int lbase = (int)lgid.size();
                for (int ti = 0; ti < T; ti++) {
                    TileDesc td;
                    td.scene = si;
                    td.e_begin = s.elem_begin + plan.t_ebegin[ti];
                    td.e_count = plan.t_ecount[ti];
                    td.ln_begin = lbase + plan.t_lnbegin[ti];
                    td.ln_count = plan.t_lncount[ti];
                    td.first = (ti == 0);
                    td.pslot = (int)tiles_h.size();
                    td.pad = 0;
                    tiles_h.push_back(td);
                }
                tile_count[si] = T;
                for (int i = 0; i < nE; i++) ltet_h[s.elem_begin + i] = plan.ltet[i];
                const int ibase = (int)linc.size();
                for (size_t j = 0; j < plan.ln_gid.size(); j++) {
                    lgid.push_back(plan.ln_gid[j] + s.node_begin);
                    lptr.push_back(ibase + plan.ln_inc_ptr[j + 1]);
                    lps.push_back(plan.ln_pslot[j] < 0 ? -1 : plan.ln_pslot[j] + slot_base);
                }
                linc.insert(linc.end(), plan.ln_inc.begin(), plan.ln_inc.end());
                // shared / element-free nodes
                std::vector<int> shared = plan.sh_gid, sptr = plan.sh_ptr;
                if (nE == 0) {   // no elements: every node is finished by the finalize kernel
                    shared.resize(s.n_nodes);
                    for (int i = 0; i < s.n_nodes; i++) shared[i] = i;
                    sptr.assign(s.n_nodes + 1, 0);
                }
                const int shbase = (int)shg.size();
                for (size_t j = 0; j < shared.size(); j++) {
                    shg.push_back(shared[j] + s.node_begin);
                    shp.push_back(slot_base + sptr[j + 1]);
                }
                fin_begin[si] = (int)fin_h.size();
                for (int j0 = 0; j0 < (int)shared.size(); j0 += kBlock)
                    fin_h.push_back(FinBlk{si, shbase + j0, shbase + std::min((int)shared.size(), j0 + kBlock), 0,
                                           j0 / kBlock, 0, 0, 0});
                fin_count[si] = (int)fin_h.size() - fin_begin[si];
                slot_base += sptr.back();
                shared_slots[si] = sptr.back();
                nbk_begin[si] = (int)nb_h.size();
                const int per = kBlock * kNodesPerThread;
                for (int n0 = 0; n0 < s.n_nodes; n0 += per)
                    nb_h.push_back(NodeBlk{si, s.node_begin + n0, s.node_begin + std::min(s.n_nodes, n0 + per),
                                           (int)nb_h.size(), n0 == 0, 0, 0, 0});
                nbk_count[si] = (int)nb_h.size() - nbk_begin[si];
            }
            // scene record
            InstDev& in = inst_h[si];
            std::memset(&in, 0, sizeof(in));
            in.node_begin = s.node_begin; in.node_count = s.n_nodes;
            in.elem_begin = s.elem_begin; in.elem_count = nE;
            in.nblk_begin = s.nblk_begin; in.nblk_count = s.nblk_count;
            in.mass_kind = s.mass_kind;
            in.max_iter = s.cg_iter;
            in.dt = s.dt; in.rm = s.rm; in.rs = s.rs;
            in.vertex_mass = s.mass_kind == SOFA_AMD_MASS_UNIFORM ? s.mass_value : 0.0;
            for (int k = 0; k < 3; k++) {
                in.gravity[k] = s.g[k];
                // UniformMass::addForce: mg = theGravity * m, in the state's precision
                in.mg[k] = (double)((VT)s.g[k] * (VT)in.vertex_mass);
            }
            const double h = s.dt, tr = 1.0;
            in.k_rhs = h * tr + s.rs;
            in.m_rhs = -s.rm;
            in.m_cg = 1 + tr * h * s.rm;
            in.k_cg = -tr * h * (tr * h + s.rs);
            in.tol = s.tol < 0 ? 1e-5 : s.tol;
            in.thr = s.thr < 0 ? 1e-5 : s.thr;
            in.first_solve = 1;
            in.stopped = tiled ? 0 : 1;
        }
        const int n_tiles = (int)tiles_h.size();
        for (auto& fb : fin_h) fb.pslot = n_tiles + (int)(&fb - fin_h.data());
        // incidence CSR for the exact path and the component calls: node -> position*4 + vertex,
        // visited in ascending original element id
        for (int64_t i = 0; i < Ntot; i++) cnt[i + 1] += cnt[i];
        std::vector<int> inch(4 * Etot), fillp(cnt.begin(), cnt.end() - 1);
        for (int si = 0; si < S; si++) {
            const SceneHost& s = scenes[si];
            const int nE = (int)(s.tets.size() / 4);
            for (int e = 0; e < nE; e++) {
                const int64_t ge = s.elem_begin + (int64_t)iperm[si][e];
                for (int k = 0; k < 4; k++) inch[fillp[s.node_begin + s.tets[4 * (size_t)e + k]]++] = (int)(4 * ge + k);
            }
        }
        // device buffers
        x = dalloc<VT>(3 * Ntot); v = dalloc<VT>(3 * Ntot); f = dalloc<VT>(3 * Ntot); b = dalloc<VT>(3 * Ntot);
        r = dalloc<VT>(3 * Ntot); p = dalloc<VT>(3 * Ntot); q = dalloc<VT>(3 * Ntot); xcg = dalloc<VT>(3 * Ntot);
        std::vector<VT> xv(3 * Ntot);
        for (int64_t i = 0; i < 3 * Ntot; i++) xv[i] = (VT)xh[i];
        HIPCHK(hipMemcpy(x, xv.data(), xv.size() * sizeof(VT), hipMemcpyHostToDevice));
        for (VT* vec : {v, f, b, r, p, q, xcg}) HIPCHK(hipMemset(vec, 0, 3 * Ntot * sizeof(VT)));
        upload(mass, massh); upload(fixed, fixh);
        upload(inc_ptr, cnt); upload(inc, inch);
        upload(tet, teth); upload(rot, roth); upload(rest, resth); upload(kmat, kh);
        ef = dalloc<VT>(12 * Etot);
        if (!tiled) ef2 = dalloc<VT>(12 * Etot);
        upload(inst, inst_h);
        partial = dalloc<double>(nblk_h.size());
        HIPCHK(hipMemset(partial, 0, std::max<size_t>(1, nblk_h.size()) * sizeof(double)));
        if (tiled) {
            std::vector<Ranges> rg(S);
            for (int si = 0; si < S; si++)
                rg[si] = Ranges{tile_begin[si], tile_count[si], n_tiles + fin_begin[si], fin_count[si], nbk_begin[si],
                                nbk_count[si]};
            upload(tiles_d, tiles_h);
            upload(ranges_d, rg);
            upload(ltet, ltet_h);
            upload(ln_gid, lgid); upload(ln_inc_ptr, lptr); upload(ln_pslot, lps); upload(ln_inc, linc);
            upload(sh_gid, shg); upload(sh_ptr, shp);
            sp0 = dalloc<VT>(3 * (size_t)std::max(1, slot_base));
            sp1 = dalloc<VT>(3 * (size_t)std::max(1, slot_base));
            const size_t np = n_tiles + fin_h.size();
            pbb = dalloc<double>(np); pden = dalloc<double>(np); prho = dalloc<double>(nb_h.size());
            HIPCHK(hipMemset(pbb, 0, std::max<size_t>(1, np) * sizeof(double)));
            HIPCHK(hipMemset(pden, 0, std::max<size_t>(1, np) * sizeof(double)));
            HIPCHK(hipMemset(prho, 0, std::max<size_t>(1, nb_h.size()) * sizeof(double)));
        }
        HIPCHK(hipDeviceSynchronize());
        committed = true;
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
        }
        std::unique_ptr<LaunchSet> ls(new LaunchSet());
        int maxit = 0;
        if (!tiled) {
            std::vector<Blk> eb, nb;
            for (int32_t id : want) {
                const SceneHost& sc = scenes[id];
                push_range(eb, eblk_h, sc.eblk_begin, sc.eblk_count);
                push_range(nb, nblk_h, sc.nblk_begin, sc.nblk_count);
                maxit = std::max(maxit, sc.cg_iter);
            }
            put(ls->a, eb); put(ls->b, nb);
            ls->na = (int)eb.size(); ls->nb = (int)nb.size(); ls->nc = 0;
        } else {
            std::vector<TileDesc> ta;
            std::vector<FinBlk> fa;
            std::vector<NodeBlk> na;
            for (int32_t id : want) {
                push_range(ta, tiles_h, tile_begin[id], tile_count[id]);
                push_range(fa, fin_h, fin_begin[id], fin_count[id]);
                push_range(na, nb_h, nbk_begin[id], nbk_count[id]);
                maxit = std::max(maxit, scenes[id].cg_iter);
            }
            put(ls->a, ta); put(ls->b, fa); put(ls->c, na);
            ls->na = (int)ta.size(); ls->nb = (int)fa.size(); ls->nc = (int)na.size();
        }
        void* sc = nullptr;
        put(sc, std::vector<int>(want.begin(), want.end()));
        ls->scenes = (int*)sc;
        ls->maxit = maxit;
        ls->n = n;
        cur = ls.get();
        sets[want] = std::move(ls);
        (void)s;
    }

    // one implicit Euler step of the current launch set, enqueued on `s`
    void launch_step(int n, hipStream_t s) {
        if (tiled) { launch_step_tiled(n, s); return; }
        const ElemView<R, NR> ev = view();
        const NodeArrays na = nodes();
        const Blk* eb = (const Blk*)cur->a;
        const Blk* nb = (const Blk*)cur->b;
        if (cur->na > 0) k_rhs_elem<R, NR, VT><<<cur->na, kBlock, 0, s>>>(ev, eb, inst, x, v, ef, ef2);
        if (cur->nb > 0) k_rhs_node<VT><<<cur->nb, kBlock, 0, s>>>(nb, inst, na, ef, ef2, v, f, b, r, xcg, partial);
        for (int it = 0; it < cur->maxit; it++) {
            if (cur->na > 0) k_cg_elem<R, NR, VT><<<cur->na, kBlock, 0, s>>>(ev, eb, inst, r, p, ef);
            if (cur->nb > 0) {
                k_cg_node<VT><<<cur->nb, kBlock, 0, s>>>(nb, inst, na, ef, r, p, q, partial);
                k_cg_update<VT><<<cur->nb, kBlock, 0, s>>>(nb, inst, p, q, xcg, r, partial);
            }
        }
        if (cur->nb > 0) k_euler_update<VT><<<cur->nb, kBlock, 0, s>>>(nb, inst, xcg, v, x);
        k_finalize<<<(n + 63) / 64, 64, 0, s>>>(cur->scenes, n, inst);
    }

    void launch_step_tiled(int n, hipStream_t s) {
        TileArgs<R, NR> ta;
        ta.tiles = (const TileDesc*)cur->a;
        ta.inst = inst;
        ta.ranges = ranges_d;
        ta.ev = view();
        ta.ltet = ltet;
        ta.tn = TileNodes{ln_gid, ln_inc_ptr, ln_inc, ln_pslot};
        ta.mass = mass;
        ta.fixed = fixed;
        const FinBlk* fb = (const FinBlk*)cur->b;
        const NodeBlk* nb = (const NodeBlk*)cur->c;
        const SharedNodes sn{sh_gid, sh_ptr};
        if (cur->na > 0) k_tile_rhs<R, NR, VT><<<cur->na, kBlock, 0, s>>>(ta, x, v, f, b, r, xcg, sp0, sp1, pbb);
        if (cur->nb > 0) k_fin_rhs<VT><<<cur->nb, kBlock, 0, s>>>(fb, inst, ranges_d, sn, mass, fixed, v, sp0, sp1, f, b, r, xcg, pbb);
        for (int it = 1; it <= cur->maxit; it++) {
            if (cur->na > 0)
                k_tile_apply<R, NR, VT><<<cur->na, kBlock, 0, s>>>(ta, it, r, p, q, sp0, it == 1 ? pbb : prho, pden);
            if (cur->nb > 0) k_fin_apply<VT><<<cur->nb, kBlock, 0, s>>>(fb, inst, ranges_d, sn, mass, fixed, it, r, p, q, sp0, pden);
            if (cur->nc > 0) k_update<VT><<<cur->nc, kBlock, 0, s>>>(nb, inst, ranges_d, it, p, q, xcg, r, pden, prho);
        }
        if (cur->nc > 0) k_integrate<VT><<<cur->nc, kBlock, 0, s>>>(nb, inst, xcg, v, x);
        k_step_end_tiled<<<(n + 63) / 64, 64, 0, s>>>(cur->scenes, n, inst);
    }

    void reset_graph() override {
        for (auto& kv : sets)
            if (kv.second->gexec) {
                (void)hipGraphExecDestroy(kv.second->gexec);
                kv.second->gexec = nullptr;
            }
    }

    void step(const int32_t* ids, int n, int nsteps, hipStream_t s) override {
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
        else for (int i = 0; i < (int)scenes.size(); i++) ids.push_back(i);
        prepare_set(ids.data(), (int)ids.size(), s);
        TileArgs<R, NR> ta;
        ta.tiles = (const TileDesc*)cur->a;
        ta.inst = inst;
        ta.ranges = ranges_d;
        ta.ev = view();
        ta.ltet = ltet;
        ta.tn = TileNodes{ln_gid, ln_inc_ptr, ln_inc, ln_pslot};
        ta.mass = mass;
        ta.fixed = fixed;
        for (int i = 0; i < reps; i++)
            if (cur->na > 0) k_tile_apply<R, NR, VT><<<cur->na, kBlock, 0, s>>>(ta, 2, r, p, q, sp0, prho, pden);
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
        out[0] = tile_count[scene];
        for (int t = 0; t < tile_count[scene]; t++) out[1] += tiles_h[tile_begin[scene] + t].ln_count;
        for (int b = 0; b < fin_count[scene]; b++) {
            const FinBlk& fb = fin_h[fin_begin[scene] + b];
            out[2] += fb.end - fb.begin;
        }
        out[3] = shared_slots[scene];
        out[4] = 4 * out[5];
    }

    void read_vec(const VT* src, int scene, double* out) {
        const SceneHost& sc = scenes[scene];
        std::vector<VT> tmp(3 * (size_t)sc.n_nodes);
        HIPCHK(hipMemcpy(tmp.data(), src + 3 * (int64_t)sc.node_begin, tmp.size() * sizeof(VT), hipMemcpyDeviceToHost));
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
            HIPCHK(hipMemcpy(x + 3 * (int64_t)sc.node_begin, tmp.data(), tmp.size() * sizeof(VT), hipMemcpyHostToDevice));
        }
        if (vi) {
            for (size_t i = 0; i < tmp.size(); i++) tmp[i] = (VT)vi[i];
            HIPCHK(hipMemcpy(v + 3 * (int64_t)sc.node_begin, tmp.data(), tmp.size() * sizeof(VT), hipMemcpyHostToDevice));
        }
    }
    void stats(int scene, sofa_amd_cg_stats* out) override {
        InstDev d;
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(&d, inst + scene, sizeof(d), hipMemcpyDeviceToHost));
        out->iterations = d.last_iters;
        out->end_reason = d.end_reason;
        out->step_count = d.step_count;
        out->ir_sweeps = 0;
        out->err = d.err;
        out->den = d.den;
        out->time = d.time;
    }
    void element(int scene, int e, double* rot9, double* k3, double* rest12) override {
        HIPCHK(hipDeviceSynchronize());
        const int64_t ge = scenes[scene].elem_begin + (int64_t)iperm[scene][e];
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
        if (nE > 0) {
            const int grid = (int)((nE + kBlock - 1) / kBlock);
            k_elem_add_force<R, NR, VT><<<grid, kBlock, 0, s>>>(view(), sc.elem_begin, sc.elem_begin + nE,
                                                               sc.node_begin, (const VT*)xd, ef);
        }
        const int gn = (sc.n_nodes + kBlock - 1) / kBlock;
        if (gn > 0)
            k_gather<VT><<<gn, kBlock, 0, s>>>(inc_ptr, inc, sc.node_begin, sc.node_begin + sc.n_nodes, ef,
                                               (VT*)fd, 0);
        HIPCHK(hipGetLastError());
    }
    void add_dforce(int scene, const void* dxd, void* dfd, double kf, hipStream_t s) override {
        const SceneHost& sc = scenes[scene];
        const int64_t nE = (int64_t)sc.tets.size() / 4;
        if (nE > 0) {
            const int grid = (int)((nE + kBlock - 1) / kBlock);
            // addDForce: kFactor = (Real) mparams->kFactor(); applied as `KJtD *= fact`
            k_elem_add_dforce<R, NR, VT><<<grid, kBlock, 0, s>>>(view(), sc.elem_begin, sc.elem_begin + nE,
                                                                sc.node_begin, (const VT*)dxd, (R)kf, ef);
        }
        const int gn = (sc.n_nodes + kBlock - 1) / kBlock;
        if (gn > 0)
            k_gather<VT><<<gn, kBlock, 0, s>>>(inc_ptr, inc, sc.node_begin, sc.node_begin + sc.n_nodes, ef,
                                               (VT*)dfd, 1);
        HIPCHK(hipGetLastError());
    }
};

#if defined(SOFA_AMD_TU_FP32)
EngineBase* make_engine_fp32() {
    auto* e = new EngineT<float, 6, float>();
    e->tiled = true;
    return e;
}
#else
EngineBase* make_engine_fp32();
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
    if (eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine already committed");
    SceneHost s;
    s.n_nodes = n_nodes;
    s.rest.assign(positions, positions + 3 * (size_t)n_nodes);
    for (double d : s.rest)
        if (!std::isfinite(d)) return fail(SOFA_AMD_EINVAL, "non-finite position");
    eng->impl->scenes.push_back(std::move(s));
    *out_scene = (int)eng->impl->scenes.size() - 1;
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_scene_set_fem(sofa_amd_engine eng, int scene, int n_tets, const int32_t* tets, const double* young,
                           int n_young, const double* poisson, int n_poisson) {
    GUARD_BEGIN
    SceneHost* s = scene_of(eng, scene);
    if (!s || n_tets < 0 || (n_tets > 0 && !tets)) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine already committed");
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

int sofa_amd_scene_set_mass(sofa_amd_engine eng, int scene, int kind, double value) {
    SceneHost* s = scene_of(eng, scene);
    if (!s || (kind != SOFA_AMD_MASS_UNIFORM && kind != SOFA_AMD_MASS_DIAGONAL) || !(value > 0))
        return fail(SOFA_AMD_EINVAL, "invalid mass");
    if (eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine already committed");
    s->mass_kind = kind;
    s->mass_value = value;
    return SOFA_AMD_OK;
}

int sofa_amd_scene_set_gravity(sofa_amd_engine eng, int scene, double gx, double gy, double gz) {
    SceneHost* s = scene_of(eng, scene);
    if (!s) return fail(SOFA_AMD_EINVAL, "invalid scene");
    if (eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine already committed");
    s->g[0] = gx; s->g[1] = gy; s->g[2] = gz;
    return SOFA_AMD_OK;
}

int sofa_amd_scene_set_fixed(sofa_amd_engine eng, int scene, const int32_t* indices, int n) {
    SceneHost* s = scene_of(eng, scene);
    if (!s || n < 0 || (n > 0 && !indices)) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine already committed");
    for (int i = 0; i < n; i++)
        if (indices[i] < 0 || indices[i] >= s->n_nodes) return fail(SOFA_AMD_EINVAL, "fixed index out of range");
    s->fixed.assign(indices, indices + n);
    return SOFA_AMD_OK;
}

int sofa_amd_scene_set_solver(sofa_amd_engine eng, int scene, double dt, double rs, double rm, int iters,
                              double tol, double thr) {
    SceneHost* s = scene_of(eng, scene);
    if (!s || !(dt > 0) || iters < 0) return fail(SOFA_AMD_EINVAL, "invalid solver parameters");
    if (eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine already committed");
    s->dt = dt; s->rs = rs; s->rm = rm; s->cg_iter = iters; s->tol = tol; s->thr = thr;
    return SOFA_AMD_OK;
}

int sofa_amd_engine_commit(sofa_amd_engine eng) {
    GUARD_BEGIN
    if (!eng || !eng->impl) return fail(SOFA_AMD_EINVAL, "null engine");
    if (eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine already committed");
    if (eng->impl->scenes.empty()) return fail(SOFA_AMD_EINVAL, "no scene");
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

int sofa_amd_scene_read(sofa_amd_engine eng, int scene, double* x, double* v, double* f) {
    GUARD_BEGIN
    if (!scene_of(eng, scene)) return fail(SOFA_AMD_EINVAL, "invalid scene");
    if (!eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine not committed");
    eng->impl->read(scene, x, v, f);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_scene_write(sofa_amd_engine eng, int scene, const double* x, const double* v) {
    GUARD_BEGIN
    if (!scene_of(eng, scene)) return fail(SOFA_AMD_EINVAL, "invalid scene");
    if (!eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine not committed");
    eng->impl->write(scene, x, v);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_scene_stats(sofa_amd_engine eng, int scene, sofa_amd_cg_stats* out) {
    GUARD_BEGIN
    if (!scene_of(eng, scene) || !out) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine not committed");
    eng->impl->stats(scene, out);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_scene_element(sofa_amd_engine eng, int scene, int elem, double* rot9, double* k3, double* rest12) {
    GUARD_BEGIN
    SceneHost* s = scene_of(eng, scene);
    if (!s || !rot9 || !k3 || !rest12 || elem < 0 || elem >= (int)(s->tets.size() / 4))
        return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine not committed");
    eng->impl->element(scene, elem, rot9, k3, rest12);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_fem_add_force(sofa_amd_engine eng, int scene, const void* x_dev, void* f_dev, void* stream) {
    GUARD_BEGIN
    if (!scene_of(eng, scene) || !x_dev || !f_dev) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine not committed");
    HIPCHK(hipSetDevice(eng->impl->device));
    eng->impl->add_force(scene, x_dev, f_dev, stream ? (hipStream_t)stream : eng->impl->stream);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_fem_add_dforce(sofa_amd_engine eng, int scene, const void* dx_dev, void* df_dev, double k_factor,
                            void* stream) {
    GUARD_BEGIN
    if (!scene_of(eng, scene) || !dx_dev || !df_dev) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine not committed");
    HIPCHK(hipSetDevice(eng->impl->device));
    eng->impl->add_dforce(scene, dx_dev, df_dev, k_factor, stream ? (hipStream_t)stream : eng->impl->stream);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_engine_bench_apply(sofa_amd_engine eng, int scene, int reps, void* stream) {
    GUARD_BEGIN
    if (!eng || !eng->impl || (scene >= 0 && !scene_of(eng, scene)) || reps < 0) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine not committed");
    HIPCHK(hipSetDevice(eng->impl->device));
    eng->impl->bench_apply(scene, reps, stream ? (hipStream_t)stream : eng->impl->stream);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_scene_tiling(sofa_amd_engine eng, int scene, int32_t* out7) {
    GUARD_BEGIN
    if (!scene_of(eng, scene) || !out7) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine not committed");
    eng->impl->tiling(scene, out7);
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
