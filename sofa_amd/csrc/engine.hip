// engine.hip -- host side of libsofa_amd.so: scene tables, element initialisation, device
// buffers, launch schedule and the C ABI declared in include/sofa_amd.h.
//
// Element initialisation restates TetrahedronFEMForceField::reinit for method "large"
// (TetrahedronFEMForceField.inl:1391-1545): computeMaterialStiffness (:256-291), initLarge
// (:835-868) with computeRotationLarge (:755-778).  It runs once per scene on the host in
// double, in the reference's operation order (this file is compiled with contraction off), and
// the result is rounded to the element precision of the engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sofa_amd.h"
#include "fem_kernels.h"

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
static inline V3 sub(const V3& a, const V3& b) { return {{a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]}}; }
static inline V3 cross(const V3& a, const V3& b) {
    return {{a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2], a.v[0] * b.v[1] - a.v[1] * b.v[0]}};
}
static inline double dot(const V3& a, const V3& b) {
    double r = a.v[0] * b.v[0];
    r += a.v[1] * b.v[1];
    r += a.v[2] * b.v[2];
    return r;
}
static inline V3 normalized(V3 a) {
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
    return std::abs(dot(cross(sub(n1, n0), sub(n2, n0)), sub(n3, n0)) / 6.0);
}

struct ElemInit {
    double rot[9];     // rotations[e] = _initialRotations[e] (transpose of R_0_1)
    double rest[9];    // rotated rest coordinates b, c, d (a = 0)
    double k[3];       // K00, K01, K33 after / (36 V)
    double vol;
};

static bool init_element(const std::vector<V3>& P, const int32_t* t, double E, double nu, ElemInit& out) {
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
    virtual ~EngineBase() {}
    virtual void commit() = 0;
    virtual void step(const int32_t* ids, int n, int nsteps, hipStream_t s) = 0;
    virtual void read(int scene, double* x, double* v, double* f) = 0;
    virtual void write(int scene, const double* x, const double* v) = 0;
    virtual void stats(int scene, sofa_amd_cg_stats* out) = 0;
    virtual void element(int scene, int e, double* rot9, double* k3, double* rest12) = 0;
    virtual void add_force(int scene, const double* x, double* f, hipStream_t s) = 0;
    virtual void add_dforce(int scene, const double* dx, double* df, double k, hipStream_t s) = 0;
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
    int64_t Ntot = 0, Etot = 0;
    // node vectors (AoS xyz)
    VT *x = nullptr, *v = nullptr, *f = nullptr, *b = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *xcg = nullptr;
    double* mass = nullptr;
    int8_t* fixed = nullptr;
    int *inc_ptr = nullptr, *inc = nullptr;
    // element record
    int4* tet = nullptr;
    R *rot = nullptr, *rest = nullptr, *kmat = nullptr;
    VT *ef = nullptr, *ef2 = nullptr;
    double* ef_d = nullptr;   // double contributions for the component-level calls
    // scene records and block lists
    InstDev* inst = nullptr;
    std::vector<InstDev> inst_h;
    std::vector<Blk> eblk_h, nblk_h;
    double* partial = nullptr;
    // launch-set cache
    std::vector<int32_t> set_ids;
    Blk *set_eblk = nullptr, *set_nblk = nullptr;
    int* set_scenes = nullptr;
    int set_ne = 0, set_nn = 0, set_cap_e = 0, set_cap_n = 0, set_cap_s = 0, set_maxit = 0;

    ~EngineT() override {
        void* ptrs[] = {x, v, f, b, r, p, q, xcg, mass, fixed, inc_ptr, inc, tet, rot, rest, kmat, ef, ef2,
                        inst, partial, set_eblk, set_nblk, set_scenes};
        if (ef_d && (void*)ef_d != (void*)ef) (void)hipFree(ef_d);
        for (void* ptr : ptrs)
            if (ptr) (void)hipFree(ptr);
    }

    ElemView<R, NR> view() const { return ElemView<R, NR>{tet, rot, rest, kmat, Etot}; }
    NodeArrays nodes() const { return NodeArrays{inc_ptr, inc, mass, fixed}; }

    void commit() override {
        HIPCHK(hipSetDevice(device));
        // placement
        Ntot = 0; Etot = 0;
        eblk_h.clear(); nblk_h.clear();
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
        std::vector<double> xh(3 * Ntot), massh(Ntot, 0.0);
        std::vector<int8_t> fixh(Ntot, 0);
        std::vector<int4> teth(Etot);
        std::vector<R> roth(9 * Etot), resth(NR * Etot), kh(3 * Etot);
        std::vector<int> cnt(Ntot + 1, 0);
        for (int si = 0; si < S; si++) {
            SceneHost& s = scenes[si];
            const int nE = (int)(s.tets.size() / 4);
            std::vector<host::V3> P(s.n_nodes);
            for (int i = 0; i < s.n_nodes; i++) {
                for (int k = 0; k < 3; k++) P[i].v[k] = s.rest[3 * i + k];
                for (int k = 0; k < 3; k++) xh[3 * (s.node_begin + (int64_t)i) + k] = s.rest[3 * i + k];
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
                const int64_t ge = s.elem_begin + (int64_t)e;
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
            // blocks
            s.eblk_begin = (int)eblk_h.size();
            for (int e0 = 0; e0 < nE; e0 += kBlock)
                eblk_h.push_back(Blk{si, s.elem_begin + e0, s.elem_begin + std::min(nE, e0 + kBlock), 0});
            s.eblk_count = (int)eblk_h.size() - s.eblk_begin;
            s.nblk_begin = (int)nblk_h.size();
            for (int n0 = 0; n0 < s.n_nodes; n0 += kBlock)
                nblk_h.push_back(Blk{si, s.node_begin + n0, s.node_begin + std::min(s.n_nodes, n0 + kBlock),
                                     (int)nblk_h.size()});
            s.nblk_count = (int)nblk_h.size() - s.nblk_begin;
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
            in.stopped = 1;
        }
        // incidence CSR: slots e*4+j sorted by node, then by element (counting sort keeps
        // ascending slot order because elements are visited in order)
        for (int64_t i = 0; i < Ntot; i++) cnt[i + 1] += cnt[i];
        std::vector<int> inch(4 * Etot), fillp(cnt.begin(), cnt.end() - 1);
        for (int64_t ge = 0; ge < Etot; ge++) {
            const int4 t = teth[ge];
            const int nid[4] = {t.x, t.y, t.z, t.w};
            for (int k = 0; k < 4; k++) inch[fillp[nid[k]]++] = (int)(4 * ge + k);
        }
        // device buffers
        auto up = [&](auto* dst, const auto& src) {
            HIPCHK(hipMemcpy(dst, src.data(), src.size() * sizeof(src[0]), hipMemcpyHostToDevice));
        };
        x = dalloc<VT>(3 * Ntot); v = dalloc<VT>(3 * Ntot); f = dalloc<VT>(3 * Ntot); b = dalloc<VT>(3 * Ntot);
        r = dalloc<VT>(3 * Ntot); p = dalloc<VT>(3 * Ntot); q = dalloc<VT>(3 * Ntot); xcg = dalloc<VT>(3 * Ntot);
        mass = dalloc<double>(Ntot); fixed = dalloc<int8_t>(Ntot);
        inc_ptr = dalloc<int>(Ntot + 1); inc = dalloc<int>(4 * Etot);
        tet = dalloc<int4>(Etot); rot = dalloc<R>(9 * Etot); rest = dalloc<R>(NR * Etot); kmat = dalloc<R>(3 * Etot);
        ef = dalloc<VT>(12 * Etot); ef2 = dalloc<VT>(12 * Etot);
        if constexpr (sizeof(VT) == sizeof(double)) ef_d = (double*)ef;
        inst = dalloc<InstDev>(S); partial = dalloc<double>(nblk_h.size());
        std::vector<VT> xv(3 * Ntot);
        for (int64_t i = 0; i < 3 * Ntot; i++) xv[i] = (VT)xh[i];
        up(x, xv);
        HIPCHK(hipMemset(v, 0, 3 * Ntot * sizeof(VT)));
        HIPCHK(hipMemset(f, 0, 3 * Ntot * sizeof(VT)));
        HIPCHK(hipMemset(b, 0, 3 * Ntot * sizeof(VT)));
        HIPCHK(hipMemset(r, 0, 3 * Ntot * sizeof(VT)));
        HIPCHK(hipMemset(p, 0, 3 * Ntot * sizeof(VT)));
        HIPCHK(hipMemset(q, 0, 3 * Ntot * sizeof(VT)));
        HIPCHK(hipMemset(xcg, 0, 3 * Ntot * sizeof(VT)));
        up(mass, massh); up(fixed, fixh);
        up(inc_ptr, cnt); up(inc, inch);
        up(tet, teth); up(rot, roth); up(rest, resth); up(kmat, kh);
        up(inst, inst_h);
        HIPCHK(hipMemset(partial, 0, std::max<size_t>(1, nblk_h.size()) * sizeof(double)));
        HIPCHK(hipDeviceSynchronize());
        committed = true;
    }

    void prepare_set(const int32_t* ids, int n, hipStream_t s) {
        std::vector<int32_t> want(ids, ids + n);
        if (want == set_ids && set_eblk) return;
        std::vector<Blk> eb, nb;
        int maxit = 0;
        for (int32_t id : want) {
            const SceneHost& sc = scenes[id];
            eb.insert(eb.end(), eblk_h.begin() + sc.eblk_begin, eblk_h.begin() + sc.eblk_begin + sc.eblk_count);
            nb.insert(nb.end(), nblk_h.begin() + sc.nblk_begin, nblk_h.begin() + sc.nblk_begin + sc.nblk_count);
            maxit = std::max(maxit, sc.cg_iter);
        }
        HIPCHK(hipStreamSynchronize(s));   // the previous lists may still be in use
        if ((int)eb.size() > set_cap_e) {
            if (set_eblk) HIPCHK(hipFree(set_eblk));
            set_cap_e = (int)eb.size();
            set_eblk = dalloc<Blk>(set_cap_e);
        }
        if ((int)nb.size() > set_cap_n) {
            if (set_nblk) HIPCHK(hipFree(set_nblk));
            set_cap_n = (int)nb.size();
            set_nblk = dalloc<Blk>(set_cap_n);
        }
        if (n > set_cap_s) {
            if (set_scenes) HIPCHK(hipFree(set_scenes));
            set_cap_s = n;
            set_scenes = dalloc<int>(set_cap_s);
        }
        if (!eb.empty()) HIPCHK(hipMemcpy(set_eblk, eb.data(), eb.size() * sizeof(Blk), hipMemcpyHostToDevice));
        if (!nb.empty()) HIPCHK(hipMemcpy(set_nblk, nb.data(), nb.size() * sizeof(Blk), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(set_scenes, want.data(), n * sizeof(int), hipMemcpyHostToDevice));
        set_ids = want;
        set_ne = (int)eb.size();
        set_nn = (int)nb.size();
        set_maxit = maxit;
    }

    void step(const int32_t* ids, int n, int nsteps, hipStream_t s) override {
        prepare_set(ids, n, s);
        const ElemView<R, NR> ev = view();
        const NodeArrays na = nodes();
        for (int st = 0; st < nsteps; st++) {
            if (set_ne > 0) k_rhs_elem<R, NR, VT><<<set_ne, kBlock, 0, s>>>(ev, set_eblk, inst, x, v, ef, ef2);
            if (set_nn > 0) k_rhs_node<VT><<<set_nn, kBlock, 0, s>>>(set_nblk, inst, na, ef, ef2, v, f, b, r, xcg, partial);
            for (int it = 0; it < set_maxit; it++) {
                if (set_ne > 0) k_cg_elem<R, NR, VT><<<set_ne, kBlock, 0, s>>>(ev, set_eblk, inst, r, p, ef);
                if (set_nn > 0) {
                    k_cg_node<VT><<<set_nn, kBlock, 0, s>>>(set_nblk, inst, na, ef, r, p, q, partial);
                    k_cg_update<VT><<<set_nn, kBlock, 0, s>>>(set_nblk, inst, p, q, xcg, r, partial);
                }
            }
            if (set_nn > 0) k_euler_update<VT><<<set_nn, kBlock, 0, s>>>(set_nblk, inst, xcg, v, x);
            k_finalize<<<(n + 63) / 64, 64, 0, s>>>(set_scenes, n, inst);
        }
        HIPCHK(hipGetLastError());
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
        const int64_t ge = scenes[scene].elem_begin + (int64_t)e;
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

    double* component_buffer() {
        if (!ef_d) ef_d = dalloc<double>(12 * Etot);
        return ef_d;
    }
    void add_force(int scene, const double* xd, double* fd, hipStream_t s) override {
        const SceneHost& sc = scenes[scene];
        double* efd = component_buffer();
        const int64_t nE = (int64_t)sc.tets.size() / 4;
        if (nE > 0) {
            const int grid = (int)((nE + kBlock - 1) / kBlock);
            k_elem_add_force<R, NR, double><<<grid, kBlock, 0, s>>>(view(), sc.elem_begin, sc.elem_begin + nE,
                                                                   sc.node_begin, xd, efd);
        }
        const int gn = (sc.n_nodes + kBlock - 1) / kBlock;
        if (gn > 0)
            k_gather<double><<<gn, kBlock, 0, s>>>(inc_ptr, inc, sc.node_begin, sc.node_begin + sc.n_nodes,
                                                   efd, fd, 0);
        HIPCHK(hipGetLastError());
    }
    void add_dforce(int scene, const double* dxd, double* dfd, double kf, hipStream_t s) override {
        const SceneHost& sc = scenes[scene];
        double* efd = component_buffer();
        const int64_t nE = (int64_t)sc.tets.size() / 4;
        if (nE > 0) {
            const int grid = (int)((nE + kBlock - 1) / kBlock);
            // addDForce: kFactor = (Real) mparams->kFactor(); applied as `KJtD *= fact`
            k_elem_add_dforce<R, NR, double><<<grid, kBlock, 0, s>>>(view(), sc.elem_begin, sc.elem_begin + nE,
                                                                    sc.node_begin, dxd, (R)kf, efd);
        }
        const int gn = (sc.n_nodes + kBlock - 1) / kBlock;
        if (gn > 0)
            k_gather<double><<<gn, kBlock, 0, s>>>(inc_ptr, inc, sc.node_begin, sc.node_begin + sc.n_nodes,
                                                   efd, dfd, 1);
        HIPCHK(hipGetLastError());
    }
};

}  // namespace sofa_amd

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
    if (precision == SOFA_AMD_FP64) impl.reset(new EngineT<double, 9, double>());
    else if (precision == SOFA_AMD_FP32) impl.reset(new EngineT<float, 6, float>());
    else return fail(SOFA_AMD_EINVAL, "unsupported precision mode");
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

int sofa_amd_fem_add_force(sofa_amd_engine eng, int scene, const double* x_dev, double* f_dev, void* stream) {
    GUARD_BEGIN
    if (!scene_of(eng, scene) || !x_dev || !f_dev) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine not committed");
    HIPCHK(hipSetDevice(eng->impl->device));
    eng->impl->add_force(scene, x_dev, f_dev, stream ? (hipStream_t)stream : eng->impl->stream);
    return SOFA_AMD_OK;
    GUARD_END
}

int sofa_amd_fem_add_dforce(sofa_amd_engine eng, int scene, const double* dx_dev, double* df_dev, double k_factor,
                            void* stream) {
    GUARD_BEGIN
    if (!scene_of(eng, scene) || !dx_dev || !df_dev) return fail(SOFA_AMD_EINVAL, "invalid argument");
    if (!eng->impl->committed) return fail(SOFA_AMD_ESTATE, "engine not committed");
    HIPCHK(hipSetDevice(eng->impl->device));
    eng->impl->add_dforce(scene, dx_dev, df_dev, k_factor, stream ? (hipStream_t)stream : eng->impl->stream);
    return SOFA_AMD_OK;
    GUARD_END
}

}  // extern "C"
