// fem_device.h -- gfx950 device code of the corotational tetrahedral FEM step.
//
// Element math restates TetrahedronFEMForceField (method "large"):
//   computeRotationLarge              TetrahedronFEMForceField.inl:755-778
//   accumulateForceLarge              TetrahedronFEMForceField.inl:870-986
//   applyStiffnessCorotational        TetrahedronFEMForceField.inl:1193-1237
//   computeForce (both forms)         TetrahedronFEMForceField.inl:293-521
//   computeStrainDisplacement / pd    TetrahedronFEMForceField.inl:134-209
// with the same operation order, so that with contraction off the FP64 kernels produce the
// reference's bits.  The strain-displacement matrix J is never stored: its 12 distinct
// entries ("groups" g0..g11, each repeated on three (row, column) positions) are recomputed
// from the element's rotated rest coordinates.  Exact layout (FP64): the 9 rest coordinates
// b, c, d of the local frame (a = 0), so every pseudo-determinant is evaluated exactly as the
// reference does, rounding noise included.  Compact layout (FP32): the 6 coordinates that are
// non-zero in exact arithmetic (b0, c0, c1, d0, d1, d2); g6 = g9 = g10 = 0 there.
//
// Element arrays are structure-of-arrays (one coalesced stream per component); node vectors
// are AoS xyz triples, the layout SOFA hands over (VecTypes.h:200-234), which also keeps the
// 4 node gathers of an element at one or two cache lines each.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include <type_traits>

namespace sofa_amd {

constexpr int kBlock = 256;   // threads per block for every kernel (4 wave64)

// ---------------------------------------------------------------------------------------
// per-scene device record; one per scene, written by the last block of each reduction
// ---------------------------------------------------------------------------------------
struct InstDev {
    // ranges in the flattened index spaces
    int node_begin, node_count, elem_begin, elem_count;
    int nblk_begin, nblk_count;      // node blocks of this scene (reduction partials)
    int mass_kind;                    // 0 uniform, 1 diagonal
    int max_iter;
    double dt, rm, rs;
    double vertex_mass;
    double gravity[3];
    double mg[3];                     // gravity * vertexMass (UniformMass.inl:486-487)
    double k_rhs, m_rhs;              // (h + rs), -rm           (EulerImplicitSolver.cpp:152-154)
    double k_cg, m_cg;                // -h(h + rs), 1 + h rm    (EulerImplicitSolver.cpp:169-171)
    double tol, thr;
    // CG scalars (CGLinearSolver.inl:86-270)
    double normb, rho, rho_1, den, alpha, beta, err;
    int iter;                         // current nb_iter
    int stopped;                      // 1 once the loop has exited
    int end_reason;                   // 0 iterations, 1 tolerance, 2 threshold, 3 den==0, 4 |b|==0
    int first_solve;                  // timeStepCount == 0
    int step_count;
    int last_iters;
    double time;
    unsigned int counter;             // arrivals of the current reduction (exact path)
    int pad_;
    double rho_slot[2];               // rho of iterations k-1 and k (tiled path), by k & 1
    // iterative refinement (MIXED engines; the scene record of the inner FP32 CG is separate)
    int max_sweeps;                   // inner solves per step at most
    int ir_sweeps;                    // inner solves done in the last step
    int cg_total;                     // inner CG iterations summed over the sweeps of a step
    int keep_dir;                     // inner record: sweeps continue the previous search direction
    int sweep;                        // inner record: current sweep (1-based)
    int pad3_;
    double rho_last;                  // inner record: r.r that built the current direction
    double ks[3];                     // one material for the whole scene: K00, K01, K33 times
                                      // 36 V (computeMaterialStiffness before the volume division)
    // EulerImplicitSolver options (EulerImplicitSolver.cpp:40-51,136-173,246-299)
    double b_scale;                   // the right-hand side's factor: h, or 1 with firstOrder
    double vdamp;                     // exp(-h vdamping), 1 without damping
    int first_order;                  // v = dv instead of v += dv
    int project_velocity;             // FixedProjectiveConstraint activate_projectVelocity
    // PlaneForceField (PlaneForceField.inl:86-98), after the FEM in scene order
    double plane_n[3], plane_d, plane_k, plane_damp, plane_max;
    int plane_on, plane_bilateral;
    // tiled engines in the reference's summation order: |b|^2, p.q and r.r of the current
    // iteration, each summed node after node by k_seqdot (MechanicalObject::vDot's order)
    double dots[3];
    // LinearMovementProjectiveConstraint (LinearMovementProjectiveConstraint.inl:160-330), one
    // per scene: lm_n constrained nodes (global ids), the mask each has without the movement's bit,
    // their positions at the first projection (x0, captured then), lm_nk key times followed by
    // 3 lm_nk movements
    const int* lm_node;
    const int8_t* lm_base;
    double* lm_x0;
    const double* lm_keys;
    int lm_n, lm_nk, lm_relative, lm_x0_set;
    // PCGLinearSolver (PCGLinearSolver.inl:178-262) in place of CGLinearSolver: 0 none, 1 with a
    // JacobiPreconditioner, 2 with a BlockJacobiPreconditioner; b.b of the step and the tolerance
    // on r.z (tolerance b.b).  warm_start: CGLinearSolver warmStart (CGLinearSolver.inl:94-105),
    // the solve starting from the previous step's solution
    int pcg, warm_start;
    double bnorm2, pcg_tol;
    // segmented reference-order dot products: boundaries where no candidate matched (counted
    // since the scene was placed)
    int dot_fallbacks, pad_fb;
};

// which entry of InstDev::dots a k_seqdot launch writes
constexpr int kDotBB = 0, kDotPQ = 1, kDotRR = 2;

// MechanicalObject::vDot in the reference's order (MechanicalObject.inl:2334-2356): the scene's
// per-node terms nd[node_begin + i], i = 0 .. n - 1 (each already the node's three-term dot
// product), added one after the other from 0.0.  One wave per scene of the launch set.  The terms
// pass through LDS in chunks of kSeqChunk: the wave loads the next chunk (coalesced) while it sums
// the current one, and every lane runs the same chain of dependent adds over the chunk, reading
// it from LDS (broadcast) in register groups of kSeqGroup terms, the next group's reads issued
// before the current group's adds, so that the chain waits mostly on the FP64 add latency (about
// one add per 6.7 cycles) and little on LDS (measured per term: 2.8 ns for a chain over registers
// alone, 4.2 ns with groups of 16, 3.7 ns with 32, 4.1 ns with 64).  The last chunk is padded with +0.0, which leaves the sum
// unchanged: s + 0.0 == s for every s but -0.0, and a sum that starts at +0.0 is never -0.0 under
// round-to-nearest.  The scene's scalar lands in InstDev::dots[which]; a stopped scene is skipped.
constexpr int kSeqChunk = 1024, kSeqGroup = 32;
__device__ __forceinline__ void seq_group_load(const double* c, double g[kSeqGroup]) {
    const double2* c2 = reinterpret_cast<const double2*>(c);
#pragma unroll
    for (int k = 0; k < kSeqGroup / 2; k++) {
        const double2 t = c2[k];
        g[2 * k] = t.x;
        g[2 * k + 1] = t.y;
    }
}
// the compiler may not move instructions across this point: the next group's LDS reads, issued
// before it, stay ahead of the current group's adds after it
__device__ __forceinline__ void seq_fence() { __builtin_amdgcn_sched_barrier(0); }
__device__ __forceinline__ double seq_group_add(double s, const double g[kSeqGroup]) {
#pragma unroll
    for (int k = 0; k < kSeqGroup; k++) s = s + g[k];
    return s;
}
// the chain of k_seqdot: s + t[begin] + t[begin + 1] + ... + t[end - 1], one dependent add after
// the other from the start value s (which may differ from lane to lane), the terms staged through
// `buf` as described above.  Called by every lane of a 64-thread workgroup (it synchronizes).
__device__ __forceinline__ double seq_chain(const double* __restrict__ t, int begin, int end, double s,
                                            double (*buf)[kSeqChunk]) {
    const int lane = threadIdx.x;
    constexpr int PER = kSeqChunk / 64;
    double reg[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = begin + k * 64 + lane;
        reg[k] = i < end ? t[i] : 0.0;
    }
    int round = 0;
    for (int base = begin; base < end; base += kSeqChunk, round ^= 1) {
#pragma unroll
        for (int k = 0; k < PER; k++) buf[round][k * 64 + lane] = reg[k];
        const int nb = base + kSeqChunk;
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = nb + k * 64 + lane;
            reg[k] = i < end ? t[i] : 0.0;
        }
        __syncthreads();
        const double* c = buf[round];
        double ga[kSeqGroup], gb[kSeqGroup];
        seq_group_load(c, ga);
#pragma unroll
        for (int g = 0; g < kSeqChunk; g += 2 * kSeqGroup) {
            seq_group_load(c + g + kSeqGroup, gb);
            seq_fence();
            s = seq_group_add(s, ga);
            seq_fence();
            // the next pair's first group (after the last pair: the chunk's first, unused)
            seq_group_load(c + ((g + 2 * kSeqGroup) & (kSeqChunk - 1)), ga);
            seq_fence();
            s = seq_group_add(s, gb);
            seq_fence();
        }
        __syncthreads();   // the buffer of this round is rewritten two rounds later
    }
    return s;
}
static __global__ __launch_bounds__(64) void k_seqdot(const int* __restrict__ scenes, InstDev* __restrict__ inst,
                                                       const double* __restrict__ nd, int which) {
    __shared__ __align__(16) double buf[2][kSeqChunk];
    InstDev* ins = &inst[scenes[blockIdx.x]];
    if (ins->stopped) return;
    const double s = seq_chain(nd + ins->node_begin, 0, ins->node_count, 0.0, buf);
    if (threadIdx.x == 0) ins->dots[which] = s;
}

// ---------------------------------------------------------------------------------------
// The same chain in K segments run at once, bit for bit.  A segment's chain can only start once
// the previous segment's sum s_k is known; instead, W waves run segment k from 64 W candidate
// start values at once, one per lane: the doubles whose ordered integer images (dbl_ord) lie
// within 32 W of that of a prediction of s_k (the terms before the segment each rounded as the
// chain would round them, k_seqdot_model).  Afterwards one wave per scene walks
// the boundaries: s_0 is segment 0's sum from +0.0; at boundary k the candidate equal to s_k
// (compared bit for bit) gives s_{k+1} as the sequential chain from s_k would; when no candidate
// equals s_k the wave runs segment k itself from s_k.  Either way every add of the reference's
// chain happens in its order on its exact operands: the result is the sequential sum's bits.
// Segments are aligned on the 64 prediction chunks of a scene (each n / 64 terms rounded up to a
// multiple of 64; k_seqdot_part, k_seqdot_model); scenes below a node count run the single chain.
// ---------------------------------------------------------------------------------------
constexpr int kSpecChunks = 64;   // prediction chunks per scene
// per scene record: [0] segment 0's sum; then per boundary k = 1 .. K - 1: the prediction, then
// the 64 W candidates' sums (1 + (K - 1)(64 W + 1) doubles: K <= 8 with W <= 4, or K = 16 with
// W = 1)
constexpr int kSpecMaxSeg = 8, kSpecMaxWaves = 4;
constexpr int kSpecStride = 1 + (kSpecMaxSeg - 1) * (64 * kSpecMaxWaves + 1);
struct SpecDots {
    double* part;       // kSpecChunks per scene record: plain chunk sums
    double* model;      // kSpecChunks per scene record: chunk sums with the chain's roundings modelled
    double* seg;        // kSpecStride per scene record
    int K, W, min_n;    // segments (a power of two dividing kSpecChunks), candidate waves, least nodes
};
// a monotone integer image of a double (+0.0 and -0.0 both 0) and its inverse (to +0.0)
__device__ __forceinline__ long long dbl_ord(double x) {
    const long long b = __double_as_longlong(x);
    return b >= 0 ? b : -(b & 0x7fffffffffffffffLL);
}
__device__ __forceinline__ double ord_dbl(long long o) {
    return o >= 0 ? __longlong_as_double(o) : __longlong_as_double((-o) | (long long)0x8000000000000000ULL);
}
// candidate j of a boundary (wrapping integer arithmetic: the same bits wherever it is computed)
__device__ __forceinline__ double spec_candidate(double pred, int j, int W) {
    return ord_dbl((long long)((unsigned long long)dbl_ord(pred) + (unsigned long long)(long long)(j - 32 * W)));
}
__device__ __forceinline__ int spec_chunk(int n) { return ((n + kSpecChunks - 1) / kSpecChunks + 63) & ~63; }
__device__ __forceinline__ int spec_bound(int n, int k, int K) {   // first term of segment k
    return min(k * (kSpecChunks / K) * spec_chunk(n), n);
}
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
// prediction chunks: one 64-thread workgroup per (scene, chunk), a plain sum (any order)
static __global__ __launch_bounds__(64) void k_seqdot_part(const int* __restrict__ scenes, const InstDev* __restrict__ inst,
                                                            const double* __restrict__ nd, SpecDots sd) {
    const int sc = scenes[blockIdx.x / kSpecChunks], c = blockIdx.x % kSpecChunks;
    const InstDev* ins = &inst[sc];
    const int n = ins->node_count;
    if (ins->stopped || n < sd.min_n) return;
    const int cs = spec_chunk(n);
    const int i0 = min(c * cs, n), i1 = min(i0 + cs, n);
    const double* t = nd + ins->node_begin;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    int i = i0 + (int)threadIdx.x;
    for (; i + 192 < i1; i += 256) {
        a[0] += t[i]; a[1] += t[i + 64]; a[2] += t[i + 128]; a[3] += t[i + 192];
    }
    for (; i < i1; i += 64) a[0] += t[i];
    const double v = wave_sum_f64((a[0] + a[1]) + (a[2] + a[3]));
    if (threadIdx.x == 0) sd.part[(size_t)sc * kSpecChunks + c] = v;
}
// the prediction: a chunk's terms each rounded as the chain rounds them.  Adding t to a running
// sum s that is a multiple of u = ulp(s + t) gives s + rint(t / u) u (round to nearest, ties to
// even in t / u in place of the sum's last bit) as long as s + t stays in s's binade: terms under
// half an ulp are absorbed, the others rounded to the sum's grid.  The running sum is taken as the
// plain prefix (the chunk sums before the chunk, then a scan of the chunk), which decides only
// the grid u; the chunk's modelled sum goes to sd.model.  One 64-thread workgroup per (scene,
// chunk), after k_seqdot_part.
// t rounded to the grid of a running sum s: rint(t / u) u with u = ulp(s), scaled by powers of
// two (exact); every double is on the grid of a subnormal or zero s
__device__ __forceinline__ double spec_round(double t, double s) {
    const int eb = (int)((__double_as_longlong(s) >> 52) & 0x7ff);
    if (eb == 0) return t;
    const int ue = eb - 1075;   // u = 2^ue
    return ldexp(rint(ldexp(t, -ue)), ue);
}
static __global__ __launch_bounds__(64) void k_seqdot_model(const int* __restrict__ scenes, const InstDev* __restrict__ inst,
                                                             const double* __restrict__ nd, SpecDots sd) {
    const int sc = scenes[blockIdx.x / kSpecChunks], c = blockIdx.x % kSpecChunks;
    const InstDev* ins = &inst[sc];
    const int n = ins->node_count;
    if (ins->stopped || n < sd.min_n) return;
    const int lane = threadIdx.x;
    const int cs = spec_chunk(n);
    const int i0 = min(c * cs, n), i1 = min(i0 + cs, n);
    const double* t = nd + ins->node_begin;
    double run = wave_sum_f64(lane < c ? sd.part[(size_t)sc * kSpecChunks + lane] : 0.0);
    double acc = 0.0;
    constexpr int G = 8;   // groups of 64 terms loaded ahead (all in flight), then scanned
    for (int g0 = i0; g0 < i1; g0 += G * 64) {
        double tg[G];
#pragma unroll
        for (int k = 0; k < G; k++) {
            const int i = g0 + k * 64 + lane;
            tg[k] = i < i1 ? t[i] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < G; k++) {
            const double ti = tg[k];
            double x = ti;   // inclusive scan of the group's terms
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const double y = __shfl_up(x, d, 64);
                if (lane >= d) x += y;
            }
            acc += wave_sum_f64(spec_round(ti, run + x));
            run += __shfl(x, 63, 64);
        }
    }
    if (lane == 0) sd.model[(size_t)sc * kSpecChunks + c] = acc;
}
// the segments: 1 + (K - 1) W workgroups of 64 per scene (slot 0: segment 0 from +0.0; slot
// 1 + (k - 1) W + w: segment k from candidates w 64 .. w 64 + 63); a scene below min_n nodes
// runs its whole chain in slot 0 and writes the result
static __global__ __launch_bounds__(64) void k_seqdot_spec(const int* __restrict__ scenes, InstDev* __restrict__ inst,
                                                            const double* __restrict__ nd, SpecDots sd, int which) {
    __shared__ __align__(16) double buf[2][kSeqChunk];
    const int slots = 1 + (sd.K - 1) * sd.W;
    const int sc = scenes[blockIdx.x / slots], slot = blockIdx.x % slots;
    InstDev* ins = &inst[sc];
    if (ins->stopped) return;
    const int n = ins->node_count;
    const double* t = nd + ins->node_begin;
    const int lane = threadIdx.x;
    if (n < sd.min_n) {
        if (slot != 0) return;
        const double s = seq_chain(t, 0, n, 0.0, buf);
        if (lane == 0) ins->dots[which] = s;
        return;
    }
    double* out = sd.seg + (size_t)sc * kSpecStride;
    if (slot == 0) {
        const double s = seq_chain(t, 0, spec_bound(n, 1, sd.K), 0.0, buf);
        if (lane == 0) out[0] = s;
        return;
    }
    const int k = 1 + (slot - 1) / sd.W, w = (slot - 1) % sd.W;
    const int nc = k * (kSpecChunks / sd.K);   // chunks before the segment
    const double pred = wave_sum_f64(lane < nc ? sd.model[(size_t)sc * kSpecChunks + lane] : 0.0);
    const double s = seq_chain(t, spec_bound(n, k, sd.K), spec_bound(n, k + 1, sd.K),
                               spec_candidate(pred, w * 64 + lane, sd.W), buf);
    double* o = out + 1 + (size_t)(k - 1) * (64 * sd.W + 1);
    if (w == 0 && lane == 0) o[0] = pred;
    o[1 + w * 64 + lane] = s;
}
// the boundaries, one wave per scene (every lane the same walk)
static __global__ __launch_bounds__(64) void k_seqdot_resolve(const int* __restrict__ scenes, InstDev* __restrict__ inst,
                                                               const double* __restrict__ nd, SpecDots sd, int which) {
    __shared__ __align__(16) double buf[2][kSeqChunk];
    const int sc = scenes[blockIdx.x];
    InstDev* ins = &inst[sc];
    const int n = ins->node_count;
    if (ins->stopped || n < sd.min_n) return;
    const double* out = sd.seg + (size_t)sc * kSpecStride;
    const double* t = nd + ins->node_begin;
    double s = out[0];
    for (int k = 1; k < sd.K; k++) {
        const double* o = out + 1 + (size_t)(k - 1) * (64 * sd.W + 1);
        const double pred = o[0];
        const long long d = (long long)((unsigned long long)dbl_ord(s) - (unsigned long long)dbl_ord(pred)) + 32LL * sd.W;
        bool hit = false;
        if (sd.W > 0 && d >= 0 && d < 64LL * sd.W) {
            const int j = (int)d;
            hit = __double_as_longlong(spec_candidate(pred, j, sd.W)) == __double_as_longlong(s);
            if (hit) s = o[1 + j];
        }
        if (!hit) {
            s = seq_chain(t, spec_bound(n, k, sd.K), spec_bound(n, k + 1, sd.K), s, buf);
            if (threadIdx.x == 0) ins->dot_fallbacks++;
        }
    }
    if (threadIdx.x == 0) ins->dots[which] = s;
}

struct Blk { int inst, begin, end, pad; };   // a block's scene and [begin, end) range

// ---------------------------------------------------------------------------------------
// scalar helpers
// ---------------------------------------------------------------------------------------
template <typename R> struct Eps;
template <> struct Eps<double> { static constexpr double v = DBL_EPSILON; };
template <> struct Eps<float> { static constexpr float v = FLT_EPSILON; };

template <typename R>
__device__ __forceinline__ R pdet(R m00, R m01, R m02, R m10, R m11, R m12) {
    // peudo_determinant_for_coef, TetrahedronFEMForceField.inl:205-209
    return m01 * m12 - m11 * m02 - m00 * m12 + m10 * m02 + m00 * m11 - m10 * m01;
}

// normalize per Vec.h:535-578: divide by sqrt(norm2) unless the norm is <= epsilon
template <typename R>
__device__ __forceinline__ void normalize3(R& x, R& y, R& z) {
#pragma clang fp contract(off)
    R n2 = x * x;
    n2 += y * y;
    n2 += z * z;
    const R n = sqrt(n2);
    if (n > Eps<R>::v) { x /= n; y /= n; z /= n; }
}

// FP32: one hardware reciprocal square root (v_rsq_f32, 1 ulp) and three multiplies instead of a
// correctly rounded square root and three divisions.  The FP32 engine's error budget is far above
// the ulp this costs; every FP32 kernel calls this same function, so a rotation recomputed from
// the same positions still has the same bits.  The FP64 engines keep the divisions.
template <>
__device__ __forceinline__ void normalize3<float>(float& x, float& y, float& z) {
#pragma clang fp contract(off)
    float n2 = x * x;
    n2 += y * y;
    n2 += z * z;
    if (n2 > Eps<float>::v * Eps<float>::v) {
        const float s = __builtin_amdgcn_rsqf(n2);
        x *= s; y *= s; z *= s;
    }
}

// PlaneForceField::addForce for one node (PlaneForceField.inl:158-205): adds the penalty force
// of a node below the plane (either side if bilateral) to f; returns whether it is in contact
template <typename VT>
__device__ __forceinline__ bool plane_force(const InstDev& in, const VT x[3], const VT v[3], VT f[3]) {
#pragma clang fp contract(off)
    const VT n[3] = {(VT)in.plane_n[0], (VT)in.plane_n[1], (VT)in.plane_n[2]};
    VT dist = x[0] * n[0];
    dist += x[1] * n[1];
    dist += x[2] * n[2];
    dist = dist - (VT)in.plane_d;
    if (!(in.plane_bilateral || dist < 0)) return false;
    const VT fi = -(VT)in.plane_k * dist;
    const VT di = -(VT)in.plane_damp * dist;
    VT force[3];
#pragma unroll
    for (int k = 0; k < 3; k++) force[k] = n[k] * fi - v[k] * di;
    VT limit = (VT)in.plane_max;
    limit *= limit;
    VT amp = force[0] * force[0];
    amp += force[1] * force[1];
    amp += force[2] * force[2];
    if (limit > 0 && amp > limit) {
        const VT sc = sqrt(limit / amp);
#pragma unroll
        for (int k = 0; k < 3; k++) force[k] *= sc;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) f[k] += force[k];
    return true;
}

// PlaneForceField::addDForce for a node in contact (PlaneForceField.inl:208-224):
// df += n (fact (dx . n)), fact = -stiffness kFactor
template <typename VT>
__device__ __forceinline__ void plane_dforce(const InstDev& in, double kfactor, const VT dx[3], VT df[3]) {
#pragma clang fp contract(off)
    const VT n[3] = {(VT)in.plane_n[0], (VT)in.plane_n[1], (VT)in.plane_n[2]};
    const VT fact = (VT)(-in.plane_k * kfactor);
    VT dn = dx[0] * n[0];
    dn += dx[1] * n[1];
    dn += dx[2] * n[2];
    const VT a = fact * dn;
#pragma unroll
    for (int k = 0; k < 3; k++) df[k] = df[k] + n[k] * a;
}

template <typename T> __device__ __forceinline__ void load3(const T* __restrict__ v, int n, T& a, T& b, T& c) {
    const T* p = v + 3 * (int64_t)n;
    a = p[0]; b = p[1]; c = p[2];
}

// ---------------------------------------------------------------------------------------
// element record view
// ---------------------------------------------------------------------------------------
template <typename R, int NR>
struct ElemView {
    const int4* __restrict__ tet;   // global node ids
    R* __restrict__ rot;            // 9 SoA streams: rotations[e](i,j) at rot[(3i+j)*E + e]
    const R* __restrict__ rest;     // NR SoA streams of rotated rest coordinates
    const R* __restrict__ kmat;     // 3 SoA streams: K00, K01, K33 (already / (36 V))
    int64_t E;                      // stride of the SoA streams
};

// the 12 distinct strain-displacement entries of one element
template <typename R, int NR>
__device__ __forceinline__ void jgroups(const ElemView<R, NR>& ev, int64_t e, R g[12], R rc[9]) {
    if constexpr (NR == 9) {
#pragma unroll
        for (int k = 0; k < 9; k++) rc[k] = ev.rest[k * ev.E + e];
        const R a0 = 0, a1 = 0, a2 = 0;
        const R b0 = rc[0], b1 = rc[1], b2 = rc[2];
        const R c0 = rc[3], c1 = rc[4], c2 = rc[5];
        const R d0 = rc[6], d1 = rc[7], d2 = rc[8];
        g[0] = -pdet(b1, c1, d1, b2, c2, d2);
        g[1] = pdet(b0, c0, d0, b2, c2, d2);
        g[2] = -pdet(b0, c0, d0, b1, c1, d1);
        g[3] = pdet(c1, d1, a1, c2, d2, a2);
        g[4] = -pdet(c0, d0, a0, c2, d2, a2);
        g[5] = pdet(c0, d0, a0, c1, d1, a1);
        g[6] = -pdet(d1, a1, b1, d2, a2, b2);
        g[7] = pdet(d0, a0, b0, d2, a2, b2);
        g[8] = -pdet(d0, a0, b0, d1, a1, b1);
        g[9] = pdet(a1, b1, c1, a2, b2, c2);
        g[10] = -pdet(a0, b0, c0, a2, b2, c2);
        g[11] = pdet(a0, b0, c0, a1, b1, c1);
    } else {
        // compact: rc = b0, c0, c1, d0, d1, d2 (b1 = b2 = c2 = 0)
        R r6[6];
#pragma unroll
        for (int k = 0; k < 6; k++) r6[k] = ev.rest[k * ev.E + e];
        const R b0 = r6[0], c0 = r6[1], c1 = r6[2], d0 = r6[3], d1 = r6[4], d2 = r6[5];
        rc[0] = b0; rc[1] = 0; rc[2] = 0; rc[3] = c0; rc[4] = c1; rc[5] = 0; rc[6] = d0; rc[7] = d1; rc[8] = d2;
        g[0] = -(c1 * d2);
        g[1] = c0 * d2 - b0 * d2;
        g[2] = -(((c0 * d1 - c1 * d0) - b0 * d1) + b0 * c1);
        g[3] = c1 * d2;
        g[4] = -(c0 * d2);
        g[5] = c0 * d1 - c1 * d0;
        g[6] = 0;
        g[7] = d2 * b0;
        g[8] = -(d1 * b0);
        g[9] = 0;
        g[10] = 0;
        g[11] = b0 * c1;
    }
}

// JtD = J^T D (12 -> 6), the reference's term order (zero groups dropped in compact layout)
template <typename R, int NR>
__device__ __forceinline__ void jt_mul(const R g[12], const R D[12], R JtD[6]) {
    if constexpr (NR == 9) {
        JtD[0] = g[0] * D[0] + g[3] * D[3] + g[6] * D[6] + g[9] * D[9];
        JtD[1] = g[1] * D[1] + g[4] * D[4] + g[7] * D[7] + g[10] * D[10];
        JtD[2] = g[2] * D[2] + g[5] * D[5] + g[8] * D[8] + g[11] * D[11];
        JtD[3] = g[1] * D[0] + g[0] * D[1] + g[4] * D[3] + g[3] * D[4] + g[7] * D[6] + g[6] * D[7] + g[10] * D[9] + g[9] * D[10];
        JtD[4] = g[2] * D[1] + g[1] * D[2] + g[5] * D[4] + g[4] * D[5] + g[8] * D[7] + g[7] * D[8] + g[11] * D[10] + g[10] * D[11];
        JtD[5] = g[2] * D[0] + g[0] * D[2] + g[5] * D[3] + g[3] * D[5] + g[8] * D[6] + g[6] * D[8] + g[11] * D[9] + g[9] * D[11];
    } else {
        JtD[0] = g[0] * D[0] + g[3] * D[3];
        JtD[1] = g[1] * D[1] + g[4] * D[4] + g[7] * D[7];
        JtD[2] = g[2] * D[2] + g[5] * D[5] + g[8] * D[8] + g[11] * D[11];
        JtD[3] = g[1] * D[0] + g[0] * D[1] + g[4] * D[3] + g[3] * D[4] + g[7] * D[6];
        JtD[4] = g[2] * D[1] + g[1] * D[2] + g[5] * D[4] + g[4] * D[5] + g[8] * D[7] + g[7] * D[8] + g[11] * D[10];
        JtD[5] = g[2] * D[0] + g[0] * D[2] + g[5] * D[3] + g[3] * D[5] + g[8] * D[6] + g[11] * D[9];
    }
}

// KJtD = K JtD, then F = J KJtD (6 -> 12)
template <typename R>
__device__ __forceinline__ void k_mul(const R k[3], const R JtD[6], R KJ[6]) {
    const R K00 = k[0], K01 = k[1], K33 = k[2];
    KJ[0] = K00 * JtD[0] + K01 * JtD[1] + K01 * JtD[2];
    KJ[1] = K01 * JtD[0] + K00 * JtD[1] + K01 * JtD[2];
    KJ[2] = K01 * JtD[0] + K01 * JtD[1] + K00 * JtD[2];
    KJ[3] = K33 * JtD[3];
    KJ[4] = K33 * JtD[4];
    KJ[5] = K33 * JtD[5];
}

template <typename R>
__device__ __forceinline__ void j_mul(const R g[12], const R KJ[6], R F[12]) {
#pragma unroll
    for (int n = 0; n < 4; n++) {
        const R gA = g[3 * n], gB = g[3 * n + 1], gC = g[3 * n + 2];
        F[3 * n + 0] = gA * KJ[0] + gB * KJ[3] + gC * KJ[5];
        F[3 * n + 1] = gB * KJ[1] + gA * KJ[3] + gC * KJ[4];
        F[3 * n + 2] = gC * KJ[2] + gB * KJ[4] + gA * KJ[5];
    }
}

// c = rot * (F[3n], F[3n+1], F[3n+2]) (Mat.h:577-587 order)
template <typename R>
__device__ __forceinline__ void rot_apply(const R rot[9], const R F[12], R c[12]) {
#pragma unroll
    for (int n = 0; n < 4; n++) {
#pragma unroll
        for (int i = 0; i < 3; i++)
            c[3 * n + i] = rot[3 * i + 0] * F[3 * n] + rot[3 * i + 1] * F[3 * n + 1] + rot[3 * i + 2] * F[3 * n + 2];
    }
}

// ---------------------------------------------------------------------------------------
// element force: accumulateForceLarge; returns the rotation and the 4 nodal contributions
// (to be ADDED to f); positions read as VT and converted to R
// ---------------------------------------------------------------------------------------
// computeRotationLarge (TetrahedronFEMForceField.inl:755-778) from the element's first three
// vertices: rot = transpose(R_0_2), the rows of R_0_2 being ex, ey, ez.  Multiplies and adds are
// never fused here, so that every kernel that needs the rotation (the force pass, and the tiled
// operator, which recomputes it from the positions instead of reading it back) gets the same bits.
template <typename R>
__device__ __forceinline__ void rotation_large(const R p0[3], const R p1[3], const R p2[3], R rot[9]) {
#pragma clang fp contract(off)
    R ex0 = p1[0] - p0[0], ex1 = p1[1] - p0[1], ex2 = p1[2] - p0[2];
    normalize3(ex0, ex1, ex2);
    R ey0 = p2[0] - p0[0], ey1 = p2[1] - p0[1], ey2 = p2[2] - p0[2];
    R ez0 = ex1 * ey2 - ex2 * ey1, ez1 = ex2 * ey0 - ex0 * ey2, ez2 = ex0 * ey1 - ex1 * ey0;
    normalize3(ez0, ez1, ez2);
    ey0 = ez1 * ex2 - ez2 * ex1;
    ey1 = ez2 * ex0 - ez0 * ex2;
    ey2 = ez0 * ex1 - ez1 * ex0;
    rot[0] = ex0; rot[1] = ey0; rot[2] = ez0;
    rot[3] = ex1; rot[4] = ey1; rot[5] = ez1;
    rot[6] = ex2; rot[7] = ey2; rot[8] = ez2;
}

// initLarge (TetrahedronFEMForceField.inl:835-868): the rest frame R_0_1 from the rest positions
// of vertices a, b, c, and the rest shape in it relative to a, in the element set-up's order
// (each coordinate a three-term dot product, then the difference to vertex a).  out: NR = 9,
// b, c and d with their rounding residue; NR = 6, b0, c0, c1, d0, d1, d2.
template <typename R, int NR>
__device__ __forceinline__ void rest_frame(const R p0[4][3], R* out) {
#pragma clang fp contract(off)
    R rot[9];
    rotation_large(p0[0], p0[1], p0[2], rot);      // rot[3 i + j] = rows[j][i]
    R ri[4][3];
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            R r = rot[i] * p0[k][0];
            r += rot[3 + i] * p0[k][1];
            r += rot[6 + i] * p0[k][2];
            ri[k][i] = r;
        }
#pragma unroll
    for (int k = 1; k < 4; k++)
#pragma unroll
        for (int i = 0; i < 3; i++) ri[k][i] -= ri[0][i];
    if (NR == 9) {
#pragma unroll
        for (int k = 1; k < 4; k++)
#pragma unroll
            for (int i = 0; i < 3; i++) out[3 * (k - 1) + i] = ri[k][i];
    } else {
        out[0] = ri[1][0];
        out[1] = ri[2][0]; out[2] = ri[2][1];
        out[3] = ri[3][0]; out[4] = ri[3][1]; out[5] = ri[3][2];
    }
}

// computeMaterialStiffness's last step for a scene of one material: K = ks / (36 V), V the
// absolute volume of the rest tetrahedron (Tetrahedron.h:55-58), in the element set-up's order
template <typename R>
__device__ __forceinline__ void stiffness_from_volume(const R p0[4][3], const double ks[3], R k[3]) {
#pragma clang fp contract(off)
    const R a0 = p0[1][0] - p0[0][0], a1 = p0[1][1] - p0[0][1], a2 = p0[1][2] - p0[0][2];
    const R b0 = p0[2][0] - p0[0][0], b1 = p0[2][1] - p0[0][1], b2 = p0[2][2] - p0[0][2];
    const R c0 = p0[3][0] - p0[0][0], c1 = p0[3][1] - p0[0][1], c2 = p0[3][2] - p0[0][2];
    const R x0 = a1 * b2 - a2 * b1, x1 = a2 * b0 - a0 * b2, x2 = a0 * b1 - a1 * b0;
    R d = x0 * c0;
    d += x1 * c1;
    d += x2 * c2;
    if constexpr (std::is_same<R, float>::value) {
        // FP32: 36 V = 6 |d|, one hardware reciprocal (1 ulp) for the three entries
        const float inv = __builtin_amdgcn_rcpf(fabsf(d) * 6.0f);
#pragma unroll
        for (int i = 0; i < 3; i++) k[i] = (float)ks[i] * inv;
    } else {
        const R vol = fabs(d / (R)6);
        const R div = vol * (R)36;
#pragma unroll
        for (int i = 0; i < 3; i++) k[i] = (R)ks[i] / div;
    }
}

// accumulateForceLarge's element part once the rest shape is known: rc the rest coordinates
// b, c, d in the rest frame (zeros included), g their strain-displacement groups, k3 the
// stiffness entries
template <typename R, int NR>
__device__ __forceinline__ void elem_force_core(const R p[4][3], const R rc[9], const R g[12], const R k3[3],
                                                R rot[9], R c[12]) {
    rotation_large(p[0], p[1], p[2], rot);
    // R_0_2 = transpose(rot)
    const R r02[3][3] = {{rot[0], rot[3], rot[6]}, {rot[1], rot[4], rot[7]}, {rot[2], rot[5], rot[8]}};
    // deforme[k] = R_0_2 * p_k, then relative to deforme[0] on the used components
    R df[4][3];
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int i = 0; i < 3; i++) df[k][i] = r02[i][0] * p[k][0] + r02[i][1] * p[k][1] + r02[i][2] * p[k][2];
    df[1][0] -= df[0][0];
    df[2][0] -= df[0][0];
    df[2][1] -= df[0][1];
    df[3][0] -= df[0][0];
    df[3][1] -= df[0][1];
    df[3][2] -= df[0][2];
    R D[12];
    D[0] = 0; D[1] = 0; D[2] = 0;
    D[3] = rc[0] - df[1][0];
    D[4] = 0; D[5] = 0;
    D[6] = rc[3] - df[2][0];
    D[7] = rc[4] - df[2][1];
    D[8] = 0;
    D[9] = rc[6] - df[3][0];
    D[10] = rc[7] - df[3][1];
    D[11] = rc[8] - df[3][2];
    R JtD[6], KJ[6], F[12];
    jt_mul<R, NR>(g, D, JtD);
    k_mul(k3, JtD, KJ);
    j_mul(g, KJ, F);
    rot_apply(rot, F, c);
}

template <typename R, int NR>
__device__ __forceinline__ void elem_force_p(const ElemView<R, NR>& ev, int64_t e, const R p[4][3], R rot[9],
                                             R c[12], R g[12]) {
    R rc[9];
    jgroups<R, NR>(ev, e, g, rc);
    const R k3[3] = {ev.kmat[e], ev.kmat[ev.E + e], ev.kmat[2 * ev.E + e]};
    elem_force_core<R, NR>(p, rc, g, k3, rot, c);
}

template <typename R, int NR, typename VT>
__device__ __forceinline__ void elem_force(const ElemView<R, NR>& ev, int64_t e, const int4 t,
                                           const VT* __restrict__ x, R rot[9], R c[12], R g[12]) {
    VT xa0, xa1, xa2, xb0, xb1, xb2, xc0, xc1, xc2, xd0, xd1, xd2;
    load3(x, t.x, xa0, xa1, xa2);
    load3(x, t.y, xb0, xb1, xb2);
    load3(x, t.z, xc0, xc1, xc2);
    load3(x, t.w, xd0, xd1, xd2);
    const R p[4][3] = {{(R)xa0, (R)xa1, (R)xa2}, {(R)xb0, (R)xb1, (R)xb2},
                       {(R)xc0, (R)xc1, (R)xc2}, {(R)xd0, (R)xd1, (R)xd2}};
    elem_force_p<R, NR>(ev, e, p, rot, c, g);
}

// ---------------------------------------------------------------------------------------
// element dforce: applyStiffnessCorotational; returns the 4 nodal contributions (to be
// SUBTRACTED from df).  dx for node k is supplied by the caller (already gathered).
// ---------------------------------------------------------------------------------------
template <typename R, int NR>
__device__ __forceinline__ void elem_dforce(const ElemView<R, NR>& ev, int64_t e, const R rot[9],
                                            const R u[12], R fact, R c[12]) {
    R X[12];
#pragma unroll
    for (int n = 0; n < 4; n++) {
#pragma unroll
        for (int j = 0; j < 3; j++)
            X[3 * n + j] = rot[0 * 3 + j] * u[3 * n] + rot[1 * 3 + j] * u[3 * n + 1] + rot[2 * 3 + j] * u[3 * n + 2];
    }
    R g[12], rc[9];
    jgroups<R, NR>(ev, e, g, rc);
    R JtD[6], KJ[6], F[12];
    jt_mul<R, NR>(g, X, JtD);
    const R k3[3] = {ev.kmat[e], ev.kmat[ev.E + e], ev.kmat[2 * ev.E + e]};
    k_mul(k3, JtD, KJ);
#pragma unroll
    for (int k = 0; k < 6; k++) KJ[k] *= fact;
    j_mul(g, KJ, F);
    rot_apply(rot, F, c);
}

// the same 12 strain-displacement groups from rest coordinates already in registers
// (NR == 9: b, c, d; NR == 6: b0, c0, c1, d0, d1, d2)
template <typename R, int NR>
__device__ __forceinline__ void jgroups_r(const R* rin, R g[12]) {
    if constexpr (NR == 9) {
        const R a0 = 0, a1 = 0, a2 = 0;
        const R b0 = rin[0], b1 = rin[1], b2 = rin[2];
        const R c0 = rin[3], c1 = rin[4], c2 = rin[5];
        const R d0 = rin[6], d1 = rin[7], d2 = rin[8];
        g[0] = -pdet(b1, c1, d1, b2, c2, d2);
        g[1] = pdet(b0, c0, d0, b2, c2, d2);
        g[2] = -pdet(b0, c0, d0, b1, c1, d1);
        g[3] = pdet(c1, d1, a1, c2, d2, a2);
        g[4] = -pdet(c0, d0, a0, c2, d2, a2);
        g[5] = pdet(c0, d0, a0, c1, d1, a1);
        g[6] = -pdet(d1, a1, b1, d2, a2, b2);
        g[7] = pdet(d0, a0, b0, d2, a2, b2);
        g[8] = -pdet(d0, a0, b0, d1, a1, b1);
        g[9] = pdet(a1, b1, c1, a2, b2, c2);
        g[10] = -pdet(a0, b0, c0, a2, b2, c2);
        g[11] = pdet(a0, b0, c0, a1, b1, c1);
    } else {
        const R b0 = rin[0], c0 = rin[1], c1 = rin[2], d0 = rin[3], d1 = rin[4], d2 = rin[5];
        g[0] = -(c1 * d2);
        g[1] = c0 * d2 - b0 * d2;
        g[2] = -(((c0 * d1 - c1 * d0) - b0 * d1) + b0 * c1);
        g[3] = c1 * d2;
        g[4] = -(c0 * d2);
        g[5] = c0 * d1 - c1 * d0;
        g[6] = 0;
        g[7] = d2 * b0;
        g[8] = -(d1 * b0);
        g[9] = 0;
        g[10] = 0;
        g[11] = b0 * c1;
    }
}

// the element force with the element record in registers (rin: NR rest coordinates)
template <typename R, int NR>
__device__ __forceinline__ void elem_force_r(const R p[4][3], const R* rin, const R k3[3], R rot[9], R c[12]) {
    R g[12], rc[9];
    jgroups_r<R, NR>(rin, g);
    if constexpr (NR == 9) {
#pragma unroll
        for (int k = 0; k < 9; k++) rc[k] = rin[k];
    } else {
        rc[0] = rin[0]; rc[1] = 0; rc[2] = 0; rc[3] = rin[1]; rc[4] = rin[2]; rc[5] = 0;
        rc[6] = rin[3]; rc[7] = rin[4]; rc[8] = rin[5];
    }
    elem_force_core<R, NR>(p, rc, g, k3, rot, c);
}

// applyStiffnessCorotational with the element record in registers
template <typename R, int NR>
__device__ __forceinline__ void elem_dforce_r(const R rot[9], const R* rin, const R k3[3], const R u[12],
                                              R fact, R c[12]) {
    R X[12];
#pragma unroll
    for (int n = 0; n < 4; n++) {
#pragma unroll
        for (int j = 0; j < 3; j++)
            X[3 * n + j] = rot[0 * 3 + j] * u[3 * n] + rot[1 * 3 + j] * u[3 * n + 1] + rot[2 * 3 + j] * u[3 * n + 2];
    }
    R g[12];
    jgroups_r<R, NR>(rin, g);
    R JtD[6], KJ[6], F[12];
    jt_mul<R, NR>(g, X, JtD);
    k_mul(k3, JtD, KJ);
#pragma unroll
    for (int k = 0; k < 6; k++) KJ[k] *= fact;
    j_mul(g, KJ, F);
    rot_apply(rot, F, c);
}

template <typename R, int NR>
__device__ __forceinline__ void load_rot(const ElemView<R, NR>& ev, int64_t e, R rot[9]) {
#pragma unroll
    for (int k = 0; k < 9; k++) rot[k] = ev.rot[k * ev.E + e];
}

// ---------------------------------------------------------------------------------------
// block reduction (fixed tree, replayed by the oracle's dot_blocked) and the last-block
// pattern: every block writes its partial, the last block to arrive (integer ticket, no
// float atomics) reduces the scene's partials in a fixed order.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v, double* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
#pragma unroll
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = sh[t] + sh[t + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// returns true in every thread of the last-arriving block of scene `ins`; `total` then holds
// the scene's sum over its node blocks
__device__ __forceinline__ bool last_block_reduce(double blockval, double* partial, InstDev* ins,
                                                  int blk, double* sh, double& total) {
    __shared__ int am_last;
    if (threadIdx.x == 0) {
        partial[blk] = blockval;
        __threadfence();
        const unsigned ticket = atomicAdd(&ins->counter, 1u);
        am_last = (ticket == (unsigned)(ins->nblk_count - 1)) ? 1 : 0;
    }
    __syncthreads();
    if (!am_last) return false;
    __threadfence();
    const int nb = ins->nblk_count;
    const int b0 = ins->nblk_begin;
    double acc = 0.0;
    for (int k = threadIdx.x; k < nb; k += kBlock) acc = acc + partial[b0 + k];
    total = block_sum(acc, sh);
    if (threadIdx.x == 0) atomicExch(&ins->counter, 0u);
    return true;
}

// the same hand-off with the reference's summation order: every thread has stored its node's
// value in nd; the last-arriving block adds nd[node_begin], nd[node_begin + 1], ... one after
// the other from 0.0, as MechanicalObject::vDot accumulates (MechanicalObject.inl:2334-2356).
// One wave does the adds (each lane of a 64-value chunk read by every lane in order), so the
// cost is one dependent add per node of the scene.
__device__ __forceinline__ bool last_block_sequential(const double* nd, InstDev* ins, double& total) {
    __shared__ int am_last;
    __threadfence();                  // this thread's nd store
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned ticket = atomicAdd(&ins->counter, 1u);
        am_last = (ticket == (unsigned)(ins->nblk_count - 1)) ? 1 : 0;
    }
    __syncthreads();
    if (!am_last) return false;
    __threadfence();
    double s = 0.0;
    if (threadIdx.x < 64) {
        const int b0 = ins->node_begin, nn = ins->node_count;
        for (int base = 0; base < nn; base += 64) {
            const int i = base + (int)threadIdx.x;
            const double v = i < nn ? nd[b0 + i] : 0.0;
            const int cnt = min(64, nn - base);
            // lane k's value read into a scalar register (constant lane index), added in order
            const int lo = __double2loint(v), hi = __double2hiint(v);
#pragma unroll
            for (int k = 0; k < 64; k++) {
                if (k < cnt)
                    s = s + __hiloint2double(__builtin_amdgcn_readlane(hi, k), __builtin_amdgcn_readlane(lo, k));
            }
        }
    }
    total = s;                        // valid in the first wave
    __syncthreads();
    if (threadIdx.x == 0) atomicExch(&ins->counter, 0u);
    return true;
}

}  // namespace sofa_amd
