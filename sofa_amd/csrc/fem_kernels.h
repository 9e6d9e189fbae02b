// fem_kernels.h -- the kernels of one implicit Euler step, batched over scenes.
//
// One launch per phase covers every scene of the launch set: element kernels walk a list of
// element blocks, node kernels a list of node blocks; each block belongs to exactly one scene
// (Blk.inst), so per-scene scalars (CG alpha, beta, stop flags) are block-uniform and a scene
// that has left its CG loop costs one early return per block.
//
// Step schedule (EulerImplicitSolver.cpp:83-300 + CGLinearSolver.inl:71-315):
//   k_rhs_elem    rotations from x, element force F(x), element K v          (1 element pass)
//   k_rhs_node    f = m g + sum F; b = h (f - rm M v - (h+rs) K v); project;
//                 r = b; x_cg = 0; |b| and rho = r.r                          (1 node pass)
//   per CG iteration:
//     k_cg_elem   p = r + beta p at the 4 gathered nodes; element K p        (1 element pass)
//     k_cg_node   p (owned node); q = (1 + h rm) M p - h(h+rs) K p; project; p.q
//     k_cg_update x_cg += alpha p; r -= alpha q; r.r
//   k_euler_update  v += x_cg; x += h v
//   k_finalize      time += dt, timeStepCount++
// Node sums run over each node's incidence list in ascending element id: the order in which
// the reference's element loop accumulates into that node.  No float atomics anywhere.
#pragma once

#include "fem_device.h"

namespace sofa_amd {

struct NodeArrays {
    const int* __restrict__ inc_ptr;   // per node [inc_ptr[n], inc_ptr[n+1]) in inc
    const int* __restrict__ inc;       // element*4 + local vertex, ascending element id
    const double* __restrict__ mass;   // per-node mass (DiagonalMass, MeshMatrixMass vertex mass)
    const int8_t* __restrict__ fixed;  // non-zero where a projective constraint clears the response:
                                       // 1 a fixed node, 2 a node driven by an oscillator
    const int* __restrict__ em_ptr;    // MeshMatrixMass edges of each node, ascending edge id:
    const int* __restrict__ em_nbr;    //   the other end
    const double* __restrict__ em_val; //   the edge mass rho V / 20 summed over the edge's tets
    int8_t* __restrict__ contact;      // PlaneForceField: 1 where the node touched the plane at the
                                       // step's force evaluation (its addDForce contact list)
};

// mass kinds of a scene
constexpr int kMassUniform = 0, kMassDiagonal = 1, kMassMatrix = 2, kMassMatrixLumped = 3;
constexpr double kLumpTet = 2.5;   // MeshMatrixMass lumping coefficient of tetrahedra (MeshMatrixMass.inl:1476-1489)

// vertex part of addMDx: UniformMass.inl:403-419, DiagonalMass.inl:536-559, MeshMatrixMass.inl:1987-2030
template <typename VT>
__device__ __forceinline__ void mass_mdx(const InstDev& in, const NodeArrays& na, int n, double factor,
                                         const VT d[3], VT out[3]) {
    if (in.mass_kind == kMassUniform) {
        VT m = (VT)in.vertex_mass;
        if (factor != 1.0) m *= (VT)factor;
        out[0] += d[0] * m; out[1] += d[1] * m; out[2] += d[2] * m;
    } else if (in.mass_kind == kMassDiagonal) {
        const VT m = (VT)na.mass[n];
        if (factor == 1.0) {
            out[0] += d[0] * m; out[1] += d[1] * m; out[2] += d[2] * m;
        } else {
            const VT fct = (VT)factor;
            out[0] += (d[0] * m) * fct; out[1] += (d[1] * m) * fct; out[2] += (d[2] * m) * fct;
        }
    } else if (in.mass_kind == kMassMatrix) {
        const VT m = (VT)na.mass[n], fct = (VT)factor;
        out[0] += (d[0] * m) * fct; out[1] += (d[1] * m) * fct; out[2] += (d[2] * m) * fct;
    } else {
        const VT m = (VT)na.mass[n], fct = (VT)factor, lc = (VT)kLumpTet;
        out[0] += ((d[0] * m) * lc) * fct; out[1] += ((d[1] * m) * lc) * fct; out[2] += ((d[2] * m) * lc) * fct;
    }
}

// gravity of the mass's addForce: UniformMass.inl:471-496 (m g), DiagonalMass.inl:1393-1413
// (g m_i), MeshMatrixMass.inl:2073-2092 (g m_i 2.5)
template <typename VT>
__device__ __forceinline__ void mass_gravity(const InstDev& in, const double* __restrict__ mass, int n, VT f[3]) {
    if (in.mass_kind == kMassUniform) {
        f[0] = (VT)0 + (VT)in.mg[0]; f[1] = (VT)0 + (VT)in.mg[1]; f[2] = (VT)0 + (VT)in.mg[2];
    } else if (in.mass_kind == kMassDiagonal) {
        const VT m = (VT)mass[n];
#pragma unroll
        for (int i = 0; i < 3; i++) f[i] = (VT)0 + (VT)in.gravity[i] * m;
    } else {
        const VT m = (VT)mass[n], lc = (VT)kLumpTet;
#pragma unroll
        for (int i = 0; i < 3; i++) f[i] = (VT)0 + ((VT)in.gravity[i] * m) * lc;
    }
}

// edge part of the consistent MeshMatrixMass addMDx (MeshMatrixMass.inl:2014-2026): for each
// edge of node n in ascending edge id, out += d[other] * (edgeMass * factor); `dof(j, v)` gives
// the value of the multiplied vector at node j
template <typename VT, typename F>
__device__ __forceinline__ void mass_edges(const NodeArrays& na, int n, double factor, F dof, VT out[3]) {
    const int s0 = na.em_ptr[n], s1 = na.em_ptr[n + 1];
    const VT fct = (VT)factor;
    for (int s = s0; s < s1; s++) {
        const VT tm = (VT)na.em_val[s] * fct;
        VT d[3];
        dof(na.em_nbr[s], d);
        out[0] += d[0] * tm; out[1] += d[1] * tm; out[2] += d[2] * tm;
    }
}

template <typename VT>
__device__ __forceinline__ double dot3d(const VT a[3], const VT b[3]) {
    double r = (double)a[0] * (double)b[0];
    r += (double)a[1] * (double)b[1];
    r += (double)a[2] * (double)b[2];
    return r;
}

// z = P^-1 r with P^-1 stored as 9 doubles per node (row-major): DiagonalMatrix::mult
// (z_i = d_i r_i, DiagonalMatrix.h:187-190) or BlockDiagonalMatrix::mul (z_j = sum_i B(i, j) r_i
// from 0, BlockDiagonalMatrix.h:202-221) -- the preconditioner solve of PCGLinearSolver
template <typename VT>
__device__ __forceinline__ void precond_apply(int kind, const double* __restrict__ P, const VT r[3], VT z[3]) {
    if (kind == 1) {
#pragma unroll
        for (int a = 0; a < 3; a++) z[a] = (VT)(P[4 * a] * (double)r[a]);
    } else {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
#pragma unroll
            for (int a = 0; a < 3; a++) acc += P[3 * a + j] * (double)r[a];
            z[j] = (VT)acc;
        }
    }
}

__device__ __forceinline__ int blk_scene(const Blk& b) { return b.inst; }

// CGLinearSolver.inl:86-160 for nb_iter = 1 (rho = r.r = b.b), once |b|^2 is known: the exact
// path's CG scalars of a scene at the start of a solve
__device__ __forceinline__ void cg_start(InstDev* ins, double total) {
    ins->normb = sqrt(total);
    ins->rho = total;
    ins->rho_1 = 0.0;
    ins->iter = 1;
    ins->stopped = 0;
    ins->end_reason = 0;
    ins->err = 1.0;
    ins->den = 0.0;
    if (ins->normb == 0.0) {
        ins->stopped = 1; ins->end_reason = 4; ins->last_iters = 0;
    } else {
        const double err = sqrt(total) / ins->normb;
        ins->err = err;
        if (err <= ins->tol && !ins->first_solve) { ins->stopped = 1; ins->end_reason = 1; ins->last_iters = 1; }
    }
}

// CGLinearSolver's first tests after a warm start: |b| from b.b, rho = r.r of r = b - A x, the
// iteration-1 tolerance test with the first step's exemption (CGLinearSolver.inl:94-105,139-170)
__device__ __forceinline__ void cg_start_warm(InstDev* ins, double bb, double rr) {
    ins->normb = sqrt(bb);
    ins->rho = rr;
    ins->rho_1 = 0.0;
    ins->iter = 1;
    ins->stopped = 0;
    ins->end_reason = 0;
    ins->err = 1.0;
    ins->den = 0.0;
    if (ins->normb == 0.0) {
        ins->stopped = 1; ins->end_reason = 4; ins->last_iters = 0;
    } else {
        const double err = sqrt(rr) / ins->normb;
        ins->err = err;
        if (err <= ins->tol && !ins->first_solve) { ins->stopped = 1; ins->end_reason = 1; ins->last_iters = 1; }
    }
}

// end of a component-level CG solve (no time step): the loop's exit record, timeStepCount++
// (CGLinearSolver.inl:300); reopen = 1 for the tiled kernels, which run with stopped == 0
static __global__ void k_solve_end(const int* __restrict__ scenes, int n, InstDev* __restrict__ inst, int reopen) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    InstDev* in = &inst[scenes[i]];
    if (!in->stopped) { in->stopped = 1; in->end_reason = 0; in->last_iters = in->max_iter; }
    in->first_solve = 0;
    in->step_count += 1;
    if (reopen) in->stopped = 0;
}

// ---------------------------------------------------------------------------------------
// RHS
// ---------------------------------------------------------------------------------------
template <typename R, int NR, typename VT>
__global__ __launch_bounds__(kBlock) void k_rhs_elem(ElemView<R, NR> ev, const Blk* __restrict__ eblk,
                                                      const InstDev* __restrict__ inst, const VT* __restrict__ x,
                                                      const VT* __restrict__ v, VT* __restrict__ ef_f,
                                                      VT* __restrict__ ef_v) {
    const Blk bk = eblk[blockIdx.x];
    const int64_t e = (int64_t)bk.begin + threadIdx.x;
    if (e >= bk.end) return;
    const InstDev& in = inst[bk.inst];
    const int4 t = ev.tet[e];
    R rot[9], c[12], g[12];
    elem_force<R, NR, VT>(ev, e, t, x, rot, c, g);
#pragma unroll
    for (int k = 0; k < 9; k++) ev.rot[k * ev.E + e] = rot[k];
#pragma unroll
    for (int k = 0; k < 12; k++) ef_f[12 * e + k] = (VT)c[k];
    // FEM part of addMBKv(b, v): kFactor = h + rs, skipped when 0 (BaseForceField.cpp:38-47)
    if (in.k_rhs != 0.0) {
        R u[12];
        const int nid[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
        for (int n = 0; n < 4; n++) {
            VT a, b, cc;
            load3(v, nid[n], a, b, cc);
            u[3 * n] = (R)a; u[3 * n + 1] = (R)b; u[3 * n + 2] = (R)cc;
        }
        elem_dforce<R, NR>(ev, e, rot, u, (R)in.k_rhs, c);
#pragma unroll
        for (int k = 0; k < 12; k++) ef_v[12 * e + k] = (VT)c[k];
    } else {
#pragma unroll
        for (int k = 0; k < 12; k++) ef_v[12 * e + k] = (VT)0;
    }
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void k_rhs_node(const Blk* __restrict__ nblk, InstDev* __restrict__ inst,
                                                      NodeArrays na, const VT* __restrict__ ef_f,
                                                      const VT* __restrict__ ef_v, const VT* __restrict__ v,
                                                      VT* __restrict__ f, VT* __restrict__ b, VT* __restrict__ r,
                                                      VT* __restrict__ xcg, double* __restrict__ partial,
                                                      double* __restrict__ nd,     // non-null: reference dot order
                                                      const VT* __restrict__ x) {
    __shared__ double sh[kBlock];
    const Blk bk = nblk[blockIdx.x];
    InstDev* ins = &inst[bk.inst];
    const int n = bk.begin + threadIdx.x;
    double bb = 0.0;
    if (n < bk.end) {
        const InstDev& in = *ins;
        VT xn[3] = {0, 0, 0}, vn[3] = {0, 0, 0};
        if (in.plane_on) {
            load3(x, n, xn[0], xn[1], xn[2]);
            load3(v, n, vn[0], vn[1], vn[2]);
        }
        // computeForce: f = 0; mass addForce (f += m g); FEM addForce (f += R F, element order)
        VT fn[3];
        mass_gravity(in, na.mass, n, fn);
        const int s0 = na.inc_ptr[n], s1 = na.inc_ptr[n + 1];
        for (int s = s0; s < s1; s++) {
            const int64_t slot = na.inc[s];
            fn[0] += ef_f[3 * slot]; fn[1] += ef_f[3 * slot + 1]; fn[2] += ef_f[3 * slot + 2];
        }
        // PlaneForceField after the FEM (scene order); its contacts are kept for addDForce
        const bool touch = in.plane_on && plane_force(in, xn, vn, fn);
        if (in.plane_on) na.contact[n] = touch ? 1 : 0;
        f[3 * (int64_t)n] = fn[0]; f[3 * (int64_t)n + 1] = fn[1]; f[3 * (int64_t)n + 2] = fn[2];
        // b = f; b += (-rm) M v (mass first); b -= (h+rs) K v (FEM, element order); b *= h
        VT bn[3] = {fn[0], fn[1], fn[2]};
        if (in.m_rhs != 0.0) {
            VT vm[3];
            load3(v, n, vm[0], vm[1], vm[2]);
            mass_mdx(in, na, n, in.m_rhs, vm, bn);
            if (in.mass_kind == kMassMatrix)
                mass_edges(na, n, in.m_rhs, [&](int j, VT d[3]) { load3(v, j, d[0], d[1], d[2]); }, bn);
        }
        for (int s = s0; s < s1; s++) {
            const int64_t slot = na.inc[s];
            bn[0] -= ef_v[3 * slot]; bn[1] -= ef_v[3 * slot + 1]; bn[2] -= ef_v[3 * slot + 2];
        }
        if (touch && in.k_rhs != 0.0) plane_dforce(in, in.k_rhs, vn, bn);
        const VT h = (VT)in.b_scale;   // h (firstOrder: 1, b = f)
        bn[0] *= h; bn[1] *= h; bn[2] *= h;
        if (na.fixed[n]) { bn[0] = 0; bn[1] = 0; bn[2] = 0; }
        const int64_t o = 3 * (int64_t)n;
        b[o] = bn[0]; b[o + 1] = bn[1]; b[o + 2] = bn[2];
        r[o] = bn[0]; r[o + 1] = bn[1]; r[o + 2] = bn[2];
        // PCG and a CG with warmStart start from the last solution
        if (!in.pcg && !in.warm_start) { xcg[o] = 0; xcg[o + 1] = 0; xcg[o + 2] = 0; }
        bb = dot3d(bn, bn);
        if (nd) nd[n] = bb;
    }
    double total;
    const bool last = nd ? last_block_sequential(nd, ins, total)
                         : last_block_reduce(block_sum(bb, sh), partial, ins, bk.pad, sh, total);
    if (last && threadIdx.x == 0) {
        if (ins->pcg || ins->warm_start) ins->bnorm2 = total;
        else cg_start(ins, total);
    }
}

// the exact path's start of CGLinearSolver::solve on a caller's right-hand side (component call):
// r = b, x = 0, |b|^2 and the first iteration's tests as after the RHS
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_cg_begin(const Blk* __restrict__ nblk, InstDev* __restrict__ inst,
                                                      const VT* __restrict__ b_in, int nb0, VT* __restrict__ b,
                                                      VT* __restrict__ r, VT* __restrict__ xcg,
                                                      double* __restrict__ partial, double* __restrict__ nd) {
    __shared__ double sh[kBlock];
    const Blk bk = nblk[blockIdx.x];
    InstDev* ins = &inst[bk.inst];
    const int n = bk.begin + threadIdx.x;
    double bb = 0.0;
    if (n < bk.end) {
        VT bn[3];
        load3(b_in, n - nb0, bn[0], bn[1], bn[2]);
        const int64_t o = 3 * (int64_t)n;
        b[o] = bn[0]; b[o + 1] = bn[1]; b[o + 2] = bn[2];
        r[o] = bn[0]; r[o + 1] = bn[1]; r[o + 2] = bn[2];
        xcg[o] = 0; xcg[o + 1] = 0; xcg[o + 2] = 0;
        bb = dot3d(bn, bn);
        if (nd) nd[n] = bb;
    }
    double total;
    const bool last = nd ? last_block_sequential(nd, ins, total)
                         : last_block_reduce(block_sum(bb, sh), partial, ins, bk.pad, sh, total);
    if (last && threadIdx.x == 0) cg_start(ins, total);
}

// ---------------------------------------------------------------------------------------
// CG iteration
// ---------------------------------------------------------------------------------------
template <typename R, int NR, typename VT>
__global__ __launch_bounds__(kBlock) void k_cg_elem(ElemView<R, NR> ev, const Blk* __restrict__ eblk,
                                                     const InstDev* __restrict__ inst, const VT* __restrict__ r_,
                                                     const VT* __restrict__ p_old, VT* __restrict__ ef,
                                                     const VT* __restrict__ z = nullptr) {
    const Blk bk = eblk[blockIdx.x];
    const InstDev& in = inst[bk.inst];
    if (in.stopped) return;
    // PCG: the direction is built from the preconditioned residual z (w = z + beta w)
    const VT* __restrict__ r = (in.pcg && z) ? z : r_;
    const int64_t e = (int64_t)bk.begin + threadIdx.x;
    if (e >= bk.end) return;
    const int4 t = ev.tet[e];
    const int first = in.iter == 1;
    const VT beta = (VT)in.beta;
    R u[12];
    const int nid[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int n = 0; n < 4; n++) {
        VT r0, r1, r2;
        load3(r, nid[n], r0, r1, r2);
        if (!first) {
            // cgstep_beta, GraphScattered: p *= beta; p += r (CGLinearSolver.cpp:40-44)
            VT p0, p1, p2;
            load3(p_old, nid[n], p0, p1, p2);
            p0 *= beta; p0 += r0;
            p1 *= beta; p1 += r1;
            p2 *= beta; p2 += r2;
            r0 = p0; r1 = p1; r2 = p2;
        }
        u[3 * n] = (R)r0; u[3 * n + 1] = (R)r1; u[3 * n + 2] = (R)r2;
    }
    R rot[9], c[12];
    load_rot(ev, e, rot);
    elem_dforce<R, NR>(ev, e, rot, u, (R)in.k_cg, c);
#pragma unroll
    for (int k = 0; k < 12; k++) ef[12 * e + k] = (VT)c[k];
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void k_cg_node(const Blk* __restrict__ nblk, InstDev* __restrict__ inst,
                                                     NodeArrays na, const VT* __restrict__ ef,
                                                     const VT* __restrict__ r_, const VT* __restrict__ p_old,
                                                     VT* __restrict__ p, VT* __restrict__ q,
                                                     double* __restrict__ partial, double* __restrict__ nd,
                                                     const VT* __restrict__ z = nullptr) {
    __shared__ double sh[kBlock];
    const Blk bk = nblk[blockIdx.x];
    InstDev* ins = &inst[bk.inst];
    if (ins->stopped) return;
    const VT* __restrict__ r = (ins->pcg && z) ? z : r_;
    const int n = bk.begin + threadIdx.x;
    double pq = 0.0;
    if (n < bk.end) {
        const InstDev& in = *ins;
        const int64_t o = 3 * (int64_t)n;
        const bool first = in.iter == 1;
        const VT beta = (VT)in.beta;
        // p = r (first iteration) or p *= beta; p += r (cgstep_beta), from the previous p
        auto pdir = [&](int j, VT d[3]) {
            load3(r, j, d[0], d[1], d[2]);
            if (!first) {
                VT po[3];
                load3(p_old, j, po[0], po[1], po[2]);
#pragma unroll
                for (int k = 0; k < 3; k++) { po[k] *= beta; po[k] += d[k]; d[k] = po[k]; }
            }
        };
        VT pn[3];
        pdir(n, pn);
        p[o] = pn[0]; p[o + 1] = pn[1]; p[o + 2] = pn[2];
        // q = 0; mass addMDx(q, p, 1 + h rm); FEM addDForce(q, p, -h(h+rs)); projectResponse(q)
        VT qn[3] = {0, 0, 0};
        if (in.m_cg != 0.0) {
            mass_mdx(in, na, n, in.m_cg, pn, qn);
            if (in.mass_kind == kMassMatrix) mass_edges(na, n, in.m_cg, pdir, qn);
        }
        const int s0 = na.inc_ptr[n], s1 = na.inc_ptr[n + 1];
        for (int s = s0; s < s1; s++) {
            const int64_t slot = na.inc[s];
            qn[0] -= ef[3 * slot]; qn[1] -= ef[3 * slot + 1]; qn[2] -= ef[3 * slot + 2];
        }
        if (in.plane_on && na.contact[n]) plane_dforce(in, in.k_cg, pn, qn);
        if (na.fixed[n]) { qn[0] = 0; qn[1] = 0; qn[2] = 0; }
        q[o] = qn[0]; q[o + 1] = qn[1]; q[o + 2] = qn[2];
        pq = dot3d(pn, qn);
        if (nd) nd[n] = pq;
    }
    double den;
    const bool last = nd ? last_block_sequential(nd, ins, den)
                         : last_block_reduce(block_sum(pq, sh), partial, ins, bk.pad, sh, den);
    if (last && threadIdx.x == 0) {
        ins->den = den;
        if (ins->pcg) {                     // PCGLinearSolver.inl:222-224: alpha = r.z / w.s
            ins->alpha = ins->rho / den;
            return;
        }
        // CGLinearSolver.inl:207-265
        const bool exempt = (ins->iter == 1 && ins->first_solve);
        if (den != 0.0) {
            if (fabs(den) <= ins->thr && !exempt) {
                ins->stopped = 1; ins->end_reason = 2; ins->last_iters = ins->iter;
            } else {
                ins->alpha = ins->rho / den;
            }
        } else {
            ins->stopped = 1; ins->end_reason = 3; ins->last_iters = ins->iter;
        }
    }
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void k_cg_update(const Blk* __restrict__ nblk, InstDev* __restrict__ inst,
                                                       const VT* __restrict__ p, const VT* __restrict__ q,
                                                       VT* __restrict__ xcg, VT* __restrict__ r,
                                                       double* __restrict__ partial, double* __restrict__ nd,
                                                       const double* __restrict__ pinv = nullptr,
                                                       VT* __restrict__ z = nullptr) {
    __shared__ double sh[kBlock];
    const Blk bk = nblk[blockIdx.x];
    InstDev* ins = &inst[bk.inst];
    if (ins->stopped) return;
    const int n = bk.begin + threadIdx.x;
    double rr = 0.0;
    if (n < bk.end) {
        // cgstep_alpha: x += p alpha; r += q (-alpha) (CGLinearSolver.cpp:46-64, vOp_v_inc_bf)
        const VT a = (VT)ins->alpha, ma = (VT)(-ins->alpha);
        const int64_t o = 3 * (int64_t)n;
        VT rn[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            xcg[o + k] += p[o + k] * a;
            rn[k] = r[o + k] + q[o + k] * ma;
            r[o + k] = rn[k];
        }
        if (ins->pcg && pinv) {             // s = P^-1 r; r_norm = r.s (PCGLinearSolver.inl:229-244)
            VT zn[3];
            precond_apply(ins->pcg, pinv + 9 * (int64_t)n, rn, zn);
            z[o] = zn[0]; z[o + 1] = zn[1]; z[o + 2] = zn[2];
            rr = dot3d(rn, zn);
        } else {
            rr = dot3d(rn, rn);
        }
        if (nd) nd[n] = rr;
    }
    double rho;
    const bool last = nd ? last_block_sequential(nd, ins, rho)
                         : last_block_reduce(block_sum(rr, sh), partial, ins, bk.pad, sh, rho);
    if (last && threadIdx.x == 0) {
        // rho_1 = rho; next iteration: rho = r.r, err test, beta = rho / rho_1
        ins->rho_1 = ins->rho;
        ins->rho = rho;
        const int it = ins->iter + 1;
        ins->iter = it;
        if (ins->pcg) {                     // while (iter <= maxIter && r_norm > tol)
            ins->err = ins->bnorm2 != 0.0 ? rho / ins->bnorm2 : 0.0;
            if (it > ins->max_iter || !(rho > ins->pcg_tol)) {
                ins->stopped = 1; ins->end_reason = it > ins->max_iter ? 0 : 1; ins->last_iters = it;
            } else {
                ins->beta = rho / ins->rho_1;
            }
            return;
        }
        if (it > ins->max_iter) {
            ins->stopped = 1; ins->end_reason = 0; ins->last_iters = ins->max_iter;
        } else {
            const double err = sqrt(rho) / ins->normb;
            ins->err = err;
            if (err <= ins->tol) {
                ins->stopped = 1; ins->end_reason = 1; ins->last_iters = it;
            } else {
                ins->beta = rho / ins->rho_1;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// integration and bookkeeping
// ---------------------------------------------------------------------------------------
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_euler_update(const Blk* __restrict__ nblk, const InstDev* __restrict__ inst,
                                                          const int8_t* __restrict__ fixed,
                                                          const VT* __restrict__ xcg, VT* __restrict__ v,
                                                          VT* __restrict__ x) {
    const Blk bk = nblk[blockIdx.x];
    const int n = bk.begin + threadIdx.x;
    if (n >= bk.end) return;
    // MechanicalObject::vMultiOp fast path (MechanicalObject.inl:2209-2230): v += dv (firstOrder:
    // v = dv), x += v h; then vdamping's v *= exp(-h d) (EulerImplicitSolver.cpp:246-299)
    const InstDev& in = inst[bk.inst];
    const VT h = (VT)in.dt;
    const int64_t o = 3 * (int64_t)n;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const VT vv = in.first_order ? xcg[o + k] : v[o + k] + xcg[o + k];
        x[o + k] = x[o + k] + vv * h;
        v[o + k] = in.vdamp != 1.0 ? vv * (VT)in.vdamp : vv;
    }
    // FixedProjectiveConstraint::projectVelocity (.inl:237-258), when activated
    if (in.project_velocity && (fixed[n] & 1)) {
#pragma unroll
        for (int k = 0; k < 3; k++) v[o + k] = 0;
    }
}

static __global__ void k_finalize(const int* __restrict__ scenes, int n, InstDev* __restrict__ inst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    InstDev* in = &inst[scenes[i]];
    if (!in->stopped) { in->stopped = 1; in->last_iters = in->max_iter; }
    in->first_solve = 0;
    in->step_count += 1;
    in->time += in->dt;
    in->lm_x0_set = 1;         // LinearMovementProjectiveConstraint's x0 captured by this step
}

// ---------------------------------------------------------------------------------------
// OscillatorProjectiveConstraint::projectPosition / projectVelocity (OscillatorProjective
// Constraint.inl:77-106) at the context time of the step start: x = mean + a sin(w t + phi),
// v = a w cos(w t + phi)
// ---------------------------------------------------------------------------------------
struct Oscillators {
    const int* __restrict__ node;        // global node id
    const double* __restrict__ mean;     // 3 per entry
    const double* __restrict__ amp;      // 3 per entry
    const double* __restrict__ w;        // pulsation
    const double* __restrict__ phase;
};

template <typename VT>
__global__ __launch_bounds__(kBlock) void k_oscillate(const Blk* __restrict__ oblk, const InstDev* __restrict__ inst,
                                                       Oscillators os, VT* __restrict__ x, VT* __restrict__ v) {
    const Blk bk = oblk[blockIdx.x];
    const int i = bk.begin + threadIdx.x;
    if (i >= bk.end) return;
    const double t = inst[bk.inst].time;
    const double w = os.w[i], ph = os.phase[i];
    const double sn = sin(w * t + ph), cs = cos(w * t + ph);
    const int64_t o = 3 * (int64_t)os.node[i];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        x[o + k] = (VT)(os.mean[3 * i + k] + os.amp[3 * i + k] * sn);
        v[o + k] = (VT)((os.amp[3 * i + k] * w) * cs);
    }
}

// ---------------------------------------------------------------------------------------
// LinearMovementProjectiveConstraint at the context time of the step start
// ---------------------------------------------------------------------------------------
// findKeyTimes (LinearMovementProjectiveConstraint.inl:296-330): true when cT lies in
// [keyTimes[k], keyTimes[k+1]) for some k (the reference's "finished && nextT != prevT"), with the
// bracketing key times and movements
__device__ inline bool lm_window(const InstDev& in, double cT, double& pT, double& nT, double pM[3], double nM[3]) {
    const int nk = in.lm_nk;
    const double* kt = in.lm_keys;
    const double* km = in.lm_keys + nk;
    if (nk == 0 || !(cT >= kt[0] && cT <= kt[nk - 1])) return false;
    nT = kt[0];
    pT = nT;
    for (int k = 0; k < nk; k++) {
        if (kt[k] <= cT) {
            pT = kt[k];
            for (int c = 0; c < 3; c++) pM[c] = km[3 * k + c];
        } else {
            nT = kt[k];
            for (int c = 0; c < 3; c++) nM[c] = km[3 * k + c];
            return nT != pT;
        }
    }
    return false;
}

// the response mask of the constrained nodes for the step starting at the scene's time: the
// movement's bit (4) while a movement is on (projectResponseT, :160-183, clears their responses)
template <bool SCENE>
__global__ __launch_bounds__(kBlock) void k_lm_mask(const Blk* __restrict__ lblk, const InstDev* __restrict__ inst,
                                                     int sid, int8_t* __restrict__ fixed) {
    int si, i;
    if (SCENE) {
        si = sid;
        i = blockIdx.x * kBlock + threadIdx.x;
        if (i >= inst[si].lm_n) return;
    } else {
        const Blk bk = lblk[blockIdx.x];
        si = bk.inst;
        i = bk.begin + threadIdx.x;
        if (i >= bk.end) return;
    }
    const InstDev& in = inst[si];
    double pT = 0, nT = 0, pM[3] = {0, 0, 0}, nM[3] = {0, 0, 0};
    const bool on = lm_window(in, in.time, pT, nT, pM, nM);
    fixed[in.lm_node[i]] = (int8_t)(in.lm_base[i] | (on ? 4 : 0));
}

// projectPosition then projectVelocity (:193-244, in the order of MechanicalProjectPosition-
// AndVelocityVisitor.cpp:35-40): x0 captured on the first projection; inside a movement window
// m = prevM + (nextM - prevM) (cT - prevT) / (nextT - prevT), x = x0 + m (relative) or m,
// v = (nextM - prevM) (1 / (nextT - prevT)).  The scene's time is the step's start; the capture
// flag is set by the step-end kernel.
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_lm_project(const Blk* __restrict__ lblk, InstDev* __restrict__ inst,
                                                        VT* __restrict__ x, VT* __restrict__ v) {
    const Blk bk = lblk[blockIdx.x];
    const int i = bk.begin + threadIdx.x;
    if (i >= bk.end) return;
    const InstDev& in = inst[bk.inst];
    const int64_t o = 3 * (int64_t)in.lm_node[i];
    double* x0 = in.lm_x0 + 3 * (int64_t)i;
    if (!in.lm_x0_set)
        for (int c = 0; c < 3; c++) x0[c] = (double)x[o + c];
    double pT = 0, nT = 0, pM[3] = {0, 0, 0}, nM[3] = {0, 0, 0};
    if (!lm_window(in, in.time, pT, nT, pM, nM)) return;
    const double dt = (in.time - pT) / (nT - pT);
    const double inv = 1.0 / (nT - pT);
    for (int c = 0; c < 3; c++) {
        const double m = pM[c] + (nM[c] - pM[c]) * dt;
        x[o + c] = (VT)(in.lm_relative ? x0[c] + m : m);
        v[o + c] = (VT)((nM[c] - pM[c]) * inv);
    }
}

// the same on one caller vector of one scene at time t (component-level projectPosition /
// projectVelocity); the caller's vector starts at the scene's first node nb0
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_lm_vec(const InstDev* __restrict__ inst, int sid, double t, int nb0,
                                                    VT* __restrict__ vec, int position) {
    const InstDev& in = inst[sid];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= in.lm_n) return;
    const int64_t o = 3 * (int64_t)(in.lm_node[i] - nb0);
    double* x0 = in.lm_x0 + 3 * (int64_t)i;
    if (position && !in.lm_x0_set)
        for (int c = 0; c < 3; c++) x0[c] = (double)vec[o + c];
    double pT = 0, nT = 0, pM[3] = {0, 0, 0}, nM[3] = {0, 0, 0};
    if (!lm_window(in, t, pT, nT, pM, nM)) return;
    const double dt = (t - pT) / (nT - pT);
    const double inv = 1.0 / (nT - pT);
    for (int c = 0; c < 3; c++) {
        const double m = pM[c] + (nM[c] - pM[c]) * dt;
        vec[o + c] = position ? (VT)(in.lm_relative ? x0[c] + m : m) : (VT)((nM[c] - pM[c]) * inv);
    }
}

static __global__ void k_lm_mark_x0(InstDev* __restrict__ inst, int sid) {
    if (threadIdx.x == 0) inst[sid].lm_x0_set = 1;
}

// ---------------------------------------------------------------------------------------
// component-level kernels (one TetrahedronFEMForceField call on caller vectors)
// ---------------------------------------------------------------------------------------
// tet holds flattened node ids; the caller's vectors start at the scene's first node `nb`
template <typename R, int NR, typename VT>
__global__ __launch_bounds__(kBlock) void k_elem_add_force(ElemView<R, NR> ev, int64_t e0, int64_t e1, int nb,
                                                            const VT* __restrict__ x, VT* __restrict__ ef) {
    const int64_t e = e0 + (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= e1) return;
    int4 t = ev.tet[e];
    t.x -= nb; t.y -= nb; t.z -= nb; t.w -= nb;
    R rot[9], c[12], g[12];
    elem_force<R, NR, VT>(ev, e, t, x, rot, c, g);
#pragma unroll
    for (int k = 0; k < 9; k++) ev.rot[k * ev.E + e] = rot[k];
#pragma unroll
    for (int k = 0; k < 12; k++) ef[12 * e + k] = (VT)c[k];
}

template <typename R, int NR, typename VT>
__global__ __launch_bounds__(kBlock) void k_elem_add_dforce(ElemView<R, NR> ev, int64_t e0, int64_t e1, int nb,
                                                             const VT* __restrict__ dx, R fact, VT* __restrict__ ef) {
    const int64_t e = e0 + (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= e1) return;
    const int4 t = ev.tet[e];
    const int nid[4] = {t.x - nb, t.y - nb, t.z - nb, t.w - nb};
    R u[12];
#pragma unroll
    for (int n = 0; n < 4; n++) {
        VT a, b, c;
        load3(dx, nid[n], a, b, c);
        u[3 * n] = (R)a; u[3 * n + 1] = (R)b; u[3 * n + 2] = (R)c;
    }
    R rot[9], c[12];
    load_rot(ev, e, rot);
    elem_dforce<R, NR>(ev, e, rot, u, fact, c);
#pragma unroll
    for (int k = 0; k < 12; k++) ef[12 * e + k] = (VT)c[k];
}

// out[n] (+/-)= sum of the node's contributions; node ids are global, `out` is indexed from
// the scene's first node
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_gather(const int* __restrict__ inc_ptr, const int* __restrict__ inc,
                                                    int n0, int n1, const VT* __restrict__ ef, VT* __restrict__ out,
                                                    int subtract) {
    const int n = n0 + blockIdx.x * kBlock + threadIdx.x;
    if (n >= n1) return;
    const int64_t o = 3 * (int64_t)(n - n0);
    VT a[3] = {out[o], out[o + 1], out[o + 2]};
    const int s0 = inc_ptr[n], s1 = inc_ptr[n + 1];
    if (subtract) {
        for (int s = s0; s < s1; s++) {
            const int64_t slot = inc[s];
            a[0] -= ef[3 * slot]; a[1] -= ef[3 * slot + 1]; a[2] -= ef[3 * slot + 2];
        }
    } else {
        for (int s = s0; s < s1; s++) {
            const int64_t slot = inc[s];
            a[0] += ef[3 * slot]; a[1] += ef[3 * slot + 1]; a[2] += ef[3 * slot + 2];
        }
    }
    out[o] = a[0]; out[o + 1] = a[1]; out[o + 2] = a[2];
}

// ---------------------------------------------------------------------------------------
// component-level kernels of the masses, the projective constraints and the state (one call of
// Mass / ProjectiveConstraintSet / MechanicalState on caller vectors of one scene; the caller's
// vectors are indexed from the scene's first node nb0, engine tables by global node id)
// ---------------------------------------------------------------------------------------
// Mass::addMDx: res += factor M dx -- UniformMass.inl:403-419 (m *= factor once, then dx m),
// DiagonalMass.inl:536-559, MeshMatrixMass.inl:1987-2047 (vertex term, then the consistent mass's
// edges in ascending edge id: per node the same order as the reference's two loops)
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_mass_mdx(const InstDev* __restrict__ inst, int sid, NodeArrays na, int nb0,
                                                      int n, const VT* __restrict__ dx, VT* __restrict__ res,
                                                      double factor) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const InstDev& in = inst[sid];
    VT d[3], o[3];
    load3(dx, i, d[0], d[1], d[2]);
    load3(res, i, o[0], o[1], o[2]);
    mass_mdx(in, na, nb0 + i, factor, d, o);
    if (in.mass_kind == kMassMatrix)
        mass_edges(na, nb0 + i, factor, [&](int j, VT e[3]) { load3(dx, j - nb0, e[0], e[1], e[2]); }, o);
    const int64_t q = 3 * (int64_t)i;
    res[q] = o[0]; res[q + 1] = o[1]; res[q + 2] = o[2];
}

// Mass::accFromF: a = M^-1 f -- UniformMass.inl:423-436 (f / m), DiagonalMass.inl:564-576
// (f / m_i), MeshMatrixMass.inl:2051-2069 (lumped only: f / (m_i 2.5))
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_mass_acc(const InstDev* __restrict__ inst, int sid,
                                                      const double* __restrict__ mass, int nb0, int n,
                                                      const VT* __restrict__ f, VT* __restrict__ a) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const InstDev& in = inst[sid];
    VT m;
    if (in.mass_kind == kMassUniform) m = (VT)in.vertex_mass;
    else if (in.mass_kind == kMassDiagonal) m = (VT)mass[nb0 + i];
    else m = (VT)mass[nb0 + i] * (VT)kLumpTet;
    const int64_t q = 3 * (int64_t)i;
#pragma unroll
    for (int k = 0; k < 3; k++) a[q + k] = f[q + k] / m;
}

// Mass::addForce (gravity): UniformMass.inl:471-496 (f += g m), DiagonalMass.inl:1393-1413
// (f += g m_i), MeshMatrixMass.inl:2073-2092 (f += g m_i 2.5, lumped coefficient in both modes)
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_mass_add_force(const InstDev* __restrict__ inst, int sid,
                                                            const double* __restrict__ mass, int nb0, int n,
                                                            VT* __restrict__ f) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const InstDev& in = inst[sid];
    const int64_t q = 3 * (int64_t)i;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        VT w;
        if (in.mass_kind == kMassUniform) w = (VT)in.mg[k];
        else if (in.mass_kind == kMassDiagonal) w = (VT)in.gravity[k] * (VT)mass[nb0 + i];
        else w = ((VT)in.gravity[k] * (VT)mass[nb0 + i]) * (VT)kLumpTet;
        f[q + k] += w;
    }
}

// ProjectiveConstraintSet::projectResponse of the scene's constraints: FixedProjectiveConstraint
// (FixedProjectiveConstraint.inl:184-207; fixAll clears every node) and
// OscillatorProjectiveConstraint (OscillatorProjectiveConstraint.inl:65-70) clear their nodes.
// velocity = 1: FixedProjectiveConstraint::projectVelocity (.inl:237-258): the fixed nodes only,
// and only when activate_projectVelocity is set
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_project(const InstDev* __restrict__ inst, int sid,
                                                     const int8_t* __restrict__ fixed, int nb0, int n,
                                                     VT* __restrict__ vec, int velocity) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int8_t m = fixed[nb0 + i];
    const bool clear = velocity ? (inst[sid].project_velocity && (m & 1)) : m != 0;
    if (!clear) return;
    const int64_t q = 3 * (int64_t)i;
    vec[q] = 0; vec[q + 1] = 0; vec[q + 2] = 0;
}

// OscillatorProjectiveConstraint::projectPosition / projectVelocity (.inl:72-106) at time t on a
// caller vector: x = mean + a sin(w t + phi) (position = 1) or v = a w cos(w t + phi)
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_oscillate_vec(Oscillators os, int o0, int o1, int nb0, double t,
                                                           VT* __restrict__ vec, int position) {
    const int i = o0 + blockIdx.x * kBlock + threadIdx.x;
    if (i >= o1) return;
    const double w = os.w[i], ph = os.phase[i];
    const int64_t q = 3 * (int64_t)(os.node[i] - nb0);
    if (position) {
        const double sn = sin(w * t + ph);
#pragma unroll
        for (int k = 0; k < 3; k++) vec[q + k] = (VT)(os.mean[3 * i + k] + os.amp[3 * i + k] * sn);
    } else {
        const double cs = cos(w * t + ph);
#pragma unroll
        for (int k = 0; k < 3; k++) vec[q + k] = (VT)((os.amp[3 * i + k] * w) * cs);
    }
}

// BaseMechanicalState::vDot on caller vectors (MechanicalObject::vDot, MechanicalObject.inl
// :2334-2356), in the reference's order: one wave computes each node's three-term dot product
// and adds them node after node from 0.0 (the k_seqdot chain, terms formed on the fly; padding
// with +0.0 leaves the sum unchanged)
template <typename VT>
__global__ __launch_bounds__(64) void k_vdot_seq(const VT* __restrict__ a, const VT* __restrict__ b, int n,
                                                  double* __restrict__ out) {
    __shared__ double buf[kSeqChunk];
    const int lane = threadIdx.x;
    constexpr int PER = kSeqChunk / 64;
    double s = 0.0;
    for (int base = 0; base < n; base += kSeqChunk) {
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = base + k * 64 + lane;
            double t = 0.0;
            if (i < n) {
                VT x[3], y[3];
                load3(a, i, x[0], x[1], x[2]);
                load3(b, i, y[0], y[1], y[2]);
                t = dot3d(x, y);
            }
            buf[k * 64 + lane] = t;
        }
        __syncthreads();
#pragma unroll 64
        for (int k = 0; k < kSeqChunk; k++) s = s + buf[k];
        __syncthreads();
    }
    if (lane == 0) *out = s;
}

// the same dot product in the FP64 engine's parallel blocked order: per 256-node block the
// fixed tree of block_sum, then the block partials added by one block as last_block_reduce
// does (each thread a strided sum, then block_sum) -- the oracle's dot_blocked
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_vdot_blocks(const VT* __restrict__ a, const VT* __restrict__ b, int n,
                                                         double* __restrict__ partial) {
    __shared__ double sh[kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double t = 0.0;
    if (i < n) {
        VT x[3], y[3];
        load3(a, i, x[0], x[1], x[2]);
        load3(b, i, y[0], y[1], y[2]);
        t = dot3d(x, y);
    }
    const double bs = block_sum(t, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = bs;
}
static __global__ __launch_bounds__(kBlock) void k_vdot_final(const double* __restrict__ partial, int nb,
                                                               double* __restrict__ out) {
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (int k = threadIdx.x; k < nb; k += kBlock) acc = acc + partial[k];
    const double total = block_sum(acc, sh);
    if (threadIdx.x == 0) *out = total;
}

// ---------------------------------------------------------------------------------------
// von Mises stress (TetrahedronFEMForceField::computeVonMisesStress, .inl:2197-2347)
// ---------------------------------------------------------------------------------------
// Per element, from the scene's current positions x, in double whatever the engine's precision
// (a post-process read once per step at most): the displacement gradient
// gradU(k,l) = sum_m shf(l+1,m) u_m[k] with shf the inverse of the rest matrix [1 x y z] per
// vertex and u_m either the corotational displacement of vertex m in the element frame
// (mode 1; rest minus current, as the reference's D) or x - x0 (mode 2); the strain
// (gradU + gradU^T)/2 (mode 1) or (gradU + gradU^T + gradU^T gradU)/2 (mode 2); the stress
// s = 2 mu strain + lambda tr(strain) I in Voigt order (xx, yy, zz, yz, xz, xy); von Mises
// sqrt(s0^2 + s1^2 + s2^2 - s0 s1 - s1 s2 - s2 s0 + 3 (s3^2 + s4^2 + s5^2)), values below 1e-10
// set to 0.  Mode 1 also stores the element rotation of x, as the reference does.  Every
// operation in the reference's order, no contraction where the translation unit honours the
// pragma (the FP64 engines': they match the reference bit for bit).
// shf: 12 values per element (rows 1..3 of the inverse, row-major), lm: lambda, mu per element,
// both in the engine's element order of the scene; vme in that order.  x and x0 are the scene's
// (node ids of the element record minus nb index them).
template <typename R, int NR, typename VT>
__global__ __launch_bounds__(kBlock) void k_von_mises_elem(ElemView<R, NR> ev, int64_t e0, int64_t e1, int nb,
                                                           const VT* __restrict__ x,
                                                           const double* __restrict__ x0,
                                                           const double* __restrict__ shf,
                                                           const double* __restrict__ lm, int mode,
                                                           double* __restrict__ vme) {
#pragma clang fp contract(off)
    const int64_t e = e0 + (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= e1) return;
    const int64_t le = e - e0;
    const int4 t = ev.tet[e];
    const int nid[4] = {t.x - nb, t.y - nb, t.z - nb, t.w - nb};
    double p[4][3];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        VT a, b, c;
        load3(x, nid[m], a, b, c);
        p[m][0] = (double)a; p[m][1] = (double)b; p[m][2] = (double)c;
    }
    double D[12];
    if (mode == 2) {
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int k = 0; k < 3; k++) D[3 * m + k] = p[m][k] - x0[3 * (int64_t)nid[m] + k];
    } else {
        R gr[12], rcr[9];
        jgroups<R, NR>(ev, e, gr, rcr);
        double rc[9];
#pragma unroll
        for (int k = 0; k < 9; k++) rc[k] = (double)rcr[k];
        double rot[9];
        rotation_large<double>(p[0], p[1], p[2], rot);
#pragma unroll
        for (int k = 0; k < 9; k++) ev.rot[k * ev.E + e] = (R)rot[k];
        // deforme[k] = R_0_2 p_k, R_0_2 = transpose(rot); then relative to deforme[0]
        double df[4][3];
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int i = 0; i < 3; i++) {
                double v = rot[i] * p[k][0];
                v = v + rot[3 + i] * p[k][1];
                df[k][i] = v + rot[6 + i] * p[k][2];
            }
        df[1][0] -= df[0][0];
        df[2][0] -= df[0][0];
        df[2][1] -= df[0][1];
#pragma unroll
        for (int i = 0; i < 3; i++) df[3][i] -= df[0][i];
        D[0] = 0; D[1] = 0; D[2] = 0;
        D[3] = rc[0] - df[1][0]; D[4] = 0; D[5] = 0;
        D[6] = rc[3] - df[2][0]; D[7] = rc[4] - df[2][1]; D[8] = 0;
        D[9] = rc[6] - df[3][0]; D[10] = rc[7] - df[3][1]; D[11] = rc[8] - df[3][2];
    }
    const double* sf = shf + 12 * le;
    double gu[3][3];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int l = 0; l < 3; l++) {
            double a = 0.0;
#pragma unroll
            for (int m = 0; m < 4; m++) a += sf[4 * l + m] * D[3 * m + k];
            gu[k][l] = a;
        }
    double st[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double a = gu[i][j] + gu[j][i];
            if (mode == 2) {
                // (gradU^T gradU)(i,j), the 3x3 product's three-term order
                double g = gu[0][i] * gu[0][j] + gu[1][i] * gu[1][j];
                g = g + gu[2][i] * gu[2][j];
                a = a + g;
            }
            st[i][j] = a * 0.5;
        }
    const double vs[6] = {st[0][0], st[1][1], st[2][2], st[1][2], st[0][2], st[0][1]};
    const double lambda = lm[2 * le], mu = lm[2 * le + 1];
    double s[6], tr = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        tr += vs[k];
        s[k] = vs[k] * 2 * mu;
    }
#pragma unroll
    for (int k = 3; k < 6; k++) s[k] = vs[k] * 2 * mu;
#pragma unroll
    for (int k = 0; k < 3; k++) s[k] += lambda * tr;
    double q = s[0] * s[0] + s[1] * s[1];
    q = q + s[2] * s[2];
    q = q - s[0] * s[1];
    q = q - s[1] * s[2];
    q = q - s[2] * s[0];
    q = q + 3 * s[3] * s[3];
    q = q + 3 * s[4] * s[4];
    q = q + 3 * s[5] * s[5];
    double v = sqrt(q);
    if (v < 1e-10) v = 0.0;
    vme[le] = v;
}

// vonMisesPerNode: the mean of the element values around each node, summed over its incidence
// list (ascending element id, getTetrahedraAroundVertex's order); 0 for a node of no element
static __global__ __launch_bounds__(kBlock) void k_von_mises_node(const int* __restrict__ inc_ptr, const int* __restrict__ inc,
                                                           int n0, int n1, int64_t e0,
                                                           const double* __restrict__ vme,
                                                           double* __restrict__ vmn) {
#pragma clang fp contract(off)
    const int n = n0 + blockIdx.x * kBlock + threadIdx.x;
    if (n >= n1) return;
    const int s0 = inc_ptr[n], s1 = inc_ptr[n + 1];
    double a = 0.0;
    for (int s = s0; s < s1; s++) a += vme[(int64_t)(inc[s] >> 2) - e0];
    if (s1 > s0) a /= (double)(s1 - s0);
    vmn[n - n0] = a;
}

// ---------------------------------------------------------------------------------------
// assembled stiffness (TetrahedronFEMForceField::addKToMatrix / buildStiffnessMatrix,
// .inl:1921-1999, with computeStiffnessMatrix, :212-232)
// ---------------------------------------------------------------------------------------
// block (n1, n2) of R J K J^T R^T of element e (computeStiffnessMatrix, TetrahedronFEMForceField
// .inl:212-232, with the last addForce's rotation): J's rows of n1 and n2 only, with R =
// rotations[e], JK = J(n1 rows) K, B = JK J(n2 rows)^T, S = R B, block = S R^T, every sum in the
// reference's dense-product order (zero terms of its 12x12 products leave its sums unchanged)
template <typename R, int NR>
__device__ inline void elem_stiffness_block(const ElemView<R, NR>& ev, int64_t e, int n1, int n2, double blk[9]) {
#pragma clang fp contract(off)
    R gr[12], rcr[9];
    jgroups<R, NR>(ev, e, gr, rcr);
    double rot[9];
#pragma unroll
    for (int k = 0; k < 9; k++) rot[k] = (double)ev.rot[k * ev.E + e];
    const double k00 = (double)ev.kmat[e], k01 = (double)ev.kmat[ev.E + e], k33 = (double)ev.kmat[2 * ev.E + e];
    double Km[6][6];
#pragma unroll
    for (int l = 0; l < 6; l++)
#pragma unroll
        for (int m = 0; m < 6; m++)
            Km[l][m] = (l < 3 && m < 3) ? (l == m ? k00 : k01) : ((l == m) ? k33 : 0.0);
    // J rows of a vertex (computeStrainDisplacement's placement of its three groups)
    double J1[3][6], J2[3][6];
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
        for (int l = 0; l < 6; l++) { J1[q][l] = 0.0; J2[q][l] = 0.0; }
    {
        const double a0 = (double)gr[3 * n1], a1 = (double)gr[3 * n1 + 1], a2 = (double)gr[3 * n1 + 2];
        J1[0][0] = a0; J1[0][3] = a1; J1[0][5] = a2;
        J1[1][3] = a0; J1[1][1] = a1; J1[1][4] = a2;
        J1[2][5] = a0; J1[2][4] = a1; J1[2][2] = a2;
        const double b0 = (double)gr[3 * n2], b1 = (double)gr[3 * n2 + 1], b2 = (double)gr[3 * n2 + 2];
        J2[0][0] = b0; J2[0][3] = b1; J2[0][5] = b2;
        J2[1][3] = b0; J2[1][1] = b1; J2[1][4] = b2;
        J2[2][5] = b0; J2[2][4] = b1; J2[2][2] = b2;
    }
    double JK[3][6];
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
        for (int m = 0; m < 6; m++) {
            double a = 0.0;
#pragma unroll
            for (int l = 0; l < 6; l++) a += J1[q][l] * Km[l][m];
            JK[q][m] = a;
        }
    double B[3][3];
#pragma unroll
    for (int q1 = 0; q1 < 3; q1++)
#pragma unroll
        for (int q2 = 0; q2 < 3; q2++) {
            double a = 0.0;
#pragma unroll
            for (int m = 0; m < 6; m++) a += JK[q1][m] * J2[q2][m];
            B[q1][q2] = a;
        }
    double S[3][3];
#pragma unroll
    for (int q1 = 0; q1 < 3; q1++)
#pragma unroll
        for (int p = 0; p < 3; p++) {
            double a = 0.0;
#pragma unroll
            for (int r = 0; r < 3; r++) a += rot[3 * q1 + r] * B[r][p];
            S[q1][p] = a;
        }
#pragma unroll
    for (int q1 = 0; q1 < 3; q1++)
#pragma unroll
        for (int q2 = 0; q2 < 3; q2++) {
            double a = 0.0;
#pragma unroll
            for (int p = 0; p < 3; p++) a += S[q1][p] * rot[3 * q2 + p];
            blk[3 * q1 + q2] = a;
        }
}

// The assembled matrix in two passes.  (1) One thread per element and vertex pair (n1, n2): the
// element block -(R J K J^T R^T)(n1, n2) -- J's rows of n1 and n2 only: with R = rotations[e]
// (the last addForce's), JK = J(n1 rows) K, B = JK J(n2 rows)^T, S = R B, block = S R^T, every
// sum in the reference's dense-product order (zero terms of the reference's 12x12 products leave
// its sums unchanged) -- into the scene's element-block buffer at (element - e0) * 16 + n1 * 4 +
// n2.  (2) One thread per 3x3 block (a, b): the sum, in ascending element id (the element loop's
// order), of block * k over the contributions slots[] lists for it (device element * 16 + n1 * 4
// + n2).  Each element record is read once (not once per block it touches), and both passes
// store through LDS so that a wave's stores are contiguous.  Computed in double, no contraction
// where the translation unit honours the pragma.
template <typename R, int NR>
__global__ __launch_bounds__(kBlock) void k_stiffness_elem(ElemView<R, NR> ev, int64_t e0, int64_t n16,
                                                            double* __restrict__ eblk) {
#pragma clang fp contract(off)
    __shared__ double st[9 * kBlock];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n16) {
        double blk[9];
        const int j = (int)(i & 15);
        elem_stiffness_block<R, NR>(ev, e0 + (i >> 4), j >> 2, j & 3, blk);
#pragma unroll
        for (int k = 0; k < 9; k++) st[9 * threadIdx.x + k] = blk[k];
    }
    __syncthreads();
    const int64_t o = 9 * (int64_t)blockIdx.x * kBlock, end = 9 * n16;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int64_t w = o + k * kBlock + threadIdx.x;
        if (w < end) eblk[w] = st[k * kBlock + threadIdx.x];
    }
}

template <typename R, int NR>
__global__ __launch_bounds__(kBlock) void k_stiffness_blocks(const int* __restrict__ slot_ptr, const int* __restrict__ slots,
                                                             int nblk, int64_t slot0, double kf,
                                                             const double* __restrict__ eblk, double* __restrict__ out) {
#pragma clang fp contract(off)
    __shared__ double st[9 * kBlock];
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b < nblk) {
        double acc[9];
#pragma unroll
        for (int i = 0; i < 9; i++) acc[i] = 0.0;
        for (int s = slot_ptr[b]; s < slot_ptr[b + 1]; s++) {
            const double* blk = eblk + 9 * ((int64_t)slots[s] - slot0);
#pragma unroll
            for (int i = 0; i < 9; i++) acc[i] += (-blk[i]) * kf;
        }
#pragma unroll
        for (int i = 0; i < 9; i++) st[9 * threadIdx.x + i] = acc[i];
    }
    __syncthreads();
    const int64_t o = 9 * (int64_t)blockIdx.x * kBlock, end = 9 * (int64_t)nblk;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int64_t w = o + k * kBlock + threadIdx.x;
        if (w < end) out[w] = st[k * kBlock + threadIdx.x];
    }
}

// ---------------------------------------------------------------------------------------
// PCGLinearSolver with a Jacobi / block-Jacobi preconditioner (PCGLinearSolver.inl:178-262,
// JacobiPreconditioner.inl:43-52, BlockJacobiPreconditioner.inl:41-52)
// ---------------------------------------------------------------------------------------
// node n's preconditioner block, inverted: mF (vertex mass) I, then kF (-R J K J^T R^T)(v, v) of
// its elements in ascending element id (MatrixLinearSystem's assembly: mass, then the FEM's
// element loop, each value scaled by its factor); the identity for a fixed node (Dirichlet row,
// MatrixLinearSystem.inl:1177-1189); then DiagonalMatrix::invert (1 / d, off-diagonal dropped) or
// Mat<3,3>::invert (invertMatrix, Mat.h:1172-1192, with the determinant of Mat.h:989-997)
template <typename R, int NR>
__device__ inline void pcg_block(const InstDev& in, const ElemView<R, NR>& ev, const int* __restrict__ inc_ptr,
                                 const int* __restrict__ inc, const double* __restrict__ mass,
                                 const int8_t* __restrict__ fixed, int n, double* __restrict__ out) {
#pragma clang fp contract(off)
    double B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (fixed[n] & 1) {
        B[0] = B[4] = B[8] = 1.0;
    } else {
        double m = in.mass_kind == kMassUniform ? in.vertex_mass : mass[n];
        if (in.mass_kind == kMassMatrixLumped) m = m * kLumpTet;
        const double mm = m * in.m_cg;
        B[0] += mm; B[4] += mm; B[8] += mm;
        for (int s = inc_ptr[n]; s < inc_ptr[n + 1]; s++) {
            const int64_t e = inc[s] >> 2;
            const int v = inc[s] & 3;
            double blk[9];
            elem_stiffness_block<R, NR>(ev, e, v, v, blk);
#pragma unroll
            for (int i = 0; i < 9; i++) B[i] += (-blk[i]) * in.k_cg;
        }
    }
    if (in.pcg == 1) {
#pragma unroll
        for (int i = 0; i < 9; i++) out[i] = (i % 4 == 0) ? 1.0 / B[i] : 0.0;
        return;
    }
    auto F = [&](int r, int c) { return B[3 * r + c]; };
    const double det = F(0, 0) * F(1, 1) * F(2, 2) + F(1, 0) * F(2, 1) * F(0, 2) + F(2, 0) * F(0, 1) * F(1, 2)
                       - F(0, 0) * F(2, 1) * F(1, 2) - F(1, 0) * F(0, 1) * F(2, 2) - F(2, 0) * F(1, 1) * F(0, 2);
    out[0] = (F(1, 1) * F(2, 2) - F(2, 1) * F(1, 2)) / det;
    out[3] = (F(1, 2) * F(2, 0) - F(2, 2) * F(1, 0)) / det;
    out[6] = (F(1, 0) * F(2, 1) - F(2, 0) * F(1, 1)) / det;
    out[1] = (F(2, 1) * F(0, 2) - F(0, 1) * F(2, 2)) / det;
    out[4] = (F(2, 2) * F(0, 0) - F(0, 2) * F(2, 0)) / det;
    out[7] = (F(2, 0) * F(0, 1) - F(0, 0) * F(2, 1)) / det;
    out[2] = (F(0, 1) * F(1, 2) - F(1, 1) * F(0, 2)) / det;
    out[5] = (F(0, 2) * F(1, 0) - F(1, 2) * F(0, 0)) / det;
    out[8] = (F(0, 0) * F(1, 1) - F(1, 0) * F(0, 1)) / det;
}

// every node of the PCG scenes of the launch set: its inverted preconditioner block (the
// rotations of this step's addForce)
template <typename R, int NR, typename BLK>
__global__ __launch_bounds__(kBlock) void k_pcg_diag(ElemView<R, NR> ev, const BLK* __restrict__ nblk,
                                                      const InstDev* __restrict__ inst, const int* __restrict__ inc_ptr,
                                                      const int* __restrict__ inc, const double* __restrict__ mass,
                                                      const int8_t* __restrict__ fixed, double* __restrict__ pinv) {
    const BLK bk = nblk[blockIdx.x];
    const InstDev& in = inst[blk_scene(bk)];
    if (!in.pcg) return;
    for (int n = bk.begin + threadIdx.x; n < bk.end; n += kBlock)
        pcg_block<R, NR>(in, ev, inc_ptr, inc, mass, fixed, n, pinv + 9 * (int64_t)n);
}

// LinearMovementProjectiveConstraint's indices are Dirichlet rows of the assembled matrix
// (applyConstraint(ZeroDirichletCondition*)) whatever the time: the identity block
static __global__ __launch_bounds__(kBlock) void k_pcg_lm_identity(const Blk* __restrict__ lblk,
                                                                    const InstDev* __restrict__ inst,
                                                                    double* __restrict__ pinv) {
    const Blk bk = lblk[blockIdx.x];
    const int i = bk.begin + threadIdx.x;
    const InstDev& in = inst[bk.inst];
    if (i >= bk.end || !in.pcg) return;
    double* P = pinv + 9 * (int64_t)in.lm_node[i];
#pragma unroll
    for (int k = 0; k < 9; k++) P[k] = (k % 4 == 0) ? 1.0 : 0.0;
}

// the PCG scenes of the launch set start a solve: iteration 1, from the solution vector
static __global__ void k_pcg_open(const int* __restrict__ scenes, int n, InstDev* __restrict__ inst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    InstDev* in = &inst[scenes[i]];
    if (!in->pcg) return;
    in->stopped = 0;
    in->iter = 1;
    in->end_reason = 0;
    in->pcg_tol = in->tol * in->bnorm2;      // tol = tolerance b.b
}

// r = M x (the operator on the solution vector, element contributions from k_cg_elem at
// iteration 1 reading x), r = -r + b (cgstep_beta(r, b, -1): r *= -1; r += b), w = P^-1 r,
// r_norm = r.w, the loop's first test (PCGLinearSolver.inl:193-215); w goes to z, from which
// iteration 1's direction is taken
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_pcg_begin(const Blk* __restrict__ nblk, InstDev* __restrict__ inst,
                                                       NodeArrays na, const VT* __restrict__ ef,
                                                       const VT* __restrict__ xcg, const VT* __restrict__ b,
                                                       VT* __restrict__ r, VT* __restrict__ z,
                                                       const double* __restrict__ pinv, double* __restrict__ partial,
                                                       double* __restrict__ nd) {
    __shared__ double sh[kBlock];
    const Blk bk = nblk[blockIdx.x];
    InstDev* ins = &inst[bk.inst];
    if (!ins->pcg || ins->stopped) return;
    const int n = bk.begin + threadIdx.x;
    double rz = 0.0;
    if (n < bk.end) {
        const InstDev& in = *ins;
        const int64_t o = 3 * (int64_t)n;
        VT xn[3];
        load3(xcg, n, xn[0], xn[1], xn[2]);
        VT qn[3] = {0, 0, 0};
        if (in.m_cg != 0.0) {
            mass_mdx(in, na, n, in.m_cg, xn, qn);
            if (in.mass_kind == kMassMatrix)
                mass_edges(na, n, in.m_cg, [&](int j, VT d[3]) { load3(xcg, j, d[0], d[1], d[2]); }, qn);
        }
        const int s0 = na.inc_ptr[n], s1 = na.inc_ptr[n + 1];
        for (int s = s0; s < s1; s++) {
            const int64_t slot = na.inc[s];
            qn[0] -= ef[3 * slot]; qn[1] -= ef[3 * slot + 1]; qn[2] -= ef[3 * slot + 2];
        }
        if (in.plane_on && na.contact[n]) plane_dforce(in, in.k_cg, xn, qn);
        if (na.fixed[n]) { qn[0] = 0; qn[1] = 0; qn[2] = 0; }
        VT rn[3], zn[3];
#pragma unroll
        for (int k = 0; k < 3; k++) { VT t = qn[k]; t *= (VT)-1.0; t += b[o + k]; rn[k] = t; r[o + k] = t; }
        precond_apply(in.pcg, pinv + 9 * (int64_t)n, rn, zn);
        z[o] = zn[0]; z[o + 1] = zn[1]; z[o + 2] = zn[2];
        rz = dot3d(rn, zn);
        if (nd) nd[n] = rz;
    }
    double rho;
    const bool last = nd ? last_block_sequential(nd, ins, rho)
                         : last_block_reduce(block_sum(rz, sh), partial, ins, bk.pad, sh, rho);
    if (last && threadIdx.x == 0) {
        ins->rho = rho;
        ins->rho_1 = 0.0;
        ins->err = ins->bnorm2 != 0.0 ? rho / ins->bnorm2 : 0.0;
        if (ins->max_iter < 1 || !(rho > ins->pcg_tol)) {
            ins->stopped = 1; ins->end_reason = 1; ins->last_iters = 1;
        }
    }
}

// the warm-start scenes of the launch set open their solve: iteration 1, so that k_cg_elem
// evaluates the element part of A x from the solution vector
static __global__ void k_warm_open(const int* __restrict__ scenes, int n, InstDev* __restrict__ inst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    InstDev* in = &inst[scenes[i]];
    if (!in->warm_start || in->pcg) return;
    in->stopped = 0;
    in->iter = 1;
    in->end_reason = 0;
}

// CGLinearSolver warmStart: r = A x (mass, the element contributions of k_cg_elem on x, plane,
// projection) then r.eq(b, r, -1.0): r = b + r (-1); rho = r.r and the first tests
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_cg_warm_begin(const Blk* __restrict__ nblk, InstDev* __restrict__ inst,
                                                           NodeArrays na, const VT* __restrict__ ef,
                                                           const VT* __restrict__ xcg, const VT* __restrict__ b,
                                                           VT* __restrict__ r, double* __restrict__ partial,
                                                           double* __restrict__ nd) {
    __shared__ double sh[kBlock];
    const Blk bk = nblk[blockIdx.x];
    InstDev* ins = &inst[bk.inst];
    if (!ins->warm_start || ins->pcg || ins->stopped) return;
    const int n = bk.begin + threadIdx.x;
    double rr = 0.0;
    if (n < bk.end) {
        const InstDev& in = *ins;
        const int64_t o = 3 * (int64_t)n;
        VT xn[3];
        load3(xcg, n, xn[0], xn[1], xn[2]);
        VT qn[3] = {0, 0, 0};
        if (in.m_cg != 0.0) {
            mass_mdx(in, na, n, in.m_cg, xn, qn);
            if (in.mass_kind == kMassMatrix)
                mass_edges(na, n, in.m_cg, [&](int j, VT d[3]) { load3(xcg, j, d[0], d[1], d[2]); }, qn);
        }
        const int s0 = na.inc_ptr[n], s1 = na.inc_ptr[n + 1];
        for (int s = s0; s < s1; s++) {
            const int64_t slot = na.inc[s];
            qn[0] -= ef[3 * slot]; qn[1] -= ef[3 * slot + 1]; qn[2] -= ef[3 * slot + 2];
        }
        if (in.plane_on && na.contact[n]) plane_dforce(in, in.k_cg, xn, qn);
        if (na.fixed[n]) { qn[0] = 0; qn[1] = 0; qn[2] = 0; }
        VT rn[3];
#pragma unroll
        for (int k = 0; k < 3; k++) { VT t = qn[k] * (VT)-1.0; rn[k] = b[o + k] + t; r[o + k] = rn[k]; }
        rr = dot3d(rn, rn);
        if (nd) nd[n] = rr;
    }
    double total;
    const bool last = nd ? last_block_sequential(nd, ins, total)
                         : last_block_reduce(block_sum(rr, sh), partial, ins, bk.pad, sh, total);
    if (last && threadIdx.x == 0) cg_start_warm(ins, ins->bnorm2, total);
}

}  // namespace sofa_amd
