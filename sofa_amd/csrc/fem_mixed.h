// fem_mixed.h -- iterative refinement for the MIXED precision mode.
//
// The implicit Euler system A dv = b (A = (1 + h rm) M - h (h + rs) K, b from the FP64 RHS
// kernels) is solved by refinement sweeps: with the FP64 correction sum x (x = 0 at the start),
//   r = b - A x            FP64 operator on the FP64 element record (k_tile_apply, iter 0)
//   stop if |r| / |b| <= tolerance, or after max_sweeps inner solves
//   A d = (float) r        FP32 CG (the FP32 tiled kernels, a scene record of its own)
//   x += d                 FP64
// Every sweep is launched for every scene of the launch set; a scene that has converged skips
// the work (the kernels test its `stopped` flag), so the step stays one fixed launch sequence
// that can be captured in a graph.  Reductions are per-block partials re-reduced in a fixed
// order, as on the tiled path; nothing here uses atomics.
#pragma once

#include "fem_kernels.h"
#include "fem_tiled.h"

namespace sofa_amd {

// r = b (sweep 1) or b - A x (q = A x from the FP64 operator); r32 = (float) r; d = 0.
// Partials of |r|^2 (FP64 values) and of |r32|^2 per node block -- with a preconditioner (zbuf
// given) r32.z instead, z = P^-1 r32 stored in zbuf (precond_z).
static __global__ __launch_bounds__(kBlock) void k_ir_residual(const NodeBlk* __restrict__ nb_,
                                                                const InstDev* __restrict__ inst, int sweep,
                                                                const double* __restrict__ b,
                                                                const double* __restrict__ q,
                                                                float* __restrict__ r32, float* __restrict__ d32,
                                                                double* __restrict__ prr, double* __restrict__ prr32,
                                                                const float* __restrict__ dinv,
                                                                const float* __restrict__ binv,
                                                                float* __restrict__ zbuf) {
    __shared__ double sh[kBlock / 64];
    const NodeBlk nb = nb_[blockIdx.x];
    if (inst[nb.scene].stopped) return;
    double rr = 0.0, rr32 = 0.0;
#pragma unroll
    for (int k = 0; k < kNodesPerThread; k++) {
        const int n = nb.begin + k * kBlock + threadIdx.x;
        if (n < nb.end) {
            double rn[3];
            ld3(b, n, rn);
            if (sweep > 1) {
                double qn[3];
                ld3(q, n, qn);
#pragma unroll
                for (int i = 0; i < 3; i++) rn[i] = rn[i] - qn[i];
            }
            const float rf[3] = {(float)rn[0], (float)rn[1], (float)rn[2]};
            const float z[3] = {0.f, 0.f, 0.f};
            st3(r32, n, rf);
            st3(d32, n, z);
            rr += dot3d(rn, rn);
            if (zbuf) {
                float zf[3];
                precond_z(inst[nb.scene].pcg, dinv, binv, n, rf, zf);
                st3(zbuf, n, zf);
                rr32 += dot3d(rf, zf);
            } else {
                rr32 += dot3d(rf, rf);
            }
        }
    }
    const double t0 = wave_block_sum<kBlock>(rr, sh);
    const double t1 = wave_block_sum<kBlock>(rr32, sh);
    if (threadIdx.x == 0) { prr[nb.pslot] = t0; prr32[nb.pslot] = t1; }
}

// One block per scene of the launch set: the refinement test of sweep `sweep` (1-based; sweep
// max_sweeps + 1 only measures the final residual).  Opens or closes the inner solve.
static __global__ __launch_bounds__(kBlock) void k_ir_control(const int* __restrict__ scenes, InstDev* __restrict__ inst,
                                                               InstDev* __restrict__ inner,
                                                               const Ranges* __restrict__ ranges, int sweep,
                                                               const double* __restrict__ prr) {
    __shared__ double sh[kBlock / 64];
    const int si = scenes[blockIdx.x];
    InstDev* o = &inst[si];
    InstDev* in = &inner[si];
    if (o->stopped) {
        if (threadIdx.x == 0) in->stopped = 1;
        return;
    }
    const Ranges rg = ranges[si];
    const double rr = reduce2(prr, rg.b0, rg.nb, 0, 0, sh);
    if (threadIdx.x != 0) return;
    if (sweep == 1) o->normb = sqrt(rr);
    const double normb = o->normb;
    o->ir_sweeps = sweep - 1;
    if (normb == 0.0) {
        o->stopped = 1; o->end_reason = 4; o->err = 0.0; in->stopped = 1;
        return;
    }
    const double err = sqrt(rr) / normb;
    o->err = err;
    if (err <= o->tol) {
        o->stopped = 1; o->end_reason = 1; in->stopped = 1;
    } else if (sweep > o->max_sweeps) {
        o->stopped = 1; o->end_reason = 0; in->stopped = 1;
    } else {
        // the next inner solve: its kernels start at iteration 1.  Keeping the direction, the
        // r.r that built it is the one of the last iteration whose direction update ran (a
        // tolerance exit in the operator kernel at iteration j stops before building p_j)
        if (in->keep_dir && sweep > 1) {
            int lj = in->last_iters;
            if (!in->stopped) lj = in->max_iter;
            else if (in->end_reason == 1) lj = lj - 1;
            in->rho_last = in->rho_slot[lj & 1];
        }
        in->sweep = sweep;
        in->stopped = 0; in->end_reason = 0; in->last_iters = 0; in->first_solve = 0;
    }
}

// x += d after an inner solve (scenes whose refinement had already stopped skip it); the
// first block of a scene adds the inner iteration count to the step's total.
static __global__ __launch_bounds__(kBlock) void k_ir_accumulate(const NodeBlk* __restrict__ nb_, InstDev* __restrict__ inst,
                                                                  InstDev* __restrict__ inner,
                                                                  const float* __restrict__ d32, double* __restrict__ x) {
    const NodeBlk nb = nb_[blockIdx.x];
    InstDev* o = &inst[nb.scene];
    if (o->stopped) return;
    if (nb.first && threadIdx.x == 0) {
        InstDev* in = &inner[nb.scene];
        // an inner loop that ran every iteration did not record its exit
        const int its = in->stopped ? in->last_iters : in->max_iter;
        o->cg_total += its;
    }
#pragma unroll
    for (int k = 0; k < kNodesPerThread; k++) {
        const int n = nb.begin + k * kBlock + threadIdx.x;
        if (n < nb.end) {
            const int64_t j = 3 * (int64_t)n;
#pragma unroll
            for (int i = 0; i < 3; i++) x[j + i] = x[j + i] + (double)d32[j + i];
        }
    }
}

// The inner FP32 CG's Jacobi preconditioner, once per step from this step's rotations: per node
// the inverse of its diagonal -- the mass term and the element blocks' diagonals (pcg_block, the
// assembly of PCGLinearSolver's JacobiPreconditioner) -- in float; 1 for the nodes of scenes
// without a preconditioner (z = r: their inner CG is the plain one) and for fixed nodes
template <typename R, int NR>
__global__ __launch_bounds__(kBlock) void k_jacobi32(ElemView<R, NR> ev, const NodeBlk* __restrict__ nblk,
                                                      const InstDev* __restrict__ inst, const int* __restrict__ inc_ptr,
                                                      const int* __restrict__ inc, const double* __restrict__ mass,
                                                      const int8_t* __restrict__ fixed, float* __restrict__ dinv,
                                                      float* __restrict__ binv) {
    const NodeBlk bk = nblk[blockIdx.x];
    const InstDev& in = inst[bk.scene];
    if (!in.pcg) return;   // z = r: neither array is read
    for (int n = bk.begin + threadIdx.x; n < bk.end; n += kBlock) {
        double B[9];
        pcg_block<R, NR>(in, ev, inc_ptr, inc, mass, fixed, n, B);
        if (in.pcg == 1) {
            const float d[3] = {(float)B[0], (float)B[4], (float)B[8]};
            st3(dinv, n, d);
        } else {
#pragma unroll
            for (int i = 0; i < 9; i++) binv[9 * (int64_t)n + i] = (float)B[i];
        }
    }
}

// LinearMovementProjectiveConstraint's nodes are Dirichlet rows: identity
static __global__ __launch_bounds__(kBlock) void k_jacobi32_lm(const Blk* __restrict__ lblk,
                                                                const InstDev* __restrict__ inst,
                                                                float* __restrict__ dinv, float* __restrict__ binv) {
    const Blk bk = lblk[blockIdx.x];
    const int i = bk.begin + threadIdx.x;
    const InstDev& in = inst[bk.inst];
    if (i >= bk.end || !in.pcg) return;
    const int64_t n = in.lm_node[i];
    if (in.pcg == 1) {
        const float one[3] = {1.f, 1.f, 1.f};
        st3(dinv, (int)n, one);
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) binv[9 * n + k] = (k % 4 == 0) ? 1.f : 0.f;
    }
}

// scene bookkeeping at the end of a MIXED step (one thread per scene of the launch set)
static __global__ void k_step_end_mixed(const int* __restrict__ scenes, int n, InstDev* __restrict__ inst,
                                        InstDev* __restrict__ inner) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    InstDev* o = &inst[scenes[i]];
    o->last_iters = o->cg_total;
    o->cg_total = 0;
    o->first_solve = 0;
    o->step_count += 1;
    o->time += o->dt;
    o->lm_x0_set = 1;
    o->stopped = 0;
    inner[scenes[i]].stopped = 1;
}

}  // namespace sofa_amd
