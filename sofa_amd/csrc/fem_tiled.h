// fem_tiled.h -- the tiled fast path: one workgroup per element tile, node gather in LDS.
//
// Elements are grouped into spatially compact tiles (Morton order of their centroids, cut so
// that a tile has at most kTileElems elements and kTileNodes distinct nodes).  One workgroup
// per tile:
//   1. stages the tile's nodes (p = r + beta p for the CG apply; x and v for the RHS) in LDS,
//   2. computes every element's 12 nodal contributions (applyStiffnessCorotational /
//      accumulateForceLarge arithmetic, shared with the exact path) into LDS,
//   3. sums them per node over the node's incidence list in ascending element id.
// A node touched by one tile only ("interior") is finished in the tile: q = (1 + h rm) m p,
// then the contributions subtracted one by one in the reference's element order, projection,
// store, and its p.q enters the tile's partial.  A node touched by several tiles ("shared")
// gets one partial sum per tile, added in ascending tile order by k_fin_apply.  Every step is
// deterministic; no atomics.
//
// Scalar reductions (|b|^2, p.q, r.r) are written as per-block partials and re-reduced, in a
// fixed order, by every block of the kernel that consumes them (the first block of a scene
// persists the scalars), so no block ever waits on or fences against another block.
#pragma once

#include "fem_device.h"

namespace sofa_amd {

#ifndef SOFA_AMD_TILE_ELEMS
#define SOFA_AMD_TILE_ELEMS 512
#endif
constexpr int kTileElems = SOFA_AMD_TILE_ELEMS;   // max elements per tile
constexpr int kTileNodes = SOFA_AMD_TILE_ELEMS;   // max distinct nodes per tile
// incidences are uint16 contribution slots (row 3 v + c of 12 rows of kTileElems)
static_assert(12 * kTileElems <= 65536, "contribution slots must fit in 16 bits");
#ifndef SOFA_AMD_NODES_PER_THREAD
#define SOFA_AMD_NODES_PER_THREAD 2
#endif
constexpr int kNodesPerThread = SOFA_AMD_NODES_PER_THREAD;   // node kernels: 256 threads x this
#ifndef SOFA_AMD_TILE_THREADS
#define SOFA_AMD_TILE_THREADS 512
#endif
constexpr int kTileThreads = SOFA_AMD_TILE_THREADS;   // threads of the MIXED inner operator kernel


struct TileDesc {
    int scene;
    int e_begin, e_count;     // permuted element range
    int ln_begin, ln_count;   // tile-local node range (into ln_* arrays)
    int flags;                // kTileFirst: first tile of its scene (persists the scene's scalars);
                              // kTileConsistent: the scene has a consistent MeshMatrixMass
    int pslot;                // slot of this tile's partial in the tile/finalize partial buffers
    int inc_begin;            // first entry of the tile's incidence lists (ln_inc)
};
constexpr int kTileFirst = 1, kTileConsistent = 2;
constexpr int kTileOneMaterial = 4;   // the scene has one E and nu: K follows from the rest volume

struct FinBlk { int scene, begin, end, pslot, bi, pad0, pad1, pad2; };  // shared-node range; bi = index in scene
struct NodeBlk { int scene, begin, end, pslot, first, pad0, pad1, pad2; };  // node range
__device__ __forceinline__ int blk_scene(const NodeBlk& b) { return b.scene; }

struct TileNodes {
    const int* __restrict__ gid;        // global node id of each tile-local node
    const int* __restrict__ inc_ptr;    // incidence range (absolute, into inc)
    const uint16_t* __restrict__ inc;   // local element slot * 4 + vertex, ascending element id
    const int* __restrict__ pslot;      // -1: interior; else shared-partial slot of (node, tile),
                                        // | kSlotOwner in the node's first tile (reference order:
                                        // the node's first per-incidence slot)
    const uint8_t* __restrict__ rank;   // reference order: per incidence of a shared node, its
                                        // slot's offset from the node's first slot (else null)
    const uint8_t* __restrict__ arank;  // reference order, CG operator: 0 for an incidence of the
                                        // node's leading run (folded into its first slot by this
                                        // tile), else its slot's offset from the node's first slot
    const uint8_t* __restrict__ lead;   // reference order, CG operator: per tile-local node, the
                                        // number of its first incidences this tile folds (else 0)
};
constexpr int kSlotOwner = 1 << 30;

struct SharedNodes {
    const int* __restrict__ gid;        // shared (and element-free) nodes
    const int* __restrict__ ptr;        // their partial slots [ptr[i], ptr[i+1]) in tile order
    const int* __restrict__ of;         // per node: its index in gid / ptr, or -1 (tile-owned)
    const uint8_t* __restrict__ acnt;   // reference order, CG operator: slots after the first
};

// per-scene slot ranges of the two partial buffers
struct Ranges {
    int t0, nt;    // tile partials
    int f0, nf;    // finalize-block partials (offset into the same buffer)
    int b0, nb;    // node-block partials (rho buffer)
};

// one DPP step of a 64-lane sum: v of the lane CTRL names in the rows ROWS enables, 0.0 where
// there is none (a double moved as its two 32-bit halves)
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_from(double v) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_update_dpp(0, hi, CTRL, ROWS, 0xf, true),
                            __builtin_amdgcn_update_dpp(0, lo, CTRL, ROWS, 0xf, true));
}

// sum over the 64 lanes of a wave in a fixed order: an inclusive scan within each row of 16
// lanes (row_shr 1, 2, 4, 8), then row 0's total into row 1 and row 2's into row 3
// (row_bcast 15), then the first half's total into the second half (row_bcast 31); lane 63 then
// holds the sum, and every lane reads it from there (the same bits in every lane)
__device__ __forceinline__ double wave_sum(double v) {
    v = v + dpp_from<0x111, 0xf>(v);
    v = v + dpp_from<0x112, 0xf>(v);
    v = v + dpp_from<0x114, 0xf>(v);
    v = v + dpp_from<0x118, 0xf>(v);
    v = v + dpp_from<0x142, 0xa>(v);
    v = v + dpp_from<0x143, 0xc>(v);
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                            __builtin_amdgcn_readlane(__double2loint(v), 63));
}

// every wave of the block: the sum of buf[a0 .. a0 + na) (lane-strided partial sums, then
// wave_sum) -- no LDS and no barrier, and the same bits in every wave and every block
__device__ __forceinline__ double wave_reduce(const double* __restrict__ buf, int a0, int na) {
    double acc = 0.0;
    for (int k = (int)(threadIdx.x & 63); k < na; k += 64) acc = acc + buf[a0 + k];
    return wave_sum(acc);
}

// fixed-order block sum of the tiled kernels (TB threads): wave_sum within each wave, then the
// wave sums added in wave order; sh holds TB / 64 doubles
template <int TB>
__device__ __forceinline__ double wave_block_sum(double v, double* sh) {
    v = wave_sum(v);
    __syncthreads();                     // a previous call's readers are done with sh
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = sh[0];
#pragma unroll
    for (int k = 1; k < TB / 64; k++) r = r + sh[k];
    return r;
}

// fixed-order sum of buf[a0 .. a0+na) then buf[c0 .. c0+nc), all threads of the block
template <int TB = kBlock>
__device__ __forceinline__ double reduce2(const double* __restrict__ buf, int a0, int na, int c0, int nc, double* sh) {
    double acc = 0.0;
    for (int k = threadIdx.x; k < na; k += TB) acc = acc + buf[a0 + k];
    for (int k = threadIdx.x; k < nc; k += TB) acc = acc + buf[c0 + k];
    return wave_block_sum<TB>(acc, sh);
}

// a scene's sum of na partials, the same bits in every block of the scene: few partials are
// added by every wave on its own (no barrier), many are split over the block (two barriers).
// WIDE = false: the launch has no scene of more than 256 partials (the block path compiled out)
constexpr int kWavePartials = 256;
template <int TB, bool WIDE = true>
__device__ __forceinline__ double scene_sum(const double* __restrict__ buf, int a0, int na, double* sh) {
    if (!WIDE || na <= kWavePartials) return wave_reduce(buf, a0, na);
    return reduce2<TB>(buf, a0, na, 0, 0, sh);
}

template <typename VT>
__device__ __forceinline__ void ld3(const VT* __restrict__ v, int n, VT o[3]) {
    const VT* q = v + 3 * (int64_t)n;
    o[0] = q[0]; o[1] = q[1]; o[2] = q[2];
}
template <typename VT>
__device__ __forceinline__ void st3(VT* __restrict__ v, int n, const VT o[3]) {
    VT* q = v + 3 * (int64_t)n;
    q[0] = o[0]; q[1] = o[1]; q[2] = o[2];
}

// stop the scene's CG loop: every block of the kernel reaches the same decision; the first
// block records it
__device__ __forceinline__ void record_stop(InstDev* ins, bool first, int reason, int iters, double err) {
    if (first && threadIdx.x == 0) {
        ins->stopped = 1;
        ins->end_reason = reason;
        ins->last_iters = iters;
        ins->err = err;
    }
}

// the scene's tile partials are folded into its finalize-block partials (finalize block bi
// of nf adds tiles [t0 + bi*ceil(nt/nf), ...)), so consumers reduce only nf values
__device__ __forceinline__ double fold_tiles(const double* __restrict__ buf, const Ranges& rg, int bi) {
    if (rg.nf <= 0) return 0.0;
    const int chunk = (rg.nt + rg.nf - 1) / rg.nf;
    const int a0 = rg.t0 + bi * chunk;
    const int a1 = min(rg.t0 + rg.nt, a0 + chunk);
    double acc = 0.0;
    for (int k = a0; k < a1; k++) acc = acc + buf[k];
    return acc;
}

template <typename R, int NR>
struct TileArgs {
    const TileDesc* __restrict__ tiles;
    InstDev* __restrict__ inst;
    const Ranges* __restrict__ ranges;
    ElemView<R, NR> ev;                 // permuted element SoA (ev.tet unused)
    const uint2* __restrict__ ltet;     // 4 x uint16 tile-local node indices per element
    TileNodes tn;
    const double* __restrict__ mass;
    const int8_t* __restrict__ fixed;
    int8_t* __restrict__ contact;        // PlaneForceField contacts of the step (per node)
    const R* __restrict__ emass;        // per element rho V / 20 (consistent MeshMatrixMass), else unused
    float* __restrict__ x_lo;           // MIXED engines: float copy of the positions (written by the RHS)
    double* __restrict__ nd;            // reference order: per-node dot terms for k_seqdot
    const float* __restrict__ dinv;     // MIXED inner CG with a Jacobi preconditioner: the inverted
                                        // diagonal per node (z = dinv r); null: plain CG
};

// consistent-mass edge terms of one element, folded into its nodal contributions (which the
// node sums subtract): vertex v of the element receives emass * factor * (sum of the other
// three vertices' values), i.e. the element's share of the edge loop of MeshMatrixMass::addMDx
template <typename R>
__device__ __forceinline__ void add_edge_mass(R mf, const R u[12], R c[12]) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const R sum = u[i] + u[3 + i] + u[6 + i] + u[9 + i];
#pragma unroll
        for (int v = 0; v < 4; v++) c[3 * v + i] -= mf * (sum - u[3 * v + i]);
    }
}

// acc (+|-)= the node's element contributions lc[slot + c kTileElems] in incidence-list order.
// Four incidences per round: their slots are read together, then their values, so a round
// waits on LDS twice instead of eight times; the adds keep the list order (same bits).
template <bool SUB, typename R, typename VT>
__device__ __forceinline__ void sum_incidences(const uint16_t* linc, int s0, int s1, const R* lc, VT acc[3]) {
    int s = s0;
    for (; s + 4 <= s1; s += 4) {
        int sl[4];
#pragma unroll
        for (int t = 0; t < 4; t++) sl[t] = linc[s + t];
        R v[4][3];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < 3; i++) v[t][i] = lc[sl[t] + i * kTileElems];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < 3; i++) {
                if (SUB) acc[i] -= (VT)v[t][i];
                else acc[i] += (VT)v[t][i];
            }
    }
    for (; s < s1; s++) {
        const int sl = linc[s];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (SUB) acc[i] -= (VT)lc[sl + i * kTileElems];
            else acc[i] += (VT)lc[sl + i * kTileElems];
        }
    }
}

// reference order: each incidence's contribution lc[slot + c kTileElems] to its own global slot
// base + rank (the node's slots run in ascending element id, the order its finalize adds them);
// rank: the tile's incidence ranks, indexed like linc
template <typename R, typename VT>
__device__ __forceinline__ void store_incidences(const uint16_t* linc, const uint8_t* __restrict__ rank, int s0,
                                                 int s1, const R* lc, VT* __restrict__ sp, int base) {
    for (int s = s0; s < s1; s++) {
        const int sl = linc[s];
        VT* o = sp + 3 * (int64_t)(base + rank[s]);
#pragma unroll
        for (int i = 0; i < 3; i++) o[i] = (VT)lc[sl + i * kTileElems];
    }
}

// acc (+|-)= the partials or slots sp[s0 .. s1) in that order: the loads four at a time, then
// their adds in slot order (the same operations as one slot after the other)
template <bool SUB, typename VT>
__device__ __forceinline__ void sum_slots(const VT* __restrict__ sp, int s0, int s1, VT acc[3]) {
    int s = s0;
    for (; s + 4 <= s1; s += 4) {
        VT t[4][3];
#pragma unroll
        for (int u = 0; u < 4; u++) ld3(sp, s + u, t[u]);
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int i = 0; i < 3; i++) {
                if (SUB) acc[i] -= t[u][i];
                else acc[i] += t[u][i];
            }
    }
    for (; s < s1; s++) {
        VT t[3];
        ld3(sp, s, t);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (SUB) acc[i] -= t[i];
            else acc[i] += t[i];
        }
    }
}

__device__ __forceinline__ void unpack_ltet(uint2 w, int l[4]) {
    l[0] = (int)(w.x & 0xffffu); l[1] = (int)(w.x >> 16); l[2] = (int)(w.y & 0xffffu); l[3] = (int)(w.y >> 16);
}

// ---------------------------------------------------------------------------------------
// RHS: rotations, element force, K v; interior nodes finished, shared nodes -> partials.
// TB threads, one element (and at most one tile-local node) per thread for TB = 512.  One LDS
// contribution buffer serves both element passes: the force pass, the node sums of f, then the
// K v pass (element record and rotation kept in registers) and the node sums of b.
// ---------------------------------------------------------------------------------------
// EXACT: the reference's summation order -- a shared node's contributions go to per-incidence
// slots (f to spf, K v to spv) instead of per-tile partials, and every node's |b|^2 term to nd
template <typename R, int NR, typename VT, int TB = 512, bool EXACT = false>
__global__ __launch_bounds__(TB) void k_tile_rhs(TileArgs<R, NR> a, const VT* __restrict__ x,
                                                  const VT* __restrict__ v, VT* __restrict__ f,
                                                  VT* __restrict__ b, VT* __restrict__ r, VT* __restrict__ xcg,
                                                  VT* __restrict__ spf, VT* __restrict__ spv,
                                                  double* __restrict__ pbb) {
    constexpr int EPT = kTileElems / TB;
    constexpr int NPT = (kTileNodes + TB - 1) / TB;
    __shared__ R lx[kTileNodes * 3];
    __shared__ R lv[kTileNodes * 3];
    __shared__ R lc[kTileElems * 12];
    __shared__ uint16_t linc[kTileElems * 4];
    __shared__ int lptr[kTileNodes + 1];
    __shared__ double sh[TB / 64];
    const TileDesc td = a.tiles[blockIdx.x];
    const InstDev& in = a.inst[td.scene];
    // element records of this thread's elements, in registers for both passes
    R rin[EPT][NR], kk[EPT][3], em[EPT], rot[EPT][9];
    uint2 lt[EPT];
#pragma unroll
    for (int m = 0; m < EPT; m++) {
        const int k = threadIdx.x + m * TB;
        if (k < td.e_count) {
            const int64_t e = (int64_t)td.e_begin + k;
            lt[m] = a.ltet[e];
#pragma unroll
            for (int i = 0; i < NR; i++) rin[m][i] = a.ev.rest[i * a.ev.E + e];
#pragma unroll
            for (int i = 0; i < 3; i++) kk[m][i] = a.ev.kmat[i * a.ev.E + e];
            em[m] = in.mass_kind == kMassMatrix ? a.emass[e] : (R)0;
        }
    }
    // node tables, incidence tables and node gathers at clamped indices (no branch: all in flight
    // together; values past the tile's counts unused), then staged in LDS
    int gid[NPT], psl[NPT];
#pragma unroll
    for (int m = 0; m < NPT; m++) {
        const int li = td.ln_begin + min((int)threadIdx.x + m * TB, td.ln_count - 1);
        gid[m] = a.tn.gid[li];
        psl[m] = a.tn.pslot[li];
    }
    const int ninc = 4 * td.e_count;
    constexpr int IPT = (4 * kTileElems + TB - 1) / TB;     // incidences per thread
    constexpr int PPT = (kTileNodes + 1 + TB - 1) / TB;     // incidence pointers per thread
    int inc_r[IPT], ptr_r[PPT];
#pragma unroll
    for (int k = 0; k < IPT; k++)
        inc_r[k] = a.tn.inc[ninc > 0 ? td.inc_begin + min((int)threadIdx.x + k * TB, ninc - 1) : 0];
#pragma unroll
    for (int k = 0; k < PPT; k++)
        ptr_r[k] = a.tn.inc_ptr[td.ln_begin + min((int)threadIdx.x + k * TB, td.ln_count)] - td.inc_begin;
    VT xt[NPT][3], vt[NPT][3];
#pragma unroll
    for (int m = 0; m < NPT; m++) {
        ld3(x, gid[m], xt[m]);
        ld3(v, gid[m], vt[m]);
    }
#pragma unroll
    for (int m = 0; m < NPT; m++) {
        const int j = threadIdx.x + m * TB;
        if (j < td.ln_count) {
            lx[3 * j] = (R)xt[m][0]; lx[3 * j + 1] = (R)xt[m][1]; lx[3 * j + 2] = (R)xt[m][2];
            if (a.x_lo) {   // every tile of a shared node writes the same value
                const float xf[3] = {(float)xt[m][0], (float)xt[m][1], (float)xt[m][2]};
                st3(a.x_lo, gid[m], xf);
            }
            lv[3 * j] = (R)vt[m][0]; lv[3 * j + 1] = (R)vt[m][1]; lv[3 * j + 2] = (R)vt[m][2];
        }
    }
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const int j = threadIdx.x + k * TB;
        if (j < ninc) linc[j] = (uint16_t)inc_r[k];
    }
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const int j = threadIdx.x + k * TB;
        if (j <= td.ln_count) lptr[j] = ptr_r[k];
    }
    __syncthreads();
    // ---- element force (accumulateForceLarge); the rotations are stored for the operator
#pragma unroll
    for (int m = 0; m < EPT; m++) {
        const int k = threadIdx.x + m * TB;
        if (k < td.e_count) {
            const int64_t e = (int64_t)td.e_begin + k;
            int l[4];
            unpack_ltet(lt[m], l);
            R p[4][3], c[12];
#pragma unroll
            for (int n = 0; n < 4; n++)
#pragma unroll
                for (int i = 0; i < 3; i++) p[n][i] = lx[3 * l[n] + i];
            elem_force_r<R, NR>(p, rin[m], kk[m], rot[m], c);
#pragma unroll
            for (int i = 0; i < 9; i++) a.ev.rot[i * a.ev.E + e] = rot[m][i];
#pragma unroll
            for (int i = 0; i < 12; i++) lc[i * kTileElems + k] = c[i];
        }
    }
    __syncthreads();
    // ---- node sums of f in ascending element id (interior nodes: with gravity)
    VT fn[NPT][3];
#pragma unroll
    for (int m = 0; m < NPT; m++) {
        const int j = threadIdx.x + m * TB;
        if (j < td.ln_count) {
            if (psl[m] < 0) mass_gravity(in, a.mass, gid[m], fn[m]);
            else { fn[m][0] = 0; fn[m][1] = 0; fn[m][2] = 0; }
            const int s0 = lptr[j], s1 = lptr[j + 1];
            if (EXACT && psl[m] >= 0) store_incidences(linc, a.tn.rank + td.inc_begin, s0, s1, lc, spf, psl[m]);
            else sum_incidences<false>(linc, s0, s1, lc, fn[m]);
        }
    }
    __syncthreads();
    // ---- K v (+ the consistent-mass edge terms) into the same buffer
    const R krhs = (R)in.k_rhs;
    const bool edge = in.mass_kind == kMassMatrix && in.m_rhs != 0.0;
#pragma unroll
    for (int m = 0; m < EPT; m++) {
        const int k = threadIdx.x + m * TB;
        if (k < td.e_count) {
            int l[4];
            unpack_ltet(lt[m], l);
            R u[12], c[12];
#pragma unroll
            for (int n = 0; n < 4; n++)
#pragma unroll
                for (int i = 0; i < 3; i++) u[3 * n + i] = lv[3 * l[n] + i];
            if (in.k_rhs != 0.0) {
                elem_dforce_r<R, NR>(rot[m], rin[m], kk[m], u, krhs, c);
            } else {
#pragma unroll
                for (int i = 0; i < 12; i++) c[i] = 0;
            }
            if (edge) add_edge_mass((R)in.m_rhs * em[m], u, c);
#pragma unroll
            for (int i = 0; i < 12; i++) lc[i * kTileElems + k] = c[i];
        }
    }
    __syncthreads();
    // ---- b = h (f - rm M v - (h + rs) K v), projected; shared nodes leave partials
    double bb = 0.0;
    const VT h = (VT)in.b_scale;   // h (firstOrder: 1, b = f)
#pragma unroll
    for (int m = 0; m < NPT; m++) {
        const int j = threadIdx.x + m * TB;
        if (j < td.ln_count) {
            const int g = gid[m];
            const int s0 = lptr[j], s1 = lptr[j + 1];
            if (psl[m] < 0) {
                const VT vn[3] = {(VT)lv[3 * j], (VT)lv[3 * j + 1], (VT)lv[3 * j + 2]};
                // PlaneForceField after the FEM (scene order); the contact is kept for addDForce
                bool touch = false;
                if (in.plane_on) {
                    const VT xn[3] = {(VT)lx[3 * j], (VT)lx[3 * j + 1], (VT)lx[3 * j + 2]};
                    touch = plane_force(in, xn, vn, fn[m]);
                    a.contact[g] = touch ? 1 : 0;
                }
                st3(f, g, fn[m]);
                VT bn[3] = {fn[m][0], fn[m][1], fn[m][2]};
                if (in.m_rhs != 0.0) mass_mdx(in, NodeArrays{nullptr, nullptr, a.mass, a.fixed}, g, in.m_rhs, vn, bn);
                sum_incidences<true>(linc, s0, s1, lc, bn);
                if (touch && in.k_rhs != 0.0) plane_dforce(in, in.k_rhs, vn, bn);
                bn[0] *= h; bn[1] *= h; bn[2] *= h;
                if (a.fixed[g]) { bn[0] = 0; bn[1] = 0; bn[2] = 0; }
                st3(b, g, bn);
                st3(r, g, bn);
                const VT z[3] = {0, 0, 0};
                if (!in.warm_start) st3(xcg, g, z);   // warmStart: x stays the previous solution
                if (EXACT) a.nd[g] = dot3d(bn, bn);
                else bb += dot3d(bn, bn);
            } else if (EXACT) {
                store_incidences(linc, a.tn.rank + td.inc_begin, s0, s1, lc, spv, psl[m]);
            } else {
                const int ps = psl[m] & (kSlotOwner - 1);
                VT pv[3] = {0, 0, 0};
                sum_incidences<false>(linc, s0, s1, lc, pv);
                st3(spf, ps, fn[m]);
                st3(spv, ps, pv);
            }
        }
    }
    if (!EXACT) {
        const double tot = wave_block_sum<TB>(bb, sh);
        if (threadIdx.x == 0) pbb[td.pslot] = tot;
    }
}

template <typename VT, bool EXACT = false>
__global__ __launch_bounds__(kBlock) void k_fin_rhs(const FinBlk* __restrict__ fb_, const InstDev* __restrict__ inst,
                                                     const Ranges* __restrict__ ranges, SharedNodes sn, const double* __restrict__ mass,
                                                     const int8_t* __restrict__ fixed, const VT* __restrict__ v,
                                                     const VT* __restrict__ spf, const VT* __restrict__ spv,
                                                     VT* __restrict__ f, VT* __restrict__ b, VT* __restrict__ r,
                                                     VT* __restrict__ xcg, double* __restrict__ pbb,
                                                     const VT* __restrict__ x, int8_t* __restrict__ contact,
                                                     double* __restrict__ nd = nullptr) {
    __shared__ double sh[kBlock / 64];
    const FinBlk fb = fb_[blockIdx.x];
    const InstDev& in = inst[fb.scene];
    const int j = fb.begin + threadIdx.x;
    double bb = 0.0;
    if (j < fb.end) {
        const int g = sn.gid[j];
        VT fn[3], vn[3];
        ld3(v, g, vn);
        mass_gravity(in, mass, g, fn);
        const int s0 = sn.ptr[j], s1 = sn.ptr[j + 1];
        sum_slots<false>(spf, s0, s1, fn);
        // PlaneForceField after the FEM (scene order); the contact is kept for addDForce
        bool touch = false;
        if (in.plane_on) {
            VT xn[3];
            ld3(x, g, xn);
            touch = plane_force(in, xn, vn, fn);
            contact[g] = touch ? 1 : 0;
        }
        st3(f, g, fn);
        VT bn[3] = {fn[0], fn[1], fn[2]};
        if (in.m_rhs != 0.0) mass_mdx(in, NodeArrays{nullptr, nullptr, mass, fixed}, g, in.m_rhs, vn, bn);
        sum_slots<true>(spv, s0, s1, bn);
        if (touch && in.k_rhs != 0.0) plane_dforce(in, in.k_rhs, vn, bn);
        const VT h = (VT)in.b_scale;   // h (firstOrder: 1, b = f)
        bn[0] *= h; bn[1] *= h; bn[2] *= h;
        if (fixed[g]) { bn[0] = 0; bn[1] = 0; bn[2] = 0; }
        st3(b, g, bn);
        st3(r, g, bn);
        const VT z[3] = {0, 0, 0};
        if (!in.warm_start) st3(xcg, g, z);   // warmStart: x stays the previous solution
        bb = dot3d(bn, bn);
        if (EXACT) nd[g] = bb;
    }
    if (EXACT) return;
    const double own = wave_block_sum<kBlock>(bb, sh);
    if (threadIdx.x == 0) pbb[fb.pslot] = fold_tiles(pbb, ranges[fb.scene], fb.bi) + own;
}

// ---------------------------------------------------------------------------------------
// CG iteration `iter` (1-based): K1 tile apply, K2 shared finalize, K3 update.
// iter = 0 is a plain operator application q = A r (no CG scalars, p untouched), used for
// the FP64 residual of iterative refinement.
// ---------------------------------------------------------------------------------------
// PLANE = false: no scene of the launch has a PlaneForceField (its per-node terms compiled out).
// GEO: how much element geometry is recomputed from node positions instead of read back, with
// the functions and operation order of the element set-up and no fused multiply-adds (so the
// same bits).  0: none (rotation, rest coordinates, K loaded).  1: the rotation, from the
// current positions x (the ones the force pass used: x changes only after the solve) -- 9 reals
// less per element, 3 more per tile-local node.  2: also the rotated rest coordinates from the
// rest positions x0 (initLarge, TetrahedronFEMForceField.inl:835-868) and, for a scene of one
// material, K from the rest volume -- NR (+ 3) reals less per element, 3 more per local node.
// EXACT: the reference's summation order -- |b|^2 and r.r come from k_seqdot (InstDev::dots),
// interior nodes write their p.q term to nd; a shared node's leading run (its first incidences in
// ascending element id, when they lie in this tile) is summed here after the mass term into the
// node's first slot, every later incidence goes to a slot of its own (finished by k_fin_exact);
// no tile partial of p.q is formed.
template <typename R, int NR, typename VT, int TB = kBlock, int GEO = 0, bool WIDE = true, bool PLANE = true,
          bool EXACT = false>
__global__ __launch_bounds__(TB) void k_tile_apply(TileArgs<R, NR> a, int iter, const VT* __restrict__ r,
                                                    VT* __restrict__ p, VT* __restrict__ q, VT* __restrict__ sp,
                                                    const double* __restrict__ pin,   // |b|^2 (iter 1) or r.r partials
                                                    double* __restrict__ pden,
                                                    int pin_nodes,     // 1: iter-1 |b|^2 partials per node block
                                                    const VT* __restrict__ x, const VT* __restrict__ x0) {
    constexpr int EPT = kTileElems / TB;                   // elements per thread
    constexpr int NPT = (kTileNodes + TB - 1) / TB;        // tile-local nodes per thread
    static_assert(kTileElems * 12 >= kTileNodes * 6, "positions are staged in the contribution buffer");
    __shared__ R lu[kTileNodes * 3];
    __shared__ R lc[kTileElems * 12];
    __shared__ uint16_t linc[kTileElems * 4];
    __shared__ int lptr[kTileNodes + 1];
    __shared__ double sh[TB / 64];
    const TileDesc td = a.tiles[blockIdx.x];
    const bool consistent = (td.flags & kTileConsistent) != 0;
    const bool kvol = GEO == 2 && (td.flags & kTileOneMaterial) != 0;   // K recomputed, not loaded
    // ---- 1. issue every load that depends on the tile descriptor only: the element records of
    //         this thread (kept in registers), the tile-local node ids / slots / masks, the
    //         incidence tables, the scene record
    R rot[EPT][9], rin[EPT][NR], kk[EPT][3], em[EPT];
    uint2 lt[EPT];
#pragma unroll
    for (int m = 0; m < EPT; m++) {
        const int k = threadIdx.x + m * TB;
        if (k < td.e_count) {
            const int64_t e = (int64_t)td.e_begin + k;
            lt[m] = a.ltet[e];
            if (GEO == 0) {
#pragma unroll
                for (int i = 0; i < 9; i++) rot[m][i] = a.ev.rot[i * a.ev.E + e];
            }
            if (GEO < 2) {
#pragma unroll
                for (int i = 0; i < NR; i++) rin[m][i] = a.ev.rest[i * a.ev.E + e];
            }
            if (!kvol) {
#pragma unroll
                for (int i = 0; i < 3; i++) kk[m][i] = a.ev.kmat[i * a.ev.E + e];
            }
            em[m] = consistent ? a.emass[e] : (R)0;
        }
    }
    // tile-local node tables: every thread loads, at its node or (past the tile's count) at the
    // tile's last node, which every tile has -- no branch, so the compiler keeps all these loads
    // and the gathers that depend on them in flight together; values past the count are unused
    int gid[NPT], psl[NPT], lead[NPT];
    int8_t fx[NPT], ct[NPT];
#pragma unroll
    for (int m = 0; m < NPT; m++) {
        const int j = min((int)threadIdx.x + m * TB, td.ln_count - 1);
        const int li = td.ln_begin + j;
        gid[m] = a.tn.gid[li];
        psl[m] = a.tn.pslot[li];
        lead[m] = EXACT ? a.tn.lead[li] : 0;
    }
    InstDev* ins = &a.inst[td.scene];
    // a refinement sweep that keeps the previous search direction continues CG at its iteration 1
    const bool cont = iter == 1 && ins->keep_dir && ins->sweep > 1;
    // the incidence tables into registers first: their loads are in flight while the node ids
    // arrive and the node gathers below are issued; they go to LDS after those
    const int ninc = 4 * td.e_count;
    constexpr int IPT = (4 * kTileElems + TB - 1) / TB;     // incidences per thread
    constexpr int PPT = (kTileNodes + 1 + TB - 1) / TB;     // incidence pointers per thread
    // (clamped like the node tables; a tile without elements reads entry 0, which every buffer has)
    int inc_r[IPT], ptr_r[PPT];
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const int j = threadIdx.x + k * TB;
        inc_r[k] = a.tn.inc[ninc > 0 ? td.inc_begin + min(j, ninc - 1) : 0];
    }
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const int j = min((int)threadIdx.x + k * TB, td.ln_count);
        ptr_r[k] = a.tn.inc_ptr[td.ln_begin + j] - td.inc_begin;
    }
    // ---- 2. the tile's nodes: r and p (dependent on the node ids), masks; issued before the
    //         scene's stop flag is tested (in-bounds loads, unused for a stopped scene)
    VT rr_[NPT][3], po_[NPT][3], xx_[NPT][3], x0_[NPT][3];
    float dv_[NPT][3];
    // p is read from iteration 2 on (or continuing a direction); without a p (a plain operator
    // application) the loads read r instead, unused
    const VT* __restrict__ pl = p ? p : r;
    const bool plane = PLANE && ins->plane_on;
#pragma unroll
    for (int m = 0; m < NPT; m++) {
        ld3(r, gid[m], rr_[m]);
        if (a.dinv) ld3(a.dinv, gid[m], dv_[m]);
        ld3(pl, gid[m], po_[m]);
        if (GEO >= 1) ld3(x, gid[m], xx_[m]);
        if (GEO == 2) ld3(x0, gid[m], x0_[m]);
        fx[m] = a.fixed[gid[m]];
        if (PLANE) {
            const int8_t c = a.contact[gid[m]];
            ct[m] = plane ? c : 0;
        } else {
            ct[m] = 0;
        }
    }
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const int j = threadIdx.x + k * TB;
        if (j < ninc) linc[j] = (uint16_t)inc_r[k];
    }
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const int j = threadIdx.x + k * TB;
        if (j <= td.ln_count) lptr[j] = ptr_r[k];
    }
    if (ins->stopped) return;
    const Ranges rg = a.ranges[td.scene];
    const bool first = (td.flags & kTileFirst) != 0;
    // ---- 3. scalars: rho, the tolerance test, beta (CGLinearSolver.inl:130-196)
    double beta = 0.0;
    if (cont) {
        // the refinement replaced r by the true residual: beta from its r.r and the r.r that
        // built p; the tolerance is relative to the first sweep's |r|
        const double rho = scene_sum<TB, WIDE>(pin, rg.b0, rg.nb, sh);
        const double err = sqrt(rho) / ins->normb;
        if (rho == 0.0 || err <= ins->tol) { record_stop(ins, first, 1, 1, err); return; }
        beta = rho / ins->rho_last;
        if (first && threadIdx.x == 0) {
            ins->rho_slot[1] = rho; ins->beta = beta; ins->iter = 1; ins->err = err;
        }
    } else if (iter == 1) {
        const double bb = EXACT ? ins->dots[kDotBB]
                          : pin_nodes ? scene_sum<TB, WIDE>(pin, rg.b0, rg.nb, sh)
                          : rg.nf > 0 ? scene_sum<TB, WIDE>(pin, rg.f0, rg.nf, sh) : scene_sum<TB, WIDE>(pin, rg.t0, rg.nt, sh);
        const double normb = sqrt(bb);
        if (normb == 0.0) { record_stop(ins, first, 4, 0, 1.0); return; }
        // rho = r.r: b.b from a cold start; warmStart (reference order): the r.r of r = b - A x
        const double rho = EXACT && ins->warm_start ? ins->dots[kDotRR] : bb;
        const double err = sqrt(rho) / normb;
        if (err <= ins->tol && !ins->first_solve) { record_stop(ins, first, 1, 1, err); return; }
        if (first && threadIdx.x == 0) {
            ins->normb = normb; ins->rho_slot[1] = rho; ins->iter = 1; ins->err = err;
        }
    } else if (iter > 1) {
        if (iter > ins->max_iter) { record_stop(ins, first, 0, ins->max_iter, ins->err); return; }
        const double rho = EXACT ? ins->dots[kDotRR] : scene_sum<TB, WIDE>(pin, rg.b0, rg.nb, sh);
        const double err = sqrt(rho) / ins->normb;
        if (err <= ins->tol) { record_stop(ins, first, 1, iter, err); return; }
        const double rho_1 = ins->rho_slot[(iter - 1) & 1];
        beta = rho / rho_1;
        if (first && threadIdx.x == 0) {
            ins->rho_slot[iter & 1] = rho; ins->beta = beta; ins->iter = iter; ins->err = err;
        }
    }
    // ---- 4. stage p = r + beta p, or z + beta p with z = dinv r (interior nodes: this tile owns
    //         p and writes it back)
    const VT vb = (VT)beta;
#pragma unroll
    for (int m = 0; m < NPT; m++) {
        const int j = threadIdx.x + m * TB;
        if (j < td.ln_count) {
            VT pp[3] = {rr_[m][0], rr_[m][1], rr_[m][2]};
            if (a.dinv) {
#pragma unroll
                for (int i = 0; i < 3; i++) pp[i] = pp[i] * (VT)dv_[m][i];
            }
            if (iter > 1 || cont) {
#pragma unroll
                for (int i = 0; i < 3; i++) { VT t = po_[m][i] * vb; t += pp[i]; pp[i] = t; }
            }
            if (iter > 0 && psl[m] < 0) st3(p, gid[m], pp);
            lu[3 * j] = (R)pp[0]; lu[3 * j + 1] = (R)pp[1]; lu[3 * j + 2] = (R)pp[2];
            if (GEO >= 1) {
#pragma unroll
                for (int i = 0; i < 3; i++) lc[3 * j + i] = (R)xx_[m][i];
            }
            if (GEO == 2) {
#pragma unroll
                for (int i = 0; i < 3; i++) lc[3 * kTileNodes + 3 * j + i] = (R)x0_[m][i];
            }
        }
    }
    __syncthreads();
    if (GEO >= 1) {   // geometry from the staged positions, then the buffer is free again
        const R* lx = lc;
        const R* lx0 = lc + 3 * kTileNodes;
#pragma unroll
        for (int m = 0; m < EPT; m++) {
            const int k = threadIdx.x + m * TB;
            if (k < td.e_count) {
                int l[4];
                unpack_ltet(lt[m], l);
                R p[3][3];
#pragma unroll
                for (int v = 0; v < 3; v++)
#pragma unroll
                    for (int i = 0; i < 3; i++) p[v][i] = lx[3 * l[v] + i];
                rotation_large(p[0], p[1], p[2], rot[m]);
                if (GEO == 2) {
                    R p0[4][3];
#pragma unroll
                    for (int v = 0; v < 4; v++)
#pragma unroll
                        for (int i = 0; i < 3; i++) p0[v][i] = lx0[3 * l[v] + i];
                    rest_frame<R, NR>(p0, rin[m]);
                    if (kvol) stiffness_from_volume(p0, ins->ks, kk[m]);
                }
            }
        }
        __syncthreads();
    }
    // ---- 5. element K p into LDS
    const R kcg = (R)ins->k_cg;
    const R mcg = (R)ins->m_cg;
#pragma unroll
    for (int m = 0; m < EPT; m++) {
        const int k = threadIdx.x + m * TB;
        if (k < td.e_count) {
            int l[4];
            unpack_ltet(lt[m], l);
            R u[12], c[12];
#pragma unroll
            for (int n = 0; n < 4; n++)
#pragma unroll
                for (int i = 0; i < 3; i++) u[3 * n + i] = lu[3 * l[n] + i];
            elem_dforce_r<R, NR>(rot[m], rin[m], kk[m], u, kcg, c);
            if (consistent) add_edge_mass(mcg * em[m], u, c);
#pragma unroll
            for (int i = 0; i < 12; i++) lc[i * kTileElems + k] = c[i];
        }
    }
    __syncthreads();
    // ---- 6. per-node sums in ascending element id, from LDS only
    const InstDev& in = *ins;
    double pq = 0.0;
#pragma unroll
    for (int m = 0; m < NPT; m++) {
        const int j = threadIdx.x + m * TB;
        if (j < td.ln_count) {
            const int s0 = lptr[j], s1 = lptr[j + 1];
            if (psl[m] < 0) {
                const VT pp[3] = {(VT)lu[3 * j], (VT)lu[3 * j + 1], (VT)lu[3 * j + 2]};
                VT qn[3] = {0, 0, 0};
                if (in.m_cg != 0.0) mass_mdx(in, NodeArrays{nullptr, nullptr, a.mass, a.fixed}, gid[m], in.m_cg, pp, qn);
                sum_incidences<true>(linc, s0, s1, lc, qn);
                if (PLANE && ct[m]) plane_dforce(in, in.k_cg, pp, qn);
                if (fx[m]) { qn[0] = 0; qn[1] = 0; qn[2] = 0; }
                st3(q, gid[m], qn);
                if (EXACT) a.nd[gid[m]] = dot3d(pp, qn);
                else pq += dot3d(pp, qn);
            } else if (EXACT) {
                int s = s0;
                if (lead[m] > 0) {
                    // this tile holds the node's leading run: the node's sum starts here (the mass
                    // term, then the run's contributions one by one) and goes to its first slot
                    const VT pp[3] = {(VT)lu[3 * j], (VT)lu[3 * j + 1], (VT)lu[3 * j + 2]};
                    VT acc[3] = {0, 0, 0};
                    if (in.m_cg != 0.0) mass_mdx(in, NodeArrays{nullptr, nullptr, a.mass, a.fixed}, gid[m], in.m_cg, pp, acc);
                    s = s0 + lead[m];
                    sum_incidences<true>(linc, s0, s, lc, acc);
                    st3(sp, psl[m], acc);
                }
                store_incidences(linc, a.tn.arank + td.inc_begin, s, s1, lc, sp, psl[m]);
            } else {
                VT acc[3] = {0, 0, 0};
                sum_incidences<true>(linc, s0, s1, lc, acc);
                st3(sp, psl[m] & (kSlotOwner - 1), acc);
                // p.q is linear in q: a shared node's term is split over its tiles (p . this
                // tile's partial), its mass and plane terms added by its first tile, so that the
                // tile partials hold all of p.q and no later kernel has to add to it
                if (!fx[m]) {
                    const VT pp[3] = {(VT)lu[3 * j], (VT)lu[3 * j + 1], (VT)lu[3 * j + 2]};
                    pq += dot3d(pp, acc);
                    if ((psl[m] & kSlotOwner) && (in.m_cg != 0.0 || (PLANE && ct[m]))) {
                        VT mq[3] = {0, 0, 0};
                        if (in.m_cg != 0.0)
                            mass_mdx(in, NodeArrays{nullptr, nullptr, a.mass, a.fixed}, gid[m], in.m_cg, pp, mq);
                        if (PLANE && ct[m]) plane_dforce(in, in.k_cg, pp, mq);
                        pq += dot3d(pp, mq);
                    }
                }
            }
        }
    }
    if (!EXACT) {
        const double tot = wave_block_sum<TB>(pq, sh);
        if (threadIdx.x == 0) pden[td.pslot] = tot;
    }
}

// reference order, after the tile apply: the shared nodes' p = r + beta p, q = (1 + h rm) M p
// minus their element contributions one by one in ascending element id (continued from the sum
// the leading run's tile left in the node's first slot), plane and projection, and their p.q
// terms for k_seqdot (the work k_update does for them in tile order otherwise)
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_fin_exact(const FinBlk* __restrict__ fb_, InstDev* __restrict__ inst,
                                                       SharedNodes sn, const double* __restrict__ mass,
                                                       const int8_t* __restrict__ fixed, int iter,
                                                       const VT* __restrict__ r, VT* __restrict__ p,
                                                       VT* __restrict__ q, const VT* __restrict__ sp,
                                                       double* __restrict__ nd, const int8_t* __restrict__ contact) {
    const FinBlk fb = fb_[blockIdx.x];
    const InstDev& in = inst[fb.scene];
    if (in.stopped) return;
    const int j = fb.begin + threadIdx.x;
    if (j >= fb.end) return;
    const int g = sn.gid[j];
    VT pp[3];
    ld3(r, g, pp);
    if (iter > 1) {
        const VT vb = (VT)in.beta;
        VT po[3];
        ld3(p, g, po);
#pragma unroll
        for (int i = 0; i < 3; i++) { VT t = po[i] * vb; t += pp[i]; pp[i] = t; }
    }
    st3(p, g, pp);
    // the node's first slot holds its sum up to the end of its leading run (the mass term and
    // that run, added by the run's tile); the rest of its contributions follow, one per slot
    const int s0 = sn.ptr[j], s1 = s0 + 1 + (int)sn.acnt[j];
    VT qn[3];
    ld3(sp, s0, qn);
    sum_slots<true>(sp, s0 + 1, s1, qn);
    if (in.plane_on && contact[g]) plane_dforce(in, in.k_cg, pp, qn);
    if (fixed[g]) { qn[0] = 0; qn[1] = 0; qn[2] = 0; }
    st3(q, g, qn);
    nd[g] = dot3d(pp, qn);
}

// CGLinearSolver warmStart in the reference order (CGLinearSolver.inl:95-99): after the operator
// of iteration 0 on the previous solution (q = A x, shared nodes finished by k_fin_exact),
// r.eq(b, q, -1.0): r = b + q (-1) per component, and each node's r.r term for k_seqdot
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_warm_begin_exact(const NodeBlk* __restrict__ nb_,
                                                              const InstDev* __restrict__ inst,
                                                              const VT* __restrict__ b, const VT* __restrict__ q,
                                                              VT* __restrict__ r, double* __restrict__ nd) {
    const NodeBlk nb = nb_[blockIdx.x];
    const InstDev& in = inst[nb.scene];
    if (!in.warm_start || in.stopped) return;
    const VT m1 = (VT)(-1.0);
#pragma unroll
    for (int k = 0; k < kNodesPerThread; k++) {
        const int n = nb.begin + k * kBlock + threadIdx.x;
        if (n < nb.end) {
            VT bn[3], qn[3], rn[3];
            ld3(b, n, bn);
            ld3(q, n, qn);
#pragma unroll
            for (int i = 0; i < 3; i++) rn[i] = bn[i] + qn[i] * m1;
            st3(r, n, rn);
            nd[n] = dot3d(rn, rn);
        }
    }
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void k_fin_apply(const FinBlk* __restrict__ fb_, InstDev* __restrict__ inst,
                                                       const Ranges* __restrict__ ranges, SharedNodes sn,
                                                       const double* __restrict__ mass,
                                                       const int8_t* __restrict__ fixed, int iter,
                                                       const VT* __restrict__ r, VT* __restrict__ p,
                                                       VT* __restrict__ q, const VT* __restrict__ sp,
                                                       double* __restrict__ pden,
                                                       const int8_t* __restrict__ contact) {
    __shared__ double sh[kBlock / 64];
    const FinBlk fb = fb_[blockIdx.x];
    const InstDev& in = inst[fb.scene];
    if (in.stopped) return;
    const int j = fb.begin + threadIdx.x;
    double pq = 0.0;
    if (j < fb.end) {
        const int g = sn.gid[j];
        VT pp[3];
        ld3(r, g, pp);
        if (iter > 1) {
            const VT vb = (VT)in.beta;
            VT po[3];
            ld3(p, g, po);
#pragma unroll
            for (int i = 0; i < 3; i++) { po[i] *= vb; po[i] += pp[i]; pp[i] = po[i]; }
        }
        if (iter > 0) st3(p, g, pp);
        VT qn[3] = {0, 0, 0};
        if (in.m_cg != 0.0) mass_mdx(in, NodeArrays{nullptr, nullptr, mass, fixed}, g, in.m_cg, pp, qn);
        const int s0 = sn.ptr[j], s1 = sn.ptr[j + 1];
        sum_slots<false>(sp, s0, s1, qn);
        if (in.plane_on && contact[g]) plane_dforce(in, in.k_cg, pp, qn);
        if (fixed[g]) { qn[0] = 0; qn[1] = 0; qn[2] = 0; }
        st3(q, g, qn);
        pq = dot3d(pp, qn);
    }
    const double own = wave_block_sum<kBlock>(pq, sh);
    if (threadIdx.x == 0) pden[fb.pslot] = fold_tiles(pden, ranges[fb.scene], fb.bi) + own;
}

// z = P^-1 r of a MIXED inner CG's preconditioner (the scene's kind: 0 none, 1 the inverted
// diagonal dinv, 3 per node, 2 the inverted 3x3 block binv, 9 per node, row-major), in float
template <typename VT>
__device__ __forceinline__ void precond_z(int kind, const float* __restrict__ dinv, const float* __restrict__ binv,
                                          int n, const VT r[3], VT z[3]) {
    if (kind == 1) {
        float dv[3];
        ld3(dinv, n, dv);
#pragma unroll
        for (int i = 0; i < 3; i++) z[i] = r[i] * (VT)dv[i];
    } else if (kind == 2) {
        const float* B = binv + 9 * (int64_t)n;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            VT t = r[0] * (VT)B[3 * i];
            t += r[1] * (VT)B[3 * i + 1];
            t += r[2] * (VT)B[3 * i + 2];
            z[i] = t;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 3; i++) z[i] = r[i];
    }
}

// CG iteration `iter` after the tile apply.  Nodes shared by several tiles are finished here:
// p = r + beta p, q = (1 + h rm) M p + their tile partials in tile order, projection (the work
// CGLinearSolver's p update and GraphScatteredMatrix::apply do for them).  Then den = p.q from
// the tile partials (which already hold every node's term), the threshold test and alpha
// (CGLinearSolver.inl:207-265), x += alpha p, r -= alpha q and the r.r partials.
// EXACT: every node's p and q are final (k_fin_exact finished the shared ones), den is the
// reference-order p.q of k_seqdot, and each node's r.r term goes to nd.  zbuf (MIXED inner CG
// with a preconditioner): it holds z = P^-1 r of the previous residual (shared nodes' p = z +
// beta p); each node's new z (precond_z with dinv / binv) replaces it, r.z partials in place of r.r
template <typename VT, bool EXACT = false>
__global__ __launch_bounds__(kBlock) void k_update(const NodeBlk* __restrict__ nb_, InstDev* __restrict__ inst,
                                                    const Ranges* __restrict__ ranges, int iter, SharedNodes sn,
                                                    const VT* __restrict__ sp, const double* __restrict__ mass,
                                                    const int8_t* __restrict__ fixed, VT* __restrict__ p,
                                                    const VT* __restrict__ q, VT* __restrict__ xcg, VT* __restrict__ r,
                                                    const double* __restrict__ pden, double* __restrict__ prho,
                                                    const int8_t* __restrict__ contact, double* __restrict__ nd = nullptr,
                                                    const float* __restrict__ dinv = nullptr,
                                                    const float* __restrict__ binv = nullptr, VT* __restrict__ zbuf = nullptr) {
    __shared__ double sh[kBlock / 64];
    const NodeBlk nb = nb_[blockIdx.x];
    InstDev* ins = &inst[nb.scene];
    // node data first (registers), then the reduction
    VT pv[kNodesPerThread][3], qv[kNodesPerThread][3], xv[kNodesPerThread][3], rv[kNodesPerThread][3];
    VT zv[kNodesPerThread][3];
    int shk[kNodesPerThread];
    // every load at a clamped node index (no branch: all in flight together; a node past the
    // block's end repeats its last node, unused), p and q also where the shared-node path below
    // replaces them
#pragma unroll
    for (int k = 0; k < kNodesPerThread; k++) {
        const int n = min(nb.begin + k * kBlock + (int)threadIdx.x, nb.end - 1);
        shk[k] = EXACT ? -1 : sn.of[n];
        ld3(xcg, n, xv[k]);
        ld3(r, n, rv[k]);
        if (zbuf) ld3(zbuf, n, zv[k]);
        ld3(p, n, pv[k]);
        ld3(q, n, qv[k]);
    }
    if (ins->stopped) return;
    const bool cont = iter == 1 && ins->keep_dir && ins->sweep > 1;
    const Ranges rg = ranges[nb.scene];
    const bool first = nb.first != 0;
    const double den = EXACT ? ins->dots[kDotPQ] : scene_sum<kBlock>(pden, rg.t0, rg.nt, sh);
    const bool exempt = (iter == 1 && ins->first_solve);
    if (den == 0.0) { record_stop(ins, first, 3, iter, ins->err); if (first && threadIdx.x == 0) ins->den = den; return; }
    if (fabs(den) <= ins->thr && !exempt) {
        record_stop(ins, first, 2, iter, ins->err);
        if (first && threadIdx.x == 0) ins->den = den;
        return;
    }
    const double alpha = ins->rho_slot[iter & 1] / den;
    if (first && threadIdx.x == 0) { ins->den = den; ins->alpha = alpha; }
    const InstDev& in = *ins;
    const VT vb = (VT)in.beta;
    const VT va = (VT)alpha, vma = (VT)(-alpha);
    double rr = 0.0;
#pragma unroll
    for (int k = 0; k < kNodesPerThread; k++) {
        const int n = nb.begin + k * kBlock + threadIdx.x;
        if (n < nb.end) {
            if (shk[k] >= 0) {
                VT pp[3] = {rv[k][0], rv[k][1], rv[k][2]};
                if (zbuf) {
#pragma unroll
                    for (int i = 0; i < 3; i++) pp[i] = zv[k][i];
                }
                if (iter > 1 || cont) {
#pragma unroll
                    for (int i = 0; i < 3; i++) { VT t = pv[k][i] * vb; t += pp[i]; pp[i] = t; }
                }
                st3(p, n, pp);
                VT qn[3] = {0, 0, 0};
                if (in.m_cg != 0.0) mass_mdx(in, NodeArrays{nullptr, nullptr, mass, fixed}, n, in.m_cg, pp, qn);
                const int s0 = sn.ptr[shk[k]], s1 = sn.ptr[shk[k] + 1];
                sum_slots<false>(sp, s0, s1, qn);
                if (in.plane_on && contact[n]) plane_dforce(in, in.k_cg, pp, qn);
                if (fixed[n]) { qn[0] = 0; qn[1] = 0; qn[2] = 0; }
#pragma unroll
                for (int i = 0; i < 3; i++) { pv[k][i] = pp[i]; qv[k][i] = qn[i]; }
            }
            VT xn[3], rn[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                xn[i] = xv[k][i] + pv[k][i] * va;
                rn[i] = rv[k][i] + qv[k][i] * vma;
            }
            st3(xcg, n, xn);
            st3(r, n, rn);
            if (EXACT) {
                nd[n] = dot3d(rn, rn);
            } else if (zbuf) {   // z = P^-1 r for the next iteration, and r.z
                VT zn[3];
                precond_z(ins->pcg, dinv, binv, n, rn, zn);
                st3(zbuf, n, zn);
                rr += dot3d(rn, zn);
            } else {
                rr += dot3d(rn, rn);
            }
        }
    }
    if (EXACT) return;
    const double tot = wave_block_sum<kBlock>(rr, sh);
    if (threadIdx.x == 0) prho[nb.pslot] = tot;
}

// start of a component-level CG solve on the tiled kernels: r = b = the caller's right-hand side,
// x = 0, and |b|^2 as per-node-block partials (read by the tile apply of iteration 1 with
// pin_nodes = 1) or, in the reference order, as per-node terms for k_seqdot
template <typename VT, bool EXACT = false>
__global__ __launch_bounds__(kBlock) void k_cg_begin_tiled(const NodeBlk* __restrict__ nb_, const VT* __restrict__ b_in,
                                                            int nb0, VT* __restrict__ b, VT* __restrict__ r,
                                                            VT* __restrict__ xcg, double* __restrict__ prho,
                                                            double* __restrict__ nd) {
    __shared__ double sh[kBlock / 64];
    const NodeBlk nb = nb_[blockIdx.x];
    double bb = 0.0;
#pragma unroll
    for (int k = 0; k < kNodesPerThread; k++) {
        const int n = nb.begin + k * kBlock + threadIdx.x;
        if (n < nb.end) {
            VT bn[3];
            ld3(b_in, n - nb0, bn);
            st3(b, n, bn);
            st3(r, n, bn);
            const VT z[3] = {0, 0, 0};
            st3(xcg, n, z);
            if (EXACT) nd[n] = dot3d(bn, bn);
            else bb += dot3d(bn, bn);
        }
    }
    if (EXACT) return;
    const double tot = wave_block_sum<kBlock>(bb, sh);
    if (threadIdx.x == 0) prho[nb.pslot] = tot;
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void k_integrate(const NodeBlk* __restrict__ nb_, InstDev* __restrict__ inst,
                                                       const int8_t* __restrict__ fixed,
                                                       const VT* __restrict__ xcg, VT* __restrict__ v,
                                                       VT* __restrict__ x) {
    const NodeBlk nb = nb_[blockIdx.x];
    const InstDev& in = inst[nb.scene];
    const VT h = (VT)in.dt;
    // v += dv (firstOrder: v = dv), x += v h, then vdamping's v *= exp(-h d)
    // (EulerImplicitSolver.cpp:246-299)
    const bool first_order = in.first_order != 0, damp = in.vdamp != 1.0;
    const VT vd = (VT)in.vdamp;
#pragma unroll
    for (int k = 0; k < kNodesPerThread; k++) {
        const int n = nb.begin + k * kBlock + threadIdx.x;
        if (n < nb.end) {
            const int64_t o = 3 * (int64_t)n;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const VT vv = first_order ? xcg[o + i] : v[o + i] + xcg[o + i];
                x[o + i] = x[o + i] + vv * h;
                v[o + i] = damp ? vv * vd : vv;
            }
            // FixedProjectiveConstraint::projectVelocity (.inl:237-258), when activated
            if (in.project_velocity && (fixed[n] & 1)) {
#pragma unroll
                for (int i = 0; i < 3; i++) v[o + i] = 0;
            }
        }
    }
}

// scene bookkeeping after the update (one thread per scene of the launch set)
static __global__ void k_step_end_tiled(const int* __restrict__ scenes, int n, InstDev* __restrict__ inst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    InstDev* in = &inst[scenes[i]];
    if (!in->stopped) { in->stopped = 1; in->end_reason = 0; in->last_iters = in->max_iter; }
    in->first_solve = 0;
    in->step_count += 1;
    in->time += in->dt;
    in->lm_x0_set = 1;
    in->stopped = 0;   // ready for the next solve
}

}  // namespace sofa_amd
