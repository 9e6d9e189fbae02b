// tiling.h -- host planning of element tiles for the LDS-gather kernels (fem_tiled.h).
//
// Elements are ordered by the Morton code of their rest centroid in element units (each axis
// scaled by the median element extent along it; ties by element id), or kept in element-id order
// (for meshes numbered slab by slab, such as generated grids: there a node's elements fall in one
// or two consecutive tiles, each holding a contiguous part of the node's list), and cut
// greedily into tiles of at most `max_elems` elements and `max_nodes` distinct nodes.  For
// each tile-local node the incidence list is sorted by ORIGINAL element id: the order in which
// the reference's element loop accumulates into that node (TetrahedronFEMForceField.inl:
// 1630-1634).  An incidence is the LDS slot of the element's contribution to the node: the
// kernels keep the 12 contribution components of a tile's elements as 12 rows of max_elems
// (row 3 v + c, column = tile-local element), so vertex v of local element k is the slot
// 3 v max_elems + k and its component c sits c max_elems further.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace sofa_amd {

struct TilePlan {
    std::vector<int> t_ebegin, t_ecount, t_lnbegin, t_lncount;   // per tile (scene-local offsets)
    std::vector<int> perm;          // permuted position -> original element id
    std::vector<int> iperm;         // original element id -> permuted position
    std::vector<uint2> ltet;        // per permuted element: 4 x uint16 tile-local node index
    std::vector<int> ln_gid;        // tile-local node -> scene-local node id
    std::vector<int> ln_inc_ptr;    // size(ln_gid) + 1
    std::vector<uint16_t> ln_inc;   // contribution slot: 3 max_elems vertex + local element
    std::vector<int> ln_pslot;      // -1 interior, else scene-local shared-partial slot
    std::vector<int> sh_gid;        // nodes finished by the finalize kernel (ascending)
    std::vector<int> sh_ptr;        // size(sh_gid) + 1
    std::vector<uint8_t> inc_rank;  // reference order only (plan_exact): per incidence of a shared
                                    // node, its rank in the node's list in ascending element id
    std::vector<uint8_t> inc_arank; // reference order, CG operator: 0 for an incidence of the node's
                                    // leading run (folded into the node's first slot by its tile),
                                    // else its slot's offset from the node's first slot
    std::vector<uint8_t> sh_acnt;   // reference order, CG operator: per shared node, its slots after
                                    // the first (the incidences after the leading run)
    std::vector<uint8_t> ln_lead;   // reference order, CG operator: per tile-local node, how many of
                                    // its first incidences its tile folds (the node's leading run
                                    // when the tile holds it, else 0)
    int n_slots = 0;
};

static inline uint64_t spread3(uint64_t v) {   // 21 bits -> every third bit
    v &= 0x1fffffULL;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

enum TileOrder { kOrderMorton = 0, kOrderElementId = 1 };
static inline TilePlan plan_tiles(int n_nodes, const double* rest, const int32_t* tets, int E, int max_elems,
                                  int max_nodes, TileOrder order = kOrderMorton) {
    TilePlan P;
    P.perm.resize(E);
    std::iota(P.perm.begin(), P.perm.end(), 0);
    if (order == kOrderMorton) {
        // Morton order of element centroids
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int i = 0; i < n_nodes; i++)
            for (int k = 0; k < 3; k++) {
                lo[k] = std::min(lo[k], rest[3 * i + k]);
                hi[k] = std::max(hi[k], rest[3 * i + k]);
            }
        // the curve runs in element units: each axis is scaled by the typical element extent along
        // it (median over a sample of elements), so that meshes of flat or long elements (a thin
        // beam cut into many thin layers) still get compact tiles
        double scale[3] = {1, 1, 1};
        {
            const int stride = std::max(1, E / 65536);
            std::vector<double> w[3];
            for (int e = 0; e < E; e += stride)
                for (int k = 0; k < 3; k++) {
                    double a = 1e300, b = -1e300;
                    for (int v = 0; v < 4; v++) {
                        const double c = rest[3 * tets[4 * e + v] + k];
                        a = std::min(a, c);
                        b = std::max(b, c);
                    }
                    w[k].push_back(b - a);
                }
            for (int k = 0; k < 3; k++) {
                if (w[k].empty()) continue;
                std::nth_element(w[k].begin(), w[k].begin() + w[k].size() / 2, w[k].end());
                const double m = w[k][w[k].size() / 2];
                if (m > 0) scale[k] = m;
            }
            if (!(scale[0] > 0 && scale[1] > 0 && scale[2] > 0)) scale[0] = scale[1] = scale[2] = 1;
        }
        double ext = 0;
        for (int k = 0; k < 3; k++) ext = std::max(ext, (hi[k] - lo[k]) / scale[k]);
        if (!(ext > 0)) ext = 1;
        std::vector<uint64_t> code(E);
        for (int e = 0; e < E; e++) {
            uint64_t c = 0;
            for (int k = 0; k < 3; k++) {
                double s = 0;
                for (int v = 0; v < 4; v++) s += rest[3 * tets[4 * e + v] + k];
                double u = (s * 0.25 - lo[k]) / scale[k] / ext;
                u = std::min(std::max(u, 0.0), 1.0);
                c |= spread3((uint64_t)(u * 2097151.0)) << k;
            }
            code[e] = c;
        }
        std::stable_sort(P.perm.begin(), P.perm.end(), [&](int a, int b) { return code[a] < code[b]; });
    }
    P.iperm.assign(E, 0);
    for (int i = 0; i < E; i++) P.iperm[P.perm[i]] = i;

    // greedy cut
    std::vector<int> stamp(n_nodes, -1);
    int t = -1, ecount = 0, ncount = 0;
    for (int i = 0; i < E; i++) {
        const int e = P.perm[i];
        int fresh = 0;
        for (int v = 0; v < 4; v++) {
            const int n = tets[4 * e + v];
            bool dup = false;
            for (int w = 0; w < v; w++) dup |= (tets[4 * e + w] == n);
            if (!dup && stamp[n] != t) fresh++;
        }
        if (t < 0 || ecount + 1 > max_elems || ncount + fresh > max_nodes) {
            t++;
            P.t_ebegin.push_back(i);
            P.t_ecount.push_back(0);
            ecount = 0;
            ncount = 0;
            fresh = 0;
            for (int v = 0; v < 4; v++) {
                const int n = tets[4 * e + v];
                bool dup = false;
                for (int w = 0; w < v; w++) dup |= (tets[4 * e + w] == n);
                if (!dup) fresh++;
            }
        }
        for (int v = 0; v < 4; v++) stamp[tets[4 * e + v]] = t;
        ecount++;
        ncount += fresh;
        P.t_ecount[t] = ecount;
    }
    const int T = (int)P.t_ebegin.size();

    // tile-local nodes, local tets, incidences
    std::vector<int> loc(n_nodes, -1), tiles_of(n_nodes, 0);
    P.ltet.resize(E);
    P.t_lnbegin.resize(T);
    P.t_lncount.resize(T);
    P.ln_inc_ptr.push_back(0);
        for (int ti = 0; ti < T; ti++) {
        const int e0 = P.t_ebegin[ti], ec = P.t_ecount[ti];
        std::vector<int> nodes;
        nodes.reserve(4 * ec);
        for (int k = 0; k < ec; k++)
            for (int v = 0; v < 4; v++) nodes.push_back(tets[4 * P.perm[e0 + k] + v]);
        std::sort(nodes.begin(), nodes.end());
        nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
        // incidence count of each node in this tile, then the nodes by descending count (ties
        // by node id): the per-node loops over incident elements run one node per thread, and
        // a wave of nodes of like counts runs about as many iterations as its mean, not its max
        std::vector<int> cnt(nodes.size(), 0);
        for (int k = 0; k < ec; k++)
            for (int v = 0; v < 4; v++) {
                const int n = tets[4 * P.perm[e0 + k] + v];
                cnt[std::lower_bound(nodes.begin(), nodes.end(), n) - nodes.begin()]++;
            }
        {
            std::vector<int> order(nodes.size());
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cnt[a] > cnt[b]; });
            std::vector<int> by_count(nodes.size());
            for (size_t j = 0; j < order.size(); j++) by_count[j] = nodes[order[j]];
            nodes.swap(by_count);
        }
        if ((int)nodes.size() > max_nodes) throw std::logic_error("tile node budget exceeded");
        P.t_lnbegin[ti] = (int)P.ln_gid.size();
        P.t_lncount[ti] = (int)nodes.size();
        for (size_t j = 0; j < nodes.size(); j++) {
            loc[nodes[j]] = (int)j;
            tiles_of[nodes[j]]++;
        }
        std::vector<std::vector<std::pair<int, int>>> per(nodes.size());
        for (int k = 0; k < ec; k++) {
            const int e = P.perm[e0 + k];
            uint32_t l[4];
            for (int v = 0; v < 4; v++) {
                l[v] = (uint32_t)loc[tets[4 * e + v]];
                per[l[v]].push_back({e, v * 3 * max_elems + k});
            }
            P.ltet[e0 + k] = make_uint2(l[0] | (l[1] << 16), l[2] | (l[3] << 16));
        }
        for (size_t j = 0; j < nodes.size(); j++) {
            auto& lst = per[j];
            std::sort(lst.begin(), lst.end());
            for (auto& pr : lst) P.ln_inc.push_back((uint16_t)pr.second);
            P.ln_inc_ptr.push_back((int)P.ln_inc.size());
            P.ln_gid.push_back(nodes[j]);
        }
    }
    // nodes of no element get tiles of their own (no elements, up to max_nodes nodes), where
    // they are finished like any node owned by one tile (the mass term only)
    {
        std::vector<int> free_nodes;
        for (int n = 0; n < n_nodes; n++)
            if (tiles_of[n] == 0) free_nodes.push_back(n);
        for (size_t j0 = 0; j0 < free_nodes.size(); j0 += (size_t)max_nodes) {
            const size_t j1 = std::min(free_nodes.size(), j0 + (size_t)max_nodes);
            P.t_ebegin.push_back(E);
            P.t_ecount.push_back(0);
            P.t_lnbegin.push_back((int)P.ln_gid.size());
            P.t_lncount.push_back((int)(j1 - j0));
            for (size_t j = j0; j < j1; j++) {
                P.ln_gid.push_back(free_nodes[j]);
                P.ln_inc_ptr.push_back((int)P.ln_inc.size());
                tiles_of[free_nodes[j]] = 1;
            }
        }
    }
    const int TT = (int)P.t_ebegin.size();
    // shared nodes (in >1 tile) go to the finalize kernel
    std::vector<int> shidx(n_nodes, -1), fill(n_nodes, 0);
    P.sh_ptr.push_back(0);
    for (int n = 0; n < n_nodes; n++) {
        if (tiles_of[n] > 1) {
            shidx[n] = (int)P.sh_gid.size();
            P.sh_gid.push_back(n);
            P.sh_ptr.push_back(P.sh_ptr.back() + tiles_of[n]);
        }
    }
    P.n_slots = P.sh_ptr.back();
    P.ln_pslot.assign(P.ln_gid.size(), -1);
    for (int ti = 0; ti < TT; ti++) {
        for (int j = 0; j < P.t_lncount[ti]; j++) {
            const int li = P.t_lnbegin[ti] + j;
            const int n = P.ln_gid[li];
            if (shidx[n] >= 0) P.ln_pslot[li] = P.sh_ptr[shidx[n]] + fill[n]++;
        }
    }
    return P;
}

// The reference's summation order for the nodes shared by several tiles.  A per-tile partial
// sum changes the association of a node's element contributions ((m + c1 + c2) + (c3 + c4)
// instead of m + c1 + c2 + c3 + c4), so here every incidence of a shared node gets a slot of its
// own.  A shared node's slots are consecutive, in ascending ORIGINAL element id, at
// sh_ptr[k] .. sh_ptr[k + 1]: the order in which the finalize kernel adds them after the mass
// term -- TetrahedronFEMForceField's element loop (TetrahedronFEMForceField.inl:1630-1634) -- so
// that it reads each node's contributions as one contiguous run.  A tile writes its element's
// contribution to the node at ln_pslot (the node's first slot) + inc_rank of the incidence.
// Interior nodes keep their in-tile sums, which are already in that order.
//
// The CG operator folds each shared node's leading run -- the longest prefix of its list that
// lies in one tile -- in that tile: the tile starts the node's sum (the mass term, then those
// contributions one by one) and writes it to the node's first slot; the rest of the list takes
// the following slots, one per incidence.  The finalize kernel continues the sum from the first
// slot: the same operations in the same order.
static inline void plan_exact(TilePlan& P, int n_nodes, int max_elems) {
    const int TT = (int)P.t_ebegin.size();
    std::vector<int> shidx(n_nodes, -1);
    for (size_t j = 0; j < P.sh_gid.size(); j++) shidx[P.sh_gid[j]] = (int)j;
    struct Inc { int elem, c, tile, li; bool operator<(const Inc& o) const { return elem < o.elem || (elem == o.elem && c < o.c); } };
    std::vector<std::vector<Inc>> lists(P.sh_gid.size());   // (original element, incidence, tile, local node)
    for (int ti = 0; ti < TT; ti++) {
        for (int j = 0; j < P.t_lncount[ti]; j++) {
            const int li = P.t_lnbegin[ti] + j;
            const int k = shidx[P.ln_gid[li]];
            if (k < 0) continue;
            for (int c = P.ln_inc_ptr[li]; c < P.ln_inc_ptr[li + 1]; c++) {
                const int local_elem = (int)P.ln_inc[c] % max_elems;   // slot = 3 max_elems v + element
                lists[k].push_back({P.perm[P.t_ebegin[ti] + local_elem], c, ti, li});
            }
        }
    }
    P.inc_rank.assign(P.ln_inc.size(), 0);
    P.inc_arank.assign(P.ln_inc.size(), 0);
    P.sh_acnt.assign(P.sh_gid.size(), 0);
    P.ln_lead.assign(P.ln_gid.size(), 0);
    P.sh_ptr.assign(1, 0);
    for (size_t k = 0; k < lists.size(); k++) {
        auto& lst = lists[k];
        if (lst.size() > 256) throw std::invalid_argument("a node with more than 256 tetrahedra");
        std::sort(lst.begin(), lst.end());
        size_t lead = 1;
        while (lead < lst.size() && lst[lead].tile == lst[0].tile) lead++;
        for (size_t r = 0; r < lst.size(); r++) {
            P.inc_rank[lst[r].c] = (uint8_t)r;
            P.inc_arank[lst[r].c] = (uint8_t)(r < lead ? 0 : r - lead + 1);
        }
        P.sh_acnt[k] = (uint8_t)(lst.size() - lead);
        P.ln_lead[lst[0].li] = (uint8_t)lead;
        P.sh_ptr.push_back(P.sh_ptr.back() + (int)lst.size());
    }
    for (int ti = 0; ti < TT; ti++)
        for (int j = 0; j < P.t_lncount[ti]; j++) {
            const int li = P.t_lnbegin[ti] + j;
            const int k = shidx[P.ln_gid[li]];
            if (k >= 0) P.ln_pslot[li] = P.sh_ptr[k];
        }
    P.n_slots = P.sh_ptr.back();
}

// device bytes one operator launch moves for the plan's replication (every tile-local node reads
// its id, slot, incidence pointer and mask (13 B) and r and p (6 reals); every partial slot is
// written by a tile and read by the finalize kernel (6 reals))
static inline double plan_cost(const TilePlan& P, int real_bytes) {
    return (double)P.ln_gid.size() * (13.0 + 6.0 * real_bytes) + (double)P.n_slots * 6.0 * real_bytes;
}

// the cheaper of the two element orders for this mesh (exact: the reference-order slot layout
// of plan_exact, one slot per incidence of a shared node); ties keep the Morton order
static inline TilePlan plan_best(int n_nodes, const double* rest, const int32_t* tets, int E, int max_elems,
                                 int max_nodes, bool exact, int real_bytes) {
    TilePlan a = plan_tiles(n_nodes, rest, tets, E, max_elems, max_nodes, kOrderMorton);
    TilePlan b = plan_tiles(n_nodes, rest, tets, E, max_elems, max_nodes, kOrderElementId);
    if (exact) {
        plan_exact(a, n_nodes, max_elems);
        plan_exact(b, n_nodes, max_elems);
    }
    return plan_cost(b, real_bytes) < plan_cost(a, real_bytes) ? std::move(b) : std::move(a);
}

}  // namespace sofa_amd
