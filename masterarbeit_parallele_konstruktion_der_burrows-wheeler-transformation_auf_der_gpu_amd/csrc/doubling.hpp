// Group refinement ("prefix doubling") shared by the BWT pipelines.
//
// State: elements V[0..m) in an order in which equal keys K form contiguous groups, each
// group occupying the global sorted positions G[...] of its members.  A round
//   1. assigns rank[x] = 1 + global position of the group head and SA[G[i]] = V[i],
//   2. keeps the groups that can still split (>= 2 members and Policy::open),
//   3. orders each kept group by the rank of Policy::next(x) (0 when there is none):
//      groups of up to kSmallGroup members in place (every member counts the members ordered
//      before it), larger ones by one radix sort of (group, rank) keys.
// The loop stops when no group can split.  Policies:
//   - strings of a family (PFP dictionary / parse): next(x) = x + h inside the segment,
//     open while the compared prefix length h has not run past the strings (then equal
//     strings stay in one group); h doubles every round;
//   - chains (Lyndon): next(x) = jump pointer, always open; pointers are squared per round.
#pragma once

#include "prims.hpp"

namespace mibwt {

constexpr uint32_t kNoNext = 0xFFFFFFFFu;
constexpr uint32_t kSmallGroup = 256;

static __global__ void k_dbl_headflag(const uint64_t* __restrict__ K, uint32_t m, uint32_t* __restrict__ hf)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    hf[i] = (i == 0 || K[i] != K[i - 1]) ? i : 0;
}

template <typename Policy>
__global__ void k_dbl_update(const uint64_t* __restrict__ K, const uint32_t* __restrict__ V,
                             const uint32_t* __restrict__ G, const uint32_t* __restrict__ hidx,
                             uint32_t m, uint32_t* __restrict__ rank, uint32_t* __restrict__ SA,
                             uint32_t* __restrict__ keep, Policy pol)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t x = V[i];
    const uint32_t hi = hidx[i];
    rank[x] = G[hi] + 1;
    SA[G[i]] = x;
    const bool single = (hi == i) && (i + 1 == m || K[i + 1] != K[i]);
    keep[i] = (!single && pol.open(V[hi])) ? 1u : 0u;
}

static __global__ void k_dbl_compact(const uint32_t* __restrict__ V, const uint32_t* __restrict__ G,
                                     const uint32_t* __restrict__ keep, const uint32_t* __restrict__ pos,
                                     uint32_t m, uint32_t* __restrict__ V2, uint32_t* __restrict__ G2)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    if (keep[i]) {
        V2[pos[i]] = V[i];
        G2[pos[i]] = G[i];
    }
}

static __global__ void k_iota(uint32_t* __restrict__ p, uint32_t m)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) p[i] = i;
}

// group heads of the compacted active array: members of one group share their rank
static __global__ void k_grp_heads(const uint32_t* __restrict__ V, const uint32_t* __restrict__ rank, uint32_t m,
                                   uint32_t* __restrict__ hf)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    hf[i] = (i == 0 || rank[V[i]] != rank[V[i - 1]]) ? 1u : 0u;
}

// gofs[g] = first index of group g (0-based), gofs[ngroups] = m
static __global__ void k_grp_offsets(const uint32_t* __restrict__ hf, const uint32_t* __restrict__ gid1, uint32_t m,
                                     uint32_t* __restrict__ gofs)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    if (hf[i]) gofs[gid1[i] - 1] = i;
    if (i == m - 1) gofs[gid1[i]] = m;
}

template <typename Policy>
__global__ void k_key2(const uint32_t* __restrict__ V, uint32_t m, const uint32_t* __restrict__ rank,
                       Policy pol, uint32_t* __restrict__ key2)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t y = pol.next(V[i]);
    key2[i] = (y == kNoNext) ? 0u : rank[y];
}

// Small groups: each member counts the members ordered before it (key, then index) and moves
// to that slot of its group; members of large groups are only flagged.
static __global__ void k_seg_small(const uint32_t* __restrict__ V, const uint32_t* __restrict__ key2,
                                   const uint32_t* __restrict__ gid1, const uint32_t* __restrict__ gofs,
                                   uint32_t m, uint32_t small_max, uint32_t* __restrict__ V2,
                                   uint64_t* __restrict__ K2, uint32_t* __restrict__ large)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t g = gid1[i] - 1;
    const uint32_t s = gofs[g], e = gofs[g + 1];
    if (e - s > small_max) {
        large[i] = 1;
        return;
    }
    large[i] = 0;
    const uint32_t k = key2[i];
    uint32_t r = 0;
    for (uint32_t j = s; j < e; j++) {
        const uint32_t kj = key2[j];
        r += (kj < k) || (kj == k && j < i);
    }
    V2[s + r] = V[i];
    K2[s + r] = ((uint64_t)(g + 1) << 32) | k;
}

static __global__ void k_seg_large_gather(const uint32_t* __restrict__ V, const uint32_t* __restrict__ key2,
                                          const uint32_t* __restrict__ gid1, const uint32_t* __restrict__ large,
                                          const uint32_t* __restrict__ lpos, uint32_t m, int B,
                                          uint64_t* __restrict__ LK, uint32_t* __restrict__ LV,
                                          uint32_t* __restrict__ LI)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || !large[i]) return;
    const uint32_t p = lpos[i];
    LK[p] = ((uint64_t)gid1[i] << B) | key2[i];
    LV[p] = V[i];
    LI[p] = i;
}

static __global__ void k_seg_large_scatter(const uint64_t* __restrict__ LK, const uint32_t* __restrict__ LV,
                                           const uint32_t* __restrict__ LI, uint32_t cnt, int B,
                                           uint32_t* __restrict__ V2, uint64_t* __restrict__ K2)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cnt) return;
    const uint32_t i = LI[j];
    const uint64_t k = LK[j];
    V2[i] = LV[j];
    K2[i] = ((k >> B) << 32) | (k & ((1ull << B) - 1));
}

struct DoublingResult {
    int rounds = 0;
    double bytes = 0;
    uint64_t large_items = 0;
};

// Refine the groups of an initial order.  K0/V0 hold the initial keys/elements of N items,
// already sorted (equal keys = one group); K1/V1 are scratch of N items.  `advance(pol, Vact,
// m)` is called after each round's secondary keys were taken, with the still-active elements.
template <typename Policy, typename Advance>
DoublingResult refine_groups(Storage& S, uint64_t* K0, uint32_t* V0, uint32_t* V1, uint32_t N, Policy& pol,
                             Advance advance, uint32_t* rank, uint32_t* SA)
{
    DoublingResult res;
    uint32_t* G0 = S.buf<uint32_t>(B_ACT_POS, N);
    uint32_t* G1 = S.buf<uint32_t>(B_ACT_POS2, N);
    uint32_t* hidx = S.buf<uint32_t>(B_ACT_HEAD, N);
    uint32_t* keep = S.buf<uint32_t>(B_ACT_FLAG, N);
    uint32_t* pos = S.buf<uint32_t>(B_ACT_X, N);
    uint32_t* Vc = S.buf<uint32_t>(B_ACT_X2, N);
    uint32_t* key2 = S.buf<uint32_t>(B_DB_KEY2, N);
    uint32_t* gid1 = S.buf<uint32_t>(B_DB_GID, N);
    uint32_t* gofs = S.buf<uint32_t>(B_DB_GOFS, N + 1);
    uint32_t* large = S.buf<uint32_t>(B_DB_LARGE, N);
    uint32_t* sc = S.buf<uint32_t>(B_SCALARS, SC_COUNT_);

    const int B = bits_for((uint64_t)N + 1);
    uint32_t m = N;
    k_iota<<<grid_for(m), kBlock, 0, S.stream>>>(G0, m);
    MIBWT_LAUNCH_CHECK();
    uint64_t* K = K0;
    uint32_t* V = V0;
    for (;;) {
        // ---- ranks and SA slots of the current order; which groups stay active
        k_dbl_headflag<<<grid_for(m), kBlock, 0, S.stream>>>(K, m, hidx);
        MIBWT_LAUNCH_CHECK();
        inclusive_max<uint32_t>(S, hidx, hidx, m);
        k_dbl_update<Policy><<<grid_for(m), kBlock, 0, S.stream>>>(K, V, G0, hidx, m, rank, SA, keep, pol);
        MIBWT_LAUNCH_CHECK();
        exclusive_sum<uint32_t>(S, keep, pos, m, sc + SC_ACTIVE);
        uint32_t m2 = 0;
        S.read(&m2, sc + SC_ACTIVE);
        res.rounds++;
        res.bytes += (double)m * (8 * 2 + 4 * 8);
        if (m2 == 0) break;
        k_dbl_compact<<<grid_for(m), kBlock, 0, S.stream>>>(V, G0, keep, pos, m, Vc, G1);
        MIBWT_LAUNCH_CHECK();
        std::swap(G0, G1);
        m = m2;
        // ---- group structure of the active elements
        k_grp_heads<<<grid_for(m), kBlock, 0, S.stream>>>(Vc, rank, m, keep);
        MIBWT_LAUNCH_CHECK();
        inclusive_sum<uint32_t>(S, keep, gid1, m, sc + SC_GROUPS);
        k_grp_offsets<<<grid_for(m), kBlock, 0, S.stream>>>(keep, gid1, m, gofs);
        MIBWT_LAUNCH_CHECK();
        // ---- order each group by the rank of the next position
        k_key2<Policy><<<grid_for(m), kBlock, 0, S.stream>>>(Vc, m, rank, pol, key2);
        MIBWT_LAUNCH_CHECK();
        advance(pol, Vc, m);
        k_seg_small<<<grid_for(m), kBlock, 0, S.stream>>>(Vc, key2, gid1, gofs, m, kSmallGroup, V1, K0, large);
        MIBWT_LAUNCH_CHECK();
        exclusive_sum<uint32_t>(S, large, pos, m, sc + SC_LARGE);
        uint32_t hv[SC_COUNT_];
        S.read(hv, sc, SC_COUNT_);
        const uint32_t nl = hv[SC_LARGE];
        res.bytes += (double)m * (4 * 9 + 8);
        if (nl > 0) {
            uint64_t* LK0 = S.buf<uint64_t>(B_LG_K0, nl);
            uint64_t* LK1 = S.buf<uint64_t>(B_LG_K1, nl);
            uint32_t* LV0 = S.buf<uint32_t>(B_LG_V0, nl);
            uint32_t* LV1 = S.buf<uint32_t>(B_LG_V1, nl);
            uint32_t* LI = S.buf<uint32_t>(B_LG_I, nl);
            k_seg_large_gather<<<grid_for(m), kBlock, 0, S.stream>>>(Vc, key2, gid1, large, pos, m, B, LK0, LV0, LI);
            MIBWT_LAUNCH_CHECK();
            const int kb = B + bits_for(hv[SC_GROUPS]);
            radix_sort<uint64_t, uint32_t>(S, LK0, LK1, LV0, LV1, nl, 0, kb);
            k_seg_large_scatter<<<grid_for(nl), kBlock, 0, S.stream>>>(LK0, LV0, LI, nl, B, V1, K0);
            MIBWT_LAUNCH_CHECK();
            res.large_items += nl;
            res.bytes += (double)nl * (12 * 2 * ((kb + 7) / 8) + 32);
        }
        // the compaction reads V and writes Vc, the refinement reads Vc and writes V1
        K = K0;
        V = V1;
    }
    return res;
}

}  // namespace mibwt
