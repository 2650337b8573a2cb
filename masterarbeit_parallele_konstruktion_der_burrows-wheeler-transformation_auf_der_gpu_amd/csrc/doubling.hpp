// Group refinement ("prefix doubling") shared by the BWT pipelines.
//
// State: the active elements V[0..m) in an order in which the members of a group are
// contiguous; per element the group record (first slot Hs, global start Gs, size Gz).  A round
//   1. takes the secondary key of every member: the rank of Policy::next(x) (0 if none);
//   2. orders each group by that key: groups of up to kSmallGroup members in place (every
//      member counts the members ordered before it), larger ones by one radix sort of
//      (group start, key) pairs; the same kernels write the new rank (1 + global position of
//      the sub-group's first member), the SA slot and the sub-group record of every member;
//   3. keeps the sub-groups that can still split (>= 2 members and Policy::open) and compacts
//      them, which also yields next round's counts in one host read-back.
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

// ---- initial groups: runs of equal keys in the initially sorted order

static __global__ void k_dbl_headflag(const uint64_t* __restrict__ K, uint32_t m, uint32_t* __restrict__ hf)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    hf[i] = (i == 0 || K[i] != K[i - 1]) ? 1u : 0u;
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

// Element i of the initial order: rank and SA slot, whether its group stays active, and the
// group size (the first compaction reads the rest of the group record from gid1 / gofs).
template <typename Policy>
__global__ void k_dbl_init(const uint32_t* __restrict__ V, const uint32_t* __restrict__ gid1,
                           const uint32_t* __restrict__ gofs, uint32_t m, uint32_t* __restrict__ rank,
                           uint32_t* __restrict__ SA, uint32_t* __restrict__ keep, uint32_t* __restrict__ NGz,
                           Policy pol)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t g = gid1[i] - 1;
    const uint32_t s = gofs[g], z = gofs[g + 1] - s;
    const uint32_t x = V[i];
    rank[x] = s + 1;
    SA[i] = x;
    keep[i] = (z > 1 && pol.open(x)) ? 1u : 0u;
    NGz[i] = z;
}

// first compaction: the group record of a kept element from the initial groups
static __global__ void k_dbl_compact_init(const uint32_t* __restrict__ keep, const uint32_t* __restrict__ pos,
                                          const uint32_t* __restrict__ V0, const uint32_t* __restrict__ gid1,
                                          const uint32_t* __restrict__ gofs, const uint32_t* __restrict__ NGz,
                                          uint32_t m, uint32_t* __restrict__ V, uint32_t* __restrict__ Hs,
                                          uint32_t* __restrict__ Gs, uint32_t* __restrict__ Gz)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m || !keep[p]) return;
    const uint32_t o = pos[p];
    const uint32_t s = gofs[gid1[p] - 1];
    V[o] = V0[p];
    Hs[o] = pos[s];
    Gs[o] = s;
    Gz[o] = NGz[p];
}

// ---- one refinement round over the m active elements (groups contiguous)

template <typename Policy>
__global__ void k_key2(const uint32_t* __restrict__ V, uint32_t m, const uint32_t* __restrict__ rank,
                       Policy pol, uint32_t* __restrict__ key2)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t y = pol.next(V[i]);
    key2[i] = (y == kNoNext) ? 0u : rank[y];
}

// Small groups: each member counts the members ordered before it (key, then slot), moves to
// that slot and records its new sub-group (members with an equal key): rank = global start of
// the sub-group + 1, SA slot, and the sub-group record for the compaction.  A small group that
// has a member in the block lies inside [b0 - kSmallGroup, b0 + kBlock + kSmallGroup), so the
// keys of that window are staged in LDS and every count reads LDS (members of one group read
// the same words: broadcasts).  Members of large groups are only flagged.
static_assert(kSmallGroup <= kBlock, "the LDS key window assumes small groups of at most one block");
template <typename Policy>
__global__ __launch_bounds__(kBlock) void k_seg_small(const uint32_t* __restrict__ V, const uint32_t* __restrict__ Hs,
                                                     const uint32_t* __restrict__ Gs, const uint32_t* __restrict__ Gz,
                                                     const uint32_t* __restrict__ key2, uint32_t m,
                                                     uint32_t small_max, uint32_t* __restrict__ rank,
                                                     uint32_t* __restrict__ SA, uint32_t* __restrict__ keep,
                                                     uint32_t* __restrict__ NV, uint32_t* __restrict__ NHs,
                                                     uint32_t* __restrict__ NGs, uint32_t* __restrict__ NGz,
                                                     uint32_t* __restrict__ large, Policy pol)
{
    constexpr uint32_t kWin = kBlock + 2 * kSmallGroup;
    __shared__ uint32_t lk[kWin];
    const uint32_t b0 = blockIdx.x * kBlock;
    const int64_t w0 = (int64_t)b0 - kSmallGroup;  // slot of lk[0]
    for (uint32_t t = threadIdx.x; t < kWin; t += kBlock) {
        const int64_t j = w0 + t;
        lk[t] = (j >= 0 && j < (int64_t)m) ? key2[j] : 0u;
    }
    __syncthreads();
    const uint32_t i = b0 + threadIdx.x;
    if (i >= m) return;
    const uint32_t z = Gz[i];
    if (z > small_max || z > kSmallGroup) {
        large[i] = 1;
        return;
    }
    large[i] = 0;
    const uint32_t s = Hs[i];
    const uint32_t k = lk[i - w0];
    const uint32_t ls = (uint32_t)((int64_t)s - w0), li = (uint32_t)((int64_t)i - w0);
    // members before this one and from it on, each counting keys < k and <= k (two compares and
    // two carry-adds per key): lt = all smaller keys, eq_before = equal keys before it
    uint32_t lt1 = 0, le1 = 0, lt2 = 0, le2 = 0;
    for (uint32_t j = ls; j < li; j++) {
        const uint32_t kj = lk[j];
        lt1 += (kj < k) ? 1u : 0u;
        le1 += (kj <= k) ? 1u : 0u;
    }
    for (uint32_t j = li; j < ls + z; j++) {
        const uint32_t kj = lk[j];
        lt2 += (kj < k) ? 1u : 0u;
        le2 += (kj <= k) ? 1u : 0u;
    }
    const uint32_t lt = lt1 + lt2, eq_before = le1 - lt1, eq = eq_before + (le2 - lt2);
    const uint32_t x = V[i];
    const uint32_t gs = Gs[i];
    const uint32_t p = s + lt + eq_before;
    rank[x] = gs + lt + 1;
    SA[gs + lt + eq_before] = x;
    keep[p] = (eq > 1 && pol.open(x)) ? 1u : 0u;
    NV[p] = x;
    NHs[p] = s + lt;
    NGs[p] = gs + lt;
    NGz[p] = eq;
}

// Large groups: (group start slot, key) pairs of the flagged members in slot order.
static __global__ void k_seg_large_gather(const uint32_t* __restrict__ V, const uint32_t* __restrict__ Hs,
                                          const uint32_t* __restrict__ key2, const uint32_t* __restrict__ large,
                                          const uint32_t* __restrict__ lpos, uint32_t m, int B,
                                          uint64_t* __restrict__ LK, uint32_t* __restrict__ LV)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || !large[i]) return;
    const uint32_t q = lpos[i];
    LK[q] = ((uint64_t)Hs[i] << B) | key2[i];
    LV[q] = V[i];
}

static __global__ void k_lg_heads(const uint64_t* __restrict__ LK, uint32_t nl, uint32_t* __restrict__ hq)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nl) return;
    hq[q] = (q == 0 || LK[q] != LK[q - 1]) ? q : 0u;
}

// size of each sub-group, stored at its first index
static __global__ void k_lg_sizes(const uint64_t* __restrict__ LK, const uint32_t* __restrict__ qh, uint32_t nl,
                                  uint32_t* __restrict__ sz)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nl) return;
    if (q + 1 == nl || LK[q] != LK[q + 1]) sz[qh[q]] = q - qh[q] + 1;
}

template <typename Policy>
__global__ void k_lg_update(const uint64_t* __restrict__ LK, const uint32_t* __restrict__ LV,
                            const uint32_t* __restrict__ qh, const uint32_t* __restrict__ sz,
                            const uint32_t* __restrict__ lpos, const uint32_t* __restrict__ Gs, uint32_t nl, int B,
                            uint32_t* __restrict__ rank, uint32_t* __restrict__ SA, uint32_t* __restrict__ keep,
                            uint32_t* __restrict__ NV, uint32_t* __restrict__ NHs, uint32_t* __restrict__ NGs,
                            uint32_t* __restrict__ NGz, Policy pol)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nl) return;
    const uint32_t x = LV[q];
    const uint32_t s = (uint32_t)(LK[q] >> B);  // group start slot (a member of the group)
    const uint32_t q0 = lpos[s];                // first index of the group among the large ones
    const uint32_t lt = qh[q] - q0, c = sz[qh[q]];
    const uint32_t gs = Gs[s];
    const uint32_t p = s + (q - q0);
    rank[x] = gs + lt + 1;
    SA[gs + (q - q0)] = x;
    keep[p] = (c > 1 && pol.open(x)) ? 1u : 0u;
    NV[p] = x;
    NHs[p] = s + lt;
    NGs[p] = gs + lt;
    NGz[p] = c;
}

// Compaction positions of the kept members: pos = exclusive scan of keep (chained tiles, the
// total into *active); per tile the number of kept members whose sub-group is large next round
// (lpart, summed by k_sum_tiles).
static __global__ __launch_bounds__(kBlock) void k_dbl_keep_scan(const uint32_t* __restrict__ keep,
                                                                const uint32_t* __restrict__ NGz, uint32_t m,
                                                                uint32_t small_max, uint32_t* __restrict__ pos,
                                                                uint32_t* __restrict__ active,
                                                                uint32_t* __restrict__ lpart, Chain c, uint32_t tiles)
{
    __shared__ uint32_t slot;
    __shared__ uint32_t lds[kWaves];
    const uint32_t tile = chain_ticket(c, tiles, &slot);
    const uint64_t base = (uint64_t)tile * kChainTile + (uint64_t)threadIdx.x * kChainItems;
    uint32_t v[kChainItems];
    uint32_t nl = 0;
    if (base + kChainItems <= m) {
#pragma unroll
        for (int q = 0; q < kChainItems / 4; q++) {
            const uint4 a = *reinterpret_cast<const uint4*>(keep + base + 4 * q);
            const uint4 z = *reinterpret_cast<const uint4*>(NGz + base + 4 * q);
            v[4 * q] = a.x;
            v[4 * q + 1] = a.y;
            v[4 * q + 2] = a.z;
            v[4 * q + 3] = a.w;
            nl += (a.x && z.x > small_max) + (a.y && z.y > small_max) + (a.z && z.z > small_max) +
                  (a.w && z.w > small_max);
        }
    } else {
#pragma unroll
        for (int i = 0; i < kChainItems; i++) {
            const bool in = base + i < m;
            v[i] = in ? keep[base + i] : 0u;
            nl += (v[i] && NGz[base + i] > small_max) ? 1u : 0u;
        }
    }
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < kChainItems; i++) acc += v[i];
    uint32_t tot;
    uint32_t run = block_exclusive_scan(acc, 0u, OpSum{}, lds, &tot);
    const uint32_t pre = chain_prefix(c, tile, tot, OpSum{}, 0u, &slot);
    run += pre;
    if (base + kChainItems <= m) {
#pragma unroll
        for (int q = 0; q < kChainItems / 4; q++) {
            uint32_t o[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                o[j] = run;
                run += v[4 * q + j];
            }
            *reinterpret_cast<uint4*>(pos + base + 4 * q) = make_uint4(o[0], o[1], o[2], o[3]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < kChainItems; i++) {
            if (base + i < m) pos[base + i] = run;
            run += v[i];
        }
    }
    uint32_t ltot;
    (void)block_exclusive_scan(nl, 0u, OpSum{}, lds, &ltot);
    if (threadIdx.x == 0) {
        lpart[tile] = ltot;
        if (tile == tiles - 1) *active = pre + tot;
    }
}

static __global__ __launch_bounds__(kBlock) void k_sum_tiles(const uint32_t* __restrict__ part, uint32_t tiles,
                                                            uint32_t* __restrict__ out)
{
    __shared__ uint32_t lds[kWaves];
    uint32_t acc = 0;
    for (uint32_t t = threadIdx.x; t < tiles; t += kBlock) acc += part[t];
    uint32_t tot;
    (void)block_exclusive_scan(acc, 0u, OpSum{}, lds, &tot);
    if (threadIdx.x == 0) *out = tot;
}

// compaction of the kept members (whole sub-groups are kept or dropped, so a sub-group's
// first slot is kept with it and maps through pos as well)
static __global__ void k_dbl_compact(const uint32_t* __restrict__ keep, const uint32_t* __restrict__ pos,
                                     const uint32_t* __restrict__ NV, const uint32_t* __restrict__ NHs,
                                     const uint32_t* __restrict__ NGs, const uint32_t* __restrict__ NGz, uint32_t m,
                                     uint32_t* __restrict__ V, uint32_t* __restrict__ Hs, uint32_t* __restrict__ Gs,
                                     uint32_t* __restrict__ Gz)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m || !keep[p]) return;
    const uint32_t o = pos[p];
    V[o] = NV[p];
    Hs[o] = pos[NHs[p]];
    Gs[o] = NGs[p];
    Gz[o] = NGz[p];
}

struct DoublingResult {
    int rounds = 0;
    double bytes = 0;
    uint64_t large_items = 0;
};

// Refine the groups of an initial order.  K0/V0 hold the initial keys/elements of N items,
// already sorted (equal keys = one group); K1/V1 are scratch of N items.  `advance(pol, Vact,
// m)` is called after each round's secondary keys were taken, with the still-active elements.
// On return rank[x] = 1 + global position of x's group head and SA is the sorted order.
//
// Per round: secondary keys, one kernel that sorts every small group in place and directly
// writes ranks, SA slots and the sub-group records, a radix sort for large groups, then one
// chained scan of the kept flags (which also counts next round's large-group members) and the
// compaction; one host read-back per round (active and large counts).
template <typename Policy, typename Advance>
DoublingResult refine_groups(Storage& S, uint64_t* K0, uint32_t* V0, uint32_t* V1, uint32_t N, Policy& pol,
                             Advance advance, uint32_t* rank, uint32_t* SA)
{
    DoublingResult res;
    uint32_t* V = S.buf<uint32_t>(B_ACT_X2, N);
    uint32_t* Hs = S.buf<uint32_t>(B_ACT_POS, N);
    uint32_t* Gs = S.buf<uint32_t>(B_ACT_POS2, N);
    uint32_t* Gz = S.buf<uint32_t>(B_ACT_HEAD, N);
    uint32_t* NV = V1;
    uint32_t* NHs = S.buf<uint32_t>(B_DB_GID, N);
    uint32_t* NGs = S.buf<uint32_t>(B_DB_GOFS, N + 1);
    uint32_t* NGz = S.buf<uint32_t>(B_DB_NGZ, N);
    uint32_t* keep = S.buf<uint32_t>(B_ACT_FLAG, N);
    uint32_t* pos = S.buf<uint32_t>(B_ACT_X, N);
    uint32_t* key2 = S.buf<uint32_t>(B_DB_KEY2, N);
    uint32_t* large = S.buf<uint32_t>(B_DB_LARGE, N);
    uint32_t* sc = S.buf<uint32_t>(B_SCALARS, SC_COUNT_);
    const int B = bits_for((uint64_t)N + 1);

    // initial groups: runs of equal keys (gid1 lives in `large` until the first compaction)
    uint32_t* gid1 = large;
    uint32_t* gofs = S.buf<uint32_t>(B_DB_GOFS0, N + 1);
    {
        uint32_t* hf = keep;
        k_dbl_headflag<<<grid_for(N), kBlock, 0, S.stream>>>(K0, N, hf);
        MIBWT_LAUNCH_CHECK();
        inclusive_sum<uint32_t>(S, hf, gid1, N);
        k_grp_offsets<<<grid_for(N), kBlock, 0, S.stream>>>(hf, gid1, N, gofs);
        MIBWT_LAUNCH_CHECK();
        k_dbl_init<Policy><<<grid_for(N), kBlock, 0, S.stream>>>(V0, gid1, gofs, N, rank, SA, keep, NGz, pol);
        MIBWT_LAUNCH_CHECK();
        res.bytes += (double)N * (8 * 2 + 4 * 9);
    }
    bool first = true;
    uint32_t m = N;
    for (;;) {
        // ---- compaction of the kept sub-groups; next round's counts
        {
            const uint32_t tiles = std::max<uint32_t>(1, ceil_div(m, (uint32_t)kChainTile));
            uint32_t* lpart = S.buf<uint32_t>(B_DB_LPART, tiles);
            const Chain c = S.chain(tiles);
            k_dbl_keep_scan<<<tiles, kBlock, 0, S.stream>>>(keep, NGz, m, kSmallGroup, pos, sc + SC_ACTIVE, lpart, c,
                                                            tiles);
            MIBWT_LAUNCH_CHECK();
            k_sum_tiles<<<1, kBlock, 0, S.stream>>>(lpart, tiles, sc + SC_LARGE);
            MIBWT_LAUNCH_CHECK();
        }
        uint32_t hv[SC_COUNT_];
        S.read(hv, sc, SC_COUNT_);
        const uint32_t m2 = hv[SC_ACTIVE], nl = hv[SC_LARGE];
        res.rounds++;
        if (m2 == 0) break;
        if (first) {
            k_dbl_compact_init<<<grid_for(m), kBlock, 0, S.stream>>>(keep, pos, V0, gid1, gofs, NGz, m, V, Hs, Gs, Gz);
            first = false;
        } else {
            k_dbl_compact<<<grid_for(m), kBlock, 0, S.stream>>>(keep, pos, NV, NHs, NGs, NGz, m, V, Hs, Gs, Gz);
        }
        MIBWT_LAUNCH_CHECK();
        res.bytes += (double)m * (4 * 12) + (double)m2 * 16;
        m = m2;
        // ---- order each group by the rank of the next position
        k_key2<Policy><<<grid_for(m), kBlock, 0, S.stream>>>(V, m, rank, pol, key2);
        MIBWT_LAUNCH_CHECK();
        advance(pol, V, m);
        k_seg_small<Policy><<<grid_for(m), kBlock, 0, S.stream>>>(V, Hs, Gs, Gz, key2, m, kSmallGroup, rank, SA, keep,
                                                                  NV, NHs, NGs, NGz, large, pol);
        MIBWT_LAUNCH_CHECK();
        res.bytes += (double)m * (4 * 12);
        if (nl > 0) {
            uint32_t* lpos = pos;
            exclusive_sum<uint32_t>(S, large, lpos, m);
            uint64_t* LK0 = S.buf<uint64_t>(B_LG_K0, nl);
            uint64_t* LK1 = S.buf<uint64_t>(B_LG_K1, nl);
            uint32_t* LV0 = S.buf<uint32_t>(B_LG_V0, nl);
            uint32_t* LV1 = S.buf<uint32_t>(B_LG_V1, nl);
            uint32_t* qh = S.buf<uint32_t>(B_LG_I, nl);
            uint32_t* sz = S.buf<uint32_t>(B_LG_SZ, nl);
            k_seg_large_gather<<<grid_for(m), kBlock, 0, S.stream>>>(V, Hs, key2, large, lpos, m, B, LK0, LV0);
            MIBWT_LAUNCH_CHECK();
            const int kb = B + bits_for(m);
            radix_sort<uint64_t, uint32_t>(S, LK0, LK1, LV0, LV1, nl, 0, kb);
            k_lg_heads<<<grid_for(nl), kBlock, 0, S.stream>>>(LK0, nl, qh);
            MIBWT_LAUNCH_CHECK();
            inclusive_max<uint32_t>(S, qh, qh, nl);
            k_lg_sizes<<<grid_for(nl), kBlock, 0, S.stream>>>(LK0, qh, nl, sz);
            MIBWT_LAUNCH_CHECK();
            k_lg_update<Policy><<<grid_for(nl), kBlock, 0, S.stream>>>(LK0, LV0, qh, sz, lpos, Gs, nl, B, rank, SA,
                                                                       keep, NV, NHs, NGs, NGz, pol);
            MIBWT_LAUNCH_CHECK();
            res.large_items += nl;
            res.bytes += (double)nl * (12 * 2 * ((kb + 7) / 8) + 48);
        }
    }
    return res;
}

}  // namespace mibwt
