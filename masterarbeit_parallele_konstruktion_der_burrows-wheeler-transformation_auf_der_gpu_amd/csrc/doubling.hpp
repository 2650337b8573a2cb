// Group refinement ("prefix doubling") shared by the BWT pipelines.
//
// State: the active elements V[0..m) in an order in which the members of a group are
// contiguous; per element the group record (first slot Hs, global start Gs, size Gz).  A round
//   1. takes the secondary key of every member: the rank of Policy::next(x) (0 if none);
//   2. orders each group by that key: groups of up to kSmallGroup members in place (every
//      member counts the members ordered before it), larger ones by one radix sort of
//      (group start, key) pairs; the same kernels write the new rank (1 + global position of
//      the sub-group's first member; every member of a group holds 1 + the group's start, so
//      the members of its first sub-group keep their rank and are not written), the SA slot
//      and the sub-group record of every member;
//   3. keeps the sub-groups that can still split (>= 2 members and Policy::open) and compacts
//      them, which also yields next round's counts in one host read-back.
// The loop stops when no group can split.  A policy gives next(x, r) and open(x, r), where r
// is a per-element value that travels with the element through the rounds when
// Policy::kCarry (taken once from rem(x) at the start, so that no round gathers it):
//   - strings of a family (PFP dictionary / parse): next(x) = x + h inside the segment,
//     open while the compared prefix length h has not run past the strings (then equal
//     strings stay in one group); h doubles every round; for the dictionary r is the length
//     of the rest of the phrase;
//   - chains (Lyndon): next(x) = jump pointer, always open; pointers are squared per round.
// An element is placed once, in the round in which its group stops (kept elements are placed
// again later): pol.place(SA, slot, x, group head) writes SA[slot] = x, or whatever the
// caller wants in sorted order instead.
#pragma once

#include "prims.hpp"

namespace mibwt {

constexpr uint32_t kNoNext = 0xFFFFFFFFu;

// default placement: the suffix array
struct PlaceSA {
    __device__ __forceinline__ void operator()(uint32_t* SA, uint32_t pos, uint32_t x, uint32_t) const
    {
        SA[pos] = x;
    }
};
constexpr uint32_t kSmallGroup = 256;

// A policy may define first(SA, slot, x, group start, group size, r&): the start-of-refinement
// work of one element (its carried value r, its placement when it is not kept; returns whether
// it is kept), for policies whose carried value and placement come from one read.
template <typename P, typename = void>
struct HasFirst : std::false_type {};
template <typename P>
struct HasFirst<P, std::void_t<decltype(&P::first)>> : std::true_type {};

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

// gofs[g] = first index of group g (0-based) from the inclusive head counts gid1 alone,
// gofs[ngroups] = m
static __global__ void k_grp_offsets_gid(const uint32_t* __restrict__ gid1, uint32_t m, uint32_t* __restrict__ gofs)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t g = gid1[i];
    if (i == 0 || gid1[i - 1] != g) gofs[g - 1] = i;
    if (i == m - 1) gofs[g] = m;
}

// Element i of the initial order: rank and SA slot, whether its group stays active, and the
// group size (the first compaction reads the rest of the group record from gid1 / gofs).  With
// `pairs` the rank is not scattered but written as pairs[i] = x << 32 | rank, in order.  `pairs`
// may be the sorted key array, which a policy's first() can still read at i (not restrict: the
// read stays before the write).
template <typename Policy>
__global__ void k_dbl_init(const uint32_t* __restrict__ V, const uint32_t* __restrict__ gid1,
                           const uint32_t* __restrict__ gofs, uint32_t m, uint32_t* __restrict__ rank,
                           uint64_t* pairs, uint32_t* __restrict__ SA, uint32_t* __restrict__ keep,
                           uint32_t* __restrict__ NGz, uint32_t* __restrict__ NR, Policy pol)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t g = gid1[i] - 1;
    const uint32_t s = gofs[g], z = gofs[g + 1] - s;
    const uint32_t x = V[i];
    uint32_t r = 0;
    bool kp;
    if constexpr (HasFirst<Policy>::value) {
        kp = pol.first(SA, i, x, s, z, r);
    } else {
        if constexpr (Policy::kCarry) r = pol.rem(x);
        kp = z > 1 && pol.open(x, r);
        if (!kp) pol.place(SA, i, x, s);
    }
    if constexpr (Policy::kCarry) NR[i] = r;
    if (pairs) pairs[i] = ((uint64_t)x << 32) | (s + 1);
    else rank[x] = s + 1;
    keep[i] = kp ? 1u : 0u;
    NGz[i] = z;
}

// rank[x] = r for the (x << 32 | r) pairs, partitioned by the high bits of x so that the writes of
// neighbouring blocks fall into one window of the rank array.  Workgroups go to the 8 XCDs in
// turn (workgroup b to XCD b mod 8), and each XCD has its own L2: the blocks are renumbered so
// that XCD q takes the q-th eighth of the pairs, a contiguous range of windows, and its L2
// gathers whole lines of them instead of eight XCDs writing parts of the same lines.  The grid
// is a multiple of 8.
static __global__ void k_dbl_rank_pairs(const uint64_t* __restrict__ P, uint32_t m, uint32_t* __restrict__ rank)
{
    const uint32_t per = gridDim.x >> 3;
    const uint32_t b = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    const uint32_t i = b * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint64_t p = P[i];
    rank[(uint32_t)(p >> 32)] = (uint32_t)p;
}

// The first placement writes the rank of every element, a permutation of [0, m).  Scattered
// directly, each 4-byte write costs a random access to HBM once the rank array outgrows the
// last-level cache; from this many elements on, the (position, rank) pairs are first partitioned
// by the top kRankPartBits bits of the position (two radix passes) and then written window by
// window.  Measured with a random permutation on MI355X: 369 M elements 14.6 ms direct against
// 1.9 + 4.7 ms with one 8-bit pass, 64 M elements 0.88 ms against 0.35 + 0.49 ms; in the 256 MiB
// order-2 Markov text's dictionary sort (352 M positions) 16 bits in two passes take 1.1 ms less
// than 8 bits in one (12 bits: the same as 16).
constexpr int kRankPartBits = 16;
constexpr uint32_t kRankPartMin = 96u << 20;

// ---- one refinement round over the m active elements (groups contiguous)

// The slot count is read from the device (*mdev, written by the compaction: the round can be
// enqueued before the host knows the count, with a grid sized for an upper bound) or, for a round
// over an uncompacted layout, given by the host (mdev == nullptr, m_host).  In such a layout the
// slots of finished elements are holes: group size 0 (`live` = the group sizes, else nullptr).
__device__ __forceinline__ uint32_t slot_count(const uint32_t* mdev, uint32_t m_host)
{
    return mdev ? *mdev : m_host;
}

template <typename Policy>
__global__ void k_key2(const uint32_t* __restrict__ V, const uint32_t* __restrict__ R,
                       const uint32_t* __restrict__ mdev, uint32_t m_host, const uint32_t* __restrict__ live,
                       const uint32_t* __restrict__ rank, Policy pol, uint32_t* __restrict__ key2)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slot_count(mdev, m_host)) return;
    // all loads of the slot at once (a hole's element and rest are stale but in bounds)
    const uint32_t z = live ? live[i] : 1u;
    uint32_t r = 0;
    if constexpr (Policy::kCarry) r = R[i];
    const uint32_t x = V[i];
    if (z == 0) return;  // a hole: no key
    const uint32_t y = pol.next(x, r);
    key2[i] = (y == kNoNext) ? 0u : rank[y];
}

// Small groups (at most kSmallGroup members) are sorted by the block that holds their first
// slot: such a group lies inside [b0, b0 + 2 * kBlock).  Every member stages (key << 32 | slot)
// in LDS and counts the members of its group with a smaller value (one 64-bit compare per
// member, 16-byte LDS reads, broadcast within a group), which is its place in the sorted
// group; sub-group starts are then a max-scan of the heads of equal-key runs, sub-group ends
// are recorded by the last member of each run.  Outputs are written in sorted slot order:
// rank = global start of the sub-group + 1, the SA slot, the sub-group record and the keep
// flag.  Members of larger groups are only flagged (per slot of the block's own range).
static_assert(kSmallGroup <= kBlock, "a small group must fit in the block's extended range");

// number of staged values in wk[a, a + z) below c; wk has at least a + z + 1 entries when
// a + z is odd (the pair-aligned reads cover one neighbour on each side, subtracted again)
__device__ __forceinline__ uint32_t count_below(const uint64_t* wk, uint32_t a, uint32_t z, uint64_t c)
{
    uint32_t j = a & ~1u;
    const uint32_t e = (a + z + 1) & ~1u;
    uint32_t n = 0;
    for (; j + 8 <= e; j += 8) {
        const ulonglong2 q0 = *reinterpret_cast<const ulonglong2*>(wk + j);
        const ulonglong2 q1 = *reinterpret_cast<const ulonglong2*>(wk + j + 2);
        const ulonglong2 q2 = *reinterpret_cast<const ulonglong2*>(wk + j + 4);
        const ulonglong2 q3 = *reinterpret_cast<const ulonglong2*>(wk + j + 6);
        n += (q0.x < c) + (q0.y < c) + (q1.x < c) + (q1.y < c) + (q2.x < c) + (q2.y < c) + (q3.x < c) +
             (q3.y < c);
    }
    for (; j < e; j += 2) {
        const ulonglong2 q = *reinterpret_cast<const ulonglong2*>(wk + j);
        n += (q.x < c) + (q.y < c);
    }
    if (a & 1u) n -= (wk[a - 1] < c) ? 1u : 0u;
    if ((a + z) & 1u) n -= (wk[a + z] < c) ? 1u : 0u;
    return n;
}

template <typename Policy>
__global__ __launch_bounds__(kBlock) void k_seg_small(const uint32_t* __restrict__ V, const uint32_t* __restrict__ Hs,
                                                     const uint32_t* __restrict__ Gs, const uint32_t* __restrict__ Gz,
                                                     const uint32_t* __restrict__ key2, const uint32_t* __restrict__ R,
                                                     const uint32_t* __restrict__ mdev, uint32_t m_host,
                                                     uint32_t small_max, uint32_t* __restrict__ rank,
                                                     uint32_t* __restrict__ SA, uint32_t* __restrict__ keep,
                                                     uint32_t* __restrict__ NV, uint32_t* __restrict__ NHs,
                                                     uint32_t* __restrict__ NGs, uint32_t* __restrict__ NGz,
                                                     uint32_t* __restrict__ NR, uint32_t* __restrict__ large,
                                                     uint32_t* __restrict__ part, Policy pol)
{
    constexpr uint32_t kSpan = 2 * kBlock;
    constexpr uint32_t kSr = Policy::kCarry ? kSpan : 1;
    __shared__ __align__(16) uint64_t wk[kSpan];
    __shared__ uint32_t sk[kSpan];    // key at sorted slot
    __shared__ uint32_t sx[kSpan];    // element at sorted slot
    __shared__ uint32_t send[kSpan];  // end of the sub-group that starts at a slot
    __shared__ uint32_t sr[kSr];      // carried value at sorted slot
    __shared__ uint32_t tail[4];      // group crossing b0 + kBlock: start, size, global start, end
    __shared__ uint32_t red[2 * kWaves];
    const uint32_t b0 = blockIdx.x * kBlock;
    const uint32_t t = threadIdx.x;
    const uint32_t m = slot_count(mdev, m_host);
    if (b0 >= m) {  // the whole block (grid sized for an upper bound)
        if (t == 0) part[blockIdx.x] = 0;
        return;
    }

    // ---- own slots: classification and staging (a hole, group size 0, only clears its flags)
    const uint32_t i0 = b0 + t;
    bool own0 = false;
    uint32_t a0 = 0, z0 = 0, gs0 = 0, x0 = 0, k0 = 0, r0 = 0;
    if (i0 < m) {
        const uint32_t s = Hs[i0];  // loaded with the size (a hole's is stale, unused)
        z0 = Gz[i0];
        const bool lg = z0 > small_max || z0 > kSmallGroup;
        large[i0] = lg ? 1u : 0u;
        if (z0 == 0) {
            keep[i0] = 0;
            NGz[i0] = 0;
        }
        own0 = z0 != 0 && !lg && s >= b0;
        if (own0) {
            a0 = s - b0;
            gs0 = Gs[i0];
            x0 = V[i0];
            k0 = key2[i0];
            if constexpr (Policy::kCarry) r0 = R[i0];
            wk[t] = ((uint64_t)k0 << 32) | t;
        }
    }
    if (t == kBlock - 1) {
        const bool cross = own0 && a0 + z0 > kBlock;
        tail[0] = a0;
        tail[1] = z0;
        tail[2] = gs0;
        tail[3] = cross ? a0 + z0 : kBlock;
    }
    __syncthreads();

    // ---- the rest of the group that crosses b0 + kBlock (inside the grid's range)
    const uint32_t p1 = kBlock + t;
    const bool own1 = p1 < tail[3];
    uint32_t a1 = 0, z1 = 0, gs1 = 0, x1 = 0, k1 = 0, r1 = 0;
    if (own1) {
        a1 = tail[0];
        z1 = tail[1];
        gs1 = tail[2];
        x1 = V[b0 + p1];
        k1 = key2[b0 + p1];
        if constexpr (Policy::kCarry) r1 = R[b0 + p1];
        wk[p1] = ((uint64_t)k1 << 32) | p1;
    }
    __syncthreads();

    // ---- place in the sorted group
    if (own0) {
        const uint32_t q = a0 + count_below(wk, a0, z0, ((uint64_t)k0 << 32) | t);
        sk[q] = k0;
        sx[q] = x0;
        if constexpr (Policy::kCarry) sr[q] = r0;
    }
    if (own1) {
        const uint32_t q = a1 + count_below(wk, a1, z1, ((uint64_t)k1 << 32) | p1);
        sk[q] = k1;
        sx[q] = x1;
        if constexpr (Policy::kCarry) sr[q] = r1;
    }
    __syncthreads();

    // ---- sub-groups: heads of equal-key runs, their starts by a max-scan over both halves
    const uint32_t p0 = t;
    uint32_t kk0 = 0, kk1 = 0;
    uint32_t h0 = 0, h1 = 0;
    if (own0) {
        kk0 = sk[p0];
        h0 = (p0 == a0 || sk[p0 - 1] != kk0) ? p0 : 0u;
    }
    if (own1) {
        kk1 = sk[p1];
        h1 = (p1 == a1 || sk[p1 - 1] != kk1) ? p1 : 0u;
    }
    const uint32_t lane = t & 63, wid = t >> 6;
    uint32_t c0 = wave_inclusive_scan(h0, OpMax{});
    uint32_t c1 = wave_inclusive_scan(h1, OpMax{});
    if (lane == 63) {
        red[wid] = c0;
        red[kWaves + wid] = c1;
    }
    __syncthreads();
    uint32_t pre0 = 0, pre1 = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
        const uint32_t r0 = red[w], r1 = red[kWaves + w];
        pre1 = r0 > pre1 ? r0 : pre1;  // the whole first half precedes the second
        if (w < (int)wid) {
            pre0 = r0 > pre0 ? r0 : pre0;
            pre1 = r1 > pre1 ? r1 : pre1;
        }
    }
    const uint32_t ss0 = c0 > pre0 ? c0 : pre0;
    const uint32_t ss1 = c1 > pre1 ? c1 : pre1;
    if (own0 && (p0 + 1 == a0 + z0 || sk[p0 + 1] != kk0)) send[ss0] = p0 + 1;
    if (own1 && (p1 + 1 == a1 + z1 || sk[p1 + 1] != kk1)) send[ss1] = p1 + 1;
    __syncthreads();

    // ---- outputs in sorted slot order; a finished member leaves a hole (group size 0)
    uint32_t kept = 0;  // kept members | kept members of large sub-groups << 16
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const bool own = h ? own1 : own0;
        if (!own) continue;
        const uint32_t p = h ? p1 : p0, a = h ? a1 : a0, gs = h ? gs1 : gs0, ss = h ? ss1 : ss0;
        const uint32_t eq = send[ss] - ss, lt = ss - a;
        const uint32_t x = sx[p], g = b0 + p;
        uint32_t r = 0;
        if constexpr (Policy::kCarry) {
            r = sr[p];
            NR[g] = r;
        }
        // members of the group's first sub-group keep their rank (gs + 1): no write
        if (lt != 0) rank[x] = gs + lt + 1;
        const bool kp = eq > 1 && pol.open(x, r);
        if (!kp) pol.place(SA, gs + (p - a), x, gs + lt);
        keep[g] = kp ? 1u : 0u;
        NV[g] = x;
        NHs[g] = b0 + ss;
        NGs[g] = gs + lt;
        NGz[g] = kp ? eq : 0u;
        if (kp) kept += 1u + ((eq > small_max) ? (1u << 16) : 0u);
    }
    // this block's kept members (at most 2 * kBlock per field): the next round's counts
    uint32_t tot;
    (void)block_exclusive_scan(kept, 0u, OpSum{}, red, &tot);
    if (t == 0) part[blockIdx.x] = tot;
}

// Large groups: (group start slot, key) pairs of the flagged members in slot order, with the
// member's active slot as the value.
static __global__ void k_seg_large_gather(const uint32_t* __restrict__ Hs, const uint32_t* __restrict__ key2,
                                          const uint32_t* __restrict__ large, const uint32_t* __restrict__ lpos,
                                          uint32_t m, int B, uint64_t* __restrict__ LK, uint32_t* __restrict__ LV)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || !large[i]) return;
    const uint32_t q = lpos[i];
    LK[q] = ((uint64_t)Hs[i] << B) | key2[i];
    LV[q] = i;
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

// One member of a large group, sorted position q: its new rank, placement or sub-group record.
template <typename Policy>
__device__ __forceinline__ uint32_t lg_update_one(const uint64_t* __restrict__ LK, const uint32_t* __restrict__ LV,
                                                  const uint32_t* __restrict__ qh, const uint32_t* __restrict__ sz,
                                                  const uint32_t* __restrict__ lpos, const uint32_t* __restrict__ V,
                                                  const uint32_t* __restrict__ R, const uint32_t* __restrict__ Gs,
                                                  uint32_t q, int B, uint32_t* __restrict__ rank,
                                                  uint32_t* __restrict__ SA, uint32_t* __restrict__ keep,
                                                  uint32_t* __restrict__ NV, uint32_t* __restrict__ NHs,
                                                  uint32_t* __restrict__ NGs, uint32_t* __restrict__ NGz,
                                                  uint32_t* __restrict__ NR, uint32_t small_max, Policy pol)
{
    const uint32_t i = LV[q];
    const uint32_t x = V[i];
    uint32_t r = 0;
    const uint32_t s = (uint32_t)(LK[q] >> B);  // group start slot (a member of the group)
    const uint32_t q0 = lpos[s];                // first index of the group among the large ones
    const uint32_t lt = qh[q] - q0, c = sz[qh[q]];
    const uint32_t gs = Gs[s];
    const uint32_t p = s + (q - q0);
    if constexpr (Policy::kCarry) {
        r = R[i];
        NR[p] = r;
    }
    if (lt != 0) rank[x] = gs + lt + 1;  // the first sub-group keeps its rank
    const bool kp = c > 1 && pol.open(x, r);
    if (!kp) pol.place(SA, gs + (q - q0), x, gs + lt);
    keep[p] = kp ? 1u : 0u;
    NV[p] = x;
    NHs[p] = s + lt;
    NGs[p] = gs + lt;
    NGz[p] = kp ? c : 0u;
    return kp ? 1u + ((c > small_max) ? (1u << 16) : 0u) : 0u;
}

template <typename Policy>
__global__ void k_lg_update(const uint64_t* __restrict__ LK, const uint32_t* __restrict__ LV,
                            const uint32_t* __restrict__ qh, const uint32_t* __restrict__ sz,
                            const uint32_t* __restrict__ lpos, const uint32_t* __restrict__ V,
                            const uint32_t* __restrict__ R, const uint32_t* __restrict__ Gs, uint32_t nl, int B,
                            uint32_t* __restrict__ rank, uint32_t* __restrict__ SA, uint32_t* __restrict__ keep,
                            uint32_t* __restrict__ NV, uint32_t* __restrict__ NHs, uint32_t* __restrict__ NGs,
                            uint32_t* __restrict__ NGz, uint32_t* __restrict__ NR, uint32_t small_max,
                            uint32_t* __restrict__ extra, Policy pol)
{
    __shared__ uint32_t red[kWaves];
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t kept = 0;  // as in k_seg_small: kept | kept in large sub-groups << 16
    if (q < nl) kept = lg_update_one(LK, LV, qh, sz, lpos, V, R, Gs, q, B, rank, SA, keep, NV, NHs, NGs, NGz, NR,
                                     small_max, pol);
    uint32_t tot;
    (void)block_exclusive_scan(kept, 0u, OpSum{}, red, &tot);
    if (threadIdx.x == 0 && tot) {
        atomicAdd(extra, tot & 0xFFFFu);
        atomicAdd(extra + 1, tot >> 16);
    }
}

// Compaction of the kept members in one pass over chained tiles of kCompactTile slots: the
// block takes kCompactRounds rounds of kBlock consecutive slots, records each wavefront's kept
// mask, scans the (round, wavefront) counts, obtains the tile's offset by decoupled look-back
// and moves every kept member to its place with coalesced reads.  A kept sub-group is kept
// whole and stays contiguous, so its first slot moves to o - (p - first slot) for a member that
// moves from p to o.  The first compaction reads the initial groups (kInit: NV = initial
// order, NHs = 1-based group id, NGs = group offsets).  Also counts the kept members whose
// sub-group is large next round (lpart per tile, summed by k_sum_tiles) and the total (*active).
constexpr int kCompactRounds = 16;
constexpr uint32_t kCompactTile = kCompactRounds * kBlock;  // 4096 slots: many tiles fill the chip
template <bool kInit>
__global__ __launch_bounds__(kBlock) void k_dbl_compact(const uint32_t* __restrict__ keep, uint32_t m,
                                                       uint32_t small_max, const uint32_t* __restrict__ NV,
                                                       const uint32_t* __restrict__ NHs,
                                                       const uint32_t* __restrict__ NGs,
                                                       const uint32_t* __restrict__ NGz,
                                                       const uint32_t* __restrict__ NR, uint32_t* __restrict__ V,
                                                       uint32_t* __restrict__ Hs, uint32_t* __restrict__ Gs,
                                                       uint32_t* __restrict__ Gz, uint32_t* __restrict__ R,
                                                       uint32_t* __restrict__ active, uint32_t* __restrict__ lpart,
                                                       Chain c, uint32_t tiles)
{
    constexpr int kRounds = kCompactRounds;
    static_assert(kRounds * kWaves == 64, "the (round, wavefront) counts are scanned by one wavefront");
    __shared__ uint64_t bal[kRounds * kWaves];
    __shared__ uint32_t off[kRounds * kWaves];
    __shared__ uint32_t lds[kWaves];
    __shared__ uint32_t slot;
    __shared__ uint32_t tile_total;
    const uint32_t tile = chain_ticket(c, tiles, &slot);
    const uint64_t base = (uint64_t)tile * kCompactTile;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // all flag loads of a batch are issued before the first ballot
    constexpr int kBatch = kRounds;
    for (int it0 = 0; it0 < kRounds; it0 += kBatch) {
        uint32_t kv[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; j++) {
            const uint64_t p = base + (uint64_t)(it0 + j) * kBlock + threadIdx.x;
            kv[j] = (p < m) ? keep[p] : 0u;
        }
#pragma unroll
        for (int j = 0; j < kBatch; j++) {
            const uint64_t b = __ballot(kv[j] != 0);
            if (lane == 0) {
                bal[(it0 + j) * kWaves + w] = b;
                off[(it0 + j) * kWaves + w] = (uint32_t)__popcll(b);
            }
        }
    }
    __syncthreads();
    if (w == 0) {
        const uint32_t cnt = off[lane];
        const uint32_t inc = wave_inclusive_scan(cnt, OpSum{});
        off[lane] = inc - cnt;
        if (lane == 63) tile_total = inc;
    }
    __syncthreads();
    const uint32_t tot = tile_total;
    const uint32_t pre = chain_prefix(c, tile, tot, OpSum{}, 0u, &slot);
    const uint64_t lt = lanemask_lt();
    uint32_t nl = 0;
    // moves in batches: the loads of a batch are issued before its stores
    constexpr int kMove = 8;
    for (int it0 = 0; it0 < kRounds; it0 += kMove) {
        bool kk[kMove];
        uint32_t xv[kMove], zv[kMove], hv[kMove], gv[kMove], rv[kMove];
#pragma unroll
        for (int j = 0; j < kMove; j++) {
            const uint64_t b = bal[(it0 + j) * kWaves + w];
            kk[j] = (b >> lane) & 1u;
            const uint32_t p = (uint32_t)(base + (uint64_t)(it0 + j) * kBlock + threadIdx.x);
            xv[j] = zv[j] = hv[j] = gv[j] = rv[j] = 0;
            if (kk[j]) {
                xv[j] = NV[p];
                zv[j] = NGz[p];
                hv[j] = NHs[p];
                if constexpr (!kInit) gv[j] = NGs[p];
                if (NR) rv[j] = NR[p];
            }
        }
#pragma unroll
        for (int j = 0; j < kMove; j++) {
            if (!kk[j]) continue;
            const uint64_t b = bal[(it0 + j) * kWaves + w];
            const uint32_t p = (uint32_t)(base + (uint64_t)(it0 + j) * kBlock + threadIdx.x);
            const uint32_t o = pre + off[(it0 + j) * kWaves + w] + (uint32_t)__popcll(b & lt);
            V[o] = xv[j];
            Gz[o] = zv[j];
            nl += (zv[j] > small_max) ? 1u : 0u;
            if constexpr (kInit) {
                const uint32_t s = NGs[hv[j] - 1];
                Hs[o] = o - (p - s);
                Gs[o] = s;
            } else {
                Hs[o] = o - (p - hv[j]);
                Gs[o] = gv[j];
            }
            if (NR) R[o] = rv[j];
        }
    }
    uint32_t ltot;
    (void)block_exclusive_scan(nl, 0u, OpSum{}, lds, &ltot);
    if (threadIdx.x == 0) {
        lpart[tile] = ltot;
        if (tile == tiles - 1) *active = pre + tot;
    }
}

// Also publishes (*active, total) to the host signal words with sequence number `seq` when
// `sig` is given (stores to mapped host memory, then the sequence number with release
// semantics at system scope).
static __global__ __launch_bounds__(kBlock) void k_sum_tiles(const uint32_t* __restrict__ part, uint32_t tiles,
                                                            uint32_t* __restrict__ out,
                                                            const uint32_t* __restrict__ active = nullptr,
                                                            uint32_t* __restrict__ sig = nullptr,
                                                            uint32_t seq = 0)
{
    __shared__ uint32_t lds[kWaves];
    uint32_t acc = 0;
    for (uint32_t t = threadIdx.x; t < tiles; t += kBlock) acc += part[t];
    uint32_t tot;
    (void)block_exclusive_scan(acc, 0u, OpSum{}, lds, &tot);
    if (threadIdx.x == 0) {
        *out = tot;
        if (sig) {
            __hip_atomic_store(sig, *active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(sig + 1, tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(sig + Storage::kSigSeq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// Next round's counts from the per-block counts of k_seg_small (kept | kept in large sub-groups
// << 16): every block sums a contiguous chunk and adds it to extra[0], extra[1], which already
// hold the totals of k_lg_update; k_round_publish then copies them to the scalar slots and
// publishes them like k_sum_tiles.
constexpr uint32_t kRoundSumBlocks = 128;
static __global__ __launch_bounds__(kBlock) void k_round_sum(const uint32_t* __restrict__ part, uint32_t blocks,
                                                            uint32_t* __restrict__ extra)
{
    __shared__ uint32_t lds[kWaves];
    const uint32_t chunk = (blocks + gridDim.x - 1) / gridDim.x;
    const uint32_t lo = blockIdx.x * chunk, hi = lo + chunk < blocks ? lo + chunk : blocks;
    uint32_t a = 0, l = 0;
    for (uint32_t t = lo + threadIdx.x; t < hi; t += kBlock) {
        const uint32_t v = part[t];
        a += v & 0xFFFFu;
        l += v >> 16;
    }
    uint32_t ta, tl;
    (void)block_exclusive_scan(a, 0u, OpSum{}, lds, &ta);
    (void)block_exclusive_scan(l, 0u, OpSum{}, lds, &tl);
    if (threadIdx.x == 0) {
        if (ta) atomicAdd(extra, ta);
        if (tl) atomicAdd(extra + 1, tl);
    }
}

static __global__ void k_round_publish(const uint32_t* __restrict__ extra, uint32_t* __restrict__ active,
                                       uint32_t* __restrict__ large, uint32_t* __restrict__ sig, uint32_t seq)
{
    if (threadIdx.x != 0) return;
    const uint32_t ta = extra[0], tl = extra[1];
    *active = ta;
    *large = tl;
    if (sig) {
        __hip_atomic_store(sig, ta, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(sig + 1, tl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(sig + Storage::kSigSeq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
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
// `max_rounds` bounds the loop: a caller whose keys can be inconsistent (elements that never
// separate) gets MIBWT_HASH_COLLISION instead of an endless refinement.
// K1 (optional, N entries) is scratch for the partitioned first placement.
template <typename Policy, typename Advance>
DoublingResult refine_groups(Storage& S, uint64_t* K0, uint32_t* V0, uint32_t* V1, uint32_t N, Policy& pol,
                             Advance advance, uint32_t* rank, uint32_t* SA, int max_rounds = 1 << 30,
                             uint64_t* K1 = nullptr, uint64_t key_mask = ~0ull)
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
    uint32_t* R = nullptr;
    uint32_t* NR = nullptr;
    if constexpr (Policy::kCarry) {
        R = S.buf<uint32_t>(B_DB_R, N);
        NR = S.buf<uint32_t>(B_DB_NR, N);
    }
    const int B = bits_for((uint64_t)N + 1);

    // initial groups: runs of equal keys (gid1 lives in `large` until the first compaction)
    uint32_t* gid1 = large;
    uint32_t* gofs = S.buf<uint32_t>(B_DB_GOFS0, N + 1);
    {
        sum_scan_of(S, ScanHeads<uint64_t>{K0, key_mask}, true, gid1, N, true, nullptr);
        k_grp_offsets_gid<<<grid_for(N), kBlock, 0, S.stream>>>(gid1, N, gofs);
        MIBWT_LAUNCH_CHECK();
        // the keys are dead once the heads are flagged: K0 takes the (position, rank) pairs
        const bool part = K1 && N >= kRankPartMin;
        k_dbl_init<Policy><<<grid_for(N), kBlock, 0, S.stream>>>(V0, gid1, gofs, N, rank, part ? K0 : nullptr, SA,
                                                                  keep, NGz, NR, pol);
        MIBWT_LAUNCH_CHECK();
        res.bytes += (double)N * (8 * 2 + 4 * 9);
        if (part) {
            const int xb = bits_for((uint64_t)N - 1);
            int where = 0;
            radix_sort<uint64_t, uint32_t>(S, K0, K1, nullptr, nullptr, N, 32 + std::max(0, xb - kRankPartBits),
                                           32 + xb,
                                           static_cast<const EmitPair<uint64_t, uint32_t>*>(nullptr), &where);
            k_dbl_rank_pairs<<<(grid_for(N) + 7) & ~7u, kBlock, 0, S.stream>>>(where ? K1 : K0, N, rank);
            MIBWT_LAUNCH_CHECK();
            res.bytes += (double)N * (8 * 3 + 4);
        }
    }
    // A round normally starts with the compaction of the kept sub-groups.  For the string
    // families, while most slots stay active (an estimated kLazyActive of them), the compaction is
    // skipped: the round's outputs, written at the same slots, become the next round's layout,
    // and the slots of finished elements are holes (group size 0) that the kernels pass over;
    // the counts then come from k_seg_small / k_lg_update.  (The chain policy steps its pointers
    // over the compacted active set and always compacts.)
    constexpr double kLazyActive = 0.8;
    static const bool in_place_ok = [] {  // MIBWT_INPLACE_ROUNDS=0: compact every round
        const char* e = getenv("MIBWT_INPLACE_ROUNDS");
        return !(e && e[0] == '0');
    }();
    bool first = true, compact = true;
    uint32_t m = N;  // slots of the current layout (after a compaction: the active elements)
    uint32_t prev_active = N;
    uint32_t* part = S.buf<uint32_t>(B_DB_PART, grid_for(N));
    uint32_t* extra = S.buf<uint32_t>(B_DB_EXTRA, 2);
    uint32_t seq = 0;
    for (;;) {
        // the counts come back through the mapped signal words, or, in arena mode (no mapped
        // memory), by a plain read of the scalar slots
        uint32_t* sigd = S.arena ? nullptr : S.signal_dev();
        if (compact) {
            seq = S.arena ? 0u : S.signal_next();
            const uint32_t tiles = std::max<uint32_t>(1, ceil_div(m, kCompactTile));
            uint32_t* lpart = S.buf<uint32_t>(B_DB_LPART, tiles);
            const Chain c = S.chain(tiles);
            if (first) {
                k_dbl_compact<true><<<tiles, kBlock, 0, S.stream>>>(keep, m, kSmallGroup, V0, gid1, gofs, NGz, NR, V,
                                                                    Hs, Gs, Gz, R, sc + SC_ACTIVE, lpart, c, tiles);
                first = false;
            } else {
                k_dbl_compact<false><<<tiles, kBlock, 0, S.stream>>>(keep, m, kSmallGroup, NV, NHs, NGs, NGz, NR, V,
                                                                     Hs, Gs, Gz, R, sc + SC_ACTIVE, lpart, c, tiles);
            }
            MIBWT_LAUNCH_CHECK();
            k_sum_tiles<<<1, kBlock, 0, S.stream>>>(lpart, tiles, sc + SC_LARGE, sc + SC_ACTIVE, sigd, seq);
            MIBWT_LAUNCH_CHECK();
        } else {
            // last round's outputs are this round's layout (its counts are on their way: seq)
            std::swap(V, NV);
            std::swap(Hs, NHs);
            std::swap(Gs, NGs);
            std::swap(Gz, NGz);
            if constexpr (Policy::kCarry) std::swap(R, NR);
        }
        // ---- order each group by the rank of the next position: secondary keys and the sort of
        // the small groups; with Policy::kSpeculate enqueued before the count comes back (the
        // kernels read it on the device, grids sized for the previous count), so the device
        // does not idle while the host waits
        const uint32_t* mdev = compact ? sc + SC_ACTIVE : nullptr;
        const uint32_t* live = compact ? nullptr : Gz;
        uint32_t blocks = 0;
        auto sort_round = [&](uint32_t mb) {
            // Policy::kAdvanceFirst: the step to this round's positions is taken at the start of
            // the round (from the second on) over the compacted active set only.  For chains this
            // is enough: an element that stays active after a round points (before the step) at
            // an element that was active in that round, so it was stepped at that round's start.
            if constexpr (Policy::kAdvanceFirst)
                if (res.rounds > 1) advance(pol, V, mb);
            k_key2<Policy><<<grid_for(mb), kBlock, 0, S.stream>>>(V, R, mdev, m, live, rank, pol, key2);
            MIBWT_LAUNCH_CHECK();
            if constexpr (!Policy::kAdvanceFirst) advance(pol, V, mb);
            blocks = grid_for(mb);
            k_seg_small<Policy><<<blocks, kBlock, 0, S.stream>>>(V, Hs, Gs, Gz, key2, R, mdev, m, kSmallGroup, rank,
                                                                 SA, keep, NV, NHs, NGs, NGz, NR, large, part, pol);
            MIBWT_LAUNCH_CHECK();
        };
        if constexpr (Policy::kSpeculate) sort_round(m);
        uint32_t hv[2];
        if (S.arena) {
            uint32_t all[SC_COUNT_];
            S.read(all, sc, SC_COUNT_);
            hv[0] = all[SC_ACTIVE];
            hv[1] = all[SC_LARGE];
        } else {
            S.wait_signal(seq, hv, 2);
        }
        const uint32_t m2 = hv[0], nl = hv[1];
        res.rounds++;
        if (S.verbose)
            fprintf(stderr, "mibwt: refinement round %d: %u of %u elements active, %u in large groups (%s)\n",
                    res.rounds, m2, N, nl, compact ? "compacted" : "in place");
        MIBWT_REQUIRE(m2 == 0 || res.rounds < max_rounds, MIBWT_HASH_COLLISION, "groups do not separate");
        res.bytes += compact ? (double)m * 4 + (double)m2 * (4 * 5 + 4 * 4) : 0.0;
        if (m2 == 0) break;
        if (compact) m = m2;
        if constexpr (!Policy::kSpeculate) sort_round(m);
        res.bytes += (double)m * (4 * 12);
        fill<uint32_t>(S, extra, 2, 0u);
        if (nl > 0) {
            uint32_t* lpos = pos;
            exclusive_sum<uint32_t>(S, large, lpos, m);
            uint64_t* LK0 = S.buf<uint64_t>(B_LG_K0, nl);
            uint64_t* LK1 = S.buf<uint64_t>(B_LG_K1, nl);
            uint32_t* LV0 = S.buf<uint32_t>(B_LG_V0, nl);
            uint32_t* LV1 = S.buf<uint32_t>(B_LG_V1, nl);
            uint32_t* qh = S.buf<uint32_t>(B_LG_I, nl);
            uint32_t* sz = S.buf<uint32_t>(B_LG_SZ, nl);
            k_seg_large_gather<<<grid_for(m), kBlock, 0, S.stream>>>(Hs, key2, large, lpos, m, B, LK0, LV0);
            MIBWT_LAUNCH_CHECK();
            const int kb = B + bits_for(m);
            radix_sort<uint64_t, uint32_t>(S, LK0, LK1, LV0, LV1, nl, 0, kb);
            k_lg_heads<<<grid_for(nl), kBlock, 0, S.stream>>>(LK0, nl, qh);
            MIBWT_LAUNCH_CHECK();
            inclusive_max<uint32_t>(S, qh, qh, nl);
            k_lg_sizes<<<grid_for(nl), kBlock, 0, S.stream>>>(LK0, qh, nl, sz);
            MIBWT_LAUNCH_CHECK();
            k_lg_update<Policy><<<grid_for(nl), kBlock, 0, S.stream>>>(LK0, LV0, qh, sz, lpos, V, R, Gs, nl, B, rank,
                                                                       SA, keep, NV, NHs, NGs, NGz, NR, kSmallGroup,
                                                                       extra, pol);
            MIBWT_LAUNCH_CHECK();
            res.large_items += nl;
            res.bytes += (double)nl * (12 * 2 * ((kb + 7) / 8) + 48);
        }
        // the next round in place while most slots stay active: its active count is estimated
        // from this round's shrink (m2 of prev_active) against the m slots of this layout
        const double next_active = (double)m2 * ((double)m2 / (double)prev_active);
        prev_active = m2;
        compact = !(Policy::kSpeculate && in_place_ok && next_active >= kLazyActive * (double)m);
        if (!compact) {
            seq = S.arena ? 0u : S.signal_next();
            k_round_sum<<<std::min(kRoundSumBlocks, blocks), kBlock, 0, S.stream>>>(part, blocks, extra);
            MIBWT_LAUNCH_CHECK();
            k_round_publish<<<1, 64, 0, S.stream>>>(extra, sc + SC_ACTIVE, sc + SC_LARGE, sigd, seq);
            MIBWT_LAUNCH_CHECK();
        }
    }
    return res;
}

}  // namespace mibwt
