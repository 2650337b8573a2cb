// Lyndon-grammar BWT on gfx950.
//
// Reference: Lyndon_Grammar_BWT.cu (build_Lyndon_Grammar :2592, Sort_Lyndon_Grammar :3621,
// Build_BWT :4628).  Every position i of the text has a longest Lyndon prefix
// P_i = T[i .. i+lam[i]); these are the maximal nodes of the Lyndon forest (stack words and
// right children, the reference's rules).  The suffix order is the order of (P_i, suffix at
// i+lam[i]): Lyndon prefixes compare lexicographically (a proper prefix is smaller), equal
// prefixes are ordered by the suffix that follows, which starts at the next smaller suffix
// and is therefore smaller.  The reference builds the grammar block-locally and merges it
// right to left (:647-2062), deduplicates rules by Karp-Rabin fingerprints (:2183-2588),
// sorts rules (:3129-4083) and induces the BWT over rank groups (:4114-4896).  Here:
//
//  1. rules    Karp-Rabin prefix fingerprints (mod 2^61-1) of the text in one device scan.
//  2. lyndon   Lyndon array as next-smaller-suffix distances: every 256-byte block chases
//              NSS pointers right to left; positions without a smaller suffix in the block
//              stay open and are resolved by a merge tree over neighbouring ranges, where an
//              open position bisects the right range's prefix minima.  Suffix comparisons
//              use the fingerprints for long common prefixes (O(log lcp) per comparison).
//  3. rsort    positions ordered by P_i: radix sort of the first symbols packed with dense
//              alphabet codes, then groups that tie on that prefix are ordered with a
//              fingerprint-LCP comparator (equal fingerprints over the whole prefix = the same
//              rule); the final order is checked pair by pair byte-exactly and the
//              fingerprint base is re-drawn if a collision ever broke it.
//  4. induce   groups of equal P_i are refined by the rank of the following suffix, along
//              the next-smaller-suffix chain with pointer jumping (segmented group refinement
//              shared with the prefix-free-parsing pipeline).
//  5. output   O[j] = T[SA[j]-1] with the wrap to T[n-1].
#include <chrono>
#include <cstdlib>

#include "doubling.hpp"
#include "text.hpp"

namespace mibwt {

bool bwt_inverts_to(Storage& S, const uint8_t* dO, int64_t n, int64_t index, const uint8_t* dT);

constexpr uint32_t kLynBlock = 256;   // characters per block of the local pass
// Label tie groups of more members go to the pivot partition (linear work per round; a group of
// equal labels settles in one round) instead of the comparator counting sort (quadratic in the
// group size, but one pass); MIBWT_LAB_SMALL overrides it (2..256).  Measured on the 256 MiB DNA
// collection: 256 -> rsort 83 ms, 64 -> 97 ms, 16 -> 148 ms, 4 -> 174 ms (more pivot rounds,
// each a radix sort, outweigh the cheaper comparisons).
constexpr uint32_t kLabSmall = 256;
constexpr uint64_t kM61 = (1ull << 61) - 1;

// ------------------------------------------------------------------- 1. Lyndon array

// longest common prefix of T[a..] and T[b..] (a < b), at most cap and never past n
__device__ __forceinline__ uint32_t lcp_text(const uint8_t* __restrict__ T, uint64_t n, uint32_t a, uint32_t b,
                                             uint32_t cap)
{
    const uint64_t lim64 = ((uint64_t)cap < n - b) ? (uint64_t)cap : n - b;
    const uint32_t lim = (uint32_t)lim64;
    uint32_t k = 0;
    while (k + 8 <= lim) {
        const uint64_t x = load8(T, a + k) ^ load8(T, b + k);
        if (x) return k + (uint32_t)(__builtin_ctzll(x) >> 3);
        k += 8;
    }
    while (k < lim && T[a + k] == T[b + k]) k++;
    return k;
}

// ------------------------------------------------------------------- 2. fingerprints

__device__ __forceinline__ uint64_t mulmod61(uint64_t a, uint64_t b)
{
    const uint64_t lo = a * b, hi = __umul64hi(a, b);
    uint64_t r = (lo & kM61) + (lo >> 61) + (hi << 3);
    r = (r & kM61) + (r >> 61);
    return r >= kM61 ? r - kM61 : r;
}

__device__ __forceinline__ uint64_t addmod61(uint64_t a, uint64_t b)
{
    const uint64_t r = a + b;
    return r >= kM61 ? r - kM61 : r;
}

struct HashPair {
    uint64_t h, p;  // polynomial value and base^length, both mod 2^61-1
};

__device__ __forceinline__ HashPair shfl_up_val(HashPair x, int o)
{
    HashPair y;
    y.h = __shfl_up(x.h, o, 64);
    y.p = __shfl_up(x.p, o, 64);
    return y;
}

struct OpHash {
    __device__ __host__ HashPair operator()(HashPair a, HashPair b) const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return HashPair{addmod61(mulmod61(a.h, b.p), b.h), mulmod61(a.p, b.p)};
#else
        return a;  // host instantiation is never executed
#endif
    }
};

__global__ void k_hash_items(const uint8_t* __restrict__ T, uint64_t n, uint64_t base, HashPair* __restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = HashPair{(uint64_t)T[i] + 1, base};
}

struct PrefixHash {
    const HashPair* pre;  // pre[k] = (H[k+1], base^(k+1))
    uint64_t mask;        // bits of the fingerprints compared (all 61 but in collision tests)
    // fingerprint of T[a..a+len)
    __device__ __forceinline__ uint64_t sub(uint32_t a, uint32_t len) const
    {
        if (len == 0) return 0;
        const uint64_t hb = pre[a + len - 1].h;
        if (a == 0) return hb & mask;
        const uint64_t ha = pre[a - 1].h;
        const uint64_t pw = pre[len - 1].p;
        const uint64_t x = mulmod61(ha, pw);
        return (hb >= x ? hb - x : hb + kM61 - x) & mask;
    }
};

// ------------------------------------------------------------------- 1. Lyndon array
//
// lam[i] = NSS[i] - i, NSS[i] = the nearest j > i whose suffix is smaller (n if none).
// Suffixes are compared through their longest common prefix: up to 32 bytes directly, then by
// galloping and bisection on the prefix fingerprints, so the long repeats of a collection cost
// O(log lcp) instead of O(lcp).
struct SuffixCmp {
    const uint8_t* T;
    uint64_t n;
    PrefixHash ph;
    __device__ uint32_t lcp(uint32_t a, uint32_t b) const
    {
        return lcp_upto(a, b, (uint32_t)(n - (a > b ? a : b)));  // the comparable length
    }
    // longest common prefix of the suffixes at a and b, at most m (m <= n - max(a, b))
    __device__ uint32_t lcp_upto(uint32_t a, uint32_t b, uint32_t m) const
    {
        uint32_t k = 0;
        while (k < 32 && k + 8 <= m) {
            const uint64_t x = load8(T, a + k) ^ load8(T, b + k);
            if (x) return k + (uint32_t)(__builtin_ctzll(x) >> 3);
            k += 8;
        }
        if (k < 32) {
            while (k < m && T[a + k] == T[b + k]) k++;
            return k;
        }
        // prefixes agree on [0, lo); find the first disagreement in [lo, hi)
        uint32_t lo = k, step = 32, hi;
        for (;;) {
            if (step > m - lo) {
                hi = m + 1;
                break;
            }
            if (ph.sub(a + lo, step) == ph.sub(b + lo, step)) {
                lo += step;
                step *= 2;
            } else {
                hi = lo + step;
                break;
            }
        }
        // largest l in [lo, hi) with equal prefixes of length l
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (ph.sub(a, mid) == ph.sub(b, mid)) lo = mid;
            else hi = mid;
        }
        return lo;
    }
    // suffix(a) < suffix(b) (a != b); a suffix that is a prefix of the other is smaller
    __device__ bool less(uint32_t a, uint32_t b) const
    {
        const uint32_t l = lcp(a, b);
        if ((uint64_t)a + l == n) return true;
        if ((uint64_t)b + l == n) return false;
        return T[a + l] < T[b + l];
    }
};

constexpr uint32_t kNoNss = 0xFFFFFFFFu;

// One thread per 256-byte block, right to left: NSS by pointer chasing inside the block.
// Positions without a smaller suffix inside the block stay open (kNoNss).  Per block: the
// open positions in ascending order (O, increasing suffixes) and the chain of prefix minima
// from the block start (M = s, NSS[s], NSS[NSS[s]], ..., decreasing suffixes).
__global__ void k_nss_local(SuffixCmp sc, uint32_t nb, uint32_t* __restrict__ nss, uint32_t* __restrict__ O,
                            uint32_t* __restrict__ oc, uint32_t* __restrict__ M, uint32_t* __restrict__ mc)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const uint32_t s = b * kLynBlock;
    const uint32_t e = (uint32_t)(((uint64_t)s + kLynBlock < sc.n) ? s + kLynBlock : sc.n);
    for (uint32_t i = e; i-- > s;) {
        uint32_t j = i + 1;
        while (j < e && !sc.less(j, i)) j = nss[j];  // kNoNss ends the walk as well
        nss[i] = (j < e) ? j : kNoNss;
    }
    uint32_t c = 0;
    for (uint32_t i = s; i < e; i++)
        if (nss[i] == kNoNss) O[s + c++] = i;
    oc[b] = c;
    c = 0;
    for (uint32_t m = s; m != kNoNss; m = nss[m]) M[s + c++] = m;
    mc[b] = c;
}

// One wavefront per pair of neighbouring ranges L, R of `size` blocks.  An open position i
// of L finds its NSS among R's prefix minima: they decrease, so the first one below suffix(i)
// is found by bisection.  The merged range keeps L's unresolved positions followed by R's
// open ones, and L's prefix minima followed by the prefix minima of R below min(L).
__global__ __launch_bounds__(64) void k_nss_merge(SuffixCmp sc, uint32_t nb, uint32_t size,
                                                  uint32_t* __restrict__ nss, const uint32_t* __restrict__ Oin,
                                                  const uint32_t* __restrict__ ocin, const uint32_t* __restrict__ Min,
                                                  const uint32_t* __restrict__ mcin, uint32_t* __restrict__ Oout,
                                                  uint32_t* __restrict__ ocout, uint32_t* __restrict__ Mout,
                                                  uint32_t* __restrict__ mcout)
{
    const int lane = threadIdx.x;
    const uint64_t a = (uint64_t)blockIdx.x * 2 * size;
    if (a >= nb) return;
    const uint64_t rb = a + size;
    const uint64_t sL = a * kLynBlock;
    const uint32_t oL = ocin[a], mL = mcin[a];
    const uint32_t* OL = Oin + sL;
    const uint32_t* ML = Min + sL;
    if (rb >= nb) {
        for (uint32_t q = lane; q < oL; q += 64) Oout[sL + q] = OL[q];
        for (uint32_t q = lane; q < mL; q += 64) Mout[sL + q] = ML[q];
        if (lane == 0) {
            ocout[a] = oL;
            mcout[a] = mL;
        }
        return;
    }
    const uint64_t sR = rb * kLynBlock;
    const uint32_t oR = ocin[rb], mR = mcin[rb];
    const uint32_t* OR = Oin + sR;
    const uint32_t* MR = Min + sR;
    // bisection: first t in [0, mR) with suffix(MR[t]) < suffix(x), mR if none
    auto first_below = [&](uint32_t x) -> uint32_t {
        uint32_t lo = 0, hi = mR;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (sc.less(MR[mid], x)) hi = mid;
            else lo = mid + 1;
        }
        return lo;
    };
    // L's open positions: resolve or keep (ordered compaction)
    uint32_t kept = 0;
    for (uint32_t q0 = 0; q0 < oL; q0 += 64) {
        const uint32_t q = q0 + lane;
        bool keep = false;
        if (q < oL) {
            const uint32_t i = OL[q];
            const uint32_t t = first_below(i);
            if (t < mR) nss[i] = MR[t];
            else keep = true;
        }
        const uint64_t bal = __ballot(keep);
        if (keep) Oout[sL + kept + (uint32_t)__popcll(bal & lanemask_lt())] = OL[q];
        kept += (uint32_t)__popcll(bal);
    }
    for (uint32_t q = lane; q < oR; q += 64) Oout[sL + kept + q] = OR[q];
    // prefix minima: all of L's, then R's below min(L) = the last of L's
    const uint32_t tstar = first_below(ML[mL - 1]);  // every lane computes the same value
    for (uint32_t q = lane; q < mL; q += 64) Mout[sL + q] = ML[q];
    for (uint32_t q = tstar + lane; q < mR; q += 64) Mout[sL + mL + (q - tstar)] = MR[q];
    if (lane == 0) {
        ocout[a] = kept + oR;
        mcout[a] = mL + (mR - tstar);
    }
}

__global__ void k_nss_to_lam(uint32_t* __restrict__ a, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = a[i];
    a[i] = (j == kNoNss ? (uint32_t)n : j) - (uint32_t)i;
}

// lam[] of T[0..n) (the padded text copy T, prefix fingerprints ph).  Returns the number of
// merge levels.
int lyndon_array(Storage& S, const uint8_t* T, uint64_t n, PrefixHash ph, uint32_t* lam)
{
    const uint32_t nb = ceil_div(n, kLynBlock);
    uint32_t* oa = S.buf<uint32_t>(B_LY_LIST0, (size_t)nb * kLynBlock);
    uint32_t* ob = S.buf<uint32_t>(B_LY_LIST1, (size_t)nb * kLynBlock);
    uint32_t* ma = S.buf<uint32_t>(B_LY_RES, (size_t)nb * kLynBlock);
    uint32_t* mb = S.buf<uint32_t>(B_LY_JMP0, (size_t)nb * kLynBlock);
    uint32_t* oca = S.buf<uint32_t>(B_LY_CNT0, nb);
    uint32_t* ocb = S.buf<uint32_t>(B_LY_CNT1, nb);
    uint32_t* mca = S.buf<uint32_t>(B_LY_HF, nb);
    uint32_t* mcb = S.buf<uint32_t>(B_LY_HF2, nb);
    const SuffixCmp sc{T, n, ph};
    k_nss_local<<<ceil_div(nb, 64), 64, 0, S.stream>>>(sc, nb, lam, oa, oca, ma, mca);
    MIBWT_LAUNCH_CHECK();
    int levels = 0;
    for (uint64_t size = 1; size < nb; size *= 2) {
        const uint32_t pairs = (uint32_t)((nb + 2 * size - 1) / (2 * size));
        k_nss_merge<<<pairs, 64, 0, S.stream>>>(sc, nb, (uint32_t)size, lam, oa, oca, ma, mca, ob, ocb, mb, mcb);
        MIBWT_LAUNCH_CHECK();
        std::swap(oa, ob);
        std::swap(oca, ocb);
        std::swap(ma, mb);
        std::swap(mca, mcb);
        levels++;
    }
    k_nss_to_lam<<<grid_for(n), kBlock, 0, S.stream>>>(lam, n);
    MIBWT_LAUNCH_CHECK();
    return levels;
}

// Prefix fingerprints pre[k] = (H(T[0..k]), base^(k+1)) mod 2^61-1 by one device scan.
HashPair* prefix_hashes(Storage& S, const uint8_t* T, uint64_t n, uint64_t base)
{
    HashPair* pre = S.buf<HashPair>(B_LY_HASH, n);
    k_hash_items<<<grid_for(n), kBlock, 0, S.stream>>>(T, n, base, pre);
    MIBWT_LAUNCH_CHECK();
    device_scan<HashPair, OpHash>(S, pre, pre, n, true, HashPair{0, 1}, OpHash{}, nullptr, B_LY_SCAN0);
    return pre;
}

// ------------------------------------------------------------------- 3. rule order

// Key of position i: the first `per` symbols of P_i as dense codes (0 past its end).
__global__ void k_lyn_initkey(const uint8_t* __restrict__ T, uint64_t n, const uint32_t* __restrict__ lam,
                              const uint32_t* __restrict__ code, int bits, int per, uint64_t* __restrict__ K,
                              uint32_t* __restrict__ V)
{
    __shared__ uint32_t cs[256];
    cs[threadIdx.x] = code[threadIdx.x];  // blockDim == 256
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t l = lam[i];
    uint64_t key = 0, w = 0;
    for (int k = 0; k < per; k++) {
        if ((k & 7) == 0) w = (i + k < n) ? load8(T, i + k) : 0ull;
        const uint64_t s = ((uint32_t)k < l) ? cs[(w >> (8 * (k & 7))) & 0xFF] : 0ull;
        key = (key << bits) | s;
    }
    K[i] = key;
    V[i] = (uint32_t)i;
}

// fingerprint of every label P_i (compared before any longest-common-prefix search: most tie
// groups of a collection are copies of one label)
__global__ void k_label_hash(PrefixHash ph, const uint32_t* __restrict__ lam, uint64_t n, uint64_t* __restrict__ F)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) F[i] = ph.sub((uint32_t)i, lam[i]);
}

// Common prefix of T[x..x+m) and T[y..y+m) by fingerprints, knowing that the first `known`
// symbols agree: the whole range first, then bisection over the candidate lengths.
__device__ __forceinline__ uint32_t label_lcp(const PrefixHash& ph, uint32_t x, uint32_t y, uint32_t known,
                                              uint32_t m)
{
    uint32_t lo = known < m ? known : m;
    if (ph.sub(x, m) == ph.sub(y, m)) return m;
    uint32_t hi = m - 1;  // largest candidate LCP
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo + 1) / 2;
        if (ph.sub(x, mid) == ph.sub(y, mid)) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct LabelCmp {
    const uint8_t* T;
    const uint32_t* lam;
    PrefixHash ph;
    const uint64_t* F;  // label fingerprints
    // -1 / 0 / +1 for P_x against P_y, knowing their first `known` symbols agree
    __device__ int operator()(uint32_t x, uint32_t y, uint32_t known) const
    {
        if (x == y) return 0;
        const uint32_t lx = lam[x], ly = lam[y];
        const uint32_t m = lx < ly ? lx : ly;
        if (lx == ly && F[x] == F[y]) return 0;
        const uint32_t lo = label_lcp(ph, x, y, known, m);
        if (lo < m) return T[x + lo] < T[y + lo] ? -1 : 1;
        return lx < ly ? -1 : (lx > ly ? 1 : 0);
    }
};

// Groups of the current label order: gid1 (1-based, inclusive scan of heads) and gofs.
__global__ void k_lab_heads_from_keys(const uint64_t* __restrict__ K, uint64_t n, uint32_t* __restrict__ hf)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    hf[i] = (i == 0 || K[i] != K[i - 1]) ? 1u : 0u;
}

// Per element: is its group a tie group (>= 2 members whose labels all reach past `known`
// symbols), and is that group small or large?  cls: 0 = settled, 1 = small tie, 2 = large tie.
__global__ void k_lab_classify(const uint32_t* __restrict__ V, const uint32_t* __restrict__ gid1,
                               const uint32_t* __restrict__ gofs, const uint32_t* __restrict__ lam, uint64_t n,
                               uint32_t known, uint32_t small, const uint8_t* __restrict__ settled,
                               uint8_t* __restrict__ cls, uint32_t* __restrict__ large)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t g = gid1[i] - 1;
    const uint32_t s = gofs[g], e = gofs[g + 1];
    uint8_t c = 0;
    if (e - s >= 2 && lam[V[s]] >= known && !settled[s]) c = (e - s > small) ? 2 : 1;
    cls[i] = c;
    large[i] = (c == 2) ? 1u : 0u;
}

// Small tie groups: comparator counting sort; settled groups are copied.  hf receives the
// heads of the refined order (first of every run of equal labels).
// label length and fingerprint of every position in the current order (the comparisons of a
// tie group then read them coalesced; most members of a collection's groups are equal labels)
__global__ void k_lab_gather(const uint32_t* __restrict__ V, const uint32_t* __restrict__ lam,
                             const uint64_t* __restrict__ F, uint64_t n, uint32_t* __restrict__ GL,
                             uint64_t* __restrict__ GF)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = V[i];
    GL[i] = lam[x];
    GF[i] = F[x];
}

__global__ void k_lab_small(const uint32_t* __restrict__ V, const uint32_t* __restrict__ gid1,
                            const uint32_t* __restrict__ gofs, const uint8_t* __restrict__ cls, uint64_t n,
                            uint32_t known, LabelCmp cmp, const uint32_t* __restrict__ GL,
                            const uint64_t* __restrict__ GF, uint32_t* __restrict__ V2, uint32_t* __restrict__ hf)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t g = gid1[i] - 1;
    const uint32_t s = gofs[g], e = gofs[g + 1];
    const uint32_t x = V[i];
    if (cls[i] == 2) return;  // large groups are written by the chunk pass
    if (cls[i] == 0) {
        V2[i] = x;
        hf[i] = (i == s) ? 1u : 0u;
        return;
    }
    const uint32_t lx = GL[i];
    const uint64_t fx = GF[i];
    uint32_t before = 0, equal_before = 0;
    for (uint32_t j = s; j < e; j++) {
        if (j == i) continue;
        const int c = (GL[j] == lx && GF[j] == fx) ? 0 : cmp(V[j], x, known);
        before += (c < 0) || (c == 0 && j < (uint32_t)i);
        equal_before += (c == 0 && j < (uint32_t)i);
    }
    V2[s + before] = x;
    hf[s + before] = (equal_before == 0) ? 1u : 0u;
}

// Large tie groups are split around a pivot, their first member p: a member x is compared
// with p through the fingerprint LCP l of the two labels (both agree on the first `known`
// symbols) and falls into one of three classes, LESS, EQUAL (the same label) or GREATER.
// Ordered inside its group, LESS members rise with l (they agree longer with p) and GREATER
// members fall with l; members with the same class, l and next symbol tie and form the
// sub-groups of the next round, EQUAL members form a settled group.  Key (low 11 + lbits
// bits): class (2) | l - known, clamped to lbits bits | next symbol + 1 (9, 0 = the label ends
// at l); the dense id of the large group goes above it (k_lab_pivot_ids).  Every round removes
// at least the pivot's class from each group, so a group of g labels needs at most g rounds,
// and a run such as a^k b (labels a^j b that all share long prefixes) is sorted in one.

__global__ void k_lab_pivot_gather(PrefixHash ph, const uint8_t* __restrict__ T, uint64_t n,
                                   const uint32_t* __restrict__ V, const uint32_t* __restrict__ gid1,
                                   const uint32_t* __restrict__ gofs, const uint32_t* __restrict__ large,
                                   const uint32_t* __restrict__ lpos, const uint32_t* __restrict__ lam, uint32_t known,
                                   int lbits, uint64_t* __restrict__ LK, uint32_t* __restrict__ LV, uint32_t* __restrict__ LI,
                                   uint32_t* __restrict__ LG)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !large[i]) return;
    const uint32_t g = gid1[i] - 1;
    const uint32_t x = V[i], p = V[gofs[g]];
    const uint32_t lx = lam[x], lp = lam[p];
    const uint32_t m = lx < lp ? lx : lp;
    const uint32_t l = (x == p) ? m : label_lcp(ph, x, p, known, m);
    uint64_t cls, lpart = 0, ch = 0;
    if (l == m && lx == lp) {
        cls = 1;  // EQUAL
    } else {
        const uint32_t cx = (l < lx) ? (uint32_t)T[x + l] + 1u : 0u;  // 0: x ends at l
        const uint32_t cp = (l < lp) ? (uint32_t)T[p + l] + 1u : 0u;
        const uint32_t lmax = (uint32_t)((1ull << lbits) - 1);
        const uint32_t d = (l > known) ? l - known : 0u;
        const uint32_t lc = d < lmax ? d : lmax;
        if (cx < cp) {
            cls = 0;  // LESS: rises with l
            lpart = lc;
        } else {
            cls = 2;  // GREATER: falls with l
            lpart = lmax - lc;
        }
        ch = cx;
    }
    const uint32_t q = lpos[i];
    LK[q] = (cls << (9 + lbits)) | (lpart << 9) | ch;
    LV[q] = x;
    LI[q] = (uint32_t)i;
    LG[q] = g;
}

// dense id of every large group, above the pivot key: flag the first member of each group
__global__ void k_lab_pivot_heads(const uint32_t* __restrict__ LG, uint32_t nl, uint32_t* __restrict__ hq)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nl) hq[q] = (q == 0 || LG[q] != LG[q - 1]) ? 1u : 0u;
}

__global__ void k_lab_pivot_ids(const uint32_t* __restrict__ id1, uint32_t nl, int low_bits, uint64_t* __restrict__ LK)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nl) LK[q] |= (uint64_t)(id1[q] - 1) << low_bits;
}

// Sorted large members back to the group slots (a group occupies the same slot range before and
// after), heads where the key changes, EQUAL sub-groups marked settled.
__global__ void k_lab_pivot_scatter(const uint64_t* __restrict__ LK, const uint32_t* __restrict__ LV,
                                    const uint32_t* __restrict__ LI, uint32_t cnt, int lbits, uint32_t* __restrict__ V2,
                                    uint32_t* __restrict__ hf, uint8_t* __restrict__ settled)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cnt) return;
    const uint32_t i = LI[j];
    V2[i] = LV[j];
    hf[i] = (j == 0 || LK[j] != LK[j - 1]) ? 1u : 0u;
    if (((LK[j] >> (9 + lbits)) & 3u) == 1u) settled[i] = 1;
}

// check of the final label order: neighbours must be non-decreasing, and equal exactly when
// they share a label group.  The common prefix of two labels is found through the fingerprints
// (O(log lcp), so long runs of equal labels cost no more than short ones); a fingerprint
// collision that slips through is caught by the exact inversion check of the finished BWT.
__global__ void k_lab_verify(SuffixCmp sc, const uint32_t* __restrict__ V, const uint32_t* __restrict__ hf,
                             const uint32_t* __restrict__ lam, uint32_t* __restrict__ bad)
{
    const uint64_t n = sc.n;
    const uint8_t* T = sc.T;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 || i >= n) return;
    const uint32_t a = V[i - 1], b = V[i];
    const uint32_t la = lam[a], lb = lam[b];
    const uint32_t m = la < lb ? la : lb;
    const uint32_t k = (a == b) ? m : sc.lcp_upto(a, b, m);
    int c;
    if (k < m) c = T[a + k] < T[b + k] ? -1 : 1;
    else c = la < lb ? -1 : (la > lb ? 1 : 0);
    const bool head = hf[i] != 0;
    if (c > 0 || (c == 0) == head) atomicOr(bad, 1u);
}

__global__ void k_u32_to_key(const uint32_t* __restrict__ a, uint64_t n, uint64_t* __restrict__ K)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) K[i] = a[i];
}

// ------------------------------------------------------------------- 4. induction

struct ChainPolicy {
    static constexpr bool kCarry = false;
    static constexpr bool kSpeculate = false;   // advancing launches pointer jumping over the count
    static constexpr bool kAdvanceFirst = true;  // step only the elements that stay active
    const uint32_t* jmp;
    uint32_t n;
    __device__ __forceinline__ uint32_t rem(uint32_t) const { return 0; }
    __device__ __forceinline__ uint32_t next(uint32_t x, uint32_t) const
    {
        const uint32_t y = jmp[x];
        return y < n ? y : kNoNext;
    }
    __device__ __forceinline__ bool open(uint32_t, uint32_t) const { return true; }
    __device__ __forceinline__ void place(uint32_t* SA, uint32_t pos, uint32_t x, uint32_t) const { SA[pos] = x; }
};

__global__ void k_jump_init(const uint32_t* __restrict__ lam, uint64_t n, uint32_t* __restrict__ a,
                            uint32_t* __restrict__ b)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = (uint32_t)(i + lam[i]);
    a[i] = j;
    b[i] = j;
}

// pointer jumping for the active elements: jmp2 = jmp o jmp
__global__ void k_jump(const uint32_t* __restrict__ act, uint32_t m, const uint32_t* __restrict__ jmp,
                       uint32_t* __restrict__ jmp2, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t x = act[i];
    const uint32_t y = jmp[x];
    jmp2[x] = (y < n) ? jmp[y] : n;
}

// ------------------------------------------------------------------- 5. output

__global__ void k_bwt_from_sa(const uint32_t* __restrict__ SA, const uint8_t* __restrict__ T, uint64_t n,
                              uint8_t* __restrict__ O, uint32_t* __restrict__ pidx)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const uint32_t p = SA[g];
    if (p >= n) return;  // never for a permutation; an unwritten byte fails the inversion check
    O[g] = T[p == 0 ? n - 1 : p - 1];
    if (p == 0) *pidx = (uint32_t)g + 1;
}

// ------------------------------------------------------------------- driver

// Fingerprint base in [2^32, 2^61 - 1) drawn from a seed.
static uint64_t lyndon_base(uint64_t seed)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (1ull << 32) + z % (kM61 - (1ull << 32));
}

// One build attempt with fingerprint base `base` (fingerprints compared under `hmask`): the
// phases rules .. output.  Returns the primary index, or 0 when the label order check found a
// fingerprint collision.
static uint32_t lyndon_attempt(Storage& S, const uint8_t* T, uint64_t n, uint64_t base, uint64_t hmask,
                               uint32_t* codes, uint32_t* sigma_io, uint8_t* dO)
{
    uint32_t* sc = S.buf<uint32_t>(B_SCALARS, SC_COUNT_);
    // ---- 1. prefix fingerprints
    int ph = S.phase_begin("rules");
    HashPair* pre = prefix_hashes(S, T, n, base);
    const PrefixHash phh{pre, hmask};
    S.phase_end(ph, (double)n * (1 + 16 * 4));

    // ---- 2. Lyndon array
    ph = S.phase_begin("lyndon");
    uint32_t* lam = S.buf<uint32_t>(B_LY_LAM, n);
    const int levels = lyndon_array(S, T, n, phh, lam);
    alphabet_codes(S, T, n, codes);
    uint32_t sigma = 0;
    S.read(&sigma, codes + 256);
    *sigma_io = sigma;
    S.phase_end(ph, (double)n * (1 + 4 + 4 + 1) * 2);
    S.stats.doubling_rounds_dict = levels;

    // ---- 3. label order
    ph = S.phase_begin("rsort");
    const int bits = bits_for((uint64_t)sigma + 1);
    const int per = 64 / bits;
    uint64_t* K0 = S.buf<uint64_t>(B_SORT_K0, n);
    uint64_t* K1 = S.buf<uint64_t>(B_SORT_K1, n);
    uint32_t* V0 = S.buf<uint32_t>(B_SORT_V0, n);
    uint32_t* V1 = S.buf<uint32_t>(B_SORT_V1, n);
    uint32_t* hf = S.buf<uint32_t>(B_LY_HF, n);
    uint32_t* gid1 = S.buf<uint32_t>(B_DB_GID, n);
    uint32_t* gofs = S.buf<uint32_t>(B_DB_GOFS, n + 1);
    uint32_t* large = S.buf<uint32_t>(B_DB_LARGE, n);
    uint32_t* lpos = S.buf<uint32_t>(B_ACT_X, n);
    uint8_t* cls = S.buf<uint8_t>(B_LY_CLS, n);
    uint32_t* Vl = S.buf<uint32_t>(B_LY_VL, n);
    k_lyn_initkey<<<grid_for(n), kBlock, 0, S.stream>>>(T, n, lam, codes, bits, per, K0, V0);
    MIBWT_LAUNCH_CHECK();
    radix_sort<uint64_t, uint32_t>(S, K0, K1, V0, V1, n, 0, bits * per);
    k_lab_heads_from_keys<<<grid_for(n), kBlock, 0, S.stream>>>(K0, n, hf);
    MIBWT_LAUNCH_CHECK();
    double rbytes = (double)n * 24 * ((bits * per + 7) / 8);
    {
        uint64_t* LF = S.buf<uint64_t>(B_LY_LABH, n);
        k_label_hash<<<grid_for(n), kBlock, 0, S.stream>>>(phh, lam, n, LF);
        MIBWT_LAUNCH_CHECK();
        LabelCmp cmp{T, lam, phh, LF};
        // working copies: V (order) and hf (heads) of the packed-key order
        uint32_t* V = Vl;
        MIBWT_HIP(hipMemcpyAsync(V, V0, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, S.stream));
        uint32_t* H = S.buf<uint32_t>(B_LY_HF2, n);
        MIBWT_HIP(hipMemcpyAsync(H, hf, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, S.stream));
        const uint32_t known = (uint32_t)per;  // every tie group agrees on its first `per` symbols
        uint32_t small = kLabSmall;
        if (const char* e = getenv("MIBWT_LAB_SMALL")) {
            const int v = atoi(e);
            if (v >= 2 && v <= 256) small = (uint32_t)v;
        }
        uint8_t* settled = S.buf<uint8_t>(B_LY_SETTLED, n);
        fill<uint8_t>(S, settled, n, (uint8_t)0);
        for (int round = 0;; round++) {
            MIBWT_REQUIRE(round < 1 << 20, LIBCUBWT_GPU_ERROR, "label refinement does not converge");
            inclusive_sum<uint32_t>(S, H, gid1, n, sc + SC_GROUPS);
            k_grp_offsets<<<grid_for(n), kBlock, 0, S.stream>>>(H, gid1, (uint32_t)n, gofs);
            MIBWT_LAUNCH_CHECK();
            k_lab_classify<<<grid_for(n), kBlock, 0, S.stream>>>(V, gid1, gofs, lam, n, known, small, settled, cls,
                                                                 large);
            MIBWT_LAUNCH_CHECK();
            exclusive_sum<uint32_t>(S, large, lpos, n, sc + SC_LARGE);
            uint32_t hv[SC_COUNT_];
            S.read(hv, sc, SC_COUNT_);
            const uint32_t nl = hv[SC_LARGE];
            uint32_t* V2 = (V == Vl) ? V1 : Vl;
            uint32_t* GL = S.buf<uint32_t>(B_LY_GL, n);
            uint64_t* GF = S.buf<uint64_t>(B_LY_GF, n);
            k_lab_gather<<<grid_for(n), kBlock, 0, S.stream>>>(V, lam, LF, n, GL, GF);
            MIBWT_LAUNCH_CHECK();
            k_lab_small<<<grid_for(n), kBlock, 0, S.stream>>>(V, gid1, gofs, cls, n, known, cmp, GL, GF, V2, H);
            MIBWT_LAUNCH_CHECK();
            if (nl > 0) {
                uint64_t* LK0 = S.buf<uint64_t>(B_LG_K0, nl);
                uint64_t* LK1 = S.buf<uint64_t>(B_LG_K1, nl);
                uint32_t* LV0 = S.buf<uint32_t>(B_LG_V0, nl);
                uint32_t* LV1 = S.buf<uint32_t>(B_LG_V1, nl);
                uint32_t* LI = S.buf<uint32_t>(B_LG_I, nl);
                uint32_t* LG = S.buf<uint32_t>(B_LG_SZ, nl);
                uint32_t* hq = S.buf<uint32_t>(B_LY_SCAN0, nl);
                // large groups hold more than `small` members each, so their dense ids take gb
                // bits; the l part takes what is left of the 64-bit key (at most 31 bits)
                const int gb = bits_for((uint64_t)nl / (small + 1));
                const int lbits = std::min(31, 64 - 11 - gb);
                k_lab_pivot_gather<<<grid_for(n), kBlock, 0, S.stream>>>(phh, T, n, V, gid1, gofs, large, lpos, lam,
                                                                         known, lbits, LK0, LV0, LI, LG);
                MIBWT_LAUNCH_CHECK();
                k_lab_pivot_heads<<<grid_for(nl), kBlock, 0, S.stream>>>(LG, nl, hq);
                MIBWT_LAUNCH_CHECK();
                inclusive_sum<uint32_t>(S, hq, hq, nl, sc + SC_GROUPS);
                k_lab_pivot_ids<<<grid_for(nl), kBlock, 0, S.stream>>>(hq, nl, 11 + lbits, LK0);
                MIBWT_LAUNCH_CHECK();
                radix_sort<uint64_t, uint32_t>(S, LK0, LK1, LV0, LV1, nl, 0, 11 + lbits + gb);
                k_lab_pivot_scatter<<<grid_for(nl), kBlock, 0, S.stream>>>(LK0, LV0, LI, nl, lbits, V2, H, settled);
                MIBWT_LAUNCH_CHECK();
                rbytes += (double)nl * 24 * 8;
            }
            V = V2;
            if (nl == 0) break;
        }
        // check of the label order
        fill<uint32_t>(S, sc + SC_COLLISION, 1, 0u);
        k_lab_verify<<<grid_for(n), kBlock, 0, S.stream>>>(SuffixCmp{T, n, phh}, V, H, lam, sc + SC_COLLISION);
        MIBWT_LAUNCH_CHECK();
        uint32_t bad = 0;
        S.read(&bad, sc + SC_COLLISION);
        if (bad != 0) {
            S.phase_end(ph, rbytes);
            return 0;
        }
        if (V != Vl) MIBWT_HIP(hipMemcpyAsync(Vl, V, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, S.stream));
        MIBWT_HIP(hipMemcpyAsync(hf, H, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, S.stream));
    }
    // label group ids as the initial keys of the induction
    inclusive_sum<uint32_t>(S, hf, gid1, n, sc + SC_GROUPS);
    k_u32_to_key<<<grid_for(n), kBlock, 0, S.stream>>>(gid1, n, K0);
    MIBWT_LAUNCH_CHECK();
    uint32_t ngroups = 0;
    S.read(&ngroups, sc + SC_GROUPS);
    S.phase_end(ph, rbytes);
    S.stats.distinct_phrases = ngroups;

    // ---- 4. induction along the next-smaller-suffix chain
    ph = S.phase_begin("induce");
    uint32_t* jA = S.buf<uint32_t>(B_LY_JMP0, n);
    uint32_t* jB = S.buf<uint32_t>(B_LY_JMP1, n);
    k_jump_init<<<grid_for(n), kBlock, 0, S.stream>>>(lam, n, jA, jB);
    MIBWT_LAUNCH_CHECK();
    uint32_t* rank = S.buf<uint32_t>(B_DT_RANK, n);
    uint32_t* SA = S.buf<uint32_t>(B_DT_SA, n);
    ChainPolicy pol{jA, (uint32_t)n};
    uint32_t* jcur = jA;
    uint32_t* jnext = jB;
    auto advance = [&](ChainPolicy& p, const uint32_t* act, uint32_t m) {
        k_jump<<<grid_for(m), kBlock, 0, S.stream>>>(act, m, jcur, jnext, (uint32_t)n);
        MIBWT_LAUNCH_CHECK();
        std::swap(jcur, jnext);
        p.jmp = jcur;
    };
    // pointers double per round, so after 34 rounds every chain has left the text: a group that
    // is still active then holds equal elements, which only inconsistent (colliding) label
    // fingerprints produce
    DoublingResult rr;
    try {
        rr = refine_groups(S, K0, Vl, V1, (uint32_t)n, pol, advance, rank, SA, 40);
    } catch (const Error& e) {
        if (e.code != MIBWT_HASH_COLLISION) throw;
        MIBWT_HIP(hipStreamSynchronize(S.stream));
        S.phase_end(ph, 0);
        return 0;
    }
    S.phase_end(ph, rr.bytes);
    S.stats.doubling_rounds_parse = rr.rounds;

    // ---- 5. output
    ph = S.phase_begin("output");
    fill<uint32_t>(S, sc + SC_PIDX, 1, 0u);
    k_bwt_from_sa<<<grid_for(n), kBlock, 0, S.stream>>>(SA, T, n, dO, sc + SC_PIDX);
    MIBWT_LAUNCH_CHECK();
    S.phase_end(ph, (double)n * 6);
    uint32_t pidx = 0;
    S.read(&pidx, sc + SC_PIDX);
    return (pidx >= 1 && pidx <= n) ? pidx : 0xFFFFFFFFu;  // an invalid index fails the inversion
}


int64_t lyndon_bwt_run(Storage& S, const uint8_t* src, bool src_host, uint8_t* dO, int64_t n64)
{
    MIBWT_REQUIRE(n64 >= 1 && n64 < (int64_t)INT32_MAX - (1 << 20), LIBCUBWT_BAD_PARAMETER, "n out of range");
    const uint64_t n = (uint64_t)n64;
    S.phases_reset();
    S.stats.n = n64;
    apply_tuning(S, n);
    if (!S.t_begin) {
        MIBWT_HIP(hipEventCreate(&S.t_begin));
        MIBWT_HIP(hipEventCreate(&S.t_end));
    }
    uint8_t* T;
    if (src_host) {
        T = stage_text(S, src, true, n);
        MIBWT_HIP(hipEventRecord(S.t_begin, S.stream));
    } else {
        MIBWT_HIP(hipEventRecord(S.t_begin, S.stream));
        T = stage_text(S, src, false, n);
    }

    // The suffix comparisons behind the Lyndon array and the label order trust Karp-Rabin
    // fingerprints; the finished BWT is therefore inverted on the device and compared with the
    // text, and a mismatch (a fingerprint collision) repeats the build with a new base.  The base
    // is drawn per call.  MIBWT_LYN_HASH_BITS narrows the first attempt's fingerprints (tests).
    uint64_t mask0 = ~0ull;
    if (const char* hb = getenv("MIBWT_LYN_HASH_BITS")) {
        const int v = atoi(hb);
        if (v >= 1 && v < 64) mask0 = (1ull << v) - 1;
    }
    S.lyn_calls++;
    uint64_t base = lyndon_base(S.lyn_calls ^ (uint64_t)std::chrono::steady_clock::now().time_since_epoch().count());
    uint32_t* codes = S.buf<uint32_t>(B_ALPHA, 272);
    uint32_t sigma = 0;
    uint32_t pidx = 0;
    for (int attempt = 0;; attempt++) {
        MIBWT_REQUIRE(attempt < 4, MIBWT_HASH_COLLISION, "Lyndon fingerprint collisions persist");
        if (attempt > 0) base = lyndon_base(base * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull);
        const uint64_t hmask = (attempt == 0) ? mask0 : ~0ull;
        S.stats.retries = attempt;  // the timing covers every attempt
        pidx = lyndon_attempt(S, T, n, base, hmask, codes, &sigma, dO);
        if (pidx == 0) continue;  // the label order check failed: new base
        const int ph = S.phase_begin("verify");
        const bool ok = bwt_inverts_to(S, dO, (int64_t)n, pidx, T);
        S.phase_end(ph, (double)n * 40);
        if (ok) break;
    }
    MIBWT_HIP(hipEventRecord(S.t_end, S.stream));
    S.phases_collect();
    float tot = 0;
    MIBWT_HIP(hipEventElapsedTime(&tot, S.t_begin, S.t_end));
    S.stats.total_ms = tot;
    return (int64_t)pidx;
}

// Lyndon array of a host text (diagnostics and parity tests of the local/merge passes).
int64_t lyndon_array_run(Storage& S, const uint8_t* src, int64_t n64, uint32_t* lam_host)
{
    MIBWT_REQUIRE(n64 >= 1 && n64 < (int64_t)INT32_MAX - (1 << 20), LIBCUBWT_BAD_PARAMETER, "n out of range");
    const uint64_t n = (uint64_t)n64;
    uint8_t* T = stage_text(S, src, true, n);
    uint32_t* lam = S.buf<uint32_t>(B_LY_LAM, n);
    HashPair* pre = prefix_hashes(S, T, n, 0x1F3D5B79A2C4E681ull % kM61);
    const int levels = lyndon_array(S, T, n, PrefixHash{pre, ~0ull}, lam);
    MIBWT_HIP(hipMemcpyAsync(lam_host, lam, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, S.stream));
    MIBWT_HIP(hipStreamSynchronize(S.stream));
    return levels;
}

}  // namespace mibwt
