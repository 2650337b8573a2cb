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
//  2. lyndon   Lyndon array as next-smaller-suffix distances: every 64-byte block chases
//              NSS pointers right to left over its text staged in LDS; positions without a smaller suffix in the block
//              stay open and are resolved by a merge tree over neighbouring ranges, where an
//              open position bisects the right range's prefix minima.  Suffix comparisons
//              use the fingerprints for long common prefixes (O(log lcp) per comparison).
//  3. labels   the grammar: one 64-bit key per position (short labels by their symbols, long
//              ones by fingerprint), one radix sort, the distinct labels with their counts, and
//              for each distinct label its children (the Lyndon factorization of the label
//              without its first symbol), runs of equal children and up-contexts.
//  4. rsort    the distinct labels in lexicographic order: radix sort of the first symbols, tie
//              groups ordered with a fingerprint-LCP comparator, checked pair by pair.
//  5. induce   induction over rule rank groups: items are child slots of distinct labels, each
//              standing for all occurrences of its label; groups are refined by the next run of
//              labels, and a group whose items share their preceding character fills its BWT
//              range with that character at once (see section 4).
//  6. verify   the BWT is inverted on the device and compared with the text; a mismatch (a
//              fingerprint collision) repeats the build with a new fingerprint base.
// The position-level induction (every position ordered by its label, groups refined along the
// next-smaller-suffix chain with pointer jumping) remains as the fallback when the grammar
// induction outgrows its budget.
#include <chrono>
#include <cstdlib>

#include "doubling.hpp"
#include "text.hpp"

namespace mibwt {

bool bwt_inverts_to(Storage& S, const uint8_t* dO, int64_t n, int64_t index, const uint8_t* dT);

// Characters per block of the local pass (MIBWT_LYN_BLOCK overrides: 16..256).  Measured on the
// 1 GiB DNA collection (local pass + merge tree with eight lanes per pair on the wide levels):
// 32 -> 25.4 ms, 64 -> 21.4 ms, 128 -> 23.0 ms, 256 -> 43 ms (the workgroups' LDS, 10 / 18 / 34 KiB
// at 64 / 128 / 256, bounds the resident wavefronts of this latency-bound pass).
constexpr uint32_t kLynBlock = 64;
// Label tie groups of more members go to the pivot partition (linear work per round; a group of
// equal labels settles in one round) instead of the comparator counting sort (quadratic in the
// group size, but one pass); MIBWT_LAB_SMALL overrides it (2..256).  Ordering every position
// (many copies of each label), measured on the 256 MiB DNA collection: 256 -> rsort 83 ms,
// 64 -> 97 ms, 16 -> 148 ms, 4 -> 174 ms (more pivot rounds, each a radix sort, outweigh the
// cheaper comparisons).  Ordering distinct labels (no equal members, so every pivot round
// splits well): 4 -> 5.7 ms, 8 -> 5.8, 16 -> 7.0, 32 -> 9.2, 256 -> 11.6.
constexpr uint32_t kLabSmall = 256;      // every position
constexpr uint32_t kLabSmallDistinct = 4;  // distinct labels
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

// Prefix fingerprints straight from the text, reduce-then-scan with 16 bytes per thread (one
// 16-byte load): a run of bytes c_1..c_k hashes to (sum (c_i + 1) base^(k-i), base^k), so each
// byte costs one modular multiply while the thread runs over it, and runs combine with OpHash.
constexpr int kPhItems = 16;
constexpr int kPhTile = kBlock * kPhItems;

__device__ __forceinline__ HashPair ph_run(const uint8_t* __restrict__ T, uint64_t n, uint64_t b0, uint64_t base,
                                           uint64_t base16, uint4& w)
{
    w = *reinterpret_cast<const uint4*>(T + b0);  // the padded copy reads zeros past n
    const uint32_t cnt = (b0 >= n) ? 0u : (uint32_t)((n - b0 < (uint64_t)kPhItems) ? n - b0 : kPhItems);
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
    uint64_t h = 0, p = 1;
#pragma unroll
    for (uint32_t i = 0; i < (uint32_t)kPhItems; i++) {
        if (i < cnt) {
            h = addmod61(mulmod61(h, base), (uint64_t)((ws[i >> 2] >> (8 * (i & 3))) & 0xFFu) + 1);
            if (cnt < (uint32_t)kPhItems) p = mulmod61(p, base);
        }
    }
    return HashPair{h, cnt == (uint32_t)kPhItems ? base16 : p};
}

__global__ __launch_bounds__(kBlock) void k_ph_reduce(const uint8_t* __restrict__ T, uint64_t n, uint64_t base,
                                                     uint64_t base16, HashPair* __restrict__ partial)
{
    __shared__ HashPair lds[kWaves];
    uint4 w;
    const HashPair acc = ph_run(T, n, (uint64_t)blockIdx.x * kPhTile + (uint64_t)threadIdx.x * kPhItems, base, base16, w);
    HashPair tot;
    (void)block_exclusive_scan(acc, HashPair{0, 1}, OpHash{}, lds, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// pre[k] = H(T[0..k]) from the tiles' exclusive prefixes.  Each thread's values go through LDS
// (rows of kPhItems + 1 words against bank conflicts) so that the tile leaves as one coalesced
// write.
__global__ __launch_bounds__(kBlock) void k_ph_down(const uint8_t* __restrict__ T, uint64_t n, uint64_t base,
                                                   uint64_t base16, const HashPair* __restrict__ partial,
                                                   uint64_t* __restrict__ pre)
{
    __shared__ HashPair lds[kWaves];
    __shared__ uint64_t out[kBlock * (kPhItems + 1)];
    const uint64_t t0 = (uint64_t)blockIdx.x * kPhTile;
    const uint64_t b0 = t0 + (uint64_t)threadIdx.x * kPhItems;
    uint4 w;
    const HashPair acc = ph_run(T, n, b0, base, base16, w);
    HashPair tot;
    HashPair run = block_exclusive_scan(acc, HashPair{0, 1}, OpHash{}, lds, &tot);
    run = OpHash{}(partial[blockIdx.x], run);
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (uint32_t i = 0; i < (uint32_t)kPhItems; i++) {
        run.h = addmod61(mulmod61(run.h, base), (uint64_t)((ws[i >> 2] >> (8 * (i & 3))) & 0xFFu) + 1);
        out[threadIdx.x * (kPhItems + 1) + i] = run.h;
    }
    __syncthreads();
    const uint32_t valid = (n - t0 < (uint64_t)kPhTile) ? (uint32_t)(n - t0) : (uint32_t)kPhTile;
    for (uint32_t q = threadIdx.x; q < valid; q += kBlock) pre[t0 + q] = out[(q / kPhItems) * (kPhItems + 1) + q % kPhItems];
}

// Powers of the base in two tables of 2^16 entries: base^i and base^(2^16 i), so base^len for
// any 32-bit len is one entry or the product of two (1 MiB in all, cache resident).
constexpr uint32_t kPowBits = 16;
constexpr uint32_t kPowSize = 1u << kPowBits;

struct PrefixHash {
    const uint64_t* pre;  // pre[k] = H(T[0..k])
    const uint64_t* pw;   // pw[i] = base^i, pw[kPowSize + j] = base^(kPowSize j)
    uint64_t mask;        // bits of the fingerprints compared (all 61 but in collision tests)
    __device__ __forceinline__ uint64_t power(uint32_t len) const
    {
        const uint64_t lo = pw[len & (kPowSize - 1)];
        return (len < kPowSize) ? lo : mulmod61(lo, pw[kPowSize + (len >> kPowBits)]);
    }
    // fingerprint of T[a..a+len)
    __device__ __forceinline__ uint64_t sub(uint32_t a, uint32_t len) const
    {
        if (len == 0) return 0;
        const uint64_t hb = pre[a + len - 1];
        if (a == 0) return hb & mask;
        const uint64_t x = mulmod61(pre[a - 1], power(len));
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

// Local pass: one thread per block of bs <= kNssMaxBlock characters, 64 blocks per workgroup.
// The workgroup stages its blocks' text in LDS, each block followed by the first kNssLook bytes
// of the next one (one coalesced read of the text), and keeps each block's NSS values as 8-bit
// offsets (0: no smaller suffix inside the block) in LDS; the neighbours compared inside a block
// almost always differ within 16 bytes, so a comparison is a few LDS reads, and only 16 equal
// bytes fall back to the global comparator.  Per-thread rows are padded to an odd number of
// words so that the 64 lanes' rows start in different LDS banks.
constexpr uint32_t kNssMaxBlock = 256;
constexpr uint32_t kNssLook = 16;
__host__ __device__ constexpr uint32_t nss_text_row(uint32_t bs) { return bs + kNssLook + 4; }
__host__ __device__ constexpr uint32_t nss_off_row(uint32_t bs) { return bs + 4; }

// 8 bytes of an LDS row from byte offset o (rows are word aligned and hold o + 12 bytes)
__device__ __forceinline__ uint64_t lds_read8(const uint8_t* row, uint32_t o)
{
    const uint32_t* w = reinterpret_cast<const uint32_t*>(row + (o & ~3u));
    const uint32_t sh = o & 3u, w0 = w[0], w1 = w[1], w2 = w[2];
    return ((uint64_t)__builtin_amdgcn_alignbyte(w2, w1, sh) << 32) | __builtin_amdgcn_alignbyte(w1, w0, sh);
}

// One thread per block, right to left: NSS by pointer chasing inside the block; lam[i] receives
// NSS(i) - i.  Positions without a smaller suffix inside the block stay open (kNoNss, set by the
// merge tree or, for the positions still open at its end, by k_nss_last).  Per block: the open positions
// in ascending order (O, increasing suffixes) and the chain of prefix minima from the block start
// (M = s, NSS[s], NSS[NSS[s]], ..., decreasing suffixes).  Launched with 64 threads and
// 64 * (nss_text_row(bs) + nss_off_row(bs)) bytes of dynamic LDS; bs is a power of two in [16, 256].
__global__ __launch_bounds__(64) void k_nss_local(SuffixCmp sc, uint32_t nb, uint32_t bs, int lbs,
                                                  uint32_t* __restrict__ lam, uint32_t* __restrict__ O,
                                                  uint32_t* __restrict__ oc, uint32_t* __restrict__ M,
                                                  uint32_t* __restrict__ mc)
{
    extern __shared__ __align__(16) uint8_t nss_lds[];
    const uint32_t TS = nss_text_row(bs), NS = nss_off_row(bs);
    uint8_t* tx = nss_lds;
    uint8_t* of = nss_lds + 64 * TS;
    const uint32_t lane = threadIdx.x;
    const uint64_t b0 = (uint64_t)blockIdx.x * 64;
    const uint64_t s0 = b0 * bs;
    // stage T[s0, s0 + 64 bs + kNssLook) in 16-byte pieces (every piece that holds text bytes
    // lies inside the kPad zero bytes after the text)
    const uint32_t span = 64 * bs + kNssLook;
    for (uint32_t q = lane * 16; q < span; q += 64 * 16) {
        uint4 v = make_uint4(0, 0, 0, 0);
        if (s0 + q + 16 <= sc.n + kPad) v = *reinterpret_cast<const uint4*>(sc.T + s0 + q);
        const uint32_t t = q >> lbs, k = q & (bs - 1);
        if (t < 64) {
            uint32_t* d = reinterpret_cast<uint32_t*>(tx + t * TS + k);
            d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
        }
        if (t >= 1 && k == 0) {  // the previous block's look-ahead
            uint32_t* d = reinterpret_cast<uint32_t*>(tx + (t - 1) * TS + bs);
            d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
        }
    }
    __syncthreads();
    const uint64_t b = b0 + lane;
    const uint8_t* mt = tx + lane * TS;
    uint8_t* mo = of + lane * NS;
    uint32_t len = 0;
    if (b < nb) {
        const uint64_t s = b * bs;
        len = (uint32_t)((s + bs < sc.n) ? bs : sc.n - s);
        // One flat loop of single steps (a comparison, or the end of a walk) over the whole block,
        // so that a lane whose position needs a long walk does not hold the others back at
        // every position: the lanes' total step counts are nearly equal.
        uint32_t k = len - 1, jk = len;
        for (;;) {
            bool done = jk >= len;  // no smaller suffix inside the block
            uint32_t off = 0;
            if (!done) {
                // suffix(s + jk) < suffix(s + k)?  m = comparable length = n - (s + jk)
                const uint64_t m = sc.n - (s + jk);
                int res = -1;  // -1: undecided within the staged bytes
#pragma unroll
                for (uint32_t l0 = 0; l0 < kNssLook; l0 += 8) {
                    if (res < 0) {
                        const uint64_t xi = lds_read8(mt, k + l0), xj = lds_read8(mt, jk + l0);
                        const uint64_t d = xi ^ xj;
                        const uint32_t l = l0 + (d ? (uint32_t)(__builtin_ctzll(d) >> 3) : 8u);
                        if (l >= m) res = 1;  // suffix s + jk ends first: it is smaller
                        else if (d) {
                            const uint32_t sh = 8 * (l - l0);
                            res = ((xj >> sh) & 0xFFu) < ((xi >> sh) & 0xFFu) ? 1 : 0;
                        }
                    }
                }
                if (res < 0) res = sc.less((uint32_t)(s + jk), (uint32_t)(s + k)) ? 1 : 0;
                if (res) {
                    off = jk - k;
                    done = true;
                } else {
                    const uint32_t o2 = mo[jk];
                    if (o2 == 0) done = true;
                    else jk += o2;
                }
            }
            if (done) {
                mo[k] = (uint8_t)off;
                if (k == 0) break;
                k--;
                jk = k + 1;
            }
        }
    }
    __syncthreads();
    // NSS distances (lam) of the workgroup's positions, coalesced; kNoNss for the open ones
    const uint64_t tot = (sc.n - s0 < (uint64_t)64 * bs) ? sc.n - s0 : (uint64_t)64 * bs;
    for (uint32_t q = lane; q < tot; q += 64) {
        const uint32_t o = of[(q >> lbs) * NS + (q & (bs - 1))];
        lam[s0 + q] = o ? o : kNoNss;
    }
    if (b >= nb) return;
    const uint32_t s = (uint32_t)(b * bs);
    uint32_t c = 0;
    for (uint32_t k = 0; k < len; k++)
        if (mo[k] == 0) O[s + c++] = s + k;
    oc[b] = c;
    c = 0;
    for (uint32_t k = 0;;) {
        M[s + c++] = s + k;
        const uint32_t o = mo[k];
        if (o == 0) break;
        k += o;
    }
    mc[b] = c;
}

// G lanes (8 or 64) per pair of neighbouring ranges L, R of `size` blocks, 64 / G pairs per
// wavefront.  An open position i of L finds its NSS among R's prefix minima: they decrease, so the
// first one below suffix(i) is found by bisection (lam[i] = NSS(i) - i).  The merged range keeps L's unresolved
// positions followed by R's open ones, and L's prefix minima followed by the prefix minima of R
// below min(L).  The lists are short on most texts (a handful of entries per range), so the
// many pairs of the low levels share wavefronts.
template <int G>
__global__ __launch_bounds__(64) void k_nss_merge(SuffixCmp sc, uint32_t nb, uint32_t bs, uint32_t size,
                                                  uint32_t* __restrict__ lam, const uint32_t* __restrict__ Oin,
                                                  const uint32_t* __restrict__ ocin, const uint32_t* __restrict__ Min,
                                                  const uint32_t* __restrict__ mcin, uint32_t* __restrict__ Oout,
                                                  uint32_t* __restrict__ ocout, uint32_t* __restrict__ Mout,
                                                  uint32_t* __restrict__ mcout)
{
    static_assert(G == 8 || G == 64, "lanes per pair");
    const int lane = threadIdx.x & (G - 1);
    const int grp = threadIdx.x / G;
    const uint64_t a = ((uint64_t)blockIdx.x * (64 / G) + grp) * 2 * size;
    if (a >= nb) return;
    const uint64_t rb = a + size;
    const uint64_t sL = a * bs;
    const uint32_t oL = ocin[a], mL = mcin[a];
    const uint32_t* OL = Oin + sL;
    const uint32_t* ML = Min + sL;
    if (rb >= nb) {
        for (uint32_t q = lane; q < oL; q += G) Oout[sL + q] = OL[q];
        for (uint32_t q = lane; q < mL; q += G) Mout[sL + q] = ML[q];
        if (lane == 0) {
            ocout[a] = oL;
            mcout[a] = mL;
        }
        return;
    }
    const uint64_t sR = rb * bs;
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
    // L's open positions: resolve or keep (ordered compaction within the group's lanes)
    const uint32_t gshift = (uint32_t)(grp * G);
    const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << gshift);
    uint32_t kept = 0;
    for (uint32_t q0 = 0; q0 < oL; q0 += G) {
        const uint32_t q = q0 + lane;
        bool keep = false;
        if (q < oL) {
            const uint32_t i = OL[q];
            const uint32_t t = first_below(i);
            if (t < mR) lam[i] = MR[t] - i;
            else keep = true;
        }
        const uint64_t bal = __ballot(keep) & gmask;
        if (keep) Oout[sL + kept + (uint32_t)__popcll(bal & lanemask_lt())] = OL[q];
        kept += (uint32_t)__popcll(bal);
    }
    for (uint32_t q = lane; q < oR; q += G) Oout[sL + kept + q] = OR[q];
    // prefix minima: all of L's, then R's below min(L) = the last of L's
    const uint32_t tstar = first_below(ML[mL - 1]);  // every lane of the group computes the same value
    for (uint32_t q = lane; q < mL; q += G) Mout[sL + q] = ML[q];
    for (uint32_t q = tstar + lane; q < mR; q += G) Mout[sL + mL + (q - tstar)] = MR[q];
    if (lane == 0) {
        ocout[a] = kept + oR;
        mcout[a] = mL + (mR - tstar);
    }
}

// the positions still open after the last level (no smaller suffix follows): lam = n - i
__global__ void k_nss_last(const uint32_t* __restrict__ O, const uint32_t* __restrict__ oc, uint64_t n,
                           uint32_t* __restrict__ lam)
{
    const uint32_t m = oc[0];
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < m; q += gridDim.x * blockDim.x)
        lam[O[q]] = (uint32_t)n - O[q];
}

// lam[] of T[0..n) (the padded text copy T, prefix fingerprints ph).  Returns the number of
// merge levels.
int lyndon_array(Storage& S, const uint8_t* T, uint64_t n, PrefixHash ph, uint32_t* lam)
{
    uint32_t bs = kLynBlock;
    if (const char* e = getenv("MIBWT_LYN_BLOCK")) {
        const int v = atoi(e);
        if (v >= 16 && v <= (int)kNssMaxBlock && (v & (v - 1)) == 0) bs = (uint32_t)v;
    }
    int lbs = 0;
    while ((1u << lbs) < bs) lbs++;
    const uint32_t nb = ceil_div(n, bs);
    uint32_t* oa = S.buf<uint32_t>(B_LY_LIST0, (size_t)nb * bs);
    uint32_t* ob = S.buf<uint32_t>(B_LY_LIST1, (size_t)nb * bs);
    uint32_t* ma = S.buf<uint32_t>(B_LY_RES, (size_t)nb * bs);
    uint32_t* mb = S.buf<uint32_t>(B_LY_JMP0, (size_t)nb * bs);
    uint32_t* oca = S.buf<uint32_t>(B_LY_CNT0, nb);
    uint32_t* ocb = S.buf<uint32_t>(B_LY_CNT1, nb);
    uint32_t* mca = S.buf<uint32_t>(B_LY_HF, nb);
    uint32_t* mcb = S.buf<uint32_t>(B_LY_HF2, nb);
    const SuffixCmp sc{T, n, ph};
    const size_t lds = (size_t)64 * (nss_text_row(bs) + nss_off_row(bs));
    k_nss_local<<<ceil_div(nb, 64), 64, lds, S.stream>>>(sc, nb, bs, lbs, lam, oa, oca, ma, mca);
    MIBWT_LAUNCH_CHECK();
    int levels = 0;
    for (uint64_t size = 1; size < nb; size *= 2) {
        const uint32_t pairs = (uint32_t)((nb + 2 * size - 1) / (2 * size));
        // eight lanes per pair while the pairs are many, a wavefront per pair above
        if (pairs >= (1u << 16))
            k_nss_merge<8><<<ceil_div(pairs, 8), 64, 0, S.stream>>>(sc, nb, bs, (uint32_t)size, lam, oa, oca, ma, mca,
                                                                    ob, ocb, mb, mcb);
        else
            k_nss_merge<64><<<pairs, 64, 0, S.stream>>>(sc, nb, bs, (uint32_t)size, lam, oa, oca, ma, mca, ob, ocb,
                                                        mb, mcb);
        MIBWT_LAUNCH_CHECK();
        std::swap(oa, ob);
        std::swap(oca, ocb);
        std::swap(ma, mb);
        std::swap(mca, mcb);
        levels++;
    }
    k_nss_last<<<256, kBlock, 0, S.stream>>>(oa, oca, n, lam);
    MIBWT_LAUNCH_CHECK();
    return levels;
}

// pw[i] = b^i and pw[kPowSize + i] = b^(kPowSize i), i < kPowSize (b16 = b^kPowSize)
__global__ void k_ph_pow(uint64_t b, uint64_t b16, uint64_t* __restrict__ pw)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kPowSize) return;
    uint64_t r0 = 1, r1 = 1, x0 = b, x1 = b16;
    for (uint32_t e = i; e; e >>= 1) {
        if (e & 1) {
            r0 = mulmod61(r0, x0);
            r1 = mulmod61(r1, x1);
        }
        x0 = mulmod61(x0, x0);
        x1 = mulmod61(x1, x1);
    }
    pw[i] = r0;
    pw[kPowSize + i] = r1;
}

// Prefix fingerprints pre[k] = H(T[0..k]) mod 2^61-1: per-tile fingerprints of the text, their
// exclusive scan, and the tiles' prefixes (the text is read twice, pre written once); the power
// tables of the base beside them.
static uint64_t powmod61(uint64_t b, uint64_t e)
{
    unsigned __int128 r = 1, x = b % kM61;
    while (e) {
        if (e & 1) r = (r * x) % kM61;
        x = (x * x) % kM61;
        e >>= 1;
    }
    return (uint64_t)r;
}

PrefixHash prefix_hashes(Storage& S, const uint8_t* T, uint64_t n, uint64_t base, uint64_t hmask)
{
    uint64_t* pre = S.buf<uint64_t>(B_LY_HASH, n);
    uint64_t* pw = S.buf<uint64_t>(B_LY_POW, 2 * kPowSize);
    k_ph_pow<<<kPowSize / kBlock, kBlock, 0, S.stream>>>(base, powmod61(base, kPowSize), pw);
    MIBWT_LAUNCH_CHECK();
    const uint32_t tiles = ceil_div(n, (uint64_t)kPhTile);
    HashPair* partial = S.buf<HashPair>(B_LY_SCAN0, tiles);
    const uint64_t base16 = powmod61(base, kPhItems);
    k_ph_reduce<<<tiles, kBlock, 0, S.stream>>>(T, n, base, base16, partial);
    MIBWT_LAUNCH_CHECK();
    device_scan<HashPair, OpHash>(S, partial, partial, tiles, false, HashPair{0, 1}, OpHash{}, nullptr, B_LY_SCAN1);
    k_ph_down<<<tiles, kBlock, 0, S.stream>>>(T, n, base, base16, partial, pre);
    MIBWT_LAUNCH_CHECK();
    return PrefixHash{pre, pw, hmask};
}

// ------------------------------------------------------------------- 3. label order
//
// The labels being ordered are strings of the text: item x stands for
// T[start(x) .. start(x) + len(x)).  The position-level induction orders every position by its
// longest Lyndon prefix (PosLab); the grammar orders each distinct label once, read at a
// representative occurrence (RepLab).
struct PosLab {
    const uint32_t* lam;
    __device__ __forceinline__ uint32_t start(uint32_t x) const { return x; }
    __device__ __forceinline__ uint32_t len(uint32_t x) const { return lam[x]; }
};
struct RepLab {
    const uint32_t* rep;
    const uint32_t* lam;
    __device__ __forceinline__ uint32_t start(uint32_t x) const { return rep[x]; }
    __device__ __forceinline__ uint32_t len(uint32_t x) const { return lam[rep[x]]; }
};

// Key of item x: the first `per` symbols of its label as dense codes (0 past its end).
template <typename Lab>
__global__ void k_lyn_initkey(const uint8_t* __restrict__ T, uint64_t n, Lab lab, uint32_t m,
                              const uint32_t* __restrict__ code, int bits, int per, uint64_t* __restrict__ K,
                              uint32_t* __restrict__ V)
{
    __shared__ uint32_t cs[256];
    cs[threadIdx.x] = code[threadIdx.x];  // blockDim == 256
    __syncthreads();
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= m) return;
    const uint32_t i = lab.start((uint32_t)x), l = lab.len((uint32_t)x);
    uint64_t key = 0, w = 0;
    for (int k = 0; k < per; k++) {
        if ((k & 7) == 0) w = (i + k < n) ? load8(T, i + k) : 0ull;
        const uint64_t s = ((uint32_t)k < l) ? cs[(w >> (8 * (k & 7))) & 0xFF] : 0ull;
        key = (key << bits) | s;
    }
    K[x] = key;
    V[x] = (uint32_t)x;
}

// fingerprint of every label (compared before any longest-common-prefix search: most tie
// groups of a collection are copies of one label)
template <typename Lab>
__global__ void k_label_hash(PrefixHash ph, Lab lab, uint32_t m, uint64_t* __restrict__ F)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < m) F[x] = ph.sub(lab.start(x), lab.len(x));
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

template <typename Lab>
struct LabelCmp {
    const uint8_t* T;
    Lab lab;
    PrefixHash ph;
    const uint64_t* F;  // label fingerprints
    // -1 / 0 / +1 for label x against label y, knowing their first `known` symbols agree
    __device__ int operator()(uint32_t x, uint32_t y, uint32_t known) const
    {
        if (x == y) return 0;
        const uint32_t lx = lab.len(x), ly = lab.len(y);
        const uint32_t m = lx < ly ? lx : ly;
        if (lx == ly && F[x] == F[y]) return 0;
        const uint32_t sx = lab.start(x), sy = lab.start(y);
        const uint32_t lo = label_lcp(ph, sx, sy, known, m);
        if (lo < m) return T[sx + lo] < T[sy + lo] ? -1 : 1;
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
template <typename Lab>
__global__ void k_lab_classify(const uint32_t* __restrict__ V, const uint32_t* __restrict__ gid1,
                               const uint32_t* __restrict__ gofs, Lab lab, uint32_t m, uint32_t known, uint32_t small,
                               const uint8_t* __restrict__ settled, uint8_t* __restrict__ cls,
                               uint32_t* __restrict__ large)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t g = gid1[i] - 1;
    const uint32_t s = gofs[g], e = gofs[g + 1];
    uint8_t c = 0;
    if (e - s >= 2 && lab.len(V[s]) >= known && !settled[s]) c = (e - s > small) ? 2 : 1;
    cls[i] = c;
    large[i] = (c == 2) ? 1u : 0u;
}

// label length and fingerprint of every item in the current order (the comparisons of a tie
// group then read them coalesced; most members of a collection's groups are equal labels)
template <typename Lab>
__global__ void k_lab_gather(const uint32_t* __restrict__ V, Lab lab, const uint64_t* __restrict__ F, uint32_t m,
                             uint32_t* __restrict__ GL, uint64_t* __restrict__ GF)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t x = V[i];
    GL[i] = lab.len(x);
    GF[i] = F[x];
}

// Small tie groups: comparator counting sort; settled groups are copied.  hf receives the
// heads of the refined order (first of every run of equal labels).
template <typename Lab>
__global__ void k_lab_small(const uint32_t* __restrict__ V, const uint32_t* __restrict__ gid1,
                            const uint32_t* __restrict__ gofs, const uint8_t* __restrict__ cls, uint32_t m,
                            uint32_t known, LabelCmp<Lab> cmp, const uint32_t* __restrict__ GL,
                            const uint64_t* __restrict__ GF, uint32_t* __restrict__ V2, uint32_t* __restrict__ hf)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t g = gid1[i] - 1;
    const uint32_t s = gofs[g], e = gofs[g + 1];
    const uint32_t x = V[i];
    if (cls[i] == 2) return;  // large groups are written by the pivot pass
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
        before += (c < 0) || (c == 0 && j < i);
        equal_before += (c == 0 && j < i);
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
template <typename Lab>
__global__ void k_lab_pivot_gather(PrefixHash ph, const uint8_t* __restrict__ T, uint32_t m,
                                   const uint32_t* __restrict__ V, const uint32_t* __restrict__ gid1,
                                   const uint32_t* __restrict__ gofs, const uint32_t* __restrict__ large,
                                   const uint32_t* __restrict__ lpos, Lab lab, uint32_t known, int lbits,
                                   uint64_t* __restrict__ LK, uint32_t* __restrict__ LV, uint32_t* __restrict__ LI,
                                   uint32_t* __restrict__ LG)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || !large[i]) return;
    const uint32_t g = gid1[i] - 1;
    const uint32_t x = V[i], p = V[gofs[g]];
    const uint32_t lx = lab.len(x), lp = lab.len(p);
    const uint32_t sx = lab.start(x), sp = lab.start(p);
    const uint32_t mm = lx < lp ? lx : lp;
    const uint32_t l = (x == p) ? mm : label_lcp(ph, sx, sp, known, mm);
    uint64_t cls, lpart = 0, ch = 0;
    if (l == mm && lx == lp) {
        cls = 1;  // EQUAL
    } else {
        const uint32_t cx = (l < lx) ? (uint32_t)T[sx + l] + 1u : 0u;  // 0: x ends at l
        const uint32_t cp = (l < lp) ? (uint32_t)T[sp + l] + 1u : 0u;
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
    LI[q] = i;
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
template <typename Lab>
__global__ void k_lab_verify(SuffixCmp sc, Lab lab, uint32_t m, const uint32_t* __restrict__ V,
                             const uint32_t* __restrict__ hf, uint32_t* __restrict__ bad)
{
    const uint8_t* T = sc.T;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 || i >= m) return;
    const uint32_t a = V[i - 1], b = V[i];
    const uint32_t la = lab.len(a), lb = lab.len(b);
    const uint32_t sa = lab.start(a), sb = lab.start(b);
    const uint32_t mm = la < lb ? la : lb;
    const uint32_t k = (sa == sb) ? mm : sc.lcp_upto(sa, sb, mm);
    int c;
    if (k < mm) c = T[sa + k] < T[sb + k] ? -1 : 1;
    else c = la < lb ? -1 : (la > lb ? 1 : 0);
    const bool head = hf[i] != 0;
    if (c > 0 || (c == 0) == head) atomicOr(bad, 1u);
}

__global__ void k_u32_to_key(const uint32_t* __restrict__ a, uint64_t n, uint64_t* __restrict__ K)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) K[i] = a[i];
}

// Orders m labels.  On return *v_out holds the items in label order and *hf_out flags the first
// item of every run of equal labels.  Returns false when the final neighbour check fails (a
// fingerprint collision broke the order).  Labels that agree on their first `per` symbols (the
// packed initial key) are refined in rounds: tie groups of up to `small` members by the
// comparator counting sort, larger ones by the pivot partition.
template <typename Lab>
static bool label_order(Storage& S, const uint8_t* T, uint64_t n, const PrefixHash& phh, Lab lab, uint32_t m,
                        const uint32_t* codes, uint32_t sigma, uint32_t small, uint32_t** v_out, uint32_t** hf_out,
                        double* bytes)
{
    uint32_t* sc = S.buf<uint32_t>(B_SCALARS, SC_COUNT_);
    const int bits = bits_for((uint64_t)sigma + 1);
    const int per = 64 / bits;
    uint64_t* K0 = S.buf<uint64_t>(B_SORT_K0, m);
    uint64_t* K1 = S.buf<uint64_t>(B_SORT_K1, m);
    uint32_t* V0 = S.buf<uint32_t>(B_SORT_V0, m);
    uint32_t* V1 = S.buf<uint32_t>(B_SORT_V1, m);
    uint32_t* hf = S.buf<uint32_t>(B_LY_HF, m);
    uint32_t* gid1 = S.buf<uint32_t>(B_DB_GID, m);
    uint32_t* gofs = S.buf<uint32_t>(B_DB_GOFS, (size_t)m + 1);
    uint32_t* large = S.buf<uint32_t>(B_DB_LARGE, m);
    uint32_t* lpos = S.buf<uint32_t>(B_ACT_X, m);
    uint8_t* cls = S.buf<uint8_t>(B_LY_CLS, m);
    uint32_t* Vl = S.buf<uint32_t>(B_LY_VL, m);
    k_lyn_initkey<Lab><<<grid_for(m), kBlock, 0, S.stream>>>(T, n, lab, m, codes, bits, per, K0, V0);
    MIBWT_LAUNCH_CHECK();
    radix_sort<uint64_t, uint32_t>(S, K0, K1, V0, V1, m, 0, bits * per);
    k_lab_heads_from_keys<<<grid_for(m), kBlock, 0, S.stream>>>(K0, m, hf);
    MIBWT_LAUNCH_CHECK();
    double rbytes = (double)m * 24 * ((bits * per + 7) / 8);
    uint64_t* LF = S.buf<uint64_t>(B_LY_LABH, m);
    k_label_hash<Lab><<<grid_for(m), kBlock, 0, S.stream>>>(phh, lab, m, LF);
    MIBWT_LAUNCH_CHECK();
    LabelCmp<Lab> cmp{T, lab, phh, LF};
    // working copies: V (order) and H (heads) of the packed-key order
    uint32_t* V = Vl;
    MIBWT_HIP(hipMemcpyAsync(V, V0, sizeof(uint32_t) * m, hipMemcpyDeviceToDevice, S.stream));
    uint32_t* H = S.buf<uint32_t>(B_LY_HF2, m);
    MIBWT_HIP(hipMemcpyAsync(H, hf, sizeof(uint32_t) * m, hipMemcpyDeviceToDevice, S.stream));
    const uint32_t known = (uint32_t)per;  // every tie group agrees on its first `per` symbols
    if (const char* e = getenv("MIBWT_LAB_SMALL")) {
        const int v = atoi(e);
        if (v >= 2 && v <= 256) small = (uint32_t)v;
    }
    uint8_t* settled = S.buf<uint8_t>(B_LY_SETTLED, m);
    fill<uint8_t>(S, settled, m, (uint8_t)0);
    for (int round = 0;; round++) {
        MIBWT_REQUIRE(round < 1 << 20, LIBCUBWT_GPU_ERROR, "label refinement does not converge");
        inclusive_sum<uint32_t>(S, H, gid1, m, sc + SC_GROUPS);
        k_grp_offsets<<<grid_for(m), kBlock, 0, S.stream>>>(H, gid1, m, gofs);
        MIBWT_LAUNCH_CHECK();
        k_lab_classify<Lab><<<grid_for(m), kBlock, 0, S.stream>>>(V, gid1, gofs, lab, m, known, small, settled, cls,
                                                                   large);
        MIBWT_LAUNCH_CHECK();
        exclusive_sum<uint32_t>(S, large, lpos, m, sc + SC_LARGE);
        uint32_t hv[SC_COUNT_];
        S.read_fast(hv, sc, SC_COUNT_);
        const uint32_t nl = hv[SC_LARGE];
        uint32_t* V2 = (V == Vl) ? V1 : Vl;
        uint32_t* GL = S.buf<uint32_t>(B_LY_GL, m);
        uint64_t* GF = S.buf<uint64_t>(B_LY_GF, m);
        k_lab_gather<Lab><<<grid_for(m), kBlock, 0, S.stream>>>(V, lab, LF, m, GL, GF);
        MIBWT_LAUNCH_CHECK();
        k_lab_small<Lab><<<grid_for(m), kBlock, 0, S.stream>>>(V, gid1, gofs, cls, m, known, cmp, GL, GF, V2, H);
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
            k_lab_pivot_gather<Lab><<<grid_for(m), kBlock, 0, S.stream>>>(phh, T, m, V, gid1, gofs, large, lpos, lab,
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
    k_lab_verify<Lab><<<grid_for(m), kBlock, 0, S.stream>>>(SuffixCmp{T, n, phh}, lab, m, V, H, sc + SC_COLLISION);
    MIBWT_LAUNCH_CHECK();
    uint32_t bad = 0;
    S.read_fast(&bad, sc + SC_COLLISION);
    *bytes += rbytes;
    if (bad != 0) return false;
    if (V != Vl) MIBWT_HIP(hipMemcpyAsync(Vl, V, sizeof(uint32_t) * m, hipMemcpyDeviceToDevice, S.stream));
    MIBWT_HIP(hipMemcpyAsync(hf, H, sizeof(uint32_t) * m, hipMemcpyDeviceToDevice, S.stream));
    *v_out = Vl;
    *hf_out = hf;
    return true;
}

// ------------------------------------------------------------------- 4. the grammar
//
// Every position i starts exactly one maximal node of the Lyndon forest, its label
// P_i = T[i .. i + lam[i]) (the reference's rules: right children and stack words).  A label M of
// length >= 2 has the children c_1 = i + 1, c_{k+1} = c_k + lam[c_k] < i + lam[i]: the Lyndon
// factorization of M without its first symbol, non-increasing.  The children's labels are labels
// themselves, so the distinct labels with their child lists are the grammar; the roots (positions
// that no earlier label covers: the Lyndon factorization of T) are the stack.
//
// The suffix at a position is the sequence of labels along its next-smaller-suffix chain (its
// Lyndon factorization), and suffixes compare like these sequences, a run of k equal labels
// comparing as one symbol (label, k): more copies is larger, since the label after a run is
// smaller.  A run never reaches past the end of the label whose children it belongs to, because
// what follows a label M is at most M and M is smaller than each of its children; so the run
// length from every child slot is a property of the label alone.
//
// Slots: child k of distinct label d is slot off[d] + k; root s is slot C + s; slot Z = C + S is
// the end of the text.  An item is a slot standing for the positions of that child in every
// occurrence of its label (cnt[d] of them; one for a root slot), all preceded by the same
// character.  After the run that starts at slot q an item continues at nxt[q], the slot after
// the run in the same label (or the next root), or, when the run ends its label d, at each of
// d's up-contexts: the slots that follow d where d occurs as a child.  d occurs exactly at the
// ends of the last-child chains (v, lastchild(v), ...) of the children v that are not the last
// in their label (the context is the next child slot) and of the roots (the next root slot).
// The counts of the up-contexts of d add up to cnt[d].

constexpr uint32_t kNone = 0xFFFFFFFFu;

// Slots of B_GR_SC.
enum GrSlot : int {
    GR_D = 0,     // distinct labels
    GR_S,         // roots
    GR_C,         // child slots
    GR_UT,        // up-context pairs
    GR_NR,        // distinct run keys
    GR_ERR,       // an inconsistent grammar (a child label missing, a run past the end): colliding keys
    GR_M,         // items of the next round
    GR_G,         // their groups
    GR_RUNS,      // runs written this round
    GR_RB,        // bytes of those runs
    GR_NG,        // groups of this round
    GR_PIDX,      // primary index (BWT position of suffix 0, plus one)
    GR_NA,        // positions whose label key fits its high word
    GR_DA,        // distinct labels among them
    GR_DB,        // distinct labels among the others
    GR_COUNT_ = 16
};

// Identity of the label at position i as one 64-bit key.  Labels of at most per32 symbols: the
// number 1 d_0 d_1 .. d_{l-1} in base `radix` (d_k = dense symbol code - 2, radix = max(sigma, 2)),
// below 2^32, one number per string.  Labels of at most `per` symbols: their codes (0 past the
// end) packed `bits` bits each from the top, exact; the first code is at least 2, so such a key is
// at least 2^(bits (per - 1) + 1) >= 2^33.  Longer labels: their fingerprint mixed with the length,
// top bit set.  Equal labels always get equal keys, and the three kinds never meet.
struct LabelKey {
    const uint8_t* T;
    uint64_t n;
    const uint32_t* lam;
    PrefixHash ph;
    int bits, per;
    uint32_t per32, radix;
    __device__ __forceinline__ uint64_t operator()(uint32_t i, const uint32_t* cs) const
    {
        const uint32_t l = lam[i];
        return (l <= per32) ? short_key(i, l, cs) : long_key(i, l, cs);
    }
    // labels of l <= per32 symbols
    __device__ __forceinline__ uint32_t short_key(uint32_t i, uint32_t l, const uint32_t* cs) const
    {
        uint32_t key = 1;
        uint64_t w = 0;
        for (uint32_t k = 0; k < l; k++) {
            if ((k & 7) == 0) w = load8(T, (uint64_t)i + k);
            key = key * radix + (cs[(w >> (8 * (k & 7))) & 0xFF] - 2);
        }
        return key;
    }
    // labels of l > per32 symbols
    __device__ __forceinline__ uint64_t long_key(uint32_t i, uint32_t l, const uint32_t* cs) const
    {
        if (l <= (uint32_t)per) {
            // the l symbols, then zero codes up to `per` (l <= per < 64, so i + l <= n)
            uint64_t key = 0, w = 0;
            for (uint32_t k = 0; k < l; k++) {
                if ((k & 7) == 0) w = load8(T, (uint64_t)i + k);
                key = (key << bits) | cs[(w >> (8 * (k & 7))) & 0xFF];
            }
            return key << (bits * ((uint32_t)per - l));
        }
        const uint64_t f = ph.sub(i, l) ^ ((uint64_t)l * 0x9E3779B97F4A7C15ull);
        return (1ull << 63) | (f >> 1);
    }
};

// index of `key` in the sorted distinct keys U[0..D); sets *err when it is missing
__device__ __forceinline__ uint32_t key_lookup(const uint64_t* __restrict__ U, uint32_t D, uint64_t key,
                                               uint32_t* err)
{
    uint32_t lo = 0, hi = D;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (U[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= D || U[lo] != key) {
        atomicOr(err, 1u);
        return 0;
    }
    return lo;
}

// Keys of labels of at most `per32` symbols are below 2^32 and sorted as 32-bit keys, the others
// as 64-bit keys: the scan input "position i has a short label".
struct ScanShort {
    const uint32_t* lam;
    uint32_t per32;
    __device__ __forceinline__ uint32_t operator()(uint64_t i) const { return lam[i] <= per32 ? 1u : 0u; }
    __device__ __forceinline__ uint4 load4(uint64_t i) const
    {
        const uint4 v = *reinterpret_cast<const uint4*>(lam + i);
        return make_uint4(v.x <= per32, v.y <= per32, v.z <= per32, v.w <= per32);
    }
};

// the key of every position into the short list (32 bits; fa = exclusive scan of the flags) or
// the long list (64 bits), with the position as value
// The key work is split by label kind so that a wavefront runs one kind of loop: this pass keys
// the short labels and lists the positions of the others, k_gl_keys_long keys those.
__global__ void k_gl_keys(LabelKey lk, const uint32_t* __restrict__ code, uint32_t per32,
                          const uint32_t* __restrict__ fa, uint32_t* __restrict__ KA, uint32_t* __restrict__ VA,
                          uint32_t* __restrict__ VB)
{
    __shared__ uint32_t cs[256];
    cs[threadIdx.x] = code[threadIdx.x];  // blockDim == 256
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= lk.n) return;
    const uint32_t l = lk.lam[i];
    const uint32_t a = fa[i];
    if (l <= per32) {
        KA[a] = lk.short_key((uint32_t)i, l, cs);
        VA[a] = (uint32_t)i;
    } else {
        VB[i - a] = (uint32_t)i;
    }
}

__global__ void k_gl_keys_long(LabelKey lk, const uint32_t* __restrict__ code, const uint32_t* __restrict__ VB,
                               uint32_t nB, uint64_t* __restrict__ KB)
{
    __shared__ uint32_t cs[256];
    cs[threadIdx.x] = code[threadIdx.x];  // blockDim == 256
    __syncthreads();
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nB) return;
    const uint32_t i = VB[j];
    KB[j] = lk.long_key(i, lk.lam[i], cs);
}

// Short labels whose keys lie below 2^24 are counted in a table indexed by the key instead of
// being sorted: per key the occurrences and a representative occurrence.  Keys below kCountBins
// (on DNA the labels of up to 5 symbols, 80% of the positions) are counted per workgroup in LDS
// first, kCountPer positions per thread, and leave as one device atomic per key and workgroup
// (representative: the smallest position); the others go to the table directly, one atomic
// count and a plain store of the position (any occurrence serves: a label's string, children
// and order are the same at each).  (Combining the lanes of a wavefront with equal small keys
// first, one ballot per key bit, measured no faster: the kernel waits on memory, not on LDS
// atomics.)  The long labels' positions are listed as in k_gl_keys.
constexpr int kTabBitsMax = 24;
constexpr uint32_t kCountBins = 2048;
constexpr int kCountPer = 8;
__global__ __launch_bounds__(kBlock) void k_gl_count(LabelKey lk, const uint32_t* __restrict__ code,
                                                    const uint32_t* __restrict__ fa, uint32_t* __restrict__ tcnt,
                                                    uint32_t* __restrict__ trep, uint32_t* __restrict__ VB)
{
    __shared__ uint32_t cs[256];
    __shared__ uint32_t bc[kCountBins], br[kCountBins];
    cs[threadIdx.x] = code[threadIdx.x];  // blockDim == 256
    for (uint32_t b = threadIdx.x; b < kCountBins; b += kBlock) {
        bc[b] = 0;
        br[b] = 0xFFFFFFFFu;
    }
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * (kBlock * kCountPer);
    for (int j = 0; j < kCountPer; j++) {
        const uint64_t i = base + (uint64_t)j * kBlock + threadIdx.x;
        const bool valid = i < lk.n;
        const uint32_t l = valid ? lk.lam[i] : 0u;
        const bool sh = valid && l <= lk.per32;
        uint32_t key = 0;
        if (sh) key = lk.short_key((uint32_t)i, l, cs);
        else if (valid) VB[i - fa[i]] = (uint32_t)i;
        const bool small = sh && key < kCountBins;
        if (sh && !small) {
            atomicAdd(&tcnt[key], 1u);
            trep[key] = (uint32_t)i;  // any occurrence: the label's string is the same at each
        }
        if (small) {
            atomicAdd(&bc[key], 1u);
            atomicMin(&br[key], (uint32_t)i);
        }
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < kCountBins; b += kBlock)
        if (bc[b]) {
            atomicAdd(&tcnt[b], bc[b]);
            atomicMin(&trep[b], br[b]);
        }
}

// the scan input "table entry k is used"
struct ScanNonzero {
    const uint32_t* c;
    __device__ __forceinline__ uint32_t operator()(uint64_t i) const { return c[i] ? 1u : 0u; }
    __device__ __forceinline__ uint4 load4(uint64_t i) const
    {
        const uint4 v = *reinterpret_cast<const uint4*>(c + i);
        return make_uint4(v.x != 0, v.y != 0, v.z != 0, v.w != 0);
    }
};

// the used table entries in key order: representative, count, key
__global__ void k_gl_tab_dedup(const uint32_t* __restrict__ tcnt, const uint32_t* __restrict__ trep,
                               const uint32_t* __restrict__ tpos, uint32_t K, uint32_t* __restrict__ rep,
                               uint32_t* __restrict__ cnt, uint64_t* __restrict__ U)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const uint32_t c = tcnt[k];
    if (!c) return;
    const uint32_t d = tpos[k];
    rep[d] = trep[k];
    cnt[d] = c;
    U[d] = k;
}

// per distinct label d (the runs of equal sorted keys): representative, count, 64-bit key
template <typename K>
__global__ void k_gl_dedup(const K* __restrict__ Ks, const uint32_t* __restrict__ V, const uint32_t* __restrict__ gofs,
                           uint32_t D, uint32_t* __restrict__ rep, uint32_t* __restrict__ cnt,
                           uint64_t* __restrict__ U)
{
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= D) return;
    const uint32_t j = gofs[d];
    rep[d] = V[j];
    cnt[d] = gofs[d + 1] - j;
    U[d] = (uint64_t)Ks[j];
}

// number of keys in the sorted X[0..m) below x
__device__ __forceinline__ uint32_t count_below64(const uint64_t* __restrict__ X, uint32_t m, uint64_t x)
{
    uint32_t lo = 0, hi = m;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (X[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// the two sorted lists of distinct labels (their keys never coincide: short keys are below 2^32,
// the others at least 2^33) merged into one sorted list
__global__ void k_gl_merge(const uint64_t* __restrict__ UA, const uint32_t* __restrict__ RA,
                           const uint32_t* __restrict__ CA, uint32_t DA, const uint64_t* __restrict__ UB,
                           const uint32_t* __restrict__ RB, const uint32_t* __restrict__ CB, uint32_t DB,
                           uint64_t* __restrict__ U, uint32_t* __restrict__ rep, uint32_t* __restrict__ cnt)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= DA + DB) return;
    if (t < DA) {
        const uint64_t k = UA[t];
        const uint32_t o = t + count_below64(UB, DB, k);
        U[o] = k;
        rep[o] = RA[t];
        cnt[o] = CA[t];
    } else {
        const uint32_t d = t - DA;
        const uint64_t k = UB[d];
        const uint32_t o = d + count_below64(UA, DA, k);
        U[o] = k;
        rep[o] = RB[d];
        cnt[o] = CB[d];
    }
}

// The roots are the starts of the Lyndon factorization of T: 0, then r + lam[r] (the first factor
// of a suffix is its longest Lyndon prefix).  One thread follows that chain for at most `limit`
// factors and writes their number, or kNone when the text has more (then the prefix maxima of
// i + lam[i] below find them in parallel).
constexpr uint32_t kRootWalk = 1024;
__global__ void k_gl_rootwalk(const uint32_t* __restrict__ lam, uint64_t n, uint32_t limit,
                              uint32_t* __restrict__ roots, uint32_t* __restrict__ count)
{
    if (threadIdx.x != 0) return;
    uint64_t k = 0;
    uint32_t c = 0;
    while (k < n && c < limit) {
        roots[c++] = (uint32_t)k;
        k += lam[k];
    }
    *count = (k >= n) ? c : kNone;
}

__global__ void k_gl_ends(const uint32_t* __restrict__ lam, uint64_t n, uint32_t* __restrict__ E)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) E[i] = (uint32_t)i + lam[i];
}

// root flags from the prefix maxima PM of i + lam[i]: no label starting before k reaches past k
__global__ void k_gl_rootflag(const uint32_t* __restrict__ PM, uint64_t n, uint32_t* __restrict__ rf)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) rf[k] = (k == 0 || PM[k - 1] <= k) ? 1u : 0u;
}

__global__ void k_gl_roots(const uint32_t* __restrict__ PM, const uint32_t* __restrict__ rpos, uint64_t n,
                           uint32_t* __restrict__ roots)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && (k == 0 || PM[k - 1] <= k)) roots[rpos[k]] = (uint32_t)k;
}

// number of children of each distinct label
__global__ void k_gl_nchild(const uint32_t* __restrict__ rep, const uint32_t* __restrict__ lam, uint32_t D,
                            uint32_t* __restrict__ t)
{
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= D) return;
    const uint32_t p = rep[d], e = p + lam[p];
    uint32_t c = 0;
    for (uint32_t x = p + 1; x < e; x += lam[x]) c++;
    t[d] = c;
}

// child slots of each distinct label: child label, preceding character, owner; last child
__global__ void k_gl_children(LabelKey lk, const uint32_t* __restrict__ code, const uint32_t* __restrict__ rep,
                              const uint32_t* __restrict__ off, uint32_t D, const uint64_t* __restrict__ U,
                              uint32_t* __restrict__ chl, uint8_t* __restrict__ chc, uint32_t* __restrict__ cont,
                              uint32_t* __restrict__ last, uint32_t* __restrict__ err)
{
    __shared__ uint32_t cs[256];
    cs[threadIdx.x] = code[threadIdx.x];
    __syncthreads();
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= D) return;
    const uint32_t p = rep[d], e = p + lk.lam[p];
    uint32_t q = off[d], l = kNone;
    for (uint32_t x = p + 1; x < e; x += lk.lam[x], q++) {
        l = key_lookup(U, D, lk(x, cs), err);
        chl[q] = l;
        chc[q] = lk.T[x - 1];
        cont[q] = d;
    }
    last[d] = l;
}

// root slots C + s
__global__ void k_gl_rootslots(LabelKey lk, const uint32_t* __restrict__ code, const uint32_t* __restrict__ roots,
                               uint32_t S, uint32_t C, const uint64_t* __restrict__ U, uint32_t D,
                               uint32_t* __restrict__ chl, uint8_t* __restrict__ chc, uint32_t* __restrict__ cont,
                               uint32_t* __restrict__ err)
{
    __shared__ uint32_t cs[256];
    cs[threadIdx.x] = code[threadIdx.x];
    __syncthreads();
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const uint32_t k = roots[s];
    chl[C + s] = key_lookup(U, D, lk(k, cs), err);
    chc[C + s] = k ? lk.T[k - 1] : lk.T[lk.n - 1];
    cont[C + s] = kNone;
}

// run breaks, reversed for a prefix-maximum scan: RV[Z-1-q] = Z-1-q where a run ends at slot q
__global__ void k_gl_brk(const uint32_t* __restrict__ chl, const uint32_t* __restrict__ cont, uint32_t Z,
                         uint32_t* __restrict__ RV)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Z) return;
    const bool brk = (q + 1 == Z) || cont[q] != cont[q + 1] || chl[q] != chl[q + 1];
    const uint32_t r = Z - 1 - q;
    RV[r] = brk ? r : 0u;
}

// per slot: run length (in place of the scanned breaks), next slot, positions per item
__global__ void k_gl_runs(uint32_t* __restrict__ RV, const uint32_t* __restrict__ cont, const uint32_t* __restrict__ cnt,
                          uint32_t Z, uint32_t* __restrict__ mrun, uint32_t* __restrict__ nxt, uint32_t* __restrict__ ccnt)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > Z) return;
    if (q == Z) {  // the end of the text
        nxt[Z] = kNone;
        ccnt[Z] = 1;
        return;
    }
    const uint32_t re = Z - 1 - RV[Z - 1 - q];  // last slot of the run starting at q
    const uint32_t c = cont[q], nx = re + 1;
    mrun[q] = re - q + 1;
    if (c == kNone) nxt[q] = nx;  // the next root, or the end
    else nxt[q] = (nx < Z && cont[nx] == c) ? nx : kNone;
    ccnt[q] = (c == kNone) ? 1u : cnt[c];
}

// Up-contexts: a slot q that is a root or not the last child of its label sends every label on
// the last-child chain of its child to slot q + 1.  Pass 1 counts, pass 2 writes the pairs.
__device__ __forceinline__ bool up_source(const uint32_t* __restrict__ cont, uint32_t q, uint32_t Z)
{
    const uint32_t c = cont[q];
    return c == kNone || (q + 1 < Z && cont[q + 1] == c);
}

__global__ void k_gl_upcount(const uint32_t* __restrict__ chl, const uint32_t* __restrict__ cont,
                             const uint32_t* __restrict__ off, const uint32_t* __restrict__ last, uint32_t Z,
                             uint32_t* __restrict__ ucnt)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Z) return;
    uint32_t c = 0;
    if (up_source(cont, q, Z))
        for (uint32_t v = chl[q]; off[v + 1] > off[v]; v = last[v]) c++;
    ucnt[q] = c;
}

__global__ void k_gl_upfill(const uint32_t* __restrict__ chl, const uint32_t* __restrict__ cont,
                            const uint32_t* __restrict__ off, const uint32_t* __restrict__ last, uint32_t Z,
                            const uint32_t* __restrict__ uoff, uint32_t* __restrict__ UK, uint32_t* __restrict__ UV)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Z || !up_source(cont, q, Z)) return;
    uint32_t o = uoff[q];
    for (uint32_t v = chl[q]; off[v + 1] > off[v]; v = last[v], o++) {
        UK[o] = v;
        UV[o] = q + 1;
    }
}

// list offsets of the up-contexts sorted by label: uo[v] = first pair of label v, uo[D] = UT
__global__ void k_gl_csr(const uint32_t* __restrict__ UK, uint32_t UT, uint32_t D, uint32_t* __restrict__ uo)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= UT) return;
    const uint32_t a = (j == 0) ? 0u : UK[j - 1] + 1u;
    for (uint32_t v = a; v <= UK[j]; v++) uo[v] = j;
    if (j == UT - 1)
        for (uint32_t v = UK[j] + 1; v <= D; v++) uo[v] = UT;
}

__global__ void k_gl_lrank(const uint32_t* __restrict__ V, const uint32_t* __restrict__ rk, uint32_t D,
                           uint32_t* __restrict__ lrank)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < D) lrank[V[j]] = rk[j];
}

// run key of every slot: (rank of its child's label, run length), to be ranked densely
__global__ void k_gl_rkinit(const uint32_t* __restrict__ chl, const uint32_t* __restrict__ mrun,
                            const uint32_t* __restrict__ lrank, uint32_t Z, int mb, uint64_t* __restrict__ K,
                            uint32_t* __restrict__ V)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Z) return;
    K[q] = ((uint64_t)lrank[chl[q]] << mb) | mrun[q];
    V[q] = q;
}

__global__ void k_gl_rkey(const uint32_t* __restrict__ V, const uint32_t* __restrict__ rk, uint32_t Z,
                          uint32_t* __restrict__ rkey)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < Z) rkey[V[j]] = rk[j];  // 1-based; the end of the text keeps key 0
    if (j == 0) rkey[Z] = 0;
}

// ---- items

__global__ void k_it_init(const uint8_t* __restrict__ chc, uint32_t Z, uint32_t C, uint32_t* __restrict__ IQ,
                          uint32_t* __restrict__ IG, uint32_t* __restrict__ ID, uint32_t* __restrict__ ICH)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Z) return;
    IQ[q] = q;
    IG[q] = 0;
    ID[q] = 0;
    ICH[q] = (uint32_t)chc[q] | ((q == C) ? 0x100u : 0u);  // slot C: the root at position 0
}

__global__ void k_it_hasrule(const uint32_t* __restrict__ IQ, const uint32_t* __restrict__ ID, uint32_t m,
                             uint32_t C, uint32_t* __restrict__ HR)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m && IQ[i] < C) HR[ID[i]] = 1;
}

// Sort key: the item's group, then its next factor run.  A group with only root items (and the
// end) orders them by position alone: a later root's suffix is smaller (it lies on the earlier
// one's next-smaller-suffix chain).
__global__ void k_it_key(const uint32_t* __restrict__ IQ, const uint32_t* __restrict__ ID, uint32_t m, uint32_t C,
                         uint32_t Z, const uint32_t* __restrict__ HR, const uint32_t* __restrict__ rkey, int kb,
                         uint64_t* __restrict__ K, uint32_t* __restrict__ V)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t q = IQ[i], g = ID[i];
    const uint32_t k = (q >= C && !HR[g]) ? (Z - q) : rkey[q];
    K[i] = ((uint64_t)g << kb) | k;
    V[i] = i;
}

__global__ void k_it_gather(const uint32_t* __restrict__ V, uint32_t m, const uint32_t* __restrict__ IQ,
                            const uint32_t* __restrict__ IG, const uint32_t* __restrict__ ICH,
                            const uint32_t* __restrict__ ccnt, uint32_t* __restrict__ SQ, uint32_t* __restrict__ SG,
                            uint32_t* __restrict__ SCH, uint32_t* __restrict__ CN)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const uint32_t i = V[j], q = IQ[i];
    SQ[j] = q;
    SG[j] = IG[i];
    SCH[j] = ICH[i];
    CN[j] = ccnt[q];
}

// first item of each old group (OF) and head flags of the new groups (NH)
__global__ void k_it_heads(const uint64_t* __restrict__ K, uint32_t m, int kb, uint32_t* __restrict__ OF,
                           uint32_t* __restrict__ NH)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const uint64_t k = K[j];
    const bool first = (j == 0) || (K[j - 1] >> kb) != (k >> kb);
    if (first) OF[k >> kb] = j;
    NH[j] = (j == 0 || K[j - 1] != k) ? 1u : 0u;
}

// New groups: BWT start (GST) and first item (GFI) of each; GEXP flags the groups that must be
// refined further (two characters, or the item of suffix 0 among several items).
__global__ void k_it_groups(const uint64_t* __restrict__ K, uint32_t m, int kb, const uint32_t* __restrict__ SG,
                            const uint32_t* __restrict__ P, const uint32_t* __restrict__ OF,
                            const uint32_t* __restrict__ NG1, const uint32_t* __restrict__ SCH,
                            uint32_t* __restrict__ GST, uint32_t* __restrict__ GFI, uint32_t* __restrict__ GEXP)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const uint32_t ng = NG1[j] - 1;
    const bool head = (j == 0) || NG1[j - 1] != NG1[j];
    if (head) {
        GST[ng] = SG[j] + P[j] - P[OF[K[j] >> kb]];
        GFI[ng] = j;
    } else {
        const uint32_t c = SCH[j], pc = SCH[j - 1];
        if (((c ^ pc) & 0xFFu) || ((c | pc) & 0x100u)) GEXP[ng] = 1;
    }
}

// Per item: the expansion count of a refined group's item, the run record of a finished group's
// first item; the primary index when suffix 0 finishes.
__global__ void k_it_flags(uint32_t m, const uint32_t* __restrict__ NG1, const uint32_t* __restrict__ GEXP,
                           const uint32_t* __restrict__ SQ, const uint32_t* __restrict__ SCH,
                           const uint32_t* __restrict__ nxt, const uint32_t* __restrict__ cont,
                           const uint32_t* __restrict__ uo, uint32_t Z, const uint32_t* __restrict__ GST,
                           uint32_t* __restrict__ RF, uint32_t* __restrict__ EC, uint32_t* __restrict__ sc)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const uint32_t ng = NG1[j] - 1;
    const bool head = (j == 0) || NG1[j - 1] != NG1[j];
    const bool exp = GEXP[ng] != 0;
    RF[j] = (!exp && head) ? 1u : 0u;
    uint32_t e = 0;
    if (exp) {
        const uint32_t q = SQ[j];
        if (q >= Z) {
            sc[GR_ERR] = 1;  // the end of the text is never refined further
        } else if (nxt[q] != kNone) {
            e = 1;
        } else {
            const uint32_t c = cont[q];
            e = uo[c + 1] - uo[c];
        }
    } else if (SCH[j] & 0x100u) {
        sc[GR_PIDX] = GST[ng] + 1;
    }
    EC[j] = e;
}

// run records of the finished groups: start, length (the group's count), character; RIDX is the
// exclusive scan of the run flags
__global__ void k_it_runs(uint32_t m, const uint32_t* __restrict__ RIDX, const uint32_t* __restrict__ NG1,
                          const uint32_t* __restrict__ GEXP, const uint32_t* __restrict__ GST,
                          const uint32_t* __restrict__ GFI, const uint32_t* __restrict__ P,
                          const uint32_t* __restrict__ SCH, const uint32_t* __restrict__ sc,
                          uint32_t* __restrict__ RST, uint32_t* __restrict__ RLEN, uint8_t* __restrict__ RCH)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const uint32_t ng = NG1[j] - 1;
    const bool head = (j == 0) || NG1[j - 1] != NG1[j];
    if (!head || GEXP[ng]) return;
    const uint32_t NG = sc[GR_NG];
    const uint32_t end = (ng + 1 < NG) ? GFI[ng + 1] : m;
    const uint32_t r = RIDX[j];
    RST[r] = GST[ng];
    RLEN[r] = P[end] - P[j];
    RCH[r] = (uint8_t)(SCH[j] & 0xFFu);
}

// Writes the finished runs, 16 bytes of their concatenation per thread and step (grid-stride;
// the total comes from the device).
__global__ void k_it_fill(const uint32_t* __restrict__ ROFF, const uint32_t* __restrict__ RST,
                          const uint8_t* __restrict__ RCH, uint32_t nruns, const uint32_t* __restrict__ total,
                          uint8_t* __restrict__ O, uint64_t n, uint32_t* __restrict__ err)
{
    const uint64_t bytes = *total;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * 16;
    for (uint64_t b0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 16; b0 < bytes; b0 += stride) {
        uint32_t lo = 0, hi = nruns;  // last run with ROFF <= b0
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (ROFF[mid] <= b0) lo = mid;
            else hi = mid;
        }
        uint32_t r = lo;
        const uint64_t b1 = (b0 + 16 < bytes) ? b0 + 16 : bytes;
        for (uint64_t b = b0; b < b1; b++) {
            while (r + 1 < nruns && ROFF[r + 1] <= b) r++;
            const uint64_t at = RST[r] + (b - ROFF[r]);
            if (at < n) O[at] = RCH[r];
            else *err = 1;  // a run past the end: never for consistent groups
        }
    }
}

// Items of the next round: output o belongs to the item whose expansion range holds it.
__global__ void k_it_expand(uint32_t m, uint32_t mo, const uint32_t* __restrict__ EOFF, const uint32_t* __restrict__ SQ,
                            const uint32_t* __restrict__ SCH, const uint32_t* __restrict__ NG1,
                            const uint32_t* __restrict__ GST, const uint32_t* __restrict__ XID,
                            const uint32_t* __restrict__ nxt, const uint32_t* __restrict__ cont,
                            const uint32_t* __restrict__ uo, const uint32_t* __restrict__ UV, uint32_t* __restrict__ OQ,
                            uint32_t* __restrict__ OG, uint32_t* __restrict__ OID, uint32_t* __restrict__ OCH)
{
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= mo) return;
    uint32_t lo = 0, hi = m;  // last item with EOFF <= o
    while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (EOFF[mid] <= o) lo = mid;
        else hi = mid;
    }
    const uint32_t i = lo, q = SQ[i], ng = NG1[i] - 1;
    const uint32_t nq = nxt[q];
    OQ[o] = (nq != kNone) ? nq : UV[uo[cont[q]] + (o - EOFF[i])];
    OG[o] = GST[ng];
    OID[o] = XID[ng];
    OCH[o] = SCH[i];
}

// The grammar build: labels, their order, the induction.  Returns the primary index, 0 when the
// label order check found a fingerprint collision, or kNone when the induction outgrew its
// budget (the caller then runs the position-level induction).
static uint32_t lyndon_grammar(Storage& S, const uint8_t* T, uint64_t n, const PrefixHash& phh, const uint32_t* lam,
                               const uint32_t* codes, uint32_t sigma, uint8_t* dO)
{
    uint32_t* gsc = S.buf<uint32_t>(B_GR_SC, GR_COUNT_);
    fill<uint32_t>(S, gsc, GR_COUNT_, 0u);
    uint32_t hv[GR_COUNT_];

    // ---- distinct labels
    int ph = S.phase_begin("labels");
    const int bits = bits_for((uint64_t)sigma + 1);
    const int per = 63 / bits;
    // short labels: keys below 2^24 (three radix passes) when that holds 8 symbols, else below 2^32
    const uint32_t radix = std::max<uint32_t>(sigma, 2);
    auto fit = [&](uint64_t lim) {
        uint32_t L = 0;
        for (uint64_t v = 2ull * radix; v <= lim; v *= radix) L++;
        return L;
    };
    uint32_t per32 = fit(1ull << 24);
    if (per32 < 8) per32 = fit(1ull << 32);
    uint64_t amax = 2;
    for (uint32_t k = 0; k < per32; k++) amax *= radix;
    const int bitsA = bits_for(amax - 1);  // keys below 2 radix^per32
    const LabelKey lk{T, n, lam, phh, bits, per, per32, radix};
    uint32_t* fa = S.buf<uint32_t>(B_GL_FA, n);
    sum_scan_of(S, ScanShort{lam, per32}, (uintptr_t)lam % 16 == 0, fa, n, false, gsc + GR_NA);
    S.read_fast(hv, gsc, GR_COUNT_);
    const uint32_t nA = hv[GR_NA], nB = (uint32_t)n - nA;
    uint64_t* K0 = S.buf<uint64_t>(B_SORT_K0, nB);
    uint64_t* K1 = S.buf<uint64_t>(B_SORT_K1, nB);
    uint32_t* V0 = S.buf<uint32_t>(B_SORT_V0, nB);
    uint32_t* V1 = S.buf<uint32_t>(B_SORT_V1, nB);
    // short labels: counted in a table of 2^bitsA keys, or (wider keys) sorted as 32-bit keys
    const bool table = bitsA <= kTabBitsMax;
    const uint32_t tabK = table ? (1u << bitsA) : 0u;
    uint32_t *KA0 = nullptr, *VA0 = nullptr, *tcnt = nullptr, *trep = nullptr, *tpos = nullptr;
    if (table) {
        tcnt = S.buf<uint32_t>(B_GL_KA0, tabK);
        trep = S.buf<uint32_t>(B_GL_VA0, tabK);
        tpos = S.buf<uint32_t>(B_GL_KA1, tabK);
        fill<uint32_t>(S, tcnt, tabK, 0u);
        fill<uint32_t>(S, trep, tabK, 0xFFFFFFFFu);
        k_gl_count<<<ceil_div(n, (uint64_t)kBlock * kCountPer), kBlock, 0, S.stream>>>(lk, codes, fa, tcnt, trep, V0);
        MIBWT_LAUNCH_CHECK();
    } else {
        KA0 = S.buf<uint32_t>(B_GL_KA0, nA);
        VA0 = S.buf<uint32_t>(B_GL_VA0, nA);
        k_gl_keys<<<grid_for(n), kBlock, 0, S.stream>>>(lk, codes, per32, fa, KA0, VA0, V0);
        MIBWT_LAUNCH_CHECK();
    }
    if (nB) k_gl_keys_long<<<grid_for(nB), kBlock, 0, S.stream>>>(lk, codes, V0, nB, K0);
    MIBWT_LAUNCH_CHECK();
    uint32_t* gid1 = S.buf<uint32_t>(B_DB_GID, n + 4);
    uint32_t* gofs = S.buf<uint32_t>(B_DB_GOFS, n + 8);
    if (table) {
        sum_scan_of(S, ScanNonzero{tcnt}, true, tpos, tabK, false, gsc + GR_DA);
    } else {
        uint32_t* KA1 = S.buf<uint32_t>(B_GL_KA1, nA);
        uint32_t* VA1 = S.buf<uint32_t>(B_GL_VA1, nA);
        int whereA = 0;  // an odd number of passes leaves the short list in buffer 1
        radix_sort<uint32_t, uint32_t>(S, KA0, KA1, VA0, VA1, nA, 0, bitsA,
                                       static_cast<const EmitPair<uint32_t, uint32_t>*>(nullptr), &whereA);
        if (whereA) {
            std::swap(KA0, KA1);
            std::swap(VA0, VA1);
        }
        // distinct keys: inclusive counts of the run heads, then the run starts
        sum_scan_of(S, ScanHeads<uint32_t>{KA0}, true, gid1, nA, true, gsc + GR_DA);
        if (nA) k_grp_offsets_gid<<<grid_for(nA), kBlock, 0, S.stream>>>(gid1, nA, gofs);
        MIBWT_LAUNCH_CHECK();
    }
    radix_sort<uint64_t, uint32_t>(S, K0, K1, V0, V1, nB, 0, 64);
    // the long list's group ids and starts behind the short list's (16-byte aligned)
    const uint32_t offB = table ? 0u : ((nA + 3) & ~3u);
    uint32_t* gidB = gid1 + offB;
    uint32_t* gofsB = gofs + offB + 4;
    sum_scan_of(S, ScanHeads<uint64_t>{K0}, true, gidB, nB, true, gsc + GR_DB);
    if (nB) k_grp_offsets_gid<<<grid_for(nB), kBlock, 0, S.stream>>>(gidB, nB, gofsB);
    MIBWT_LAUNCH_CHECK();
    // roots: the factorization chain, or the prefix maxima of i + lam[i] for many factors
    uint32_t* roots = S.buf<uint32_t>(B_GR_ROOT, kRootWalk);
    k_gl_rootwalk<<<1, 64, 0, S.stream>>>(lam, n, kRootWalk, roots, gsc + GR_S);
    MIBWT_LAUNCH_CHECK();
    S.read_fast(hv, gsc, GR_COUNT_);
    if (hv[GR_S] == kNone) {
        uint32_t* PM = S.buf<uint32_t>(B_LY_JMP0, n);
        uint32_t* rpos = S.buf<uint32_t>(B_LY_JMP1, n);
        k_gl_ends<<<grid_for(n), kBlock, 0, S.stream>>>(lam, n, PM);
        MIBWT_LAUNCH_CHECK();
        inclusive_max<uint32_t>(S, PM, PM, n);
        k_gl_rootflag<<<grid_for(n), kBlock, 0, S.stream>>>(PM, n, rpos);
        MIBWT_LAUNCH_CHECK();
        exclusive_sum<uint32_t>(S, rpos, rpos, n, gsc + GR_S);
        S.read_fast(hv, gsc, GR_COUNT_);
        roots = S.buf<uint32_t>(B_GR_ROOT, hv[GR_S]);
        k_gl_roots<<<grid_for(n), kBlock, 0, S.stream>>>(PM, rpos, n, roots);
        MIBWT_LAUNCH_CHECK();
    }
    const uint32_t DA = hv[GR_DA], DB = hv[GR_DB], NS = hv[GR_S];
    const uint32_t D = DA + DB;
    uint32_t* repA = S.buf<uint32_t>(B_GL_REPA, DA);
    uint32_t* cntA = S.buf<uint32_t>(B_GL_CNTA, DA);
    uint64_t* UA = S.buf<uint64_t>(B_GL_UA, DA);
    uint32_t* repB = S.buf<uint32_t>(B_GL_REPB, DB);
    uint32_t* cntB = S.buf<uint32_t>(B_GL_CNTB, DB);
    uint64_t* UB = S.buf<uint64_t>(B_GL_UB, DB);
    if (table) k_gl_tab_dedup<<<grid_for(tabK), kBlock, 0, S.stream>>>(tcnt, trep, tpos, tabK, repA, cntA, UA);
    else k_gl_dedup<uint32_t><<<grid_for(DA), kBlock, 0, S.stream>>>(KA0, VA0, gofs, DA, repA, cntA, UA);
    MIBWT_LAUNCH_CHECK();
    k_gl_dedup<uint64_t><<<grid_for(DB), kBlock, 0, S.stream>>>(K0, V0, gofsB, DB, repB, cntB, UB);
    MIBWT_LAUNCH_CHECK();
    uint32_t* rep = S.buf<uint32_t>(B_GR_REP, D);
    uint32_t* cnt = S.buf<uint32_t>(B_GR_CNT, D);
    uint64_t* U = S.buf<uint64_t>(B_GR_KEY, D);
    k_gl_merge<<<grid_for(D), kBlock, 0, S.stream>>>(UA, repA, cntA, DA, UB, repB, cntB, DB, U, rep, cnt);
    MIBWT_LAUNCH_CHECK();
    uint32_t* off = S.buf<uint32_t>(B_GR_OFF, (size_t)D + 1);
    k_gl_nchild<<<grid_for(D), kBlock, 0, S.stream>>>(rep, lam, D, off);
    MIBWT_LAUNCH_CHECK();
    exclusive_sum<uint32_t>(S, off, off, D, off + D);
    MIBWT_HIP(hipMemcpyAsync(gsc + GR_C, off + D, sizeof(uint32_t), hipMemcpyDeviceToDevice, S.stream));
    S.read_fast(hv, gsc, GR_COUNT_);
    const uint32_t C = hv[GR_C];
    const uint32_t Z = C + NS;  // slots; slot Z is the end of the text
    uint32_t* last = S.buf<uint32_t>(B_GR_LAST, D);
    uint32_t* chl = S.buf<uint32_t>(B_GR_CHL, Z);
    uint8_t* chc = S.buf<uint8_t>(B_GR_CHC, Z);
    uint32_t* cont = S.buf<uint32_t>(B_GR_CONT, Z);
    k_gl_children<<<grid_for(D), kBlock, 0, S.stream>>>(lk, codes, rep, off, D, U, chl, chc, cont, last, gsc + GR_ERR);
    MIBWT_LAUNCH_CHECK();
    k_gl_rootslots<<<grid_for(NS), kBlock, 0, S.stream>>>(lk, codes, roots, NS, C, U, D, chl, chc, cont, gsc + GR_ERR);
    MIBWT_LAUNCH_CHECK();
    // runs of equal children
    uint32_t* mrun = S.buf<uint32_t>(B_GR_RUN, Z);
    uint32_t* nxt = S.buf<uint32_t>(B_GR_NXT, (size_t)Z + 1);
    uint32_t* ccnt = S.buf<uint32_t>(B_GR_CCNT, (size_t)Z + 1);
    uint32_t* RV = S.buf<uint32_t>(B_GR_UCNT, Z);
    k_gl_brk<<<grid_for(Z), kBlock, 0, S.stream>>>(chl, cont, Z, RV);
    MIBWT_LAUNCH_CHECK();
    inclusive_max<uint32_t>(S, RV, RV, Z);
    k_gl_runs<<<grid_for((uint64_t)Z + 1), kBlock, 0, S.stream>>>(RV, cont, cnt, Z, mrun, nxt, ccnt);
    MIBWT_LAUNCH_CHECK();
    // up-contexts
    uint32_t* ucnt = RV;
    k_gl_upcount<<<grid_for(Z), kBlock, 0, S.stream>>>(chl, cont, off, last, Z, ucnt);
    MIBWT_LAUNCH_CHECK();
    exclusive_sum<uint32_t>(S, ucnt, ucnt, Z, gsc + GR_UT);
    S.read_fast(hv, gsc, GR_COUNT_);
    if (hv[GR_ERR] != 0) {  // inconsistent keys: only colliding fingerprints produce them
        S.phase_end(ph, 0);
        return 0;
    }
    const uint32_t UT = hv[GR_UT];
    uint32_t* UK0 = S.buf<uint32_t>(B_GR_UK0, UT);
    uint32_t* UK1 = S.buf<uint32_t>(B_GR_UK1, UT);
    uint32_t* UV0 = S.buf<uint32_t>(B_GR_UV0, UT);
    uint32_t* UV1 = S.buf<uint32_t>(B_GR_UV1, UT);
    uint32_t* uo = S.buf<uint32_t>(B_GR_UOFF, (size_t)D + 1);
    fill<uint32_t>(S, uo, (uint64_t)D + 1, 0u);
    if (UT > 0) {
        k_gl_upfill<<<grid_for(Z), kBlock, 0, S.stream>>>(chl, cont, off, last, Z, ucnt, UK0, UV0);
        MIBWT_LAUNCH_CHECK();
        radix_sort<uint32_t, uint32_t>(S, UK0, UK1, UV0, UV1, UT, 0, bits_for(D));
        k_gl_csr<<<grid_for(UT), kBlock, 0, S.stream>>>(UK0, UT, D, uo);
        MIBWT_LAUNCH_CHECK();
    }
    S.phase_end(ph, (double)n * (4 + 8 + 4 + 4) + (double)nA * 16 * 4 + (double)nB * 24 * 8 + (double)Z * 32);
    S.stats.distinct_phrases = D;
    S.stats.dict_length = C;
    S.stats.triggers = NS;
    S.stats.mixed_entries = UT;

    // ---- label order (the rule sort over distinct labels) and run keys
    ph = S.phase_begin("rsort");
    double rbytes = 0;
    uint32_t* Vs = nullptr;
    uint32_t* Hs = nullptr;
    if (!label_order<RepLab>(S, T, n, phh, RepLab{rep, lam}, D, codes, sigma, kLabSmallDistinct, &Vs, &Hs, &rbytes)) {
        S.phase_end(ph, rbytes);
        return 0;
    }
    uint32_t* lrank = S.buf<uint32_t>(B_GR_LRANK, D);
    {
        uint32_t* rk = S.buf<uint32_t>(B_DB_GID, D);
        inclusive_sum<uint32_t>(S, Hs, rk, D);
        k_gl_lrank<<<grid_for(D), kBlock, 0, S.stream>>>(Vs, rk, D, lrank);
        MIBWT_LAUNCH_CHECK();
    }
    uint32_t* rkey = S.buf<uint32_t>(B_GR_RKEY, (size_t)Z + 1);
    {
        const int mb = bits_for(Z);
        uint64_t* RK0 = S.buf<uint64_t>(B_SORT_K0, Z);
        uint64_t* RK1 = S.buf<uint64_t>(B_SORT_K1, Z);
        uint32_t* RV0 = S.buf<uint32_t>(B_SORT_V0, Z);
        uint32_t* RV1 = S.buf<uint32_t>(B_SORT_V1, Z);
        k_gl_rkinit<<<grid_for(Z), kBlock, 0, S.stream>>>(chl, mrun, lrank, Z, mb, RK0, RV0);
        MIBWT_LAUNCH_CHECK();
        radix_sort<uint64_t, uint32_t>(S, RK0, RK1, RV0, RV1, Z, 0, bits_for(D) + mb);
        uint32_t* h = S.buf<uint32_t>(B_LY_HF, Z);
        k_dbl_headflag<<<grid_for(Z), kBlock, 0, S.stream>>>(RK0, Z, h);
        MIBWT_LAUNCH_CHECK();
        inclusive_sum<uint32_t>(S, h, h, Z, gsc + GR_NR);
        k_gl_rkey<<<grid_for(Z), kBlock, 0, S.stream>>>(RV0, h, Z, rkey);
        MIBWT_LAUNCH_CHECK();
        rbytes += (double)Z * 24 * 8;
    }
    S.read_fast(hv, gsc, GR_COUNT_);
    const uint32_t NR = hv[GR_NR];
    S.phase_end(ph, rbytes);

    // ---- induction over the items
    ph = S.phase_begin("induce");
    const int kb = bits_for(std::max<uint64_t>(NR, NS));
    uint32_t m = Z, G = 1;
    uint32_t* IQ = S.buf<uint32_t>(B_IT_Q0, m);
    uint32_t* IG = S.buf<uint32_t>(B_IT_G0, m);
    uint32_t* ID = S.buf<uint32_t>(B_IT_ID0, m);
    uint32_t* ICH = S.buf<uint32_t>(B_IT_CH0, m);
    int cur = 0;
    k_it_init<<<grid_for(m), kBlock, 0, S.stream>>>(chc, Z, C, IQ, IG, ID, ICH);
    MIBWT_LAUNCH_CHECK();
    uint64_t budget = 8 * (uint64_t)n + (1u << 22);
    if (const char* e = getenv("MIBWT_LYN_ITEM_BUDGET")) budget = strtoull(e, nullptr, 10);  // tests
    uint64_t items_total = 0;
    double ibytes = 0;
    int rounds = 0;
    bool over = false;
    while (m > 0) {
        items_total += m;
        if (items_total > budget || rounds >= 4096) {
            over = true;
            break;
        }
        rounds++;
        uint32_t* HR = S.buf<uint32_t>(B_IT_HR, G);
        fill<uint32_t>(S, HR, G, 0u);
        k_it_hasrule<<<grid_for(m), kBlock, 0, S.stream>>>(IQ, ID, m, C, HR);
        MIBWT_LAUNCH_CHECK();
        uint64_t* IK0 = S.buf<uint64_t>(B_SORT_K0, m);
        uint64_t* IK1 = S.buf<uint64_t>(B_SORT_K1, m);
        uint32_t* IV0 = S.buf<uint32_t>(B_SORT_V0, m);
        uint32_t* IV1 = S.buf<uint32_t>(B_SORT_V1, m);
        k_it_key<<<grid_for(m), kBlock, 0, S.stream>>>(IQ, ID, m, C, Z, HR, rkey, kb, IK0, IV0);
        MIBWT_LAUNCH_CHECK();
        const int sbits = bits_for(G - 1) + kb;
        radix_sort<uint64_t, uint32_t>(S, IK0, IK1, IV0, IV1, m, 0, sbits);
        uint32_t* SQ = S.buf<uint32_t>(B_IT_SQ, m);
        uint32_t* SG = S.buf<uint32_t>(B_IT_SG, m);
        uint32_t* SCH = S.buf<uint32_t>(B_IT_SCH, m);
        uint32_t* P = S.buf<uint32_t>(B_IT_P, (size_t)m + 1);
        k_it_gather<<<grid_for(m), kBlock, 0, S.stream>>>(IV0, m, IQ, IG, ICH, ccnt, SQ, SG, SCH, P);
        MIBWT_LAUNCH_CHECK();
        exclusive_sum<uint32_t>(S, P, P, m, P + m);
        uint32_t* OF = S.buf<uint32_t>(B_IT_OF, G);
        uint32_t* NG1 = S.buf<uint32_t>(B_IT_NH, m);
        k_it_heads<<<grid_for(m), kBlock, 0, S.stream>>>(IK0, m, kb, OF, NG1);
        MIBWT_LAUNCH_CHECK();
        inclusive_sum<uint32_t>(S, NG1, NG1, m, gsc + GR_NG);
        uint32_t* GST = S.buf<uint32_t>(B_IT_GST, m);
        uint32_t* GFI = S.buf<uint32_t>(B_IT_GFI, m);
        uint32_t* GEXP = S.buf<uint32_t>(B_IT_GEXP, m);
        fill<uint32_t>(S, GEXP, m, 0u);
        k_it_groups<<<grid_for(m), kBlock, 0, S.stream>>>(IK0, m, kb, SG, P, OF, NG1, SCH, GST, GFI, GEXP);
        MIBWT_LAUNCH_CHECK();
        uint32_t* RF = S.buf<uint32_t>(B_IT_RF, m);   // run flags, then run indices
        uint32_t* EC = S.buf<uint32_t>(B_IT_EC, m);   // expansion counts, then output offsets
        k_it_flags<<<grid_for(m), kBlock, 0, S.stream>>>(m, NG1, GEXP, SQ, SCH, nxt, cont, uo, Z, GST, RF, EC, gsc);
        MIBWT_LAUNCH_CHECK();
        // dense ids of the refined groups (GEXP beyond the last group is 0), run indices, offsets
        uint32_t* XID = S.buf<uint32_t>(B_IT_XID, m);
        exclusive_sum<uint32_t>(S, GEXP, XID, m, gsc + GR_G);
        exclusive_sum<uint32_t>(S, RF, RF, m, gsc + GR_RUNS);
        exclusive_sum<uint32_t>(S, EC, EC, m, gsc + GR_M);
        uint32_t* RST = S.buf<uint32_t>(B_IT_RST, m);
        uint32_t* RLEN = S.buf<uint32_t>(B_IT_RLEN, m);  // lengths, then offsets
        uint8_t* RCH = S.buf<uint8_t>(B_IT_RCH, m);
        k_it_runs<<<grid_for(m), kBlock, 0, S.stream>>>(m, RF, NG1, GEXP, GST, GFI, P, SCH, gsc, RST, RLEN, RCH);
        MIBWT_LAUNCH_CHECK();
        S.read_fast(hv, gsc, GR_COUNT_);
        if (hv[GR_ERR] != 0) {  // an inconsistent grammar (colliding fingerprints): new base
            S.phase_end(ph, ibytes);
            return 0;
        }
        const uint32_t nruns = hv[GR_RUNS], mo = hv[GR_M], Go = hv[GR_G];
        if (nruns > 0) {
            exclusive_sum<uint32_t>(S, RLEN, RLEN, nruns, gsc + GR_RB);
            const uint32_t fb = std::max<uint32_t>(1, std::min<uint32_t>(2048, ceil_div(nruns, 16)));
            k_it_fill<<<fb, kBlock, 0, S.stream>>>(RLEN, RST, RCH, nruns, gsc + GR_RB, dO, n, gsc + GR_ERR);
            MIBWT_LAUNCH_CHECK();
        }
        if (S.verbose)
            fprintf(stderr, "[mibwt] lyndon induction round %d: %u items in %u groups, %u runs, %u items next\n",
                    rounds, m, G, nruns, mo);
        ibytes += (double)m * (24 * ((sbits + 7) / 8) + 64) + (double)mo * 28 + (double)nruns * 12;
        if (mo == 0) {
            m = 0;
            break;
        }
        const int nx = cur ^ 1;
        uint32_t* OQ = S.buf<uint32_t>(nx ? B_IT_Q1 : B_IT_Q0, mo);
        uint32_t* OG = S.buf<uint32_t>(nx ? B_IT_G1 : B_IT_G0, mo);
        uint32_t* OID = S.buf<uint32_t>(nx ? B_IT_ID1 : B_IT_ID0, mo);
        uint32_t* OCH = S.buf<uint32_t>(nx ? B_IT_CH1 : B_IT_CH0, mo);
        k_it_expand<<<grid_for(mo), kBlock, 0, S.stream>>>(m, mo, EC, SQ, SCH, NG1, GST, XID, nxt, cont, uo, UV0, OQ, OG,
                                                            OID, OCH);
        MIBWT_LAUNCH_CHECK();
        IQ = OQ;
        IG = OG;
        ID = OID;
        ICH = OCH;
        cur = nx;
        m = mo;
        G = Go;
    }
    S.stats.phrases = (int64_t)items_total;
    S.stats.doubling_rounds_parse = rounds;
    S.phase_end(ph, ibytes);
    if (over) return kNone;
    S.read_fast(hv, gsc, GR_COUNT_);
    if (hv[GR_ERR] != 0) return 0;  // a run past the end: an inconsistent grammar, new base
    const uint32_t pidx = hv[GR_PIDX];
    return (pidx >= 1 && pidx <= n) ? pidx : 0xFFFFFFFEu;  // an invalid index fails the inversion
}

// ------------------------------------------------------------------- 5. position-level induction

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

// The position-level induction: every position ordered by its label, then groups of equal labels
// refined by the rank at the next smaller suffix with pointer jumping.  It runs when the grammar
// induction outgrows its budget (or when MIBWT_LYN_INDUCTION=positions asks for it).  Returns the
// primary index, or 0 when a fingerprint collision broke the label order.
static uint32_t lyndon_positions(Storage& S, const uint8_t* T, uint64_t n, const PrefixHash& phh,
                                 const uint32_t* lam, const uint32_t* codes, uint32_t sigma, uint8_t* dO)
{
    uint32_t* sc = S.buf<uint32_t>(B_SCALARS, SC_COUNT_);
    // ---- 3. label order of every position
    int ph = S.phase_begin("rsort");
    double rbytes = 0;
    uint32_t* Vl = nullptr;
    uint32_t* hf = nullptr;
    if (!label_order<PosLab>(S, T, n, phh, PosLab{lam}, (uint32_t)n, codes, sigma, kLabSmall, &Vl, &hf, &rbytes)) {
        S.phase_end(ph, rbytes);
        return 0;
    }
    // label group ids as the initial keys of the induction
    uint32_t* gid1 = S.buf<uint32_t>(B_DB_GID, n);
    uint64_t* K0 = S.buf<uint64_t>(B_SORT_K0, n);
    uint32_t* V1 = S.buf<uint32_t>(B_SORT_V1, n);
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
    return (pidx >= 1 && pidx <= n) ? pidx : 0xFFFFFFFEu;  // an invalid index fails the inversion
}

// One build attempt with fingerprint base `base` (fingerprints compared under `hmask`):
// fingerprints, Lyndon array, then the grammar induction (or the position-level one).  Returns
// the primary index, or 0 when the label order check found a fingerprint collision.
static uint32_t lyndon_attempt(Storage& S, const uint8_t* T, uint64_t n, uint64_t base, uint64_t hmask,
                               uint32_t* codes, uint8_t* dO, bool positions)
{
    // ---- 1. prefix fingerprints
    int ph = S.phase_begin("rules");
    const PrefixHash phh = prefix_hashes(S, T, n, base, hmask);
    S.phase_end(ph, (double)n * (2 + 8));

    // ---- 2. Lyndon array
    ph = S.phase_begin("lyndon");
    uint32_t* lam = S.buf<uint32_t>(B_LY_LAM, n);
    const int levels = lyndon_array(S, T, n, phh, lam);
    alphabet_codes(S, T, n, codes);
    uint32_t sigma = 0;
    S.read_fast(&sigma, codes + 256);
    S.phase_end(ph, (double)n * (1 + 4 + 4 + 1) * 2);
    S.stats.doubling_rounds_dict = levels;
    S.stats.alpha_groups = sigma;

    if (!positions) {
        const uint32_t r = lyndon_grammar(S, T, n, phh, lam, codes, sigma, dO);
        if (r != kNone) {
            S.stats.induction = 1;
            return r;
        }
        if (S.verbose) fprintf(stderr, "[mibwt] lyndon: grammar induction over budget, position-level induction\n");
    }
    S.stats.induction = 2;
    return lyndon_positions(S, T, n, phh, lam, codes, sigma, dO);
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

    // The Lyndon array, the label keys and the label order trust Karp-Rabin fingerprints; the
    // finished BWT is therefore inverted on the device and compared with the text, and a mismatch
    // (a fingerprint collision) repeats the build with a new base.  The base is drawn per call.
    // MIBWT_LYN_HASH_BITS narrows the first attempt's fingerprints (tests);
    // MIBWT_LYN_INDUCTION=positions selects the position-level induction.
    uint64_t mask0 = ~0ull;
    if (const char* hb = getenv("MIBWT_LYN_HASH_BITS")) {
        const int v = atoi(hb);
        if (v >= 1 && v < 64) mask0 = (1ull << v) - 1;
    }
    const char* ind = getenv("MIBWT_LYN_INDUCTION");
    const bool positions = ind && strcmp(ind, "positions") == 0;
    S.lyn_calls++;
    uint64_t base = lyndon_base(S.lyn_calls ^ (uint64_t)std::chrono::steady_clock::now().time_since_epoch().count());
    uint32_t* codes = S.buf<uint32_t>(B_ALPHA, 272);
    uint32_t pidx = 0;
    for (int attempt = 0;; attempt++) {
        MIBWT_REQUIRE(attempt < 4, MIBWT_HASH_COLLISION, "Lyndon fingerprint collisions persist");
        if (attempt > 0) base = lyndon_base(base * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull);
        const uint64_t hmask = (attempt == 0) ? mask0 : ~0ull;
        S.stats.retries = attempt;  // the timing covers every attempt
        pidx = lyndon_attempt(S, T, n, base, hmask, codes, dO, positions);
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
    const int levels = lyndon_array(S, T, n, prefix_hashes(S, T, n, 0x1F3D5B79A2C4E681ull % kM61, ~0ull), lam);
    MIBWT_HIP(hipMemcpyAsync(lam_host, lam, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, S.stream));
    MIBWT_HIP(hipStreamSynchronize(S.stream));
    return levels;
}

}  // namespace mibwt
