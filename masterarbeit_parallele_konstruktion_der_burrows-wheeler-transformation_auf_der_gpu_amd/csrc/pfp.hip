// Prefix-free-parsing BWT on gfx950.
//
// Reference: Prefix_Free_Parsing_W=5_P=16_kommentiert.cu (Compute_D :2415, Compute_S :4330).
// The trigger set, the phrases and therefore the parse are exactly the reference's; the
// later phases compute the same BWT by a different, GPU-shaped route:
//
//  1. scan      trigger windows (value mod P == 0, or equal to T[0..W-1], never the window
//               at 0), one rolling Karp-Rabin value per thread over 16 window ends staged
//               in LDS, 16-bit flag masks, then an ordered emit of the end positions
//               (reference Compute_D1/D2 :443-571, D3/D4 :573-592).
//  2. phrases   64-bit fingerprint per phrase occurrence.
//  3. dict      LSD radix sort of the fingerprints; equal neighbours are compared byte by
//               byte (a collision re-seeds the hash); dense phrase ids, one representative
//               occurrence and the frequency per distinct phrase.
//  4. dsort     suffix sorting of the dictionary (all suffixes of all distinct phrases) by
//               prefix doubling with segment limits; equal suffixes of different phrases
//               end up in one group.  These groups are the reference's S set
//               (:2509-2592) and their order its S sort (:2767-3240); the order of the
//               whole phrases is the lexicographic phrase order (:1990-2164, 616-691).
//  5. psort     suffix array of the parse by prefix doubling (:2342-2412, 787-807).
//  6. assemble  each dictionary suffix alpha of length > W owns cnt(alpha) consecutive BWT
//               slots.  Groups whose preceding characters agree are filled directly (the
//               reference's single-character intervals, :3337-3350).  Mixed groups and
//               whole phrases are ordered by the rank of the next parse suffix
//               (:4217-4248): entries (alpha, char) are generated in parse-suffix order and
//               stably radix sorted by alpha; the last radix pass writes the BWT bytes.
//
// The terminator is virtual: positions >= n of the extended text T.$^W compare below every
// byte, so all 256 byte values are accepted (the reference needs bytes >= 2).
#include <cstdlib>

#include <cmath>

#include "doubling.hpp"
#include "text.hpp"

namespace mibwt {

constexpr int kTrigPerThread = 16;   // window ends per thread
constexpr int kTrigTile = kBlock * kTrigPerThread;
constexpr int kMaxW = 64;

// --------------------------------------------------------------------- helpers

// The texts of one build.  One text or many (a batch): text t lies at base[t] of the staged
// buffer, n_t = len[t] bytes followed by zero bytes up to the next text, which starts at the
// next multiple of kTextAlign at least kTextGap bytes later (text 0 at 0).  Each text has its
// own virtual terminator $^W: its last phrase ends at base + n_t + W - 1, and that end carries
// kSentinel in the phrase-end array E.  The BWT of text t lands at obase[t] of the output (the
// texts' BWTs back to back).
constexpr uint32_t kSentinel = 0x80000000u;
constexpr uint32_t kTextAlign = 8192;  // = the trigger scan's sub-tile (kScanTileW)
constexpr uint32_t kTextGap = 64;      // >= W: the sentinel end stays inside the text's range

struct TextSet {
    uint32_t count;
    const uint32_t* base;   // [count]
    const uint32_t* len;    // [count]
    const uint32_t* obase;  // [count + 1]
    const uint32_t* ptext;  // [d] text of every phrase occurrence (nullptr for one text)
    uint32_t n1;            // one text: its length (base == nullptr)
    __device__ __forceinline__ uint32_t text_of(uint32_t k) const { return ptext ? ptext[k] : 0u; }
};

// start of the text that follows the text whose last phrase ends at sentinel end `se`
__device__ __host__ __forceinline__ uint32_t next_text_base(uint32_t se, int W)
{
    const uint32_t text_end = se - (uint32_t)W + 1u;
    return (text_end + kTextGap + kTextAlign - 1) / kTextAlign * kTextAlign;
}

// Phrase k of the parse: [start, end] in the extended text; tl = its bytes inside the text
// (the last phrase of a text has W terminator symbols after them).
__device__ __forceinline__ void phrase_bounds(const uint32_t* __restrict__ E, uint32_t k, int W,
                                              uint32_t& start, uint32_t& end, uint32_t& tl)
{
    const uint32_t ek = E[k];
    end = ek & ~kSentinel;
    if (k == 0) {
        start = 0;
    } else {
        const uint32_t ep = E[k - 1];
        start = (ep & kSentinel) ? next_text_base(ep & ~kSentinel, W) : ep - (uint32_t)W + 1u;
    }
    tl = (ek & kSentinel) ? end - (uint32_t)W + 1u - start : end - start + 1u;
}

// whether occurrence k is the first phrase of its text
__device__ __forceinline__ bool first_of_text(const uint32_t* __restrict__ E, uint32_t k)
{
    return k == 0 || (E[k - 1] & kSentinel);
}

// 8 text bytes at word index k of a stream starting sh/8 bytes into word w[0].
__device__ __forceinline__ uint64_t funnel8(uint64_t lo, uint64_t hi, int sh)
{
    return sh ? ((lo >> sh) | (hi << (64 - sh))) : lo;
}

// Fingerprint of the phrase at s of `len` symbols, the first `tl` of them text bytes, the rest
// terminators: 8-byte chunks at offsets c = 0, 8, ... of the text bytes, each mixed with its
// offset, then the number of terminator symbols.  The chunks are streamed as aligned words, 32
// bytes per step with the four loads issued together (the padded text copy makes the
// over-read safe).
__device__ uint64_t phrase_hash(const uint8_t* __restrict__ T, uint32_t s, uint32_t len, uint32_t tl, uint64_t seed)
{
    const uint32_t L = tl;
    uint64_t h = seed ^ ((uint64_t)len * 0x9E3779B97F4A7C15ull);
    const uint64_t* w = reinterpret_cast<const uint64_t*>(T + (s & ~7u));
    const int sh = (int)(s & 7u) * 8;
    uint64_t prev = w[0];
    for (uint32_t i = 0; i < L; i += 32) {
        const uint32_t q = i >> 3;
        uint64_t a[4];
#pragma unroll
        for (int j = 0; j < 4; j++) a[j] = w[q + 1 + j];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t c = i + 8 * j;
            if (c < L) {
                uint64_t u = funnel8(j ? a[j - 1] : prev, a[j], sh);
                if (L - c < 8) u &= (~0ull) >> (64 - 8 * (L - c));
                h = mix64(h ^ (u + 0x2545F4914F6CDD1Dull * (c + 1)));
            }
        }
        prev = a[3];
    }
    if (len > tl) h = mix64(h ^ 0xA24BAED4963EE407ull ^ (uint64_t)(len - tl));
    return h;
}

// Byte equality of two phrases (start, length, text bytes).  Both are streamed as aligned
// words, 32 bytes per step with all eight loads issued together (the padded text copy makes the
// over-read safe).
__device__ bool phrase_equal(const uint8_t* __restrict__ T, uint32_t s1, uint32_t len1, uint32_t tl1, uint32_t s2,
                             uint32_t len2, uint32_t tl2)
{
    if (len1 != len2 || tl1 != tl2) return false;
    const uint32_t len = tl1;
    const uint64_t* w1 = reinterpret_cast<const uint64_t*>(T + (s1 & ~7u));
    const uint64_t* w2 = reinterpret_cast<const uint64_t*>(T + (s2 & ~7u));
    const int sh1 = (int)(s1 & 7u) * 8, sh2 = (int)(s2 & 7u) * 8;
    uint64_t p1 = w1[0], p2 = w2[0];
    for (uint32_t i = 0; i < len; i += 32) {
        const uint32_t q = i >> 3;
        uint64_t a[4], b[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            a[j] = w1[q + 1 + j];
            b[j] = w2[q + 1 + j];
        }
        uint64_t diff = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t u = funnel8(j ? a[j - 1] : p1, a[j], sh1);
            const uint64_t v = funnel8(j ? b[j - 1] : p2, b[j], sh2);
            const uint32_t c = i + 8 * j;
            const uint64_t m = (c >= len) ? 0ull : ((len - c >= 8) ? ~0ull : ((~0ull) >> (64 - 8 * (len - c))));
            diff |= (u ^ v) & m;
        }
        if (diff) return false;
        p1 = a[3];
        p2 = b[3];
    }
    return true;
}

// --------------------------------------------------------------------- 1. trigger scan

struct TrigTabs {
    uint16_t in_tab[256];   // c mod P
    uint16_t out_tab[256];  // -(c * 2^W) mod P
};

__device__ __forceinline__ uint32_t modP(uint32_t x, uint32_t P)
{
    // x < 4P
    x = (x >= 2 * P) ? x - 2 * P : x;
    x = (x >= P) ? x - P : x;
    return x;
}

// Flags of the 16 window ends [e0, e0+16) of one thread as a 16-bit mask; LDS `txt` holds
// the tile's text starting at tile_base - kMaxW.
__device__ __forceinline__ uint32_t trig_mask16(const uint8_t* txt, const uint8_t* first,
                                                const TrigTabs& tb, uint32_t h0, int W, uint32_t P,
                                                uint64_t e0, uint64_t n, uint32_t loc)
{
    // loc = local index of e0 in txt
    uint32_t h = 0;
    for (int k = 0; k < W; k++) h = modP(2 * h + tb.in_tab[txt[loc - W + 1 + k]], P);
    uint32_t mask = 0;
#pragma unroll
    for (int i = 0; i < kTrigPerThread; i++) {
        if (i > 0) {
            const uint32_t l = loc + i;
            h = modP(2 * h + tb.in_tab[txt[l]] + tb.out_tab[txt[l - W]], P);
        }
        const uint64_t e = e0 + i;
        bool trig = (h == 0);
        if (!trig && h == h0) {
            bool eq = true;
            for (int k = 0; k < W && eq; k++) eq = (txt[loc + i - W + 1 + k] == first[k]);
            trig = eq;
        }
        if (trig && e >= (uint64_t)W && e < n) mask |= 1u << i;
    }
    return mask;
}

__device__ void trig_load_tile(const uint8_t* __restrict__ T, uint64_t n, uint64_t tile_base,
                               uint8_t* txt, uint8_t* first, TrigTabs& tb, uint32_t& h0_out, int W,
                               uint32_t P, uint32_t* h0_lds)
{
    // text [tile_base - kMaxW, tile_base + kTrigTile) in 16-byte pieces
    const int pieces = (kTrigTile + kMaxW) / 16;
    for (int q = threadIdx.x; q < pieces; q += kBlock) {
        const int64_t g = (int64_t)tile_base - kMaxW + (int64_t)q * 16;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (g >= 0 && (uint64_t)g < n + kPad - 16) v = *reinterpret_cast<const uint4*>(T + g);
        *reinterpret_cast<uint4*>(txt + q * 16) = v;
    }
    if (threadIdx.x < kMaxW) first[threadIdx.x] = T[threadIdx.x];
    {
        const uint32_t c = threadIdx.x;  // kBlock == 256
        uint32_t p2w = 1 % P;
        for (int k = 0; k < W; k++) p2w = (p2w * 2) % P;
        tb.in_tab[c] = (uint16_t)(c % P);
        tb.out_tab[c] = (uint16_t)((P - (c * p2w) % P) % P);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t h = 0;
        for (int k = 0; k < W; k++) h = modP(2 * h + tb.in_tab[first[k]], P);
        *h0_lds = h;
    }
    __syncthreads();
    h0_out = *h0_lds;
}

// Pass 1: per-thread 16-bit trigger masks and per-tile counts.
__global__ __launch_bounds__(kBlock) void k_trig_mask(const uint8_t* __restrict__ T, uint64_t n, int W,
                                                     uint32_t P, uint16_t* __restrict__ masks,
                                                     uint32_t* __restrict__ tile_cnt)
{
    __shared__ __align__(16) uint8_t txt[kTrigTile + kMaxW];
    __shared__ uint8_t first[kMaxW];
    __shared__ TrigTabs tb;
    __shared__ uint32_t h0s;
    __shared__ uint32_t wsum[kWaves];
    const uint64_t tile_base = (uint64_t)blockIdx.x * kTrigTile;
    uint32_t h0;
    trig_load_tile(T, n, tile_base, txt, first, tb, h0, W, P, &h0s);
    const uint32_t loc = kMaxW + threadIdx.x * kTrigPerThread;
    const uint32_t m = trig_mask16(txt, first, tb, h0, W, P, tile_base + threadIdx.x * kTrigPerThread, n, loc);
    masks[(uint64_t)blockIdx.x * kBlock + threadIdx.x] = (uint16_t)m;
    uint32_t tot;
    (void)block_exclusive_scan((uint32_t)__popc(m), 0u, OpSum{}, wsum, &tot);
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = tot;
}

// Pass 2: ordered emit of trigger end positions.
__global__ __launch_bounds__(kBlock) void k_trig_emit(const uint16_t* __restrict__ masks,
                                                     const uint32_t* __restrict__ tile_off,
                                                     uint32_t* __restrict__ E)
{
    __shared__ uint32_t wsum[kWaves];
    const uint32_t m = masks[(uint64_t)blockIdx.x * kBlock + threadIdx.x];
    uint32_t tot;
    uint32_t off = block_exclusive_scan((uint32_t)__popc(m), 0u, OpSum{}, wsum, &tot) + tile_off[blockIdx.x];
    const uint32_t e0 = blockIdx.x * kTrigTile + threadIdx.x * kTrigPerThread;
    uint32_t mm = m;
    while (mm) {
        const int b = __ffs(mm) - 1;
        E[off++] = e0 + b;
        mm &= mm - 1;
    }
}

// Single-pass trigger scan.  The window value V = sum_k T[e-W+1+k] * 2^(W-1-k) is kept
// exactly (32 bits for W <= 24, 64 bits for W <= 56) and rolled per window end:
// V' = 2V + T[e] - T[e-W] * 2^W.  V mod P == 0 is tested without division: for
// P = 2^s * q (q odd), P | V  <=>  rotr(V * q^-1, s) <= floor((2^bits - 1) / P); for q = 1 it
// is a mask test.  The first-window rule needs V == V(T[0..W-1]), confirmed byte by byte.
// Each thread owns 32 consecutive window ends (its bytes arrive by two 16-byte loads), the
// outgoing bytes T[e-W] come from an LDS copy of the tile, and the tile's trigger count is
// chained to its predecessors by decoupled look-back, so the end positions are written in
// order in the same pass.
constexpr int kScanPer = 32;
constexpr int kScanTileW = kBlock * kScanPer;  // window ends per tile
constexpr int kScanHalo = 64;                   // bytes staged before the tile
constexpr int kScanMaxW = 56;
// Sub-tiles per chained tile: a launch shape (Storage::tune.scan_sub, 8 by default; smaller for
// short texts, whose few tiles would otherwise leave most of the chip idle).

struct TrigScan {
    const uint8_t* T;
    uint64_t n;       // staged length (all texts with their gaps, a multiple of kTextAlign)
    uint32_t n1;      // one text: its length (sub_text == nullptr)
    const uint32_t* sub_text;  // text of every sub-tile (nullptr: one text at 0)
    const uint32_t* tbase;     // per text: start and length
    const uint32_t* tlen;
    int W;
    uint32_t P;
    uint32_t pshift;  // P = 2^pshift * q, q odd
    uint64_t qinv;    // q^-1 mod 2^64
    uint64_t plim;    // floor((2^bits - 1) / P) for the value width in use
    uint32_t inv24;   // odd P: P^-1 mod 2^24
    uint32_t plim24;  // odd P: floor((2^24 - 1) / P) << 8 | 0xFF
    uint32_t tiles;
    uint32_t cap;     // capacity of E
    uint32_t* E;
    uint32_t* total;  // number of triggers (written by the last tile)
    Chain chain;
};

// Divisibility of the window value by P: kDiv 1 (P = 2^s) a mask test; kDiv 2 (odd P, window
// values below 2^24, i.e. W <= 16) P | v <=> v * P^-1 mod 2^24 <= floor((2^24 - 1) / P), with the
// full-rate 24-bit multiply (v_mul_u32_u24) whose low 24 bits are compared after a shift by 8
// (plim24 holds the bound << 8 | 0xFF); kDiv 0: P = 2^s q, rotr(v * q^-1, s) <= floor((2^b - 1) / P).
constexpr int kDivGeneral = 0, kDivPow2 = 1, kDivOdd24 = 2;
template <typename V, int kDiv>
__device__ __forceinline__ bool trig_divisible(V v, const TrigScan& p)
{
    if constexpr (kDiv == kDivPow2) return (v & (V)(p.P - 1)) == 0;
    if constexpr (kDiv == kDivOdd24) {
        // the compiler turns a 24-bit multiply of a small value into the quarter-rate 32-bit one
        uint32_t x;
        asm("v_mul_u32_u24 %0, %1, %2" : "=v"(x) : "v"((uint32_t)v), "s"(p.inv24));
        return (x << 8) <= p.plim24;
    }
    const V x = v * (V)p.qinv;
    V r;
    if (sizeof(V) == 4) r = (V)__builtin_rotateright32((uint32_t)x, p.pshift);
    else r = (V)__builtin_rotateright64((uint64_t)x, p.pshift);
    return r <= (V)p.plim;
}

// Trigger mask of the 32 window ends [e0, e0+32) of one thread; `txt` holds the current
// sub-tile's bytes from kScanHalo before its first window end (LDS), a/b the thread's own 32
// bytes; window ends lo <= e < hi of the sub-tile's text count, and its sentinel end `sent` is
// always set.
template <typename V, int kDiv>
__device__ __forceinline__ uint32_t trig_mask32(const uint8_t* txt, uint32_t loc, const uint4& a, const uint4& b,
                                                const TrigScan& p, V v0, const uint8_t* first, uint64_t e0,
                                                uint64_t lo, uint64_t hi, uint64_t sent)
{
    const int W = p.W;
    // value of the window ending at e0 - 1
    V v = 0;
    for (int k = 0; k < W; k++) v = (v << 1) + (V)txt[loc - W + k];
    // outgoing bytes T[e0 - W + i], i < 32, re-aligned from dword reads
    const uint32_t ob = loc - (uint32_t)W;
    const uint32_t oa = ob & ~3u, sh = ob & 3u;
    uint32_t ow[9];
#pragma unroll
    for (int m = 0; m < 9; m++) ow[m] = *reinterpret_cast<const uint32_t*>(txt + oa + 4 * m);
    uint32_t out[8];
#pragma unroll
    for (int m = 0; m < 8; m++) out[m] = __builtin_amdgcn_alignbyte(ow[m + 1], ow[m], sh);
    const uint32_t in[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};

    uint32_t mh = 0, mc = 0;  // bit 31 - i: window e0 + i (hash rule / value equals the first window's)
#pragma unroll
    for (int i = 0; i < kScanPer; i++) {
        const uint32_t cin = (in[i >> 2] >> (8 * (i & 3))) & 0xFFu;
        const uint32_t cout = (out[i >> 2] >> (8 * (i & 3))) & 0xFFu;
        v = (v << 1) + (V)cin - ((V)cout << W);
        if constexpr (kDiv == kDivOdd24) {
            // each test's lane mask goes straight into its bit mask as the carry of
            // mask + mask (two instructions per test instead of a compare, a select and a share of
            // the shifts and ors that the compiler otherwise emits)
            uint32_t x;
            asm("v_mul_u32_u24 %0, %1, %2" : "=v"(x) : "v"((uint32_t)v), "s"(p.inv24));
            x <<= 8;
            asm volatile("v_cmp_ge_u32_e32 vcc, %1, %2\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc"
                         : "+v"(mh)
                         : "s"(p.plim24), "v"(x)
                         : "vcc");
            asm volatile("v_cmp_eq_u32_e32 vcc, %1, %2\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc"
                         : "+v"(mc)
                         : "s"((uint32_t)v0), "v"((uint32_t)v)
                         : "vcc");
        } else {
            mh = (mh << 1) + (trig_divisible<V, kDiv>(v, p) ? 1u : 0u);
            mc = (mc << 1) + ((v == v0) ? 1u : 0u);
        }
    }
    uint32_t mask = __brev(mh);
    uint32_t cand = __brev(mc) & ~mask;
    // first-window rule: confirm the candidates byte by byte
    while (cand) {
        const int i = __ffs(cand) - 1;
        cand &= cand - 1;
        bool eq = true;
        for (int k = 0; k < W && eq; k++) eq = (txt[loc + i - W + 1 + k] == first[k]);
        if (eq) mask |= 1u << i;
    }
    // window ends lo <= e < hi (W <= e - base < n_t), then the sentinel end
    const uint64_t ia = lo > e0 ? lo - e0 : 0, ib = hi > e0 ? hi - e0 : 0;
    const uint32_t below_b = ib >= 32 ? ~0u : ((1u << (uint32_t)ib) - 1u);
    const uint32_t below_a = ia >= 32 ? ~0u : ((1u << (uint32_t)ia) - 1u);
    mask &= below_b & ~below_a;
    if (sent >= e0 && sent < e0 + kScanPer) mask |= 1u << (uint32_t)(sent - e0);
    return mask;
}

// Lanes of a 4-byte word that are zero, as 4 bits (exact per lane: no borrow between lanes).
__device__ __forceinline__ uint32_t zero_lanes4(uint32_t x)
{
    const uint32_t z = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
    return ((z >> 7) * 0x01020408u) >> 24;  // bits 7, 15, 23, 31 -> bits 0..3
}

// The same mask for P = 2^s with s <= 5, four window ends per 32-bit word and no window value:
// mod 2^s only the last min(W, s) bytes count, byte T[e-j] with its low s-j bits and weight 2^j,
// so the lanes of sum_j (w_j & low_{s-j}) << j (w_j = the word of the bytes j positions back) are
// the hash values; each lane stays below 5 * 32 < 256, so no carry leaves its byte.  First-window
// candidates are the lanes whose last min(W, 4) bytes equal the first window's (then confirmed
// byte by byte).
__device__ __forceinline__ uint32_t trig_mask32_swar(const uint8_t* txt, uint32_t loc, const uint4& a, const uint4& b,
                                                     const TrigScan& p, const uint8_t* first, uint64_t e0,
                                                     uint64_t lo, uint64_t hi, uint64_t sent)
{
    const int W = p.W;
    const uint32_t sh = p.pshift;
    const int nt = W < (int)sh ? W : (int)sh;  // hash terms
    const int nf = W < 4 ? W : 4;              // bytes compared with the first window
    uint32_t keep[5], fb[4];
#pragma unroll
    for (int j = 0; j < 5; j++) keep[j] = (j < nt) ? 0x01010101u * ((1u << (sh - j)) - 1u) : 0u;
#pragma unroll
    for (int j = 0; j < 4; j++) fb[j] = (j < nf) ? 0x01010101u * first[W - 1 - j] : 0u;
    const uint32_t low = 0x01010101u * ((1u << sh) - 1u);
    const uint32_t in[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t prev = *reinterpret_cast<const uint32_t*>(txt + loc - 4);  // bytes e0-4 .. e0-1
    uint32_t mh = 0, mc = 0;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const uint32_t w0 = in[m];
        const uint32_t w1 = __builtin_amdgcn_alignbyte(w0, prev, 3);  // lane i: byte e-1
        const uint32_t w2 = __builtin_amdgcn_alignbyte(w0, prev, 2);
        const uint32_t w3 = __builtin_amdgcn_alignbyte(w0, prev, 1);
        const uint32_t ws[5] = {w0, w1, w2, w3, prev};  // prev: lane i = byte e-4
        uint32_t h = 0;
#pragma unroll
        for (int j = 0; j < 5; j++) h += (ws[j] & keep[j]) << j;
        h &= low;
        uint32_t ne = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) ne |= (j < nf) ? (ws[j] ^ fb[j]) : 0u;
        mh |= zero_lanes4(h) << (4 * m);
        mc |= zero_lanes4(ne) << (4 * m);
        prev = w0;
    }
    uint32_t mask = mh;
    uint32_t cand = mc & ~mask;
    while (cand) {
        const int i = __ffs(cand) - 1;
        cand &= cand - 1;
        bool eq = true;
        for (int k = 0; k < W && eq; k++) eq = (txt[loc + i - W + 1 + k] == first[k]);
        if (eq) mask |= 1u << i;
    }
    const uint64_t ia = lo > e0 ? lo - e0 : 0, ib = hi > e0 ? hi - e0 : 0;
    const uint32_t below_b = ib >= 32 ? ~0u : ((1u << (uint32_t)ib) - 1u);
    const uint32_t below_a = ia >= 32 ? ~0u : ((1u << (uint32_t)ia) - 1u);
    mask &= below_b & ~below_a;
    if (sent >= e0 && sent < e0 + kScanPer) mask |= 1u << (uint32_t)(sent - e0);
    return mask;
}

// One tile = kScanSub sub-tiles of kScanTileW window ends.  Each sub-tile is staged in LDS
// (own bytes by two 16-byte loads per thread plus a 64-byte halo) and turned into one 32-bit
// mask per thread; the masks wait in LDS until the tile's offset is known.  Eight sub-tiles
// per tile balance the number of chained tiles (one ticket each) against the LDS that limits
// the resident blocks (measured faster than 2, 4, 6, 12, 16 or 32 on 256 MiB and 1 GiB texts).
template <typename V, int kDiv, int kScanSub, bool kSwar = false>
__global__ __launch_bounds__(kBlock) void k_trig_scan(TrigScan p)
{
    __shared__ __align__(16) uint8_t txt[kScanHalo + kScanTileW + 16];
    __shared__ uint8_t first[64];
    __shared__ uint32_t masks[kScanSub][kBlock];
    __shared__ uint32_t sub_sent[kScanSub];  // sentinel end of each sub-tile's text
    __shared__ V v0s;
    __shared__ uint32_t slot;
    __shared__ uint32_t wsum[kWaves];
    const int W = p.W;
    const uint32_t tile = chain_ticket(p.chain, p.tiles, &slot);
    const uint32_t loc = kScanHalo + threadIdx.x * kScanPer;
    uint32_t cnt = 0;
    uint32_t cur_text = 0xFFFFFFFFu;
#pragma unroll 1
    for (int sub = 0; sub < kScanSub; sub++) {
        const uint64_t gsub = (uint64_t)tile * kScanSub + sub;
        const uint64_t sb = gsub * kScanTileW;
        const uint64_t e0 = sb + (uint64_t)threadIdx.x * kScanPer;
        uint4 a = make_uint4(0, 0, 0, 0), b = make_uint4(0, 0, 0, 0);
        if (e0 < p.n) {
            a = *reinterpret_cast<const uint4*>(p.T + e0);
            b = *reinterpret_cast<const uint4*>(p.T + e0 + 16);
        }
        // the sub-tile's text (uniform across the block)
        uint32_t t = 0, base = 0, nt = p.n1;
        if (p.sub_text && sb < p.n) {
            t = p.sub_text[gsub];
            base = p.tbase[t];
            nt = p.tlen[t];
        }
        if (sub) __syncthreads();  // the previous sub-tile's LDS reads are done
        if (t != cur_text && threadIdx.x < 64) {
            // value of the text's first window T[base .. base+W-1], reduced over wavefront 0
            const int k = threadIdx.x;
            const uint8_t c = (k < W && (uint32_t)k < nt) ? p.T[base + k] : 0;
            first[k] = c;
            V x = (k < W) ? ((V)c << (W - 1 - k)) : (V)0;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                if (sizeof(V) == 4) x += (V)__shfl_xor((uint32_t)x, o, 64);
                else x += (V)__shfl_xor((unsigned long long)x, o, 64);
            }
            if (k == 0) v0s = x;
        }
        cur_text = t;
        const uint64_t sent = (uint64_t)base + nt + W - 1;
        if (threadIdx.x == 0) sub_sent[sub] = (uint32_t)sent;
        *reinterpret_cast<uint4*>(txt + loc) = a;
        *reinterpret_cast<uint4*>(txt + loc + 16) = b;
        if (threadIdx.x < kScanHalo / 16) {
            const int64_t g = (int64_t)sb - kScanHalo + (int64_t)threadIdx.x * 16;
            uint4 h = make_uint4(0, 0, 0, 0);
            if (g >= 0 && (uint64_t)g < p.n + kPad) h = *reinterpret_cast<const uint4*>(p.T + g);
            *reinterpret_cast<uint4*>(txt + threadIdx.x * 16) = h;
        } else if (threadIdx.x == kScanHalo / 16) {
            *reinterpret_cast<uint4*>(txt + kScanHalo + kScanTileW) = make_uint4(0, 0, 0, 0);
        }
        __syncthreads();
        uint32_t m = 0;
        if (sb < p.n) {
            if constexpr (kSwar)
                m = trig_mask32_swar(txt, loc, a, b, p, first, e0, (uint64_t)base + W, (uint64_t)base + nt, sent);
            else
                m = trig_mask32<V, kDiv>(txt, loc, a, b, p, v0s, first, e0, (uint64_t)base + W, (uint64_t)base + nt,
                                          sent);
        }
        masks[sub][threadIdx.x] = m;
        cnt += (uint32_t)__popc(m);
    }
    // order inside the tile: sub-tile major, then thread; per-sub-tile offsets from scans
    uint32_t tot_tile;
    (void)block_exclusive_scan(cnt, 0u, OpSum{}, wsum, &tot_tile);
    const uint32_t base = chain_prefix(p.chain, tile, tot_tile, OpSum{}, 0u, &slot);
    if (tile == p.tiles - 1 && threadIdx.x == 0) *p.total = base + tot_tile;
    uint32_t run = base;
#pragma unroll 1
    for (int sub = 0; sub < kScanSub; sub += 2) {
        // two sub-tiles per scan: counts are at most 32 * kBlock < 2^16
        const uint32_t m0 = masks[sub][threadIdx.x], m1 = (sub + 1 < kScanSub) ? masks[sub + 1][threadIdx.x] : 0u;
        const uint32_t pair = (uint32_t)__popc(m0) | ((uint32_t)__popc(m1) << 16);
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(pair, 0u, OpSum{}, wsum, &tot);
        const uint32_t tot0 = tot & 0xFFFFu, tot1 = tot >> 16;
        for (int h = 0; h < 2; h++) {
            uint32_t o = run + (h ? tot0 + (ex >> 16) : (ex & 0xFFFFu));
            uint32_t mm = h ? m1 : m0;
            const uint64_t e0 = ((uint64_t)tile * kScanSub + sub + h) * kScanTileW + (uint64_t)threadIdx.x * kScanPer;
            const uint32_t sent = (sub + h < kScanSub) ? sub_sent[sub + h] : 0u;
            while (mm) {
                const int i = __ffs(mm) - 1;
                mm &= mm - 1;
                const uint32_t e = (uint32_t)(e0 + i);
                if (o < p.cap) p.E[o] = e | (e == sent ? kSentinel : 0u);
                o++;
            }
        }
        run += tot0 + tot1;
    }
}

// Modular inverse of an odd q modulo 2^64 (Newton iteration).
inline uint64_t inverse_mod_2_64(uint64_t q)
{
    uint64_t x = q;  // correct to 3 bits
    for (int i = 0; i < 5; i++) x *= 2 - q * x;
    return x;
}

// Phrase ends in order into E (capacity cap): the trigger ends of every text and its sentinel
// end (flagged kSentinel); *total_dev receives their number, which may exceed cap (then E holds
// the first cap of them).  n = staged length (a multiple of kTextAlign); one text of n1 bytes
// (sub_text == nullptr) or many (sub_text/tbase/tlen on the device).  Requires W <= kScanMaxW.
uint32_t trig_scan(Storage& S, const uint8_t* T, uint64_t n, uint32_t n1, const uint32_t* sub_text,
                   const uint32_t* tbase, const uint32_t* tlen, int W, uint32_t P, uint32_t* E, uint32_t cap,
                   uint32_t* total_dev)
{
    TrigScan p{};
    p.T = T;
    p.n = n;
    p.n1 = n1;
    p.sub_text = sub_text;
    p.tbase = tbase;
    p.tlen = tlen;
    p.W = W;
    p.P = P;
    uint32_t s = 0, q = P;
    while ((q & 1u) == 0) {
        q >>= 1;
        s++;
    }
    p.pshift = s;
    p.qinv = inverse_mod_2_64(q);
    const bool narrow = W <= 24;
    p.plim = narrow ? (uint64_t)(0xFFFFFFFFull / P) : (~0ull) / P;
    const int sub = S.tune.scan_sub;
    p.tiles = (uint32_t)ceil_div(n, (uint64_t)kScanTileW * sub);
    p.cap = cap;
    p.E = E;
    p.total = total_dev;
    p.chain = S.chain(p.tiles);
    const bool pow2 = (q == 1);
    // odd P and windows of at most 16 bytes (values below 2^24): the 24-bit test
    const bool odd24 = (s == 0) && W <= 16;
    p.inv24 = (uint32_t)(p.qinv & 0xFFFFFFu);
    p.plim24 = ((uint32_t)(0xFFFFFFu / P) << 8) | 0xFFu;
    // P = 2^s with s <= 5: four window ends per word (MIBWT_SCAN_SWAR=0 takes the rolling value)
    const char* sw = getenv("MIBWT_SCAN_SWAR");
    const bool swar = pow2 && s <= 5 && !(sw && atoi(sw) == 0);
    auto launch = [&](auto sub_tag) {
        constexpr int kSub = decltype(sub_tag)::value;
        if (swar) k_trig_scan<uint32_t, kDivPow2, kSub, true><<<p.tiles, kBlock, 0, S.stream>>>(p);
        else if (narrow && pow2) k_trig_scan<uint32_t, kDivPow2, kSub><<<p.tiles, kBlock, 0, S.stream>>>(p);
        else if (odd24) k_trig_scan<uint32_t, kDivOdd24, kSub><<<p.tiles, kBlock, 0, S.stream>>>(p);
        else if (narrow) k_trig_scan<uint32_t, kDivGeneral, kSub><<<p.tiles, kBlock, 0, S.stream>>>(p);
        else if (pow2) k_trig_scan<uint64_t, kDivPow2, kSub><<<p.tiles, kBlock, 0, S.stream>>>(p);
        else k_trig_scan<uint64_t, kDivGeneral, kSub><<<p.tiles, kBlock, 0, S.stream>>>(p);
    };
    if (sub == 1) launch(std::integral_constant<int, 1>{});
    else if (sub == 2) launch(std::integral_constant<int, 2>{});
    else if (sub == 4) launch(std::integral_constant<int, 4>{});
    else launch(std::integral_constant<int, 8>{});
    MIBWT_LAUNCH_CHECK();
    return p.tiles;
}

// --------------------------------------------------------------------- 2/3. dictionary

// Also records the character before each phrase (the last character of its text for the
// first phrase of a text), which the whole-phrase BWT entries need later: it lies in the previous
// phrase's bytes, which a neighbouring lane reads at the same time.  With many texts the text id
// goes into the fingerprint: equal strings of different texts are different phrases.
__global__ void k_phrase_hash(const uint8_t* __restrict__ T, TextSet ts, int W, const uint32_t* __restrict__ E,
                              uint32_t d, uint64_t seed, uint64_t hmask, uint64_t* __restrict__ H,
                              uint32_t* __restrict__ V, uint8_t* __restrict__ pc)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d) return;
    uint32_t s, e, tl;
    phrase_bounds(E, k, W, s, e, tl);
    uint64_t sd = seed;
    if (ts.ptext) sd ^= mix64(0x51ED270B27A1F3ADull + ts.ptext[k]);
    H[k] = phrase_hash(T, s, e - s + 1, tl, sd) & hmask;
    V[k] = k;
    if (first_of_text(E, k)) {
        const uint32_t t = ts.text_of(k);
        pc[k] = T[ts.base ? ts.base[t] + ts.len[t] - 1 : ts.n1 - 1];
    } else {
        pc[k] = T[s - 1];
    }
}

// Phrase fingerprints are first sorted by their low 48 bits (6 radix passes).  Neighbours
// with equal sort keys are taken as equal phrases: different 64-bit fingerprints under one sort
// key (bit 1 of *collision) repeat the sort on all 64 bits; every occurrence is then compared
// byte by byte with its phrase's representative (k_dict_verify), and equal fingerprints of
// different phrases (bit 0) re-seed the hash.
constexpr int kDictKeyBits = 48;

__global__ void k_dict_mark(const uint64_t* __restrict__ H, uint32_t d, uint64_t key_mask,
                            uint32_t* __restrict__ flag, uint32_t* __restrict__ collision)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d) return;
    uint32_t f = 1;
    if (j > 0) {
        const uint64_t a = H[j], b = H[j - 1];
        if ((a & key_mask) == (b & key_mask)) {
            f = 0;
            if (a != b) atomicOr(collision, 2u);
        }
    }
    flag[j] = f;
}

// Occurrence k (text order: its own bytes stream) against the representative of its phrase
// (few distinct phrases, cache resident); a mismatch, or a representative of another text, is
// a fingerprint collision (bit 0).
__global__ void k_dict_verify(const uint8_t* __restrict__ T, TextSet ts, int W, const uint32_t* __restrict__ E,
                              uint32_t d, const uint32_t* __restrict__ pid, const uint32_t* __restrict__ dstart,
                              const uint32_t* __restrict__ dlen, const uint32_t* __restrict__ dtl,
                              const uint32_t* __restrict__ dtext, uint32_t* __restrict__ collision)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d) return;
    uint32_t s, e, tl;
    phrase_bounds(E, k, W, s, e, tl);
    const uint32_t u = pid[k];
    const uint32_t s2 = dstart[u];
    if (ts.ptext && ts.ptext[k] != dtext[u]) atomicOr(collision, 1u);
    else if (s2 != s && !phrase_equal(T, s, e - s + 1, tl, s2, dlen[u], dtl[u])) atomicOr(collision, 1u);
}

// ids[j] = inclusive count of heads (1-based id of the distinct phrase at sorted position j)
__global__ void k_dict_assign(const uint32_t* __restrict__ ids, const uint32_t* __restrict__ V,
                              const uint32_t* __restrict__ flag, const uint32_t* __restrict__ E, TextSet ts,
                              uint32_t d, int W, uint32_t* __restrict__ pid, uint32_t* __restrict__ head,
                              uint32_t* __restrict__ du_start, uint32_t* __restrict__ du_len,
                              uint32_t* __restrict__ du_tl, uint32_t* __restrict__ du_text)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d) return;
    const uint32_t u = ids[j] - 1;
    const uint32_t k = V[j];
    pid[k] = u;
    if (flag[j]) {
        uint32_t s, e, tl;
        phrase_bounds(E, k, W, s, e, tl);
        head[u] = j;
        du_start[u] = s;
        du_len[u] = e - s + 1;
        du_tl[u] = tl;
        du_text[u] = ts.text_of(k);
    }
}

__global__ void k_dict_freq(const uint32_t* __restrict__ head, uint32_t du, uint32_t d,
                            uint32_t* __restrict__ freq)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= du) return;
    freq[u] = ((u + 1 < du) ? head[u + 1] : d) - head[u];
}

// --------------------------------------------------------------------- 4/5. prefix doubling

// Segment limits of a string family.  Dictionary: the suffix at DT position x ends at the
// end of its phrase.  Parse: one string of length d.
// Per dictionary position x: its assembly record (see kRecAlpha) and the end (exclusive) of
// its phrase, side by side so that one gather by x serves both.
struct alignas(16) DictPos {
    uint64_t rec;
    uint32_t lim;
    uint32_t pad;
};

struct LimDict {
    static constexpr bool kCarry = true;  // the rest length travels with the element
    const DictPos* pos;
    // Payload above the initial key bits (pay == nullptr: none), from bit pshift: with fbits > 0
    // the record: the frequency in fbits bits (all ones: too large, read the DictPos), the
    // preceding character (8 bits), then the alpha and whole-phrase flags; with rbits > 0 the
    // rest length of the suffix in rbits bits (all ones: too large).
    const uint64_t* pay = nullptr;
    int pshift = 64;
    int fbits = 0;
    int rbits = 0;
    __device__ __forceinline__ uint32_t operator()(uint32_t x) const { return pos[x].lim; }
};
struct LimParse {
    static constexpr bool kCarry = false;
    uint32_t d;
    __device__ __forceinline__ uint32_t operator()(uint32_t) const { return d; }
};

__global__ void k_dt_owner(const uint64_t* __restrict__ off, const uint32_t* __restrict__ len,
                           uint32_t du, uint32_t* __restrict__ owner)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= du) return;
    const uint64_t o = off[u];
    for (uint32_t j = 0; j < len[u]; j++) owner[o + j] = u;
}

// Initial key of DT position x: `per` symbols of `bits` bits each, most significant first;
// symbol code 0 = past the phrase end, 1 = terminator, 2.. = dense code of the byte (codes
// follow byte order, so key order is string order).

// Initial parse key: `per` consecutive phrase ranks (+1; 0 past the end) as digits in base
// du + 1.
__global__ void k_parse_initkey(const uint32_t* __restrict__ lexrank, const uint32_t* __restrict__ pid,
                                uint32_t d, uint64_t base, int per, uint64_t* __restrict__ K, uint32_t* __restrict__ V)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d) return;
    uint64_t key = 0;
    for (int i = 0; i < per; i++) {
        const uint64_t s = (k + i < d) ? (uint64_t)lexrank[pid[k + i]] + 1 : 0;
        key = key * base + s;
    }
    K[k] = key;
    V[k] = k;
}

// Family of strings compared h symbols at a time: suffix x continues at x + h inside its
// segment [.., lim(x)); a group of equal h-prefixes is final once its strings are shorter
// than h (they are then equal).
// With Lim::kCarry the element carries r = lim(x) - x (the rest of its string) and no round
// reads lim.
template <typename Lim, typename Place = PlaceSA>
struct StringPolicy {
    static constexpr bool kCarry = Lim::kCarry;
    static constexpr bool kSpeculate = true;      // advancing is host-only (h doubles)
    static constexpr bool kAdvanceFirst = false;  // open() of this round uses the doubled h
    Lim lim;
    uint32_t h;
    Place place_fn;
    __device__ __forceinline__ void place(uint32_t* SA, uint32_t pos, uint32_t x, uint32_t head) const
    {
        place_fn(SA, pos, x, head);
    }
    __device__ __forceinline__ uint32_t rem(uint32_t x) const { return lim(x) - x; }
    __device__ __forceinline__ uint32_t next(uint32_t x, uint32_t r) const
    {
        if constexpr (kCarry) return h < r ? x + h : kNoNext;
        else return ((uint64_t)x + h < lim(x)) ? x + h : kNoNext;
    }
    __device__ __forceinline__ bool open(uint32_t x, uint32_t r) const
    {
        if constexpr (kCarry) return r >= h;
        else return lim(x) - x >= h;
    }
};

template <typename Lim, typename Place>
struct PolicyFor {
    using type = StringPolicy<Lim, Place>;
};

// Prefix doubling over N suffix positions: K0/V0 hold the initial keys (bits [0, key_bits),
// the first h0 symbols with a terminator padding) and positions; on exit rank[x] = 1 +
// position of x's group head and SA = the sorted order (equal strings share a group).
template <typename Lim, typename Place = PlaceSA>
DoublingResult prefix_doubling(Storage& S, uint64_t* K0, uint64_t* K1, uint32_t* V0, uint32_t* V1, uint32_t N,
                               int key_bits, uint32_t h0, Lim lim, uint32_t* rank, uint32_t* SA,
                               Place place = Place{})
{
    // the sorted pairs stay in whichever buffer the last pass wrote (no copy back)
    int where = 0;
    radix_sort<uint64_t, uint32_t>(S, K0, K1, V0, V1, N, 0, key_bits,
                                   static_cast<const EmitPair<uint64_t, uint32_t>*>(nullptr), &where);
    if (where) {
        std::swap(K0, K1);
        std::swap(V0, V1);
    }
    using Pol = typename PolicyFor<Lim, Place>::type;
    Pol pol{};
    pol.lim = lim;
    if constexpr (std::is_same_v<Lim, LimDict>)
        if (pol.lim.pay) pol.lim.pay = K0;  // the record payload travelled with the keys
    pol.h = h0;
    pol.place_fn = place;
    const uint64_t key_mask = key_bits >= 64 ? ~0ull : (1ull << key_bits) - 1;
    DoublingResult r =
        refine_groups(S, K0, V0, V1, N, pol, [](Pol& p, const uint32_t*, uint32_t) { p.h *= 2; }, rank, SA,
                      1 << 30, K1, key_mask);
    r.bytes += (double)N * 24 * ((key_bits + 7) / 8);
    return r;
}

// --------------------------------------------------------------------- 5. phrase order

// Lexicographic rank of phrase u: the number of whole phrases placed before its group head h
// (whole phrases are distinct strings, so no other whole phrase shares its group): word
// prefix counts of the placement bitmap plus the bits of h's word below h.
__global__ void k_popc_words(const uint32_t* __restrict__ bits, uint32_t words, uint32_t* __restrict__ cnt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words) cnt[i] = (uint32_t)__popc(bits[i]);
}

__global__ void k_lexrank(const uint32_t* __restrict__ rank, const uint64_t* __restrict__ off,
                          const uint32_t* __restrict__ bits, const uint32_t* __restrict__ wscan, uint32_t du,
                          uint32_t* __restrict__ lexrank)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= du) return;
    const uint32_t h = rank[(uint32_t)off[u]] - 1;
    lexrank[u] = wscan[h >> 5] + (uint32_t)__popc(bits[h >> 5] & ((1u << (h & 31)) - 1u));
}

// --------------------------------------------------------------------- 6. BWT assembly

// Per dictionary position x (dictionary order, coalesced): a packed record
//   bits  0..31  frequency of the phrase,
//   bits 32..39  character preceding the suffix inside the phrase (j >= 1),
//   bit  40      suffix longer than W (an alpha, owns BWT slots),
//   bit  41      whole phrase (j == 0).
constexpr uint64_t kRecAlpha = 1ull << 40, kRecWhole = 1ull << 41;

// Per dictionary position x (dictionary order): the initial sort key (the first `per` symbols
// of its suffix as digits in base `base` = sigma + 2: 0 past the phrase end, 1 terminator, then
// the dense codes in byte order; with many texts the text id of its phrase times tpow =
// base^per above them), V = x, and its DictPos (record: frequency, preceding character, flags;
// phrase end).  Mixed-radix digits order like the symbol strings and need fewer key bits than
// whole bits per symbol (DNA: 21 symbols in 55 bits instead of 63, one radix pass less).  The
// text id keeps the suffixes of different texts in different groups.
__global__ void k_dt_init(const uint8_t* __restrict__ T, const uint32_t* __restrict__ owner,
                          const uint64_t* __restrict__ off, const uint32_t* __restrict__ len,
                          const uint32_t* __restrict__ tlen, const uint32_t* __restrict__ dtext,
                          const uint32_t* __restrict__ start, const uint32_t* __restrict__ freq,
                          const uint32_t* __restrict__ code, uint32_t base, int per, uint64_t tpow, uint32_t L,
                          int W, int pshift, int fbits, int rbits, uint64_t* __restrict__ K,
                          uint32_t* __restrict__ V,
                          DictPos* __restrict__ dp)
{
    __shared__ uint32_t cs[256];
    cs[threadIdx.x] = code[threadIdx.x];  // blockDim == 256
    __syncthreads();
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= L) return;
    const uint32_t u = owner[x];
    const uint32_t o = (uint32_t)off[u];
    const uint32_t j = x - o;
    const uint32_t rem = len[u] - j;
    const uint32_t tr = (tlen[u] > j) ? tlen[u] - j : 0u;  // text bytes left in the suffix
    const uint64_t tp = (uint64_t)start[u] + j;
    uint64_t key = 0;
    uint64_t w = 0;
    for (int i = 0; i < per; i++) {
        if ((i & 7) == 0) w = ((uint32_t)i < tr) ? load8(T, tp + i) : 0ull;
        uint64_t s;
        if ((uint32_t)i >= rem) s = 0;
        else if ((uint32_t)i >= tr) s = 1;
        else s = cs[(w >> (8 * (i & 7))) & 0xFF];
        key = key * base + s;
    }
    if (tpow) key += (uint64_t)dtext[u] * tpow;  // the text id above the symbols
    uint64_t r = freq[u];
    if (rem > (uint32_t)W) r |= kRecAlpha;
    if (j == 0) r |= kRecWhole;
    else if (rem > (uint32_t)W) r |= (uint64_t)T[tp - 1] << 32;
    if (pshift < 64 && fbits) {
        // the record rides above the key bits through the sort (see LimDict)
        const uint32_t fm = (1u << fbits) - 1u, f = (uint32_t)r;
        const uint64_t pl = (uint64_t)(f < fm ? f : fm) | (((r >> 32) & 0xFFu) << fbits) |
                            ((uint64_t)((r & kRecAlpha) != 0) << (fbits + 8)) |
                            ((uint64_t)((r & kRecWhole) != 0) << (fbits + 9));
        key |= pl << pshift;
    } else if (pshift < 64) {
        // the rest length instead
        const uint32_t rm = (1u << rbits) - 1u;
        key |= (uint64_t)(rem < rm ? rem : rm) << pshift;
    }
    K[x] = key;
    V[x] = x;
    dp[x] = DictPos{r, o + len[u], 0};
}

// Final placement of a dictionary position by the suffix sort: its record, group head and
// whole-phrase flag go straight to its sorted slot (no suffix array to gather through later);
// the whole-phrase flags in sorted order give the lexicographic phrase order.
struct alignas(16) SortedRec {
    uint64_t r;     // record of the dictionary position placed here
    uint32_t head;  // sorted position of its group head
    uint32_t x;     // the dictionary position
};

struct PlaceRec {
    const DictPos* dp;
    SortedRec* out;
    uint32_t* wbits;  // bit per sorted position: a whole phrase is placed there
    __device__ __forceinline__ void put(uint32_t pos, uint32_t x, uint32_t head, uint64_t r) const
    {
        out[pos] = SortedRec{r, head, x};
        if (r & kRecWhole) atomicOr(wbits + (pos >> 5), 1u << (pos & 31));
    }
    __device__ __forceinline__ void operator()(uint32_t*, uint32_t pos, uint32_t x, uint32_t head) const
    {
        put(pos, x, head, dp[x].rec);
    }
};

// Dictionary suffixes: at the start one 16-byte read of a position's DictPos gives both its
// rest length and, when it is placed at once, its record.  A position alone in its initial group
// whose record came through the sort with its key (LimDict::pay) is placed without that read:
// in a dictionary of mostly distinct phrases that is nearly every position, and the read is a
// random access per position.  Where only the rest length fits above the key, a position that
// stays in a larger group takes it from there (its record is read when it is placed).
struct DictPolicy : StringPolicy<LimDict, PlaceRec> {
    __device__ __forceinline__ bool first(uint32_t*, uint32_t pos, uint32_t x, uint32_t s, uint32_t z,
                                          uint32_t& r) const
    {
        if (lim.pay && lim.rbits && z > 1) {
            const uint32_t rm = (1u << lim.rbits) - 1u;
            const uint32_t rr = (uint32_t)(lim.pay[pos] >> lim.pshift) & rm;
            if (rr != rm && open(x, rr)) {
                r = rr;
                return true;
            }
        }
        if (lim.pay && lim.fbits && z == 1) {
            const uint64_t w = lim.pay[pos] >> lim.pshift;
            const uint32_t fm = (1u << lim.fbits) - 1u, f = (uint32_t)w & fm;
            if (f != fm) {
                uint64_t rec = (uint64_t)f | (((w >> lim.fbits) & 0xFFu) << 32);
                if ((w >> (lim.fbits + 8)) & 1u) rec |= kRecAlpha;
                if ((w >> (lim.fbits + 9)) & 1u) rec |= kRecWhole;
                place_fn.put(pos, x, s, rec);
                r = 0;
                return false;
            }
        }
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        // rec lo, rec hi, lim, pad in one streaming read (no reuse: the next read of this
        // position comes rounds later)
        const u32x4 d = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(lim.pos + x));
        r = d.z - x;
        const bool kp = z > 1 && open(x, r);
        if (!kp) place_fn.put(pos, x, s, ((uint64_t)d.y << 32) | d.x);
        return kp;
    }
};
template <>
struct PolicyFor<LimDict, PlaceRec> {
    using type = DictPolicy;
};

// a group is mixed when a member's preceding character differs from its head's
__global__ void k_asm_mix(const SortedRec* __restrict__ RH, uint32_t L, uint8_t* __restrict__ gmix)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= L) return;
    const SortedRec a = RH[g];
    const uint64_t r = a.r;
    const uint32_t h = a.head;
    if (!(r & kRecAlpha) || (r & kRecWhole) || h == g) return;
    if (((r >> 32) & 0xFF) != ((RH[h].r >> 32) & 0xFF)) gmix[h] = 1;
}

// slots per sorted position (cnt), slots written through ordered entries (ecnt) and, for the
// positions whose slots are filled directly (single-character groups), 0x100 | character (fc).
// The heads of groups that receive entries are flagged (gflag, zeroed by the caller; every
// writer stores 1): numbered densely they key the entry sort with few bits.  The members of
// mixed groups are flagged at their dictionary positions (mxf, zeroed by the caller), which
// groups them by phrase for the entry generation.
__global__ void k_asm_counts(const SortedRec* __restrict__ RH, const uint8_t* __restrict__ gmix, uint32_t L,
                             uint32_t* __restrict__ cnt, uint32_t* __restrict__ ecnt, uint16_t* __restrict__ fc,
                             uint8_t* __restrict__ gflag, uint8_t* __restrict__ mxf)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= L) return;
    const SortedRec a = RH[g];
    const uint64_t r = a.r;
    const uint32_t f = (uint32_t)r;
    const bool alpha = (r & kRecAlpha) != 0;
    const bool mixed = alpha && !(r & kRecWhole) && gmix[a.head];
    const bool ordered = mixed || (alpha && (r & kRecWhole));
    if (mixed) mxf[a.x] = 1;
    cnt[g] = alpha ? f : 0u;
    ecnt[g] = ordered ? f : 0u;
    fc[g] = (alpha && !ordered) ? (uint16_t)(0x100u | ((uint32_t)(r >> 32) & 0xFFu)) : (uint16_t)0;
    if (ordered && f) gflag[a.head] = 1;
}

// Direct fill of the single-character groups, one wavefront per 64 * kFillPer consecutive
// sorted positions: their slot ranges are consecutive in the output, so the lanes sweep the wave's
// span together in 16-byte granules and look the runs up in LDS (one bisection per granule,
// then a forward walk).  A granule whose bytes are all fill bytes of this wave is one 16-byte
// store; otherwise its fill bytes are stored one by one (neighbouring waves and the later
// entry pass own the other bytes).  A run is [goff[g], goff[g + 1]) for the positions that
// fc marks.
constexpr int kFillPer = 4;  // sorted positions per lane: a wavefront covers 256 runs
__global__ __launch_bounds__(kBlock) void k_asm_fill(const uint32_t* __restrict__ goff,
                                                    const uint16_t* __restrict__ fc, uint32_t L, uint64_t n,
                                                    uint8_t* __restrict__ O)
{
    constexpr int kWavePos = 64 * kFillPer;
    __shared__ uint32_t rs[kBlock * kFillPer], rl[kBlock * kFillPer];
    __shared__ uint8_t rc[kBlock * kFillPer];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t wb = (uint32_t)w * kWavePos;  // this wave's runs in LDS
    const uint64_t g0 = (uint64_t)blockIdx.x * kBlock * kFillPer + wb;
    const bool aligned = ((uintptr_t)O & 15u) == 0;
#pragma unroll
    for (int i = 0; i < kFillPer; i++) {
        const uint64_t g = g0 + (uint64_t)i * 64 + lane;
        uint32_t st = 0, ln = 0;
        uint8_t ch = 0;
        if (g < L) {
            st = goff[g];
            const uint32_t f = fc[g];
            if (f) {
                ln = ((g + 1 < L) ? goff[g + 1] : (uint32_t)n) - st;
                ch = (uint8_t)f;
            }
        }
        rs[wb + i * 64 + lane] = st;
        rl[wb + i * 64 + lane] = ln;
        rc[wb + i * 64 + lane] = ch;
    }
    __syncthreads();
    // span of this wave: from its first run start to the end of its last run
    const uint32_t nv = (g0 >= L) ? 0u : (uint32_t)((L - g0) < (uint64_t)kWavePos ? (L - g0) : kWavePos);
    if (nv == 0) return;
    const uint32_t lo = rs[wb];
    const uint32_t hi_run = rs[wb + nv - 1] + rl[wb + nv - 1];
    const uint32_t hi = (hi_run < n) ? hi_run : (uint32_t)n;
    if (hi <= lo) return;
    for (uint32_t q = (lo >> 4) + lane; q <= ((hi - 1) >> 4); q += 64) {
        const uint32_t p0 = q << 4;
        const uint32_t first = (p0 > lo) ? p0 : lo;
        // last run start <= first
        int a = 0, b = (int)nv - 1;
        while (a < b) {
            const int m = (a + b + 1) >> 1;
            if (rs[wb + m] <= first) a = m;
            else b = m - 1;
        }
        // the fill bytes of the runs that overlap the granule (usually one or two), as a
        // 16-bit byte mask and four words
        uint32_t bytes[4] = {0, 0, 0, 0};
        uint32_t mask = 0;
        const uint32_t gend = (p0 + 16 < hi) ? p0 + 16 : hi;
        for (int r = a; r < (int)nv; r++) {
            const uint32_t st = rs[wb + r];
            if (st >= gend) break;
            const uint32_t ln = rl[wb + r];
            const uint32_t b0 = (st > first ? st : first) - p0;
            const uint32_t e = st + ln < gend ? st + ln : gend;
            if (ln == 0 || e <= p0 + b0) continue;
            const uint32_t m = ((1u << (e - p0)) - 1u) & ~((1u << b0) - 1u);
            mask |= m;
            const uint32_t c4 = (uint32_t)rc[wb + r] * 0x01010101u;
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const uint32_t nib = (m >> (4 * d)) & 0xFu;  // 4 byte flags -> byte mask
                bytes[d] |= c4 & ((((nib * 0x00204081u) & 0x01010101u)) * 0xFFu);
            }
        }
        if (mask == 0) continue;
        // a granule inside this wavefront's span holds only its own positions' slots: the bytes
        // that are not fill bytes belong to ordered groups, which the entry sort writes later,
        // so one 16-byte store may carry anything there
        if (aligned && (mask == 0xFFFFu || (p0 >= lo && p0 + 16 <= hi))) {
            *reinterpret_cast<uint4*>(O + p0) = make_uint4(bytes[0], bytes[1], bytes[2], bytes[3]);
        } else {
            while (mask) {
                const int k = __ffs(mask) - 1;
                mask &= mask - 1;
                O[p0 + k] = (uint8_t)(bytes[k >> 2] >> (8 * (k & 3)));
            }
        }
    }
}

// BWT entries: group (sorted position of the dictionary group head) above the output byte.
// 64-bit entries keep bit 8 for the primary-index mark (group from bit 9); 32-bit entries
// (groups below 2^24) hold group << 8 | byte and the primary index is counted separately.
template <typename EntT>
struct Ent {
    static constexpr int kShift = sizeof(EntT) == 4 ? 8 : 9;
    __device__ __forceinline__ static EntT make(uint32_t group, uint32_t c)
    {
        return ((EntT)group << kShift) | (EntT)c;
    }
    __device__ __forceinline__ static uint32_t group(EntT e) { return (uint32_t)(e >> kShift); }
};

template <typename EntT>
__global__ void k_mx_write(const DictPos* __restrict__ dp, const uint32_t* __restrict__ rank,
                           const uint8_t* __restrict__ mxf, const uint32_t* __restrict__ mxs,
                           const uint32_t* __restrict__ cid, uint32_t L, EntT* __restrict__ mx)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= L || !mxf[x]) return;
    mx[mxs[x]] = Ent<EntT>::make(cid[rank[x] - 1], (uint32_t)(dp[x].rec >> 32) & 0xFFu);
}

// slot offset minus entry offset of each entry group, by its dense id
__global__ void k_ent_delta(const uint8_t* __restrict__ gflag, const uint32_t* __restrict__ cid,
                            const uint32_t* __restrict__ goff, const uint32_t* __restrict__ escan, uint32_t L,
                            uint32_t* __restrict__ deltac)
{
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h < L && gflag[h]) deltac[cid[h]] = goff[h] - escan[h];
}

// per distinct phrase: first mixed entry, number of mixed entries, whole-phrase group
struct PhraseRec {
    uint32_t mx_begin, mx_count, alpha0, pad;
};

__global__ void k_phrase_rec(const uint64_t* __restrict__ off, const uint32_t* __restrict__ len,
                             const uint32_t* __restrict__ mxs, const uint32_t* __restrict__ rank,
                             const uint32_t* __restrict__ cid, uint32_t du, PhraseRec* __restrict__ pr)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= du) return;
    const uint32_t o = (uint32_t)off[u];
    const uint32_t a = mxs[o], b = mxs[o + len[u]];
    pr[u] = PhraseRec{a, b - a, cid[rank[o] - 1], 0};
}

// Per occurrence k (parse order, sequential reads): its whole-phrase entry and its phrase's
// run of mixed entries.
template <typename EntT>
struct alignas(16) OccRec {
    EntT whole;
    uint32_t mx_begin, mx_count;
};

// whole-phrase entry of occurrence k: its phrase's group above the character before it; with
// 64-bit entries the first phrase of every text carries the primary-index mark (bit 8)
template <typename EntT>
__device__ __forceinline__ EntT whole_phrase_entry(const uint32_t* __restrict__ E, const uint8_t* __restrict__ pchar,
                                                   uint32_t k, uint32_t alpha0)
{
    EntT v = Ent<EntT>::make(alpha0, pchar[k]);
    if (sizeof(EntT) == 8 && first_of_text(E, k)) v |= (EntT)0x100;
    return v;
}

template <typename EntT>
__global__ void k_occ_rec(const uint32_t* __restrict__ pid, uint32_t d, const PhraseRec* __restrict__ pr,
                          const uint8_t* __restrict__ pchar, const uint32_t* __restrict__ E,
                          OccRec<EntT>* __restrict__ occ)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d) return;
    const PhraseRec q = pr[pid[k]];
    occ[k] = OccRec<EntT>{whole_phrase_entry<EntT>(E, pchar, k, q.alpha0), q.mx_begin, q.mx_count};
}

// Per parse-suffix rank r (one gather of the occurrence k = psa[r]-1): the number of entries
// (its whole-phrase entry and one per mixed suffix of its phrase; none for the suffix that
// starts at phrase 0, psa[r] == 0) and, in rank order, its whole-phrase entry and first
// mixed entry.
template <typename EntT>
struct alignas(sizeof(EntT) == 4 ? 8 : 16) RankRec {
    EntT whole;
    uint32_t mx_begin;
};

template <typename EntT>
__global__ void k_q_count(const uint32_t* __restrict__ psa, uint32_t d, const OccRec<EntT>* __restrict__ occ,
                          uint32_t* __restrict__ qc, RankRec<EntT>* __restrict__ qrec)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= d) return;
    const uint32_t k1 = psa[r];
    if (k1 == 0) {
        qc[r] = 0;
        return;
    }
    const OccRec<EntT> o = occ[k1 - 1];
    qc[r] = 1u + o.mx_count;
    qrec[r] = RankRec<EntT>{o.whole, o.mx_begin};
}

// The primary-index entry (whole phrase of occurrence 0): its group and its index among the
// entries (the first entry of parse-suffix rank prank[1] - 1, or the last entry if d == 1).
__global__ void k_q_special(const uint32_t* __restrict__ prank, const uint32_t* __restrict__ pid, uint32_t d,
                            const PhraseRec* __restrict__ pr, const uint32_t* __restrict__ qo, uint32_t last,
                            uint32_t* __restrict__ special)
{
    special[0] = pr[pid[0]].alpha0;
    special[1] = (d >= 2) ? qo[prank[1] - 1] : last;
}

// Load-balanced generation of the entries: every block writes kQBlock consecutive entries.
// k_q_blocks records, per block, the parse-suffix rank whose entries contain the block's first
// output.  The block stages the offsets of the following ranks and, for the ranks whose
// entries reach into the block, their whole-phrase entry and first mixed entry (coalesced
// reads in rank order); each output o then finds its rank r by bisection over the offsets and takes
// j = o - qo[r]: j == 0 is the whole-phrase entry, j >= 1 the (j-1)-th mixed suffix of the
// phrase.  Writes are fully coalesced.
constexpr int kQBlock = 1024;

__global__ void k_q_blocks(const uint32_t* __restrict__ qo, const uint32_t* __restrict__ qc, uint32_t d,
                           uint32_t* __restrict__ rb)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= d || qc[r] == 0) return;
    const uint32_t a = qo[r], e = a + qc[r];  // outputs [a, e)
    for (uint32_t b = (a + kQBlock - 1) / kQBlock; (uint64_t)b * kQBlock < e; b++) rb[b] = r;
}

template <typename EntT>
__global__ __launch_bounds__(kBlock) void k_q_expand(const RankRec<EntT>* __restrict__ qrec, uint32_t d,
                                                    const EntT* __restrict__ mx,
                                                    const uint32_t* __restrict__ qo, const uint32_t* __restrict__ rb,
                                                    uint32_t total, EntT* __restrict__ ent,
                                                    const uint32_t* __restrict__ special, uint32_t* __restrict__ part)
{
    constexpr uint32_t kCap = kQBlock + 2;
    __shared__ uint32_t lq[kCap];
    __shared__ uint32_t lmx[kCap];
    __shared__ EntT lwe[kCap];
    __shared__ uint32_t red[kWaves];
    const uint32_t o0 = blockIdx.x * kQBlock;
    const uint32_t oend = (total - o0 < (uint32_t)kQBlock) ? total : o0 + kQBlock;
    const uint32_t r0 = rb[blockIdx.x];
    // every rank but one (the suffix starting at phrase 0) has at least one entry, so
    // kQBlock + 2 ranks cover the block's outputs
    const uint32_t cnt = (d - r0 < kCap) ? d - r0 : kCap;
    for (uint32_t t = threadIdx.x; t < cnt; t += kBlock) {
        const uint32_t q = qo[r0 + t];
        lq[t] = q;
        if (q < oend) {
            // ranks without entries carry no record; they are never chosen below
            const RankRec<EntT> rr = qrec[r0 + t];
            lmx[t] = rr.mx_begin;
            lwe[t] = rr.whole;
        }
    }
    __syncthreads();
    uint32_t g0 = 0, s0 = 0, before = 0;
    if (part) {
        g0 = special[0];
        s0 = special[1];
    }
    for (uint32_t t = threadIdx.x; t < (uint32_t)kQBlock; t += kBlock) {
        const uint32_t o = o0 + t;
        if (o >= total) break;
        // last local rank with qo <= o
        uint32_t a = 0, b = cnt - 1;
        while (a < b) {
            const uint32_t m = (a + b + 1) >> 1;
            if (lq[m] <= o) a = m;
            else b = m - 1;
        }
        // ranks without entries share their successor's offset: the bisection takes the last
        // of equal offsets, a rank with entries
        const uint32_t j = o - lq[a];
        const EntT v = (j == 0) ? lwe[a] : mx[lmx[a] + j - 1];
        ent[o] = v;
        if (part) {
            // entries that the stable sort places before the primary-index entry
            const uint32_t g = Ent<EntT>::group(v);
            before += (g < g0 || (g == g0 && o < s0)) ? 1u : 0u;
        }
    }
    if (part) {
        uint32_t tot;
        (void)block_exclusive_scan(before, 0u, OpSum{}, red, &tot);
        if (threadIdx.x == 0) part[blockIdx.x] = tot;
    }
}

// The last occurrence has no next parse suffix; its whole-phrase group is a singleton.  (With
// many texts the last phrase of every other text is reached from the first parse suffix of the
// next text; its group is a singleton as well, so its place among the entries does not matter.)
template <typename EntT>
__global__ void k_q_last(const uint32_t* __restrict__ pid, uint32_t d, const PhraseRec* __restrict__ pr,
                         const uint32_t* __restrict__ E, const uint8_t* __restrict__ pchar, uint32_t pos,
                         EntT* __restrict__ ent)
{
    const uint32_t k = d - 1;
    ent[pos] = whole_phrase_entry<EntT>(E, pchar, k, pr[pid[k]].alpha0);
}

// Primary index from 32-bit entries: 1 + delta(group) + number of entries sorted before the
// primary-index entry (block counts of k_q_expand, summed by k_q_count_sum into *acc, plus the
// last entry).
constexpr int kPidxPer = 16;
__global__ __launch_bounds__(kBlock) void k_q_count_sum(const uint32_t* __restrict__ part, uint32_t blocks,
                                                       uint32_t* __restrict__ acc)
{
    __shared__ uint32_t red[kWaves];
    const uint32_t b0 = (blockIdx.x * kBlock + threadIdx.x) * kPidxPer;
    uint32_t v = 0;
#pragma unroll
    for (int i = 0; i < kPidxPer; i++)
        if (b0 + i < blocks) v += part[b0 + i];
    uint32_t tot;
    (void)block_exclusive_scan(v, 0u, OpSum{}, red, &tot);
    if (threadIdx.x == 0 && tot) atomicAdd(acc, tot);
}

__global__ void k_q_pidx(const uint32_t* __restrict__ acc, const uint32_t* __restrict__ ent, uint32_t last,
                         const uint32_t* __restrict__ special, const uint32_t* __restrict__ delta,
                         uint32_t* __restrict__ pidx)
{
    const uint32_t g0 = special[0], s0 = special[1];
    const uint32_t g = ent[last] >> 8;
    const uint32_t tot = *acc + ((last != s0 && (g < g0 || (g == g0 && last < s0))) ? 1u : 0u);
    *pidx = delta[g0] + tot + 1;
}

// last radix pass of the entry sort: slot = sorted position + delta(alpha); a marked entry (the
// first phrase of a text) gives that text's primary index: 1 + its slot inside the text's BWT,
// the text found by bisection over the output offsets obase[0..texts]
struct EmitBwt {
    static constexpr double bytes_per_item = 1 + 4;  // BWT byte written, delta read
    static constexpr bool kValues = false;
    const uint32_t* delta;
    uint8_t* O;
    uint32_t* pidx;  // [texts]
    uint64_t n;
    const uint32_t* obase;  // [texts + 1]; nullptr for one text
    uint32_t texts;
    __device__ __forceinline__ void val(uint64_t, uint32_t) const {}
    __device__ __forceinline__ void key(uint64_t pos, uint64_t key) const
    {
        const uint64_t slot = pos + delta[(uint32_t)(key >> 9)];
        if (slot >= n) return;  // guard: a slot past the output is a bug caught by the checks
        O[slot] = (uint8_t)(key & 0xFF);
        if (key & 0x100) {
            if (!obase) {  // one text: its BWT starts at 0
                pidx[0] = (uint32_t)slot + 1;
                return;
            }
            uint32_t a = 0, b = texts - 1;  // last text with obase <= slot
            while (a < b) {
                const uint32_t m = (a + b + 1) >> 1;
                if (obase[m] <= slot) a = m;
                else b = m - 1;
            }
            pidx[a] = (uint32_t)(slot - obase[a]) + 1;
        }
    }
};

// the same for 32-bit entries (the primary index comes from k_q_pidx)
struct EmitBwt32 {
    static constexpr double bytes_per_item = 1 + 4;
    static constexpr bool kValues = false;
    const uint32_t* delta;
    uint8_t* O;
    uint64_t n;
    __device__ __forceinline__ void val(uint64_t, uint32_t) const {}
    __device__ __forceinline__ void key(uint64_t pos, uint32_t key) const
    {
        const uint64_t slot = pos + delta[key >> 8];
        if (slot >= n) return;  // guard: a slot past the output is a bug caught by the checks
        O[slot] = (uint8_t)(key & 0xFF);
    }
};

__global__ void k_u32_to_u64(const uint32_t* __restrict__ a, uint64_t* __restrict__ b, uint32_t m)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) b[i] = a[i];
    if (i == m) b[i] = 0;
}

// --------------------------------------------------------------------- driver

// Bits of the packed initial keys of the dictionary and parse sorts (MIBWT_INIT_KEY_BITS
// overrides both, MIBWT_PARSE_KEY_BITS the parse sort alone, for tuning runs).
static int init_key_bits(const Storage& S, bool parse = false)
{
    int v = S.tune.key_bits;
    const char* e = getenv("MIBWT_INIT_KEY_BITS");
    if (parse && getenv("MIBWT_PARSE_KEY_BITS")) e = getenv("MIBWT_PARSE_KEY_BITS");
    if (e) {
        const int x = atoi(e);
        if (x >= 8 && x <= 64) v = x;
    }
    return v;
}

// Mixed-radix initial keys: the most digits of base `base` that fit, with a leading field of
// `lead` values (1: none) above them, in keys below 2^cap.  Returns the digit count (at least
// 1); *key_bits receives the width of the largest key and *lead_pow base^digits (0 without a
// leading field).
static int mixed_radix_digits(uint64_t base, uint64_t lead, int cap, int* key_bits, uint64_t* lead_pow)
{
    const unsigned __int128 lim = (unsigned __int128)1 << (cap < 64 ? cap : 64);
    unsigned __int128 pw = 1;
    int per = 0;
    while (per < 64 && pw * base * lead <= lim) {
        pw *= base;
        per++;
    }
    if (per == 0) {
        per = 1;
        pw = base;
    }
    const unsigned __int128 top = pw * lead - 1;
    MIBWT_REQUIRE(top >> 64 == 0, LIBCUBWT_BAD_PARAMETER, "initial keys exceed 64 bits");
    *key_bits = std::max(1, bits_for((uint64_t)top));
    *lead_pow = lead > 1 ? (uint64_t)pw : 0;
    return per;
}

// Host description of the texts of one build: one text of n1 bytes at 0, or `count` texts laid
// out as TextSet describes (device arrays tbase, tlen, obase, sub_text).
struct BuildTexts {
    uint32_t count = 1;
    uint64_t staged = 0;  // staged length (a multiple of kTextAlign, >= every text end + kTextGap)
    uint64_t total = 0;   // sum of the text lengths = output length
    uint32_t n1 = 0;
    const uint32_t* tbase = nullptr;
    const uint32_t* tlen = nullptr;
    const uint32_t* obase = nullptr;
    const uint32_t* sub_text = nullptr;
    uint32_t* pidx = nullptr;  // [count] primary indices (device, count > 1)
};

inline uint64_t staged_length_one(uint64_t n) { return (n + kTextGap + kTextAlign - 1) / kTextAlign * kTextAlign; }

// flags[k] = 1 where phrase k ends a text (its end carries kSentinel)
__global__ void k_sent_flags(const uint32_t* __restrict__ E, uint32_t d, uint32_t* __restrict__ f)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < d) f[k] = E[k] >> 31;
}

// Phrase ends of the staged texts into B_TRIG: every trigger end and every text's sentinel end
// (flagged), in text order; returns their number d (the parse length).
uint32_t pfp_scan(Storage& S, const uint8_t* T, const BuildTexts& bt, int W, uint32_t P, uint32_t** E_out,
                  double* bytes)
{
    uint32_t* sc = S.buf<uint32_t>(B_SCALARS, SC_COUNT_);
    const uint64_t n = bt.staged;
    uint32_t t = 0;
    uint32_t* E = nullptr;
    double by = 0;
    if (W <= kScanMaxW) {
        // single pass; E is sized for the usual trigger density and the scan is repeated
        // with the exact size when a text has more triggers
        uint32_t cap = (uint32_t)std::min<uint64_t>(n, n / 8 + 65536);
        for (int pass = 0;; pass++) {
            E = S.buf<uint32_t>(B_TRIG, (size_t)cap + 1);
            const int kt = S.kt_begin(2, (double)n);
            trig_scan(S, T, n, bt.n1, bt.sub_text, bt.tbase, bt.tlen, W, P, E, cap, sc + SC_TRIGGERS);
            S.kt_end(kt);
            uint32_t hv[SC_COUNT_];
            S.read_fast(hv, sc, SC_COUNT_);
            t = hv[SC_TRIGGERS];
            if (kt >= 0) S.kt_bytes[kt] = (double)bt.total + 4.0 * std::min(t, cap);
            by += (double)bt.total + 4.0 * std::min(t, cap);
            if (t <= cap) break;
            MIBWT_REQUIRE(pass == 0, LIBCUBWT_GPU_ERROR, "trigger count changed between scans");
            cap = t;
        }
    } else {
        // windows wider than the exact-value range (one text only): two passes over per-thread
        // masks, then the sentinel end
        MIBWT_REQUIRE(bt.count == 1, LIBCUBWT_BAD_PARAMETER, "batched builds need W <= 56");
        const uint64_t n1 = bt.n1;
        const uint32_t tiles = ceil_div(n1 + 1, kTrigTile);
        uint16_t* masks = S.buf<uint16_t>(B_TRIG_MASK, (size_t)tiles * kBlock);
        uint32_t* tcnt = S.buf<uint32_t>(B_TILE_CNT, tiles + 1);
        k_trig_mask<<<tiles, kBlock, 0, S.stream>>>(T, n1, W, P, masks, tcnt);
        MIBWT_LAUNCH_CHECK();
        exclusive_sum<uint32_t>(S, tcnt, tcnt, tiles, sc + SC_TRIGGERS);
        uint32_t hv[SC_COUNT_];
        S.read_fast(hv, sc, SC_COUNT_);
        t = hv[SC_TRIGGERS];
        E = S.buf<uint32_t>(B_TRIG, (size_t)t + 1);
        k_trig_emit<<<tiles, kBlock, 0, S.stream>>>(masks, tcnt, E);
        MIBWT_LAUNCH_CHECK();
        fill<uint32_t>(S, E + t, 1, (uint32_t)(n1 + W - 1) | kSentinel);
        t += 1;
        by = (double)n1 + 2.0 * tiles * kBlock * 2 + 4.0 * t;
    }
    *E_out = E;
    *bytes = by;
    return t;
}

// Trigger end positions of a host text (the reference's Compute_D1/D2 output, PFPk:443-571):
// returns their number; E receives the first min(t, cap) of them.
int64_t pfp_triggers_run(Storage& S, const uint8_t* src, int64_t n64, int W, int P, uint32_t* Eh, int64_t cap)
{
    MIBWT_REQUIRE(n64 >= 1 && n64 < (int64_t)INT32_MAX - (1 << 20), LIBCUBWT_BAD_PARAMETER, "n out of range");
    MIBWT_REQUIRE(W >= 1 && W <= kMaxW && P >= 1 && P <= 65535, LIBCUBWT_BAD_PARAMETER, "bad W/P");
    const uint64_t n = (uint64_t)n64;
    S.phases_reset();
    uint8_t* T = stage_text(S, src, true, n);
    BuildTexts bt;
    bt.count = 1;
    bt.staged = staged_length_one(n);
    bt.total = n;
    bt.n1 = (uint32_t)n;
    uint32_t* E = nullptr;
    double bytes = 0;
    const uint32_t t = pfp_scan(S, T, bt, W, (uint32_t)P, &E, &bytes) - 1;  // without the sentinel end
    const int64_t m = std::min<int64_t>((int64_t)t, cap);
    if (m > 0) {
        MIBWT_HIP(hipMemcpyAsync(Eh, E, sizeof(uint32_t) * m, hipMemcpyDeviceToHost, S.stream));
        MIBWT_HIP(hipStreamSynchronize(S.stream));
    }
    S.stats.triggers = t;
    return (int64_t)t;
}

// rank[k] = lexicographic rank of the phrase of occurrence k
__global__ void k_occ_lexrank(const uint32_t* __restrict__ lexrank, const uint32_t* __restrict__ pid, uint32_t d,
                              uint32_t* __restrict__ out)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < d) out[k] = lexrank[pid[k]];
}

// The build of the staged texts T (B_TEXT) into dO: one text (bt.count == 1, returns its
// primary index) or a batch (returns 0; the primary indices land in pidx_out[count]).
int64_t pfp_build(Storage& S, uint8_t* T, const BuildTexts& bt, uint8_t* dO, int W, int P, PfpParseOut* parse_out,
                  int64_t* pidx_out)
{
    const uint64_t n = bt.total;
    const bool multi = bt.count > 1;
    const bool timed = !S.arena;  // arena mode records no events
    apply_tuning(S, bt.staged);
    uint32_t* sc = S.buf<uint32_t>(B_SCALARS, SC_COUNT_);
    // ---- 1. trigger scan
    int ph = S.phase_begin("scan");
    uint32_t* codes = S.buf<uint32_t>(B_ALPHA, 272);
    uint32_t sigma = 0;
    double scan_bytes = 0;
    uint32_t* E = nullptr;
    const uint32_t d = pfp_scan(S, T, bt, W, (uint32_t)P, &E, &scan_bytes);
    TextSet ts{bt.count, bt.tbase, bt.tlen, bt.obase, nullptr, bt.n1};
    if (multi) {
        // text of every phrase occurrence: the number of sentinel ends before it
        uint32_t* ptext = S.buf<uint32_t>(B_PTEXT, d);
        k_sent_flags<<<grid_for(d), kBlock, 0, S.stream>>>(E, d, ptext);
        MIBWT_LAUNCH_CHECK();
        exclusive_sum<uint32_t>(S, ptext, ptext, d);
        ts.ptext = ptext;
    }
    S.phase_end(ph, scan_bytes);
    S.stats.triggers = d - bt.count;
    S.stats.phrases = d;

    // ---- 2/3. phrase fingerprints and the dictionary
    ph = S.phase_begin("dict");
    uint64_t* K0 = S.buf<uint64_t>(B_SORT_K0, d);
    uint64_t* K1 = S.buf<uint64_t>(B_SORT_K1, d);
    uint32_t* V0 = S.buf<uint32_t>(B_SORT_V0, d);
    uint32_t* V1 = S.buf<uint32_t>(B_SORT_V1, d);
    uint32_t* flag = S.buf<uint32_t>(B_PH_FLAG, d);
    uint32_t du = 0;
    uint64_t seed = 0x6A09E667F3BCC909ull;
    int key_bits = kDictKeyBits;
    if (const char* kb = getenv("MIBWT_DICT_KEY_BITS")) {
        // narrower first sorts exercise the widening path (tests)
        const int v = atoi(kb);
        if (v >= 8 && v <= 64) key_bits = v & ~7;
    }
    int hash_bits = 64;
    if (const char* hb = getenv("MIBWT_DICT_HASH_BITS")) {
        // narrow first-attempt fingerprints collide and exercise the re-seeding path (tests)
        const int v = atoi(hb);
        if (v >= 1 && v <= 64) hash_bits = v;
    }
    int dict_passes = 0;
    uint32_t* pid = nullptr;
    uint32_t *head = nullptr, *dstart = nullptr, *dlen = nullptr, *dtl = nullptr, *dtext = nullptr, *freq = nullptr;
    uint64_t* doff = nullptr;
    uint64_t L64 = 0;
    uint8_t* pchar = S.buf<uint8_t>(B_PH_CHAR, d);
    for (int attempt = 0;; attempt++) {
        MIBWT_REQUIRE(attempt < 6, MIBWT_HASH_COLLISION, "phrase fingerprint collisions persist");
        const uint64_t hmask = (attempt == 0 && hash_bits < 64) ? ((1ull << hash_bits) - 1) : ~0ull;
        k_phrase_hash<<<grid_for(d), kBlock, 0, S.stream>>>(T, ts, W, E, d, seed, hmask, K0, V0, pchar);
        MIBWT_LAUNCH_CHECK();
        radix_sort<uint64_t, uint32_t>(S, K0, K1, V0, V1, d, 0, key_bits);
        dict_passes += key_bits / 8;
        fill<uint32_t>(S, sc + SC_COLLISION, 1, 0u);
        const uint64_t key_mask = (key_bits >= 64) ? ~0ull : ((1ull << key_bits) - 1);
        k_dict_mark<<<grid_for(d), kBlock, 0, S.stream>>>(K0, d, key_mask, flag, sc + SC_COLLISION);
        MIBWT_LAUNCH_CHECK();
        inclusive_sum<uint32_t>(S, flag, K1 ? reinterpret_cast<uint32_t*>(K1) : nullptr, d, sc + SC_DISTINCT);
        uint32_t hv[SC_COUNT_];
        S.read_fast(hv, sc, SC_COUNT_);
        du = hv[SC_DISTINCT];
        if (hv[SC_COLLISION] & 2u) {  // distinct fingerprints under one sort key: sort all 64 bits
            key_bits = 64;
            continue;
        }
        const uint32_t* ids = reinterpret_cast<const uint32_t*>(K1);
        pid = S.buf<uint32_t>(B_PH_PID, d);
        head = S.buf<uint32_t>(B_DU_HEAD, du);
        dstart = S.buf<uint32_t>(B_DU_START, du);
        dlen = S.buf<uint32_t>(B_DU_LEN, du);
        dtl = S.buf<uint32_t>(B_DU_TL, du);
        dtext = S.buf<uint32_t>(B_DU_TEXT, du);
        freq = S.buf<uint32_t>(B_DU_FREQ, du);
        k_dict_assign<<<grid_for(d), kBlock, 0, S.stream>>>(ids, V0, flag, E, ts, d, W, pid, head, dstart, dlen, dtl,
                                                             dtext);
        MIBWT_LAUNCH_CHECK();
        k_dict_verify<<<grid_for(d), kBlock, 0, S.stream>>>(T, ts, W, E, d, pid, dstart, dlen, dtl, dtext,
                                                             sc + SC_COLLISION);
        MIBWT_LAUNCH_CHECK();
        k_dict_freq<<<grid_for(du), kBlock, 0, S.stream>>>(head, du, d, freq);
        MIBWT_LAUNCH_CHECK();
        uint64_t* dlen64 = S.buf<uint64_t>(B_DU_LEN64, du + 1);
        doff = S.buf<uint64_t>(B_DU_OFF, du + 1);
        k_u32_to_u64<<<grid_for(du + 1), kBlock, 0, S.stream>>>(dlen, dlen64, du);
        MIBWT_LAUNCH_CHECK();
        exclusive_sum<uint64_t>(S, dlen64, doff, du + 1);
        // alphabet of the dictionary = alphabet of the text (the phrases cover it)
        alphabet_codes(S, T, bt.staged, codes, dstart, dtl, du);
        // one read-back: collision flag, dictionary length, alphabet size
        MIBWT_HIP(hipMemcpyAsync(sc + SC_DICTLEN, doff + du, sizeof(uint64_t), hipMemcpyDeviceToDevice, S.stream));
        MIBWT_HIP(hipMemcpyAsync(sc + SC_SIGMA, codes + 256, sizeof(uint32_t), hipMemcpyDeviceToDevice, S.stream));
        S.read_fast(hv, sc, SC_COUNT_);
        if (hv[SC_COLLISION] & 1u) {  // equal fingerprints of different phrases: new seed
            seed = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
            continue;
        }
        L64 = (uint64_t)hv[SC_DICTLEN] | ((uint64_t)hv[SC_DICTLEN + 1] << 32);
        sigma = hv[SC_SIGMA];
        break;
    }
    MIBWT_REQUIRE(L64 < (uint64_t)INT32_MAX - 2, MIBWT_INDEX_RANGE, "dictionary exceeds the 32-bit index range");
    const uint32_t L = (uint32_t)L64;
    S.phase_end(ph, (double)n + (double)W * d + 16.0 * d + 8.0 * d + 24.0 * d * dict_passes + 12.0 * d);
    S.stats.distinct_phrases = du;
    S.stats.dict_length = L;

    // ---- 4. dictionary suffix sort
    ph = S.phase_begin("dsort");
    uint32_t* owner = S.buf<uint32_t>(B_DT_OWNER, L);
    uint32_t* drank = S.buf<uint32_t>(B_DT_RANK, L);
    k_dt_owner<<<grid_for(du), kBlock, 0, S.stream>>>(doff, dlen, du, owner);
    MIBWT_LAUNCH_CHECK();
    K0 = S.buf<uint64_t>(B_SORT_K0, L);
    K1 = S.buf<uint64_t>(B_SORT_K1, L);
    V0 = S.buf<uint32_t>(B_SORT_V0, L);
    V1 = S.buf<uint32_t>(B_SORT_V1, L);
    // symbols per initial key: dense codes of the bytes that occur (the reference's small-
    // alphabet variant packs two symbols per radix digit, PFPs16:995-1002; here a 64-bit key
    // holds as many codes as fit, 21 for DNA)
    // initial key width: whole radix passes trade against one more doubling round; with many
    // texts the text id takes the digits above the symbols
    int dkey_bits = 0;
    uint64_t tpow = 0;
    int dcap = init_key_bits(S);
    // A dictionary of mostly distinct phrases (little repetition) is nearly sorted by its first
    // few symbols: enough of them for L distinct keys plus two, in whole radix digits, instead of
    // the tuned width (the 256 MiB order-2 Markov text: 64 -> 48 bits, dsort 61.6 -> 55.7 ms).
    if (!getenv("MIBWT_INIT_KEY_BITS") && 2 * (uint64_t)du > (uint64_t)d) {
        const double lg = std::log2((double)sigma + 2);
        const int need = (int)std::ceil((std::ceil(std::log2((double)L + 1) / lg) + 2) * lg);
        dcap = std::min(dcap, std::max(16, (need + 7) / 8 * 8));
    }
    // The records of the dictionary positions ride above the key bits through the initial sort
    // when the flags, the preceding character and at least 4 frequency bits fit there (keys of
    // mostly distinct phrases: the order-2 Markov text's 47-bit keys leave 7 frequency bits);
    // else the rest lengths when at least 5 bits are free.  A tuned width that leaves fewer
    // gives up 8 bits of key for them (DNA: 21 symbols in 55 bits instead of 24 in 63, 9 bits of
    // rest length; the 1 GiB collection at W=10,P=71: dictionary sort 10.9 -> 10.1 ms, at
    // W=5,P=16 4.70 -> 4.43 ms; the narrower key alone changed nothing).
    // MIBWT_DICT_PAYLOAD: 0 = never, k > 0 = at most k frequency bits, -k = the rest lengths in
    // at most k bits.
    int fcap = 24, rcap = 16;
    bool pay_on = true, rest_only = false;
    if (const char* e = getenv("MIBWT_DICT_PAYLOAD")) {
        const int v = atoi(e);
        pay_on = v != 0;
        if (v > 0) fcap = std::min(fcap, v);
        if (v < 0) {
            rest_only = true;
            rcap = std::min(rcap, -v);
        }
    }
    int aper = mixed_radix_digits((uint64_t)sigma + 2, multi ? bt.count : 1, dcap, &dkey_bits, &tpow);
    if (pay_on && 64 - dkey_bits < 5 && dcap > 24 && !getenv("MIBWT_INIT_KEY_BITS"))
        aper = mixed_radix_digits((uint64_t)sigma + 2, multi ? bt.count : 1, dcap - 8, &dkey_bits, &tpow);
    int pshift = 64, fbits = 0, rbits = 0;
    {
        const bool on = pay_on;
        const int room = 64 - dkey_bits;
        const int fb = std::min(room - 10, fcap);
        const int rb = std::min(room, rcap);
        if (on && !rest_only && fb >= std::min(4, fcap)) {
            pshift = dkey_bits;
            fbits = fb;
        } else if (on && rb >= std::min(5, rcap)) {
            pshift = dkey_bits;
            rbits = rb;
        }
    }
    DictPos* dp = reinterpret_cast<DictPos*>(S.buf<uint4>(B_DT_POS, L));
    k_dt_init<<<grid_for(L), kBlock, 0, S.stream>>>(T, owner, doff, dlen, dtl, dtext, dstart, freq, codes,
                                                     (uint32_t)sigma + 2, aper, tpow, L, W, pshift, fbits, rbits,
                                                     K0, V0, dp);
    MIBWT_LAUNCH_CHECK();
    // per dictionary position a record (frequency, preceding character, flags); the sort
    // places it with its group head in sorted order (RH) and marks the whole phrases (wbits)
    SortedRec* RH = reinterpret_cast<SortedRec*>(S.buf<uint4>(B_AS_R, L));
    const uint32_t wwords = (uint32_t)ceil_div((uint64_t)L, (uint64_t)32);
    uint32_t* wbits = S.buf<uint32_t>(B_AS_WBITS, wwords);
    fill<uint32_t>(S, wbits, wwords, 0u);
    LimDict limd{dp};
    if (pshift < 64) {
        limd.pay = K0;  // prefix_doubling points it at the buffer that holds the sorted keys
        limd.pshift = pshift;
        limd.fbits = fbits;
        limd.rbits = rbits;
    }
    DoublingResult rd = prefix_doubling<LimDict, PlaceRec>(S, K0, K1, V0, V1, L, dkey_bits, (uint32_t)aper, limd,
                                                           drank, nullptr,
                                                           PlaceRec{dp, RH, wbits});
    S.phase_end(ph, rd.bytes);
    S.stats.doubling_rounds_dict = rd.rounds;

    // ---- 5. phrase order and parse suffix array
    ph = S.phase_begin("psort");
    uint32_t* wscan = S.buf<uint32_t>(B_AS_WSCAN, wwords);
    k_popc_words<<<grid_for(wwords), kBlock, 0, S.stream>>>(wbits, wwords, wscan);
    MIBWT_LAUNCH_CHECK();
    exclusive_sum<uint32_t>(S, wscan, wscan, wwords);
    uint32_t* lexrank = S.buf<uint32_t>(B_LEXRANK, du);
    k_lexrank<<<grid_for(du), kBlock, 0, S.stream>>>(drank, doff, wbits, wscan, du, lexrank);
    MIBWT_LAUNCH_CHECK();
    K0 = S.buf<uint64_t>(B_SORT_K0, d);
    K1 = S.buf<uint64_t>(B_SORT_K1, d);
    V0 = S.buf<uint32_t>(B_SORT_V0, d);
    V1 = S.buf<uint32_t>(B_SORT_V1, d);
    int pkey_bits = 0;
    uint64_t unused_pow = 0;
    const int pper = mixed_radix_digits((uint64_t)du + 1, 1, init_key_bits(S, true), &pkey_bits, &unused_pow);
    k_parse_initkey<<<grid_for(d), kBlock, 0, S.stream>>>(lexrank, pid, d, (uint64_t)du + 1, pper, K0, V0);
    MIBWT_LAUNCH_CHECK();
    uint32_t* prank = S.buf<uint32_t>(B_P_RANK, d);
    uint32_t* psa = S.buf<uint32_t>(B_P_SA, d);
    DoublingResult rp = prefix_doubling<LimParse>(S, K0, K1, V0, V1, d, pkey_bits, (uint32_t)pper, LimParse{d},
                                                  prank, psa);
    S.phase_end(ph, rp.bytes);
    S.stats.doubling_rounds_parse = rp.rounds;
    if (parse_out) {
        // the intermediate results of the phrase order and the parse suffix sort, then stop
        const uint64_t m = std::min<uint64_t>(d, (uint64_t)parse_out->cap);
        uint32_t* occ = S.buf<uint32_t>(B_Q_CNT, d);
        k_occ_lexrank<<<grid_for(d), kBlock, 0, S.stream>>>(lexrank, pid, d, occ);
        MIBWT_LAUNCH_CHECK();
        if (m > 0) {
            MIBWT_HIP(hipMemcpyAsync(parse_out->rank, occ, sizeof(uint32_t) * m, hipMemcpyDeviceToHost, S.stream));
            MIBWT_HIP(hipMemcpyAsync(parse_out->psa, psa, sizeof(uint32_t) * m, hipMemcpyDeviceToHost, S.stream));
        }
        MIBWT_HIP(hipStreamSynchronize(S.stream));
        parse_out->distinct = du;
        return (int64_t)d;
    }

    // ---- 6. BWT assembly
    ph = S.phase_begin("assemble");
    uint8_t* gmix = S.buf<uint8_t>(B_G_MIX, L);
    fill<uint8_t>(S, gmix, L, (uint8_t)0);
    k_asm_mix<<<grid_for(L), kBlock, 0, S.stream>>>(RH, L, gmix);
    MIBWT_LAUNCH_CHECK();
    uint32_t* goff = S.buf<uint32_t>(B_G_OFF, L);
    uint32_t* ecnt = S.buf<uint32_t>(B_G_ECNT, L);
    uint16_t* fc = S.buf<uint16_t>(B_G_FILLC, L);
    uint8_t* gflag = S.buf<uint8_t>(B_E_GFLAG, L + 16);
    uint32_t* cid = S.buf<uint32_t>(B_E_CID, L);
    uint8_t* mxf = S.buf<uint8_t>(B_MX_FLAG, L + 16);
    uint32_t* mxs = S.buf<uint32_t>(B_MX_SCAN, L + 1);
    fill<uint8_t>(S, gflag, L, (uint8_t)0);
    fill<uint8_t>(S, mxf, L + 1, (uint8_t)0);
    k_asm_counts<<<grid_for(L), kBlock, 0, S.stream>>>(RH, gmix, L, goff, ecnt, fc, gflag, mxf);
    MIBWT_LAUNCH_CHECK();
    // dense ids of the entry groups
    sum_scan_of(S, ScanBytes{gflag}, true, cid, L, false, sc + SC_NGROUPS);
    exclusive_sum<uint32_t>(S, goff, goff, L, sc + SC_TOTAL);
    exclusive_sum<uint32_t>(S, ecnt, ecnt, L);
    uint32_t hv0[SC_COUNT_];
    S.read_fast(hv0, sc, SC_COUNT_);
    MIBWT_REQUIRE(hv0[SC_TOTAL] == (uint32_t)n, LIBCUBWT_GPU_ERROR, "BWT slot count mismatch");
    const uint32_t ngroups = hv0[SC_NGROUPS];
    const int gbits = std::max(1, bits_for(ngroups > 0 ? ngroups - 1 : 0));
    k_asm_fill<<<ceil_div((uint64_t)L, (uint64_t)kBlock * kFillPer), kBlock, 0, S.stream>>>(goff, fc, L, n, dO);
    MIBWT_LAUNCH_CHECK();
    uint32_t* deltac = S.buf<uint32_t>(B_E_DELTAC, std::max<uint32_t>(1, ngroups));
    k_ent_delta<<<grid_for(L), kBlock, 0, S.stream>>>(gflag, cid, goff, ecnt, L, deltac);
    MIBWT_LAUNCH_CHECK();
    sum_scan_of(S, ScanBytes{mxf}, true, mxs, L + 1, false, sc + SC_ENTRIES);
    uint32_t nmx = 0;
    S.read_fast(&nmx, sc + SC_ENTRIES);
    PhraseRec* prec = S.buf<PhraseRec>(B_AS_PREC, du);
    uint32_t* qc = S.buf<uint32_t>(B_Q_CNT, d);
    uint32_t* qo = S.buf<uint32_t>(B_Q_OFF, d);
    uint32_t ne = 0;
    // entries: (group, byte) in parse-suffix order, then one stable sort by group whose last
    // pass writes the BWT bytes; 32-bit entries when the groups fit in 24 bits
    auto assemble = [&](auto tag) {
        using EntT = decltype(tag);
        constexpr bool k32 = sizeof(EntT) == 4;
        EntT* mx = reinterpret_cast<EntT*>(S.buf<uint64_t>(B_MX_ENT, nmx + 1));
        k_mx_write<EntT><<<grid_for(L), kBlock, 0, S.stream>>>(dp, drank, mxf, mxs, cid, L, mx);
        MIBWT_LAUNCH_CHECK();
        k_phrase_rec<<<grid_for(du), kBlock, 0, S.stream>>>(doff, dlen, mxs, drank, cid, du, prec);
        MIBWT_LAUNCH_CHECK();
        OccRec<EntT>* occ = reinterpret_cast<OccRec<EntT>*>(S.buf<uint4>(B_Q_OCC, d));
        RankRec<EntT>* qrec = reinterpret_cast<RankRec<EntT>*>(S.buf<uint4>(B_Q_RREC, k32 ? (d + 1) / 2 : d));
        k_occ_rec<EntT><<<grid_for(d), kBlock, 0, S.stream>>>(pid, d, prec, pchar, E, occ);
        MIBWT_LAUNCH_CHECK();
        k_q_count<EntT><<<grid_for(d), kBlock, 0, S.stream>>>(psa, d, occ, qc, qrec);
        MIBWT_LAUNCH_CHECK();
        exclusive_sum<uint32_t>(S, qc, qo, d, sc + SC_ENTRIES);
        uint32_t hv[SC_COUNT_];
        S.read_fast(hv, sc, SC_COUNT_);
        ne = hv[SC_ENTRIES] + 1;
        // the histograms from the group counts are used only if they count every entry
        EntT* ent0 = reinterpret_cast<EntT*>(S.buf<uint64_t>(B_SORT_K0, ne));
        EntT* ent1 = reinterpret_cast<EntT*>(S.buf<uint64_t>(B_SORT_K1, ne));
        const uint32_t nq = ne - 1;  // entries of the parse suffixes; the last occurrence adds one
        const uint32_t qblocks = nq > 0 ? ceil_div(nq, (uint32_t)kQBlock) : 0;
        uint32_t* special = sc + SC_SPECIAL;
        uint32_t* part = nullptr;
        if (k32) {
            part = S.buf<uint32_t>(B_Q_PART, qblocks + 1);
            k_q_special<<<1, 1, 0, S.stream>>>(prank, pid, d, prec, qo, ne - 1, special);
            MIBWT_LAUNCH_CHECK();
        }
        if (nq > 0) {
            uint32_t* rb = S.buf<uint32_t>(B_Q_BLK, qblocks);
            k_q_blocks<<<grid_for(d), kBlock, 0, S.stream>>>(qo, qc, d, rb);
            MIBWT_LAUNCH_CHECK();
            k_q_expand<EntT><<<qblocks, kBlock, 0, S.stream>>>(qrec, d, mx, qo, rb, nq, ent0, special, part);
            MIBWT_LAUNCH_CHECK();
        }
        k_q_last<EntT><<<1, 1, 0, S.stream>>>(pid, d, prec, E, pchar, ne - 1, ent0);
        MIBWT_LAUNCH_CHECK();
        if constexpr (k32) {
            fill<uint32_t>(S, sc + SC_TOTAL, 1, 0u);
            if (qblocks > 0) {
                k_q_count_sum<<<ceil_div(qblocks, (uint32_t)(kBlock * kPidxPer)), kBlock, 0, S.stream>>>(part, qblocks,
                                                                                                     sc + SC_TOTAL);
                MIBWT_LAUNCH_CHECK();
            }
            k_q_pidx<<<1, 1, 0, S.stream>>>(sc + SC_TOTAL, ent0, ne - 1, special, deltac, sc + SC_PIDX);
            MIBWT_LAUNCH_CHECK();
            EmitBwt32 em{deltac, dO, n};
            radix_sort<uint32_t, uint32_t, EmitBwt32>(S, ent0, ent1, nullptr, nullptr, ne, 8, 8 + gbits, &em);
        } else {
            uint32_t* pidx_dev = multi ? bt.pidx : sc + SC_PIDX;
            fill<uint32_t>(S, pidx_dev, bt.count, 0u);
            EmitBwt em{deltac, dO, pidx_dev, n, multi ? bt.obase : nullptr, bt.count};
            radix_sort<uint64_t, uint32_t, EmitBwt>(S, ent0, ent1, nullptr, nullptr, ne, 9, 9 + gbits, &em);
        }
    };
    // MIBWT_WIDE_ENTRIES=1 takes the 64-bit entries (with the in-sort primary index) at any size
    // (batches always: the primary-index marks go through the sort)
    const bool wide = multi || gbits > 24 || (getenv("MIBWT_WIDE_ENTRIES") && atoi(getenv("MIBWT_WIDE_ENTRIES")) == 1);
    if (!wide) assemble(uint32_t{});
    else assemble(uint64_t{});
    const double esize = wide ? 8.0 : 4.0;
    S.phase_end(ph, 2.0 * n + 8.0 * L * 6 + esize * ne * 2 * ((gbits + 7) / 8) + 4.0 * d * 3);
    S.stats.alpha_groups = L;
    S.stats.mixed_entries = ne;

    if (timed) MIBWT_HIP(hipEventRecord(S.t_end, S.stream));
    uint32_t pidx = 0;
    if (multi) {
        std::vector<uint32_t> h(bt.count);
        MIBWT_HIP(hipMemcpyAsync(h.data(), bt.pidx, sizeof(uint32_t) * bt.count, hipMemcpyDeviceToHost, S.stream));
        MIBWT_HIP(hipStreamSynchronize(S.stream));
        for (uint32_t t = 0; t < bt.count; t++) pidx_out[t] = h[t];
    } else {
        S.read(&pidx, sc + SC_PIDX);
    }
    S.phases_collect();
    float tot = 0;
    if (timed) MIBWT_HIP(hipEventElapsedTime(&tot, S.t_begin, S.t_end));
    S.stats.total_ms = tot;
    S.last_W = W;
    S.last_P = P;
    S.last_sort_items = d;
    if (multi) return 0;
    MIBWT_REQUIRE(pidx >= 1 && pidx <= n, LIBCUBWT_GPU_ERROR, "primary index not produced");
    return (int64_t)pidx;
}

int64_t pfp_bwt_run(Storage& S, const uint8_t* src, bool src_host, uint8_t* dO, int64_t n64, int W, int P,
                    PfpParseOut* parse_out)
{
    MIBWT_REQUIRE(n64 >= 1 && n64 < (int64_t)INT32_MAX - (1 << 20), LIBCUBWT_BAD_PARAMETER, "n out of range");
    MIBWT_REQUIRE(W >= 1 && W <= kMaxW && P >= 1 && P <= 65535, LIBCUBWT_BAD_PARAMETER, "bad W/P");
    const uint64_t n = (uint64_t)n64;
    S.phases_reset();
    S.stats.n = n64;
    S.stats.W = W;
    S.stats.P = P;
    const bool timed = !S.arena;  // arena mode creates and records no events
    if (timed && !S.t_begin) {
        MIBWT_HIP(hipEventCreate(&S.t_begin));
        MIBWT_HIP(hipEventCreate(&S.t_end));
    }
    auto mark_begin = [&]() {
        if (timed) MIBWT_HIP(hipEventRecord(S.t_begin, S.stream));
    };
    // the text: a padded private copy (src == nullptr: already staged by the caller); a copy
    // from device memory is part of the timed build, an upload from the host is not
    uint8_t* T;
    if (src == nullptr) {
        T = S.buf<uint8_t>(B_TEXT, staged_text_bytes(n));
        mark_begin();
    } else if (src_host) {
        T = stage_text(S, src, true, n);
        mark_begin();
    } else {
        mark_begin();
        T = stage_text(S, src, false, n);
    }
    BuildTexts bt;
    bt.count = 1;
    bt.staged = staged_length_one(n);
    bt.total = n;
    bt.n1 = (uint32_t)n;
    const int64_t idx = pfp_build(S, T, bt, dO, W, P, parse_out, nullptr);
    S.last_text = src_host ? nullptr : src;
    S.last_n = n64;
    return idx;
}

// Staged layout of a batch: text t at base[t] (see TextSet), sub_text[] = text of every
// kTextAlign-byte sub-tile.
__global__ void k_stage_texts(const uint8_t* __restrict__ src, const uint32_t* __restrict__ sub_text,
                              const uint32_t* __restrict__ tbase, const uint32_t* __restrict__ tlen,
                              const uint32_t* __restrict__ obase, uint64_t staged, uint8_t* __restrict__ T)
{
    const uint64_t q = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 16;  // 16 staged bytes
    if (q >= staged) return;
    const uint32_t t = sub_text[q / kTextAlign];
    const uint64_t b = tbase[t], len = tlen[t], so = obase[t];
    uint32_t w[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        uint32_t v = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint64_t l = q + 4 * j + i - b;  // q >= b: a sub-tile never starts before its text
            if (l < len) v |= (uint32_t)src[so + l] << (8 * i);
        }
        w[j] = v;
    }
    *reinterpret_cast<uint4*>(T + q) = make_uint4(w[0], w[1], w[2], w[3]);
}

// PFP BWT of k device texts at dT + offs[i] (offs[k] = end of the last) into dO + offs[i]
// (reference format), in builds of consecutive texts whose staged layout stays below
// `chunk_bytes`; index[i] receives text i's primary index.  W = P = 0: the selector chooses per
// build on the staged texts.  A build whose dictionary outgrows the index range or the memory
// is split in two.
void pfp_batch_range(Storage& S, const uint8_t* dT, const int64_t* offs, int64_t a, int64_t b, uint8_t* dO, int W,
                     int P, int64_t* index);
void selector_choose_device(Storage& S, const uint8_t* T, uint64_t n, int* W, int* P, float* pred, uint8_t* copy);

int64_t pfp_batch_run(Storage& S, const uint8_t* dT, const int64_t* offs, int64_t k, uint8_t* dO, int W, int P,
                      int64_t* index, uint64_t chunk_bytes)
{
    int64_t a = 0;
    while (a < k) {
        // consecutive texts up to the chunk budget (at least one)
        uint64_t staged = 0;
        int64_t b = a;
        while (b < k) {
            const uint64_t nt = (uint64_t)(offs[b + 1] - offs[b]);
            const uint64_t next = (staged + nt + kTextGap + kTextAlign - 1) / kTextAlign * kTextAlign;
            if (b > a && next > chunk_bytes) break;
            staged = next;
            b++;
        }
        pfp_batch_range(S, dT, offs, a, b, dO, W, P, index);
        a = b;
    }
    return LIBCUBWT_NO_ERROR;
}

void pfp_batch_range(Storage& S, const uint8_t* dT, const int64_t* offs, int64_t a, int64_t b, uint8_t* dO, int W,
                     int P, int64_t* index)
{
    const int64_t k = b - a;
    if (k == 1) {
        // one text: the single-text build (no batch layout)
        int w = W, p = P;
        const uint64_t n = (uint64_t)(offs[a + 1] - offs[a]);
        if (W == 0) {
            const uint64_t bytes = staged_text_bytes(n);
            uint8_t* T = S.buf<uint8_t>(B_TEXT, bytes);
            MIBWT_HIP(hipMemsetAsync(T + n, 0, bytes - n, S.stream));
            selector_choose_device(S, dT + offs[a], n, &w, &p, nullptr, T);
            index[a] = pfp_bwt_run(S, nullptr, false, dO + offs[a], (int64_t)n, w, p, nullptr);
        } else {
            index[a] = pfp_bwt_run(S, dT + offs[a], false, dO + offs[a], (int64_t)n, w, p, nullptr);
        }
        return;
    }
    // host tables of the layout
    std::vector<uint32_t> tb(k), tn(k), ob(k + 1);
    uint64_t staged = 0;
    for (int64_t t = 0; t < k; t++) {
        const uint64_t nt = (uint64_t)(offs[a + t + 1] - offs[a + t]);
        tb[t] = (uint32_t)staged;
        tn[t] = (uint32_t)nt;
        ob[t] = (uint32_t)(offs[a + t] - offs[a]);
        staged = (staged + nt + kTextGap + kTextAlign - 1) / kTextAlign * kTextAlign;
    }
    ob[k] = (uint32_t)(offs[b] - offs[a]);
    MIBWT_REQUIRE(staged < (uint64_t)INT32_MAX - (1u << 20), LIBCUBWT_BAD_PARAMETER, "batch build too large");
    const uint64_t subs = staged / kTextAlign;
    std::vector<uint32_t> st(subs);
    for (int64_t t = 0; t < k; t++) {
        const uint64_t s0 = tb[t] / kTextAlign;
        const uint64_t s1 = (t + 1 < k) ? tb[t + 1] / kTextAlign : subs;
        for (uint64_t q = s0; q < s1; q++) st[q] = (uint32_t)t;
    }
    uint32_t* dtab = S.buf<uint32_t>(B_BT_TABLES, 3 * (size_t)k + 1 + subs);
    uint32_t* d_tbase = dtab;
    uint32_t* d_tlen = dtab + k;
    uint32_t* d_obase = dtab + 2 * k;
    uint32_t* d_sub = dtab + 3 * k + 1;
    MIBWT_HIP(hipMemcpyAsync(d_tbase, tb.data(), 4 * k, hipMemcpyHostToDevice, S.stream));
    MIBWT_HIP(hipMemcpyAsync(d_tlen, tn.data(), 4 * k, hipMemcpyHostToDevice, S.stream));
    MIBWT_HIP(hipMemcpyAsync(d_obase, ob.data(), 4 * (k + 1), hipMemcpyHostToDevice, S.stream));
    MIBWT_HIP(hipMemcpyAsync(d_sub, st.data(), 4 * subs, hipMemcpyHostToDevice, S.stream));
    uint32_t* d_pidx = S.buf<uint32_t>(B_BT_PIDX, (size_t)k);
    S.phases_reset();
    S.stats.n = (int64_t)ob[k];
    const bool timed = !S.arena;
    if (timed && !S.t_begin) {
        MIBWT_HIP(hipEventCreate(&S.t_begin));
        MIBWT_HIP(hipEventCreate(&S.t_end));
    }
    if (timed) MIBWT_HIP(hipEventRecord(S.t_begin, S.stream));
    uint8_t* T = S.buf<uint8_t>(B_TEXT, staged + 8192);
    MIBWT_HIP(hipMemsetAsync(T + staged, 0, 8192, S.stream));
    k_stage_texts<<<grid_for(ceil_div(staged, 16)), kBlock, 0, S.stream>>>(dT + offs[a], d_sub, d_tbase, d_tlen,
                                                                           d_obase, staged, T);
    MIBWT_LAUNCH_CHECK();
    int w = W, p = P;
    if (W == 0) selector_choose_device(S, T, staged, &w, &p, nullptr, nullptr);
    S.stats.W = w;
    S.stats.P = p;
    BuildTexts bt;
    bt.count = (uint32_t)k;
    bt.staged = staged;
    bt.total = ob[k];
    bt.tbase = d_tbase;
    bt.tlen = d_tlen;
    bt.obase = d_obase;
    bt.sub_text = d_sub;
    bt.pidx = d_pidx;
    try {
        pfp_build(S, T, bt, dO + offs[a], w, p, nullptr, index + a);
    } catch (const Error& e) {
        // a dictionary beyond the 32-bit index range or the memory: two smaller builds
        if (e.code != MIBWT_INDEX_RANGE && e.code != LIBCUBWT_GPU_NOT_ENOUGH_MEMORY) throw;
        (void)hipGetLastError();
        MIBWT_HIP(hipStreamSynchronize(S.stream));
        const int64_t m = a + k / 2;
        pfp_batch_range(S, dT, offs, a, m, dO, W, P, index);
        pfp_batch_range(S, dT, offs, m, b, dO, W, P, index);
    }
}

__global__ __launch_bounds__(kBlock) void k_copy16(const uint4* __restrict__ a, uint4* __restrict__ b, uint64_t m)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < m) b[i] = a[i];
}

// Plain device copy of `bytes` bytes, `reps` times: the reference point for hardware byte
// counters (it reads and writes exactly `bytes` each).
double time_copy(Storage& S, uint64_t bytes, int reps)
{
    const uint64_t m = bytes / 16;
    uint4* a = S.buf<uint4>(B_SORT_K0, m);
    uint4* b = S.buf<uint4>(B_SORT_K1, m);
    hipEvent_t e0, e1;
    MIBWT_HIP(hipEventCreate(&e0));
    MIBWT_HIP(hipEventCreate(&e1));
    MIBWT_HIP(hipEventRecord(e0, S.stream));
    for (int i = 0; i < reps; i++) k_copy16<<<ceil_div(m, kBlock), kBlock, 0, S.stream>>>(a, b, m);
    MIBWT_LAUNCH_CHECK();
    MIBWT_HIP(hipEventRecord(e1, S.stream));
    MIBWT_HIP(hipEventSynchronize(e1));
    float ms = 0;
    MIBWT_HIP(hipEventElapsedTime(&ms, e0, e1));
    MIBWT_HIP(hipEventDestroy(e0));
    MIBWT_HIP(hipEventDestroy(e1));
    return ms / reps;
}

__global__ void k_rand_pairs(uint64_t* __restrict__ K, uint32_t* __restrict__ V, uint64_t m, uint64_t seed)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    K[i] = mix64(seed + 0x9E3779B97F4A7C15ull * (i + 1));
    V[i] = (uint32_t)i;
}

// out-of-order neighbours (keys) and values that did not travel with their key
__global__ void k_sorted_check(const uint64_t* __restrict__ K, const uint32_t* __restrict__ V, uint64_t m,
                               uint64_t seed, uint32_t* __restrict__ bad)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const bool b = (i + 1 < m && K[i] > K[i + 1]) || K[i] != mix64(seed + 0x9E3779B97F4A7C15ull * ((uint64_t)V[i] + 1));
    if (b) atomicAdd(bad, 1u);
}

// LSD radix sort of m uniformly random (64-bit key, 32-bit value) pairs, `reps` times: mean
// duration of one digit pass (8 per sort, all of its kernels) and its algorithmic bytes (24 per
// pair).  The result of every sort is checked (sorted keys, values with their keys).
double time_radix(Storage& S, uint64_t m, int reps, double* bytes)
{
    MIBWT_REQUIRE(m >= 1 && m < (1ull << 31), LIBCUBWT_BAD_PARAMETER, "radix timing: item count out of range");
    // the launch shape of the storage's last build, or on a fresh storage that of a build of m bytes
    if (S.last_n == 0) apply_tuning(S, m);
    uint64_t* K0 = S.buf<uint64_t>(B_SORT_K0, m);
    uint64_t* K1 = S.buf<uint64_t>(B_SORT_K1, m);
    uint32_t* V0 = S.buf<uint32_t>(B_SORT_V0, m);
    uint32_t* V1 = S.buf<uint32_t>(B_SORT_V1, m);
    uint32_t* sc = S.buf<uint32_t>(B_SCALARS, SC_COUNT_);
    hipEvent_t e0, e1;
    MIBWT_HIP(hipEventCreate(&e0));
    MIBWT_HIP(hipEventCreate(&e1));
    double total = 0;
    uint32_t bad = 0;
    for (int r = 0; r <= reps; r++) {  // the first sort warms the buffers and is not timed
        const uint64_t seed = 0x5EED0000ull + (uint64_t)r;
        k_rand_pairs<<<grid_for(m), kBlock, 0, S.stream>>>(K0, V0, m, seed);
        MIBWT_LAUNCH_CHECK();
        MIBWT_HIP(hipEventRecord(e0, S.stream));
        radix_sort<uint64_t, uint32_t>(S, K0, K1, V0, V1, m, 0, 64);
        MIBWT_HIP(hipEventRecord(e1, S.stream));
        fill<uint32_t>(S, sc + SC_COLLISION, 1, 0u);
        k_sorted_check<<<grid_for(m), kBlock, 0, S.stream>>>(K0, V0, m, seed, sc + SC_COLLISION);
        MIBWT_LAUNCH_CHECK();
        MIBWT_HIP(hipEventSynchronize(e1));
        float ms = 0;
        MIBWT_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (r > 0) total += ms;
        uint32_t b = 0;
        S.read(&b, sc + SC_COLLISION);
        bad += b;
    }
    MIBWT_HIP(hipEventDestroy(e0));
    MIBWT_HIP(hipEventDestroy(e1));
    MIBWT_REQUIRE(bad == 0, LIBCUBWT_GPU_ERROR, "radix timing: the sort's output is not sorted");
    *bytes = 24.0 * (double)m;
    return total / reps / 8.0;
}

// Re-run the trigger scan kernel `reps` times on the text of the last call.
double pfp_time_scan(Storage& S, int reps, double* bytes)
{
    const uint64_t n = (uint64_t)S.last_n;
    MIBWT_REQUIRE(S.last_W >= 1 && S.last_W <= kScanMaxW, LIBCUBWT_BAD_PARAMETER, "no single-pass scan for this W");
    uint8_t* T = S.buf<uint8_t>(B_TEXT, staged_text_bytes(n));
    uint32_t* sc = S.buf<uint32_t>(B_SCALARS, SC_COUNT_);
    const uint32_t cap = (uint32_t)S.stats.triggers + 1;  // the triggers and the text's sentinel end
    uint32_t* E = S.buf<uint32_t>(B_TRIG, cap);
    hipEvent_t a, b;
    MIBWT_HIP(hipEventCreate(&a));
    MIBWT_HIP(hipEventCreate(&b));
    const uint64_t staged = staged_length_one(n);
    trig_scan(S, T, staged, (uint32_t)n, nullptr, nullptr, nullptr, S.last_W, (uint32_t)S.last_P, E, cap,
              sc + SC_TRIGGERS);
    MIBWT_HIP(hipEventRecord(a, S.stream));
    for (int i = 0; i < reps; i++)
        trig_scan(S, T, staged, (uint32_t)n, nullptr, nullptr, nullptr, S.last_W, (uint32_t)S.last_P, E, cap,
                  sc + SC_TRIGGERS);
    MIBWT_HIP(hipEventRecord(b, S.stream));
    MIBWT_HIP(hipEventSynchronize(b));
    float ms = 0;
    MIBWT_HIP(hipEventElapsedTime(&ms, a, b));
    MIBWT_HIP(hipEventDestroy(a));
    MIBWT_HIP(hipEventDestroy(b));
    uint32_t t = 0;
    S.read(&t, sc + SC_TRIGGERS);
    MIBWT_REQUIRE((int64_t)t == S.stats.triggers + 1, LIBCUBWT_GPU_ERROR, "trigger count differs between runs");
    *bytes = (double)n + 4.0 * t;
    return ms / reps;
}

}  // namespace mibwt
