// Learned (W,P) selector for prefix-free parsing.
//
// The reference picks (W,P) per input by hand (Benchmark_Skript_Linux.sh:14-73, default
// W=5,P=16 at :70-71) and compiles one binary per pair (Prefix_Free_Parsing_W=5_P=16_kommentiert.cu:19-20).
// Here a small MLP (64 -> 128 -> 128 -> 3, bf16 weights, fp32 accumulation) predicts, for every
// candidate of the grid W in {5,10,20,35} x P in kGridP, the dictionary length, the parse
// length and the build time per byte (natural logarithms) from cheap statistics of the text;
// the candidate with the smallest predicted time is chosen.  The BWT does not depend on (W,P),
// so the output stays bit-exact whatever the choice.
//
// Features (all computed on the GPU in one pass over a 1 MiB sample plus one content-defined
// sketch of the whole text):
//   text:      log n, log sigma, order-0 and order-1 entropy of the sample, distinct fractions
//              of content-sampled 16- and 64-mers (repetitiveness);
//   candidate: one-hot W, P/100, log P, 1/P, trigger density of (W,P) on the sample, the
//              implied mean phrase length and parse length.
// Training uses hand-written forward / backward / Adam kernels on runs of the pipeline itself.
#include "text.hpp"

#include <dlfcn.h>

#include <cmath>
#include <mutex>
#include <string>

namespace mibwt {

constexpr int kGridW[4] = {5, 10, 20, 35};
constexpr int kGridP[10] = {11, 13, 17, 23, 31, 41, 53, 61, 71, 81};
constexpr int kNumW = 4, kNumP = 10, kNumGrid = kNumW * kNumP;
// candidates: the grid, then the reference default W=5,P=16 (Benchmark_Skript_Linux.sh:70-71)
constexpr int kNumCand = kNumGrid + 1;
constexpr int kDefaultCand = kNumGrid;
inline int cand_W(int c) { return c < kNumGrid ? kGridW[c / kNumP] : 5; }
inline int cand_P(int c) { return c < kNumGrid ? kGridP[c % kNumP] : 16; }
constexpr int kIn = 64, kHid = 128, kOut = 3;
constexpr int kSampleChunks = 64, kChunk = 16384;  // 1 MiB sample

// ------------------------------------------------------------------------ bf16

__device__ __host__ inline uint16_t f32_to_bf16(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    const uint32_t r = 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)((u + r) >> 16);
}
__device__ __host__ inline float bf16_to_f32(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// ------------------------------------------------------------------------ features

// Divisibility by the odd grid moduli without division: v % P == 0 <=> v * P^-1 <= lim
// (mod 2^32 for W <= 20, mod 2^64 for W = 35).
struct GridDiv {
    uint32_t inv32[kNumP], lim32[kNumP];
    uint64_t inv64[kNumP], lim64[kNumP];
};

// Per-chunk partial counts of the sample: trigger counts of the 41 candidates, sampled
// positions, and 1 + the last text byte when the chunk holds position n-1 (else 0).
constexpr int kSpTrig = 0, kSpPos = kNumCand, kSpLast = kNumCand + 1, kSpWords = kNumCand + 2;
constexpr int kSampleBlock = 1024;                   // threads per chunk
constexpr int kSamplePer = kChunk / kSampleBlock;    // 16 consecutive positions per thread
constexpr int kSampleHalo = 48;                      // bytes loaded before a thread's positions

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Sample statistics over up to kSampleChunks chunks of kChunk bytes at a stride that is a
// multiple of 16 (one block per chunk): the byte-pair histogram (LDS, packed 16-bit counters,
// then added to hist2) and the trigger count of every candidate from rolled exact window values.
// Bytes outside [0, n) read as 0; windows ending before position W-1 are not counted.
__global__ __launch_bounds__(kSampleBlock) void k_feat_sample(const uint8_t* __restrict__ T, uint64_t n,
                                                             uint64_t stride, GridDiv gd,
                                                             uint32_t* __restrict__ hist2,
                                                             uint32_t* __restrict__ part)
{
    __shared__ uint32_t h2[65536 / 2];
    __shared__ uint32_t tc[(kNumCand + 1) / 2 + 2];
    {
        uint4* z = reinterpret_cast<uint4*>(h2);
        for (int i = threadIdx.x; i < 65536 / 8; i += kSampleBlock) z[i] = make_uint4(0, 0, 0, 0);
        if (threadIdx.x < (kNumCand + 1) / 2 + 2) tc[threadIdx.x] = 0;
    }
    __syncthreads();
    const uint64_t p0 = (uint64_t)blockIdx.x * stride + (uint64_t)threadIdx.x * kSamplePer;
    // bytes T[p0-48 .. p0+15] in w[0..15], T[p0+16] in nx (0 outside the text)
    uint32_t w[16];
    uint32_t nx = 0;
    if (p0 >= kSampleHalo && p0 + kSamplePer + 1 <= n) {
        const uint4* v = reinterpret_cast<const uint4*>(T + p0 - kSampleHalo);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint4 x = v[i];
            w[4 * i] = x.x;
            w[4 * i + 1] = x.y;
            w[4 * i + 2] = x.z;
            w[4 * i + 3] = x.w;
        }
        nx = T[p0 + kSamplePer];
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            uint32_t x = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const uint64_t p = p0 - kSampleHalo + 4 * i + b;  // wraps for p0 < 48: then >= n
                if (p0 + 4 * i + b >= kSampleHalo && p < n) x |= (uint32_t)T[p] << (8 * b);
            }
            w[i] = x;
        }
        if (p0 + kSamplePer < n) nx = T[p0 + kSamplePer];
    }
    auto byte_at = [&](int i) -> uint32_t { return (w[i >> 2] >> (8 * (i & 3))) & 0xFFu; };
    uint32_t cnt[kNumCand];
#pragma unroll
    for (int c = 0; c < kNumCand; c++) cnt[c] = 0;
    uint32_t v5 = 0, v10 = 0, v20 = 0;
    uint64_t v35 = 0;
#pragma unroll
    for (int k = 35; k >= 1; k--) {  // windows ending at p0 - 1
        const uint32_t c = byte_at(kSampleHalo - k);
        if (k <= 5) v5 = (v5 << 1) + c;
        if (k <= 10) v10 = (v10 << 1) + c;
        if (k <= 20) v20 = (v20 << 1) + c;
        v35 = (v35 << 1) + c;
    }
    uint32_t pos = 0, last = 0;
#pragma unroll
    for (int q = 0; q < kSamplePer; q++) {
        const uint64_t p = p0 + q;
        if (p < n) {
            const uint32_t c = byte_at(kSampleHalo + q);
            v5 = (v5 << 1) + c - (byte_at(kSampleHalo + q - 5) << 5);
            v10 = (v10 << 1) + c - (byte_at(kSampleHalo + q - 10) << 10);
            v20 = (v20 << 1) + c - (byte_at(kSampleHalo + q - 20) << 20);
            v35 = (v35 << 1) + c - ((uint64_t)byte_at(kSampleHalo + q - 35) << 35);
            if (p + 1 < n) {
                const uint32_t pr = (c << 8) | (q + 1 < kSamplePer ? byte_at(kSampleHalo + q + 1) : nx);
                atomicAdd(&h2[pr >> 1], 1u << (16 * (pr & 1u)));
            } else {
                last = c + 1;
            }
            pos++;
#pragma unroll
            for (int pi = 0; pi < kNumP; pi++) {
                if (p >= 5) cnt[0 * kNumP + pi] += (v5 * gd.inv32[pi] <= gd.lim32[pi]) ? 1u : 0u;
                if (p >= 10) cnt[1 * kNumP + pi] += (v10 * gd.inv32[pi] <= gd.lim32[pi]) ? 1u : 0u;
                if (p >= 20) cnt[2 * kNumP + pi] += (v20 * gd.inv32[pi] <= gd.lim32[pi]) ? 1u : 0u;
                if (p >= 35) cnt[3 * kNumP + pi] += (v35 * gd.inv64[pi] <= gd.lim64[pi]) ? 1u : 0u;
            }
            if (p >= 5) cnt[kDefaultCand] += ((v5 & 15u) == 0) ? 1u : 0u;
        }
    }
    // counts are at most 16 per thread: two per word in 16-bit halves, summed over the wave
    // (at most 1024), then over the block's waves (at most 16384)
#pragma unroll
    for (int c = 0; c < kNumCand + 1; c += 2) {
        const uint32_t lo = (c < kNumCand) ? cnt[c] : pos;
        const uint32_t hi = (c + 1 < kNumCand) ? cnt[c + 1] : (c + 1 == kNumCand ? pos : 0u);
        const uint32_t s = wave_sum(lo | (hi << 16));
        if ((threadIdx.x & 63) == 0 && s) atomicAdd(&tc[c / 2], s);
    }
    if (last) tc[(kNumCand + 1) / 2] = last;  // one thread of the grid at most
    __syncthreads();
    if (threadIdx.x < kNumCand + 1) {
        const uint32_t word = tc[threadIdx.x / 2];
        part[(uint64_t)blockIdx.x * kSpWords + threadIdx.x] = (threadIdx.x & 1) ? (word >> 16) : (word & 0xFFFFu);
    }
    if (threadIdx.x == 0) part[(uint64_t)blockIdx.x * kSpWords + kSpLast] = tc[(kNumCand + 1) / 2];
    for (int i = threadIdx.x; i < 65536 / 2; i += kSampleBlock) {
        const uint32_t x = h2[i];
        if (x & 0xFFFFu) atomicAdd(&hist2[2 * i], x & 0xFFFFu);
        if (x >> 16) atomicAdd(&hist2[2 * i + 1], x >> 16);
    }
}

constexpr int kSketchPer = 64;                              // positions per thread
constexpr uint64_t kSketchBlock = (uint64_t)kBlock * kSketchPer;
constexpr int kSketchQueue = 512;                           // selected positions per block in LDS

// insert a 32-bit fingerprint into an epoch-tagged open-addressing table (load factor at
// most 1/4); true if new.  The probe sequence is bounded: a fingerprint that finds no slot
// within kMaxProbe counts as new.
constexpr int kMaxProbe = 256;
__device__ __forceinline__ bool sketch_insert(uint64_t* __restrict__ tab, uint32_t mask, uint32_t epoch, uint32_t fp)
{
    const uint64_t mine = ((uint64_t)epoch << 32) | fp;
    uint32_t h = (uint32_t)(mix64(fp) & mask);
    for (int probe = 0; probe < kMaxProbe; probe++) {
        uint64_t cur = __hip_atomic_load(tab + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (;;) {
            if ((uint32_t)(cur >> 32) != epoch) {
                const uint64_t seen = atomicCAS((unsigned long long*)(tab + h), (unsigned long long)cur,
                                                (unsigned long long)mine);
                if (seen == cur) return true;
                cur = seen;
                continue;
            }
            if ((uint32_t)cur == fp) return false;
            break;
        }
        h = (h + 1) & mask;
    }
    return true;
}

// selection hash of the 8 bytes at a position (lo = bytes 0..3, hi = bytes 4..7, little endian):
// the position is selected when the top 10 bits are zero (about 1 in 1024 windows)
__device__ __forceinline__ bool sketch_selected(uint32_t lo, uint32_t hi)
{
    return ((lo ^ (hi * 0x85EBCA6Bu)) * 0x9E3779B1u) >> 22 == 0;
}

// 32-bit fingerprints of the 16 and 64 bytes at p (p + 64 <= n)
__device__ __forceinline__ void sketch_fingerprints(const uint8_t* __restrict__ T, uint64_t p, uint32_t& f16,
                                                    uint32_t& f64)
{
    uint64_t h = 0x9E3779B97F4A7C15ull;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        h = mix64(h ^ load8(T, p + 8 * k));
        if (k == 1) f16 = (uint32_t)(h >> 32);
    }
    f64 = (uint32_t)(h >> 32);
}

// Content-defined sketch of the whole text: the selected positions p (p + 64 <= n) contribute
// 32-bit fingerprints of the 16 and 64 bytes that start there; the distinct ones are counted
// with two hash sets.  A block first collects its selected positions in LDS, then inserts them
// with one lane each.  Per-block totals (selected, distinct 16, distinct 64), no global counter.
// With `copy`, the block also writes the bytes it reads to copy[0..n) (the build's private
// text copy is filled by the same pass over the text).
__global__ __launch_bounds__(kBlock) void k_feat_sketch(const uint8_t* __restrict__ T, uint64_t n,
                                                       uint64_t* __restrict__ tab16, uint64_t* __restrict__ tab64,
                                                       uint32_t mask, uint32_t epoch, uint32_t* __restrict__ part,
                                                       uint8_t* __restrict__ copy)
{
    __shared__ uint32_t qn, red[3];
    __shared__ uint32_t queue[kSketchQueue];
    if (threadIdx.x < 3) red[threadIdx.x] = 0;
    if (threadIdx.x == 0) qn = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kSketchBlock;
    const uint64_t p0 = base + (uint64_t)threadIdx.x * kSketchPer;
    uint32_t d16 = 0, d64 = 0;
    auto take = [&](uint64_t p) {
        const uint32_t slot = atomicAdd(&qn, 1u);
        if (slot < kSketchQueue) {
            queue[slot] = (uint32_t)(p - base);
        } else {  // queue full: insert directly
            uint32_t f16, f64;
            sketch_fingerprints(T, p, f16, f64);
            d16 += sketch_insert(tab16, mask, epoch, f16) ? 1u : 0u;
            d64 += sketch_insert(tab64, mask, epoch, f64) ? 1u : 0u;
        }
    };
    if (p0 + kSketchPer + 8 <= n) {
        // words x[0..17] = T[p0 .. p0+71]; p0 is a multiple of 64
        uint32_t x[18];
        const uint4* v = reinterpret_cast<const uint4*>(T + p0);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint4 y = v[i];
            x[4 * i] = y.x;
            x[4 * i + 1] = y.y;
            x[4 * i + 2] = y.z;
            x[4 * i + 3] = y.w;
        }
        const uint2 y = *reinterpret_cast<const uint2*>(T + p0 + kSketchPer);
        x[16] = y.x;
        x[17] = y.y;
        if (copy) {
            uint4* c = reinterpret_cast<uint4*>(copy + p0);
#pragma unroll
            for (int i = 0; i < 4; i++) c[i] = make_uint4(x[4 * i], x[4 * i + 1], x[4 * i + 2], x[4 * i + 3]);
        }
        uint64_t selmask = 0;
#pragma unroll
        for (int j = 0; j < kSketchPer; j++) {
            const int i = j >> 2, s = j & 3;
            const uint32_t lo = s ? __builtin_amdgcn_alignbyte(x[i + 1], x[i], s) : x[i];
            const uint32_t hi = s ? __builtin_amdgcn_alignbyte(x[i + 2], x[i + 1], s) : x[i + 1];
            selmask |= (uint64_t)(sketch_selected(lo, hi) ? 1u : 0u) << j;
        }
        while (selmask) {
            const int j = __builtin_ctzll(selmask);
            selmask &= selmask - 1;
            if (p0 + j + 64 <= n) take(p0 + j);
        }
    } else if (p0 < n) {  // tail: byte loads inside the text
        if (copy)
            for (uint64_t q = p0; q < p0 + kSketchPer && q < n; q++) copy[q] = T[q];
        for (int j = 0; j < kSketchPer && p0 + j + 64 <= n; j++) {
            const uint64_t p = p0 + j;
            uint32_t lo = 0, hi = 0;
            for (int b = 0; b < 4; b++) {
                lo |= (uint32_t)T[p + b] << (8 * b);
                hi |= (uint32_t)T[p + 4 + b] << (8 * b);
            }
            if (sketch_selected(lo, hi)) take(p);
        }
    }
    __syncthreads();
    const uint32_t queued = min(qn, (uint32_t)kSketchQueue);
    for (uint32_t i = threadIdx.x; i < queued; i += kBlock) {
        uint32_t f16, f64;
        sketch_fingerprints(T, base + queue[i], f16, f64);
        d16 += sketch_insert(tab16, mask, epoch, f16) ? 1u : 0u;
        d64 += sketch_insert(tab64, mask, epoch, f64) ? 1u : 0u;
    }
    if (d16) atomicAdd(&red[1], d16);
    if (d64) atomicAdd(&red[2], d64);
    __syncthreads();
    if (threadIdx.x == 0) red[0] = qn;
    if (threadIdx.x < 3) part[(uint64_t)blockIdx.x * 3 + threadIdx.x] = red[threadIdx.x];
}

// block-wide sum of a double over kSampleBlock threads (all threads get the result)
__device__ double block_sum_f64(double v, double* lds)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0;
    for (int i = 0; i < kSampleBlock / 64; i++) s += lds[i];
    return s;
}

// Pair counts, spread over kPairBlocks blocks of kSampleBlock / 256 rows each: per block the
// sums of v and v * log2(v) over its counts (then H = (V log2 T - sum v log2 v) / T for T
// pairs), and the row sums (first-byte counts), each summed in a fixed order.
constexpr int kPairBlocks = 65536 / kSampleBlock;
__global__ __launch_bounds__(kSampleBlock) void k_feat_pairent(const uint32_t* __restrict__ hist2,
                                                              double* __restrict__ epart,
                                                              uint32_t* __restrict__ rows)
{
    constexpr int kRows = kSampleBlock / 256, kWavesPerRow = 256 / 64;
    __shared__ double red[kSampleBlock / 64];
    __shared__ uint32_t ws[kSampleBlock / 64];
    const uint32_t v = hist2[blockIdx.x * kSampleBlock + threadIdx.x];
    const uint32_t rv = wave_sum(v);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = rv;
    const double dv = (double)v;
    const double a = block_sum_f64(dv, red);
    const double b = block_sum_f64(v ? dv * log2(dv) : 0.0, red);
    if (threadIdx.x == 0) {
        epart[2 * blockIdx.x] = a;
        epart[2 * blockIdx.x + 1] = b;
    }
    if (threadIdx.x < kRows) {
        uint32_t r = 0;
        for (int w = 0; w < kWavesPerRow; w++) r += ws[threadIdx.x * kWavesPerRow + w];
        rows[blockIdx.x * kRows + threadIdx.x] = r;
    }
}

// Norm of the features: x_k -> (x_k - mean_k) / std_k; null for raw features.
struct FeatNorm {
    const float* mean;
    const float* std;
};

// One block: totals of the sample and sketch partials, order-0 and order-1 entropies of the
// sample, and the feature vector of every candidate (raw into X; normalised into Xn if given).
__global__ __launch_bounds__(kSampleBlock) void k_feat_finish(const uint32_t* __restrict__ spart, uint32_t chunks,
                                                             const uint32_t* __restrict__ hist2,
                                                             const uint32_t* __restrict__ kpart, uint32_t kblocks,
                                                             const double* __restrict__ epart,
                                                             const uint32_t* __restrict__ rows, uint64_t n,
                                                             FeatNorm norm, float* __restrict__ X,
                                                             float* __restrict__ Xn)
{
    __shared__ uint32_t tot[kSpWords];
    __shared__ uint32_t h1[256];
    __shared__ uint32_t sk[3];
    __shared__ double red[kSampleBlock / 64];
    const int t = threadIdx.x;
    if (t < kSpWords) tot[t] = 0;
    if (t < 256) h1[t] = rows[t];
    if (t < 3) sk[t] = 0;
    __syncthreads();
    // chunk partials: all threads load at once (integer sums, any order)
    for (uint32_t i = t; i < chunks * (uint32_t)kSpWords; i += kSampleBlock) atomicAdd(&tot[i % kSpWords], spart[i]);
    __syncthreads();
    if (t == 0 && tot[kSpLast]) h1[(tot[kSpLast] - 1) & 0xFFu]++;  // position n-1 has no pair
    uint32_t a0 = 0, a1 = 0, a2 = 0;
#pragma unroll 8
    for (uint32_t b = t; b < kblocks; b += kSampleBlock) {
        a0 += kpart[(uint64_t)b * 3];
        a1 += kpart[(uint64_t)b * 3 + 1];
        a2 += kpart[(uint64_t)b * 3 + 2];
    }
    a0 = wave_sum(a0);
    a1 = wave_sum(a1);
    a2 = wave_sum(a2);
    if ((t & 63) == 0) {
        atomicAdd(&sk[0], a0);
        atomicAdd(&sk[1], a1);
        atomicAdd(&sk[2], a2);
    }
    __syncthreads();
    // entropies (bits): H0 of the bytes, Hpair of the byte pairs
    const double tot1 = (double)tot[kSpPos];
    const double tot2 = tot1 - (tot[kSpLast] ? 1.0 : 0.0);
    double e0 = 0, s1 = 0;
    if (t < 256 && h1[t]) {
        const double p = h1[t] / tot1;
        e0 = -p * log2(p);
        s1 = 1;
    }
    const double H0 = block_sum_f64(e0, red);
    const double sigma = block_sum_f64(s1, red);
    // pair entropy from the block partials, summed in block order
    double sv = 0, svl = 0;
    for (int b = 0; b < kPairBlocks; b++) {
        sv += epart[2 * b];
        svl += epart[2 * b + 1];
    }
    const double Hpair = (tot2 > 0) ? (sv * log2(tot2) - svl) / tot2 : 0.0;
    if (t >= kNumCand) return;
    const double H1 = fmax(0.0, Hpair - H0);
    const uint32_t m = sk[0];
    const double rep16 = m ? (double)sk[1] / m : 1.0;
    const double rep64 = m ? (double)sk[2] / m : 1.0;
    const double pos = fmax(1.0, tot1);
    float x[kIn];
#pragma unroll
    for (int k = 0; k < kIn; k++) x[k] = 0.f;
    const int c = t;
    const int W = c < kNumGrid ? kGridW[c / kNumP] : 5;
    const int P = c < kNumGrid ? kGridP[c % kNumP] : 16;
    const int wi = c < kNumGrid ? c / kNumP : 0;
    const double dens = (tot[kSpTrig + c] + 1.0) / (pos + 1.0);
    x[0] = (float)(log2((double)n) / 32.0);
    x[1] = (float)(log2(fmax(1.0, sigma)) / 8.0);
    x[2] = (float)(H0 / 8.0);
    x[3] = (float)(H1 / 8.0);
    x[4] = (float)rep16;
    x[5] = (float)rep64;
    x[6] = (float)log(rep16 + 1e-6);
    x[7] = (float)log(rep64 + 1e-6);
#pragma unroll
    for (int k = 0; k < kNumW; k++) x[8 + k] = (k == wi) ? 1.f : 0.f;
    x[12] = (float)(P / 100.0);
    x[13] = (float)(log((double)P) / 5.0);
    x[14] = (float)(1.0 / P);
    x[15] = (float)(W / 35.0);
    x[16] = (float)dens;
    x[17] = (float)log(dens);
    x[18] = (float)(log(1.0 / dens + W) / 5.0);          // mean phrase length
    x[19] = (float)(log((double)n * dens + 1.0) / 32.0);  // parse length
    x[20] = (float)(rep64 / dens / 100.0);                // distinct content x phrase length
    x[21] = (float)(log(rep64 + 1e-6) + log(1.0 / dens + W));
#pragma unroll
    for (int k = 0; k < kIn; k++) {
        X[c * kIn + k] = x[k];
        if (Xn) Xn[c * kIn + k] = (x[k] - norm.mean[k]) / norm.std[k];
    }
}

// Features of all candidates into the device array X[kNumCand][kIn] (and normalised into Xn).
void wp_features_device(Storage& S, const uint8_t* T, uint64_t n, float* X, FeatNorm norm, float* Xn,
                        uint8_t* copy = nullptr)
{
    uint32_t* hist2 = S.buf<uint32_t>(B_SEL_HIST2, 65536);
    fill<uint32_t>(S, hist2, 65536, 0u);
    GridDiv gd;
    for (int pi = 0; pi < kNumP; pi++) {
        const uint32_t P = (uint32_t)kGridP[pi];
        uint32_t x32 = P;
        for (int i = 0; i < 5; i++) x32 *= 2 - P * x32;
        gd.inv32[pi] = x32;
        gd.lim32[pi] = 0xFFFFFFFFu / P;
        uint64_t x = P;
        for (int i = 0; i < 6; i++) x *= 2 - (uint64_t)P * x;
        gd.inv64[pi] = x;
        gd.lim64[pi] = ~0ull / P;
    }
    const uint64_t stride = (n > (uint64_t)kChunk * kSampleChunks) ? (n / kSampleChunks) & ~15ull : kChunk;
    const uint32_t chunks = (uint32_t)std::min<uint64_t>(kSampleChunks, ceil_div(n, (uint64_t)kChunk));
    uint32_t* spart = S.buf<uint32_t>(B_SEL_COUNTS, (size_t)kSampleChunks * kSpWords);
    k_feat_sample<<<chunks, kSampleBlock, 0, S.stream>>>(T, n, stride, gd, hist2, spart);
    MIBWT_LAUNCH_CHECK();
    // sketch of the whole text: two epoch-tagged hash sets of 2^k >= n/256 slots (about n/1024
    // positions are selected, so the load factor stays below 1/4)
    uint32_t bits = 12;
    while ((1ull << bits) < n / 256 && bits < 30) bits++;
    const uint64_t slots = 1ull << bits;
    const size_t have = ((size_t)B_SEL_SKETCH < S.bufs.size()) ? S.bufs[B_SEL_SKETCH].cap : 0;
    uint64_t* tab = S.buf<uint64_t>(B_SEL_SKETCH, 2 * slots);
    if (have != S.bufs[B_SEL_SKETCH].cap || S.sketch_epoch == 0 || S.sketch_epoch == 0xFFFFFFFFu) {
        MIBWT_HIP(hipMemsetAsync(tab, 0, S.bufs[B_SEL_SKETCH].cap, S.stream));
        S.sketch_epoch = 0;
    }
    S.sketch_epoch++;
    const uint32_t blocks = (uint32_t)ceil_div(n, kSketchBlock);
    uint32_t* kpart = S.buf<uint32_t>(B_SEL_PART, (size_t)blocks * 3);
    k_feat_sketch<<<blocks, kBlock, 0, S.stream>>>(T, n, tab, tab + slots, (uint32_t)(slots - 1), S.sketch_epoch,
                                                   kpart, copy);
    MIBWT_LAUNCH_CHECK();
    double* epart = S.buf<double>(B_SEL_EPART, 2 * kPairBlocks);
    uint32_t* rows = S.buf<uint32_t>(B_SEL_ROWS, 256);
    k_feat_pairent<<<kPairBlocks, kSampleBlock, 0, S.stream>>>(hist2, epart, rows);
    MIBWT_LAUNCH_CHECK();
    k_feat_finish<<<1, kSampleBlock, 0, S.stream>>>(spart, chunks, hist2, kpart, blocks, epart, rows, n, norm, X,
                                                    Xn);
    MIBWT_LAUNCH_CHECK();
}

// Raw features of all candidates into host memory X[kNumCand][kIn].
void wp_features(Storage& S, const uint8_t* T, uint64_t n, float* X)
{
    float* dX = S.buf<float>(B_SEL_XRAW, (size_t)kNumCand * kIn);
    wp_features_device(S, T, n, dX, FeatNorm{nullptr, nullptr}, nullptr);
    MIBWT_HIP(hipMemcpyAsync(X, dX, sizeof(float) * kNumCand * kIn, hipMemcpyDeviceToHost, S.stream));
    MIBWT_HIP(hipStreamSynchronize(S.stream));
}

// ------------------------------------------------------------------------ MLP kernels

struct MlpDev {
    // bf16 weights used by forward/backward, fp32 biases
    const uint16_t* W1;  // [kHid][kIn]
    const float* b1;
    const uint16_t* W2;  // [kHid][kHid]
    const float* b2;
    const uint16_t* W3;  // [kOut][kHid]
    const float* b3;
};

// one thread per (sample, output unit) of a layer: Y = act(W X + b)
template <int IN, int OUT, bool RELU>
__global__ void k_layer_fwd(const float* __restrict__ X, const uint16_t* __restrict__ Wt, const float* __restrict__ b,
                            uint32_t B, float* __restrict__ Y)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * OUT) return;
    const uint32_t s = t / OUT, o = t % OUT;
    float acc = b[o];
    const float* x = X + (uint64_t)s * IN;
    const uint16_t* w = Wt + (uint64_t)o * IN;
#pragma unroll 8
    for (int k = 0; k < IN; k++) acc = fmaf(bf16_to_f32(w[k]), x[k], acc);
    Y[t] = RELU ? fmaxf(acc, 0.f) : acc;
}

// output gradient of the mean squared error over B samples and kOut targets; loss summed
__global__ void k_mse_grad(const float* __restrict__ Y, const float* __restrict__ Tg, uint32_t B,
                           float* __restrict__ dY, float* __restrict__ loss)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * kOut) return;
    const float e = Y[t] - Tg[t];
    dY[t] = 2.f * e / (float)(B * kOut);
    atomicAdd(loss, e * e / (float)(B * kOut));
}

// dX[s][k] = sum_o dY[s][o] W[o][k], masked by relu'(Xact) when the input was a relu output
template <int IN, int OUT>
__global__ void k_layer_bwd_x(const float* __restrict__ dY, const uint16_t* __restrict__ Wt,
                              const float* __restrict__ Xact, uint32_t B, float* __restrict__ dX)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * IN) return;
    const uint32_t s = t / IN, k = t % IN;
    float acc = 0.f;
    for (int o = 0; o < OUT; o++) acc = fmaf(dY[(uint64_t)s * OUT + o], bf16_to_f32(Wt[(uint64_t)o * IN + k]), acc);
    dX[t] = (Xact[t] > 0.f) ? acc : 0.f;
}

// dW[o][k] = sum_s dY[s][o] X[s][k]; db[o] = sum_s dY[s][o]  (fp32 accumulation)
template <int IN, int OUT>
__global__ void k_layer_bwd_w(const float* __restrict__ dY, const float* __restrict__ X, uint32_t B,
                              float* __restrict__ dW, float* __restrict__ db)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= OUT * (IN + 1)) return;
    const uint32_t o = t / (IN + 1), k = t % (IN + 1);
    float acc = 0.f;
    if (k < IN) {
        for (uint32_t s = 0; s < B; s++) acc = fmaf(dY[(uint64_t)s * OUT + o], X[(uint64_t)s * IN + k], acc);
        dW[(uint64_t)o * IN + k] = acc;
    } else {
        for (uint32_t s = 0; s < B; s++) acc += dY[(uint64_t)s * OUT + o];
        db[o] = acc;
    }
}

// Adam on fp32 master parameters; the bf16 copy used by the kernels is refreshed
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, uint16_t* __restrict__ pb, uint32_t count, float lr, float b1c,
                       float b2c)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float a = 0.9f * m1[i] + 0.1f * gi;
    const float v = 0.999f * m2[i] + 0.001f * gi * gi;
    m1[i] = a;
    m2[i] = v;
    const float mh = a / b1c, vh = v / b2c;
    const float np = p[i] - lr * mh / (sqrtf(vh) + 1e-8f);
    p[i] = np;
    if (pb) pb[i] = f32_to_bf16(np);
}

// ------------------------------------------------------------------------ model

constexpr int kW1 = kHid * kIn, kW2 = kHid * kHid, kW3 = kOut * kHid;
constexpr int kParams = kW1 + kHid + kW2 + kHid + kW3 + kOut;  // fp32 master layout: W1 b1 W2 b2 W3 b3

// An ensemble of `members` networks sharing one input / output normalisation; the selector
// averages their predictions.
constexpr int kMaxMembers = 16;
struct Model {
    bool loaded = false;
    uint64_t version = 0;  // distinct for every model loaded or trained in this process
    int members = 1;
    float xmean[kIn], xstd[kIn], ymean[kOut], ystd[kOut];
    std::vector<float> params;  // members x fp32 master layout (weights bf16-representable)
};

static Model g_model;
static std::mutex g_model_mu;
static uint64_t g_model_version = 0;  // guarded by g_model_mu

static const uint32_t kMagic = 0x4D424C50;   // "PLBM": one network
static const uint32_t kMagicE = 0x45424C50;  // "PLBE": ensemble, member count after the sizes

bool model_save(const Model& M, const char* path)
{
    FILE* f = fopen(path, "wb");
    if (!f) return false;
    const uint32_t hdr[5] = {kMagicE, (uint32_t)kIn, (uint32_t)kHid, (uint32_t)kOut, (uint32_t)M.members};
    bool ok = fwrite(hdr, 4, 5, f) == 5;
    ok = ok && fwrite(M.xmean, 4, kIn, f) == (size_t)kIn && fwrite(M.xstd, 4, kIn, f) == (size_t)kIn;
    ok = ok && fwrite(M.ymean, 4, kOut, f) == (size_t)kOut && fwrite(M.ystd, 4, kOut, f) == (size_t)kOut;
    const size_t np = (size_t)kParams * M.members;
    ok = ok && M.params.size() == np && fwrite(M.params.data(), 4, np, f) == np;
    return (fclose(f) == 0) && ok;
}

bool model_load(Model& M, const char* path)
{
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    uint32_t hdr[4];
    bool ok = fread(hdr, 4, 4, f) == 4 && (hdr[0] == kMagic || hdr[0] == kMagicE) && hdr[1] == (uint32_t)kIn &&
              hdr[2] == (uint32_t)kHid && hdr[3] == (uint32_t)kOut;
    uint32_t members = 1;
    if (ok && hdr[0] == kMagicE) ok = fread(&members, 4, 1, f) == 1 && members >= 1 && members <= kMaxMembers;
    M.members = (int)members;
    M.params.assign((size_t)kParams * members, 0.f);
    ok = ok && fread(M.xmean, 4, kIn, f) == (size_t)kIn && fread(M.xstd, 4, kIn, f) == (size_t)kIn;
    ok = ok && fread(M.ymean, 4, kOut, f) == (size_t)kOut && fread(M.ystd, 4, kOut, f) == (size_t)kOut;
    ok = ok && fread(M.params.data(), 4, M.params.size(), f) == M.params.size();
    fclose(f);
    M.loaded = ok;
    return ok;
}

// Device copies of one parameter vector (fp32 master layout) for the kernels.
struct MlpBufs {
    float* p;       // fp32 master [kParams]
    uint16_t* pb;   // bf16 copy [kParams] (weights only are read as bf16)
    float* g;       // gradients [kParams]
    MlpDev dev() const
    {
        return MlpDev{pb, p + kW1, pb + kW1 + kHid, p + kW1 + kHid + kW2, pb + kW1 + kHid + kW2 + kHid,
                      p + kW1 + kHid + kW2 + kHid + kW3};
    }
};

static __global__ void k_to_bf16(const float* __restrict__ p, uint16_t* __restrict__ pb, uint32_t count)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) pb[i] = f32_to_bf16(p[i]);
}

// forward of B normalised samples; activations in H1, H2, output in Y
void mlp_forward(Storage& S, const MlpDev& m, const float* X, uint32_t B, float* H1, float* H2, float* Y)
{
    k_layer_fwd<kIn, kHid, true><<<grid_for((uint64_t)B * kHid), kBlock, 0, S.stream>>>(X, m.W1, m.b1, B, H1);
    k_layer_fwd<kHid, kHid, true><<<grid_for((uint64_t)B * kHid), kBlock, 0, S.stream>>>(H1, m.W2, m.b2, B, H2);
    k_layer_fwd<kHid, kOut, false><<<grid_for((uint64_t)B * kOut), kBlock, 0, S.stream>>>(H2, m.W3, m.b3, B, Y);
    MIBWT_LAUNCH_CHECK();
}

// loss and gradients (fp32 master layout) of one batch
void mlp_backward(Storage& S, const MlpDev& m, const float* X, const float* Tg, uint32_t B, float* H1, float* H2,
                  float* Y, float* dY, float* dH2, float* dH1, float* g, float* loss_dev)
{
    mlp_forward(S, m, X, B, H1, H2, Y);
    k_mse_grad<<<grid_for((uint64_t)B * kOut), kBlock, 0, S.stream>>>(Y, Tg, B, dY, loss_dev);
    k_layer_bwd_w<kHid, kOut><<<grid_for(kOut * (kHid + 1)), kBlock, 0, S.stream>>>(dY, H2, B, g + kW1 + kHid + kW2 + kHid,
                                                                                 g + kW1 + kHid + kW2 + kHid + kW3);
    k_layer_bwd_x<kHid, kOut><<<grid_for((uint64_t)B * kHid), kBlock, 0, S.stream>>>(dY, m.W3, H2, B, dH2);
    k_layer_bwd_w<kHid, kHid><<<grid_for(kHid * (kHid + 1)), kBlock, 0, S.stream>>>(dH2, H1, B, g + kW1 + kHid,
                                                                                 g + kW1 + kHid + kW2);
    k_layer_bwd_x<kHid, kHid><<<grid_for((uint64_t)B * kHid), kBlock, 0, S.stream>>>(dH2, m.W2, H1, B, dH1);
    k_layer_bwd_w<kIn, kHid><<<grid_for(kHid * (kIn + 1)), kBlock, 0, S.stream>>>(dH1, X, B, g, g + kW1);
    MIBWT_LAUNCH_CHECK();
}

// ------------------------------------------------------------------------ training / inference

struct MlpWork {
    float *X, *Tg, *H1, *H2, *Y, *dY, *dH2, *dH1, *loss;
};

static MlpWork mlp_work(Storage& S, uint32_t B)
{
    MlpWork w;
    w.X = S.buf<float>(B_SEL_X, (size_t)B * kIn);
    w.Tg = S.buf<float>(B_SEL_T, (size_t)B * kOut);
    w.H1 = S.buf<float>(B_SEL_H1, (size_t)B * kHid);
    w.H2 = S.buf<float>(B_SEL_H2, (size_t)B * kHid);
    w.Y = S.buf<float>(B_SEL_Y, (size_t)B * kOut);
    w.dY = S.buf<float>(B_SEL_DY, (size_t)B * kOut);
    w.dH2 = S.buf<float>(B_SEL_DH2, (size_t)B * kHid);
    w.dH1 = S.buf<float>(B_SEL_DH1, (size_t)B * kHid);
    w.loss = S.buf<float>(B_SEL_LOSS, 4);
    return w;
}

static MlpBufs mlp_params(Storage& S, const float* host_params)
{
    MlpBufs m;
    m.p = S.buf<float>(B_SEL_P, kParams);
    m.pb = S.buf<uint16_t>(B_SEL_PB, kParams);
    m.g = S.buf<float>(B_SEL_G, kParams);
    MIBWT_HIP(hipMemcpyAsync(m.p, host_params, sizeof(float) * kParams, hipMemcpyHostToDevice, S.stream));
    k_to_bf16<<<grid_for(kParams), kBlock, 0, S.stream>>>(m.p, m.pb, kParams);
    MIBWT_LAUNCH_CHECK();
    return m;
}

// gradients of the mean squared error with respect to the fp32 master layout (the kernels
// compute with the bf16-rounded weights)
float mlp_grad(Storage& S, const float* params, const float* X, const float* Y, uint32_t B, float* grads)
{
    MlpBufs m = mlp_params(S, params);
    MlpWork w = mlp_work(S, B);
    MIBWT_HIP(hipMemcpyAsync(w.X, X, sizeof(float) * B * kIn, hipMemcpyHostToDevice, S.stream));
    MIBWT_HIP(hipMemcpyAsync(w.Tg, Y, sizeof(float) * B * kOut, hipMemcpyHostToDevice, S.stream));
    fill<float>(S, w.loss, 1, 0.f);
    mlp_backward(S, m.dev(), w.X, w.Tg, B, w.H1, w.H2, w.Y, w.dY, w.dH2, w.dH1, m.g, w.loss);
    MIBWT_HIP(hipMemcpyAsync(grads, m.g, sizeof(float) * kParams, hipMemcpyDeviceToHost, S.stream));
    float loss = 0;
    S.read(&loss, w.loss);
    return loss;
}

void mlp_predict_raw(Storage& S, const float* params, const float* X, uint32_t B, float* Y)
{
    MlpBufs m = mlp_params(S, params);
    MlpWork w = mlp_work(S, B);
    MIBWT_HIP(hipMemcpyAsync(w.X, X, sizeof(float) * B * kIn, hipMemcpyHostToDevice, S.stream));
    mlp_forward(S, m.dev(), w.X, B, w.H1, w.H2, w.Y);
    MIBWT_HIP(hipMemcpyAsync(Y, w.Y, sizeof(float) * B * kOut, hipMemcpyDeviceToHost, S.stream));
    MIBWT_HIP(hipStreamSynchronize(S.stream));
}

static uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Train one network on the normalised samples X, Y (B rows, shuffled by `seed`) from a He
// initialisation drawn from `seed`; returns the final full-batch loss, params[kParams] out.
static float mlp_train_member(Storage& S, const std::vector<float>& Xn, const std::vector<float>& Yn, uint32_t B,
                              int epochs, uint32_t mb, float lr, uint32_t seed, float* params)
{
    std::vector<uint32_t> perm(B);
    for (uint32_t i = 0; i < B; i++) perm[i] = i;
    uint64_t st = seed * 0x2545F4914F6CDD1Dull + 1;
    for (uint32_t i = B; i > 1; i--) std::swap(perm[i - 1], perm[splitmix(st) % i]);
    std::vector<float> X((size_t)B * kIn), Y((size_t)B * kOut);
    for (uint32_t i = 0; i < B; i++) {
        for (int k = 0; k < kIn; k++) X[(size_t)i * kIn + k] = Xn[(size_t)perm[i] * kIn + k];
        for (int k = 0; k < kOut; k++) Y[(size_t)i * kOut + k] = Yn[(size_t)perm[i] * kOut + k];
    }
    // He initialisation, weights rounded to bf16
    std::vector<float> P0(kParams, 0.f);
    auto init = [&](size_t off, int count, int fan_in) {
        const double sd = std::sqrt(2.0 / fan_in);
        for (int i = 0; i < count; i++) {
            const double u1 = (splitmix(st) >> 11) * (1.0 / 9007199254740992.0) + 1e-12;
            const double u2 = (splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
            P0[off + i] = bf16_to_f32(f32_to_bf16((float)(z * sd)));
        }
    };
    init(0, kW1, kIn);
    init(kW1 + kHid, kW2, kHid);
    init(kW1 + kHid + kW2 + kHid, kW3, kHid);
    MlpBufs m = mlp_params(S, P0.data());
    float* m1 = S.buf<float>(B_SEL_M1, kParams);
    float* m2 = S.buf<float>(B_SEL_M2, kParams);
    fill<float>(S, m1, kParams, 0.f);
    fill<float>(S, m2, kParams, 0.f);
    MlpWork w = mlp_work(S, B);
    MIBWT_HIP(hipMemcpyAsync(w.X, X.data(), sizeof(float) * X.size(), hipMemcpyHostToDevice, S.stream));
    MIBWT_HIP(hipMemcpyAsync(w.Tg, Y.data(), sizeof(float) * Y.size(), hipMemcpyHostToDevice, S.stream));
    if (mb == 0 || mb > B) mb = B;
    const uint32_t steps_per_epoch = (B + mb - 1) / mb;
    float* H1 = S.buf<float>(B_SEL_H1B, (size_t)mb * kHid);
    float* H2 = S.buf<float>(B_SEL_H2B, (size_t)mb * kHid);
    float* Yb = S.buf<float>(B_SEL_YB, (size_t)mb * kOut);
    int t = 0;
    for (int ep = 0; ep < epochs; ep++) {
        for (uint32_t sidx = 0; sidx < steps_per_epoch; sidx++) {
            const uint32_t off = sidx * mb;
            const uint32_t bs = (off + mb <= B) ? mb : B - off;
            mlp_backward(S, m.dev(), w.X + (size_t)off * kIn, w.Tg + (size_t)off * kOut, bs, H1, H2, Yb, w.dY, w.dH2,
                         w.dH1, m.g, w.loss);
            t++;
            const float b1c = 1.f - std::pow(0.9f, (float)t), b2c = 1.f - std::pow(0.999f, (float)t);
            // weights keep a bf16 copy; biases are read in fp32
            k_adam<<<grid_for(kParams), kBlock, 0, S.stream>>>(m.p, m.g, m1, m2, m.pb, kParams, lr, b1c, b2c);
            MIBWT_LAUNCH_CHECK();
        }
    }
    // final full-batch loss
    fill<float>(S, w.loss, 1, 0.f);
    mlp_forward(S, m.dev(), w.X, B, w.H1, w.H2, w.Y);
    k_mse_grad<<<grid_for((uint64_t)B * kOut), kBlock, 0, S.stream>>>(w.Y, w.Tg, B, w.dY, w.loss);
    MIBWT_LAUNCH_CHECK();
    float loss = 0;
    S.read(&loss, w.loss);
    // the model file stores bf16-representable weights
    MIBWT_HIP(hipMemcpyAsync(params, m.p, sizeof(float) * kParams, hipMemcpyDeviceToHost, S.stream));
    MIBWT_HIP(hipStreamSynchronize(S.stream));
    for (int i = 0; i < kParams; i++) params[i] = bf16_to_f32(f32_to_bf16(params[i]));
    return loss;
}

// Train an ensemble of `members` networks (seeds seed, seed+1, ...) on B samples (raw features /
// targets) with one shared normalisation; writes the model file; returns the members' mean
// final training loss (normalised targets).
float mlp_train(Storage& S, const float* Xraw, const float* Yraw, uint32_t B, int epochs, uint32_t mb, float lr,
                uint32_t seed, const char* path, int members)
{
    MIBWT_REQUIRE(members >= 1 && members <= kMaxMembers, LIBCUBWT_BAD_PARAMETER, "ensemble size out of range");
    Model M;
    M.members = members;
    // normalisation
    for (int k = 0; k < kIn; k++) {
        double s1 = 0, s2 = 0;
        for (uint32_t i = 0; i < B; i++) {
            s1 += Xraw[(size_t)i * kIn + k];
            s2 += (double)Xraw[(size_t)i * kIn + k] * Xraw[(size_t)i * kIn + k];
        }
        const double mu = s1 / B, var = s2 / B - mu * mu;
        M.xmean[k] = (float)mu;
        M.xstd[k] = (float)(var > 1e-12 ? std::sqrt(var) : 1.0);
    }
    for (int k = 0; k < kOut; k++) {
        double s1 = 0, s2 = 0;
        for (uint32_t i = 0; i < B; i++) {
            s1 += Yraw[(size_t)i * kOut + k];
            s2 += (double)Yraw[(size_t)i * kOut + k] * Yraw[(size_t)i * kOut + k];
        }
        const double mu = s1 / B, var = s2 / B - mu * mu;
        M.ymean[k] = (float)mu;
        M.ystd[k] = (float)(var > 1e-12 ? std::sqrt(var) : 1.0);
    }
    std::vector<float> Xn((size_t)B * kIn), Yn((size_t)B * kOut);
    for (uint32_t i = 0; i < B; i++) {
        for (int k = 0; k < kIn; k++) Xn[(size_t)i * kIn + k] = (Xraw[(size_t)i * kIn + k] - M.xmean[k]) / M.xstd[k];
        for (int k = 0; k < kOut; k++) Yn[(size_t)i * kOut + k] = (Yraw[(size_t)i * kOut + k] - M.ymean[k]) / M.ystd[k];
    }
    M.params.assign((size_t)kParams * members, 0.f);
    double loss = 0;
    for (int e = 0; e < members; e++)
        loss += mlp_train_member(S, Xn, Yn, B, epochs, mb, lr, seed + (uint32_t)e, M.params.data() + (size_t)e * kParams);
    M.loaded = true;
    if (path && !model_save(M, path)) throw Error(LIBCUBWT_BAD_PARAMETER, "cannot write the selector model");
    std::lock_guard<std::mutex> lk(g_model_mu);
    M.version = ++g_model_version;
    g_model = M;
    return (float)(loss / members);
}

bool selector_load(const char* path)
{
    Model M;
    if (!model_load(M, path)) return false;
    std::lock_guard<std::mutex> lk(g_model_mu);
    M.version = ++g_model_version;
    g_model = M;
    return true;
}

static void ensure_default_model()
{
    static std::once_flag once;
    std::call_once(once, []() {
        const char* env = getenv("MIBWT_SELECTOR");
        if (env && *env) {
            selector_load(env);
            return;
        }
        Dl_info info;
        if (dladdr((void*)&selector_load, &info) && info.dli_fname) {
            std::string p = info.dli_fname;
            const size_t k = p.rfind('/');
            p = (k == std::string::npos ? std::string(".") : p.substr(0, k)) + "/../selector_model.bin";
            selector_load(p.c_str());
        }
    });
}

// Result of one device-side choice: the chosen candidate and the predicted log(ms per byte)
// of every candidate (ensemble mean, smoothed along P).
struct SelPick {
    uint32_t best;
    float pred[kNumCand];
};

// One wave: the ensemble's mean de-normalised predicted time of each candidate, smoothed over
// each grid candidate and its P-neighbours of the same W (a choice in a flat stretch of the
// grid is robust to the error of any single prediction), argmin (lowest index on ties, NaN
// never chosen).  Y holds `members` blocks of kNumCand x kOut outputs.
__global__ void k_sel_pick(const float* __restrict__ Y, int members, const float* __restrict__ ynorm /* mean[3] std[3] */,
                           SelPick* __restrict__ out)
{
    __shared__ float v[64];
    const uint32_t c = threadIdx.x;
    float m = 0.f;
    if (c < kNumCand) {
        for (int e = 0; e < members; e++) m += Y[((size_t)e * kNumCand + c) * kOut + 2];
        m = (m / members) * ynorm[kOut + 2] + ynorm[2];
    }
    v[c] = m;
    __syncthreads();
    float t = INFINITY;
    if (c < kNumCand) {
        t = m;
        if (c < kNumGrid) {
            const uint32_t pi = c % kNumP;
            float acc = m;
            int cnt = 1;
            if (pi > 0) {
                acc += v[c - 1];
                cnt++;
            }
            if (pi + 1 < kNumP) {
                acc += v[c + 1];
                cnt++;
            }
            t = acc / cnt;
        }
        out->pred[c] = t;
        if (!(t == t)) t = INFINITY;
    }
    uint32_t bi = c;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float to = __shfl_xor(t, off);
        const uint32_t io = __shfl_xor(bi, off);
        if (to < t || (to == t && io < bi)) {
            t = to;
            bi = io;
        }
    }
    if (c == 0) out->best = (t == INFINITY) ? 0u : bi;
}

// Inference of every member of the ensemble in one launch: block (member, kFwdRows samples)
// keeps the samples' activations in LDS and evaluates each unit with the same fmaf chain as
// k_layer_fwd, so the outputs equal the layer-by-layer forward bit for bit.
constexpr int kFwdRows = 8;
struct Nets {
    MlpDev dev[kMaxMembers];
};

__global__ __launch_bounds__(kBlock) void k_mlp_forward_all(const float* __restrict__ X, uint32_t B, Nets nets,
                                                           float* __restrict__ Y)
{
    __shared__ float xs[kFwdRows][kIn];
    __shared__ float h1[kFwdRows][kHid];
    __shared__ float h2[kFwdRows][kHid];
    // hidden-layer weights transposed in LDS ([input][unit], rows padded to an odd number of
    // words so that the transposing stores do not collide): the lanes of a wavefront read
    // consecutive units of one input
    constexpr int kRow = kHid + 2;
    __shared__ uint16_t w1t[kIn * kRow];
    __shared__ uint16_t w2t[kHid * kRow];
    const MlpDev m = nets.dev[blockIdx.y];
    const uint32_t s0 = blockIdx.x * kFwdRows;
    const uint32_t rows = (B - s0 < (uint32_t)kFwdRows) ? B - s0 : (uint32_t)kFwdRows;
    for (uint32_t i = threadIdx.x; i < rows * kIn; i += kBlock) xs[i / kIn][i % kIn] = X[(uint64_t)s0 * kIn + i];
    for (uint32_t i = threadIdx.x; i < kHid * kIn; i += kBlock) w1t[(i % kIn) * kRow + i / kIn] = m.W1[i];
    for (uint32_t i = threadIdx.x; i < kHid * kHid; i += kBlock) w2t[(i % kHid) * kRow + i / kHid] = m.W2[i];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < rows * kHid; i += kBlock) {
        const uint32_t r = i / kHid, o = i % kHid;
        float acc = m.b1[o];
#pragma unroll 8
        for (int k = 0; k < kIn; k++) acc = fmaf(bf16_to_f32(w1t[k * kRow + o]), xs[r][k], acc);
        h1[r][o] = fmaxf(acc, 0.f);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < rows * kHid; i += kBlock) {
        const uint32_t r = i / kHid, o = i % kHid;
        float acc = m.b2[o];
#pragma unroll 8
        for (int k = 0; k < kHid; k++) acc = fmaf(bf16_to_f32(w2t[k * kRow + o]), h1[r][k], acc);
        h2[r][o] = fmaxf(acc, 0.f);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < rows * kOut; i += kBlock) {
        const uint32_t r = i / kOut, o = i % kOut;
        float acc = m.b3[o];
        const uint16_t* w = m.W3 + (uint64_t)o * kHid;
#pragma unroll 8
        for (int k = 0; k < kHid; k++) acc = fmaf(bf16_to_f32(w[k]), h2[r][k], acc);
        Y[((uint64_t)blockIdx.y * B + s0 + r) * kOut + o] = acc;
    }
}

// The loaded model's parameters, bf16 weights and normalisation on this Storage's device;
// uploaded again only when another model was loaded or trained since.
struct ModelDev {
    MlpDev dev[kMaxMembers];
    int members;
    const float* xnorm;  // mean[kIn] std[kIn]
    const float* ynorm;  // mean[kOut] std[kOut]
};

static ModelDev model_on_device(Storage& S, const Model& M)
{
    const size_t np = (size_t)kParams * M.members;
    float* p = S.buf<float>(B_SEL_MP, np);
    uint16_t* pb = S.buf<uint16_t>(B_SEL_MPB, np);
    float* nrm = S.buf<float>(B_SEL_NORM, 2 * kIn + 2 * kOut);
    if (S.sel_model_version != M.version && !M.params.empty()) {
        std::vector<float> h(2 * kIn + 2 * kOut);
        std::copy(M.xmean, M.xmean + kIn, h.begin());
        std::copy(M.xstd, M.xstd + kIn, h.begin() + kIn);
        std::copy(M.ymean, M.ymean + kOut, h.begin() + 2 * kIn);
        std::copy(M.ystd, M.ystd + kOut, h.begin() + 2 * kIn + kOut);
        MIBWT_HIP(hipMemcpyAsync(p, M.params.data(), sizeof(float) * np, hipMemcpyHostToDevice, S.stream));
        MIBWT_HIP(hipMemcpyAsync(nrm, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice, S.stream));
        k_to_bf16<<<grid_for(np), kBlock, 0, S.stream>>>(p, pb, (uint32_t)np);
        MIBWT_LAUNCH_CHECK();
        MIBWT_HIP(hipStreamSynchronize(S.stream));  // the host copies above are temporaries
        S.sel_model_version = M.version;
    }
    ModelDev md{};
    md.members = M.members;
    for (int e = 0; e < M.members; e++) {
        MlpBufs m{p + (size_t)e * kParams, pb + (size_t)e * kParams, nullptr};
        md.dev[e] = m.dev();
    }
    md.xnorm = nrm;
    md.ynorm = nrm + 2 * kIn;
    return md;
}

// Choose (W,P) for the text resident at T (padded device copy); fixed W=5,P=16 without a model.
// Features, normalisation, the ensemble's networks and the choice run on the device; one
// read-back.
// With `copy`, the text is also copied to copy[0..n) (by the sketch pass, or by a plain copy
// when no model is loaded).
void selector_choose_device(Storage& S, const uint8_t* T, uint64_t n, int* W, int* P, float* pred_out,
                            uint8_t* copy)
{
    ensure_default_model();
    // the model's header only (its parameters are uploaded under the lock when this storage
    // does not hold them yet): no per-call copy of the parameters
    Model M;
    ModelDev md{};
    {
        std::lock_guard<std::mutex> lk(g_model_mu);
        M.loaded = g_model.loaded;
        M.version = g_model.version;
        M.members = g_model.members;
        if (M.loaded) md = model_on_device(S, g_model);
    }
    *W = 5;
    *P = 16;
    if (!M.loaded) {
        if (copy) MIBWT_HIP(hipMemcpyAsync(copy, T, n, hipMemcpyDeviceToDevice, S.stream));
        return;
    }
    MlpWork w = mlp_work(S, kNumCand);
    float* Xraw = S.buf<float>(B_SEL_XRAW, (size_t)kNumCand * kIn);
    float* Yall = S.buf<float>(B_SEL_YALL, (size_t)kNumCand * kOut * md.members);
    wp_features_device(S, T, n, Xraw, FeatNorm{md.xnorm, md.xnorm + kIn}, w.X, copy);
    Nets nets{};
    for (int e = 0; e < md.members; e++) nets.dev[e] = md.dev[e];
    k_mlp_forward_all<<<dim3(ceil_div((uint32_t)kNumCand, (uint32_t)kFwdRows), md.members), kBlock, 0, S.stream>>>(
        w.X, kNumCand, nets, Yall);
    MIBWT_LAUNCH_CHECK();
    SelPick* pick = S.buf<SelPick>(B_SEL_PICK, 1);
    k_sel_pick<<<1, 64, 0, S.stream>>>(Yall, md.members, md.ynorm, pick);
    MIBWT_LAUNCH_CHECK();
    SelPick h;
    S.read_fast(&h, pick);
    if (pred_out)
        for (int c = 0; c < kNumCand; c++) pred_out[c] = h.pred[c];
    const int best = (int)std::min<uint32_t>(h.best, kNumCand - 1);
    *W = cand_W(best);
    *P = cand_P(best);
}

bool selector_has_model()
{
    ensure_default_model();
    std::lock_guard<std::mutex> lk(g_model_mu);
    return g_model.loaded;
}

int wp_grid(int* W, int* P)
{
    for (int c = 0; c < kNumCand; c++) {
        W[c] = cand_W(c);
        P[c] = cand_P(c);
    }
    return kNumCand;
}

}  // namespace mibwt
