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

struct SampleCounts {
    uint32_t hist1[256];
    uint32_t trig[kNumCand];
    uint32_t positions;
    uint32_t sketch[3];  // selected positions, distinct 16-mer and 64-mer fingerprints
};

// Divisibility by the odd grid moduli without division: v % P == 0 <=> v * P^-1 <= lim
// (mod 2^32 for W <= 20, mod 2^64 for W = 35).
struct GridDiv {
    uint32_t inv32[kNumP], lim32[kNumP];
    uint64_t inv64[kNumP], lim64[kNumP];
};

constexpr int kSampleParts = 4;                          // blocks per chunk
constexpr int kSamplePer = kChunk / (kBlock * kSampleParts);  // 16 consecutive positions per thread

// Sample statistics over kSampleChunks chunks of kChunk bytes (kSampleParts blocks each): byte and
// byte-pair histograms (the pair histogram lives in LDS as packed 16-bit counters), and the
// trigger count of every grid candidate from rolled exact window values.
__global__ __launch_bounds__(kBlock) void k_feat_sample(const uint8_t* __restrict__ T, uint64_t n,
                                                       uint64_t stride, GridDiv gd, uint32_t* __restrict__ hist2,
                                                       SampleCounts* __restrict__ out)
{
    __shared__ uint32_t h2[65536 / 2];
    __shared__ uint32_t h1[256];
    __shared__ uint32_t tc[kNumCand];
    for (int i = threadIdx.x; i < 65536 / 2; i += kBlock) h2[i] = 0;
    h1[threadIdx.x] = 0;
    if (threadIdx.x < kNumCand) tc[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t cstart = (uint64_t)(blockIdx.x / kSampleParts) * stride;
    const uint64_t p0 = cstart + (uint64_t)(blockIdx.x % kSampleParts) * (kChunk / kSampleParts) +
                        (uint64_t)threadIdx.x * kSamplePer;
    uint32_t cnt[kNumCand];
#pragma unroll
    for (int c = 0; c < kNumCand; c++) cnt[c] = 0;
    uint32_t v5 = 0, v10 = 0, v20 = 0;
    uint64_t v35 = 0;
    // windows ending before p0 (bytes before the text read as 0, windows there are skipped)
    for (int k = 35; k >= 1; k--) {
        const uint32_t c = (p0 >= (uint64_t)k && p0 - k < n) ? T[p0 - k] : 0u;
        if (k <= 5) v5 = (v5 << 1) + c;
        if (k <= 10) v10 = (v10 << 1) + c;
        if (k <= 20) v20 = (v20 << 1) + c;
        v35 = (v35 << 1) + c;
    }
    uint32_t pos = 0;
    for (int q = 0; q < kSamplePer; q++) {
        const uint64_t p = p0 + q;
        if (p >= n) break;
        const uint32_t c = T[p];
        const uint32_t c5 = p >= 5 ? T[p - 5] : 0u, c10 = p >= 10 ? T[p - 10] : 0u;
        const uint32_t c20 = p >= 20 ? T[p - 20] : 0u, c35 = p >= 35 ? T[p - 35] : 0u;
        v5 = (v5 << 1) + c - (c5 << 5);
        v10 = (v10 << 1) + c - (c10 << 10);
        v20 = (v20 << 1) + c - (c20 << 20);
        v35 = (v35 << 1) + c - ((uint64_t)c35 << 35);
        atomicAdd(&h1[c], 1u);
        if (p + 1 < n) {
            const uint32_t pr = (c << 8) | T[p + 1];
            atomicAdd(&h2[pr >> 1], 1u << (16 * (pr & 1u)));
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
#pragma unroll
    for (int c = 0; c < kNumCand; c++)
        if (cnt[c]) atomicAdd(&tc[c], cnt[c]);
    if (pos) atomicAdd(&out->positions, pos);
    __syncthreads();
    if (h1[threadIdx.x]) atomicAdd(&out->hist1[threadIdx.x], h1[threadIdx.x]);
    if (threadIdx.x < kNumCand && tc[threadIdx.x]) atomicAdd(&out->trig[threadIdx.x], tc[threadIdx.x]);
    for (int i = threadIdx.x; i < 65536 / 2; i += kBlock) {
        const uint32_t w = h2[i];
        if (w & 0xFFFFu) atomicAdd(&hist2[2 * i], w & 0xFFFFu);
        if (w >> 16) atomicAdd(&hist2[2 * i + 1], w >> 16);
    }
}

constexpr int kSketchPer = 64;                              // positions per thread
constexpr uint64_t kSketchBlock = (uint64_t)kBlock * kSketchPer;

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

// Content-defined sketch of the whole text: positions whose 8-byte window hashes into the
// top 1/1024 contribute 32-bit fingerprints of the 16 and 64 bytes that start there; the
// distinct ones are counted with two hash sets.  Per-block totals, no global counter.
__global__ __launch_bounds__(kBlock) void k_feat_sketch(const uint8_t* __restrict__ T, uint64_t n,
                                                       uint64_t* __restrict__ tab16, uint64_t* __restrict__ tab64,
                                                       uint32_t mask, uint32_t epoch, uint32_t* __restrict__ part)
{
    __shared__ uint32_t red[3];
    if (threadIdx.x < 3) red[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t p0 = (uint64_t)blockIdx.x * kSketchBlock + (uint64_t)threadIdx.x * kSketchPer;
    uint32_t sel = 0, d16 = 0, d64 = 0;
    if (p0 < n) {
        const uint64_t* w = reinterpret_cast<const uint64_t*>(T + p0);  // p0 is a multiple of 64
        uint64_t cur = w[0];
        for (int q = 0; q < kSketchPer; q += 8) {
            const uint64_t nxt = w[q / 8 + 1];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const uint64_t p = p0 + q + r;
                const uint64_t win = r ? ((cur >> (8 * r)) | (nxt << (64 - 8 * r))) : cur;
                if (((win * 0x9E3779B97F4A7C15ull) >> 54) == 0 && p + 64 <= n) {
                    uint64_t h = 0x9E3779B97F4A7C15ull, a = 0;
                    for (int k = 0; k < 8; k++) {
                        h = mix64(h ^ load8(T, p + 8 * k));
                        if (k == 1) a = h;
                    }
                    sel++;
                    d16 += sketch_insert(tab16, mask, epoch, (uint32_t)(a >> 32)) ? 1u : 0u;
                    d64 += sketch_insert(tab64, mask, epoch, (uint32_t)(h >> 32)) ? 1u : 0u;
                }
            }
            cur = nxt;
        }
    }
    if (sel) {
        atomicAdd(&red[0], sel);
        atomicAdd(&red[1], d16);
        atomicAdd(&red[2], d64);
    }
    __syncthreads();
    if (threadIdx.x < 3) part[(uint64_t)blockIdx.x * 3 + threadIdx.x] = red[threadIdx.x];
}

__global__ __launch_bounds__(kBlock) void k_sum3(const uint32_t* __restrict__ part, uint32_t blocks,
                                                uint32_t* __restrict__ out)
{
    __shared__ uint32_t lds[kWaves];
    for (int c = 0; c < 3; c++) {
        uint32_t acc = 0;
        for (uint32_t b = threadIdx.x; b < blocks; b += kBlock) acc += part[(uint64_t)b * 3 + c];
        uint32_t tot;
        (void)block_exclusive_scan(acc, 0u, OpSum{}, lds, &tot);
        if (threadIdx.x == 0) out[c] = tot;
    }
}

static uint32_t inv_odd32(uint32_t q)
{
    uint32_t x = q;
    for (int i = 0; i < 5; i++) x *= 2 - q * x;
    return x;
}

// Features of all candidates (host side assembly after one read-back).
void wp_features(Storage& S, const uint8_t* T, uint64_t n, float* X /* [kNumCand][kIn] */)
{
    uint32_t* hist2 = S.buf<uint32_t>(B_SEL_HIST2, 65536);
    SampleCounts* sc = S.buf<SampleCounts>(B_SEL_COUNTS, 1);
    fill<uint32_t>(S, hist2, 65536, 0u);
    fill<uint32_t>(S, reinterpret_cast<uint32_t*>(sc), sizeof(SampleCounts) / 4, 0u);
    GridDiv gd;
    for (int pi = 0; pi < kNumP; pi++) {
        const uint32_t P = (uint32_t)kGridP[pi];
        gd.inv32[pi] = inv_odd32(P);
        gd.lim32[pi] = 0xFFFFFFFFu / P;
        uint64_t x = P;
        for (int i = 0; i < 6; i++) x *= 2 - (uint64_t)P * x;
        gd.inv64[pi] = x;
        gd.lim64[pi] = ~0ull / P;
    }
    const uint64_t stride = (n > (uint64_t)kChunk * kSampleChunks) ? n / kSampleChunks : kChunk;
    const uint32_t chunks = (uint32_t)std::min<uint64_t>(kSampleChunks, ceil_div(n, (uint64_t)kChunk));
    k_feat_sample<<<chunks * kSampleParts, kBlock, 0, S.stream>>>(T, n, stride, gd, hist2, sc);
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
    uint32_t* part = S.buf<uint32_t>(B_SEL_PART, (size_t)blocks * 3);
    k_feat_sketch<<<blocks, kBlock, 0, S.stream>>>(T, n, tab, tab + slots, (uint32_t)(slots - 1), S.sketch_epoch,
                                                   part);
    MIBWT_LAUNCH_CHECK();
    k_sum3<<<1, kBlock, 0, S.stream>>>(part, blocks, sc->sketch);
    MIBWT_LAUNCH_CHECK();
    SampleCounts h;
    S.read(&h, sc);
    const uint32_t m = h.sketch[0];
    uint32_t dist[2] = {h.sketch[1], h.sketch[2]};

    std::vector<uint32_t> h2(65536);
    S.read(h2.data(), hist2, 65536);
    // order-0 / order-1 entropies
    double tot = 0, H0 = 0, sigma = 0;
    for (int c = 0; c < 256; c++) tot += h.hist1[c];
    for (int c = 0; c < 256; c++)
        if (h.hist1[c]) {
            sigma += 1;
            const double p = h.hist1[c] / tot;
            H0 -= p * std::log2(p);
        }
    double tot2 = 0, Hpair = 0;
    for (uint32_t v : h2) tot2 += v;
    for (uint32_t v : h2)
        if (v) {
            const double p = v / tot2;
            Hpair -= p * std::log2(p);
        }
    const double H1 = std::max(0.0, Hpair - H0);
    const double rep16 = m ? (double)dist[0] / m : 1.0;
    const double rep64 = m ? (double)dist[1] / m : 1.0;
    const double pos = std::max(1u, h.positions);
    for (int c = 0; c < kNumCand; c++) {
        {
            float* x = X + c * kIn;
            for (int k = 0; k < kIn; k++) x[k] = 0.f;
            const int W = cand_W(c), P = cand_P(c);
            int wi = 0;
            while (kGridW[wi] != W) wi++;
            const double dens = (h.trig[c] + 1.0) / (pos + 1.0);
            x[0] = (float)(std::log2((double)n) / 32.0);
            x[1] = (float)(std::log2(std::max(1.0, sigma)) / 8.0);
            x[2] = (float)(H0 / 8.0);
            x[3] = (float)(H1 / 8.0);
            x[4] = (float)rep16;
            x[5] = (float)rep64;
            x[6] = (float)std::log(rep16 + 1e-6);
            x[7] = (float)std::log(rep64 + 1e-6);
            x[8 + wi] = 1.f;
            x[12] = (float)(P / 100.0);
            x[13] = (float)(std::log((double)P) / 5.0);
            x[14] = (float)(1.0 / P);
            x[15] = (float)(W / 35.0);
            x[16] = (float)dens;
            x[17] = (float)(std::log(dens));
            x[18] = (float)(std::log(1.0 / dens + W) / 5.0);          // mean phrase length
            x[19] = (float)(std::log((double)n * dens + 1.0) / 32.0);  // parse length
            x[20] = (float)(rep64 / dens / 100.0);                     // distinct content x phrase length
            x[21] = (float)(std::log(rep64 + 1e-6) + std::log(1.0 / dens + W));
        }
    }
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

struct Model {
    bool loaded = false;
    float xmean[kIn], xstd[kIn], ymean[kOut], ystd[kOut];
    std::vector<float> params;  // fp32 master layout (weights bf16-representable)
};

static Model g_model;
static std::mutex g_model_mu;

static const uint32_t kMagic = 0x4D424C50;  // "PLBM"

bool model_save(const Model& M, const char* path)
{
    FILE* f = fopen(path, "wb");
    if (!f) return false;
    const uint32_t hdr[4] = {kMagic, (uint32_t)kIn, (uint32_t)kHid, (uint32_t)kOut};
    bool ok = fwrite(hdr, 4, 4, f) == 4;
    ok = ok && fwrite(M.xmean, 4, kIn, f) == (size_t)kIn && fwrite(M.xstd, 4, kIn, f) == (size_t)kIn;
    ok = ok && fwrite(M.ymean, 4, kOut, f) == (size_t)kOut && fwrite(M.ystd, 4, kOut, f) == (size_t)kOut;
    ok = ok && fwrite(M.params.data(), 4, kParams, f) == (size_t)kParams;
    return (fclose(f) == 0) && ok;
}

bool model_load(Model& M, const char* path)
{
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    uint32_t hdr[4];
    bool ok = fread(hdr, 4, 4, f) == 4 && hdr[0] == kMagic && hdr[1] == (uint32_t)kIn && hdr[2] == (uint32_t)kHid &&
              hdr[3] == (uint32_t)kOut;
    M.params.assign(kParams, 0.f);
    ok = ok && fread(M.xmean, 4, kIn, f) == (size_t)kIn && fread(M.xstd, 4, kIn, f) == (size_t)kIn;
    ok = ok && fread(M.ymean, 4, kOut, f) == (size_t)kOut && fread(M.ystd, 4, kOut, f) == (size_t)kOut;
    ok = ok && fread(M.params.data(), 4, kParams, f) == (size_t)kParams;
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

// Train on B samples (raw features / targets); writes the model file; returns the final
// training loss (normalised targets).
float mlp_train(Storage& S, const float* Xraw, const float* Yraw, uint32_t B, int epochs, uint32_t mb, float lr,
                uint32_t seed, const char* path)
{
    Model M;
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
    // shuffled, normalised copy
    std::vector<uint32_t> perm(B);
    for (uint32_t i = 0; i < B; i++) perm[i] = i;
    uint64_t st = seed * 0x2545F4914F6CDD1Dull + 1;
    for (uint32_t i = B; i > 1; i--) std::swap(perm[i - 1], perm[splitmix(st) % i]);
    std::vector<float> X((size_t)B * kIn), Y((size_t)B * kOut);
    for (uint32_t i = 0; i < B; i++) {
        for (int k = 0; k < kIn; k++) X[(size_t)i * kIn + k] = (Xraw[(size_t)perm[i] * kIn + k] - M.xmean[k]) / M.xstd[k];
        for (int k = 0; k < kOut; k++) Y[(size_t)i * kOut + k] = (Yraw[(size_t)perm[i] * kOut + k] - M.ymean[k]) / M.ystd[k];
    }
    // He initialisation, weights rounded to bf16
    M.params.assign(kParams, 0.f);
    auto init = [&](size_t off, int count, int fan_in) {
        const double sd = std::sqrt(2.0 / fan_in);
        for (int i = 0; i < count; i++) {
            const double u1 = (splitmix(st) >> 11) * (1.0 / 9007199254740992.0) + 1e-12;
            const double u2 = (splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
            M.params[off + i] = bf16_to_f32(f32_to_bf16((float)(z * sd)));
        }
    };
    init(0, kW1, kIn);
    init(kW1 + kHid, kW2, kHid);
    init(kW1 + kHid + kW2 + kHid, kW3, kHid);
    MlpBufs m = mlp_params(S, M.params.data());
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
    MIBWT_HIP(hipMemcpyAsync(M.params.data(), m.p, sizeof(float) * kParams, hipMemcpyDeviceToHost, S.stream));
    MIBWT_HIP(hipStreamSynchronize(S.stream));
    for (int i = 0; i < kParams; i++) M.params[i] = bf16_to_f32(f32_to_bf16(M.params[i]));
    M.loaded = true;
    if (path && !model_save(M, path)) throw Error(LIBCUBWT_BAD_PARAMETER, "cannot write the selector model");
    std::lock_guard<std::mutex> lk(g_model_mu);
    g_model = M;
    return loss;
}

bool selector_load(const char* path)
{
    Model M;
    if (!model_load(M, path)) return false;
    std::lock_guard<std::mutex> lk(g_model_mu);
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

// Choose (W,P) for the text resident at T (padded device copy); fixed W=5,P=16 without a model.
void selector_choose_device(Storage& S, const uint8_t* T, uint64_t n, int* W, int* P, float* pred_out)
{
    ensure_default_model();
    Model M;
    {
        std::lock_guard<std::mutex> lk(g_model_mu);
        M = g_model;
    }
    *W = 5;
    *P = 16;
    if (!M.loaded) return;
    std::vector<float> X((size_t)kNumCand * kIn);
    wp_features(S, T, n, X.data());
    for (int c = 0; c < kNumCand; c++)
        for (int k = 0; k < kIn; k++) X[(size_t)c * kIn + k] = (X[(size_t)c * kIn + k] - M.xmean[k]) / M.xstd[k];
    std::vector<float> Y((size_t)kNumCand * kOut);
    mlp_predict_raw(S, M.params.data(), X.data(), kNumCand, Y.data());
    int best = 0;
    float bt = 1e30f;
    for (int c = 0; c < kNumCand; c++) {
        const float t = Y[(size_t)c * kOut + 2] * M.ystd[2] + M.ymean[2];
        if (pred_out) pred_out[c] = t;
        if (t < bt) {
            bt = t;
            best = c;
        }
    }
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
