// Text access shared by the BWT pipelines: the private padded text copy, unaligned 8-byte
// loads, a 64-bit mixer and the symbol-presence / dense-code kernels.
#pragma once

#include "prims.hpp"

namespace mibwt {

constexpr int kPad = 64;  // zero bytes after the text copy

// 8 bytes of the padded text starting at any position p (little endian).
__device__ __forceinline__ uint64_t load8(const uint8_t* __restrict__ T, uint64_t p)
{
    const uint64_t a = p & ~7ull;
    const uint64_t* w = reinterpret_cast<const uint64_t*>(T + a);
    const uint64_t w0 = w[0];
    const int sh = (int)(p & 7) * 8;
    if (sh == 0) return w0;
    const uint64_t w1 = w[1];
    return (w0 >> sh) | (w1 << (64 - sh));
}

__device__ __forceinline__ uint64_t mix64(uint64_t h)
{
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}

// Symbol presence bitmaps (256 bits as 8 words).  A thread keeps its bitmap in registers
// (one compare and select per word, no dynamic register indexing), the wavefront ORs the
// bitmaps with lane shuffles, and each block writes one partial bitmap (no global atomics:
// same-address device atomics serialize); k_alpha_merge ORs the partials.
// Presence flags of the byte values in LDS: a byte is looked up first and written only when
// not yet seen (reads of one address broadcast; concurrent writes all store 1).  Sixteen
// lookups are issued before any write.
__device__ __forceinline__ void seen_add16(uint8_t* seen, const uint4& v, uint32_t valid /* bit i: byte i */)
{
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    uint32_t f[16];
#pragma unroll
    for (int i = 0; i < 16; i++) f[i] = seen[(w[i >> 2] >> (8 * (i & 3))) & 0xFFu];
#pragma unroll
    for (int i = 0; i < 16; i++)
        if (((valid >> i) & 1u) && !f[i]) seen[(w[i >> 2] >> (8 * (i & 3))) & 0xFFu] = 1;
}

// the block's flags as a 256-bit map: part[block][8]
__device__ __forceinline__ void seen_block_store(const uint8_t* seen, uint32_t* __restrict__ part)
{
    __syncthreads();
    const uint64_t b = __ballot(seen[threadIdx.x] != 0);  // kBlock == 256: wave w holds bytes 64w..64w+63
    if ((threadIdx.x & 63) == 0) {
        const uint32_t w = threadIdx.x >> 6;
        part[(uint64_t)blockIdx.x * 8 + 2 * w] = (uint32_t)b;
        part[(uint64_t)blockIdx.x * 8 + 2 * w + 1] = (uint32_t)(b >> 32);
    }
}

// bytes of T[0..n), 64 per thread
static __global__ __launch_bounds__(kBlock) void k_alpha_text(const uint8_t* __restrict__ T, uint64_t n,
                                                             uint32_t* __restrict__ part)
{
    static_assert(kBlock == 256, "one flag per thread");
    __shared__ uint8_t seen[256];
    seen[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) * 64;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint64_t p = base + q * 16;
        if (p >= n) break;
        const uint4 v = *reinterpret_cast<const uint4*>(T + p);
        const uint32_t cnt = (n - p) < 16 ? (uint32_t)(n - p) : 16u;
        seen_add16(seen, v, cnt >= 16 ? 0xFFFFu : ((1u << cnt) - 1u));
    }
    seen_block_store(seen, part);
}

// bytes of the ranges T[start[u] .. start[u] + len[u]) clipped to n, one range per thread,
// read as aligned 16-byte words, four per step issued together (T is the padded copy)
static __global__ __launch_bounds__(kBlock) void k_alpha_ranges(const uint8_t* __restrict__ T, uint64_t n,
                                                               const uint32_t* __restrict__ start,
                                                               const uint32_t* __restrict__ len, uint32_t m,
                                                               uint32_t* __restrict__ part)
{
    __shared__ uint8_t seen[256];
    seen[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < m) {
        const uint64_t s = start[u];
        const uint64_t e = (s + len[u] < n) ? s + len[u] : n;
        for (uint64_t p = s & ~15ull; p < e; p += 64) {
            uint4 v[4];
#pragma unroll
            for (int j = 0; j < 4; j++) v[j] = *reinterpret_cast<const uint4*>(T + p + 16 * j);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                // bytes q = p + 16 j + i with s <= q < e
                const uint64_t q0 = p + 16 * j;
                uint32_t valid = 0xFFFFu;
                if (q0 < s) valid &= 0xFFFFu << (uint32_t)(s - q0);
                if (q0 + 16 > e) valid &= (q0 >= e) ? 0u : ((1u << (uint32_t)(e - q0)) - 1u);
                seen_add16(seen, v[j], valid);
            }
        }
    }
    seen_block_store(seen, part);
}

static __global__ __launch_bounds__(kBlock) void k_alpha_merge(const uint32_t* __restrict__ part, uint32_t blocks,
                                                              uint32_t* __restrict__ used)
{
    __shared__ uint32_t acc[8];
    if (threadIdx.x < 8) acc[threadIdx.x] = 0;
    __syncthreads();
    uint32_t bm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < blocks; b += kBlock)
#pragma unroll
        for (int k = 0; k < 8; k++) bm[k] |= part[(uint64_t)b * 8 + k];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        uint32_t x = bm[k];
        for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o, 64);
        if ((threadIdx.x & 63) == 0 && x) atomicOr(&acc[k], x);
    }
    __syncthreads();
    if (threadIdx.x < 8) used[threadIdx.x] = acc[threadIdx.x];
}

// dense symbol codes 2.. in byte order; codes[256] = number of distinct bytes
static __global__ void k_alpha_codes(const uint32_t* __restrict__ used, uint32_t* __restrict__ codes)
{
    const uint32_t c = threadIdx.x;  // 256 threads
    uint32_t below = 0;
    for (uint32_t w = 0; w < (c >> 5); w++) below += __popc(used[w]);
    below += __popc(used[c >> 5] & ((1u << (c & 31)) - 1u));
    codes[c] = below + 2;
    if (c == 255) codes[256] = below + ((used[7] >> 31) & 1u);
}

// Dense codes of the bytes that occur in T[0..n) (or, with ranges, in those ranges):
// codes[c] for c < 256, codes[256] = sigma.  codes[260..268) holds the presence bitmap.
inline void alphabet_codes(Storage& S, const uint8_t* T, uint64_t n, uint32_t* codes,
                           const uint32_t* start = nullptr, const uint32_t* len = nullptr, uint32_t m = 0)
{
    const uint32_t blocks = start ? ceil_div(m ? m : 1, kBlock) : ceil_div(ceil_div(n, 64), kBlock);
    uint32_t* part = S.buf<uint32_t>(B_ALPHA_PART, (size_t)blocks * 8);
    if (start) k_alpha_ranges<<<blocks, kBlock, 0, S.stream>>>(T, n, start, len, m, part);
    else k_alpha_text<<<blocks, kBlock, 0, S.stream>>>(T, n, part);
    MIBWT_LAUNCH_CHECK();
    k_alpha_merge<<<1, kBlock, 0, S.stream>>>(part, blocks, codes + 260);
    MIBWT_LAUNCH_CHECK();
    k_alpha_codes<<<1, 256, 0, S.stream>>>(codes + 260, codes);
    MIBWT_LAUNCH_CHECK();
}

// Private padded copy of the text in B_TEXT: T[0..n) followed by zero bytes to the next
// 8192-byte multiple at least 64 bytes past n (the trigger scan's staged length), then 8192
// more (tile loads and unaligned word loads stay inside the allocation).
inline uint64_t staged_text_bytes(uint64_t n) { return ((n + 64 + 8191) / 8192) * 8192 + 8192; }

inline uint8_t* stage_text(Storage& S, const uint8_t* src, bool src_host, uint64_t n)
{
    const uint64_t bytes = staged_text_bytes(n);
    uint8_t* T = S.buf<uint8_t>(B_TEXT, bytes);
    MIBWT_HIP(hipMemsetAsync(T + n, 0, bytes - n, S.stream));
    MIBWT_HIP(hipMemcpyAsync(T, src, n, src_host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, S.stream));
    return T;
}

}  // namespace mibwt
