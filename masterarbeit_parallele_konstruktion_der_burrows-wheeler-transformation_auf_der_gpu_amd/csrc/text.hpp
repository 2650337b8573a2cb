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

// symbol presence bitmap of the text (256 bits)
static __global__ __launch_bounds__(kBlock) void k_alpha_used(const uint8_t* __restrict__ T, uint64_t n,
                                                             uint32_t* __restrict__ used)
{
    __shared__ uint32_t bm[8];
    if (threadIdx.x < 8) bm[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) * 16;
    if (base < n) {
        uint4 v = *reinterpret_cast<const uint4*>(T + base);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
        const int cnt = (n - base) < 16 ? (int)(n - base) : 16;
        for (int i = 0; i < cnt; i++) {
            const uint32_t c = (w[i >> 2] >> (8 * (i & 3))) & 0xFF;
            atomicOr(&bm[c >> 5], 1u << (c & 31));
        }
    }
    __syncthreads();
    if (threadIdx.x < 8 && bm[threadIdx.x]) atomicOr(&used[threadIdx.x], bm[threadIdx.x]);
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

// Dense codes of the bytes that occur in T[0..n): codes[c] for c < 256, codes[256] = sigma.
// Uses codes[260..268) as scratch.  Returns nothing; read codes[256] for sigma.
inline void alphabet_codes(Storage& S, const uint8_t* T, uint64_t n, uint32_t* codes)
{
    fill<uint32_t>(S, codes + 260, 8, 0u);
    k_alpha_used<<<ceil_div(ceil_div(n, 16), kBlock), kBlock, 0, S.stream>>>(T, n, codes + 260);
    MIBWT_LAUNCH_CHECK();
    k_alpha_codes<<<1, 256, 0, S.stream>>>(codes + 260, codes);
    MIBWT_LAUNCH_CHECK();
}

}  // namespace mibwt
