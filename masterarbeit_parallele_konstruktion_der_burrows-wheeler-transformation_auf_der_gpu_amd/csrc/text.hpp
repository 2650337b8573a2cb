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

// symbol presence bitmap of the text (256 bits): each thread ORs its 64 bytes into a
// register bitmap, the wave reduces it with lane shuffles, one atomic per wave and word
static __global__ __launch_bounds__(kBlock) void k_alpha_used(const uint8_t* __restrict__ T, uint64_t n,
                                                             uint32_t* __restrict__ used)
{
    uint32_t bm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint64_t base = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) * 64;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint64_t p = base + q * 16;
        if (p >= n) break;
        const uint4 v = *reinterpret_cast<const uint4*>(T + p);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
        const int cnt = (n - p) < 16 ? (int)(n - p) : 16;
        for (int i = 0; i < cnt; i++) {
            const uint32_t c = (w[i >> 2] >> (8 * (i & 3))) & 0xFF;
            bm[c >> 5] |= 1u << (c & 31);
        }
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
        uint32_t x = bm[k];
        for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o, 64);
        bm[k] = x;
    }
    if ((threadIdx.x & 63) < 8) {
        const int k = threadIdx.x & 63;
        uint32_t x = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) x = (j == k) ? bm[j] : x;
        if (x) atomicOr(&used[k], x);
    }
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
    k_alpha_used<<<ceil_div(ceil_div(n, 64), kBlock), kBlock, 0, S.stream>>>(T, n, codes + 260);
    MIBWT_LAUNCH_CHECK();
    k_alpha_codes<<<1, 256, 0, S.stream>>>(codes + 260, codes);
    MIBWT_LAUNCH_CHECK();
}

// Private padded copy of the text in B_TEXT: T[0..n) followed by zero bytes up to a 4096-byte
// multiple plus kPad (tile loads and unaligned word loads stay inside the allocation).
inline uint64_t staged_text_bytes(uint64_t n) { return ((n + 1 + 4095) / 4096) * 4096 + kPad + 4096; }

inline uint8_t* stage_text(Storage& S, const uint8_t* src, bool src_host, uint64_t n)
{
    const uint64_t bytes = staged_text_bytes(n);
    uint8_t* T = S.buf<uint8_t>(B_TEXT, bytes);
    MIBWT_HIP(hipMemsetAsync(T + n, 0, bytes - n, S.stream));
    MIBWT_HIP(hipMemcpyAsync(T, src, n, src_host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, S.stream));
    return T;
}

}  // namespace mibwt
