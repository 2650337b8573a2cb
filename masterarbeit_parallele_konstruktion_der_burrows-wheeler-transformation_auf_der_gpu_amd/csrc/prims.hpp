// Device-wide primitives written for CDNA (64-lane wavefronts): block/wave scans, a
// reduce-then-scan device scan, and a stable LSD radix sort whose in-tile ranking uses
// wave-level multi-split (8 ballots per item and digit) and LDS staging so that each digit
// run leaves the tile as one coalesced write.
#pragma once

#include "common.hpp"

namespace mibwt {

constexpr int kBlock = 256;  // 4 wavefronts
constexpr int kWaves = kBlock / 64;

struct OpSum {
    template <typename T>
    __device__ __host__ T operator()(T a, T b) const { return a + b; }
};
struct OpMax {
    template <typename T>
    __device__ __host__ T operator()(T a, T b) const { return a > b ? a : b; }
};

__device__ __forceinline__ uint64_t lanemask_lt()
{
    const int lane = threadIdx.x & 63;
    return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}

// Lane shuffle; aggregate value types overload it next to their definition.
template <typename T>
__device__ __forceinline__ T shfl_up_val(T x, int o)
{
    return __shfl_up(x, o, 64);
}

// Inclusive scan across one wavefront.
template <typename T, typename Op>
__device__ __forceinline__ T wave_inclusive_scan(T x, Op op)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T y = shfl_up_val(x, o);
        if (lane >= o) x = op(y, x);
    }
    return x;
}

// Exclusive scan of one value per thread over a kBlock-thread block.  `lds` needs kWaves
// entries.  Returns the exclusive prefix and writes the block total to *total.
template <typename T, typename Op>
__device__ __forceinline__ T block_exclusive_scan(T v, T identity, Op op, T* lds, T* total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    T inc = wave_inclusive_scan(v, op);
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    T wave_prefix = identity, tot = identity;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
        T s = lds[w];
        if (w < wid) wave_prefix = op(wave_prefix, s);
        tot = op(tot, s);
    }
    T ex = shfl_up_val(inc, 1);
    if (lane == 0) ex = identity;
    *total = tot;
    __syncthreads();
    return op(wave_prefix, ex);
}

// ----------------------------------------------------------------------------- device scan

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

template <typename T, typename Op>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const T* __restrict__ in, T* __restrict__ partial,
                                                       uint64_t n, T identity, Op op)
{
    __shared__ T lds[kWaves];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    T acc = identity;
#pragma unroll
    for (int i = 0; i < kScanItems; i++)
        if (base + i < n) acc = op(acc, in[base + i]);
    T tot;
    (void)block_exclusive_scan(acc, identity, op, lds, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

template <typename T, typename Op>
__global__ __launch_bounds__(kBlock) void k_scan_down(const T* __restrict__ in, T* __restrict__ out,
                                                     const T* __restrict__ partial_scanned,
                                                     uint64_t n, T identity, Op op,
                                                     int inclusive, T* __restrict__ total_out)
{
    __shared__ T lds[kWaves];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    T v[kScanItems];
    T acc = identity;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        v[i] = (base + i < n) ? in[base + i] : identity;
        acc = op(acc, v[i]);
    }
    T tot;
    T run = block_exclusive_scan(acc, identity, op, lds, &tot);
    T off = partial_scanned ? partial_scanned[blockIdx.x] : identity;
    run = op(off, run);
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        if (base + i < n) {
            T nxt = op(run, v[i]);
            out[base + i] = inclusive ? nxt : run;
            run = nxt;
        }
    }
    if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) {
        // the last thread of the last block holds the running value past the end
        *total_out = run;
    }
}

// Device-wide scan.  `out` may alias `in`.  If total_dev != nullptr the reduction of all
// n items is written there.  Scratch comes from buffers tmp_id, tmp_id+1, ... (recursion).
template <typename T, typename Op>
void device_scan(Storage& S, const T* in, T* out, uint64_t n, bool inclusive, T identity, Op op,
                 T* total_dev, int tmp_id = B_SCAN_TMP, int depth = 0)
{
    if (n == 0) {
        if (total_dev) MIBWT_HIP(hipMemcpyAsync(total_dev, &identity, sizeof(T), hipMemcpyHostToDevice, S.stream));
        return;
    }
    uint32_t tiles = ceil_div(n, kScanTile);
    if (tiles == 1) {
        k_scan_down<T, Op><<<1, kBlock, 0, S.stream>>>(in, out, nullptr, n, identity, op,
                                                        inclusive ? 1 : 0, total_dev);
        MIBWT_LAUNCH_CHECK();
        return;
    }
    MIBWT_REQUIRE(depth < 4, LIBCUBWT_BAD_PARAMETER, "scan too large");
    // partials live in a per-depth region of the scratch buffer
    T* partial = S.buf<T>(tmp_id + depth, tiles);
    k_scan_reduce<T, Op><<<tiles, kBlock, 0, S.stream>>>(in, partial, n, identity, op);
    MIBWT_LAUNCH_CHECK();
    device_scan<T, Op>(S, partial, partial, tiles, false, identity, op, nullptr, tmp_id, depth + 1);
    k_scan_down<T, Op><<<tiles, kBlock, 0, S.stream>>>(in, out, partial, n, identity, op,
                                                        inclusive ? 1 : 0, total_dev);
    MIBWT_LAUNCH_CHECK();
}

template <typename T>
void exclusive_sum(Storage& S, const T* in, T* out, uint64_t n, T* total_dev = nullptr)
{
    device_scan<T, OpSum>(S, in, out, n, false, T(0), OpSum{}, total_dev);
}
template <typename T>
void inclusive_sum(Storage& S, const T* in, T* out, uint64_t n, T* total_dev = nullptr)
{
    device_scan<T, OpSum>(S, in, out, n, true, T(0), OpSum{}, total_dev);
}
template <typename T>
void inclusive_max(Storage& S, const T* in, T* out, uint64_t n)
{
    device_scan<T, OpMax>(S, in, out, n, true, T(0), OpMax{}, nullptr);
}

// ----------------------------------------------------------------------------- radix sort

constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortItems = 16;                    // items per thread
constexpr int kSortTile = kBlock * kSortItems;    // 4096 items per tile
constexpr int kWaveTile = 64 * kSortItems;        // 1024 items per wavefront

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K key, int shift, uint32_t mask)
{
    return (uint32_t)(key >> shift) & mask;
}

template <typename K>
__global__ __launch_bounds__(kBlock) void k_radix_hist(const K* __restrict__ keys, uint64_t n, int shift,
                                                      uint32_t mask, uint32_t* __restrict__ hist,
                                                      uint32_t tiles)
{
    __shared__ uint32_t cnt[kWaves][kRadix];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kWaves * kRadix; i += kBlock) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile + (uint64_t)w * kWaveTile + lane;
#pragma unroll
    for (int i = 0; i < kSortItems; i++) {
        uint64_t idx = base + (uint64_t)i * 64;
        if (idx < n) atomicAdd(&cnt[w][digit_of(keys[idx], shift, mask)], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kRadix; d += kBlock) {
        uint32_t s = 0;
#pragma unroll
        for (int q = 0; q < kWaves; q++) s += cnt[q][d];
        hist[(uint64_t)d * tiles + blockIdx.x] = s;
    }
}

// Default emitter: writes the sorted key/value pair.
template <typename K, typename V>
struct EmitPair {
    static constexpr double bytes_per_item = sizeof(K) + sizeof(V);
    K* keys_out;
    V* vals_out;
    __device__ __forceinline__ void operator()(uint64_t pos, K key, V val) const
    {
        keys_out[pos] = key;
        if (vals_out) vals_out[pos] = val;
    }
};

template <typename K, typename V, typename Emit>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const K* __restrict__ keys, const V* __restrict__ vals,
                                                         uint64_t n, int shift, uint32_t mask,
                                                         const uint32_t* __restrict__ hist_scanned,
                                                         uint32_t tiles, Emit emit)
{
    __shared__ uint32_t wcnt[kWaves][kRadix];
    __shared__ uint32_t dstart[kRadix];
    __shared__ uint32_t dglob[kRadix];
    __shared__ uint32_t scan_lds[kWaves];
    __shared__ K skeys[kSortTile];
    __shared__ V svals[kSortTile];

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kWaves * kRadix; i += kBlock) (&wcnt[0][0])[i] = 0;
    __syncthreads();

    const uint64_t tile_base = (uint64_t)blockIdx.x * kSortTile;
    const uint64_t base = tile_base + (uint64_t)w * kWaveTile + lane;
    K k[kSortItems];
    V v[kSortItems];
    uint16_t rank[kSortItems];
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int i = 0; i < kSortItems; i++) {
        const uint64_t idx = base + (uint64_t)i * 64;
        const bool valid = idx < n;
        k[i] = valid ? keys[idx] : K(0);
        v[i] = (valid && vals) ? vals[idx] : V(0);
        const uint32_t d = digit_of(k[i], shift, mask);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t old = 0;
        const int leader = valid ? (__ffsll((unsigned long long)peers) - 1) : 0;
        if (valid && lane == leader) old = atomicAdd(&wcnt[w][d], (uint32_t)__popcll(peers));
        old = __shfl(old, leader, 64);
        rank[i] = valid ? (uint16_t)(old + (uint32_t)__popcll(peers & lt)) : (uint16_t)0;
    }
    __syncthreads();
    // per digit: wave prefixes, tile count, tile-local start, global start
    {
        const int d = threadIdx.x;  // kBlock == kRadix
        uint32_t run = 0;
#pragma unroll
        for (int q = 0; q < kWaves; q++) {
            uint32_t c = wcnt[q][d];
            wcnt[q][d] = run;
            run += c;
        }
        uint32_t tot;
        uint32_t st = block_exclusive_scan(run, 0u, OpSum{}, scan_lds, &tot);
        dstart[d] = st;
        dglob[d] = hist_scanned[(uint64_t)d * tiles + blockIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSortItems; i++) {
        const uint64_t idx = base + (uint64_t)i * 64;
        if (idx < n) {
            const uint32_t d = digit_of(k[i], shift, mask);
            const uint32_t p = dstart[d] + wcnt[w][d] + rank[i];
            skeys[p] = k[i];
            svals[p] = v[i];
        }
    }
    __syncthreads();
    const uint32_t valid_in_tile = (uint32_t)((n - tile_base) < (uint64_t)kSortTile ? (n - tile_base) : kSortTile);
    for (uint32_t p = threadIdx.x; p < valid_in_tile; p += kBlock) {
        const K key = skeys[p];
        const uint32_t d = digit_of(key, shift, mask);
        emit((uint64_t)dglob[d] + (p - dstart[d]), key, svals[p]);
    }
}

// Stable LSD radix sort of (key, value) pairs by key bits [lo, hi).  keys/vals hold the
// input in buffer 0 and are sorted in place (ping-ponging through buffer 1); on return the
// sorted data is in buffer 0 again.  vals0/vals1 may be nullptr (keys only).  If `last`
// is given, the final pass calls it instead of writing buffer 0 (and buffer 0 is left
// undefined).
template <typename K, typename V, typename LastEmit = EmitPair<K, V>>
void radix_sort(Storage& S, K* keys0, K* keys1, V* vals0, V* vals1, uint64_t n, int lo, int hi,
                const LastEmit* last = nullptr)
{
    if (n == 0 || hi <= lo) {
        if (last && n) {
            // nothing to sort: still route the data through the emitter with one trivial pass
            lo = 0;
            hi = 1;
        } else {
            return;
        }
    }
    MIBWT_REQUIRE(n < (1ull << 32), LIBCUBWT_BAD_PARAMETER, "radix_sort: too many items");
    const uint32_t tiles = ceil_div(n, kSortTile);
    uint32_t* hist = S.buf<uint32_t>(B_SORT_HIST, (size_t)kRadix * tiles);
    int passes = (hi - lo + kRadixBits - 1) / kRadixBits;
    K* kin = keys0;
    K* kout = keys1;
    V* vin = vals0;
    V* vout = vals1;
    for (int p = 0; p < passes; p++) {
        const int shift = lo + p * kRadixBits;
        const int nb = (hi - shift) < kRadixBits ? (hi - shift) : kRadixBits;
        const uint32_t mask = (1u << nb) - 1u;
        k_radix_hist<K><<<tiles, kBlock, 0, S.stream>>>(kin, n, shift, mask, hist, tiles);
        MIBWT_LAUNCH_CHECK();
        exclusive_sum<uint32_t>(S, hist, hist, (uint64_t)kRadix * tiles);
        const bool is_last = (p == passes - 1);
        const double vbytes = vin ? (double)sizeof(V) : 0.0;
        if (is_last && last) {
            const int kt = S.kt_begin(1, (double)n * (sizeof(K) + vbytes) + LastEmit::bytes_per_item * (double)n);
            k_radix_scatter<K, V, LastEmit><<<tiles, kBlock, 0, S.stream>>>(kin, vin, n, shift, mask,
                                                                           hist, tiles, *last);
            S.kt_end(kt);
        } else {
            const int kt = S.kt_begin(1, 2.0 * (double)n * (sizeof(K) + vbytes));
            EmitPair<K, V> e{kout, vout};
            k_radix_scatter<K, V, EmitPair<K, V>><<<tiles, kBlock, 0, S.stream>>>(kin, vin, n, shift,
                                                                                mask, hist, tiles, e);
            S.kt_end(kt);
        }
        MIBWT_LAUNCH_CHECK();
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (!last && kin != keys0) {
        MIBWT_HIP(hipMemcpyAsync(keys0, kin, sizeof(K) * n, hipMemcpyDeviceToDevice, S.stream));
        if (vals0 && vin) MIBWT_HIP(hipMemcpyAsync(vals0, vin, sizeof(V) * n, hipMemcpyDeviceToDevice, S.stream));
    }
}

// ----------------------------------------------------------------------------- small kernels

template <typename T>
__global__ void k_fill(T* p, uint64_t n, T v)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename T>
void fill(Storage& S, T* p, uint64_t n, T v)
{
    if (!n) return;
    k_fill<T><<<ceil_div(n, kBlock), kBlock, 0, S.stream>>>(p, n, v);
    MIBWT_LAUNCH_CHECK();
}

inline uint32_t grid_for(uint64_t n) { return ceil_div(n ? n : 1, kBlock); }

}  // namespace mibwt
