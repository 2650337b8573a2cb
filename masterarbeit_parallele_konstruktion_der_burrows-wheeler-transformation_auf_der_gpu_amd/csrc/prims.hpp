// Device-wide primitives written for CDNA (64-lane wavefronts): block/wave scans, a
// reduce-then-scan device scan, and a stable LSD radix sort whose in-tile ranking uses
// wave-level multi-split (8 ballots per item and digit) and LDS staging so that each digit
// run leaves the tile as one coalesced write.
#pragma once

#include <type_traits>

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

// ----------------------------------------------------------------------------- chained scan

// Ticket of this tile in launch order (all threads of the block call it).
__device__ __forceinline__ uint32_t chain_ticket(const Chain& c, uint32_t tiles, uint32_t* lds)
{
    if (threadIdx.x == 0) {
        const uint32_t t = atomicAdd(c.ticket, 1u);
        // every other tile of this launch has drawn its ticket: reset for the next launch
        if (t == tiles - 1) __hip_atomic_store(c.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *lds = t;
    }
    __syncthreads();
    const uint32_t t = *lds;
    __syncthreads();
    return t;
}

__device__ __forceinline__ void chain_publish(const Chain& c, uint32_t tile, uint32_t state, uint32_t v)
{
    const uint64_t w = ((uint64_t)c.epoch << 34) | ((uint64_t)state << 32) | v;
    __hip_atomic_store(c.status + tile, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of `tile` from its predecessors, read 64 tiles at a time by one wavefront
// (all 64 lanes call it).  Tiles with earlier tickets are running or done, so each wait ends.
template <typename Op>
__device__ uint32_t chain_lookback(const Chain& c, uint32_t tile, Op op, uint32_t identity)
{
    const int lane = threadIdx.x & 63;
    uint32_t excl = identity;
    int64_t j = (int64_t)tile - 1;
    for (;;) {
        const int64_t k = j - lane;
        uint32_t state = 2, v = identity;
        if (k >= 0) {
            uint64_t w = __hip_atomic_load(c.status + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            while ((uint32_t)(w >> 34) != c.epoch) {
                __builtin_amdgcn_s_sleep(1);
                w = __hip_atomic_load(c.status + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            state = (uint32_t)(w >> 32) & 3u;
            v = (uint32_t)w;
        }
        const uint64_t incl = __ballot(state == 2);
        // nearest inclusive prefix: lanes 0..stop contribute
        const int stop = incl ? (__ffsll((unsigned long long)incl) - 1) : 63;
        uint32_t x = (lane <= stop) ? v : identity;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x = op(x, (uint32_t)__shfl_xor(x, o, 64));
        excl = op(excl, x);
        if (incl) break;
        j -= 64;
    }
    return excl;
}

// Exclusive prefix of this tile given its total; every thread of the block calls it and gets
// the prefix.  `lds` is one word of shared memory.
template <typename Op>
__device__ __forceinline__ uint32_t chain_prefix(const Chain& c, uint32_t tile, uint32_t total, Op op,
                                                 uint32_t identity, uint32_t* lds)
{
    if (threadIdx.x < 64) {
        uint32_t excl = identity;
        if (tile == 0) {
            if (threadIdx.x == 0) chain_publish(c, 0, 2, total);
        } else {
            if (threadIdx.x == 0) chain_publish(c, tile, 1, total);
            excl = chain_lookback(c, tile, op, identity);
            if (threadIdx.x == 0) chain_publish(c, tile, 2, op(excl, total));
        }
        if (threadIdx.x == 0) *lds = excl;
    }
    __syncthreads();
    const uint32_t r = *lds;
    __syncthreads();
    return r;
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

// Single-pass scan of 32-bit values: each tile scans its items, takes its prefix from its
// predecessors by decoupled look-back and writes its outputs, one read and one write per item.
// `out` may alias `in` (every thread reads its items before writing them).
// Chained tiles are large (kChainItems per thread, 16384 per tile): every tile draws one
// ticket from a single device-wide counter, and such atomics serialize at roughly 11 ns each
// on MI355X, so a tile must carry far more than 11 ns of memory traffic.
// The items per thread are a launch shape (Storage::tune.chain_items: 64 by default, 16 for
// short arrays, whose tiles would otherwise be too few to fill the chip).
// The items come through `in`: in(i) is item i, in.load4(i) items i..i+3 (i a multiple of 4).
// ScanArray reads a plain array; other inputs compute the items on the fly (a flag of each
// element, say), which saves materialising them.
struct ScanArray {
    const uint32_t* p;
    __device__ __forceinline__ uint32_t operator()(uint64_t i) const { return p[i]; }
    __device__ __forceinline__ uint4 load4(uint64_t i) const { return *reinterpret_cast<const uint4*>(p + i); }
};

// Byte items (flags): a quarter of the bytes of ScanArray to read (and to clear beforehand).
struct ScanBytes {
    const uint8_t* p;
    __device__ __forceinline__ uint32_t operator()(uint64_t i) const { return p[i]; }
    __device__ __forceinline__ uint4 load4(uint64_t i) const
    {
        const uint32_t w = *reinterpret_cast<const uint32_t*>(p + i);
        return make_uint4(w & 0xFFu, (w >> 8) & 0xFFu, (w >> 16) & 0xFFu, w >> 24);
    }
};

template <typename Op, bool kVec, int kChainItems, typename In = ScanArray>
__global__ __launch_bounds__(kBlock) void k_scan_chain(In in, uint32_t* out, uint64_t n,
                                                      uint32_t identity, Op op, int inclusive,
                                                      uint32_t* __restrict__ total_out, Chain c, uint32_t tiles)
{
    constexpr int kChainTile = kBlock * kChainItems;
    __shared__ uint32_t slot;
    __shared__ uint32_t lds[kWaves];
    const uint32_t tile = chain_ticket(c, tiles, &slot);
    const uint64_t base = (uint64_t)tile * kChainTile + (uint64_t)threadIdx.x * kChainItems;
    uint32_t v[kChainItems];
    const bool full = base + kChainItems <= n;
    if (kVec && full) {
#pragma unroll
        for (int q = 0; q < kChainItems / 4; q++) {
            const uint4 a = in.load4(base + 4 * q);
            v[4 * q] = a.x;
            v[4 * q + 1] = a.y;
            v[4 * q + 2] = a.z;
            v[4 * q + 3] = a.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < kChainItems; i++) v[i] = (base + i < n) ? in(base + i) : identity;
    }
    uint32_t acc = identity;
#pragma unroll
    for (int i = 0; i < kChainItems; i++) acc = op(acc, v[i]);
    uint32_t tot;
    uint32_t run = block_exclusive_scan(acc, identity, op, lds, &tot);
    const uint32_t pre = chain_prefix(c, tile, tot, op, identity, &slot);
    run = op(pre, run);
#pragma unroll
    for (int i = 0; i < kChainItems; i++) {
        const uint32_t nxt = op(run, v[i]);
        v[i] = inclusive ? nxt : run;
        run = nxt;
    }
    if (kVec && full) {
#pragma unroll
        for (int q = 0; q < kChainItems / 4; q++)
            *reinterpret_cast<uint4*>(out + base + 4 * q) = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < kChainItems; i++)
            if (base + i < n) out[base + i] = v[i];
    }
    if (total_out && tile == tiles - 1 && threadIdx.x == 0) *total_out = op(pre, tot);
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
    if constexpr (std::is_same<T, uint32_t>::value) {
        const int items = (S.tune.chain_items == 16) ? 16 : 64;
        tiles = ceil_div(n, (uint64_t)kBlock * items);
        const Chain c = S.chain(tiles);
        const uint32_t* in32 = reinterpret_cast<const uint32_t*>(in);
        uint32_t* out32 = reinterpret_cast<uint32_t*>(out);
        uint32_t* tot32 = reinterpret_cast<uint32_t*>(total_dev);
        const bool vec = ((uintptr_t)in % 16 == 0) && ((uintptr_t)out % 16 == 0);
        const int inc = inclusive ? 1 : 0;
        const uint32_t id = (uint32_t)identity;
        const ScanArray a{in32};
        if (items == 16) {
            if (vec) k_scan_chain<Op, true, 16><<<tiles, kBlock, 0, S.stream>>>(a, out32, n, id, op, inc, tot32, c, tiles);
            else k_scan_chain<Op, false, 16><<<tiles, kBlock, 0, S.stream>>>(a, out32, n, id, op, inc, tot32, c, tiles);
        } else {
            if (vec) k_scan_chain<Op, true, 64><<<tiles, kBlock, 0, S.stream>>>(a, out32, n, id, op, inc, tot32, c, tiles);
            else k_scan_chain<Op, false, 64><<<tiles, kBlock, 0, S.stream>>>(a, out32, n, id, op, inc, tot32, c, tiles);
        }
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

// The scan input "sorted key i starts a run of equal keys" (K of 4 or 8 bytes).
template <typename K>
struct ScanHeads {
    const K* key;
    K mask = (K)~(K)0;  // key bits that count (higher bits carry data along with the key)
    __device__ __forceinline__ bool diff(K a, K b) const { return ((a ^ b) & mask) != 0; }
    __device__ __forceinline__ uint32_t operator()(uint64_t i) const
    {
        return (i == 0 || diff(key[i], key[i - 1])) ? 1u : 0u;
    }
    __device__ __forceinline__ uint4 load4(uint64_t i) const
    {
        const K a = key[i], b = key[i + 1], c = key[i + 2], d = key[i + 3];
        const uint32_t h0 = (i == 0 || diff(a, key[i - 1])) ? 1u : 0u;
        return make_uint4(h0, diff(b, a), diff(c, b), diff(d, c));
    }
};

template <typename T>
void fill(Storage& S, T* p, uint64_t n, T v);

// Sum scan of n items computed by `in` (a ScanArray-like input whose load4 reads 16-byte aligned
// words when `aligned`): out[i] = the sum of the items before i (inclusive: up to i), and the total
// into total_dev.  One chained launch for any n.
template <typename In>
void sum_scan_of(Storage& S, In in, bool aligned, uint32_t* out, uint64_t n, bool inclusive, uint32_t* total_dev)
{
    if (n == 0) {
        if (total_dev) fill<uint32_t>(S, total_dev, 1, 0u);
        return;
    }
    const int items = (S.tune.chain_items == 16) ? 16 : 64;
    const uint32_t tiles = ceil_div(n, (uint64_t)kBlock * items);
    const Chain c = S.chain(tiles);
    const bool vec = aligned && ((uintptr_t)out % 16 == 0);
    const int inc = inclusive ? 1 : 0;
    if (items == 16) {
        if (vec) k_scan_chain<OpSum, true, 16, In><<<tiles, kBlock, 0, S.stream>>>(in, out, n, 0u, OpSum{}, inc, total_dev, c, tiles);
        else k_scan_chain<OpSum, false, 16, In><<<tiles, kBlock, 0, S.stream>>>(in, out, n, 0u, OpSum{}, inc, total_dev, c, tiles);
    } else {
        if (vec) k_scan_chain<OpSum, true, 64, In><<<tiles, kBlock, 0, S.stream>>>(in, out, n, 0u, OpSum{}, inc, total_dev, c, tiles);
        else k_scan_chain<OpSum, false, 64, In><<<tiles, kBlock, 0, S.stream>>>(in, out, n, 0u, OpSum{}, inc, total_dev, c, tiles);
    }
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
//
// Stable LSD radix sort, 8-bit digits, three kernels per digit pass:
//   - k_radix_up counts the digits of every tile (4096 items of 8-byte keys, 8192 of 4-byte
//     keys), digit-major, so that one flat exclusive scan (k_scan_chain) of the counts gives
//     every (digit, tile) its output start;
//   - k_radix_pass ranks its tile with wavefront multi-split (8 ballots per item) and scatters
//     keys, then values, through LDS so that every digit run leaves the tile as one contiguous
//     write.
// The tiles need no look-back: a single kernel per pass whose tiles learn the counts of earlier
// tiles by decoupled look-back spent about half of its time waiting for them (device-coherent
// status reads cross the XCDs), while the extra read of the keys by k_radix_up costs a third of
// a pass's algorithmic bytes.

constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortItems = 16;                    // items per thread (8-byte keys)
constexpr int kMaxPasses = 8;

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K key, int shift, uint32_t mask)
{
    return (uint32_t)(key >> shift) & mask;
}

// Default emitter: writes the sorted key/value pair.
template <typename K, typename V>
struct EmitPair {
    static constexpr double bytes_per_item = sizeof(K) + sizeof(V);
    static constexpr bool kValues = true;
    K* keys_out;
    V* vals_out;
    __device__ __forceinline__ void key(uint64_t pos, K k) const { keys_out[pos] = k; }
    __device__ __forceinline__ void val(uint64_t pos, V v) const { vals_out[pos] = v; }
};

// Keys per thread: 16 for 8-byte and 4-byte keys.  (32 for 4-byte keys, the same 32 KiB of
// staging, needs over 190 VGPRs: two wavefronts per SIMD, and its passes ran 2-3x slower per key
// than the 8-byte ones.)
template <typename K>
constexpr int sort_items() { return kSortItems; }

// Tile of block b in launches without tickets.  The blocks of a launch are dealt to the 8 XCDs
// in turn (block b to XCD b mod 8); XCD x is given the contiguous tiles
// [x q + min(x, r), ...) with q = tiles / 8, r = tiles mod 8, so that neighbouring tiles, whose
// digit runs meet in the same output lines, run behind one L2.  (Any other dispatch order only
// changes which XCD runs a tile.)
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t tiles)
{
    constexpr uint32_t kXcd = 8;
    const uint32_t x = b % kXcd, j = b / kXcd, q = tiles / kXcd, r = tiles % kXcd;
    return x * q + (x < r ? x : r) + j;
}

// Digit pass, first half: per tile the count of every digit, stored digit-major
// (cnt[d * tiles + tile]) so that one flat exclusive scan of cnt gives every (digit, tile) the
// output position of its first item.  A wavefront whose valid items share one digit adds them
// with one LDS atomic (presorted high digits), otherwise every item adds itself.
template <typename K, int Items>
__global__ __launch_bounds__(kBlock) void k_radix_up(const K* __restrict__ keys, uint64_t n, int shift, uint32_t mask,
                                                    uint32_t tiles, uint32_t* __restrict__ cnt)
{
    constexpr int Tile = kBlock * Items, WaveTile = 64 * Items;
    __shared__ uint32_t h[kWaves][kRadix];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kWaves * kRadix; i += kBlock) (&h[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = xcd_tile(blockIdx.x, tiles);
    const uint64_t base = (uint64_t)tile * Tile + (uint64_t)w * WaveTile + lane;
    K k[Items];
#pragma unroll
    for (int i = 0; i < Items; i++) {
        const uint64_t idx = base + (uint64_t)i * 64;
        k[i] = (idx < n) ? keys[idx] : K(0);
    }
#pragma unroll
    for (int i = 0; i < Items; i++) {
        const bool valid = base + (uint64_t)i * 64 < n;
        const uint32_t d = digit_of(k[i], shift, mask);
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
        const uint64_t vm = __ballot(valid);
        const uint64_t other = __ballot(valid && d != d0);
        if (other == 0) {
            if (lane == 0 && vm) atomicAdd(&h[w][d0], (uint32_t)__popcll(vm));
        } else if (valid) {
            atomicAdd(&h[w][d], 1u);
        }
    }
    __syncthreads();
    const uint32_t d = threadIdx.x;  // kBlock == kRadix
    cnt[(uint64_t)d * tiles + tile] = h[0][d] + h[1][d] + h[2][d] + h[3][d];
}

// Digit pass, second half: `start` holds the scanned tile counts of k_radix_up
// (start[d * tiles + tile] = output position of the tile's first item with digit d).
template <typename K, typename V, typename Emit, int Items = sort_items<K>()>
__global__ __launch_bounds__(kBlock) void k_radix_pass(const K* __restrict__ keys, const V* __restrict__ vals,
                                                      uint64_t n, int shift, uint32_t mask,
                                                      const uint32_t* __restrict__ start, uint32_t tiles, Emit emit)
{
    static_assert(sizeof(V) <= sizeof(K), "values are staged in the key buffer");
    static_assert(Items % 2 == 0, "slots are packed in pairs");
    constexpr int Tile = kBlock * Items, WaveTile = 64 * Items;
    static_assert(Tile <= 65536, "16-bit slots");
    // per-wave digit counters; once the slots are known the same bytes hold the digit of
    // every slot (4 KiB of counters, Tile bytes of digits), next to Tile keys of staging:
    // 38 KiB of LDS for 8-byte keys (4 blocks per CU)
    constexpr int kCntWords = kWaves * kRadix, kDigWords = Tile / 4;
    __shared__ uint32_t wcnt_dig[kCntWords > kDigWords ? kCntWords : kDigWords];
    uint32_t(*wcnt)[kRadix] = reinterpret_cast<uint32_t(*)[kRadix]>(wcnt_dig);
    uint8_t* sdig = reinterpret_cast<uint8_t*>(wcnt_dig);
    __shared__ uint32_t dstart[kRadix];
    __shared__ uint32_t doff[kRadix];  // output position - tile slot, per digit (mod 2^32)
    __shared__ uint32_t scan_lds[kWaves];
    __shared__ K stage[Tile];

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kWaves * kRadix; i += kBlock) (&wcnt[0][0])[i] = 0;
    const uint32_t tile = xcd_tile(blockIdx.x, tiles);
    __syncthreads();

    const uint64_t tile_base = (uint64_t)tile * Tile;
    const uint64_t base = tile_base + (uint64_t)w * WaveTile + lane;
    K k[Items];
    uint32_t rk[Items / 2];  // two 16-bit ranks (later: tile slots) per register
#pragma unroll
    for (int i = 0; i < Items; i++) {
        const uint64_t idx = base + (uint64_t)i * 64;
        const bool valid = idx < n;
        k[i] = valid ? keys[idx] : K(0);
    }
#pragma unroll
    for (int i = 0; i < Items; i++) {
        const bool valid = base + (uint64_t)i * 64 < n;
        const uint32_t d = digit_of(k[i], shift, mask);
        // lanes with the same digit: per bit, the lanes that agree with this lane's bit.  The
        // bit is sign-extended to a 0 / all-ones word so that each half of the mask is one
        // three-input boolean op: peers &= ~(ballot ^ bitmask).
        const uint64_t vm = __ballot(valid);
        uint32_t plo = (uint32_t)vm, phi = (uint32_t)(vm >> 32);
#pragma unroll
        for (int b = 0; b < kRadixBits; b++) {
            const uint32_t mk = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
            const uint64_t m = __ballot(mk != 0u);
            plo &= ~((uint32_t)m ^ mk);
            phi &= ~((uint32_t)(m >> 32) ^ mk);
        }
        uint32_t old = 0;
        const int leader = plo ? (__ffs(plo) - 1) : (31 + __ffs(phi));
        if (valid && lane == leader) old = atomicAdd(&wcnt[w][d], (uint32_t)(__popc(plo) + __popc(phi)));
        old = __shfl(old, valid ? leader : 0, 64);
        const uint32_t below = __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));
        const uint32_t r = valid ? (old + below) : 0u;
        if (i & 1) rk[i >> 1] |= r << 16;
        else rk[i >> 1] = r;
    }
    __syncthreads();
    {
        const int d = threadIdx.x;  // kBlock == kRadix
        uint32_t run = 0;
#pragma unroll
        for (int q = 0; q < kWaves; q++) {
            const uint32_t cq = wcnt[q][d];
            wcnt[q][d] = run;
            run += cq;
        }
        uint32_t tot;
        const uint32_t ds = block_exclusive_scan(run, 0u, OpSum{}, scan_lds, &tot);
        dstart[d] = ds;
        doff[d] = start[(uint64_t)d * tiles + tile] - ds;
    }
    __syncthreads();
    // ranks become the items' slots in the tile
#pragma unroll
    for (int i = 0; i < Items; i++) {
        const uint64_t idx = base + (uint64_t)i * 64;
        if (idx < n) {
            const uint32_t d = digit_of(k[i], shift, mask);
            const uint32_t r = (i & 1) ? (rk[i >> 1] >> 16) : (rk[i >> 1] & 0xFFFFu);
            const uint32_t p = dstart[d] + wcnt[w][d] + r;
            rk[i >> 1] = (i & 1) ? ((rk[i >> 1] & 0xFFFFu) | (p << 16)) : ((rk[i >> 1] & 0xFFFF0000u) | p);
        }
    }
    __syncthreads();  // the counters are dead from here on: their bytes take the slot digits
    // keys go through LDS in slot order
#pragma unroll
    for (int i = 0; i < Items; i++) {
        if (base + (uint64_t)i * 64 < n) {
            const uint32_t p = (i & 1) ? (rk[i >> 1] >> 16) : (rk[i >> 1] & 0xFFFFu);
            stage[p] = k[i];
            sdig[p] = (uint8_t)digit_of(k[i], shift, mask);
        }
    }
    __syncthreads();
    const uint32_t valid_in_tile = (uint32_t)((n - tile_base) < (uint64_t)Tile ? (n - tile_base) : Tile);
    for (uint32_t p = threadIdx.x; p < valid_in_tile; p += kBlock) emit.key(p + doff[sdig[p]], stage[p]);
    if (Emit::kValues && vals) {
        // values are read only now, which keeps the ranking phase's register footprint small
        V v[Items];
#pragma unroll
        for (int i = 0; i < Items; i++) {
            const uint64_t idx = base + (uint64_t)i * 64;
            v[i] = (idx < n) ? vals[idx] : V(0);
        }
        __syncthreads();
        V* sv = reinterpret_cast<V*>(stage);
#pragma unroll
        for (int i = 0; i < Items; i++)
            if (base + (uint64_t)i * 64 < n) sv[(i & 1) ? (rk[i >> 1] >> 16) : (rk[i >> 1] & 0xFFFFu)] = v[i];
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < valid_in_tile; p += kBlock) emit.val(p + doff[sdig[p]], sv[p]);
    }
}

template <typename T>
__global__ void k_fill(T* p, uint64_t n, T v)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Stable LSD radix sort of (key, value) pairs by key bits [lo, hi).  keys/vals hold the
// input in buffer 0 and ping-pong through buffer 1.  Without `where` the sorted data ends in
// buffer 0 (copied back after an odd number of passes); with `where` it stays where the last
// pass wrote it and *where is set to 0 or 1.  vals0/vals1 may be nullptr (keys only).  If
// `last` is given, the final pass calls it instead of writing a buffer.
template <typename K, typename V, typename LastEmit = EmitPair<K, V>>
void radix_sort(Storage& S, K* keys0, K* keys1, V* vals0, V* vals1, uint64_t n, int lo, int hi,
                const LastEmit* last = nullptr, int* where = nullptr)
{
    if (where) *where = 0;
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
    // items per thread: the launch shape of this build (8-byte keys 8 or 16, 4-byte keys twice)
    const bool half = S.tune.sort_items == 8;
    constexpr int kFull = sort_items<K>(), kHalf = sort_items<K>() / 2;
    const uint32_t tiles = ceil_div(n, (uint64_t)kBlock * (half ? kHalf : kFull));
    const int passes = (hi - lo + kRadixBits - 1) / kRadixBits;
    MIBWT_REQUIRE(passes <= kMaxPasses, LIBCUBWT_BAD_PARAMETER, "radix_sort: key too wide");
    uint32_t* start = S.buf<uint32_t>(B_SORT_TCNT, (size_t)tiles * kRadix);
    K* kin = keys0;
    K* kout = keys1;
    V* vin = vals0;
    V* vout = vals1;
    for (int p = 0; p < passes; p++) {
        const int shift = lo + p * kRadixBits;
        const int nb = (hi - shift) < kRadixBits ? (hi - shift) : kRadixBits;
        const uint32_t mask = (1u << nb) - 1u;
        const double vbytes = vin ? (double)sizeof(V) : 0.0;
        if (half)
            k_radix_up<K, kHalf><<<tiles, kBlock, 0, S.stream>>>(kin, n, shift, mask, tiles, start);
        else
            k_radix_up<K, kFull><<<tiles, kBlock, 0, S.stream>>>(kin, n, shift, mask, tiles, start);
        MIBWT_LAUNCH_CHECK();
        exclusive_sum<uint32_t>(S, start, start, (uint64_t)tiles * kRadix);
        auto launch = [&](auto emit, double bytes) {
            using E = decltype(emit);
            const int kt = S.kt_begin(1, bytes);
            if (half)
                k_radix_pass<K, V, E, kHalf><<<tiles, kBlock, 0, S.stream>>>(kin, vin, n, shift, mask, start, tiles,
                                                                            emit);
            else
                k_radix_pass<K, V, E, kFull><<<tiles, kBlock, 0, S.stream>>>(kin, vin, n, shift, mask, start, tiles,
                                                                            emit);
            S.kt_end(kt);
            MIBWT_LAUNCH_CHECK();
        };
        if (p == passes - 1 && last)
            launch(*last, (double)n * (sizeof(K) + vbytes) + LastEmit::bytes_per_item * (double)n);
        else
            launch(EmitPair<K, V>{kout, vout}, 2.0 * (double)n * (sizeof(K) + vbytes));
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (last) return;
    if (kin != keys0) {
        if (where) {
            *where = 1;
        } else {
            MIBWT_HIP(hipMemcpyAsync(keys0, kin, sizeof(K) * n, hipMemcpyDeviceToDevice, S.stream));
            if (vals0 && vin) MIBWT_HIP(hipMemcpyAsync(vals0, vin, sizeof(V) * n, hipMemcpyDeviceToDevice, S.stream));
        }
    }
}

// ----------------------------------------------------------------------------- small kernels

template <typename T>
void fill(Storage& S, T* p, uint64_t n, T v)
{
    if (!n) return;
    k_fill<T><<<ceil_div(n, kBlock), kBlock, 0, S.stream>>>(p, n, v);
    MIBWT_LAUNCH_CHECK();
}

inline uint32_t grid_for(uint64_t n) { return ceil_div(n ? n : 1, kBlock); }

}  // namespace mibwt
