// Inverse BWT on gfx950 (libcubwt_unbwt, libcubwt.cuh:91-101; the reference declares it but its
// kernels exist only as an undefined remnant, libcubwt.cu:2452-2968).
//
// The n+1 rows of the BWT matrix of T$: B[index] = '$' and the other rows hold the text's bytes
// (libcubwt format: B[r] = L[r] for r < index, L[r-1] for r > index; reference format:
// B[0] = O[index-1], B[r] = O[r-1]).  LF maps a row to the row of the preceding rotation; the
// text is T[n-1-k] = B[r_k] with r_0 = 0, r_{k+1} = LF(r_k), ending at the '$' row after n steps.
//
//  1. LF of every row, computed in row order from per-tile byte counts (the '$' row sorts
//     first): rec[r] = LF(r) | B[r] << 32, so a walk step is one 8-byte load.
//  2. The LF cycle is ranked by sampling: every kUnSample-th row is a splitter; each splitter
//     walks to the next splitter (or the '$' row) and records the number of rows and the next
//     splitter.  The splitter list is ranked through a ruling set: every kUnRule-th splitter
//     is a ruler that follows the list to the next ruler, the ruler list is ranked by pointer
//     jumping (suffix sums, a fixed number of rounds), and the rulers walk their stretch again
//     to hand every splitter its distance to the end; a second walk writes the text.  An input
//     that is not a BWT with this primary index (LF cycle from row 0 shorter than n) is
//     reported, not written.
#include "prims.hpp"

namespace mibwt {

int64_t unbwt_run(Storage& S, const uint8_t* dL, uint8_t* dU, int64_t n64, int64_t index64, bool ref_format,
                  uint32_t* dI, int64_t rate, bool own_phases = true);

constexpr uint32_t kUnSample = 64;
constexpr uint32_t kUnRule = 64;  // splitters per ruler
constexpr uint32_t kUnEnd = 0xFFFFFFFFu;
constexpr uint32_t kUnStash = 256;  // bytes kept per walk (a walk exceeds it with probability ~e^-4)

__device__ __forceinline__ uint8_t bwt_row(const uint8_t* __restrict__ L, uint64_t r, uint64_t index, bool ref)
{
    if (ref) return (r == 0) ? L[index - 1] : L[r - 1];
    return (r < index) ? L[r] : L[r - 1];
}

// LF of the n byte rows (every row but the '$' row) without sorting them: LF(r) = 1 + (rows whose
// byte is smaller) + (rows before r with the same byte), the '$' row being row 0 of the sorted
// order.  Tiles of kLfTile byte rows, one wavefront per kLfWave consecutive rows (16 per lane);
// k_lf_count writes every tile's byte counts digit-major (cnt[c * tiles + tile]), one flat
// exclusive scan turns them into each (byte, tile)'s first LF - 1, and k_lf_fill ranks the rows
// of its tile per byte (stable: the lanes with equal bytes found by one ballot per byte bit) and
// writes rec[r] = LF(r) | B[r] << 32 in row order.
constexpr int kLfPer = 16;
constexpr uint32_t kLfWave = 64 * kLfPer;
constexpr uint32_t kLfTile = kWaves * kLfWave;

// byte of byte row q (row q, or q + 1 past the '$' row)
__device__ __forceinline__ uint32_t lf_byte(const uint8_t* __restrict__ L, uint64_t q, uint64_t index, bool ref)
{
    return bwt_row(L, (q < index) ? q : q + 1, index, ref);
}

// per-wavefront stable ranking of the byte rows [w0, w0 + kLfWave): rank[k] = the rows before
// item k in the wavefront with the same byte; wcnt (LDS, 256 entries, zeroed) ends with the
// wavefront's counts
__device__ __forceinline__ void lf_rank(const uint32_t (&c)[kLfPer], const bool (&valid)[kLfPer], uint32_t* wcnt,
                                        uint32_t (&rank)[kLfPer])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < kLfPer; k++) {
        const uint64_t vm = __ballot(valid[k]);
        uint32_t plo = (uint32_t)vm, phi = (uint32_t)(vm >> 32);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint32_t mk = (uint32_t)((int32_t)(c[k] << (31 - b)) >> 31);
            const uint64_t m = __ballot(mk != 0u);
            plo &= ~((uint32_t)m ^ mk);
            phi &= ~((uint32_t)(m >> 32) ^ mk);
        }
        const int leader = plo ? (__ffs(plo) - 1) : (31 + __ffs(phi));
        uint32_t old = 0;
        if (valid[k] && lane == leader) old = atomicAdd(&wcnt[c[k]], (uint32_t)(__popc(plo) + __popc(phi)));
        old = __shfl(old, valid[k] ? leader : 0, 64);
        rank[k] = old + __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));
    }
}

__global__ __launch_bounds__(kBlock) void k_lf_count(const uint8_t* __restrict__ L, uint64_t n, uint64_t index,
                                                    int ref, uint32_t tiles, uint32_t* __restrict__ cnt)
{
    __shared__ uint32_t wcnt[kWaves][256];
    for (int i = threadIdx.x; i < kWaves * 256; i += kBlock) (&wcnt[0][0])[i] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t w0 = (uint64_t)blockIdx.x * kLfTile + (uint64_t)w * kLfWave;
    uint32_t c[kLfPer], rank[kLfPer];
    bool valid[kLfPer];
#pragma unroll
    for (int k = 0; k < kLfPer; k++) {
        const uint64_t q = w0 + (uint64_t)k * 64 + lane;
        valid[k] = q < n;
        c[k] = valid[k] ? lf_byte(L, q, index, ref != 0) : 0u;
    }
    lf_rank(c, valid, wcnt[w], rank);
    __syncthreads();
    const uint32_t b = threadIdx.x;  // kBlock == 256 bytes
    cnt[(uint64_t)b * tiles + blockIdx.x] = wcnt[0][b] + wcnt[1][b] + wcnt[2][b] + wcnt[3][b];
}

__global__ __launch_bounds__(kBlock) void k_lf_fill(const uint8_t* __restrict__ L, uint64_t n, uint64_t index,
                                                   int ref, uint32_t tiles, const uint32_t* __restrict__ start,
                                                   uint64_t* __restrict__ rec)
{
    __shared__ uint32_t wcnt[kWaves][256];
    for (int i = threadIdx.x; i < kWaves * 256; i += kBlock) (&wcnt[0][0])[i] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t w0 = (uint64_t)blockIdx.x * kLfTile + (uint64_t)w * kLfWave;
    uint32_t c[kLfPer], rank[kLfPer];
    bool valid[kLfPer];
#pragma unroll
    for (int k = 0; k < kLfPer; k++) {
        const uint64_t q = w0 + (uint64_t)k * 64 + lane;
        valid[k] = q < n;
        c[k] = valid[k] ? lf_byte(L, q, index, ref != 0) : 0u;
    }
    lf_rank(c, valid, wcnt[w], rank);
    __syncthreads();
    {
        // per byte: the tile's first LF - 1, then the earlier wavefronts' rows of that byte
        const uint32_t b = threadIdx.x;
        uint32_t run = start[(uint64_t)b * tiles + blockIdx.x];
#pragma unroll
        for (int q = 0; q < kWaves; q++) {
            const uint32_t x = wcnt[q][b];
            wcnt[q][b] = run;
            run += x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kLfPer; k++) {
        if (!valid[k]) continue;
        const uint64_t q = w0 + (uint64_t)k * 64 + lane;
        const uint64_t r = (q < index) ? q : q + 1;
        rec[r] = (uint64_t)(1 + wcnt[w][c[k]] + rank[k]) | ((uint64_t)c[k] << 32);
    }
}

// walk from splitter s to the next splitter: rows visited, next splitter (kUnEnd: the '$' row);
// with `stash`, the bytes of the first kUnStash rows of the walk and the row after them (rowk)
__global__ void k_un_walk1(const uint64_t* __restrict__ rec, uint64_t m, uint64_t index, uint32_t S,
                           uint32_t* __restrict__ nxt, uint32_t* __restrict__ w, uint32_t* __restrict__ bad,
                           uint8_t* __restrict__ stash, uint32_t* __restrict__ rowk)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const uint64_t r0 = (uint64_t)s * kUnSample;
    if (r0 == index) {  // the '$' row ends the text
        nxt[s] = kUnEnd;
        w[s] = 0;
        return;
    }
    uint64_t r = r0, steps = 0;
    uint32_t next = kUnEnd;
    // the kept bytes leave 8 at a time (aligned 8-byte stores; a walk's last word may carry bytes
    // past its end, which nothing reads)
    uint64_t* st = stash ? reinterpret_cast<uint64_t*>(stash + (uint64_t)s * kUnStash) : nullptr;
    uint64_t acc = 0;
    for (;;) {
        const uint64_t x = rec[r];
        if (st && steps < kUnStash) {
            acc |= ((x >> 32) & 0xFFull) << (8 * (steps & 7));
            if ((steps & 7) == 7) {
                st[steps >> 3] = acc;
                acc = 0;
            }
        }
        r = (uint32_t)x;
        steps++;
        if (st && steps == kUnStash) rowk[s] = (uint32_t)r;
        if (r == index) break;
        if (r % kUnSample == 0) {
            next = (uint32_t)(r / kUnSample);
            break;
        }
        if (steps > m) {  // cannot happen for a permutation; guards the loop
            atomicOr(bad, 1u);
            break;
        }
    }
    if (st && steps < kUnStash && (steps & 7)) st[steps >> 3] = acc;
    nxt[s] = next;
    w[s] = (uint32_t)steps;
}

// one pointer-jumping round on suffix sums: d(s) += d(nxt(s)), nxt(s) = nxt(nxt(s))
__global__ void k_un_jump(const uint32_t* __restrict__ nxt, const uint32_t* __restrict__ d, uint32_t S,
                          uint32_t* __restrict__ nxt2, uint32_t* __restrict__ d2)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const uint32_t t = nxt[s];
    if (t == kUnEnd) {
        nxt2[s] = kUnEnd;
        d2[s] = d[s];
        return;
    }
    d2[s] = d[s] + d[t];
    nxt2[s] = nxt[t];
}

// Ruler q (splitter q kUnRule) follows the splitter list to the next ruler or the end: the rows
// on the way (its own walk's included) and that ruler (kUnEnd: the end).  A stretch longer than
// the list (a cycle without a ruler: not a BWT) sets *bad.
__global__ void k_un_rule1(const uint32_t* __restrict__ nxt, const uint32_t* __restrict__ w, uint32_t Sn, uint32_t Rn,
                           uint32_t* __restrict__ rnext, uint32_t* __restrict__ rsum, uint32_t* __restrict__ bad)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Rn) return;
    const uint32_t s = q * kUnRule;
    uint32_t sum = w[s], t = nxt[s], steps = 0;
    while (t != kUnEnd && t % kUnRule != 0) {
        if (++steps > Sn) {
            atomicOr(bad, 1u);
            break;
        }
        sum += w[t];
        t = nxt[t];
    }
    rnext[q] = (t == kUnEnd) ? kUnEnd : t / kUnRule;
    rsum[q] = sum;
}

// Ruler q, with its rows to the end rdist[q], walks its stretch again: dist(s) = rows from
// splitter s to the end, its own included.
__global__ void k_un_rule2(const uint32_t* __restrict__ nxt, const uint32_t* __restrict__ w, uint32_t Sn, uint32_t Rn,
                           const uint32_t* __restrict__ rdist, uint32_t* __restrict__ dist)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Rn) return;
    const uint32_t s = q * kUnRule;
    uint32_t acc = rdist[q];
    dist[s] = acc;
    acc -= w[s];
    uint32_t t = nxt[s], steps = 0;
    while (t != kUnEnd && t % kUnRule != 0 && ++steps <= Sn) {
        dist[t] = acc;
        acc -= w[t];
        t = nxt[t];
    }
}

// second walk: splitter s starts at step k = n - dist(s) (dist = rows from s to the end); the
// row at step k is the rotation starting at text position n - k and is preceded by T[n-1-k].
// With U the text is written; with I (sampling rate `rate`) the rows of the positions that are
// multiples of the rate (libcubwt_bwt_aux, libcubwt.cuh:79-89: the same convention as the
// primary index, which is I[0]).
__global__ void k_un_walk2(const uint64_t* __restrict__ rec, const uint32_t* __restrict__ w,
                           const uint32_t* __restrict__ dist, uint64_t n, uint32_t S, uint8_t* __restrict__ U,
                           uint32_t* __restrict__ I, uint64_t rate)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || w[s] == 0) return;
    uint64_t k = n - dist[s];
    uint64_t r = (uint64_t)s * kUnSample;
    for (uint32_t i = 0; i < w[s]; i++) {
        const uint64_t x = rec[r];
        if (k < n) {
            if (U) U[n - 1 - k] = (uint8_t)(x >> 32);
            if (I && k > 0 && (n - k) % rate == 0) I[(n - k) / rate] = (uint32_t)r;
        }
        k++;
        r = (uint32_t)x;
    }
}

// The kept bytes of the walks to their places in the text: step i of walk s is step
// k = n - dist(s) + i of the whole walk, preceded by T[n-1-k].  A wavefront copies kUnEmitWalks
// walks one after the other, a byte per lane (a walk keeps 64 bytes on average).
constexpr uint32_t kUnEmitWalks = 16;
__global__ __launch_bounds__(256) void k_un_emit(const uint8_t* __restrict__ stash, const uint32_t* __restrict__ w,
                                                 const uint32_t* __restrict__ dist, uint64_t n, uint32_t S,
                                                 uint8_t* __restrict__ U)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t s0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * kUnEmitWalks;
    for (uint32_t s = s0; s < s0 + kUnEmitWalks && s < S; s++) {
        const uint32_t ws = w[s];
        const uint32_t cnt = ws < kUnStash ? ws : kUnStash;
        const uint64_t k0 = n - dist[s];
        const uint8_t* st = stash + (uint64_t)s * kUnStash;
        for (uint32_t i = lane; i < cnt; i += 64) {
            const uint64_t k = k0 + i;
            if (k < n) U[n - 1 - k] = st[i];
        }
    }
}

// the rest of the walks longer than kUnStash, from the row they reached
__global__ void k_un_tail(const uint64_t* __restrict__ rec, const uint32_t* __restrict__ w,
                          const uint32_t* __restrict__ dist, const uint32_t* __restrict__ rowk, uint64_t n,
                          uint32_t S, uint8_t* __restrict__ U)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || w[s] <= kUnStash) return;
    uint64_t k = n - dist[s] + kUnStash;
    uint64_t r = rowk[s];
    for (uint32_t i = kUnStash; i < w[s]; i++) {
        const uint64_t x = rec[r];
        if (k < n) U[n - 1 - k] = (uint8_t)(x >> 32);
        k++;
        r = (uint32_t)x;
    }
}

// bad |= 1 where a[i] != b[i] (16 bytes per thread; both buffers padded to 16 bytes)
__global__ void k_un_compare(const uint8_t* __restrict__ a, const uint8_t* __restrict__ b, uint64_t n,
                             uint32_t* __restrict__ bad)
{
    const uint64_t p = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 16;
    if (p >= n) return;
    bool diff = false;
    if (p + 16 <= n && ((uintptr_t)a & 15u) == 0 && ((uintptr_t)b & 15u) == 0) {
        const uint4 x = *reinterpret_cast<const uint4*>(a + p), y = *reinterpret_cast<const uint4*>(b + p);
        diff = (x.x != y.x) | (x.y != y.y) | (x.z != y.z) | (x.w != y.w);
    } else {
        for (uint64_t i = p; i < n && i < p + 16; i++) diff |= a[i] != b[i];
    }
    if (diff) atomicOr(bad, 1u);
}

// Whether the device BWT dO[0..n) (reference format, primary index `index`) inverts to the
// device text dT[0..n): the exact check of a build whose intermediate steps trust fingerprints.
// Runs on the storage stream inside the caller's phases (none of its own).
bool bwt_inverts_to(Storage& S, const uint8_t* dO, int64_t n, int64_t index, const uint8_t* dT)
{
    if (index < 1 || index > n) return false;
    uint8_t* dU = S.buf<uint8_t>(B_UN_IO, (size_t)n + 16);
    try {
        if (unbwt_run(S, dO, dU, n, index, true, nullptr, 0, false) != LIBCUBWT_NO_ERROR) return false;
    } catch (const Error& e) {
        if (e.code != LIBCUBWT_BAD_PARAMETER) throw;  // not a BWT with this index: a mismatch
        return false;
    }
    uint32_t* sc = S.buf<uint32_t>(B_SCALARS, SC_COUNT_);
    fill<uint32_t>(S, sc + SC_COLLISION, 1, 0u);
    k_un_compare<<<grid_for(ceil_div((uint64_t)n, 16)), kBlock, 0, S.stream>>>(dU, dT, (uint64_t)n, sc + SC_COLLISION);
    MIBWT_LAUNCH_CHECK();
    uint32_t bad = 1;
    S.read(&bad, sc + SC_COLLISION);
    return bad == 0;
}

// Inverse of the BWT dL[0..n) with primary index `index` into dU[0..n) (device buffers), or,
// with dI, the sampled rows dI[p / rate] of the text positions p = 0, rate, 2 rate, ...
// Returns LIBCUBWT_NO_ERROR or LIBCUBWT_BAD_PARAMETER when the input is not a BWT.  With
// `own_phases` the call records and collects its own phase timing (a standalone entry).
int64_t unbwt_run(Storage& S, const uint8_t* dL, uint8_t* dU, int64_t n64, int64_t index64, bool ref_format,
                  uint32_t* dI, int64_t rate, bool own_phases)
{
    MIBWT_REQUIRE(n64 >= 1 && n64 < (int64_t)INT32_MAX - (1 << 20), LIBCUBWT_BAD_PARAMETER, "n out of range");
    MIBWT_REQUIRE(index64 >= 1 && index64 <= n64, LIBCUBWT_BAD_PARAMETER, "primary index out of range");
    const uint64_t n = (uint64_t)n64, index = (uint64_t)index64, m = n + 1;
    if (own_phases) {
        S.phases_reset();
        apply_tuning(S, n);
    }
    auto phase_begin = [&](const char* name) { return own_phases ? S.phase_begin(name) : -1; };
    auto phase_end = [&](int id, double bytes) {
        if (id >= 0) S.phase_end(id, bytes);
    };
    uint32_t* sc = S.buf<uint32_t>(B_SCALARS, SC_COUNT_);
    int ph = phase_begin("lf");
    uint64_t* rec = S.buf<uint64_t>(B_UN_REC, m);
    fill<uint64_t>(S, rec + index, 1, 0ull);  // LF('$' row) = row 0; its byte field is unused
    const uint32_t lft = (uint32_t)ceil_div(n, (uint64_t)kLfTile);
    uint32_t* lfc = S.buf<uint32_t>(B_SORT_TCNT, (size_t)lft * 256);
    k_lf_count<<<lft, kBlock, 0, S.stream>>>(dL, n, index, ref_format ? 1 : 0, lft, lfc);
    MIBWT_LAUNCH_CHECK();
    exclusive_sum<uint32_t>(S, lfc, lfc, (uint64_t)lft * 256);
    k_lf_fill<<<lft, kBlock, 0, S.stream>>>(dL, n, index, ref_format ? 1 : 0, lft, lfc, rec);
    MIBWT_LAUNCH_CHECK();
    phase_end(ph, 2.0 * n + 8.0 * n);

    ph = phase_begin("rank");
    const uint32_t Sn = (uint32_t)ceil_div(m, (uint64_t)kUnSample);
    uint32_t* nxt0 = S.buf<uint32_t>(B_UN_NXT0, Sn);
    uint32_t* w = S.buf<uint32_t>(B_UN_W, Sn);
    uint32_t* d0 = S.buf<uint32_t>(B_UN_D0, Sn);
    fill<uint32_t>(S, sc + SC_COLLISION, 1, 0u);
    // the text (not the sampled rows) is written from the bytes kept by the first walk
    uint8_t* stash = dI ? nullptr : S.buf<uint8_t>(B_UN_STASH, (size_t)Sn * kUnStash);
    uint32_t* rowk = dI ? nullptr : S.buf<uint32_t>(B_UN_ROW, Sn);
    k_un_walk1<<<grid_for(Sn), kBlock, 0, S.stream>>>(rec, m, index, Sn, nxt0, w, sc + SC_COLLISION, stash, rowk);
    MIBWT_LAUNCH_CHECK();
    // ruling set over the splitter list (ruler 0 is splitter 0, the start of the cycle)
    const uint32_t Rn = (uint32_t)ceil_div((uint64_t)Sn, (uint64_t)kUnRule);
    uint32_t* rn0 = S.buf<uint32_t>(B_UN_RN0, Rn);
    uint32_t* rn1 = S.buf<uint32_t>(B_UN_RN1, Rn);
    uint32_t* rd0 = S.buf<uint32_t>(B_UN_RD0, Rn);
    uint32_t* rd1 = S.buf<uint32_t>(B_UN_RD1, Rn);
    k_un_rule1<<<grid_for(Rn), kBlock, 0, S.stream>>>(nxt0, w, Sn, Rn, rn0, rd0, sc + SC_COLLISION);
    MIBWT_LAUNCH_CHECK();
    for (uint64_t span = 1; span < Rn; span *= 2) {
        k_un_jump<<<grid_for(Rn), kBlock, 0, S.stream>>>(rn0, rd0, Rn, rn1, rd1);
        MIBWT_LAUNCH_CHECK();
        std::swap(rn0, rn1);
        std::swap(rd0, rd1);
    }
    // the walk from row 0 must reach the '$' row after exactly n byte rows
    uint32_t hv[3];
    MIBWT_HIP(hipMemcpyAsync(sc + SC_TOTAL, rd0, sizeof(uint32_t), hipMemcpyDeviceToDevice, S.stream));
    MIBWT_HIP(hipMemcpyAsync(sc + SC_GROUPS, rn0, sizeof(uint32_t), hipMemcpyDeviceToDevice, S.stream));
    S.read(hv, sc + SC_COLLISION, 1);
    S.read(hv + 1, sc + SC_TOTAL, 1);
    S.read(hv + 2, sc + SC_GROUPS, 1);
    MIBWT_REQUIRE(hv[0] == 0 && hv[1] == (uint32_t)n && hv[2] == kUnEnd, LIBCUBWT_BAD_PARAMETER,
                  "not a BWT with this primary index");
    k_un_rule2<<<grid_for(Rn), kBlock, 0, S.stream>>>(nxt0, w, Sn, Rn, rd0, d0);
    MIBWT_LAUNCH_CHECK();
    phase_end(ph, 16.0 * m);

    ph = phase_begin("text");
    if (dI) fill<uint32_t>(S, dI, 1, (uint32_t)index);  // position 0: the primary index
    if (dI) {
        k_un_walk2<<<grid_for(Sn), kBlock, 0, S.stream>>>(rec, w, d0, n, Sn, dU, dI, (uint64_t)rate);
    } else {
        k_un_emit<<<ceil_div(Sn, 4 * kUnEmitWalks), 256, 0, S.stream>>>(stash, w, d0, n, Sn, dU);
        MIBWT_LAUNCH_CHECK();
        k_un_tail<<<grid_for(Sn), kBlock, 0, S.stream>>>(rec, w, d0, rowk, n, Sn, dU);
    }
    MIBWT_LAUNCH_CHECK();
    phase_end(ph, 9.0 * m);
    if (own_phases) S.phases_collect();
    return LIBCUBWT_NO_ERROR;
}

}  // namespace mibwt
