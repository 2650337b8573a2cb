// Shared host/device infrastructure of the mibwt library: error handling, the per-storage
// stream and grow-only device buffers, phase timing.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mibwt.h"

namespace mibwt {

// An error that maps onto a libcubwt-style return code.
struct Error : std::runtime_error {
    int64_t code;
    Error(int64_t c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline int64_t hip_error_code(hipError_t e)
{
    switch (e) {
    case hipSuccess: return LIBCUBWT_NO_ERROR;
    case hipErrorOutOfMemory: return LIBCUBWT_GPU_NOT_ENOUGH_MEMORY;
    case hipErrorNoDevice:
    case hipErrorInvalidDevice:
    case hipErrorNoBinaryForGpu:
    case hipErrorInsufficientDriver: return LIBCUBWT_GPU_NOT_SUPPORTED;
    default: return LIBCUBWT_GPU_ERROR;
    }
}

#define MIBWT_HIP(call)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            char b_[256];                                                                  \
            snprintf(b_, sizeof(b_), "%s:%d %s: %s", __FILE__, __LINE__, #call,            \
                     hipGetErrorString(e_));                                               \
            throw ::mibwt::Error(::mibwt::hip_error_code(e_), b_);                         \
        }                                                                                  \
    } while (0)

#define MIBWT_LAUNCH_CHECK() MIBWT_HIP(hipGetLastError())

#define MIBWT_REQUIRE(cond, code, msg)                                                     \
    do {                                                                                   \
        if (!(cond)) throw ::mibwt::Error((code), (msg));                                  \
    } while (0)

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

inline int bits_for(uint64_t v)  // number of bits needed to represent values 0..v
{
    int b = 0;
    while (b < 64 && (v >> b) != 0) b++;
    return b;
}

// A grow-only device allocation owned by a Storage.
struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
};

struct Phase {
    std::string name;
    hipEvent_t start = nullptr, stop = nullptr;
    double bytes = 0;
};

// State of one single-pass ("chained") scan launch.  Tiles take tickets in launch order and
// publish status words: bits 63..34 the launch's epoch, 33..32 the state (1 = tile total,
// 2 = inclusive prefix), 31..0 the value.  Words left by older launches carry an older epoch
// and read as "not yet published", so the status array is never cleared between launches.
// The tile that draws the last ticket resets the ticket counter for the next launch.
struct Chain {
    uint64_t* status;
    uint32_t* ticket;
    uint32_t epoch;
};

// Launch shapes chosen per build (by the autotuner's table for the build's size bucket, or the
// defaults): trigger-scan sub-tiles per chained tile (1, 2, 4, 8), radix items per thread for
// 8-byte keys (8 or 16; 4-byte keys take twice as many), chained-scan items per thread (16 or
// 64), and the width of the packed initial keys of the suffix sorts (32..64 bits: fewer radix
// passes against more doubling rounds).
struct Tuning {
    int scan_sub = 8;
    int sort_items = 16;
    int chain_items = 64;
    int key_bits = 64;
};

struct Storage;
// Sets S.tune for a build of `size` bytes from the autotuner's table (tune.hip); kept while
// S.tune_pinned.
void apply_tuning(Storage& S, uint64_t size);

// Everything one libcubwt-style storage owns.  One stream; not safe for concurrent calls.
struct Storage {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t max_length = 0;
    std::vector<DevBuf> bufs;
    void* pinned = nullptr;          // small pinned host buffer for scalar read-backs
    size_t pinned_cap = 0;
    // Device-to-host signal: a kernel stores a few words into mapped pinned memory and then a
    // sequence number (system scope, release); the host polls the sequence number instead of
    // synchronising the stream, which answers a per-round count in a few microseconds.
    static constexpr int kSigWords = 63, kSigSeq = 63;
    uint32_t* sig = nullptr;      // host view
    uint32_t* sig_dev = nullptr;  // device view
    uint32_t sig_seq = 0;
    // phase timing
    std::vector<Phase> phases;
    int phase_used = 0;
    hipEvent_t t_begin = nullptr, t_end = nullptr;
    mibwt_phase_stats stats{};
    // per-launch timing of one kernel family (0 = off, 1 = radix scatter, 2 = trigger scan)
    int kt_family = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> kt_events;
    std::vector<double> kt_bytes;
    int kt_used = 0;
    // worker storages of mibwt_batch_bwt (one stream each)
    std::vector<Storage*> pool;
    // state kept for mibwt_time_kernel
    const uint8_t* last_text = nullptr;
    bool verbose = getenv("MIBWT_VERBOSE") != nullptr;  // per-round counts on stderr
    int64_t last_n = 0;
    int last_W = 0, last_P = 0;
    int64_t last_sort_items = 0;
    // chained scans: epoch of the last launch
    uint32_t chain_epoch = 0;
    // selector sketch tables: epoch of the last use
    uint32_t sketch_epoch = 0;
    // Lyndon builds on this storage (seeds the per-call fingerprint base)
    uint64_t lyn_calls = 0;
    // launch shapes of the current build; `tune_pinned` keeps them (the autotuner's trials)
    Tuning tune;
    bool tune_pinned = false;
    uint64_t sel_model_version = 0;  // selector model resident in B_SEL_MP / B_SEL_MPB / B_SEL_NORM
    // Arena mode (the allocation-free entry point): every buffer is carved from one device region
    // that the caller owns, host read-backs go through `arena_host` (no pinned or mapped memory),
    // and no phase is timed (no events).  arena_peak = high-water mark of the region.
    uint8_t* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0, arena_peak = 0;
    alignas(16) unsigned char arena_host[4096];

    // start a call in arena mode on [base, base + cap): every buffer is carved anew
    void arena_reset(uint8_t* base, size_t cap)
    {
        arena = base;
        arena_cap = cap;
        arena_used = 0;
        arena_peak = 0;
        for (DevBuf& b : bufs) b = DevBuf{};
        chain_epoch = 0;  // the chain status words are cleared again on first use
    }

    // device buffer #id with capacity >= bytes; contents are not preserved on growth
    template <typename T>
    T* buf(int id, size_t count)
    {
        size_t bytes = count * sizeof(T) + 256;
        if ((size_t)id >= bufs.size()) bufs.resize(id + 1);
        DevBuf& b = bufs[id];
        if (arena) {
            // bump allocation; a buffer that grows leaves its old region behind until the call ends
            if (b.cap >= bytes) return reinterpret_cast<T*>(b.ptr);
            const size_t off = (arena_used + 255) & ~(size_t)255;
            if (off > arena_cap || bytes > arena_cap - off)
                throw Error(LIBCUBWT_GPU_NOT_ENOUGH_MEMORY, "the scratch array is too small for this text");
            b.ptr = arena + off;
            b.cap = bytes;
            arena_used = off + bytes;
            if (arena_used > arena_peak) arena_peak = arena_used;
            return reinterpret_cast<T*>(b.ptr);
        }
        if (b.cap < bytes) {
            if (b.ptr) {
                MIBWT_HIP(hipStreamSynchronize(stream));
                MIBWT_HIP(hipFree(b.ptr));
                b.ptr = nullptr;
                b.cap = 0;
            }
            size_t want = bytes + bytes / 8;
            hipError_t e = hipMalloc(&b.ptr, want);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                b.ptr = nullptr;
                throw Error(LIBCUBWT_GPU_NOT_ENOUGH_MEMORY, "hipMalloc failed");
            }
            b.cap = want;
        }
        return reinterpret_cast<T*>(b.ptr);
    }

    // status words for a chained scan over `tiles` tiles (cleared when first allocated)
    Chain chain(uint32_t tiles);

    template <typename T>
    T* host(size_t count)
    {
        size_t bytes = count * sizeof(T);
        if (arena) {
            MIBWT_REQUIRE(bytes <= sizeof(arena_host), LIBCUBWT_GPU_ERROR, "read-back too large");
            return reinterpret_cast<T*>(arena_host);
        }
        if (pinned_cap < bytes) {
            if (pinned) MIBWT_HIP(hipHostFree(pinned));
            pinned = nullptr;
            MIBWT_HIP(hipHostMalloc(&pinned, bytes < 4096 ? 4096 : bytes, hipHostMallocDefault));
            pinned_cap = bytes < 4096 ? 4096 : bytes;
        }
        return reinterpret_cast<T*>(pinned);
    }

    uint32_t* signal_dev()
    {
        if (!sig) {
            void* p = nullptr;
            MIBWT_HIP(hipHostMalloc(&p, 4 * (kSigSeq + 1), hipHostMallocMapped | hipHostMallocCoherent));
            sig = reinterpret_cast<uint32_t*>(p);
            memset(sig, 0, 4 * (kSigSeq + 1));
            MIBWT_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&sig_dev), p, 0));
        }
        return sig_dev;
    }
    // next sequence number for a signal that is about to be enqueued (never 0)
    uint32_t signal_next()
    {
        sig_seq = (sig_seq == 0xFFFFFFFFu) ? 1u : sig_seq + 1;
        return sig_seq;
    }
    // Wait until the device published `seq`, then copy `count` signal words.  After 200 ms of
    // polling the stream is synchronised instead (which also surfaces a failed kernel); the
    // signal must be there once the stream is idle.
    void wait_signal(uint32_t seq, uint32_t* dst, int count)
    {
        const auto t0 = std::chrono::steady_clock::now();
        for (;;) {
            if (__atomic_load_n(sig + kSigSeq, __ATOMIC_ACQUIRE) == seq) break;
            if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(200)) {
                MIBWT_HIP(hipStreamSynchronize(stream));
                if (__atomic_load_n(sig + kSigSeq, __ATOMIC_ACQUIRE) != seq)
                    throw Error(LIBCUBWT_GPU_ERROR, "device signal missing after the stream completed");
                break;
            }
            __builtin_ia32_pause();
        }
        for (int i = 0; i < count; i++) dst[i] = __atomic_load_n(sig + i, __ATOMIC_RELAXED);
    }

    // read `count` device values of type T through the signal words (no stream
    // synchronisation; for read-backs inside a build, at most kSigWords words)
    template <typename T>
    void read_fast(T* dst, const T* dev, size_t count = 1);

    // read `count` device values of type T (synchronises the stream)
    template <typename T>
    void read(T* dst, const T* dev, size_t count = 1)
    {
        T* h = host<T>(count);
        MIBWT_HIP(hipMemcpyAsync(h, dev, sizeof(T) * count, hipMemcpyDeviceToHost, stream));
        MIBWT_HIP(hipStreamSynchronize(stream));
        memcpy(dst, h, sizeof(T) * count);
    }

    void phases_reset()
    {
        phase_used = 0;
        kt_used = 0;
        memset(&stats, 0, sizeof(stats));
    }
    // bracket one launch of the timed kernel family; returns -1 when timing is off
    int kt_begin(int family, double bytes)
    {
        if (kt_family != family) return -1;
        if ((int)kt_events.size() <= kt_used) {
            hipEvent_t a, b;
            MIBWT_HIP(hipEventCreate(&a));
            MIBWT_HIP(hipEventCreate(&b));
            kt_events.emplace_back(a, b);
            kt_bytes.push_back(0);
        }
        kt_bytes[kt_used] = bytes;
        MIBWT_HIP(hipEventRecord(kt_events[kt_used].first, stream));
        return kt_used++;
    }
    void kt_end(int id)
    {
        if (id >= 0) MIBWT_HIP(hipEventRecord(kt_events[id].second, stream));
    }
    int phase_begin(const char* name)
    {
        if (arena) return -1;
        if ((int)phases.size() <= phase_used) {
            phases.emplace_back();
            MIBWT_HIP(hipEventCreate(&phases.back().start));
            MIBWT_HIP(hipEventCreate(&phases.back().stop));
        }
        Phase& p = phases[phase_used];
        p.name = name;
        p.bytes = 0;
        MIBWT_HIP(hipEventRecord(p.start, stream));
        return phase_used++;
    }
    void phase_end(int id, double bytes)
    {
        if (id < 0) return;
        phases[id].bytes = bytes;
        MIBWT_HIP(hipEventRecord(phases[id].stop, stream));
    }
    void phases_collect()
    {
        MIBWT_HIP(hipStreamSynchronize(stream));
        if (arena) {
            stats.num_phases = 0;
            stats.device_bytes = (int64_t)arena_peak;
            return;
        }
        stats.num_phases = phase_used < MIBWT_MAX_PHASES ? phase_used : MIBWT_MAX_PHASES;
        for (int i = 0; i < stats.num_phases; i++) {
            float ms = 0;
            MIBWT_HIP(hipEventElapsedTime(&ms, phases[i].start, phases[i].stop));
            snprintf(stats.name[i], sizeof(stats.name[i]), "%s", phases[i].name.c_str());
            stats.ms[i] = ms;
            stats.bytes[i] = phases[i].bytes;
        }
        stats.device_bytes = 0;
        for (const DevBuf& b : bufs) stats.device_bytes += (int64_t)b.cap;
        stats.kt_launches = kt_used;
        stats.kt_ms_total = 0;
        stats.kt_bytes_total = 0;
        for (int i = 0; i < kt_used; i++) {
            float ms = 0;
            MIBWT_HIP(hipEventElapsedTime(&ms, kt_events[i].first, kt_events[i].second));
            stats.kt_ms_total += ms;
            stats.kt_bytes_total += kt_bytes[i];
        }
    }
};

// Host destination of the intermediate PFP results (phrase ranks per occurrence, parse
// suffix array), first `cap` entries each.
struct PfpParseOut {
    uint32_t* rank;
    uint32_t* psa;
    int64_t cap;
    int64_t distinct;
};

// Buffer ids (one grow-only allocation each).  Buffers that are live at the same time have
// different ids.
enum BufId : int {
    B_TEXT = 0,     // padded copy of the text
    B_TRIG_MASK,    // 16-bit trigger masks
    B_TILE_CNT,     // per-tile trigger counts
    B_TRIG,         // trigger end positions + last phrase end (= phrase ends)
    B_SCAN_TMP,     // 4 consecutive ids for scan partials (one per recursion level)
    B_SCAN_TMP1,
    B_SCAN_TMP2,
    B_SCAN_TMP3,
    B_SORT_K0,
    B_SORT_K1,
    B_SORT_V0,
    B_SORT_V1,
    B_SORT_TCNT,
    B_PH_FLAG,      // dictionary head flags / ids
    B_PH_PID,       // phrase id per occurrence
    B_PH_CHAR,      // character before each phrase
    B_DU_HEAD,
    B_DU_START,
    B_DU_LEN,
    B_DU_TL,        // text bytes of each distinct phrase
    B_DU_TEXT,      // text of each distinct phrase (batches)
    B_PTEXT,        // text of each phrase occurrence (batches)
    B_BT_TABLES,    // batch layout tables
    B_BT_PIDX,      // batch primary indices
    B_BATCH_IN,     // host batches: the texts of one build, back to back
    B_BATCH_OUT,    // host batches: their BWTs
    B_DU_LEN64,
    B_DU_FREQ,
    B_DU_OFF,
    B_LEXRANK,
    B_DT_OWNER,
    B_DT_RANK,
    B_DT_SA,
    B_DT_POS,
    B_ACT_X,
    B_ACT_POS,
    B_ACT_X2,
    B_ACT_POS2,
    B_ACT_FLAG,
    B_ACT_HEAD,
    B_P_RANK,
    B_P_SA,
    B_G_OFF,
    B_G_MIX,
    B_G_ECNT,
    B_G_FILLC,
    B_MX_FLAG,
    B_MX_SCAN,
    B_MX_ENT,
    B_Q_CNT,
    B_Q_OFF,
    B_Q_BLK,
    B_Q_PART,
    B_Q_OCC,
    B_Q_RREC,
    B_DB_LPART,
    B_DB_PART,
    B_DB_EXTRA,
    B_E_GFLAG,
    B_E_CID,
    B_E_DELTAC,
    B_UN_REC,
    B_UN_NXT0,
    B_UN_W,
    B_UN_D0,
    B_UN_IO,
    B_UN_AUX,
    B_UN_STASH,     // the first kUnStash bytes of every splitter's walk
    B_UN_ROW,       // the row each longer walk reached after kUnStash steps
    B_UN_RN0,       // rulers of the splitter list: next ruler, rows to it (two of each)
    B_UN_RN1,
    B_UN_RD0,
    B_UN_RD1,
    B_SCALARS,
    B_STAGE,
    B_DB_KEY2,
    B_DB_GID,
    B_DB_GOFS,
    B_DB_LARGE,
    B_DB_NGZ,
    B_DB_R,
    B_DB_NR,
    B_DB_GOFS0,
    B_LG_SZ,
    B_LG_K0,
    B_LG_K1,
    B_LG_V0,
    B_LG_V1,
    B_LG_I,
    B_ALPHA,
    B_ALPHA_PART,
    B_LY_RES,
    B_LY_FIN,
    B_LY_LIST0,
    B_LY_LIST1,
    B_LY_CNT0,
    B_LY_CNT1,
    B_LY_LAM,
    B_LY_HASH,
    B_LY_POW,
    B_LY_SCAN0,     // 4 consecutive ids for the fingerprint scan partials
    B_LY_SCAN1,
    B_LY_SCAN2,
    B_LY_SCAN3,
    B_LY_HF,
    B_LY_HF2,
    B_LY_CLS,
    B_LY_VL,
    B_LY_JMP0,
    B_LY_JMP1,
    B_LY_LABH,
    B_LY_GL,
    B_LY_GF,
    B_LY_SETTLED,
    // Lyndon grammar: distinct labels, child slots, up-contexts
    B_GR_SC,        // scalars of the grammar phases (GrSlot)
    B_GR_REP,       // representative occurrence of each distinct label
    B_GR_CNT,       // occurrences of each distinct label
    B_GR_KEY,       // sorted distinct label keys (lookups)
    B_GR_OFF,       // first child slot of each label (D + 1)
    B_GR_LAST,      // label of the last child of each label
    B_GR_ROOT,      // root positions (the stack)
    B_GR_CHL,       // label of each slot's child
    B_GR_CHC,       // character before each slot's child
    B_GR_CONT,      // label owning each slot (kNone: the stack)
    B_GR_NXT,       // next slot after the slot's run
    B_GR_RUN,       // run length from each slot; later the scan of the run ends
    B_GR_CCNT,      // positions an item at each slot stands for
    B_GR_RKEY,      // dense key of each slot's run (label rank, length)
    B_GR_LRANK,     // rank of each distinct label
    B_GR_UCNT,      // up-context count / offsets per slot
    B_GR_UOFF,      // up-context list offsets per label (D + 1)
    B_GR_UK0,
    B_GR_UK1,
    B_GR_UV0,
    B_GR_UV1,
    B_GL_FA,        // label keys of one word: flags, then their positions in the short list
    B_GL_KA0,       // short label keys (high word) and positions, sorted
    B_GL_KA1,
    B_GL_VA0,
    B_GL_VA1,
    B_GL_REPA,      // distinct short labels: representative, count, key
    B_GL_CNTA,
    B_GL_UA,
    B_GL_REPB,      // distinct long labels: representative, count, key
    B_GL_CNTB,
    B_GL_UB,
    // Lyndon grammar: items of the induction (double-buffered) and per-round arrays
    B_IT_Q0,
    B_IT_Q1,
    B_IT_G0,
    B_IT_G1,
    B_IT_ID0,
    B_IT_ID1,
    B_IT_CH0,
    B_IT_CH1,
    B_IT_SQ,
    B_IT_SG,
    B_IT_SCH,
    B_IT_P,
    B_IT_NH,
    B_IT_OF,
    B_IT_GST,
    B_IT_GFI,
    B_IT_GEXP,
    B_IT_XID,
    B_IT_RF,
    B_IT_RST,
    B_IT_RLEN,
    B_IT_RCH,
    B_IT_EC,
    B_IT_HR,
    B_AS_R,
    B_AS_WBITS,
    B_AS_WSCAN,
    B_AS_PREC,
    B_CHAIN_STATUS,
    B_CHAIN_TICKET,
    B_SEL_SKETCH,
    B_SEL_PART,
    B_SEL_EPART,
    B_SEL_ROWS,
    B_SEL_HIST2,
    B_SEL_COUNTS,
    B_SEL_X,
    B_SEL_T,
    B_SEL_H1,
    B_SEL_H2,
    B_SEL_Y,
    B_SEL_DY,
    B_SEL_DH2,
    B_SEL_DH1,
    B_SEL_LOSS,
    B_SEL_P,
    B_SEL_PB,
    B_SEL_G,
    B_SEL_M1,
    B_SEL_M2,
    B_SEL_H1B,
    B_SEL_H2B,
    B_SEL_YB,
    B_SEL_XRAW,
    B_SEL_MP,
    B_SEL_MPB,
    B_SEL_NORM,
    B_SEL_PICK,
    B_SEL_YALL,
    B_COUNT_
};

// Slots of the B_SCALARS buffer.
enum ScalarSlot : int {
    SC_TRIGGERS = 0,
    SC_DISTINCT,
    SC_COLLISION,
    SC_ACTIVE,
    SC_PIDX,
    SC_ENTRIES,
    SC_TOTAL,
    SC_GROUPS,
    SC_LARGE,
    SC_SIGMA,
    SC_DICTLEN,        // two slots: 64-bit dictionary length
    SC_SPECIAL = SC_DICTLEN + 2,  // two slots: group and entry index of the primary-index entry
    SC_NGROUPS = SC_SPECIAL + 2,  // groups that receive BWT entries
    SC_COUNT_ = 16
};

// One thread stores `words` words to the mapped signal memory, then the sequence number.
static __global__ void k_signal_copy(const uint32_t* __restrict__ src, uint32_t words, uint32_t* __restrict__ sig,
                                     uint32_t seq)
{
    if (threadIdx.x != 0) return;
    for (uint32_t i = 0; i < words; i++) __hip_atomic_store(sig + i, src[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(sig + Storage::kSigSeq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <typename T>
void Storage::read_fast(T* dst, const T* dev, size_t count)
{
    static_assert(sizeof(T) % 4 == 0, "whole words");
    const uint32_t words = (uint32_t)(count * sizeof(T) / 4);
    if (words > (uint32_t)kSigWords || arena) {
        read(dst, dev, count);
        return;
    }
    uint32_t* sd = signal_dev();
    const uint32_t seq = signal_next();
    k_signal_copy<<<1, 64, 0, stream>>>(reinterpret_cast<const uint32_t*>(dev), words, sd, seq);
    MIBWT_HIP(hipGetLastError());
    uint32_t tmp[kSigWords];
    wait_signal(seq, tmp, (int)words);
    memcpy(dst, tmp, 4 * words);
}

inline Chain Storage::chain(uint32_t tiles)
{
    const size_t have = ((size_t)B_CHAIN_STATUS < bufs.size()) ? bufs[B_CHAIN_STATUS].cap : 0;
    if (chain_epoch == 0 || chain_epoch >= (1u << 30) - 1 || have < tiles * sizeof(uint64_t) + 256) {
        uint64_t* st = buf<uint64_t>(B_CHAIN_STATUS, tiles);
        uint32_t* tk = buf<uint32_t>(B_CHAIN_TICKET, 4);
        MIBWT_HIP(hipMemsetAsync(st, 0, bufs[B_CHAIN_STATUS].cap, stream));
        MIBWT_HIP(hipMemsetAsync(tk, 0, 4 * sizeof(uint32_t), stream));
        chain_epoch = 0;
    }
    chain_epoch++;
    return Chain{reinterpret_cast<uint64_t*>(bufs[B_CHAIN_STATUS].ptr), reinterpret_cast<uint32_t*>(bufs[B_CHAIN_TICKET].ptr),
                 chain_epoch};
}

}  // namespace mibwt
