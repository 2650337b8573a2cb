// extern "C" boundary of libmibwt.so.  Every entry point validates its arguments before any
// HIP call, maps HIP failures onto the libcubwt error codes (libcubwt.cuh:32-38,
// libcubwt.cu:149-171) and never lets a C++ exception cross the ABI.
#include "text.hpp"

#include <algorithm>
#include <atomic>
#include <mutex>
#include <thread>

namespace mibwt {
int64_t pfp_bwt_run(Storage& S, const uint8_t* src, bool src_host, uint8_t* dO, int64_t n, int W, int P,
                    PfpParseOut* parse_out = nullptr);
double pfp_time_scan(Storage& S, int reps, double* bytes);
int64_t pfp_batch_run(Storage& S, const uint8_t* dT, const int64_t* offs, int64_t k, uint8_t* dO, int W, int P,
                      int64_t* index, uint64_t chunk_bytes);
int64_t pfp_triggers_run(Storage& S, const uint8_t* src, int64_t n, int W, int P, uint32_t* E, int64_t cap);
double time_copy(Storage& S, uint64_t bytes, int reps);
double time_radix(Storage& S, uint64_t m, int reps, double* bytes);
int64_t lyndon_bwt_run(Storage& S, const uint8_t* src, bool src_host, uint8_t* dO, int64_t n);
int64_t lyndon_array_run(Storage& S, const uint8_t* src, int64_t n, uint32_t* lam_host);
int64_t unbwt_run(Storage& S, const uint8_t* dL, uint8_t* dU, int64_t n, int64_t index, bool ref_format,
                  uint32_t* dI, int64_t rate, bool own_phases = true);
void selector_choose_device(Storage& S, const uint8_t* T, uint64_t n, int* W, int* P, float* pred,
                            uint8_t* copy = nullptr);
bool selector_load(const char* path);
bool selector_has_model();
void wp_features(Storage& S, const uint8_t* T, uint64_t n, float* X);
int wp_grid(int* W, int* P);
float mlp_grad(Storage& S, const float* params, const float* X, const float* Y, uint32_t B, float* grads);
void mlp_predict_raw(Storage& S, const float* params, const float* X, uint32_t B, float* Y);
float mlp_train(Storage& S, const float* X, const float* Y, uint32_t B, int epochs, uint32_t mb, float lr,
                uint32_t seed, const char* path, int members);
int autotune_run(Storage& S, uint64_t max_bytes, int reps, const char* path);
bool tuning_load(const char* path);
void tuning_use_table(bool on);
Tuning tuning_for(uint64_t size);
}  // namespace mibwt

using namespace mibwt;

namespace {

constexpr int64_t kMaxLength = (int64_t)INT32_MAX - (1 << 20);

// Build of the staged text with the selector's (W,P); a choice whose dictionary outgrows the
// 32-bit index range or the device memory falls back to the reference default W=5,P=16.
int64_t pfp_bwt_selected(Storage& S, uint8_t* dO, int64_t n, int W, int P)
{
    try {
        return pfp_bwt_run(S, nullptr, false, dO, n, W, P);
    } catch (const Error& e) {
        if ((e.code != MIBWT_INDEX_RANGE && e.code != LIBCUBWT_GPU_NOT_ENOUGH_MEMORY) || (W == 5 && P == 16)) throw;
        (void)hipGetLastError();
        return pfp_bwt_run(S, nullptr, false, dO, n, 5, 16);
    }
}

// Orders the storage stream after the work already enqueued on a caller's stream (nothing for
// NULL).  The event is destroyed as soon as the wait is enqueued, also when a later step of
// the entry point throws.
struct StreamOrder {
    explicit StreamOrder(Storage& S, void* stream)
    {
        hipStream_t caller = reinterpret_cast<hipStream_t>(stream);
        if (!caller) return;
        hipEvent_t ev = nullptr;
        MIBWT_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        const hipError_t e1 = hipEventRecord(ev, caller);
        const hipError_t e2 = (e1 == hipSuccess) ? hipStreamWaitEvent(S.stream, ev, 0) : e1;
        (void)hipEventDestroy(ev);
        MIBWT_HIP(e2);
    }
};

// Staged bytes per batched build: texts are combined up to this much of the staged layout
// (MIBWT_BATCH_CHUNK overrides it, in bytes; 0 builds every text alone).
uint64_t batch_chunk_bytes()
{
    uint64_t v = (uint64_t)1 << 30;
    if (const char* e = getenv("MIBWT_BATCH_CHUNK")) v = strtoull(e, nullptr, 10);
    return v;
}

template <typename F>
int64_t guarded(F&& f)
{
    try {
        return f();
    } catch (const Error& e) {
        if (getenv("MIBWT_VERBOSE")) fprintf(stderr, "mibwt: %s\n", e.what());
        (void)hipGetLastError();
        return e.code;
    } catch (const std::bad_alloc&) {
        return LIBCUBWT_NOT_ENOUGH_MEMORY;
    } catch (...) {
        return LIBCUBWT_GPU_ERROR;
    }
}

// Convert the reference format O (device) into the libcubwt format L (host):
// L[0] = T[n-1], L[1..i-1] = O[0..i-2], L[i..n-1] = O[i..n-1]  (libcubwt.cu:2242-2254).
void copy_out_cubwt(Storage& S, const uint8_t* dO, uint8_t* L, uint8_t wrap, int64_t n, int64_t index)
{
    if (index > 1)
        MIBWT_HIP(hipMemcpyAsync(L + 1, dO, (size_t)(index - 1), hipMemcpyDeviceToHost, S.stream));
    if (n > index)
        MIBWT_HIP(hipMemcpyAsync(L + index, dO + index, (size_t)(n - index), hipMemcpyDeviceToHost, S.stream));
    MIBWT_HIP(hipStreamSynchronize(S.stream));
    L[0] = wrap;
}

int64_t host_bwt(void* storage, const uint8_t* T, uint8_t* L, int64_t n, int method, int W, int P,
                 int64_t min_n)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || !T || !L || n < min_n || n > S->max_length) return LIBCUBWT_BAD_PARAMETER;
    const bool learned = (method == MIBWT_METHOD_PFP && W == 0 && P == 0);
    if (method == MIBWT_METHOD_PFP && !learned && (W < 1 || W > 64 || P < 1 || P > 65535))
        return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        uint8_t* dO = S->buf<uint8_t>(B_STAGE, (size_t)n);
        const uint8_t wrap = T[n - 1];
        int64_t idx;
        if (learned) {
            // upload once; the (W,P) choice reads the staged text, the build reuses it
            uint8_t* dT = stage_text(*S, T, true, (uint64_t)n);
            int w = 5, p = 16;
            selector_choose_device(*S, dT, (uint64_t)n, &w, &p, nullptr);
            idx = pfp_bwt_selected(*S, dO, n, w, p);
        } else {
            idx = (method == MIBWT_METHOD_PFP) ? pfp_bwt_run(*S, T, true, dO, n, W, P)
                                               : lyndon_bwt_run(*S, T, true, dO, n);
        }
        if (idx < 0) return idx;
        copy_out_cubwt(*S, dO, L, wrap, n, idx);  // wrap read before L (may alias T) is written
        return idx;
    });
}

// ---- the allocation-free entry point
//
// mibwt_pfp_bwt_device_noalloc carves every buffer of the PFP pipeline from `scratch`, a
// file-scope device array that each GPU running the library holds until the process ends.  In
// host code the array's name only stands for the symbol: its device address on the current
// device comes from hipGetSymbolAddress, once per device at the first call, and only that
// address is handed to kernels and copies.  Calls on one GPU are serialised by a mutex (they
// share the array, which keeps the last call's contents).
constexpr int64_t kNoAllocMaxLength = 1 << 20;          // texts up to 1 MiB
constexpr size_t kScratchBytes = (size_t)320 << 20;     // 320 MiB per GPU
constexpr int kNoAllocDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

struct NoAllocDevice {
    std::mutex mu;
    bool ready = false;
    uint8_t* ptr = nullptr;  // device address of `scratch` on this device
    Storage S;               // arena-mode state (host side), set up at the first call
};
NoAllocDevice g_noalloc[kNoAllocDevices];

}  // namespace

extern "C" {

int64_t libcubwt_allocate_device_storage(void** device_storage, int64_t max_length)
{
    if (!device_storage || max_length < 0 || max_length >= kMaxLength) return LIBCUBWT_BAD_PARAMETER;
    *device_storage = nullptr;
    Storage* S = new (std::nothrow) Storage();
    if (!S) return LIBCUBWT_NOT_ENOUGH_MEMORY;
    int64_t rc = guarded([&]() -> int64_t {
        MIBWT_HIP(hipGetDevice(&S->device));
        hipDeviceProp_t prop;
        MIBWT_HIP(hipGetDeviceProperties(&prop, S->device));
        MIBWT_REQUIRE(prop.warpSize == 64, LIBCUBWT_GPU_NOT_SUPPORTED, "wave64 device required");
        MIBWT_HIP(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
        S->max_length = max_length;
        // text and output staging up front, like libcubwt's arena; the phase buffers grow
        // on first use and are then reused
        (void)S->buf<uint8_t>(B_TEXT, staged_text_bytes((uint64_t)max_length));
        (void)S->buf<uint8_t>(B_STAGE, (size_t)max_length + 64);
        (void)S->host<uint64_t>(64);
        return LIBCUBWT_NO_ERROR;
    });
    if (rc != LIBCUBWT_NO_ERROR) {
        libcubwt_free_device_storage(S);
        return rc;
    }
    *device_storage = S;
    return LIBCUBWT_NO_ERROR;
}

int64_t libcubwt_free_device_storage(void* device_storage)
{
    Storage* S = reinterpret_cast<Storage*>(device_storage);
    if (!S) return LIBCUBWT_BAD_PARAMETER;
    for (Storage* sub : S->pool) libcubwt_free_device_storage(sub);
    S->pool.clear();
    int64_t rc = guarded([&]() -> int64_t {
        if (S->stream) MIBWT_HIP(hipStreamSynchronize(S->stream));
        for (auto& b : S->bufs)
            if (b.ptr) MIBWT_HIP(hipFree(b.ptr));
        S->bufs.clear();
        if (S->pinned) MIBWT_HIP(hipHostFree(S->pinned));
        if (S->sig) MIBWT_HIP(hipHostFree(S->sig));
        for (auto& p : S->phases) {
            if (p.start) MIBWT_HIP(hipEventDestroy(p.start));
            if (p.stop) MIBWT_HIP(hipEventDestroy(p.stop));
        }
        if (S->t_begin) MIBWT_HIP(hipEventDestroy(S->t_begin));
        if (S->t_end) MIBWT_HIP(hipEventDestroy(S->t_end));
        if (S->stream) MIBWT_HIP(hipStreamDestroy(S->stream));
        return LIBCUBWT_NO_ERROR;
    });
    delete S;
    return rc;
}

int64_t libcubwt_bwt(void* device_storage, const uint8_t* T, uint8_t* L, int64_t n)
{
    Storage* S = reinterpret_cast<Storage*>(device_storage);
    if (!S || !T || !L || n < 16 || n > S->max_length) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        uint8_t* dO = S->buf<uint8_t>(B_STAGE, (size_t)n);
        const uint8_t wrap = T[n - 1];
        // upload once; the (W,P) choice reads the staged text, the build reuses it
        uint8_t* dT = stage_text(*S, T, true, (uint64_t)n);
        int W = 5, P = 16;
        selector_choose_device(*S, dT, (uint64_t)n, &W, &P, nullptr);
        const int64_t idx = pfp_bwt_selected(*S, dO, n, W, P);
        copy_out_cubwt(*S, dO, L, wrap, n, idx);
        return idx;
    });
}

int64_t libcubwt_bwt_aux(void* device_storage, const uint8_t* T, uint8_t* L, int64_t n, int64_t r, uint32_t* I)
{
    Storage* S = reinterpret_cast<Storage*>(device_storage);
    if (!S || !T || !L || !I || n < 16 || n > S->max_length) return LIBCUBWT_BAD_PARAMETER;
    // libcubwt.cu:2437: r is a power of two >= 4.  Its further limit (n + r - 1) / r <= 255 is
    // not applied: any number of samples fits the uint32 rows of I here.
    if (r < 4 || (r & (r - 1)) != 0) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        uint8_t* dO = S->buf<uint8_t>(B_STAGE, (size_t)n);
        const uint8_t wrap = T[n - 1];
        uint8_t* dT = stage_text(*S, T, true, (uint64_t)n);
        int W = 5, P = 16;
        selector_choose_device(*S, dT, (uint64_t)n, &W, &P, nullptr);
        const int64_t idx = pfp_bwt_selected(*S, dO, n, W, P);
        // sampled rows from the LF walk over the BWT just built (still in HBM)
        const int64_t cnt = (n - 1) / r + 1;
        uint32_t* dI = S->buf<uint32_t>(B_UN_AUX, (size_t)cnt);
        const int64_t rc = unbwt_run(*S, dO, nullptr, n, idx, true, dI, r);
        if (rc != LIBCUBWT_NO_ERROR) return rc;
        MIBWT_HIP(hipMemcpyAsync(I, dI, sizeof(uint32_t) * cnt, hipMemcpyDeviceToHost, S->stream));
        copy_out_cubwt(*S, dO, L, wrap, n, idx);
        return (int64_t)LIBCUBWT_NO_ERROR;
    });
}

int64_t libcubwt_unbwt(void* device_storage, const uint8_t* T, uint8_t* U, int64_t n, const int32_t* freq, int32_t i)
{
    (void)freq;  // symbol counts are taken on the device
    Storage* S = reinterpret_cast<Storage*>(device_storage);
    if (!S || !T || !U || n < 1 || n > S->max_length || i < 1 || i > n) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        uint8_t* dL = S->buf<uint8_t>(B_UN_IO, (size_t)n);
        uint8_t* dU = S->buf<uint8_t>(B_STAGE, (size_t)n);
        MIBWT_HIP(hipMemcpyAsync(dL, T, (size_t)n, hipMemcpyHostToDevice, S->stream));
        const int64_t rc = unbwt_run(*S, dL, dU, n, i, false, nullptr, 0);
        if (rc != LIBCUBWT_NO_ERROR) return rc;
        MIBWT_HIP(hipMemcpyAsync(U, dU, (size_t)n, hipMemcpyDeviceToHost, S->stream));
        MIBWT_HIP(hipStreamSynchronize(S->stream));
        return (int64_t)LIBCUBWT_NO_ERROR;
    });
}

int64_t mibwt_unbwt_device(void* storage, const uint8_t* dO, uint8_t* dU, int64_t n, int64_t index, void* stream)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || !dO || !dU || n < 1 || n > S->max_length || index < 1 || index > n) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        StreamOrder order(*S, stream);
        const int64_t rc = unbwt_run(*S, dO, dU, n, index, true, nullptr, 0);
        MIBWT_HIP(hipStreamSynchronize(S->stream));
        return rc;
    });
}

int64_t mibwt_pfp_bwt_device_noalloc(const uint8_t* dT, uint8_t* dO, int64_t n, int32_t W, int32_t P, void* stream)
{
    if (!dT || !dO || n < 1 || W < 1 || W > 64 || P < 1 || P > 65535) return LIBCUBWT_BAD_PARAMETER;
    if (n > kNoAllocMaxLength) {
        if (getenv("MIBWT_VERBOSE"))
            fprintf(stderr, "mibwt: n = %lld exceeds the %lld bytes of the allocation-free entry\n", (long long)n,
                    (long long)kNoAllocMaxLength);
        return LIBCUBWT_BAD_PARAMETER;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocDevices) return LIBCUBWT_GPU_NOT_SUPPORTED;
    NoAllocDevice& D = g_noalloc[dev];
    std::lock_guard<std::mutex> lock(D.mu);
    const int64_t rc = guarded([&]() -> int64_t {
        if (!D.ready) {
            // set-up on this device, with it current: the device address of its copy of `scratch`
            void* p = nullptr;
            MIBWT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            MIBWT_REQUIRE(p != nullptr, LIBCUBWT_GPU_ERROR, "no device address for the scratch array");
            D.ptr = reinterpret_cast<uint8_t*>(p);
            D.S.device = dev;
            D.S.bufs.resize(B_COUNT_);
            D.ready = true;
        }
        D.S.stream = reinterpret_cast<hipStream_t>(stream);
        D.S.max_length = kNoAllocMaxLength;
        D.S.arena_reset(D.ptr, kScratchBytes);
        return pfp_bwt_run(D.S, dT, false, dO, n, W, P);
    });
    if (getenv("MIBWT_VERBOSE"))
        fprintf(stderr, "mibwt: allocation-free build n = %lld: %zu of %zu scratch bytes\n", (long long)n,
                D.S.arena_peak, kScratchBytes);
    return rc;
}

int64_t mibwt_pfp_bwt(void* storage, const uint8_t* T, uint8_t* L, int64_t n, int32_t W, int32_t P)
{
    return host_bwt(storage, T, L, n, MIBWT_METHOD_PFP, W, P, 1);
}

int64_t mibwt_lyndon_bwt(void* storage, const uint8_t* T, uint8_t* L, int64_t n)
{
    return host_bwt(storage, T, L, n, MIBWT_METHOD_LYNDON, 5, 16, 1);
}

static int64_t device_bwt(void* storage, const uint8_t* dT, uint8_t* dO, int64_t n, int method, int W,
                          int P, void* stream)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || !dT || !dO || n < 1 || n > S->max_length) return LIBCUBWT_BAD_PARAMETER;
    const bool learned = (method == MIBWT_METHOD_PFP && W == 0 && P == 0);
    if (method == MIBWT_METHOD_PFP && !learned && (W < 1 || W > 64 || P < 1 || P > 65535))
        return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        StreamOrder order(*S, stream);  // after the caller's pending work
        int64_t idx;
        if (learned) {
            // the selector's pass over the caller's text also fills the private padded copy
            // that the build reads
            const uint64_t bytes = staged_text_bytes((uint64_t)n);
            uint8_t* T = S->buf<uint8_t>(B_TEXT, bytes);
            MIBWT_HIP(hipMemsetAsync(T + n, 0, bytes - (uint64_t)n, S->stream));
            int w = 5, p = 16;
            selector_choose_device(*S, dT, (uint64_t)n, &w, &p, nullptr, T);
            idx = pfp_bwt_selected(*S, dO, n, w, p);
        } else if (method == MIBWT_METHOD_PFP) {
            idx = pfp_bwt_run(*S, dT, false, dO, n, W, P);
        } else {
            idx = lyndon_bwt_run(*S, dT, false, dO, n);
        }
        return idx;
    });
}

int64_t mibwt_pfp_bwt_device(void* storage, const uint8_t* dT, uint8_t* dO, int64_t n, int32_t W, int32_t P,
                             void* stream)
{
    return device_bwt(storage, dT, dO, n, MIBWT_METHOD_PFP, W, P, stream);
}

int64_t mibwt_lyndon_bwt_device(void* storage, const uint8_t* dT, uint8_t* dO, int64_t n, void* stream)
{
    return device_bwt(storage, dT, dO, n, MIBWT_METHOD_LYNDON, 5, 16, stream);
}

int64_t mibwt_pfp_triggers(void* storage, const uint8_t* T, int64_t n, int32_t W, int32_t P, uint32_t* E,
                           int64_t cap)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || !T || (!E && cap > 0) || cap < 0 || n < 1 || n > S->max_length) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        return pfp_triggers_run(*S, T, n, W, P, E, cap);
    });
}

int64_t mibwt_pfp_parse(void* storage, const uint8_t* T, int64_t n, int32_t W, int32_t P, uint32_t* rank,
                        uint32_t* psa, int64_t cap, int64_t* distinct)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || !T || ((!rank || !psa) && cap > 0) || cap < 0 || n < 1 || n > S->max_length) return LIBCUBWT_BAD_PARAMETER;
    if (W < 1 || W > 64 || P < 1 || P > 65535) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        PfpParseOut out{rank, psa, cap, 0};
        uint8_t* dO = S->buf<uint8_t>(B_STAGE, (size_t)n);
        const int64_t d = pfp_bwt_run(*S, T, true, dO, n, W, P, &out);
        if (distinct) *distinct = out.distinct;
        return d;
    });
}

int64_t mibwt_lyndon_array(void* storage, const uint8_t* T, int64_t n, uint32_t* lam)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || !T || !lam || n < 1 || n > S->max_length) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        return lyndon_array_run(*S, T, n, lam);
    });
}

int64_t mibwt_last_phase_stats(void* storage, struct mibwt_phase_stats* out)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || !out) return LIBCUBWT_BAD_PARAMETER;
    *out = S->stats;
    return LIBCUBWT_NO_ERROR;
}

int64_t mibwt_set_kernel_timing(void* storage, const char* kernel)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S) return LIBCUBWT_BAD_PARAMETER;
    if (!kernel) S->kt_family = 0;
    else if (strcmp(kernel, "radix_scatter") == 0) S->kt_family = 1;
    else if (strcmp(kernel, "trig_mask") == 0) S->kt_family = 2;
    else return LIBCUBWT_BAD_PARAMETER;
    return LIBCUBWT_NO_ERROR;
}

int64_t mibwt_time_kernel(void* storage, const char* name, int32_t reps, double* ms, double* bytes_per_launch)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || !name || reps < 1 || !ms || !bytes_per_launch) return LIBCUBWT_BAD_PARAMETER;
    if (strcmp(name, "copy") != 0 && strcmp(name, "radix") != 0 && S->last_n < 1) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        if (strcmp(name, "scan") == 0) {
            *ms = pfp_time_scan(*S, reps, bytes_per_launch);
            return LIBCUBWT_NO_ERROR;
        }
        if (strcmp(name, "copy") == 0) {
            const uint64_t bytes = (uint64_t)(*bytes_per_launch) & ~15ull;
            if (bytes < 16) return (int64_t)LIBCUBWT_BAD_PARAMETER;
            *ms = time_copy(*S, bytes, reps);
            *bytes_per_launch = (double)bytes;
            return LIBCUBWT_NO_ERROR;
        }
        if (strcmp(name, "radix") == 0) {
            const double items = *bytes_per_launch;
            if (!(items >= 1 && items < 2147483648.0)) return (int64_t)LIBCUBWT_BAD_PARAMETER;
            *ms = time_radix(*S, (uint64_t)items, reps, bytes_per_launch);
            return LIBCUBWT_NO_ERROR;
        }
        return (int64_t)LIBCUBWT_BAD_PARAMETER;
    });
}

int64_t mibwt_batch_bwt(void* storage, const uint8_t* const* T, uint8_t* const* L, const int64_t* n, int64_t k,
                        int64_t* index, int32_t method, int32_t W, int32_t P, int32_t workers)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || !T || !L || !n || !index || k < 0 || workers < 0 || workers > 16) return LIBCUBWT_BAD_PARAMETER;
    if (method != MIBWT_METHOD_PFP && method != MIBWT_METHOD_LYNDON) return LIBCUBWT_BAD_PARAMETER;
    if (method == MIBWT_METHOD_PFP && W != 0 && (W < 1 || W > 64 || P < 1 || P > 65535)) return LIBCUBWT_BAD_PARAMETER;
    for (int64_t i = 0; i < k; i++)
        if (!T[i] || !L[i] || n[i] < 1 || n[i] > S->max_length) return LIBCUBWT_BAD_PARAMETER;
    if (k == 0) return LIBCUBWT_NO_ERROR;
    if (method == MIBWT_METHOD_PFP) {
        // PFP: consecutive texts are combined into builds of up to batch_chunk_bytes() staged
        // bytes; per build the texts go up as one dense device array, the batched pipeline runs
        // on it, and every BWT comes back in libcubwt format
        for (int64_t i = 0; i < k; i++) index[i] = LIBCUBWT_GPU_ERROR;
        return guarded([&]() -> int64_t {
            MIBWT_HIP(hipSetDevice(S->device));
            const uint64_t chunk = (W > 56) ? 0 : batch_chunk_bytes();  // wider windows: one text per build
            int64_t a = 0;
            while (a < k) {
                int64_t b = a;
                uint64_t bytes = 0;
                while (b < k && (b == a || bytes + (uint64_t)n[b] <= chunk)) bytes += (uint64_t)n[b++];
                std::vector<int64_t> offs((size_t)(b - a + 1), 0);
                for (int64_t i = a; i < b; i++) offs[(size_t)(i - a + 1)] = offs[(size_t)(i - a)] + n[i];
                uint8_t* din = S->buf<uint8_t>(B_BATCH_IN, (size_t)bytes);
                uint8_t* dout = S->buf<uint8_t>(B_BATCH_OUT, (size_t)bytes);
                for (int64_t i = a; i < b; i++)
                    MIBWT_HIP(hipMemcpyAsync(din + offs[(size_t)(i - a)], T[i], (size_t)n[i], hipMemcpyHostToDevice,
                                             S->stream));
                std::vector<uint8_t> wrap((size_t)(b - a));
                for (int64_t i = a; i < b; i++) wrap[(size_t)(i - a)] = T[i][n[i] - 1];  // before L (may alias T)
                pfp_batch_run(*S, din, offs.data(), b - a, dout, W, P, index + a, chunk);
                for (int64_t i = a; i < b; i++) {
                    // libcubwt format: L[0] = T[n-1], L[1..i-1] = O[0..i-2], L[i..n-1] = O[i..n-1]
                    const int64_t idx = index[i];
                    const uint8_t* O = dout + offs[(size_t)(i - a)];
                    if (idx > 1)
                        MIBWT_HIP(hipMemcpyAsync(L[i] + 1, O, (size_t)(idx - 1), hipMemcpyDeviceToHost, S->stream));
                    if (n[i] > idx)
                        MIBWT_HIP(hipMemcpyAsync(L[i] + idx, O + idx, (size_t)(n[i] - idx), hipMemcpyDeviceToHost,
                                                 S->stream));
                }
                MIBWT_HIP(hipStreamSynchronize(S->stream));
                for (int64_t i = a; i < b; i++) L[i][0] = wrap[(size_t)(i - a)];
                a = b;
            }
            return LIBCUBWT_NO_ERROR;
        });
    }
    // Lyndon: a pool of worker storages; default 8: with 256 mixed texts of 64 KiB-16 MiB in
    // host buffers, 4 workers reach 2.2 GB/s, 8 reach 2.55 GB/s and 16 reach 2.57 GB/s
    if (workers == 0) workers = 8;
    // one storage (stream) per worker, created on first use and kept by the parent storage
    int64_t rc = guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        while ((int)S->pool.size() < workers) {
            void* sub = nullptr;
            const int64_t r = libcubwt_allocate_device_storage(&sub, S->max_length);
            if (r != LIBCUBWT_NO_ERROR) return r;
            S->pool.push_back(reinterpret_cast<Storage*>(sub));
        }
        return LIBCUBWT_NO_ERROR;
    });
    if (rc != LIBCUBWT_NO_ERROR) return rc;
    // longest texts first, handed out through one atomic counter
    std::vector<int64_t> order((size_t)k);
    for (int64_t i = 0; i < k; i++) order[(size_t)i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return n[a] > n[b]; });
    std::atomic<int64_t> next{0};
    auto work = [&](Storage* sub) {
        (void)hipSetDevice(sub->device);
        for (;;) {
            const int64_t q = next.fetch_add(1);
            if (q >= k) break;
            const int64_t i = order[(size_t)q];
            if (method == MIBWT_METHOD_LYNDON) {
                index[i] = host_bwt(sub, T[i], L[i], n[i], MIBWT_METHOD_LYNDON, 5, 16, 1);
            } else if (W != 0) {
                index[i] = host_bwt(sub, T[i], L[i], n[i], MIBWT_METHOD_PFP, W, P, 1);
            } else {
                index[i] = guarded([&]() -> int64_t {
                    uint8_t* dO = sub->buf<uint8_t>(B_STAGE, (size_t)n[i]);
                    const uint8_t wrap = T[i][n[i] - 1];
                    uint8_t* dT = stage_text(*sub, T[i], true, (uint64_t)n[i]);
                    int w = 5, p = 16;
                    selector_choose_device(*sub, dT, (uint64_t)n[i], &w, &p, nullptr);
                    const int64_t idx = pfp_bwt_selected(*sub, dO, n[i], w, p);
                    copy_out_cubwt(*sub, dO, L[i], wrap, n[i], idx);
                    return idx;
                });
            }
        }
    };
    std::vector<std::thread> th;
    for (int w = 0; w < workers; w++) th.emplace_back(work, S->pool[(size_t)w]);
    for (auto& t : th) t.join();
    for (int64_t i = 0; i < k; i++)
        if (index[i] < 0) return index[i];
    return LIBCUBWT_NO_ERROR;
}

int64_t mibwt_batch_bwt_device(void* storage, const uint8_t* dT, const int64_t* offsets, int64_t k, uint8_t* dO,
                               int64_t* index, int32_t W, int32_t P, void* stream)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || !dT || !dO || !offsets || !index || k < 0) return LIBCUBWT_BAD_PARAMETER;
    if (!(W == 0 && P == 0) && (W < 1 || W > 64 || P < 1 || P > 65535)) return LIBCUBWT_BAD_PARAMETER;
    for (int64_t i = 0; i < k; i++) {
        const int64_t len = offsets[i + 1] - offsets[i];
        if (offsets[i] < 0 || len < 1 || len > S->max_length) return LIBCUBWT_BAD_PARAMETER;
    }
    if (k == 0) return LIBCUBWT_NO_ERROR;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        StreamOrder order(*S, stream);
        // windows wider than 56 bytes build every text alone
        const uint64_t chunk = (W > 56) ? 0 : batch_chunk_bytes();
        pfp_batch_run(*S, dT, offsets, k, dO, W, P, index, chunk);
        MIBWT_HIP(hipStreamSynchronize(S->stream));
        return LIBCUBWT_NO_ERROR;
    });
}

int64_t mibwt_select_wp_device(void* storage, const uint8_t* dT, int64_t n, int32_t* W, int32_t* P, float* pred)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || !dT || !W || !P || n < 1 || n > S->max_length) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        uint8_t* T = stage_text(*S, dT, false, (uint64_t)n);
        int w = 5, p = 16;
        selector_choose_device(*S, T, (uint64_t)n, &w, &p, pred);
        *W = w;
        *P = p;
        return selector_has_model() ? 1 : 0;
    });
}

int64_t mibwt_select_wp(const uint8_t* T, int64_t n, int32_t* W, int32_t* P)
{
    if (!T || !W || !P || n < 1) return LIBCUBWT_BAD_PARAMETER;
    void* s = nullptr;
    int64_t rc = libcubwt_allocate_device_storage(&s, n);
    if (rc != LIBCUBWT_NO_ERROR) return rc;
    Storage* S = reinterpret_cast<Storage*>(s);
    rc = guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        uint8_t* dT = stage_text(*S, T, true, (uint64_t)n);
        int w = 5, p = 16;
        selector_choose_device(*S, dT, (uint64_t)n, &w, &p, nullptr);
        *W = w;
        *P = p;
        return selector_has_model() ? 1 : 0;
    });
    libcubwt_free_device_storage(s);
    return rc;
}

int64_t mibwt_wp_features(void* storage, const uint8_t* dT, int64_t n, float* X)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || !dT || !X || n < 1 || n > S->max_length) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        uint8_t* T = stage_text(*S, dT, false, (uint64_t)n);
        wp_features(*S, T, (uint64_t)n, X);
        return LIBCUBWT_NO_ERROR;
    });
}

int64_t mibwt_wp_grid(int32_t* W, int32_t* P)
{
    if (!W || !P) return LIBCUBWT_BAD_PARAMETER;
    int w[64], p[64];
    const int count = wp_grid(w, p);
    for (int i = 0; i < count; i++) {
        W[i] = w[i];
        P[i] = p[i];
    }
    return count;
}

int64_t mibwt_selector_load(const char* path)
{
    if (!path) return LIBCUBWT_BAD_PARAMETER;
    return selector_load(path) ? LIBCUBWT_NO_ERROR : LIBCUBWT_BAD_PARAMETER;
}

int64_t mibwt_mlp_grad(void* storage, const float* params, const float* X, const float* Y, int64_t B, float* grads,
                       float* loss)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || !params || !X || !Y || !grads || !loss || B < 1 || B > (1 << 24)) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        *loss = mlp_grad(*S, params, X, Y, (uint32_t)B, grads);
        return LIBCUBWT_NO_ERROR;
    });
}

int64_t mibwt_mlp_predict(void* storage, const float* params, const float* X, int64_t B, float* Y)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || !params || !X || !Y || B < 1 || B > (1 << 24)) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        mlp_predict_raw(*S, params, X, (uint32_t)B, Y);
        return LIBCUBWT_NO_ERROR;
    });
}

int64_t mibwt_mlp_train(void* storage, const float* X, const float* Y, int64_t B, int32_t epochs, int32_t batch,
                        float lr, uint32_t seed, const char* out_path, float* final_loss)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || !X || !Y || B < 2 || B > (1 << 24) || epochs < 1 || batch < 0 || !(lr > 0)) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        const float l = mlp_train(*S, X, Y, (uint32_t)B, epochs, (uint32_t)batch, lr, seed, out_path, 1);
        if (final_loss) *final_loss = l;
        return LIBCUBWT_NO_ERROR;
    });
}

int64_t mibwt_mlp_train_ensemble(void* storage, const float* X, const float* Y, int64_t B, int32_t epochs,
                                 int32_t batch, float lr, uint32_t seed, int32_t members, const char* out_path,
                                 float* final_loss)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || !X || !Y || B < 2 || B > (1 << 24) || epochs < 1 || batch < 0 || !(lr > 0) || members < 1 ||
        members > 16)
        return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        const float l = mlp_train(*S, X, Y, (uint32_t)B, epochs, (uint32_t)batch, lr, seed, out_path, members);
        if (final_loss) *final_loss = l;
        return LIBCUBWT_NO_ERROR;
    });
}

int64_t mibwt_autotune(void* storage, int64_t max_bytes, int32_t reps, const char* out_path)
{
    Storage* S = reinterpret_cast<Storage*>(storage);
    if (!S || max_bytes < (1 << 16) || max_bytes > S->max_length || reps < 1 || reps > 16) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        return autotune_run(*S, (uint64_t)max_bytes, reps, out_path);
    });
}

int64_t mibwt_tuning_load(const char* path)
{
    if (!path) return LIBCUBWT_BAD_PARAMETER;
    return tuning_load(path) ? LIBCUBWT_NO_ERROR : LIBCUBWT_BAD_PARAMETER;
}

int64_t mibwt_tuning_use(int32_t on)
{
    tuning_use_table(on != 0);
    return LIBCUBWT_NO_ERROR;
}

int64_t mibwt_tuning_get(int64_t size, int32_t* shape)
{
    if (!shape || size < 1) return LIBCUBWT_BAD_PARAMETER;
    const Tuning t = tuning_for((uint64_t)size);
    shape[0] = t.scan_sub;
    shape[1] = t.sort_items;
    shape[2] = t.chain_items;
    shape[3] = t.key_bits;
    return LIBCUBWT_NO_ERROR;
}

int64_t mibwt_device_alloc(void** ptr, int64_t bytes)
{
    if (!ptr || bytes < 0) return LIBCUBWT_BAD_PARAMETER;
    *ptr = nullptr;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipMalloc(ptr, (size_t)(bytes > 0 ? bytes : 1)));
        return LIBCUBWT_NO_ERROR;
    });
}

int64_t mibwt_device_free(void* ptr)
{
    if (!ptr) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipFree(ptr));
        return LIBCUBWT_NO_ERROR;
    });
}

int64_t mibwt_copy(void* dst, const void* src, int64_t bytes, int32_t kind)
{
    if (!dst || !src || bytes < 0 || kind < 0 || kind > 2) return LIBCUBWT_BAD_PARAMETER;
    if (bytes == 0) return LIBCUBWT_NO_ERROR;
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                        : hipMemcpyDeviceToDevice;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipMemcpy(dst, src, (size_t)bytes, k));
        return LIBCUBWT_NO_ERROR;
    });
}

int64_t mibwt_stream_create(void** stream)
{
    if (!stream) return LIBCUBWT_BAD_PARAMETER;
    *stream = nullptr;
    return guarded([&]() -> int64_t {
        hipStream_t s = nullptr;
        MIBWT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        *stream = s;
        return LIBCUBWT_NO_ERROR;
    });
}

int64_t mibwt_stream_destroy(void* stream)
{
    if (!stream) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)));
        MIBWT_HIP(hipStreamDestroy(reinterpret_cast<hipStream_t>(stream)));
        return LIBCUBWT_NO_ERROR;
    });
}

int64_t mibwt_copy_async(void* dst, const void* src, int64_t bytes, int32_t kind, void* stream)
{
    if (!dst || !src || bytes < 0 || kind < 0 || kind > 2) return LIBCUBWT_BAD_PARAMETER;
    if (bytes == 0) return LIBCUBWT_NO_ERROR;
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                        : hipMemcpyDeviceToDevice;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, k, reinterpret_cast<hipStream_t>(stream)));
        return LIBCUBWT_NO_ERROR;
    });
}

int64_t mibwt_synchronize(void)
{
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipDeviceSynchronize());
        return LIBCUBWT_NO_ERROR;
    });
}

}  // extern "C"
