// extern "C" boundary of libmibwt.so.  Every entry point validates its arguments before any
// HIP call, maps HIP failures onto the libcubwt error codes (libcubwt.cuh:32-38,
// libcubwt.cu:149-171) and never lets a C++ exception cross the ABI.
#include "common.hpp"

namespace mibwt {
int64_t pfp_bwt_run(Storage& S, const uint8_t* src, bool src_host, uint8_t* dO, int64_t n, int W, int P);
double pfp_time_scan(Storage& S, int reps, double* bytes);
double time_copy(Storage& S, uint64_t bytes, int reps);
int64_t lyndon_bwt_run(Storage& S, const uint8_t* src, bool src_host, uint8_t* dO, int64_t n);
int64_t lyndon_array_run(Storage& S, const uint8_t* src, int64_t n, uint32_t* lam_host);
void selector_choose(const uint8_t* T, int64_t n, int* W, int* P);
}  // namespace mibwt

using namespace mibwt;

namespace {

constexpr int64_t kMaxLength = (int64_t)INT32_MAX - (1 << 20);

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
    if (method == MIBWT_METHOD_PFP && (W < 1 || W > 64 || P < 1 || P > 65535)) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        uint8_t* dO = S->buf<uint8_t>(B_STAGE, (size_t)n);
        const uint8_t wrap = T[n - 1];
        int64_t idx = (method == MIBWT_METHOD_PFP) ? pfp_bwt_run(*S, T, true, dO, n, W, P)
                                                   : lyndon_bwt_run(*S, T, true, dO, n);
        if (idx < 0) return idx;
        copy_out_cubwt(*S, dO, L, wrap, n, idx);  // wrap read before L (may alias T) is written
        return idx;
    });
}

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
        (void)S->buf<uint8_t>(B_TEXT, (size_t)max_length + 8192);
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
    int64_t rc = guarded([&]() -> int64_t {
        if (S->stream) MIBWT_HIP(hipStreamSynchronize(S->stream));
        for (auto& b : S->bufs)
            if (b.ptr) MIBWT_HIP(hipFree(b.ptr));
        S->bufs.clear();
        if (S->pinned) MIBWT_HIP(hipHostFree(S->pinned));
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
    int W = 5, P = 16;
    selector_choose(T, n, &W, &P);
    return host_bwt(device_storage, T, L, n, MIBWT_METHOD_PFP, W, P, 16);
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
    if (method == MIBWT_METHOD_PFP && (W < 1 || W > 64 || P < 1 || P > 65535)) return LIBCUBWT_BAD_PARAMETER;
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipSetDevice(S->device));
        hipStream_t caller = reinterpret_cast<hipStream_t>(stream);
        hipEvent_t ev = nullptr;
        if (caller) {
            // order the storage stream after the caller's pending work
            MIBWT_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            MIBWT_HIP(hipEventRecord(ev, caller));
            MIBWT_HIP(hipStreamWaitEvent(S->stream, ev, 0));
        }
        int64_t idx = (method == MIBWT_METHOD_PFP) ? pfp_bwt_run(*S, dT, false, dO, n, W, P)
                                                   : lyndon_bwt_run(*S, dT, false, dO, n);
        if (ev) MIBWT_HIP(hipEventDestroy(ev));
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
    if (strcmp(name, "copy") != 0 && S->last_n < 1) return LIBCUBWT_BAD_PARAMETER;
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
        return (int64_t)LIBCUBWT_BAD_PARAMETER;
    });
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

int64_t mibwt_synchronize(void)
{
    return guarded([&]() -> int64_t {
        MIBWT_HIP(hipDeviceSynchronize());
        return LIBCUBWT_NO_ERROR;
    });
}

}  // extern "C"
