/*
 * mibwt — MI355X-native Burrows-Wheeler transform builder (HIP, gfx950).
 *
 * C ABI.  Host pointers unless a function name ends in _device.  No C++ or HIP types cross
 * this boundary; streams are passed as opaque pointers (a hipStream_t, or NULL for the
 * storage's own stream).
 *
 * Two conventions for the output:
 *   - libcubwt format (libcubwt.cuh:69-77, libcubwt.cu:2242-2254): L[0] = T[n-1],
 *     L[1..i-1] and L[i..n-1] are the BWT rows with the "$" row removed, and the return
 *     value is the primary index i >= 1;
 *   - reference format (Prefix_Free_Parsing_W=5_P=16_kommentiert.cu:4776-4790):
 *     O[j] = T[(SA[j]-1) mod n] over the n non-empty suffixes in $-order.
 *     O = L[1..i-1] ++ [L[0]] ++ L[i..n-1] (libcubwt.cu:3003-3015).
 */
#ifndef MIBWT_H
#define MIBWT_H 1

#include <stdint.h>

#define LIBCUBWT_VERSION_MAJOR 1
#define LIBCUBWT_VERSION_MINOR 6
#define LIBCUBWT_VERSION_PATCH 3
#define LIBCUBWT_VERSION_STRING "1.6.3"

/* Error codes: identical values to libcubwt.cuh:32-38. */
#define LIBCUBWT_NO_ERROR 0
#define LIBCUBWT_BAD_PARAMETER -1
#define LIBCUBWT_NOT_ENOUGH_MEMORY -2
#define LIBCUBWT_GPU_ERROR -7
#define LIBCUBWT_GPU_NOT_SUPPORTED -8
#define LIBCUBWT_GPU_NOT_ENOUGH_MEMORY -9
/* mibwt-specific: a 64-bit phrase fingerprint collision survived every re-seed. */
#define MIBWT_HASH_COLLISION -20
/* mibwt-specific: the text's dictionary (sum of the distinct phrase lengths for the chosen W,P)
 * or its entry count exceeds the 32-bit index range of the device pipeline. */
#define MIBWT_INDEX_RANGE -21

/* BWT methods for mibwt_bwt / mibwt_batch_bwt. */
#define MIBWT_METHOD_PFP 0
#define MIBWT_METHOD_LYNDON 1

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- libcubwt drop-in */

/* Replaces libcubwt_allocate_device_storage (libcubwt.cuh:60, libcubwt.cu:2256).
 * Creates a storage bound to the current HIP device with its own stream.  Device buffers
 * are sized for max_length and grow on demand.  BAD_PARAMETER if device_storage is NULL or
 * max_length >= 2^31 - 2^20 (32-bit text positions). */
int64_t libcubwt_allocate_device_storage(void** device_storage, int64_t max_length);

/* Replaces libcubwt_free_device_storage (libcubwt.cuh:67, libcubwt.cu:2394). */
int64_t libcubwt_free_device_storage(void* device_storage);

/* Replaces libcubwt_bwt (libcubwt.cuh:77, libcubwt.cu:2414-2431): host T[0..n) -> host
 * L[0..n) in libcubwt format (L may alias T); returns the primary index or an error code.
 * 16 <= n <= max_length, as in libcubwt.  Computed by prefix-free parsing with the (W,P)
 * from mibwt_select_wp (fixed W=5,P=16 when the selector is not trained). */
int64_t libcubwt_bwt(void* device_storage, const uint8_t* T, uint8_t* L, int64_t n);

/* Sampled suffix-array inverse together with the BWT (libcubwt.cuh:79-89): L as libcubwt_bwt,
 * I[j] = the row of the rotation starting at text position j*r, in the convention of the
 * primary index (I[0] is the primary index); r is a power of two >= 4 (BAD_PARAMETER
 * otherwise, as libcubwt.cu:2437), I has (n-1)/r + 1 entries.  Returns LIBCUBWT_NO_ERROR or an
 * error code.  (The reference's libcubwt.cu binds it at :2433-2450.)  Deliberate difference:
 * libcubwt.cu:2437 also rejects (n + r - 1) / r > 255 samples; here any count is accepted. */
int64_t libcubwt_bwt_aux(void* device_storage, const uint8_t* T, uint8_t* L, int64_t n, int64_t r,
                         uint32_t* I);

/* Inverse BWT (libcubwt.cuh:91-101; declared but not defined by the reference's libcubwt.cu):
 * T is a BWT in libcubwt format with primary index i (1 <= i <= n), U receives the text (U may
 * alias T).  freq is accepted and not needed.  Returns LIBCUBWT_NO_ERROR, or
 * LIBCUBWT_BAD_PARAMETER when T is not a BWT with primary index i. */
int64_t libcubwt_unbwt(void* device_storage, const uint8_t* T, uint8_t* U, int64_t n, const int32_t* freq,
                       int32_t i);

/* ---------------------------------------------------------------- mibwt extensions */

/* Prefix-free-parsing BWT with runtime window W (1..64) and modulus P (1..65535), or
 * W = P = 0 for the learned selector's choice.  Replaces the compile-time variants
 * Prefix_Free_Parsing_W=<W>_P=<P>.cu (main at Prefix_Free_Parsing_W=5_P=16_kommentiert.cu:
 * 4793-4815, Compute_D :2415, Compute_S :4330) and the hand-written per-file (W,P) table of
 * Benchmark_Skript_Linux.sh:14-73.  Host buffers, libcubwt format; 1 <= n <= max_length. */
int64_t mibwt_pfp_bwt(void* storage, const uint8_t* T, uint8_t* L, int64_t n, int32_t W,
                      int32_t P);

/* Same on device buffers already resident in HBM: dT[0..n) -> dO[0..n) in REFERENCE
 * format (O), on the storage's stream (stream == NULL) or on `stream`.  Synchronous.
 * W = P = 0 lets the learned selector choose (W,P) on the resident text first (the
 * choice is reported by mibwt_last_phase_stats).  Returns the primary index. */
int64_t mibwt_pfp_bwt_device(void* storage, const uint8_t* dT, uint8_t* dO, int64_t n,
                             int32_t W, int32_t P, void* stream);

/* mibwt_pfp_bwt_device with fixed (W, P) that allocates nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  No storage: every buffer comes from
 * one file-scope device array of 320 MiB that each GPU running the library holds until the
 * process ends; texts of 1 <= n <= 1 MiB (BAD_PARAMETER beyond), W in 1..64, P in 1..65535
 * (no learned choice).  A (W, P) whose dictionary outgrows the array returns
 * LIBCUBWT_GPU_NOT_ENOUGH_MEMORY.  Calls on one GPU share the array and are serialised within a
 * process; the host side uses no pinned or mapped memory.  Runs on `stream` (NULL: the null
 * stream), synchronous; dT[0..n) -> dO[0..n) in reference format; returns the primary index. */
int64_t mibwt_pfp_bwt_device_noalloc(const uint8_t* dT, uint8_t* dO, int64_t n, int32_t W, int32_t P,
                                     void* stream);

/* Inverse BWT of a device BWT in REFERENCE format (the output of the *_device entries):
 * dO[0..n) with primary index -> dU[0..n).  Synchronous.  Returns LIBCUBWT_NO_ERROR or
 * LIBCUBWT_BAD_PARAMETER (no reference counterpart: the on-device round-trip check). */
int64_t mibwt_unbwt_device(void* storage, const uint8_t* dO, uint8_t* dU, int64_t n, int64_t index,
                           void* stream);

/* Lyndon-grammar BWT (Lyndon_Grammar_BWT.cu main :5547-5576: build_Lyndon_Grammar :2592,
 * Sort_Lyndon_Grammar :3621, Build_BWT :4628).  Host buffers, libcubwt format. */
int64_t mibwt_lyndon_bwt(void* storage, const uint8_t* T, uint8_t* L, int64_t n);

/* Device-buffer variant, reference format, like mibwt_pfp_bwt_device. */
int64_t mibwt_lyndon_bwt_device(void* storage, const uint8_t* dT, uint8_t* dO, int64_t n,
                                void* stream);

/* Trigger end positions of host T[0..n) for window W and modulus P: the ends e (W <= e < n,
 * ascending) of the windows T[e-W+1..e] whose value sum_k T[e-W+1+k]*2^(W-1-k) mod P is 0 or
 * that equal T[0..W-1] (the reference's Compute_D1/Compute_D2 output,
 * Prefix_Free_Parsing_W=5_P=16_kommentiert.cu:443-571).  E receives the first min(t, cap)
 * positions; returns t or an error code. */
int64_t mibwt_pfp_triggers(void* storage, const uint8_t* T, int64_t n, int32_t W, int32_t P,
                           uint32_t* E, int64_t cap);

/* Intermediate results of the prefix-free parse of host T[0..n) for (W, P): rank[k] = dense
 * lexicographic rank of the phrase of occurrence k (the reference's phrase order,
 * Prefix_Free_Parsing_W=5_P=16_kommentiert.cu:1990-2164 and :616-691) and psa = the suffix array
 * of the parse (:2342-2412), first min(d, cap) entries each; *distinct = number of distinct
 * phrases.  Returns the parse length d or an error code. */
int64_t mibwt_pfp_parse(void* storage, const uint8_t* T, int64_t n, int32_t W, int32_t P, uint32_t* rank,
                        uint32_t* psa, int64_t cap, int64_t* distinct);

/* Lyndon array of host T[0..n): lam[i] = length of the longest Lyndon word starting at i,
 * computed by the block-local passes and the merge tree of the Lyndon pipeline (the forest
 * the reference builds in Compute_Lyndon_Grammar_Local1/2, Lyndon_Grammar_BWT.cu:647-1868,
 * and merges in Compute_Lyndon_Grammar_Global :1910-2062).  Returns the number of merge
 * levels or an error code. */
int64_t mibwt_lyndon_array(void* storage, const uint8_t* T, int64_t n, uint32_t* lam);

/* Batched BWT of k independent texts (host buffers, libcubwt format each; the reference has
 * no batch mode, its benchmark loops over files, Benchmark_Skript_Linux.sh:7-138).
 * MIBWT_METHOD_PFP: consecutive texts are combined into builds of about 1 GiB
 * ($MIBWT_BATCH_CHUNK bytes), each uploaded as one array and built by the batched pipeline of
 * mibwt_batch_bwt_device (W = P = 0: the learned (W,P) per build); `workers` is not used.
 * MIBWT_METHOD_LYNDON: texts are handed out longest first to `workers` (0 = 8, at most 16)
 * worker storages, each with its own HIP stream and host thread.  index[i] receives the
 * primary index of text i or its error code; the return value is 0 or the first error. */
int64_t mibwt_batch_bwt(void* storage, const uint8_t* const* T, uint8_t* const* L, const int64_t* n,
                        int64_t k, int64_t* index, int32_t method, int32_t W, int32_t P, int32_t workers);

/* Batched PFP BWT of k texts resident in HBM (no counterpart in the reference, which runs one
 * process per file, Benchmark_Skript_Linux.sh:7-138): text i is dT[offsets[i] .. offsets[i+1])
 * (host offsets, nondecreasing, every text 1..max_length bytes), its BWT in REFERENCE format
 * goes to dO[offsets[i] .. offsets[i+1]) and its primary index to index[i].  Consecutive texts
 * are built together: one staged layout of up to about 1 GiB ($MIBWT_BATCH_CHUNK), every
 * kernel of the pipeline launched once per build over all of its texts (each text keeps its
 * own terminator, first-window rule and primary index; dictionary groups never mix texts).
 * W = P = 0: the learned selector chooses (W,P) per build.  On `stream` (NULL: the storage's
 * own), synchronous.  Returns LIBCUBWT_NO_ERROR or an error code. */
int64_t mibwt_batch_bwt_device(void* storage, const uint8_t* dT, const int64_t* offsets, int64_t k, uint8_t* dO,
                               int64_t* index, int32_t W, int32_t P, void* stream);

/* Per-phase statistics of the last call on this storage. */
#define MIBWT_MAX_PHASES 16
struct mibwt_phase_stats {
    int32_t num_phases;
    char name[MIBWT_MAX_PHASES][24];
    double ms[MIBWT_MAX_PHASES];         /* HIP-event time of the phase */
    double bytes[MIBWT_MAX_PHASES];      /* algorithmic HBM bytes of the phase */
    int64_t n, W, P;
    int64_t triggers, phrases, distinct_phrases, dict_length, alpha_groups;
    int64_t mixed_entries, doubling_rounds_dict, doubling_rounds_parse;
    double total_ms;                     /* HIP-event time from T resident to O resident */
    /* per-launch timing of the kernel family chosen with mibwt_set_kernel_timing */
    int64_t kt_launches;
    double kt_ms_total;                  /* sum of the launches' HIP-event durations */
    double kt_bytes_total;               /* sum of their algorithmic HBM bytes */
    int64_t retries;                     /* Lyndon: builds repeated after a fingerprint collision */
    int64_t device_bytes;                /* device memory the storage holds after the call (its peak:
                                            buffers only grow) */
    int64_t induction;                   /* Lyndon: 1 grammar induction over distinct labels,
                                            2 position-level induction (fallback) */
};
int64_t mibwt_last_phase_stats(void* storage, struct mibwt_phase_stats* out);

/* Bracket every launch of one kernel family with HIP events on the storage stream during
 * the following calls; the totals appear in mibwt_phase_stats (kt_*).  kernel:
 * "radix_scatter" (the LSD radix sort's ranking/scatter pass), "trig_mask" (the trigger
 * scan) or NULL to switch timing off. */
int64_t mibwt_set_kernel_timing(void* storage, const char* kernel);

/* Diagnostics: re-run a kernel `reps` times on the storage stream and return its mean
 * duration (ms) and algorithmic bytes per launch.  "scan": the trigger-scan kernel on the
 * text of the last PFP call; "copy": a plain 16-byte-per-thread device copy of
 * bytes_per_launch bytes (input), used to calibrate hardware byte counters; "radix": the LSD
 * radix sort of bytes_per_launch (input: an item count) uniformly random 64-bit keys with
 * 32-bit values, with the launch shape of the storage's last build (on a fresh storage: that of
 * a build of that many bytes), its mean duration per digit pass (8 per sort) and 24 bytes per
 * item (returns LIBCUBWT_GPU_ERROR if a sorted result is wrong). */
int64_t mibwt_time_kernel(void* storage, const char* name, int32_t reps, double* ms,
                          double* bytes_per_launch);

/* ---------------------------------------------------------------- learned (W,P) selector
 * The reference chooses (W,P) per input by hand (Benchmark_Skript_Linux.sh:14-73) and compiles
 * one binary per pair (Prefix_Free_Parsing_W=5_P=16_kommentiert.cu:19-20).  An MLP
 * 64->128->128->3 (bf16 weights, fp32 accumulation, hand-written HIP kernels) predicts
 * log(dictionary length/n), log(parse length/n) and log(build ms per byte) for each of 41
 * candidates: the grid W in {5,10,20,35} x P in {11,13,17,23,31,41,53,61,71,81} and the
 * reference default W=5,P=16 (last); the fastest predicted candidate is chosen (ensemble mean,
 * smoothed over each grid candidate's P-neighbours).  The model file is `selector_model.bin` next to lib/ (or $MIBWT_SELECTOR). */

/* The 41 candidates (W[i], P[i]) (arrays of at least 64); returns 41. */
int64_t mibwt_wp_grid(int32_t* W, int32_t* P);

/* Features of the 41 candidates for the device text dT[0..n): X[41][64]. */
int64_t mibwt_wp_features(void* storage, const uint8_t* dT, int64_t n, float* X);

/* (W,P) for the host text T[0..n) (a temporary storage of n bytes is allocated and freed;
 * callers that build afterwards use mibwt_select_wp_device on their storage instead).
 * Returns 1 when a trained model chose, 0 for the fixed W=5,P=16 fallback, or an error code. */
int64_t mibwt_select_wp(const uint8_t* T, int64_t n, int32_t* W, int32_t* P);

/* (W,P) for the device text dT[0..n); pred (may be NULL) receives the 41 predicted
 * log(ms per byte).  Returns 1 when a trained model chose, 0 for the fixed W=5,P=16 fallback. */
int64_t mibwt_select_wp_device(void* storage, const uint8_t* dT, int64_t n, int32_t* W, int32_t* P,
                               float* pred);

/* Load a model file (replaces the default one). */
int64_t mibwt_selector_load(const char* path);

/* MLP on raw (un-normalised) parameters in the fp32 master layout W1[128][64] b1[128]
 * W2[128][128] b2[128] W3[3][128] b3[3] (weights are used as bf16): mean-squared-error loss
 * and its gradient for a batch, and plain inference. */
int64_t mibwt_mlp_grad(void* storage, const float* params, const float* X, const float* Y, int64_t B,
                       float* grads, float* loss);
int64_t mibwt_mlp_predict(void* storage, const float* params, const float* X, int64_t B, float* Y);

/* Train on B samples (X[B][64] raw features, Y[B][3] targets) with Adam (minibatch `batch`,
 * 0 = full batch); normalisation is learned from the data; writes the model to out_path (may
 * be NULL) and installs it. */
int64_t mibwt_mlp_train(void* storage, const float* X, const float* Y, int64_t B, int32_t epochs,
                        int32_t batch, float lr, uint32_t seed, const char* out_path, float* final_loss);

/* The same for an ensemble of `members` (1..16) networks trained from seeds seed, seed+1, ...
 * with one shared normalisation; the selector averages their predicted times and smooths them
 * over each grid candidate's P-neighbours of the same W before choosing.  final_loss = the
 * members' mean. */
int64_t mibwt_mlp_train_ensemble(void* storage, const float* X, const float* Y, int64_t B, int32_t epochs,
                                 int32_t batch, float lr, uint32_t seed, int32_t members, const char* out_path,
                                 float* final_loss);

/* ---------------------------------------------------------------- launch-shape autotuner
 * The reference fixes its tiles at compile time for every input size
 * (Prefix_Free_Parsing_W=5_P=16_kommentiert.cu:16-23).  Here every build takes its launch shapes
 * (trigger-scan sub-tiles per tile, radix items per thread, chained-scan items per thread,
 * width of the packed initial sort keys) from a table with one row per length bucket
 * 2^16 .. 2^31 bytes, read at the first build from tuning_table.txt next to lib/ (or
 * $MIBWT_TUNING; "off": the defaults; $MIBWT_TUNE_FORCE="sub,items,chain,keybits" forces one
 * shape).  The BWT does not depend on the shapes. */

/* Measure the table on the device for every bucket up to max_bytes (<= the storage's
 * max_length): synthetic repetitive DNA of 3/4 of each bucket built with W=5,P=16 under each
 * candidate shape (coordinate descent, best of `reps` HIP-event timings); installs the table and
 * writes it to out_path (may be NULL).  Returns the number of buckets tuned. */
int64_t mibwt_autotune(void* storage, int64_t max_bytes, int32_t reps, const char* out_path);
/* Load a table file; use the table (on != 0) or the default shapes. */
int64_t mibwt_tuning_load(const char* path);
int64_t mibwt_tuning_use(int32_t on);
/* The shape a build of `size` bytes uses: shape[0..3] = scan sub-tiles, radix items, chained-scan
 * items, initial key bits. */
int64_t mibwt_tuning_get(int64_t size, int32_t* shape);

/* Device-memory plumbing for callers that keep texts resident in HBM without another HIP
 * runtime in the process (no counterpart in the reference, whose programs allocate inside
 * main).  Current device; synchronous copies; kind: 0 host->device, 1 device->host,
 * 2 device->device. */
int64_t mibwt_device_alloc(void** ptr, int64_t bytes);
int64_t mibwt_device_free(void* ptr);
int64_t mibwt_copy(void* dst, const void* src, int64_t bytes, int32_t kind);
int64_t mibwt_synchronize(void);
/* A non-blocking HIP stream of the library's runtime (for the `stream` arguments above), its
 * destruction (after it drained), and an asynchronous copy on it. */
int64_t mibwt_stream_create(void** stream);
int64_t mibwt_stream_destroy(void* stream);
int64_t mibwt_copy_async(void* dst, const void* src, int64_t bytes, int32_t kind, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MIBWT_H */
