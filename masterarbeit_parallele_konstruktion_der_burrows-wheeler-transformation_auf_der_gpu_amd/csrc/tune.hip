// Autotuner of the launch shapes (Tuning, common.hpp) per length bucket.
//
// The reference fixes its tiles at compile time (Prefix_Free_Parsing_W=5_P=16_kommentiert.cu:16-23:
// 512 threads, 4 elements per thread, 8192-element LDS sorts) for every input size.  Here the
// tile shapes that matter for the size of a build are chosen at run time from a table with one
// row per power-of-two length bucket (2^16 .. 2^31 bytes): the trigger scan's sub-tiles per
// chained tile, the radix items per thread, the chained-scan items per thread and the width of
// the packed initial keys of the suffix sorts.  Short builds want small tiles (a 64 KiB text
// has a single 65536-window scan tile with the default shape, so one CU would do all the
// scanning); long builds want large ones (every chained tile draws a ticket from one
// device-wide counter).
//
// mibwt_autotune measures the table on the device: for every bucket up to the requested size a
// synthetic repetitive DNA text of 3/4 of the bucket is generated on the GPU and built with
// W=5,P=16 under each candidate shape, one parameter at a time (coordinate descent from the
// defaults, every trial timed with HIP events, best of `reps`), and the table is written to a
// small text file that the library reads at the first build (tuning_table.txt next to lib/, or
// $MIBWT_TUNING; "off" keeps the defaults).  The BWT does not depend on the shapes.
#include <dlfcn.h>

#include <cstdlib>
#include <mutex>
#include <string>

#include "text.hpp"

namespace mibwt {

int64_t pfp_bwt_run(Storage& S, const uint8_t* src, bool src_host, uint8_t* dO, int64_t n, int W, int P,
                    PfpParseOut* parse_out);

namespace {

constexpr int kMinBucket = 16, kMaxBucket = 31;

struct Table {
    bool have[kMaxBucket + 1] = {};
    Tuning t[kMaxBucket + 1];
    double default_ms[kMaxBucket + 1] = {}, tuned_ms[kMaxBucket + 1] = {};  // measured by the tuner
};

std::mutex g_mu;
Table g_table;
bool g_use_table = true;

int bucket_of(uint64_t size)
{
    int b = 0;
    while (b < 63 && (1ull << b) < size) b++;
    return b < kMinBucket ? kMinBucket : (b > kMaxBucket ? kMaxBucket : b);
}

bool valid(const Tuning& t)
{
    return (t.scan_sub == 1 || t.scan_sub == 2 || t.scan_sub == 4 || t.scan_sub == 8) &&
           (t.sort_items == 8 || t.sort_items == 16) && (t.chain_items == 16 || t.chain_items == 64) &&
           t.key_bits >= 32 && t.key_bits <= 64;
}

bool table_load(const char* path, Table& out)
{
    FILE* f = fopen(path, "r");
    if (!f) return false;
    Table t;
    char line[256];
    int rows = 0;
    while (fgets(line, sizeof(line), f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        int b;
        Tuning x;
        if (sscanf(line, "%d %d %d %d %d", &b, &x.scan_sub, &x.sort_items, &x.chain_items, &x.key_bits) != 5) continue;
        if (b < kMinBucket || b > kMaxBucket || !valid(x)) continue;
        t.have[b] = true;
        t.t[b] = x;
        rows++;
    }
    fclose(f);
    if (rows == 0) return false;
    out = t;
    return true;
}

bool table_save(const char* path, const Table& t)
{
    FILE* f = fopen(path, "w");
    if (!f) return false;
    fprintf(f, "# mibwt launch-shape table: bucket (texts of 2^(b-1) < n <= 2^b bytes) scan_sub sort_items "
               "chain_items key_bits\n");
    for (int b = kMinBucket; b <= kMaxBucket; b++) {
        if (!t.have[b]) continue;
        fprintf(f, "%d %d %d %d %d", b, t.t[b].scan_sub, t.t[b].sort_items, t.t[b].chain_items, t.t[b].key_bits);
        if (t.tuned_ms[b] > 0)
            fprintf(f, "   # build of 3/4 x 2^%d bytes: default shapes %.3f ms, these %.3f ms", b, t.default_ms[b],
                    t.tuned_ms[b]);
        fprintf(f, "\n");
    }
    return fclose(f) == 0;
}

void ensure_default_table()
{
    static std::once_flag once;
    std::call_once(once, []() {
        const char* env = getenv("MIBWT_TUNING");
        std::string path;
        if (env && *env) {
            if (std::string(env) == "off") {
                g_use_table = false;
                return;
            }
            path = env;
        } else {
            Dl_info info;
            if (!dladdr((void*)&ensure_default_table, &info) || !info.dli_fname) return;
            path = info.dli_fname;
            const size_t k = path.rfind('/');
            path = (k == std::string::npos ? std::string(".") : path.substr(0, k)) + "/../tuning_table.txt";
        }
        Table t;
        if (table_load(path.c_str(), t)) {
            std::lock_guard<std::mutex> lk(g_mu);
            g_table = t;
        }
    });
}

// MIBWT_TUNE_FORCE="scan_sub,sort_items,chain_items,key_bits" forces one shape (tests)
bool forced(Tuning& t)
{
    const char* e = getenv("MIBWT_TUNE_FORCE");
    if (!e || !*e) return false;
    Tuning x;
    if (sscanf(e, "%d,%d,%d,%d", &x.scan_sub, &x.sort_items, &x.chain_items, &x.key_bits) != 4 || !valid(x))
        return false;
    t = x;
    return true;
}

// synthetic repetitive DNA: a random ACGT base of `base` bytes repeated, one substitution per
// ~1000 bytes (hash-based, so every bucket's text is reproducible)
__global__ void k_gen_dna(uint8_t* __restrict__ T, uint64_t n, uint64_t base, uint64_t seed)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t j = i % base;
    uint32_t c = (uint32_t)(mix64(seed ^ (j * 0x9E3779B97F4A7C15ull)) & 3u);
    const uint64_t h = mix64(seed + 0x632BE59BD9B4E019ull * (i + 1));
    if (h % 1000 == 0) c = (c + 1 + (uint32_t)((h >> 20) % 3)) & 3u;
    T[i] = (uint8_t)(0x54474341u >> (8 * c));  // A C G T
}

}  // namespace

void apply_tuning(Storage& S, uint64_t size)
{
    if (S.tune_pinned) return;
    Tuning t{};
    if (!forced(t)) {
        ensure_default_table();
        std::lock_guard<std::mutex> lk(g_mu);
        const int b = bucket_of(size);
        if (g_use_table && g_table.have[b]) t = g_table.t[b];
    }
    S.tune = t;
}

// the table entry a build of `size` bytes uses (after MIBWT_TUNE_FORCE and the table)
Tuning tuning_for(uint64_t size)
{
    Storage tmp;
    apply_tuning(tmp, size);
    return tmp.tune;
}

bool tuning_load(const char* path)
{
    ensure_default_table();  // the default file is read once; an explicit load then replaces it
    Table t;
    if (!table_load(path, t)) return false;
    std::lock_guard<std::mutex> lk(g_mu);
    g_table = t;
    return true;
}

void tuning_use_table(bool on)
{
    ensure_default_table();
    std::lock_guard<std::mutex> lk(g_mu);
    g_use_table = on;
}

// Measure the table for every bucket up to max_bytes (see the top of this file); returns the
// number of buckets tuned.
int autotune_run(Storage& S, uint64_t max_bytes, int reps, const char* path)
{
    MIBWT_REQUIRE(max_bytes >= (1u << 16) && (int64_t)max_bytes <= S.max_length, LIBCUBWT_BAD_PARAMETER,
                  "tuning size out of range");
    const int top = bucket_of(max_bytes);
    uint8_t* dT = S.buf<uint8_t>(B_BATCH_IN, (size_t)max_bytes);
    uint8_t* dO = S.buf<uint8_t>(B_BATCH_OUT, (size_t)max_bytes);
    Table t;
    const Tuning def{};
    const int subs[4] = {8, 4, 2, 1}, items[2] = {16, 8}, chains[2] = {64, 16}, keys[5] = {64, 56, 48, 40, 32};
    int tuned = 0;
    for (int b = kMinBucket; b <= top; b++) {
        const uint64_t n = std::min<uint64_t>(max_bytes, (3ull << b) / 4);
        k_gen_dna<<<grid_for(n), kBlock, 0, S.stream>>>(dT, n, std::max<uint64_t>(1024, n / 16),
                                                        0xB17C0DEull + (uint64_t)b);
        MIBWT_LAUNCH_CHECK();
        Tuning cur = def;
        auto trial = [&](const Tuning& x) {
            S.tune = x;
            double best = 1e30;
            for (int r = 0; r <= reps; r++) {  // the first build of a shape warms its buffers
                pfp_bwt_run(S, dT, false, dO, (int64_t)n, 5, 16, nullptr);
                if (r > 0) best = std::min(best, S.stats.total_ms);
            }
            return best;
        };
        S.tune_pinned = true;
        double base_ms = 0, best = 0;
        // a candidate replaces the current shape only when it is at least 2% faster (timing noise)
        auto better = [&](double ms) { return ms < best * 0.98; };
        try {
            best = base_ms = trial(cur);
            for (int v : subs) {
                Tuning x = cur;
                x.scan_sub = v;
                if (v == cur.scan_sub) continue;
                const double ms = trial(x);
                if (better(ms)) best = ms, cur = x;
            }
            for (int v : items) {
                Tuning x = cur;
                x.sort_items = v;
                if (v == cur.sort_items) continue;
                const double ms = trial(x);
                if (better(ms)) best = ms, cur = x;
            }
            for (int v : chains) {
                Tuning x = cur;
                x.chain_items = v;
                if (v == cur.chain_items) continue;
                const double ms = trial(x);
                if (better(ms)) best = ms, cur = x;
            }
            for (int v : keys) {
                Tuning x = cur;
                x.key_bits = v;
                if (v == cur.key_bits) continue;
                const double ms = trial(x);
                if (better(ms)) best = ms, cur = x;
            }
            if (getenv("MIBWT_VERBOSE"))
                fprintf(stderr, "mibwt: tuned bucket %d (n = %llu): %d %d %d %d, %.3f ms\n", b, (unsigned long long)n,
                        cur.scan_sub, cur.sort_items, cur.chain_items, cur.key_bits, best);
        } catch (...) {
            S.tune_pinned = false;
            throw;
        }
        S.tune_pinned = false;
        t.have[b] = true;
        t.t[b] = cur;
        t.default_ms[b] = base_ms;
        t.tuned_ms[b] = best;
        tuned++;
    }
    if (path && *path) MIBWT_REQUIRE(table_save(path, t), LIBCUBWT_BAD_PARAMETER, "cannot write the tuning table");
    ensure_default_table();
    std::lock_guard<std::mutex> lk(g_mu);
    g_table = t;
    g_use_table = true;
    return tuned;
}

}  // namespace mibwt
