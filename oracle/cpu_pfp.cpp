// Multi-threaded CPU restatement of the reference prefix-free-parsing BWT
// (Prefix_Free_Parsing_W=5_P=16_kommentiert.cu, "PFPk").  TEST INFRASTRUCTURE AND CPU BASELINE
// ONLY: the product library never links or calls it; bench.py times it beside the GPU build
// (its cpu_baseline leg) and the tests check it against the SA-IS oracle.
//
// The phases follow the reference's Compute_D (PFPk:2415) and Compute_S (PFPk:4330) with the
// threads of the host (std::thread):
//   1. triggers      windows T[e-W+1..e] with sum_k T[e-W+1+k] 2^(W-1-k) mod P == 0 or equal to
//                    T[0..W-1], W <= e < n (PFPk:443-571), rolled per window, text split in
//                    chunks; the last phrase ends at n+W-1 (PFPk:573-592);
//   2. dictionary    64-bit phrase fingerprints, parallel sort, byte-exact grouping, dense ids
//                    in lexicographic phrase order (PFPk:1990-2164, 616-691);
//   3. parse SA      prefix doubling over the parse of phrase ranks (PFPk:2342-2412);
//   4. S             all suffixes longer than W of the distinct phrases, sorted as strings;
//                    equal suffixes form one group (the reference's S set and intervals,
//                    PFPk:2509-3240);
//   5. assembly      a group whose preceding characters agree is filled directly
//                    (PFPk:3337-3350); otherwise its occurrences are merged in the order of the
//                    next parse suffix (PFPk:3394-4325).
// The terminator is virtual (smaller than every byte): all 256 byte values are accepted.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <functional>
#include <queue>
#include <thread>
#include <vector>

namespace {

int g_threads = 1;

template <typename F>
void parallel_for(int64_t n, F&& f)  // f(begin, end, thread)
{
    const int T = (int)std::max<int64_t>(1, std::min<int64_t>(g_threads, n));
    if (T == 1) {
        f(0, n, 0);
        return;
    }
    std::vector<std::thread> th;
    const int64_t chunk = (n + T - 1) / T;
    for (int t = 0; t < T; t++) {
        const int64_t a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([&, a, b, t] { f(a, b, t); });
    }
    for (auto& x : th) x.join();
}

// chunks sorted in parallel, then merged pairwise round by round (each round in parallel)
template <typename V, typename Cmp>
void parallel_sort(std::vector<V>& v, Cmp cmp)
{
    const int64_t n = (int64_t)v.size();
    int T = std::max(1, g_threads);
    while (T > 1 && n / T < 4096) T >>= 1;
    if (T <= 1) {
        std::sort(v.begin(), v.end(), cmp);
        return;
    }
    std::vector<int64_t> bound(T + 1);
    for (int t = 0; t <= T; t++) bound[t] = n * t / T;
    {
        std::vector<std::thread> th;
        for (int t = 0; t < T; t++)
            th.emplace_back([&, t] { std::sort(v.begin() + bound[t], v.begin() + bound[t + 1], cmp); });
        for (auto& x : th) x.join();
    }
    std::vector<V> tmp(v.size());
    std::vector<V>* src = &v;
    std::vector<V>* dst = &tmp;
    for (int width = 1; width < T; width *= 2) {
        std::vector<std::thread> th;
        for (int t = 0; t < T; t += 2 * width) {
            const int64_t a = bound[t], m = bound[std::min(T, t + width)], b = bound[std::min(T, t + 2 * width)];
            th.emplace_back([&, a, m, b] {
                std::merge(src->begin() + a, src->begin() + m, src->begin() + m, src->begin() + b, dst->begin() + a, cmp);
            });
        }
        for (auto& x : th) x.join();
        std::swap(src, dst);
    }
    if (src != &v) v.swap(tmp);
}

inline uint64_t mix64(uint64_t h)
{
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}

struct Text {
    const uint8_t* T;
    int64_t n;
    // symbol of the extended text T.$^W: -1 for the terminator
    int sym(int64_t i) const { return i < n ? (int)T[i] : -1; }
};

// compare the strings X[0..lx) and Y[0..ly) of the extended text starting at px, py
int cmp_str(const Text& t, int64_t px, int64_t lx, int64_t py, int64_t ly)
{
    const int64_t m = std::min(lx, ly);
    // bytes inside the text first, 8 at a time
    int64_t i = 0;
    const int64_t inside = std::max<int64_t>(0, std::min(m, std::min(t.n - px, t.n - py)));
    while (i + 8 <= inside) {
        uint64_t a, b;
        memcpy(&a, t.T + px + i, 8);
        memcpy(&b, t.T + py + i, 8);
        if (a != b) break;
        i += 8;
    }
    for (; i < m; i++) {
        const int u = t.sym(px + i), v = t.sym(py + i);
        if (u != v) return u < v ? -1 : 1;
    }
    return lx < ly ? -1 : (lx > ly ? 1 : 0);
}

}  // namespace

extern "C" {

// Reference-format BWT O[0..n) of T by the PFP pipeline above with `threads` threads (0: all
// hardware threads); returns the libcubwt primary index, or -1 on bad parameters.
int64_t cpu_pfp_bwt(const uint8_t* T, int64_t n, int W, int P, uint8_t* O, int threads, int64_t* phrases_out,
                    int64_t* distinct_out)
{
    if (!T || !O || n < 1 || W < 1 || W > 64 || P < 1 || n >= (1ll << 31) - (1ll << 20)) return -1;
    g_threads = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    const Text tx{T, n};
    // CPU_PFP_VERBOSE=1 prints the time of every phase to stderr
    const bool verbose = getenv("CPU_PFP_VERBOSE") != nullptr;
    auto t_last = std::chrono::steady_clock::now();
    auto lap = [&](const char* what) {
        if (!verbose) return;
        const auto now = std::chrono::steady_clock::now();
        fprintf(stderr, "cpu_pfp %-10s %8.3f s\n", what, std::chrono::duration<double>(now - t_last).count());
        t_last = now;
    };

    // ---- 1. triggers (PFPk:443-571): exact value mod P rolled per window end
    std::vector<std::vector<uint32_t>> part(g_threads);
    const int64_t first_e = W, last_e = n - 1;
    const int64_t ne = last_e >= first_e ? last_e - first_e + 1 : 0;
    uint64_t p2w = 1 % (uint64_t)P;
    for (int k = 0; k < W; k++) p2w = (p2w * 2) % (uint64_t)P;
    parallel_for(ne, [&](int64_t a, int64_t b, int t) {
        std::vector<uint32_t>& out = part[t];
        int64_t e = first_e + a;
        uint64_t h = 0;
        for (int64_t k = e - W + 1; k <= e; k++) h = (2 * h + T[k]) % (uint64_t)P;
        for (; e < first_e + b; e++) {
            if (e > first_e + a) h = (2 * h + T[e] + (uint64_t)P * 256 - (T[e - W] * p2w) % (uint64_t)P) % (uint64_t)P;
            bool trig = (h == 0);
            if (!trig) trig = memcmp(T + e - W + 1, T, (size_t)W) == 0;
            if (trig) out.push_back((uint32_t)e);
        }
    });
    std::vector<uint32_t> E;
    for (auto& v : part) E.insert(E.end(), v.begin(), v.end());
    part.clear();
    E.push_back((uint32_t)(n + W - 1));  // the last phrase ends in the terminator
    const int64_t d = (int64_t)E.size();
    auto pstart = [&](int64_t k) -> int64_t { return k == 0 ? 0 : (int64_t)E[k - 1] - W + 1; };
    auto plen = [&](int64_t k) -> int64_t { return (int64_t)E[k] - pstart(k) + 1; };

    lap("triggers");
    // ---- 2. dictionary: fingerprints, exact groups, lexicographic ids
    std::vector<std::pair<uint64_t, uint32_t>> hk(d);
    parallel_for(d, [&](int64_t a, int64_t b, int) {
        for (int64_t k = a; k < b; k++) {
            const int64_t s = pstart(k), l = plen(k);
            uint64_t h = mix64((uint64_t)l * 0x9E3779B97F4A7C15ull);
            for (int64_t i = 0; i < l; i++) h = mix64(h ^ (uint64_t)(tx.sym(s + i) + 1) ^ ((uint64_t)i << 40));
            hk[k] = {h, (uint32_t)k};
        }
    });
    parallel_sort(hk, [](const std::pair<uint64_t, uint32_t>& x, const std::pair<uint64_t, uint32_t>& y) {
        return x.first != y.first ? x.first < y.first : x.second < y.second;
    });
    // representatives: runs of equal fingerprints, split exactly by byte comparison
    std::vector<uint32_t> rep;        // one occurrence per distinct phrase
    std::vector<uint32_t> pid(d);     // occurrence -> distinct phrase (unordered id first)
    {
        int64_t i = 0;
        while (i < d) {
            int64_t j = i;
            while (j < d && hk[j].first == hk[i].first) j++;
            std::vector<uint32_t> reps_here;
            for (int64_t q = i; q < j; q++) {
                const uint32_t k = hk[q].second;
                uint32_t id = UINT32_MAX;
                for (uint32_t r : reps_here)
                    if (cmp_str(tx, pstart(k), plen(k), pstart(rep[r]), plen(rep[r])) == 0) id = r;
                if (id == UINT32_MAX) {
                    id = (uint32_t)rep.size();
                    rep.push_back(k);
                    reps_here.push_back(id);
                }
                pid[k] = id;
            }
            i = j;
        }
    }
    hk.clear();
    hk.shrink_to_fit();
    const int64_t du = (int64_t)rep.size();
    std::vector<uint32_t> lex(du);  // distinct ids in lexicographic order
    for (int64_t u = 0; u < du; u++) lex[u] = (uint32_t)u;
    parallel_sort(lex, [&](uint32_t a, uint32_t b) {
        return cmp_str(tx, pstart(rep[a]), plen(rep[a]), pstart(rep[b]), plen(rep[b])) < 0;
    });
    std::vector<uint32_t> lexrank(du);
    for (int64_t r = 0; r < du; r++) lexrank[lex[r]] = (uint32_t)r;
    std::vector<uint32_t> freq(du, 0);
    for (int64_t k = 0; k < d; k++) freq[pid[k]]++;

    lap("dictionary");
    // ---- 3. parse suffix array by prefix doubling (PFPk:2342-2412)
    struct Dbl {
        uint32_t r1, r2, k;
    };
    std::vector<uint32_t> rank(d), psa(d);
    parallel_for(d, [&](int64_t a, int64_t b, int) {
        for (int64_t k = a; k < b; k++) rank[k] = lexrank[pid[k]] + 1;  // 0 = past the end
    });
    {
        std::vector<Dbl> it(d);
        for (int64_t h = 1;; h *= 2) {
            parallel_for(d, [&](int64_t a, int64_t b, int) {
                for (int64_t k = a; k < b; k++) it[k] = {rank[k], (k + h < d) ? rank[k + h] : 0u, (uint32_t)k};
            });
            parallel_sort(it, [](const Dbl& x, const Dbl& y) { return x.r1 != y.r1 ? x.r1 < y.r1 : x.r2 < y.r2; });
            bool distinct = true;
            uint32_t r = 1;
            for (int64_t i = 0; i < d; i++) {
                if (i > 0) {
                    if (it[i].r1 != it[i - 1].r1 || it[i].r2 != it[i - 1].r2) r = (uint32_t)i + 1;
                    else distinct = false;
                }
                psa[i] = r;  // new rank of the item at sorted position i (scattered below)
            }
            parallel_for(d, [&](int64_t a, int64_t b, int) {
                for (int64_t i = a; i < b; i++) rank[it[i].k] = psa[i];
            });
            for (int64_t i = 0; i < d; i++) psa[i] = it[i].k;
            if (distinct || h >= d) break;
        }
    }
    lap("parse_sa");
    // occurrences of every distinct phrase in the order of their next parse suffix: walking
    // the parse suffix array, suffix k+1 at rank r belongs to occurrence k
    std::vector<int64_t> occ_off(du + 1, 0);
    for (int64_t u = 0; u < du; u++) occ_off[u + 1] = occ_off[u] + freq[u];
    std::vector<uint32_t> occ(d);
    {
        std::vector<int64_t> fill(occ_off.begin(), occ_off.end() - 1);
        for (int64_t r = 0; r < d; r++) {
            const int64_t k1 = psa[r];
            if (k1 == 0) continue;
            occ[fill[pid[k1 - 1]]++] = (uint32_t)(k1 - 1);
        }
        occ[fill[pid[d - 1]]++] = (uint32_t)(d - 1);  // the last occurrence has no next suffix
    }
    psa.clear();

    lap("occ_lists");
    // ---- 4. dictionary suffixes longer than W, sorted as strings
    // key: the first `per` symbols as dense codes of `abits` bits (0 past the suffix end,
    // 1 terminator, 2.. the bytes that occur, in byte order)
    std::vector<uint32_t> code(256, 0);
    {
        std::vector<std::vector<uint8_t>> seen(g_threads, std::vector<uint8_t>(256, 0));
        parallel_for(n, [&](int64_t a, int64_t b, int t) {
            for (int64_t i = a; i < b; i++) seen[t][T[i]] = 1;
        });
        uint32_t c = 2;
        for (int v = 0; v < 256; v++) {
            bool any = false;
            for (auto& x : seen) any = any || x[v];
            if (any) code[v] = c++;
        }
    }
    int abits = 1;
    while ((1u << abits) <= (uint32_t)(*std::max_element(code.begin(), code.end()))) abits++;
    const int per = 64 / abits;
    struct Suf {
        uint64_t key;
        uint32_t u, j;  // distinct phrase, offset
    };
    std::vector<int64_t> soff(du + 1, 0);
    for (int64_t u = 0; u < du; u++) soff[u + 1] = soff[u] + std::max<int64_t>(0, plen(rep[u]) - W);
    std::vector<Suf> S(soff[du]);
    parallel_for(du, [&](int64_t a, int64_t b, int) {
        for (int64_t u = a; u < b; u++) {
            const int64_t l = plen(rep[u]), s0 = pstart(rep[u]);
            for (int64_t j = 0; l - j > W; j++) {
                uint64_t key = 0;
                for (int64_t i = 0; i < per; i++) {
                    const int64_t q = s0 + j + i;
                    key = (key << abits) | (uint64_t)(j + i >= l ? 0u : (q >= n ? 1u : code[T[q]]));
                }
                S[soff[u] + j] = {key, (uint32_t)u, (uint32_t)j};
            }
        }
    });
    auto suf_eq_cmp = [&](const Suf& a, const Suf& b) -> int {
        if (a.key != b.key) return a.key < b.key ? -1 : 1;
        const int64_t la = plen(rep[a.u]) - a.j, lb = plen(rep[b.u]) - b.j;
        if (la <= per || lb <= per) return la < lb ? -1 : (la > lb ? 1 : 0);  // the key holds them whole
        return cmp_str(tx, pstart(rep[a.u]) + a.j + per, la - per, pstart(rep[b.u]) + b.j + per, lb - per);
    };
    parallel_sort(S, [&](const Suf& a, const Suf& b) {
        const int c = suf_eq_cmp(a, b);
        return c != 0 ? c < 0 : a.u < b.u;
    });
    const int64_t ns = (int64_t)S.size();
    lap("s_sort");
    // groups of equal suffixes and their output offsets
    std::vector<uint8_t> head(ns);
    parallel_for(ns, [&](int64_t a, int64_t b, int) {
        for (int64_t i = a; i < b; i++) head[i] = (i == 0 || suf_eq_cmp(S[i - 1], S[i]) != 0) ? 1 : 0;
    });
    std::vector<int64_t> gstart;
    for (int64_t i = 0; i < ns; i++)
        if (head[i]) gstart.push_back(i);
    const int64_t ng = (int64_t)gstart.size();
    gstart.push_back(ns);
    std::vector<int64_t> goff(ng + 1, 0);
    parallel_for(ng, [&](int64_t a, int64_t b, int) {
        for (int64_t g = a; g < b; g++) {
            int64_t c = 0;
            for (int64_t i = gstart[g]; i < gstart[g + 1]; i++) c += freq[S[i].u];
            goff[g + 1] = c;
        }
    });
    for (int64_t g = 0; g < ng; g++) goff[g + 1] += goff[g];
    if (goff[ng] != n) return -2;  // every text suffix must fall into exactly one group

    lap("s_groups");
    // ---- 5. assembly
    auto prev_char = [&](uint32_t k) -> uint8_t {  // character before occurrence k (wraps)
        const int64_t s = pstart(k);
        return s > 0 ? T[s - 1] : T[n - 1];
    };
    std::atomic<int64_t> pidx{0};
    parallel_for(ng, [&](int64_t a, int64_t b, int) {
        for (int64_t g = a; g < b; g++) {
            const int64_t s0 = gstart[g], s1 = gstart[g + 1];
            int64_t out = goff[g];
            // direct fill when no whole phrase is in the group and the preceding characters agree
            bool direct = true;
            for (int64_t i = s0; i < s1 && direct; i++) {
                if (S[i].j == 0) direct = false;
                else if (T[pstart(rep[S[i].u]) + S[i].j - 1] != T[pstart(rep[S[s0].u]) + S[s0].j - 1]) direct = false;
            }
            if (direct) {
                const uint8_t c = T[pstart(rep[S[s0].u]) + S[s0].j - 1];
                for (int64_t i = s0; i < s1; i++)
                    for (uint32_t q = 0; q < freq[S[i].u]; q++) O[out++] = c;
                continue;
            }
            // merge the members' occurrence lists by the rank of the next parse suffix
            using Item = std::pair<uint32_t, int64_t>;  // (rank of suffix k+1, member list cursor)
            struct Cur {
                int64_t pos, end;
                uint32_t u, j;
            };
            std::vector<Cur> cur;
            for (int64_t i = s0; i < s1; i++)
                cur.push_back({occ_off[S[i].u], occ_off[S[i].u + 1], S[i].u, S[i].j});
            auto key = [&](const Cur& c) -> uint32_t {
                const uint32_t k = occ[c.pos];
                return (k + 1 < (uint64_t)d) ? rank[k + 1] : 0u;  // the last occurrence: alone
            };
            std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
            for (int64_t c = 0; c < (int64_t)cur.size(); c++)
                if (cur[c].pos < cur[c].end) pq.push({key(cur[c]), c});
            while (!pq.empty()) {
                const int64_t c = pq.top().second;
                pq.pop();
                const uint32_t k = occ[cur[c].pos];
                if (cur[c].j == 0) {
                    if (k == 0) pidx = out + 1;
                    O[out++] = prev_char(k);
                } else {
                    O[out++] = T[pstart(k) + cur[c].j - 1];
                }
                if (++cur[c].pos < cur[c].end) pq.push({key(cur[c]), c});
            }
        }
    });
    lap("assembly");
    if (phrases_out) *phrases_out = d;
    if (distinct_out) *distinct_out = du;
    return pidx.load();
}

}  // extern "C"
