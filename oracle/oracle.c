/*
 * CPU oracle for the BWT-build path.  TEST INFRASTRUCTURE ONLY (see oracle.h):
 * never linked into the product library.
 *
 * Contents
 *  1. SA-IS suffix sorting (Nong, Zhang, Chan 2009) -> reference-format BWT and the
 *     libcubwt primary index; inverse BWT by LF-mapping.
 *  2. A restatement of the reference prefix-free-parsing BWT
 *     (Prefix_Free_Parsing_W=5_P=16_kommentiert.cu, "PFPk" below) with runtime W, P:
 *     trigger set, phrases, lexicographic phrase order, parse suffix array, colex order,
 *     dictionary-suffix set S with its intervals, S sort and BWT assembly.  Each step cites
 *     the reference lines it follows.  The terminator is virtual (value -1) instead of the
 *     reference's physical byte 1, so the restatement also covers bytes 0, 1 and 255.
 *  3. Duval's Lyndon factorisation and a restatement of the Lyndon-grammar BWT
 *     (Lyndon_Grammar_BWT.cu, "LYN" below): longest Lyndon prefix of every position (the
 *     maximal nodes of the Lyndon forest: stack words and right children, LYN:647-1868,
 *     1910-2062), their lexicographic ranks (the sorted grammar rules, LYN:3621-4083), and the
 *     BWT by induction over rank groups (LYN:4628-4896), here as one left-to-right induced
 *     pass: suffixes are grouped by their Lyndon prefix and, inside a group, ordered by the
 *     suffix that follows the prefix, which is smaller and therefore already placed.
 *  4. The same BWT by induction over the grammar itself (LYN:4114-4896 with runs of equal
 *     children): distinct labels with their child lists, items standing for a child slot in
 *     every occurrence of a label, groups refined by the next run of labels and filled at once
 *     when their items share the preceding character.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ------------------------------------------------------------------ 1. SA-IS */

static void sais_buckets(const int32_t* s, int32_t n, int32_t K, int32_t* bkt, int end)
{
    memset(bkt, 0, sizeof(int32_t) * (size_t)K);
    for (int32_t i = 0; i < n; i++) bkt[s[i]]++;
    int32_t sum = 0;
    for (int32_t c = 0; c < K; c++) {
        sum += bkt[c];
        bkt[c] = end ? sum : sum - bkt[c];
    }
}

#define SAIS_LMS(t, i) ((i) > 0 && (t)[i] && !(t)[(i)-1])

static void sais_induce(const int32_t* s, int32_t* SA, const uint8_t* t, int32_t n, int32_t K,
                        int32_t* bkt)
{
    sais_buckets(s, n, K, bkt, 0);
    for (int32_t i = 0; i < n; i++) {
        int32_t j = SA[i] - 1;
        if (SA[i] > 0 && !t[j]) SA[bkt[s[j]]++] = j;
    }
    sais_buckets(s, n, K, bkt, 1);
    for (int32_t i = n - 1; i >= 0; i--) {
        int32_t j = SA[i] - 1;
        if (SA[i] > 0 && t[j]) SA[--bkt[s[j]]] = j;
    }
}

/* s[0..n) over [0,K) with s[n-1] == 0 the unique smallest symbol. */
static int sais_rec(const int32_t* s, int32_t* SA, int32_t n, int32_t K)
{
    uint8_t* t = (uint8_t*)malloc((size_t)n);
    int32_t* bkt = (int32_t*)malloc(sizeof(int32_t) * (size_t)K);
    if (!t || !bkt) { free(t); free(bkt); return -1; }
    t[n - 1] = 1;
    for (int32_t i = n - 2; i >= 0; i--)
        t[i] = (s[i] < s[i + 1] || (s[i] == s[i + 1] && t[i + 1])) ? 1 : 0;

    sais_buckets(s, n, K, bkt, 1);
    for (int32_t i = 0; i < n; i++) SA[i] = -1;
    for (int32_t i = 1; i < n; i++)
        if (SAIS_LMS(t, i)) SA[--bkt[s[i]]] = i;
    sais_induce(s, SA, t, n, K, bkt);

    int32_t n1 = 0;
    for (int32_t i = 0; i < n; i++)
        if (SAIS_LMS(t, SA[i])) SA[n1++] = SA[i];
    for (int32_t i = n1; i < n; i++) SA[i] = -1;
    int32_t name = 0, prev = -1;
    for (int32_t i = 0; i < n1; i++) {
        int32_t pos = SA[i];
        int diff = 0;
        for (int32_t d = 0; d < n; d++) {
            if (prev == -1 || s[pos + d] != s[prev + d] || t[pos + d] != t[prev + d]) {
                diff = 1;
                break;
            } else if (d > 0 && (SAIS_LMS(t, pos + d) || SAIS_LMS(t, prev + d))) {
                break;
            }
        }
        if (diff) { name++; prev = pos; }
        SA[n1 + pos / 2] = name - 1;
    }
    for (int32_t i = n - 1, j = n - 1; i >= n1; i--)
        if (SA[i] >= 0) SA[j--] = SA[i];

    int32_t* s1 = SA + n - n1;
    if (name < n1) {
        if (sais_rec(s1, SA, n1, name) != 0) { free(t); free(bkt); return -1; }
    } else {
        for (int32_t i = 0; i < n1; i++) SA[s1[i]] = i;
    }
    for (int32_t i = 1, j = 0; i < n; i++)
        if (SAIS_LMS(t, i)) s1[j++] = i;
    for (int32_t i = 0; i < n1; i++) SA[i] = s1[SA[i]];
    for (int32_t i = n1; i < n; i++) SA[i] = -1;
    sais_buckets(s, n, K, bkt, 1);
    for (int32_t i = n1 - 1; i >= 0; i--) {
        int32_t j = SA[i];
        SA[i] = -1;
        SA[--bkt[s[j]]] = j;
    }
    sais_induce(s, SA, t, n, K, bkt);
    free(t);
    free(bkt);
    return 0;
}

int oracle_sa(const uint8_t* T, int64_t n, int32_t* SA)
{
    if (n <= 0) return 0;
    if (n >= INT32_MAX - 1) return -1;
    int32_t m = (int32_t)n + 1;
    int32_t* s = (int32_t*)malloc(sizeof(int32_t) * (size_t)m);
    int32_t* sa = (int32_t*)malloc(sizeof(int32_t) * (size_t)m);
    if (!s || !sa) { free(s); free(sa); return -1; }
    for (int32_t i = 0; i < m - 1; i++) s[i] = (int32_t)T[i] + 1;
    s[m - 1] = 0;
    int rc = sais_rec(s, sa, m, 257);
    if (rc == 0) memcpy(SA, sa + 1, sizeof(int32_t) * (size_t)n); /* sa[0] is the "$" row */
    free(s);
    free(sa);
    return rc;
}

int64_t oracle_bwt(const uint8_t* T, int64_t n, uint8_t* O)
{
    if (n <= 0) return 0;
    int32_t* SA = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    if (!SA) return -2;
    if (oracle_sa(T, n, SA) != 0) { free(SA); return -2; }
    int64_t index = -1;
    for (int64_t j = 0; j < n; j++) {
        int64_t p = SA[j];
        if (p == 0) index = j + 1;
        O[j] = T[p == 0 ? n - 1 : p - 1];
    }
    free(SA);
    return index;
}

void oracle_ref_to_cubwt(const uint8_t* O, int64_t n, int64_t index, uint8_t* L)
{
    /* O[index-1] holds T[n-1] (the wrap character of suffix 0). */
    uint8_t wrap = O[index - 1];
    memmove(L + 1, O, (size_t)(index - 1));
    if (L != O) memcpy(L + index, O + index, (size_t)(n - index));
    L[0] = wrap;
}

void oracle_cubwt_to_ref(const uint8_t* L, int64_t n, int64_t index, uint8_t* O)
{
    uint8_t wrap = L[0];
    memmove(O, L + 1, (size_t)(index - 1));
    if (L != O) memcpy(O + index, L + index, (size_t)(n - index));
    O[index - 1] = wrap;
}

int oracle_unbwt(const uint8_t* L, int64_t n, int64_t index, uint8_t* T)
{
    if (n <= 0) return 0;
    if (index < 1 || index > n) return -1;
    /* Rows of the (n+1)-row BWT matrix: B[0] = L[0], B[index] = $, otherwise L shifted. */
    int64_t m = n + 1;
    uint32_t* lf = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)m);
    if (!lf) return -2;
    int64_t C[257];
    memset(C, 0, sizeof(C));
    for (int64_t i = 0; i < n; i++) C[L[i] + 1]++;
    C[0] = 1; /* the "$" row */
    for (int c = 1; c < 257; c++) C[c] += C[c - 1];
    /* C[c] now = 1 + #symbols < c */
    int64_t occ[256];
    for (int c = 0; c < 256; c++) occ[c] = (c == 0) ? 1 : C[c];
    for (int64_t r = 0; r < m; r++) {
        if (r == index) { lf[r] = 0; continue; }
        uint8_t c = (r < index) ? L[r] : L[r - 1];
        lf[r] = (uint32_t)(occ[c]++);
    }
    int64_t r = 0;
    for (int64_t k = n - 1; k >= 0; k--) {
        if (r == index) { free(lf); return -1; }
        T[k] = (r < index) ? L[r] : L[r - 1];
        r = lf[r];
    }
    free(lf);
    return (r == index) ? 0 : -1;
}

/* The same inverse with `threads` POSIX threads (0: one per online CPU): LF by per-chunk symbol
 * counts, then the LF cycle is cut at every row that is a multiple of a power-of-two stride;
 * the segments between cut rows are walked in parallel twice (lengths, then the text), with
 * their order taken from the cycle starting at row 0.  Same result and error convention as
 * oracle_unbwt (the walk visits each row once; a cycle that misses the '$' row after n
 * symbols is an error). */
typedef struct {
    const uint8_t* L;
    int64_t n, index, m;
    uint32_t* lf;
    int64_t* occ;      /* per chunk: first LF slot of each symbol */
    int64_t chunk;     /* rows per chunk */
    int64_t stride;    /* power of two */
    int64_t nseg;
    int64_t* seg_len;  /* rows walked from each cut row */
    int64_t* seg_next; /* next cut row index, -1 = the '$' row */
    int64_t* seg_off;  /* output offset (from the end of the text) of each segment */
    uint8_t* T;
    int tid, threads, pass;
} unbwt_job;

static inline uint8_t un_row(const unbwt_job* j, int64_t r) { return (r < j->index) ? j->L[r] : j->L[r - 1]; }

static void* unbwt_worker(void* arg)
{
    unbwt_job* j = (unbwt_job*)arg;
    if (j->pass == 0) {  /* symbol counts of this thread's chunk */
        int64_t* cnt = j->occ + (int64_t)j->tid * 256;
        memset(cnt, 0, sizeof(int64_t) * 256);
        int64_t a = j->tid * j->chunk, b = a + j->chunk < j->m ? a + j->chunk : j->m;
        for (int64_t r = a; r < b; r++)
            if (r != j->index) cnt[un_row(j, r)]++;
    } else if (j->pass == 1) {  /* LF of this chunk's rows */
        int64_t* nx = j->occ + (int64_t)j->tid * 256;
        int64_t a = j->tid * j->chunk, b = a + j->chunk < j->m ? a + j->chunk : j->m;
        for (int64_t r = a; r < b; r++) {
            if (r == j->index) { j->lf[r] = 0; continue; }
            j->lf[r] = (uint32_t)(nx[un_row(j, r)]++);
        }
    } else {  /* pass 2: segment lengths; pass 3: the text */
        for (int64_t s = j->tid; s < j->nseg; s += j->threads) {
            int64_t r = s * j->stride;
            if (r == j->index) { j->seg_len[s] = 0; j->seg_next[s] = -1; continue; }
            int64_t k = (j->pass == 3) ? j->seg_off[s] : 0, steps = 0;
            if (j->pass == 3 && k < 0) continue;  /* not on the cycle from row 0 */
            for (;;) {
                if (j->pass == 3) {
                    if (k >= j->n) break;
                    j->T[j->n - 1 - k] = un_row(j, r);
                    k++;
                }
                r = j->lf[r];
                steps++;
                if (r == j->index || (r & (j->stride - 1)) == 0 || steps > j->m) break;
            }
            if (j->pass == 2) {
                j->seg_len[s] = steps;
                j->seg_next[s] = (r == j->index || steps > j->m) ? -1 : r / j->stride;
            }
        }
    }
    return NULL;
}

int oracle_unbwt_mt(const uint8_t* L, int64_t n, int64_t index, uint8_t* T, int threads)
{
    if (n <= 0) return 0;
    if (index < 1 || index > n) return -1;
    if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    int64_t m = n + 1;
    unbwt_job base;
    memset(&base, 0, sizeof(base));
    base.L = L; base.n = n; base.index = index; base.m = m; base.T = T; base.threads = threads;
    base.lf = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)m);
    base.occ = (int64_t*)malloc(sizeof(int64_t) * 256 * (size_t)threads);
    base.chunk = (m + threads - 1) / threads;
    int64_t stride = 1;
    while (stride * threads * 256 < m) stride <<= 1;
    base.stride = stride;
    base.nseg = (m + stride - 1) / stride;
    base.seg_len = (int64_t*)malloc(sizeof(int64_t) * (size_t)base.nseg);
    base.seg_next = (int64_t*)malloc(sizeof(int64_t) * (size_t)base.nseg);
    base.seg_off = (int64_t*)malloc(sizeof(int64_t) * (size_t)base.nseg);
    unbwt_job* jobs = (unbwt_job*)malloc(sizeof(unbwt_job) * (size_t)threads);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    int rc = 0;
    if (!base.lf || !base.occ || !base.seg_len || !base.seg_next || !base.seg_off || !jobs || !th) { rc = -2; goto done; }
    for (int pass = 0; pass < 4; pass++) {
        for (int t = 0; t < threads; t++) {
            jobs[t] = base;
            jobs[t].tid = t;
            jobs[t].pass = pass;
            pthread_create(&th[t], NULL, unbwt_worker, &jobs[t]);
        }
        for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
        if (pass == 0) {
            /* first LF slot of symbol c in chunk t: 1 + #symbols < c + #c in earlier chunks */
            int64_t run = 1;
            for (int c = 0; c < 256; c++)
                for (int t = 0; t < threads; t++) {
                    int64_t v = base.occ[(int64_t)t * 256 + c];
                    base.occ[(int64_t)t * 256 + c] = run;
                    run += v;
                }
        } else if (pass == 2) {
            /* segment order along the cycle from row 0 */
            for (int64_t s = 0; s < base.nseg; s++) base.seg_off[s] = -1;
            int64_t pos = 0, s = 0, guard = 0;
            for (;;) {
                if (base.seg_off[s] >= 0 || guard++ > base.nseg) { rc = -1; goto done; }
                base.seg_off[s] = pos;
                pos += base.seg_len[s];
                if (base.seg_next[s] < 0) break;
                s = base.seg_next[s];
            }
            if (pos != n) { rc = -1; goto done; }
        }
    }
done:
    free(base.lf); free(base.occ); free(base.seg_len); free(base.seg_next); free(base.seg_off);
    free(jobs); free(th);
    return rc;
}

/* ------------------------------------------------------------------ 2. PFP restatement */

typedef struct {
    const uint8_t* T;
    int64_t n;
    int W;
    const int32_t* end;  /* phrase end position (inclusive, extended text) */
    const int32_t* len;  /* phrase length */
    const int32_t* key;  /* secondary key (tie break) */
} pfp_ctx;

/* symbol of the extended text T . $^W; the terminator is -1 (smaller than every byte) */
static inline int sym(const pfp_ctx* c, int64_t i) { return i < c->n ? (int)c->T[i] : -1; }

int64_t oracle_pfp_triggers(const uint8_t* T, int64_t n, int W, int P, int32_t* E)
{
    /* PFPk:443-497 / 500-571: the window T[e-W+1..e] is a trigger when its value
     * sum_k T[e-W+1+k]*2^(W-1-k) mod P is 0, or when it equals T[0..W-1];
     * the window starting at position 0 is excluded; only windows inside T. */
    int64_t t = 0;
    for (int64_t e = W; e <= n - 1; e++) {
        int64_t l = e - W + 1;
        int res = 0, pot = 1, equal = 1;
        for (int64_t i = e; i >= l; i--) {
            equal = equal && (T[i] == T[i - l]);
            res = (res + pot * (int)T[i]) % P;
            pot = (pot * 2) % P;
        }
        if (res == 0 || equal) {
            if (E) E[t] = (int32_t)e;
            t++;
        }
    }
    return t;
}

/* phrase ends and lengths, PFPk:573-592 (last phrase ends at n+W-1 inside the terminator) */
static int64_t pfp_phrases(const uint8_t* T, int64_t n, int W, int P, int32_t** end_out,
                           int32_t** len_out, int64_t* t_out)
{
    int64_t t = oracle_pfp_triggers(T, n, W, P, NULL);
    int64_t d = t + 1;
    int32_t* end = (int32_t*)malloc(sizeof(int32_t) * (size_t)d);
    int32_t* len = (int32_t*)malloc(sizeof(int32_t) * (size_t)d);
    oracle_pfp_triggers(T, n, W, P, end);
    end[d - 1] = (int32_t)(n + W - 1);
    for (int64_t k = 0; k < d; k++)
        len[k] = (k > 0) ? end[k] - end[k - 1] + W : end[k] + 1;
    *end_out = end;
    *len_out = len;
    if (t_out) *t_out = t;
    return d;
}

/* lexicographic comparison of two phrases (prefix-free, so ties mean equal strings) */
static int cmp_lex(const void* a, const void* b, void* arg)
{
    const pfp_ctx* c = (const pfp_ctx*)arg;
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    int64_t px = c->end[x] - c->len[x] + 1, py = c->end[y] - c->len[y] + 1;
    int32_t m = c->len[x] < c->len[y] ? c->len[x] : c->len[y];
    for (int32_t i = 0; i < m; i++) {
        int u = sym(c, px + i), v = sym(c, py + i);
        if (u != v) return u < v ? -1 : 1;
    }
    if (c->len[x] != c->len[y]) return c->len[x] < c->len[y] ? -1 : 1;
    return 0;
}

/* colex comparison; ties (equal phrases) broken by the preceding character, PFPk:1225-1237 */
static int cmp_colex(const void* a, const void* b, void* arg)
{
    const pfp_ctx* c = (const pfp_ctx*)arg;
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    int32_t m = c->len[x] < c->len[y] ? c->len[x] : c->len[y];
    for (int32_t i = 0; i < m; i++) {
        int u = sym(c, c->end[x] - i), v = sym(c, c->end[y] - i);
        if (u != v) return u < v ? -1 : 1;
    }
    if (c->len[x] != c->len[y]) return c->len[x] < c->len[y] ? -1 : 1;
    return c->key[x] - c->key[y];
}

typedef struct {
    int32_t idx;
    int32_t r1, r2;
} dbl_item;

static int cmp_dbl(const void* a, const void* b)
{
    const dbl_item* x = (const dbl_item*)a;
    const dbl_item* y = (const dbl_item*)b;
    if (x->r1 != y->r1) return x->r1 < y->r1 ? -1 : 1;
    if (x->r2 != y->r2) return x->r2 < y->r2 ? -1 : 1;
    return 0;
}

/* suffix array of an integer string by prefix doubling (PFPk:2342-2412); beyond-end = -1 */
static void parse_suffix_array(const int32_t* s, int64_t d, int32_t* sa, int32_t* isa)
{
    dbl_item* it = (dbl_item*)malloc(sizeof(dbl_item) * (size_t)d);
    int32_t* rank = (int32_t*)malloc(sizeof(int32_t) * (size_t)d);
    for (int64_t i = 0; i < d; i++) rank[i] = s[i];
    for (int64_t h = 1;; h <<= 1) {
        for (int64_t i = 0; i < d; i++) {
            it[i].idx = (int32_t)i;
            it[i].r1 = rank[i];
            it[i].r2 = (i + h < d) ? rank[i + h] : -1;
        }
        qsort(it, (size_t)d, sizeof(dbl_item), cmp_dbl);
        int32_t r = 0;
        int distinct = 1;
        for (int64_t i = 0; i < d; i++) {
            if (i > 0 && cmp_dbl(&it[i], &it[i - 1]) != 0) r = (int32_t)i;
            else if (i > 0) distinct = 0;
            rank[it[i].idx] = r;
        }
        if (distinct || h >= d) break;
    }
    for (int64_t i = 0; i < d; i++) {
        sa[i] = it[i].idx;
        if (isa) isa[it[i].idx] = (int32_t)i;
    }
    free(it);
    free(rank);
}

/* dense lexicographic phrase ranks per occurrence (PFPk:1990-2164, 616-691) */
static int64_t pfp_lex_ranks(pfp_ctx* c, int64_t d, int32_t* rank)
{
    int32_t* ord = (int32_t*)malloc(sizeof(int32_t) * (size_t)d);
    for (int64_t k = 0; k < d; k++) ord[k] = (int32_t)k;
    qsort_r(ord, (size_t)d, sizeof(int32_t), cmp_lex, c);
    int64_t r = -1;
    for (int64_t i = 0; i < d; i++) {
        if (i == 0 || cmp_lex(&ord[i - 1], &ord[i], c) != 0) r++;
        rank[ord[i]] = (int32_t)r;
    }
    free(ord);
    return r + 1;
}

int64_t oracle_pfp_parse(const uint8_t* T, int64_t n, int W, int P, int32_t* rank_out,
                         int32_t* psa_out, int64_t* distinct)
{
    if (n <= 0 || W < 1 || P < 1) return -1;
    int32_t *end, *len;
    int64_t d = pfp_phrases(T, n, W, P, &end, &len, NULL);
    pfp_ctx c = {T, n, W, end, len, NULL};
    int32_t* rank = (int32_t*)malloc(sizeof(int32_t) * (size_t)d);
    int64_t du = pfp_lex_ranks(&c, d, rank);
    if (distinct) *distinct = du;
    if (rank_out) memcpy(rank_out, rank, sizeof(int32_t) * (size_t)d);
    if (psa_out) parse_suffix_array(rank, d, psa_out, NULL);
    free(rank);
    free(end);
    free(len);
    return d;
}

typedef struct {
    int32_t colex;  /* first colex index of the interval (S_L) */
    int32_t length; /* suffix length (S_Length_Suffix) */
    int32_t ilen;   /* interval length - 1 (S_Interval_Length) */
} s_entry;

typedef struct {
    pfp_ctx* c;
    const int32_t* colex_ord;
} s_ctx;

/* lexicographic comparison of dictionary suffixes, PFPk:2767-3240 */
static int cmp_s(const void* a, const void* b, void* arg)
{
    const s_ctx* sc = (const s_ctx*)arg;
    const s_entry* x = (const s_entry*)a;
    const s_entry* y = (const s_entry*)b;
    int32_t ox = sc->colex_ord[x->colex], oy = sc->colex_ord[y->colex];
    int64_t px = sc->c->end[ox] - x->length + 1, py = sc->c->end[oy] - y->length + 1;
    int32_t m = x->length < y->length ? x->length : y->length;
    for (int32_t i = 0; i < m; i++) {
        int u = sym(sc->c, px + i), v = sym(sc->c, py + i);
        if (u != v) return u < v ? -1 : 1;
    }
    return x->length < y->length ? -1 : (x->length > y->length ? 1 : 0);
}

typedef struct {
    int32_t nr;
    uint8_t ch;
    uint8_t is0; /* this item is text position 0 */
} mixed_item;

static int cmp_mixed(const void* a, const void* b)
{
    int32_t x = ((const mixed_item*)a)->nr, y = ((const mixed_item*)b)->nr;
    return x < y ? -1 : (x > y ? 1 : 0);
}

int64_t oracle_pfp_bwt(const uint8_t* T, int64_t n, int W, int P, uint8_t* O,
                       struct oracle_pfp_info* info)
{
    if (n <= 0 || W < 1 || P < 1) return -1;
    int64_t t;
    int32_t *end, *len;
    int64_t d = pfp_phrases(T, n, W, P, &end, &len, &t);
    pfp_ctx c = {T, n, W, end, len, NULL};

    /* lexicographic ranks and the parse suffix array; nr[k] = rank of parse suffix k+1
     * (PFPk:787-807; the last phrase gets -1) */
    int32_t* rank = (int32_t*)malloc(sizeof(int32_t) * (size_t)d);
    int64_t du = pfp_lex_ranks(&c, d, rank);
    int32_t* psa = (int32_t*)malloc(sizeof(int32_t) * (size_t)d);
    int32_t* isa = (int32_t*)malloc(sizeof(int32_t) * (size_t)d);
    parse_suffix_array(rank, d, psa, isa);
    int32_t* nr = (int32_t*)malloc(sizeof(int32_t) * (size_t)d);
    for (int64_t k = 0; k + 1 < d; k++) nr[k] = isa[k + 1];
    nr[d - 1] = -1;

    /* colex order of the occurrences with the preceding character as tie break
     * (PFPk:2166-2340; the character before phrase 0 wraps to T[n-1]) */
    int32_t* prev = (int32_t*)malloc(sizeof(int32_t) * (size_t)d);
    for (int64_t k = 0; k < d; k++) {
        int64_t p = end[k] - len[k];
        prev[k] = (p >= 0) ? T[p] : T[n - 1];
    }
    c.key = prev;
    int32_t* col = (int32_t*)malloc(sizeof(int32_t) * (size_t)d);
    for (int64_t k = 0; k < d; k++) col[k] = (int32_t)k;
    qsort_r(col, (size_t)d, sizeof(int32_t), cmp_colex, &c);

    /* L values: common suffix length with the colex predecessor; new suffixes of
     * length (max(W, L), len] enter S (PFPk:2509-2592) */
    int32_t* Lv = (int32_t*)malloc(sizeof(int32_t) * (size_t)(d + 1));
    int64_t s_size = 0;
    for (int64_t i = 0; i < d; i++) {
        int32_t x = col[i];
        int32_t lcs = 0;
        if (i > 0) {
            int32_t y = col[i - 1];
            int32_t m = len[x] < len[y] ? len[x] : len[y];
            while (lcs < m && sym(&c, end[x] - lcs) == sym(&c, end[y] - lcs)) lcs++;
        }
        Lv[i] = lcs;
        int32_t lo = lcs > W ? lcs : W;
        if (len[x] > lo) s_size += len[x] - lo;
    }
    Lv[d] = -1;
    s_entry* S = (s_entry*)malloc(sizeof(s_entry) * (size_t)(s_size > 0 ? s_size : 1));
    int64_t si = 0;
    for (int64_t i = 0; i < d; i++) {
        int32_t x = col[i];
        int32_t lo = Lv[i] > W ? Lv[i] : W;
        for (int32_t m = lo + 1; m <= len[x]; m++) {
            /* interval: colex successors sharing the suffix, up to the first L < m
             * (PFPk:2598-2764) */
            int64_t h = i + 1;
            while (h < d && Lv[h] >= m) h++;
            S[si].colex = (int32_t)i;
            S[si].length = m;
            S[si].ilen = (int32_t)(h - 1 - i);
            si++;
        }
    }

    /* lexicographic order of S (PFPk:2767-3240, loop 4426-4552) */
    s_ctx sc = {&c, col};
    qsort_r(S, (size_t)s_size, sizeof(s_entry), cmp_s, &sc);

    /* BWT assembly (PFPk:3247-4325): intervals with one preceding character are filled
     * directly, the others are ordered by the rank of the next parse suffix */
    int64_t off = 0, index = -1, mixed_iv = 0, mixed_pos = 0;
    mixed_item* tmp = (mixed_item*)malloc(sizeof(mixed_item) * (size_t)d);
    for (int64_t s = 0; s < s_size; s++) {
        int64_t lo = S[s].colex, hi = lo + S[s].ilen;
        int32_t m = S[s].length;
        int64_t cnt = hi - lo + 1;
        for (int64_t i = lo; i <= hi; i++) {
            int32_t x = col[i];
            int64_t p = (int64_t)end[x] - m;
            tmp[i - lo].nr = nr[x];
            tmp[i - lo].ch = (p >= 0) ? T[p] : T[n - 1];
            tmp[i - lo].is0 = (p < 0);
        }
        int has0 = 0;
        for (int64_t i = 0; i < cnt; i++) has0 |= tmp[i].is0;
        if (tmp[0].ch != tmp[cnt - 1].ch) {
            mixed_iv++;
            mixed_pos += cnt;
        }
        /* the order only matters for mixed intervals; the interval that holds text
         * position 0 is ordered as well so that the primary index is known */
        if (tmp[0].ch != tmp[cnt - 1].ch || has0)
            qsort(tmp, (size_t)cnt, sizeof(mixed_item), cmp_mixed);
        for (int64_t i = 0; i < cnt; i++) {
            O[off + i] = tmp[i].ch;
            if (tmp[i].is0) index = off + i + 1;
        }
        off += cnt;
    }
    free(tmp);

    if (info) {
        info->t = t;
        info->d = d;
        info->distinct = du;
        info->s_size = s_size;
        info->mixed_intervals = mixed_iv;
        info->mixed_positions = mixed_pos;
    }
    free(S); free(Lv); free(col); free(prev); free(nr); free(isa); free(psa); free(rank);
    free(end); free(len);
    return (off == n) ? index : -3;
}

/* ------------------------------------------------------------------ 3. Lyndon factorisation */

int64_t oracle_lyndon_factors(const uint8_t* T, int64_t n, int64_t* F)
{
    /* Duval's algorithm */
    int64_t cnt = 0, i = 0;
    while (i < n) {
        int64_t j = i + 1, k = i;
        while (j < n && T[k] <= T[j]) {
            if (T[k] < T[j]) k = i;
            else k++;
            j++;
        }
        while (i <= k) {
            if (F) F[cnt] = i;
            cnt++;
            i += j - k;
        }
    }
    return cnt;
}

int oracle_lyndon_array(const uint8_t* T, int64_t n, int32_t* lam)
{
    /* right to left: the word starting at i absorbs following words while it is smaller */
    for (int64_t i = n - 1; i >= 0; i--) {
        int64_t j = i + 1;
        while (j < n) {
            /* compare w = T[i..j) with t = T[j..j+lam[j]) */
            int64_t lw = j - i, lt = lam[j], k = 0;
            while (k < lw && k < lt && T[i + k] == T[j + k]) k++;
            int smaller = (k < lw && k < lt) ? (T[i + k] < T[j + k]) : (lw < lt);
            if (!smaller) break;
            j += lt;
        }
        lam[i] = (int32_t)(j - i);
    }
    return 0;
}

typedef struct {
    const uint8_t* T;
    const int32_t* lam;
} lyn_ctx;

static int cmp_lyn(const void* a, const void* b, void* arg)
{
    const lyn_ctx* c = (const lyn_ctx*)arg;
    int64_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    int64_t lx = c->lam[x], ly = c->lam[y];
    int64_t m = lx < ly ? lx : ly;
    for (int64_t k = 0; k < m; k++)
        if (c->T[x + k] != c->T[y + k]) return c->T[x + k] < c->T[y + k] ? -1 : 1;
    return lx < ly ? -1 : (lx > ly ? 1 : 0);
}

int64_t oracle_lyndon_ranks(const uint8_t* T, int64_t n, int32_t* rank)
{
    if (n <= 0) return 0;
    int32_t* lam = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* ord = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    oracle_lyndon_array(T, n, lam);
    for (int64_t i = 0; i < n; i++) ord[i] = (int32_t)i;
    lyn_ctx c = {T, lam};
    qsort_r(ord, (size_t)n, sizeof(int32_t), cmp_lyn, &c);
    int64_t r = -1;
    for (int64_t i = 0; i < n; i++) {
        if (i == 0 || cmp_lyn(&ord[i - 1], &ord[i], &c) != 0) r++;
        rank[ord[i]] = (int32_t)r;
    }
    free(lam);
    free(ord);
    return r + 1;
}

int64_t oracle_lyndon_bwt(const uint8_t* T, int64_t n, uint8_t* O)
{
    if (n <= 0) return 0;
    int32_t* lam = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* rank = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    oracle_lyndon_array(T, n, lam);
    int64_t groups = oracle_lyndon_ranks(T, n, rank);
    /* bucket starts by rank group */
    int64_t* bucket = (int64_t*)calloc((size_t)groups + 1, sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) bucket[rank[i] + 1]++;
    for (int64_t g = 0; g < groups; g++) bucket[g + 1] += bucket[g];
    /* children lists: positions i whose following suffix starts at j = i + lam[i] */
    int64_t* cstart = (int64_t*)calloc((size_t)n + 2, sizeof(int64_t));
    int32_t* child = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    for (int64_t i = 0; i < n; i++) cstart[i + lam[i] + 1]++;
    for (int64_t j = 0; j <= n; j++) cstart[j + 1] += cstart[j];
    int64_t* cfill = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 1));
    memcpy(cfill, cstart, sizeof(int64_t) * ((size_t)n + 1));
    for (int64_t i = 0; i < n; i++) child[cfill[i + lam[i]]++] = (int32_t)i;
    int32_t* SA = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    /* the empty suffix ($) comes first: place the positions whose Lyndon prefix ends at n */
    for (int64_t c = cstart[n]; c < cstart[n + 1]; c++) {
        int32_t i = child[c];
        SA[bucket[rank[i]]++] = i;
    }
    for (int64_t s = 0; s < n; s++) {
        int32_t j = SA[s];
        for (int64_t c = cstart[j]; c < cstart[j + 1]; c++) {
            int32_t i = child[c];
            SA[bucket[rank[i]]++] = i;
        }
    }
    int64_t index = -1;
    for (int64_t s = 0; s < n; s++) {
        int64_t p = SA[s];
        if (p == 0) index = s + 1;
        O[s] = T[p == 0 ? n - 1 : p - 1];
    }
    free(SA); free(cfill); free(child); free(cstart); free(bucket); free(rank); free(lam);
    return index;
}

/* ------------------------------------------------------------------ 4. grammar induction
 *
 * The labels P_i = T[i .. i + lam[i]) are the maximal nodes of the Lyndon forest (LYN's rules
 * and stack words).  Label M of length >= 2 has the children i + 1, then each next child at the
 * end of the previous one, up to i + lam[i] (the Lyndon factorization of M without its first
 * symbol).  The roots are the Lyndon factorization of T.  A suffix compares like its sequence
 * of labels along the next-smaller-suffix chain, where a run of k equal labels compares as one
 * symbol (label, k); runs never reach past the label whose children they are.  Items are slots
 * (child k of distinct label d, or root s) standing for all positions in that role; a round
 * sorts the items of every group by (label rank, run length) of their next run, fills every new
 * group whose items share their preceding character, and moves the others past the run: to
 * the next slot of the label, or, at the label's end, to every slot that follows the label
 * where it occurs as a child (its up-contexts: the slots after the last-child chains of the
 * non-final children and of the roots; LYN:4114-4256 builds the same lists as "right parents").
 */

typedef struct {
    int64_t grp, key, q;
    int32_t ch;  /* preceding character; 256 added for the item of suffix 0 */
} lg_item;

static int cmp_lg_item(const void* a, const void* b)
{
    const lg_item* x = (const lg_item*)a;
    const lg_item* y = (const lg_item*)b;
    if (x->grp != y->grp) return x->grp < y->grp ? -1 : 1;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return 0;
}

static int cmp_i64(const void* a, const void* b)
{
    const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

int64_t oracle_lyndon_grammar_bwt(const uint8_t* T, int64_t n, uint8_t* O,
                                  struct oracle_lyndon_grammar_info* info)
{
    if (n <= 0) return 0;
    int32_t* lam = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* ord = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int64_t* lid = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    oracle_lyndon_array(T, n, lam);
    /* distinct labels in lexicographic order: the label id is its rank */
    for (int64_t i = 0; i < n; i++) ord[i] = (int32_t)i;
    lyn_ctx lc = {T, lam};
    qsort_r(ord, (size_t)n, sizeof(int32_t), cmp_lyn, &lc);
    int64_t D = 0;
    int64_t* rep = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int64_t* cnt = (int64_t*)calloc((size_t)n, sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) {
        if (i == 0 || cmp_lyn(&ord[i - 1], &ord[i], &lc) != 0) rep[D++] = ord[i];
        lid[ord[i]] = D - 1;
        cnt[D - 1]++;
    }
    /* child slots of every label, then the roots */
    int64_t* off = (int64_t*)calloc((size_t)D + 1, sizeof(int64_t));
    for (int64_t d = 0; d < D; d++) {
        int64_t t = 0;
        for (int64_t c = rep[d] + 1; c < rep[d] + lam[rep[d]]; c += lam[c]) t++;
        off[d + 1] = off[d] + t;
    }
    const int64_t C = off[D];
    int64_t S = 0;
    for (int64_t c = 0; c < n; c += lam[c]) S++;
    const int64_t Z = C + S;
    int64_t* chl = (int64_t*)malloc(sizeof(int64_t) * (size_t)(Z + 1));
    int64_t* cont = (int64_t*)malloc(sizeof(int64_t) * (size_t)(Z + 1));
    int32_t* chc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(Z + 1));
    int64_t* last = (int64_t*)malloc(sizeof(int64_t) * (size_t)D);
    for (int64_t d = 0; d < D; d++) {
        int64_t q = off[d];
        last[d] = -1;
        for (int64_t c = rep[d] + 1; c < rep[d] + lam[rep[d]]; c += lam[c], q++) {
            chl[q] = lid[c];
            chc[q] = T[c - 1];
            cont[q] = d;
            last[d] = lid[c];
        }
    }
    {
        int64_t q = C;
        for (int64_t c = 0; c < n; c += lam[c], q++) {
            chl[q] = lid[c];
            chc[q] = c ? T[c - 1] : T[n - 1];
            cont[q] = -1;
        }
    }
    /* runs of equal children: length from each slot and the slot after the run */
    int64_t* mrun = (int64_t*)malloc(sizeof(int64_t) * (size_t)(Z + 1));
    int64_t* nxt = (int64_t*)malloc(sizeof(int64_t) * (size_t)(Z + 1));
    for (int64_t q = Z - 1; q >= 0; q--) {
        const int same = q + 1 < Z && cont[q + 1] == cont[q] && chl[q + 1] == chl[q];
        mrun[q] = same ? mrun[q + 1] + 1 : 1;
        const int64_t nx = q + mrun[q];
        if (cont[q] < 0) nxt[q] = nx;
        else nxt[q] = (nx < Z && cont[nx] == cont[q]) ? nx : -1;
    }
    /* dense run keys (label rank, run length); 0 is the end of the text */
    int64_t* rkey = (int64_t*)malloc(sizeof(int64_t) * (size_t)(Z + 1));
    {
        int64_t* pk = (int64_t*)malloc(sizeof(int64_t) * (size_t)(Z + 1));
        int64_t* dr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(Z + 1));  /* distinct values before k */
        for (int64_t q = 0; q < Z; q++) pk[q] = chl[q] * (Z + 1) + mrun[q];
        qsort(pk, (size_t)Z, sizeof(int64_t), cmp_i64);
        for (int64_t k = 0; k < Z; k++) dr[k] = (k == 0) ? 0 : dr[k - 1] + (pk[k] != pk[k - 1]);
        for (int64_t q = 0; q < Z; q++) {
            const int64_t v = chl[q] * (Z + 1) + mrun[q];
            int64_t lo = 0, hi = Z;
            while (lo < hi) {
                const int64_t mid = (lo + hi) / 2;
                if (pk[mid] < v) lo = mid + 1;
                else hi = mid;
            }
            rkey[q] = dr[lo] + 1;
        }
        free(dr);
        free(pk);
    }
    rkey[Z] = 0;
    /* up-contexts: CSR lists per label (uo offsets, ul targets) */
    int64_t* uo = (int64_t*)calloc((size_t)D + 1, sizeof(int64_t));
    for (int64_t q = 0; q < Z; q++)
        if (cont[q] < 0 || (q + 1 < Z && cont[q + 1] == cont[q]))
            for (int64_t v = chl[q]; off[v + 1] > off[v]; v = last[v]) uo[v + 1]++;
    for (int64_t d = 0; d < D; d++) uo[d + 1] += uo[d];
    const int64_t UT = uo[D];
    int64_t* ul = (int64_t*)malloc(sizeof(int64_t) * (size_t)(UT + 1));
    {
        int64_t* fp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(D + 1));
        memcpy(fp, uo, sizeof(int64_t) * (size_t)(D + 1));
        for (int64_t q = 0; q < Z; q++)
            if (cont[q] < 0 || (q + 1 < Z && cont[q + 1] == cont[q]))
                for (int64_t v = chl[q]; off[v + 1] > off[v]; v = last[v]) ul[fp[v]++] = q + 1;
        free(fp);
    }
    /* the induction */
    int64_t m = Z, rounds = 0, items = 0, pidx = -1;
    lg_item* it = (lg_item*)malloc(sizeof(lg_item) * (size_t)(m + 1));
    for (int64_t q = 0; q < Z; q++) {
        it[q].grp = 0;
        it[q].q = q;
        it[q].ch = chc[q] + (q == C ? 256 : 0);  /* slot C: the root at position 0 */
    }
    while (m > 0) {
        rounds++;
        items += m;
        /* keys: the next run; a group holding only root items orders them by position (a later
         * root's suffix is smaller) */
        for (int64_t i = 0; i < m; i++) it[i].key = 0;
        qsort(it, (size_t)m, sizeof(lg_item), cmp_lg_item);
        for (int64_t a = 0; a < m;) {
            int64_t b = a, rule = 0;
            while (b < m && it[b].grp == it[a].grp) rule |= it[b++].q < C;
            for (int64_t i = a; i < b; i++) it[i].key = (it[i].q >= C && !rule) ? Z - it[i].q : rkey[it[i].q];
            a = b;
        }
        qsort(it, (size_t)m, sizeof(lg_item), cmp_lg_item);
        int64_t mo = 0, cap = m + 16;
        lg_item* nx = (lg_item*)malloc(sizeof(lg_item) * (size_t)cap);
        int64_t start = 0;
        for (int64_t b = 0; b < m;) {
            /* one new group: equal (group, key); its BWT range follows the earlier new groups of
             * the same old group */
            int64_t e = b, tot = 0;
            int uni = 1;
            if (b == 0 || it[b - 1].grp != it[b].grp) start = it[b].grp;
            while (e < m && it[e].grp == it[b].grp && it[e].key == it[b].key) {
                tot += (it[e].q < C) ? cnt[cont[it[e].q]] : 1;
                if (e > b && (it[e].ch != it[e - 1].ch || it[e].ch >= 256 || it[e - 1].ch >= 256)) uni = 0;
                e++;
            }
            if (uni) {
                memset(O + start, it[b].ch & 0xFF, (size_t)tot);
                if (it[b].ch >= 256) pidx = start + 1;
            } else {
                for (int64_t i = b; i < e; i++) {
                    const int64_t q = it[i].q;
                    const int64_t k0 = (nxt[q] >= 0) ? 0 : uo[cont[q]];
                    const int64_t k1 = (nxt[q] >= 0) ? 1 : uo[cont[q] + 1];
                    for (int64_t k = k0; k < k1; k++) {
                        if (mo >= cap) {
                            cap = 2 * cap;
                            nx = (lg_item*)realloc(nx, sizeof(lg_item) * (size_t)cap);
                        }
                        nx[mo].grp = start;
                        nx[mo].q = (nxt[q] >= 0) ? nxt[q] : ul[k];
                        nx[mo].ch = it[i].ch;
                        nx[mo].key = 0;
                        mo++;
                    }
                }
            }
            start += tot;
            b = e;
        }
        free(it);
        it = nx;
        m = mo;
    }
    free(it);
    if (info) {
        info->labels = D;
        info->slots = C;
        info->roots = S;
        info->ups = UT;
        info->rounds = rounds;
        info->items = items;
    }
    free(ul); free(uo); free(rkey); free(nxt); free(mrun); free(last); free(chc); free(cont); free(chl);
    free(off); free(cnt); free(rep); free(lid); free(ord); free(lam);
    return pidx;
}
