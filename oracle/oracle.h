/*
 * CPU oracle for the BWT-build path.  TEST INFRASTRUCTURE ONLY: the product
 * library never links or calls anything declared here; only tests/, the
 * smoke check and the CPU-baseline leg of bench.py load liboracle.so, and only
 * as the checker.
 *
 * Conventions (shared with the reference programs):
 *  - the text T[0..n) is a byte string; an implicit terminator "$" that is
 *    smaller than every byte follows it ("$-order": a proper prefix sorts
 *    before its extensions);
 *  - the reference output format is O[j] = T[(SA[j]-1) mod n] over the n
 *    non-empty suffixes (Prefix_Free_Parsing_W=5_P=16_kommentiert.cu:4776-4790,
 *    Lyndon_Grammar_BWT.cu:5530-5545);
 *  - the libcubwt convention is L[0] = T[n-1], primary index i = 1 + the rank
 *    of suffix 0, L[1..i-1] = O[0..i-2], L[i..n-1] = O[i..n-1]
 *    (libcubwt.cu:2242-2254, libcubwt.cu:2989-3021).
 */
#ifndef MIBWT_ORACLE_H
#define MIBWT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Suffix array of the n non-empty suffixes of T in $-order (SA-IS). 0 on success. */
int oracle_sa(const uint8_t* T, int64_t n, int32_t* SA);

/* Reference-format BWT O[0..n); returns the libcubwt primary index (>=1), or <0 on error. */
int64_t oracle_bwt(const uint8_t* T, int64_t n, uint8_t* O);

/* Convert between the reference format O and the libcubwt format (L, index). */
void oracle_ref_to_cubwt(const uint8_t* O, int64_t n, int64_t index, uint8_t* L);
void oracle_cubwt_to_ref(const uint8_t* L, int64_t n, int64_t index, uint8_t* O);

/* Inverse BWT of the libcubwt format (L, index) -> T.  0 on success. */
int oracle_unbwt(const uint8_t* L, int64_t n, int64_t index, uint8_t* T);

/* The same with `threads` POSIX threads (0: all online CPUs); same result and return codes. */
int oracle_unbwt_mt(const uint8_t* L, int64_t n, int64_t index, uint8_t* T, int threads);

/* --- prefix-free parsing restatement (Prefix_Free_Parsing_W=5_P=16_kommentiert.cu) --- */

/* Trigger window end positions e (W <= e <= n-1), ascending; E may be NULL to count only.
 * Returns t. */
int64_t oracle_pfp_triggers(const uint8_t* T, int64_t n, int W, int P, int32_t* E);

/* Parse summary: d occurrences, D_u distinct phrases; per occurrence the dense lexicographic
 * rank of its phrase (rank[d]) and the suffix array of the parse (psa[d]); either may be NULL.
 * Returns d. */
int64_t oracle_pfp_parse(const uint8_t* T, int64_t n, int W, int P,
                         int32_t* rank, int32_t* psa, int64_t* distinct);

struct oracle_pfp_info {
    int64_t t, d, distinct, s_size, mixed_intervals, mixed_positions;
};

/* Full PFP BWT on the CPU, reference format.  Returns the libcubwt primary index. */
int64_t oracle_pfp_bwt(const uint8_t* T, int64_t n, int W, int P, uint8_t* O,
                       struct oracle_pfp_info* info);

/* --- Lyndon factorisation (Duval) and Lyndon-grammar BWT restatement --- */

/* Lyndon factorisation of T: writes factor start positions to F (may be NULL); returns count. */
int64_t oracle_lyndon_factors(const uint8_t* T, int64_t n, int64_t* F);

/* Lyndon array: lam[i] = length of the longest Lyndon word starting at i (right-to-left
 * stack loop).  0 on success. */
int oracle_lyndon_array(const uint8_t* T, int64_t n, int32_t* lam);

/* Dense lexicographic rank of every position's longest Lyndon prefix (equal strings share a
 * rank); returns the number of distinct prefixes ("rules" plus stack words). */
int64_t oracle_lyndon_ranks(const uint8_t* T, int64_t n, int32_t* rank);

/* Lyndon-grammar BWT, reference format; returns the libcubwt primary index. */
int64_t oracle_lyndon_bwt(const uint8_t* T, int64_t n, uint8_t* O);

/* Sizes of the grammar induction: distinct labels, their child slots, roots (stack words),
 * up-context pairs, induction rounds and items summed over the rounds. */
struct oracle_lyndon_grammar_info {
    int64_t labels, slots, roots, ups, rounds, items;
};

/* The Lyndon-grammar BWT by induction over distinct labels (the grammar restated with exact
 * string comparisons), reference format; returns the libcubwt primary index and fills `info`
 * (may be NULL). */
int64_t oracle_lyndon_grammar_bwt(const uint8_t* T, int64_t n, uint8_t* O,
                                  struct oracle_lyndon_grammar_info* info);

#ifdef __cplusplus
}
#endif

#endif
