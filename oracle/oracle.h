/*
 * oracle.h -- CPU restatement of libsais 2.10.4 used ONLY as test infrastructure.
 *
 * Nothing in the product library links, loads or calls this code.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, and only
 * as the checker.  Every function cites the reference lines whose output it
 * restates; all indices are int64_t so one build serves the 32- and 64-bit APIs.
 *
 * Parity pinning (what checks each function, see tests/test_oracle.py):
 *   oracle_sa          brute force (exhaustive tiny strings), the compiled
 *                      reference (oracle/_ref), closed forms (all-A, banana,
 *                      mississippi), oracle_check_sa.
 *   oracle_check_sa    planted mistakes (swapped neighbours, non-permutation).
 *   oracle_bwt         compiled reference libsais_bwt, banana/mississippi.
 *   oracle_plcp/lcp    compiled reference, direct per-pair comparison, closed forms.
 *   oracle_gsa_valid   compiled reference over all {0,1,2}-strings up to length 10.
 *   oracle_bwt_aux     compiled reference libsais_bwt_aux (U and I) over sampling
 *                      rates 2..n, hand-worked banana samples.
 *   oracle_sa_int      brute force over small integer strings, the compiled
 *                      reference libsais_int / libsais64_long.
 *   oracle_plcp_int    the compiled reference libsais_plcp_int, direct comparison.
 *   oracle_*16         the compiled reference libsais16_* on random 16-bit strings
 *                      (small and large alphabets, GSA sets), brute force on tiny ones.
 *   oracle_unbwt       round trip T -> oracle_bwt -> oracle_unbwt, the compiled
 *                      reference libsais_unbwt / libsais_unbwt_aux, banana by hand.
 */
#ifndef SAGPU_ORACLE_H
#define SAGPU_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* freq[c] = number of bytes equal to c (libsais.c:1398-1427, freq[j] = total). */
void    oracle_freq(const uint8_t *T, int64_t n, int64_t *freq);

/* 1 if T is a valid GSA input (libsais.c:7036 and the type-count test at 6886-6889). */
int     oracle_gsa_valid(const uint8_t *T, int64_t n);

/* Suffix array by SA-IS.  gsa=1: every 0 byte is a distinct separator, smaller than
 * every letter and ordered by position (libsais.c:4038-4055, 6206, 5484).
 * Returns 0, or -1 for an invalid GSA input, -2 on allocation failure. */
int     oracle_sa(const uint8_t *T, int64_t *SA, int64_t n, int gsa);

/* BWT from a suffix array (libsais.c:7110-7118): writes U[0..n-1], returns the
 * primary index (1 + row of suffix 0). */
int64_t oracle_bwt(const uint8_t *T, const int64_t *SA, uint8_t *U, int64_t n);

/* Suffix array of an integer string (libsais_int / libsais64_long, any values:
 * suffixes compare by symbol value, a proper prefix first).  Returns 0 or -2. */
int     oracle_sa_int(const int64_t *T, int64_t *SA, int64_t n);

/* PLCP of an integer string (libsais.c:8399-8415, 8263-8286). */
void    oracle_plcp_int(const int64_t *T, const int64_t *SA, int64_t *PLCP, int64_t n);

/* 16-bit alphabet (libsais16.c): SA / GSA, BWT, PLCP / PLCP-GSA and inverse BWT over
 * uint16_t symbols, with the same conventions as the byte functions. */
int     oracle_gsa_valid16(const uint16_t *T, int64_t n);
int     oracle_sa16(const uint16_t *T, int64_t *SA, int64_t n, int gsa);
int64_t oracle_bwt16(const uint16_t *T, const int64_t *SA, uint16_t *U, int64_t n);
void    oracle_plcp16(const uint16_t *T, const int64_t *SA, int64_t *PLCP, int64_t n, int gsa);
int     oracle_unbwt16(const uint16_t *U, uint16_t *out, int64_t n, int64_t p);

/* BWT with auxiliary indexes (libsais.c:7123-7146 with the sampling of
 * 4799-4819): U as oracle_bwt, and I[k] = 1 + row of suffix k*r for
 * k = 0..(n-1)/r, so I[0] is the primary index.  Returns 0. */
int     oracle_bwt_aux(const uint8_t *T, const int64_t *SA, uint8_t *U, int64_t n, int64_t r, int64_t *I);

/* Inverse BWT (libsais.c:8030-8055 unbwt / unbwt_aux output for a valid input):
 * U is the libsais BWT layout with primary index p (1-based row of suffix 0);
 * writes the original text to out[0..n-1] by the textbook LF walk from row 0.
 * Returns 0, or -2 on allocation failure. */
int     oracle_unbwt(const uint8_t *U, uint8_t *out, int64_t n, int64_t p);

/* Phi + Kasai PLCP exactly as libsais.c:8116-8142 (phi), 8167-8190 (plcp) and
 * 8215-8238 (plcp_gsa, gsa=1). */
void    oracle_plcp(const uint8_t *T, const int64_t *SA, int64_t *PLCP, int64_t n, int gsa);

/* LCP[i] = PLCP[SA[i]] (libsais.c:8311-8338). */
void    oracle_lcp(const int64_t *PLCP, const int64_t *SA, int64_t *LCP, int64_t n);

/* O(n) suffix-array verifier.  Returns 1 when SA is the (generalized) suffix array
 * of T, 0 otherwise. */
int     oracle_check_sa(const uint8_t *T, const int64_t *SA, int64_t n, int gsa);

/* 32-bit convenience wrappers (same semantics, int32 arrays). */
int     oracle_sa32(const uint8_t *T, int32_t *SA, int64_t n, int gsa);
int     oracle_check_sa32(const uint8_t *T, const int32_t *SA, int64_t n, int gsa);
int64_t oracle_bwt32(const uint8_t *T, const int32_t *SA, uint8_t *U, int64_t n);
void    oracle_plcp32(const uint8_t *T, const int32_t *SA, int32_t *PLCP, int64_t n, int gsa);
void    oracle_lcp32(const int32_t *PLCP, const int32_t *SA, int32_t *LCP, int64_t n);

#ifdef __cplusplus
}
#endif

#endif
