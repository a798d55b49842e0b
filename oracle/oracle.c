/*
 * oracle.c -- CPU restatement of the libsais 2.10.4 hot path (SA, GSA, BWT, PLCP,
 * LCP).  TEST INFRASTRUCTURE ONLY: the product library never links or calls it.
 *
 * The suffix array is built by a compact textbook SA-IS (classify S/L, place LMS,
 * induce L then S, name LMS substrings, recurse, place sorted LMS, induce again),
 * the algorithm family of libsais.c:6879-6923 (libsais_main_8u) without any of the
 * reference's block/cache/OpenMP machinery.  GSA separators are mapped to distinct
 * integer symbols ordered by position, which is the order libsais_gsa produces
 * (libsais.c:4038-4055 marker flip and first-LMS reinsertion, 6206 terminal
 * separator, 5484 no induction from separator-preceded positions).
 *
 * BWT, PLCP and LCP restate libsais.c:7110-7118, 8116-8238 and 8311-8338 line by
 * line in meaning.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

typedef int64_t I;

void oracle_freq(const uint8_t *T, int64_t n, int64_t *freq)
{
    memset(freq, 0, 256 * sizeof(int64_t));
    for (I i = 0; i < n; i++) freq[T[i]]++;
}

/* libsais.c:7036 (T[n-1] must be 0) and 6886-6889: after S/L typing, the zero
 * symbol may only occur as an S-type 0 preceded by an L-type letter, plus exactly
 * one L-type 0 (the terminal one).  For byte strings that is: T[0] != 0 and no two
 * adjacent zeros. */
int oracle_gsa_valid(const uint8_t *T, int64_t n)
{
    if (n > 0 && T[n - 1] != 0) return 0;
    if (n <= 1) return 1;
    if (T[0] == 0) return 0;
    for (I i = 0; i + 1 < n; i++)
        if (T[i] == 0 && T[i + 1] == 0) return 0;
    return 1;
}

/* ---------------------------------------------------------------- SA-IS ---- */

static void buckets(const I *s, I n, I K, I *bkt, int ends)
{
    memset(bkt, 0, (size_t)K * sizeof(I));
    for (I i = 0; i < n; i++) bkt[s[i]]++;
    I sum = 0;
    for (I c = 0; c < K; c++) {
        sum += bkt[c];
        bkt[c] = ends ? sum : sum - bkt[c];
    }
}

/* t[i] = 1 for S-type.  A virtual sentinel, smaller than every symbol, follows
 * s[n-1]; it is never stored. */
static int is_lms(const uint8_t *t, I i) { return i > 0 && t[i] && !t[i - 1]; }

static void induce(const I *s, I *SA, I n, I K, const uint8_t *t, I *bkt)
{
    buckets(s, n, K, bkt, 0);
    SA[bkt[s[n - 1]]++] = n - 1; /* induced by the virtual sentinel */
    for (I i = 0; i < n; i++) {
        I p = SA[i];
        if (p > 0 && !t[p - 1]) SA[bkt[s[p - 1]]++] = p - 1;
    }
    buckets(s, n, K, bkt, 1);
    for (I i = n - 1; i >= 0; i--) {
        I p = SA[i];
        if (p > 0 && t[p - 1]) SA[--bkt[s[p - 1]]] = p - 1;
    }
}

static int sais(const I *s, I *SA, I n, I K)
{
    if (n == 1) { SA[0] = 0; return 0; }
    uint8_t *t = (uint8_t *)malloc((size_t)n);
    I *bkt = (I *)malloc((size_t)K * sizeof(I));
    if (!t || !bkt) { free(t); free(bkt); return -2; }

    t[n - 1] = 0;
    for (I i = n - 2; i >= 0; i--) t[i] = (s[i] < s[i + 1]) || (s[i] == s[i + 1] && t[i + 1]);

    /* 1. LMS suffixes at bucket ends, then induce: sorts LMS substrings. */
    for (I i = 0; i < n; i++) SA[i] = -1;
    buckets(s, n, K, bkt, 1);
    for (I i = n - 1; i >= 1; i--)
        if (is_lms(t, i)) SA[--bkt[s[i]]] = i;
    induce(s, SA, n, K, t, bkt);

    /* 2. Compact sorted LMS positions into SA[0..m) and name the substrings. */
    I m = 0;
    for (I i = 0; i < n; i++)
        if (is_lms(t, SA[i])) SA[m++] = SA[i];
    for (I i = m; i < n; i++) SA[i] = -1;
    I names = 0, prev = -1;
    for (I i = 0; i < m; i++) {
        I pos = SA[i];
        int diff = (prev < 0);
        for (I d = 0; !diff; d++) {
            if (pos + d == n || prev + d == n) { diff = 1; break; }
            if (s[pos + d] != s[prev + d] || t[pos + d] != t[prev + d]) { diff = 1; break; }
            if (d > 0 && (is_lms(t, pos + d) || is_lms(t, prev + d))) {
                diff = !(is_lms(t, pos + d) && is_lms(t, prev + d));
                break;
            }
        }
        if (diff) { names++; prev = pos; }
        SA[m + pos / 2] = names - 1;
    }
    I j = n - 1;
    for (I i = n - 1; i >= m; i--)
        if (SA[i] >= 0) SA[j--] = SA[i];

    /* 3. Sort the reduced string (recursively when names repeat). */
    I *s1 = SA + n - m, *SA1 = SA;
    if (names < m) {
        int rc = sais(s1, SA1, m, names);
        if (rc) { free(t); free(bkt); return rc; }
    } else {
        for (I i = 0; i < m; i++) SA1[s1[i]] = i;
    }

    /* 4. Map back to text positions, place at bucket ends, final induce. */
    I k = 0;
    for (I i = 1; i < n; i++)
        if (is_lms(t, i)) s1[k++] = i;
    for (I i = 0; i < m; i++) SA1[i] = s1[SA1[i]];
    for (I i = m; i < n; i++) SA[i] = -1;
    buckets(s, n, K, bkt, 1);
    for (I i = m - 1; i >= 0; i--) {
        I p = SA[i];
        SA[i] = -1;
        SA[--bkt[s[p]]] = p;
    }
    induce(s, SA, n, K, t, bkt);

    free(t);
    free(bkt);
    return 0;
}

int oracle_sa(const uint8_t *T, int64_t *SA, int64_t n, int gsa)
{
    if (n <= 0) return 0;
    if (gsa && !oracle_gsa_valid(T, n)) return -1;
    if (n == 1) { SA[0] = 0; return 0; }
    I *s = (I *)malloc((size_t)n * sizeof(I));
    if (!s) return -2;
    I K = 256;
    if (gsa) {
        I z = 0;
        for (I i = 0; i < n; i++)
            if (T[i] == 0) s[i] = z++;
        for (I i = 0; i < n; i++)
            if (T[i] != 0) s[i] = z + T[i];
        K = z + 256;
    } else {
        for (I i = 0; i < n; i++) s[i] = T[i];
    }
    int rc = sais(s, SA, n, K);
    free(s);
    return rc;
}

/* Integer alphabet (libsais_int, libsais64_long): the order of suffixes depends
 * only on the order of the symbol values, so the values are replaced by their dense
 * ranks first and the same SA-IS runs on those. */
static int cmp_i64(const void *a, const void *b)
{
    I x = *(const I *)a, y = *(const I *)b;
    return (x > y) - (x < y);
}

int oracle_sa_int(const int64_t *T, int64_t *SA, int64_t n)
{
    if (n <= 0) return 0;
    if (n == 1) { SA[0] = 0; return 0; }
    I *v = (I *)malloc((size_t)n * sizeof(I));
    I *s = (I *)malloc((size_t)n * sizeof(I));
    if (!v || !s) { free(v); free(s); return -2; }
    memcpy(v, T, (size_t)n * sizeof(I));
    qsort(v, (size_t)n, sizeof(I), cmp_i64);
    I sigma = 0;
    for (I i = 0; i < n; i++)
        if (i == 0 || v[i] != v[i - 1]) v[sigma++] = v[i];
    for (I i = 0; i < n; i++) {
        I lo = 0, hi = sigma - 1;
        while (lo < hi) {
            I mid = (lo + hi) / 2;
            if (v[mid] < T[i]) lo = mid + 1;
            else hi = mid;
        }
        s[i] = lo;
    }
    free(v);
    int rc = sais(s, SA, n, sigma);
    free(s);
    return rc;
}

/* libsais.c:8399-8415 with compute_plcp_int (8263-8286): Phi, then Kasai over
 * integer symbols. */
void oracle_plcp_int(const int64_t *T, const int64_t *SA, int64_t *PLCP, int64_t n)
{
    if (n <= 0) return;
    if (n == 1) { PLCP[0] = 0; return; }
    I k = n;
    for (I i = 0; i < n; i++) { PLCP[SA[i]] = k; k = SA[i]; }
    I l = 0;
    for (I i = 0; i < n; i++) {
        k = PLCP[i];
        I m = n - (i > k ? i : k);
        while (l < m && T[i + l] == T[k + l]) l++;
        PLCP[i] = l;
        l -= (l != 0);
    }
}

/* 16-bit alphabet (libsais16.c): the same algorithms over uint16_t symbols.  GSA
 * validity uses the same rule as bytes (libsais16.c keeps the 8-bit checks with a
 * 16-bit zero separator). */
int oracle_gsa_valid16(const uint16_t *T, int64_t n)
{
    if (n > 0 && T[n - 1] != 0) return 0;
    if (n <= 1) return 1;
    if (T[0] == 0) return 0;
    for (I i = 0; i + 1 < n; i++)
        if (T[i] == 0 && T[i + 1] == 0) return 0;
    return 1;
}

int oracle_sa16(const uint16_t *T, int64_t *SA, int64_t n, int gsa)
{
    if (n <= 0) return 0;
    if (gsa && !oracle_gsa_valid16(T, n)) return -1;
    if (n == 1) { SA[0] = 0; return 0; }
    I *s = (I *)malloc((size_t)n * sizeof(I));
    if (!s) return -2;
    I K = 65536;
    if (gsa) {
        I z = 0;
        for (I i = 0; i < n; i++)
            if (T[i] == 0) s[i] = z++;
        for (I i = 0; i < n; i++)
            if (T[i] != 0) s[i] = z + T[i];
        K = z + 65536;
    } else {
        for (I i = 0; i < n; i++) s[i] = T[i];
    }
    int rc = sais(s, SA, n, K);
    free(s);
    return rc;
}

int64_t oracle_bwt16(const uint16_t *T, const int64_t *SA, uint16_t *U, int64_t n)
{
    if (n <= 0) return 0;
    if (n == 1) { U[0] = T[0]; return 1; }
    I j0 = -1;
    for (I j = 0; j < n; j++)
        if (SA[j] == 0) { j0 = j; break; }
    U[0] = T[n - 1];
    for (I j = 0; j < n; j++) {
        if (j == j0) continue;
        U[j + (j < j0)] = T[SA[j] - 1];
    }
    return j0 + 1;
}

void oracle_plcp16(const uint16_t *T, const int64_t *SA, int64_t *PLCP, int64_t n, int gsa)
{
    if (n <= 0) return;
    if (n == 1) { PLCP[0] = 0; return; }
    I k = n;
    for (I i = 0; i < n; i++) { PLCP[SA[i]] = k; k = SA[i]; }
    I l = 0;
    for (I i = 0; i < n; i++) {
        k = PLCP[i];
        I m = n - (i > k ? i : k);
        if (!gsa) {
            while (l < m && T[i + l] == T[k + l]) l++;
        } else {
            while (l < m && T[i + l] == T[k + l] && T[i + l] != 0) l++;
        }
        PLCP[i] = l;
        l -= (l != 0);
    }
}

int oracle_unbwt16(const uint16_t *U, uint16_t *out, int64_t n, int64_t p)
{
    if (n <= 0) return 0;
    if (n == 1) { out[0] = U[0]; return 0; }
    I *cnt = (I *)calloc(65536, sizeof(I));
    I *C = (I *)malloc(65536 * sizeof(I));
    I *seen = (I *)calloc(65536, sizeof(I));
    I *lf = (I *)malloc((size_t)(n + 1) * sizeof(I));
    if (!cnt || !C || !seen || !lf) { free(cnt); free(C); free(seen); free(lf); return -2; }
    for (I q = 0; q < n; q++) cnt[U[q]]++;
    I acc = 1;
    for (I c = 0; c < 65536; c++) { C[c] = acc; acc += cnt[c]; }
    for (I x = 0; x <= n; x++) {
        if (x == p) { lf[x] = 0; continue; }
        uint16_t c = U[x - (x > p)];
        lf[x] = C[c] + seen[c]++;
    }
    I x = 0;
    for (I t = n - 1; t >= 0; t--) {
        out[t] = U[x - (x > p)];
        x = lf[x];
        if (x == p) break;
    }
    free(cnt); free(C); free(seen); free(lf);
    return 0;
}

/* ------------------------------------------------------- SA verification ---- */

int oracle_check_sa(const uint8_t *T, const int64_t *SA, int64_t n, int gsa)
{
    if (n <= 0) return 1;
    I *isa = (I *)malloc((size_t)n * sizeof(I));
    if (!isa) return 0;
    for (I i = 0; i < n; i++) isa[i] = -1;
    int ok = 1;
    for (I j = 0; j < n && ok; j++) {
        I p = SA[j];
        if (p < 0 || p >= n || isa[p] >= 0) ok = 0;
        else isa[p] = j;
    }
    for (I j = 1; j < n && ok; j++) {
        I a = SA[j - 1], b = SA[j];
        if (T[a] < T[b]) continue;
        if (T[a] > T[b]) { ok = 0; break; }
        if (gsa && T[a] == 0) { ok = a < b; continue; }
        if (a + 1 == n) continue;           /* a is a proper prefix of b */
        if (b + 1 == n) { ok = 0; break; }  /* b is a proper prefix of a */
        ok = isa[a + 1] < isa[b + 1];
    }
    free(isa);
    return ok;
}

/* ---------------------------------------------------------- BWT / LCP ---- */

int64_t oracle_bwt(const uint8_t *T, const int64_t *SA, uint8_t *U, int64_t n)
{
    if (n <= 0) return 0;
    if (n == 1) { U[0] = T[0]; return 1; }
    I j0 = -1;
    for (I j = 0; j < n; j++)
        if (SA[j] == 0) { j0 = j; break; }
    /* libsais.c:7113-7117: U[0] = T[n-1]; rows before the primary shift right by
     * one, the primary row itself is dropped. */
    U[0] = T[n - 1];
    for (I j = 0; j < n; j++) {
        if (j == j0) continue;
        U[j + (j < j0)] = T[SA[j] - 1];
    }
    return j0 + 1;
}

int oracle_bwt_aux(const uint8_t *T, const int64_t *SA, uint8_t *U, int64_t n, int64_t r, int64_t *aux)
{
    if (n <= 1) {
        if (n == 1) U[0] = T[0];
        aux[0] = n;
        return 0;
    }
    oracle_bwt(T, SA, U, n);
    /* libsais.c:4811: when suffix p (p % r == 0) is placed at SA slot j, the
     * reference records the slot's 1-based position. */
    for (I j = 0; j < n; j++)
        if (SA[j] % r == 0) aux[SA[j] / r] = j + 1;
    return 0;
}

/* Rows 0..n of the sorted rotations of T$: row 0 starts with $, and row x >= 1 is
 * SA slot x-1.  The last column is L[0] = U[0], L[p] = $ and L[x] = U[x - (x > p)]
 * otherwise.  LF(x) = 1 + (number of bytes < L[x]) + (occurrences of L[x] in
 * L[0..x)), and walking LF from row 0 reads T backwards. */
int oracle_unbwt(const uint8_t *U, uint8_t *out, int64_t n, int64_t p)
{
    if (n <= 0) return 0;
    if (n == 1) { out[0] = U[0]; return 0; }
    I cnt[256] = {0}, C[256];
    for (I q = 0; q < n; q++) cnt[U[q]]++;
    I acc = 1;
    for (int c = 0; c < 256; c++) { C[c] = acc; acc += cnt[c]; }
    I *lf = (I *)malloc((size_t)(n + 1) * sizeof(I));
    if (!lf) return -2;
    I seen[256] = {0};
    for (I x = 0; x <= n; x++) {
        if (x == p) { lf[x] = 0; continue; }
        uint8_t c = U[x - (x > p)];
        lf[x] = C[c] + seen[c]++;
    }
    I x = 0;
    for (I t = n - 1; t >= 0; t--) {
        out[t] = U[x - (x > p)];
        x = lf[x];
        if (x == p) break;  /* only for an inconsistent (U, p): stop at the $ row */
    }
    free(lf);
    return 0;
}

void oracle_plcp(const uint8_t *T, const int64_t *SA, int64_t *PLCP, int64_t n, int gsa)
{
    if (n <= 0) return;
    if (n == 1) { PLCP[0] = 0; return; }
    I k = n;
    for (I i = 0; i < n; i++) { PLCP[SA[i]] = k; k = SA[i]; } /* phi */
    I l = 0;
    for (I i = 0; i < n; i++) {
        k = PLCP[i];
        if (!gsa) {
            I m = n - (i > k ? i : k);
            while (l < m && T[i + l] == T[k + l]) l++;
        } else {
            /* The separator stops the match; i + l and k + l stay below n for a
             * valid GSA input because T[n-1] == 0. */
            while (i + l < n && k + l < n && T[i + l] > 0 && T[i + l] == T[k + l]) l++;
        }
        PLCP[i] = l;
        l -= (l != 0);
    }
}

void oracle_lcp(const int64_t *PLCP, const int64_t *SA, int64_t *LCP, int64_t n)
{
    for (I i = 0; i < n; i++) LCP[i] = PLCP[SA[i]];
}

/* ------------------------------------------------------ 32-bit wrappers ---- */

static I *widen(const int32_t *a, I n)
{
    I *w = (I *)calloc((size_t)(n > 0 ? n : 1), sizeof(I));
    if (!w) return NULL;
    for (I i = 0; i < n; i++) w[i] = a[i];
    return w;
}

int oracle_sa32(const uint8_t *T, int32_t *SA, int64_t n, int gsa)
{
    I *w = (I *)malloc((size_t)(n > 0 ? n : 1) * sizeof(I));
    if (!w) return -2;
    int rc = oracle_sa(T, w, n, gsa);
    if (rc == 0) for (I i = 0; i < n; i++) SA[i] = (int32_t)w[i];
    free(w);
    return rc;
}

int oracle_check_sa32(const uint8_t *T, const int32_t *SA, int64_t n, int gsa)
{
    I *w = widen(SA, n);
    if (!w) return 0;
    int ok = oracle_check_sa(T, w, n, gsa);
    free(w);
    return ok;
}

int64_t oracle_bwt32(const uint8_t *T, const int32_t *SA, uint8_t *U, int64_t n)
{
    I *w = widen(SA, n);
    if (!w) return -2;
    I r = oracle_bwt(T, w, U, n);
    free(w);
    return r;
}

void oracle_plcp32(const uint8_t *T, const int32_t *SA, int32_t *PLCP, int64_t n, int gsa)
{
    I *w = widen(SA, n);
    I *p = (I *)calloc((size_t)(n > 0 ? n : 1), sizeof(I));
    if (w && p) {
        oracle_plcp(T, w, p, n, gsa);
        for (I i = 0; i < n; i++) PLCP[i] = (int32_t)p[i];
    }
    free(w);
    free(p);
}

void oracle_lcp32(const int32_t *PLCP, const int32_t *SA, int32_t *LCP, int64_t n)
{
    for (I i = 0; i < n; i++) LCP[i] = PLCP[SA[i]];
}
