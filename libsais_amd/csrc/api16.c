/*
 * api16.c -- libsais16 / libsais16x64 2.10.4 compatible entry points (C99 host code):
 * the same entry points as api.c over 16-bit symbols (freq tables of 65536 entries).
 * Argument checks and n <= 1 cases follow the reference line for line; the work
 * goes to the GPU engine through its host-buffer layer (sagpu_host_*16, engine.hip).
 * Return codes: 0 / primary index on success, -1 invalid arguments, -2 device
 * failure.  Contexts are those of api.c (libsais_create_ctx): each owns an engine
 * context, bound to the call by WITH_CTX.
 */
#define LIBSAIS_OPENMP 1 /* declare the _omp forms, which this library defines */
#include "../../include/libsais.h"
#include "../../include/libsais64.h"
#include "../../include/libsais16.h"
#include "../../include/libsais16x64.h"
#include "api_ctx.h"

#include <stddef.h>
#include <string.h>

#define HIDDEN __attribute__((visibility("hidden")))
#define ALPHABET16 65536

HIDDEN int64_t sagpu_host_sa16(const uint16_t *T, void *SA, int64_t n, int32_t gsa, void *freq, int32_t width);
HIDDEN int64_t sagpu_host_bwt16(const uint16_t *T, uint16_t *U, int64_t n, void *freq, int64_t r, void *I,
                                int32_t width);
HIDDEN int64_t sagpu_host_plcp16(const uint16_t *T, const void *SA, void *PLCP, int64_t n, int32_t gsa, int32_t width);
HIDDEN int64_t sagpu_host_unbwt16(const uint16_t *T, uint16_t *U, int64_t n, int64_t r, const void *I, int32_t width);
HIDDEN int64_t sagpu_host_lcp(const void *PLCP, const void *SA, void *LCP, int64_t n, int32_t width);

static int64_t map_rc(int64_t rc) { return rc < -2 ? -2 : rc; }

/* ------------------------------------------------------------ libsais16 -- */

/* reference src/libsais16.c:6995-7009 */
int32_t libsais16(const uint16_t *T, int32_t *SA, int32_t n, int32_t fs, int32_t *freq)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (fs < 0)) return -1;
    if (n <= 1) {
        if (freq != NULL) memset(freq, 0, ALPHABET16 * sizeof(int32_t));
        if (n == 1) { SA[0] = 0; if (freq != NULL) freq[T[0]]++; }
        return 0;
    }
    return (int32_t)map_rc(sagpu_host_sa16(T, SA, n, 0, freq, 4));
}

/* reference src/libsais16.c:7011-7025 */
int32_t libsais16_gsa(const uint16_t *T, int32_t *SA, int32_t n, int32_t fs, int32_t *freq)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (n > 0 && T[n - 1] != 0) || (fs < 0)) return -1;
    if (n <= 1) {
        if (freq != NULL) memset(freq, 0, ALPHABET16 * sizeof(int32_t));
        if (n == 1) { SA[0] = 0; if (freq != NULL) freq[T[0]]++; }
        return 0;
    }
    return (int32_t)map_rc(sagpu_host_sa16(T, SA, n, 1, freq, 4));
}

/* reference src/libsais16.c:7027-7040.  The same integer-alphabet construction as the byte library's. */
int32_t libsais16_int(int32_t *T, int32_t *SA, int32_t n, int32_t k, int32_t fs)
{
    return libsais_int(T, SA, n, k, fs);
}

/* reference src/libsais16.c:7074-7098 */
int32_t libsais16_bwt(const uint16_t *T, uint16_t *U, int32_t *A, int32_t n, int32_t fs, int32_t *freq)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || (fs < 0)) return -1;
    if (n <= 1) {
        if (freq != NULL) memset(freq, 0, ALPHABET16 * sizeof(int32_t));
        if (n == 1) { U[0] = T[0]; if (freq != NULL) freq[T[0]]++; }
        return n;
    }
    return (int32_t)map_rc(sagpu_host_bwt16(T, U, n, freq, 0, NULL, 4));
}

/* reference src/libsais16.c:7100-7123 */
int32_t libsais16_bwt_aux(const uint16_t *T, uint16_t *U, int32_t *A, int32_t n, int32_t fs, int32_t *freq, int32_t r, int32_t *I)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || (fs < 0) || (r < 2) || ((r & (r - 1)) != 0) ||
        (I == NULL))
        return -1;
    if (n <= 1) {
        if (freq != NULL) memset(freq, 0, ALPHABET16 * sizeof(int32_t));
        if (n == 1) { U[0] = T[0]; if (freq != NULL) freq[T[0]]++; }
        I[0] = n;
        return 0;
    }
    return (int32_t)map_rc(sagpu_host_bwt16(T, U, n, freq, r, I, 4));
}

/* reference src/libsais16.c:7789-7805.  The freq table is not needed: counts come from T. */
int32_t libsais16_unbwt_aux(const uint16_t *T, uint16_t *U, int32_t *A, int32_t n, const int32_t *freq, int32_t r, const int32_t *I)
{
    (void)freq;
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || ((r != n) && ((r < 2) || ((r & (r - 1)) != 0))) ||
        (I == NULL))
        return -1;
    if (n <= 1) {
        if (I[0] != n) return -1;
        if (n == 1) U[0] = T[0];
        return 0;
    }
    for (int64_t t = 0; t <= (int64_t)(n - 1) / r; ++t)
        if (I[t] <= 0 || I[t] > n) return -1;
    return (int32_t)map_rc(sagpu_host_unbwt16(T, U, n, r, I, 4));
}

/* reference src/libsais16.c:7779-7782 */
int32_t libsais16_unbwt(const uint16_t *T, uint16_t *U, int32_t *A, int32_t n, const int32_t *freq, int32_t i)
{
    return libsais16_unbwt_aux(T, U, A, n, freq, n, &i);
}

/* reference src/libsais16.c:8064-8080 */
int32_t libsais16_plcp(const uint16_t *T, const int32_t *SA, int32_t *PLCP, int32_t n)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0)) return -1;
    if (n <= 1) {
        if (n == 1) PLCP[0] = 0;
        return 0;
    }
    return (int32_t)map_rc(sagpu_host_plcp16(T, SA, PLCP, n, 0, 4));
}

/* reference src/libsais16.c:8082-8098 */
int32_t libsais16_plcp_gsa(const uint16_t *T, const int32_t *SA, int32_t *PLCP, int32_t n)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0) || (n > 0 && T[n - 1] != 0)) return -1;
    if (n <= 1) {
        if (n == 1) PLCP[0] = 0;
        return 0;
    }
    return (int32_t)map_rc(sagpu_host_plcp16(T, SA, PLCP, n, 1, 4));
}

/* reference src/libsais16.c:8100-8115 */
int32_t libsais16_lcp(const int32_t *PLCP, const int32_t *SA, int32_t *LCP, int32_t n)
{
    if ((PLCP == NULL) || (SA == NULL) || (LCP == NULL) || (n < 0)) return -1;
    if (n <= 1) {
        if (n == 1) LCP[0] = PLCP[SA[0]];
        return 0;
    }
    return (int32_t)map_rc(sagpu_host_lcp(PLCP, SA, LCP, n, 4));
}

/* ------------------------------------------------------ libsais16 contexts -- */
/* Contexts are the tokens of the byte library (device buffers are pooled inside the
 * engine); the calls behave as the context-free ones after the same checks. */

/* reference src/libsais16.c:6985-6988 */
void *libsais16_create_ctx(void) { return libsais_create_ctx(); }

/* reference src/libsais16.c:7178-7186 */
void *libsais16_create_ctx_omp(int32_t threads) { return libsais_create_ctx_omp(threads); }

/* reference src/libsais16.c:6990-6993 */
void libsais16_free_ctx(void *ctx) { libsais_free_ctx(ctx); }

/* reference src/libsais16.c:7769-7772 */
void *libsais16_unbwt_create_ctx(void) { return libsais_unbwt_create_ctx(); }

/* reference src/libsais16.c:7827-7835 */
void *libsais16_unbwt_create_ctx_omp(int32_t threads) { return libsais_unbwt_create_ctx_omp(threads); }

/* reference src/libsais16.c:7774-7777 */
void libsais16_unbwt_free_ctx(void *ctx) { libsais_unbwt_free_ctx(ctx); }

/* reference src/libsais16.c:7042-7056 */
int32_t libsais16_ctx(const void *ctx, const uint16_t *T, int32_t *SA, int32_t n, int32_t fs, int32_t *freq)
{
    if (ctx == NULL) return -1;
    WITH_CTX(ctx, int32_t, libsais16(T, SA, n, fs, freq));
}

/* reference src/libsais16.c:7058-7072 */
int32_t libsais16_gsa_ctx(const void *ctx, const uint16_t *T, int32_t *SA, int32_t n, int32_t fs, int32_t *freq)
{
    if (ctx == NULL) return -1;
    WITH_CTX(ctx, int32_t, libsais16_gsa(T, SA, n, fs, freq));
}

/* reference src/libsais16.c:7125-7149 */
int32_t libsais16_bwt_ctx(const void *ctx, const uint16_t *T, uint16_t *U, int32_t *A, int32_t n, int32_t fs,
                          int32_t *freq)
{
    if (ctx == NULL) return -1;
    WITH_CTX(ctx, int32_t, libsais16_bwt(T, U, A, n, fs, freq));
}

/* reference src/libsais16.c:7151-7174 */
int32_t libsais16_bwt_aux_ctx(const void *ctx, const uint16_t *T, uint16_t *U, int32_t *A, int32_t n, int32_t fs,
                              int32_t *freq, int32_t r, int32_t *I)
{
    if (ctx == NULL) return -1;
    WITH_CTX(ctx, int32_t, libsais16_bwt_aux(T, U, A, n, fs, freq, r, I));
}

/* reference src/libsais16.c:7807-7823 */
int32_t libsais16_unbwt_aux_ctx(const void *ctx, const uint16_t *T, uint16_t *U, int32_t *A, int32_t n,
                                const int32_t *freq, int32_t r, const int32_t *I)
{
    if (ctx == NULL) return -1;
    WITH_CTX(ctx, int32_t, libsais16_unbwt_aux(T, U, A, n, freq, r, I));
}

/* reference src/libsais16.c:7784-7787 */
int32_t libsais16_unbwt_ctx(const void *ctx, const uint16_t *T, uint16_t *U, int32_t *A, int32_t n,
                            const int32_t *freq, int32_t i)
{
    return libsais16_unbwt_aux_ctx(ctx, T, U, A, n, freq, n, &i);
}

/* ---------------------------------------------------- libsais16 OpenMP -- */
/* Thread counts are checked (< 0 rejected); the work runs on the GPU. */

/* reference src/libsais16.c:7188-7205 */
int32_t libsais16_omp(const uint16_t *T, int32_t *SA, int32_t n, int32_t fs, int32_t *freq, int32_t threads)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (fs < 0) || (threads < 0)) return -1;
    return libsais16(T, SA, n, fs, freq);
}

/* reference src/libsais16.c:7207-7224 */
int32_t libsais16_gsa_omp(const uint16_t *T, int32_t *SA, int32_t n, int32_t fs, int32_t *freq, int32_t threads)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (n > 0 && T[n - 1] != 0) || (fs < 0) || (threads < 0)) return -1;
    return libsais16_gsa(T, SA, n, fs, freq);
}

/* reference src/libsais16.c:7244-7271 */
int32_t libsais16_bwt_omp(const uint16_t *T, uint16_t *U, int32_t *A, int32_t n, int32_t fs, int32_t *freq, int32_t threads)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || (fs < 0) || (threads < 0)) return -1;
    return libsais16_bwt(T, U, A, n, fs, freq);
}

/* reference src/libsais16.c:7273-7299 */
int32_t libsais16_bwt_aux_omp(const uint16_t *T, uint16_t *U, int32_t *A, int32_t n, int32_t fs, int32_t *freq, int32_t r, int32_t *I,
                      int32_t threads)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || (fs < 0) || (r < 2) || ((r & (r - 1)) != 0) ||
        (I == NULL) || (threads < 0))
        return -1;
    return libsais16_bwt_aux(T, U, A, n, fs, freq, r, I);
}

/* reference src/libsais16.c:7842-7861 */
int32_t libsais16_unbwt_aux_omp(const uint16_t *T, uint16_t *U, int32_t *A, int32_t n, const int32_t *freq, int32_t r, const int32_t *I,
                        int32_t threads)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || ((r != n) && ((r < 2) || ((r & (r - 1)) != 0))) ||
        (I == NULL) || (threads < 0))
        return -1;
    return libsais16_unbwt_aux(T, U, A, n, freq, r, I);
}

/* reference src/libsais16.c:7837-7840 */
int32_t libsais16_unbwt_omp(const uint16_t *T, uint16_t *U, int32_t *A, int32_t n, const int32_t *freq, int32_t i, int32_t threads)
{
    return libsais16_unbwt_aux_omp(T, U, A, n, freq, n, &i, threads);
}

/* reference src/libsais16.c:8119-8138 */
int32_t libsais16_plcp_omp(const uint16_t *T, const int32_t *SA, int32_t *PLCP, int32_t n, int32_t threads)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0) || (threads < 0)) return -1;
    return libsais16_plcp(T, SA, PLCP, n);
}

/* reference src/libsais16.c:8140-8159 */
int32_t libsais16_plcp_gsa_omp(const uint16_t *T, const int32_t *SA, int32_t *PLCP, int32_t n, int32_t threads)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0) || (n > 0 && T[n - 1] != 0) || (threads < 0))
        return -1;
    return libsais16_plcp_gsa(T, SA, PLCP, n);
}

/* reference src/libsais16.c:8161-8179 */
int32_t libsais16_lcp_omp(const int32_t *PLCP, const int32_t *SA, int32_t *LCP, int32_t n, int32_t threads)
{
    if ((PLCP == NULL) || (SA == NULL) || (LCP == NULL) || (n < 0) || (threads < 0)) return -1;
    return libsais16_lcp(PLCP, SA, LCP, n);
}

/* reference src/libsais16.c:7226-7242 */
int32_t libsais16_int_omp(int32_t *T, int32_t *SA, int32_t n, int32_t k, int32_t fs, int32_t threads)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (fs < 0) || (threads < 0)) return -1;
    return libsais16_int(T, SA, n, k, fs);
}

/* ------------------------------------------------------------ libsais16x64 -- */

/* reference src/libsais16x64.c:7036-7064 */
int64_t libsais16x64(const uint16_t *T, int64_t *SA, int64_t n, int64_t fs, int64_t *freq)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (fs < 0)) return -1;
    if (n <= 1) {
        if (freq != NULL) memset(freq, 0, ALPHABET16 * sizeof(int64_t));
        if (n == 1) { SA[0] = 0; if (freq != NULL) freq[T[0]]++; }
        return 0;
    }
    return map_rc(sagpu_host_sa16(T, SA, n, 0, freq, 8));
}

/* reference src/libsais16x64.c:7066-7094 */
int64_t libsais16x64_gsa(const uint16_t *T, int64_t *SA, int64_t n, int64_t fs, int64_t *freq)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (n > 0 && T[n - 1] != 0) || (fs < 0)) return -1;
    if (n <= 1) {
        if (freq != NULL) memset(freq, 0, ALPHABET16 * sizeof(int64_t));
        if (n == 1) { SA[0] = 0; if (freq != NULL) freq[T[0]]++; }
        return 0;
    }
    return map_rc(sagpu_host_sa16(T, SA, n, 1, freq, 8));
}

/* reference src/libsais16x64.c:7096-7109.  The same integer-alphabet construction as the byte library's. */
int64_t libsais16x64_long(int64_t *T, int64_t *SA, int64_t n, int64_t k, int64_t fs)
{
    return libsais64_long(T, SA, n, k, fs);
}

/* reference src/libsais16x64.c:7111-7148 */
int64_t libsais16x64_bwt(const uint16_t *T, uint16_t *U, int64_t *A, int64_t n, int64_t fs, int64_t *freq)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || (fs < 0)) return -1;
    if (n <= 1) {
        if (freq != NULL) memset(freq, 0, ALPHABET16 * sizeof(int64_t));
        if (n == 1) { U[0] = T[0]; if (freq != NULL) freq[T[0]]++; }
        return n;
    }
    return map_rc(sagpu_host_bwt16(T, U, n, freq, 0, NULL, 8));
}

/* reference src/libsais16x64.c:7150-7187 */
int64_t libsais16x64_bwt_aux(const uint16_t *T, uint16_t *U, int64_t *A, int64_t n, int64_t fs, int64_t *freq, int64_t r, int64_t *I)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || (fs < 0) || (r < 2) || ((r & (r - 1)) != 0) ||
        (I == NULL))
        return -1;
    if (n <= 1) {
        if (freq != NULL) memset(freq, 0, ALPHABET16 * sizeof(int64_t));
        if (n == 1) { U[0] = T[0]; if (freq != NULL) freq[T[0]]++; }
        I[0] = n;
        return 0;
    }
    return map_rc(sagpu_host_bwt16(T, U, n, freq, r, I, 8));
}

/* reference src/libsais16x64.c:7789-7812.  The freq table is not needed: counts come from T. */
int64_t libsais16x64_unbwt_aux(const uint16_t *T, uint16_t *U, int64_t *A, int64_t n, const int64_t *freq, int64_t r, const int64_t *I)
{
    (void)freq;
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || ((r != n) && ((r < 2) || ((r & (r - 1)) != 0))) ||
        (I == NULL))
        return -1;
    if (n <= 1) {
        if (I[0] != n) return -1;
        if (n == 1) U[0] = T[0];
        return 0;
    }
    for (int64_t t = 0; t <= (int64_t)(n - 1) / r; ++t)
        if (I[t] <= 0 || I[t] > n) return -1;
    return map_rc(sagpu_host_unbwt16(T, U, n, r, I, 8));
}

/* reference src/libsais16x64.c:7784-7787 */
int64_t libsais16x64_unbwt(const uint16_t *T, uint16_t *U, int64_t *A, int64_t n, const int64_t *freq, int64_t i)
{
    return libsais16x64_unbwt_aux(T, U, A, n, freq, n, &i);
}

/* reference src/libsais16x64.c:8050-8066 */
int64_t libsais16x64_plcp(const uint16_t *T, const int64_t *SA, int64_t *PLCP, int64_t n)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0)) return -1;
    if (n <= 1) {
        if (n == 1) PLCP[0] = 0;
        return 0;
    }
    return map_rc(sagpu_host_plcp16(T, SA, PLCP, n, 0, 8));
}

/* reference src/libsais16x64.c:8068-8084 */
int64_t libsais16x64_plcp_gsa(const uint16_t *T, const int64_t *SA, int64_t *PLCP, int64_t n)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0) || (n > 0 && T[n - 1] != 0)) return -1;
    if (n <= 1) {
        if (n == 1) PLCP[0] = 0;
        return 0;
    }
    return map_rc(sagpu_host_plcp16(T, SA, PLCP, n, 1, 8));
}

/* reference src/libsais16x64.c:8086-8101 */
int64_t libsais16x64_lcp(const int64_t *PLCP, const int64_t *SA, int64_t *LCP, int64_t n)
{
    if ((PLCP == NULL) || (SA == NULL) || (LCP == NULL) || (n < 0)) return -1;
    if (n <= 1) {
        if (n == 1) LCP[0] = PLCP[SA[0]];
        return 0;
    }
    return map_rc(sagpu_host_lcp(PLCP, SA, LCP, n, 8));
}

/* ------------------------------------------------- libsais16x64 OpenMP -- */

/* reference src/libsais16x64.c:7191-7222 */
int64_t libsais16x64_omp(const uint16_t *T, int64_t *SA, int64_t n, int64_t fs, int64_t *freq, int64_t threads)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (fs < 0) || (threads < 0)) return -1;
    return libsais16x64(T, SA, n, fs, freq);
}

/* reference src/libsais16x64.c:7224-7255 */
int64_t libsais16x64_gsa_omp(const uint16_t *T, int64_t *SA, int64_t n, int64_t fs, int64_t *freq, int64_t threads)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (n > 0 && T[n - 1] != 0) || (fs < 0) || (threads < 0)) return -1;
    return libsais16x64_gsa(T, SA, n, fs, freq);
}

/* reference src/libsais16x64.c:7275-7315 */
int64_t libsais16x64_bwt_omp(const uint16_t *T, uint16_t *U, int64_t *A, int64_t n, int64_t fs, int64_t *freq, int64_t threads)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || (fs < 0) || (threads < 0)) return -1;
    return libsais16x64_bwt(T, U, A, n, fs, freq);
}

/* reference src/libsais16x64.c:7317-7357 */
int64_t libsais16x64_bwt_aux_omp(const uint16_t *T, uint16_t *U, int64_t *A, int64_t n, int64_t fs, int64_t *freq, int64_t r, int64_t *I,
                      int64_t threads)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || (fs < 0) || (r < 2) || ((r & (r - 1)) != 0) ||
        (I == NULL) || (threads < 0))
        return -1;
    return libsais16x64_bwt_aux(T, U, A, n, fs, freq, r, I);
}

/* reference src/libsais16x64.c:7821-7847 */
int64_t libsais16x64_unbwt_aux_omp(const uint16_t *T, uint16_t *U, int64_t *A, int64_t n, const int64_t *freq, int64_t r, const int64_t *I,
                        int64_t threads)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || ((r != n) && ((r < 2) || ((r & (r - 1)) != 0))) ||
        (I == NULL) || (threads < 0))
        return -1;
    return libsais16x64_unbwt_aux(T, U, A, n, freq, r, I);
}

/* reference src/libsais16x64.c:7816-7819 */
int64_t libsais16x64_unbwt_omp(const uint16_t *T, uint16_t *U, int64_t *A, int64_t n, const int64_t *freq, int64_t i, int64_t threads)
{
    return libsais16x64_unbwt_aux_omp(T, U, A, n, freq, n, &i, threads);
}

/* reference src/libsais16x64.c:8105-8124 */
int64_t libsais16x64_plcp_omp(const uint16_t *T, const int64_t *SA, int64_t *PLCP, int64_t n, int64_t threads)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0) || (threads < 0)) return -1;
    return libsais16x64_plcp(T, SA, PLCP, n);
}

/* reference src/libsais16x64.c:8126-8145 */
int64_t libsais16x64_plcp_gsa_omp(const uint16_t *T, const int64_t *SA, int64_t *PLCP, int64_t n, int64_t threads)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0) || (n > 0 && T[n - 1] != 0) || (threads < 0))
        return -1;
    return libsais16x64_plcp_gsa(T, SA, PLCP, n);
}

/* reference src/libsais16x64.c:8147-8165 */
int64_t libsais16x64_lcp_omp(const int64_t *PLCP, const int64_t *SA, int64_t *LCP, int64_t n, int64_t threads)
{
    if ((PLCP == NULL) || (SA == NULL) || (LCP == NULL) || (n < 0) || (threads < 0)) return -1;
    return libsais16x64_lcp(PLCP, SA, LCP, n);
}

/* reference src/libsais16x64.c:7257-7273 */
int64_t libsais16x64_long_omp(int64_t *T, int64_t *SA, int64_t n, int64_t k, int64_t fs, int64_t threads)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (fs < 0) || (threads < 0)) return -1;
    return libsais16x64_long(T, SA, n, k, fs);
}
