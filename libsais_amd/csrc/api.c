/*
 * api.c -- libsais 2.10.4 compatible entry points (C99 host code).
 *
 * Each function repeats the reference's argument checks and n <= 1 special cases in
 * the same order, then hands the work to the GPU engine through the internal
 * host-buffer layer (sagpu_host_*, engine.hip).  Return codes: 0 / primary index on
 * success, -1 invalid arguments, -2 when device memory cannot be obtained or the
 * device fails (the reference's allocation-failure code).
 */
#define LIBSAIS_OPENMP 1 /* declare the _omp forms, which this library defines */
#include "../../include/libsais.h"
#include "../../include/libsais64.h"
#include "api_ctx.h"

#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define HIDDEN __attribute__((visibility("hidden")))

HIDDEN int64_t sagpu_host_sa(const uint8_t *T, void *SA, int64_t n, int32_t gsa, void *freq, int32_t width);
HIDDEN int64_t sagpu_host_bwt(const uint8_t *T, uint8_t *U, int64_t n, void *freq, int32_t width);
HIDDEN int64_t sagpu_host_plcp(const uint8_t *T, const void *SA, void *PLCP, int64_t n, int32_t gsa, int32_t width);
HIDDEN int64_t sagpu_host_lcp(const void *PLCP, const void *SA, void *LCP, int64_t n, int32_t width);
HIDDEN int64_t sagpu_host_bwt_aux(const uint8_t *T, uint8_t *U, int64_t n, void *freq, int64_t r, void *I,
                                  int32_t width);
HIDDEN int64_t sagpu_host_unbwt(const uint8_t *T, uint8_t *U, int64_t n, int64_t r, const void *I, int32_t width);

/* Engine codes below -2 (device failures) surface as -2, the only failure code the
 * reference defines besides -1. */
static int64_t map_rc(int64_t rc) { return rc < -2 ? -2 : rc; }

/* ------------------------------------------------------------ 32-bit API -- */

/* reference src/libsais.c:7018-7032 */
int32_t libsais(const uint8_t *T, int32_t *SA, int32_t n, int32_t fs, int32_t *freq)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (fs < 0)) return -1;
    if (n <= 1) {
        if (freq != NULL) memset(freq, 0, 256 * sizeof(int32_t));
        if (n == 1) { SA[0] = 0; if (freq != NULL) freq[T[0]]++; }
        return 0;
    }
    return (int32_t)map_rc(sagpu_host_sa(T, SA, n, 0, freq, 4));
}

/* reference src/libsais.c:7034-7048 */
int32_t libsais_gsa(const uint8_t *T, int32_t *SA, int32_t n, int32_t fs, int32_t *freq)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (n > 0 && T[n - 1] != 0) || (fs < 0)) return -1;
    if (n <= 1) {
        if (freq != NULL) memset(freq, 0, 256 * sizeof(int32_t));
        if (n == 1) { SA[0] = 0; if (freq != NULL) freq[T[0]]++; }
        return 0;
    }
    return (int32_t)map_rc(sagpu_host_sa(T, SA, n, 1, freq, 4));
}

/* reference src/libsais.c:7097-7121 */
int32_t libsais_bwt(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t fs, int32_t *freq)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || (fs < 0)) return -1;
    if (n <= 1) {
        if (freq != NULL) memset(freq, 0, 256 * sizeof(int32_t));
        if (n == 1) { U[0] = T[0]; if (freq != NULL) freq[T[0]]++; }
        return n;
    }
    return (int32_t)map_rc(sagpu_host_bwt(T, U, n, freq, 4));
}

/* reference src/libsais.c:7123-7147 */
int32_t libsais_bwt_aux(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t fs, int32_t *freq, int32_t r,
                        int32_t *I)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || (fs < 0) || (r < 2) || ((r & (r - 1)) != 0) ||
        (I == NULL))
        return -1;
    if (n <= 1) {
        if (freq != NULL) memset(freq, 0, 256 * sizeof(int32_t));
        if (n == 1) { U[0] = T[0]; if (freq != NULL) freq[T[0]]++; }
        I[0] = n;
        return 0;
    }
    return (int32_t)map_rc(sagpu_host_bwt_aux(T, U, n, freq, r, I, 4));
}

/* reference src/libsais.c:8040-8056 (the freq table is not needed: counts come from T) */
int32_t libsais_unbwt_aux(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, const int32_t *freq, int32_t r,
                          const int32_t *I)
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
    return (int32_t)map_rc(sagpu_host_unbwt(T, U, n, r, I, 4));
}

/* reference src/libsais.c:8030-8033 */
int32_t libsais_unbwt(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, const int32_t *freq, int32_t i)
{
    return libsais_unbwt_aux(T, U, A, n, freq, n, &i);
}

/* reference src/libsais.c:8363-8379 */
int32_t libsais_plcp(const uint8_t *T, const int32_t *SA, int32_t *PLCP, int32_t n)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0)) return -1;
    if (n <= 1) {
        if (n == 1) PLCP[0] = 0;
        return 0;
    }
    return (int32_t)map_rc(sagpu_host_plcp(T, SA, PLCP, n, 0, 4));
}

/* reference src/libsais.c:8381-8397 */
int32_t libsais_plcp_gsa(const uint8_t *T, const int32_t *SA, int32_t *PLCP, int32_t n)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0) || (n > 0 && T[n - 1] != 0)) return -1;
    if (n <= 1) {
        if (n == 1) PLCP[0] = 0;
        return 0;
    }
    return (int32_t)map_rc(sagpu_host_plcp(T, SA, PLCP, n, 1, 4));
}

/* reference src/libsais.c:8417-8432 */
int32_t libsais_lcp(const int32_t *PLCP, const int32_t *SA, int32_t *LCP, int32_t n)
{
    if ((PLCP == NULL) || (SA == NULL) || (LCP == NULL) || (n < 0)) return -1;
    if (n <= 1) {
        if (n == 1) LCP[0] = PLCP[SA[0]];
        return 0;
    }
    return (int32_t)map_rc(sagpu_host_lcp(PLCP, SA, LCP, n, 4));
}

/* ------------------------------------------------------------ 64-bit API -- */

/* reference src/libsais64.c:7058-7086 */
int64_t libsais64(const uint8_t *T, int64_t *SA, int64_t n, int64_t fs, int64_t *freq)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (fs < 0)) return -1;
    if (n <= 1) {
        if (freq != NULL) memset(freq, 0, 256 * sizeof(int64_t));
        if (n == 1) { SA[0] = 0; if (freq != NULL) freq[T[0]]++; }
        return 0;
    }
    return map_rc(sagpu_host_sa(T, SA, n, 0, freq, 8));
}

/* reference src/libsais64.c:7088-7116 */
int64_t libsais64_gsa(const uint8_t *T, int64_t *SA, int64_t n, int64_t fs, int64_t *freq)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (n > 0 && T[n - 1] != 0) || (fs < 0)) return -1;
    if (n <= 1) {
        if (freq != NULL) memset(freq, 0, 256 * sizeof(int64_t));
        if (n == 1) { SA[0] = 0; if (freq != NULL) freq[T[0]]++; }
        return 0;
    }
    return map_rc(sagpu_host_sa(T, SA, n, 1, freq, 8));
}

/* reference src/libsais64.c:7133-7170 */
int64_t libsais64_bwt(const uint8_t *T, uint8_t *U, int64_t *A, int64_t n, int64_t fs, int64_t *freq)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || (fs < 0)) return -1;
    if (n <= 1) {
        if (freq != NULL) memset(freq, 0, 256 * sizeof(int64_t));
        if (n == 1) { U[0] = T[0]; if (freq != NULL) freq[T[0]]++; }
        return n;
    }
    return map_rc(sagpu_host_bwt(T, U, n, freq, 8));
}

/* reference src/libsais64.c:7172-7209 */
int64_t libsais64_bwt_aux(const uint8_t *T, uint8_t *U, int64_t *A, int64_t n, int64_t fs, int64_t *freq, int64_t r,
                          int64_t *I)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || (fs < 0) || (r < 2) || ((r & (r - 1)) != 0) ||
        (I == NULL))
        return -1;
    if (n <= 1) {
        if (freq != NULL) memset(freq, 0, 256 * sizeof(int64_t));
        if (n == 1) { U[0] = T[0]; if (freq != NULL) freq[T[0]]++; }
        I[0] = n;
        return 0;
    }
    return map_rc(sagpu_host_bwt_aux(T, U, n, freq, r, I, 8));
}

/* reference src/libsais64.c:8039-8064 */
int64_t libsais64_unbwt_aux(const uint8_t *T, uint8_t *U, int64_t *A, int64_t n, const int64_t *freq, int64_t r,
                            const int64_t *I)
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
    for (int64_t t = 0; t <= (n - 1) / r; ++t)
        if (I[t] <= 0 || I[t] > n) return -1;
    return map_rc(sagpu_host_unbwt(T, U, n, r, I, 8));
}

/* reference src/libsais64.c:8034-8037 */
int64_t libsais64_unbwt(const uint8_t *T, uint8_t *U, int64_t *A, int64_t n, const int64_t *freq, int64_t i)
{
    return libsais64_unbwt_aux(T, U, A, n, freq, n, &i);
}

/* reference src/libsais64.c:8302-8318 */
int64_t libsais64_plcp(const uint8_t *T, const int64_t *SA, int64_t *PLCP, int64_t n)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0)) return -1;
    if (n <= 1) {
        if (n == 1) PLCP[0] = 0;
        return 0;
    }
    return map_rc(sagpu_host_plcp(T, SA, PLCP, n, 0, 8));
}

/* reference src/libsais64.c:8320-8336 */
int64_t libsais64_plcp_gsa(const uint8_t *T, const int64_t *SA, int64_t *PLCP, int64_t n)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0) || (n > 0 && T[n - 1] != 0)) return -1;
    if (n <= 1) {
        if (n == 1) PLCP[0] = 0;
        return 0;
    }
    return map_rc(sagpu_host_plcp(T, SA, PLCP, n, 1, 8));
}

/* reference src/libsais64.c:8338-8353 */
int64_t libsais64_lcp(const int64_t *PLCP, const int64_t *SA, int64_t *LCP, int64_t n)
{
    if ((PLCP == NULL) || (SA == NULL) || (LCP == NULL) || (n < 0)) return -1;
    if (n <= 1) {
        if (n == 1) LCP[0] = PLCP[SA[0]];
        return 0;
    }
    return map_rc(sagpu_host_lcp(PLCP, SA, LCP, n, 8));
}

/* --------------------------------------------- integer alphabet (32 / 64) -- */

HIDDEN int64_t sagpu_host_sa_int(const void *T, void *SA, int64_t n, int32_t width);
HIDDEN int64_t sagpu_host_plcp_int(const int32_t *T, const void *SA, void *PLCP, int64_t n, int32_t width);

/* reference src/libsais.c:7050-7063.  T is not modified (the reference restores it). */
int32_t libsais_int(int32_t *T, int32_t *SA, int32_t n, int32_t k, int32_t fs)
{
    (void)k;
    if ((T == NULL) || (SA == NULL) || (n < 0) || (fs < 0)) return -1;
    if (n <= 1) {
        if (n == 1) SA[0] = 0;
        return 0;
    }
    return (int32_t)map_rc(sagpu_host_sa_int(T, SA, n, 4));
}

/* reference src/libsais64.c:7118-7131 */
int64_t libsais64_long(int64_t *T, int64_t *SA, int64_t n, int64_t k, int64_t fs)
{
    (void)k;
    if ((T == NULL) || (SA == NULL) || (n < 0) || (fs < 0)) return -1;
    if (n <= 1) {
        if (n == 1) SA[0] = 0;
        return 0;
    }
    return map_rc(sagpu_host_sa_int(T, SA, n, 8));
}

/* reference src/libsais.c:8399-8415 */
int32_t libsais_plcp_int(const int32_t *T, const int32_t *SA, int32_t *PLCP, int32_t n)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0)) return -1;
    if (n <= 1) {
        if (n == 1) PLCP[0] = 0;
        return 0;
    }
    return (int32_t)map_rc(sagpu_host_plcp_int(T, SA, PLCP, n, 4));
}

/* ------------------------------------------------------------- contexts -- */
/* The reference's contexts hold reusable buffers, one context per thread
 * (src/libsais.c:7008-7016, 8020-8028; include/libsais.h:54).  Here a context owns
 * an engine context of its own: a device stream and a pool of device buffers,
 * created on the context's first call.  Calls through two contexts run concurrently
 * on the device; calls without a context share the engine's slot contexts. */

static void *ctx_new(void) { return sagpu_host_token_new(); }

static void ctx_free(void *ctx) { sagpu_host_token_free(ctx); }

/* reference src/libsais.c:7008-7011 */
void *libsais_create_ctx(void) { return ctx_new(); }

/* reference src/libsais.c:7201-7209 */
void *libsais_create_ctx_omp(int32_t threads) { return threads < 0 ? NULL : ctx_new(); }

/* reference src/libsais.c:7013-7016 */
void libsais_free_ctx(void *ctx) { ctx_free(ctx); }

/* reference src/libsais.c:8020-8023 */
void *libsais_unbwt_create_ctx(void) { return ctx_new(); }

/* reference src/libsais.c:8078-8086 */
void *libsais_unbwt_create_ctx_omp(int32_t threads) { return threads < 0 ? NULL : ctx_new(); }

/* reference src/libsais.c:8025-8028 */
void libsais_unbwt_free_ctx(void *ctx) { ctx_free(ctx); }

/* reference src/libsais.c:7065-7079 */
int32_t libsais_ctx(const void *ctx, const uint8_t *T, int32_t *SA, int32_t n, int32_t fs, int32_t *freq)
{
    if (ctx == NULL) return -1;
    WITH_CTX(ctx, int32_t, libsais(T, SA, n, fs, freq));
}

/* reference src/libsais.c:7081-7095 */
int32_t libsais_gsa_ctx(const void *ctx, const uint8_t *T, int32_t *SA, int32_t n, int32_t fs, int32_t *freq)
{
    if (ctx == NULL) return -1;
    WITH_CTX(ctx, int32_t, libsais_gsa(T, SA, n, fs, freq));
}

/* reference src/libsais.c:7148-7172 */
int32_t libsais_bwt_ctx(const void *ctx, const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t fs,
                        int32_t *freq)
{
    if (ctx == NULL) return -1;
    WITH_CTX(ctx, int32_t, libsais_bwt(T, U, A, n, fs, freq));
}

/* reference src/libsais.c:7174-7197 */
int32_t libsais_bwt_aux_ctx(const void *ctx, const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t fs,
                            int32_t *freq, int32_t r, int32_t *I)
{
    if (ctx == NULL) return -1;
    WITH_CTX(ctx, int32_t, libsais_bwt_aux(T, U, A, n, fs, freq, r, I));
}

/* reference src/libsais.c:8035-8038 */
int32_t libsais_unbwt_ctx(const void *ctx, const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, const int32_t *freq,
                          int32_t i)
{
    return libsais_unbwt_aux_ctx(ctx, T, U, A, n, freq, n, &i);
}

/* reference src/libsais.c:8058-8074 */
int32_t libsais_unbwt_aux_ctx(const void *ctx, const uint8_t *T, uint8_t *U, int32_t *A, int32_t n,
                              const int32_t *freq, int32_t r, const int32_t *I)
{
    if (ctx == NULL) return -1;
    WITH_CTX(ctx, int32_t, libsais_unbwt_aux(T, U, A, n, freq, r, I));
}

/* ------------------------------------------------------- OpenMP entries -- */
/* The reference's _omp functions take a thread count (0 = OpenMP default, < 0
 * rejected); the work here runs on the GPU, so the count is only checked. */

/* reference src/libsais.c:7211-7228 */
int32_t libsais_omp(const uint8_t *T, int32_t *SA, int32_t n, int32_t fs, int32_t *freq, int32_t threads)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (fs < 0) || (threads < 0)) return -1;
    return libsais(T, SA, n, fs, freq);
}

/* reference src/libsais.c:7230-7247 */
int32_t libsais_gsa_omp(const uint8_t *T, int32_t *SA, int32_t n, int32_t fs, int32_t *freq, int32_t threads)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (n > 0 && T[n - 1] != 0) || (fs < 0) || (threads < 0)) return -1;
    return libsais_gsa(T, SA, n, fs, freq);
}

/* reference src/libsais.c:7249-7265 */
int32_t libsais_int_omp(int32_t *T, int32_t *SA, int32_t n, int32_t k, int32_t fs, int32_t threads)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (fs < 0) || (threads < 0)) return -1;
    return libsais_int(T, SA, n, k, fs);
}

/* reference src/libsais.c:7267-7294 */
int32_t libsais_bwt_omp(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t fs, int32_t *freq,
                        int32_t threads)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || (fs < 0) || (threads < 0)) return -1;
    return libsais_bwt(T, U, A, n, fs, freq);
}

/* reference src/libsais.c:7296-7322 */
int32_t libsais_bwt_aux_omp(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t fs, int32_t *freq,
                            int32_t r, int32_t *I, int32_t threads)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || (fs < 0) || (r < 2) || ((r & (r - 1)) != 0) ||
        (I == NULL) || (threads < 0))
        return -1;
    return libsais_bwt_aux(T, U, A, n, fs, freq, r, I);
}

/* reference src/libsais.c:8088-8091 */
int32_t libsais_unbwt_omp(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, const int32_t *freq, int32_t i,
                          int32_t threads)
{
    return libsais_unbwt_aux_omp(T, U, A, n, freq, n, &i, threads);
}

/* reference src/libsais.c:8093-8112 */
int32_t libsais_unbwt_aux_omp(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, const int32_t *freq, int32_t r,
                              const int32_t *I, int32_t threads)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || ((r != n) && ((r < 2) || ((r & (r - 1)) != 0))) ||
        (I == NULL) || (threads < 0))
        return -1;
    return libsais_unbwt_aux(T, U, A, n, freq, r, I);
}

/* reference src/libsais.c:8436-8455 */
int32_t libsais_plcp_omp(const uint8_t *T, const int32_t *SA, int32_t *PLCP, int32_t n, int32_t threads)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0) || (threads < 0)) return -1;
    return libsais_plcp(T, SA, PLCP, n);
}

/* reference src/libsais.c:8457-8476 */
int32_t libsais_plcp_gsa_omp(const uint8_t *T, const int32_t *SA, int32_t *PLCP, int32_t n, int32_t threads)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0) || (n > 0 && T[n - 1] != 0) || (threads < 0))
        return -1;
    return libsais_plcp_gsa(T, SA, PLCP, n);
}

/* reference src/libsais.c:8478-8497 */
int32_t libsais_plcp_int_omp(const int32_t *T, const int32_t *SA, int32_t *PLCP, int32_t n, int32_t threads)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0) || (threads < 0)) return -1;
    return libsais_plcp_int(T, SA, PLCP, n);
}

/* reference src/libsais.c:8499-8517 */
int32_t libsais_lcp_omp(const int32_t *PLCP, const int32_t *SA, int32_t *LCP, int32_t n, int32_t threads)
{
    if ((PLCP == NULL) || (SA == NULL) || (LCP == NULL) || (n < 0) || (threads < 0)) return -1;
    return libsais_lcp(PLCP, SA, LCP, n);
}

/* reference src/libsais64.c:7213-7244 */
int64_t libsais64_omp(const uint8_t *T, int64_t *SA, int64_t n, int64_t fs, int64_t *freq, int64_t threads)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (fs < 0) || (threads < 0)) return -1;
    return libsais64(T, SA, n, fs, freq);
}

/* reference src/libsais64.c:7246-7277 */
int64_t libsais64_gsa_omp(const uint8_t *T, int64_t *SA, int64_t n, int64_t fs, int64_t *freq, int64_t threads)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (n > 0 && T[n - 1] != 0) || (fs < 0) || (threads < 0)) return -1;
    return libsais64_gsa(T, SA, n, fs, freq);
}

/* reference src/libsais64.c:7279-7295 */
int64_t libsais64_long_omp(int64_t *T, int64_t *SA, int64_t n, int64_t k, int64_t fs, int64_t threads)
{
    if ((T == NULL) || (SA == NULL) || (n < 0) || (fs < 0) || (threads < 0)) return -1;
    return libsais64_long(T, SA, n, k, fs);
}

/* reference src/libsais64.c:7297-7337 */
int64_t libsais64_bwt_omp(const uint8_t *T, uint8_t *U, int64_t *A, int64_t n, int64_t fs, int64_t *freq,
                          int64_t threads)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || (fs < 0) || (threads < 0)) return -1;
    return libsais64_bwt(T, U, A, n, fs, freq);
}

/* reference src/libsais64.c:7339-7379 */
int64_t libsais64_bwt_aux_omp(const uint8_t *T, uint8_t *U, int64_t *A, int64_t n, int64_t fs, int64_t *freq,
                              int64_t r, int64_t *I, int64_t threads)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || (fs < 0) || (r < 2) || ((r & (r - 1)) != 0) ||
        (I == NULL) || (threads < 0))
        return -1;
    return libsais64_bwt_aux(T, U, A, n, fs, freq, r, I);
}

/* reference src/libsais64.c:8067-8070 */
int64_t libsais64_unbwt_omp(const uint8_t *T, uint8_t *U, int64_t *A, int64_t n, const int64_t *freq, int64_t i,
                            int64_t threads)
{
    return libsais64_unbwt_aux_omp(T, U, A, n, freq, n, &i, threads);
}

/* reference src/libsais64.c:8072-8099 */
int64_t libsais64_unbwt_aux_omp(const uint8_t *T, uint8_t *U, int64_t *A, int64_t n, const int64_t *freq, int64_t r,
                                const int64_t *I, int64_t threads)
{
    if ((T == NULL) || (U == NULL) || (A == NULL) || (n < 0) || ((r != n) && ((r < 2) || ((r & (r - 1)) != 0))) ||
        (I == NULL) || (threads < 0))
        return -1;
    return libsais64_unbwt_aux(T, U, A, n, freq, r, I);
}

/* reference src/libsais64.c:8357-8376 */
int64_t libsais64_plcp_omp(const uint8_t *T, const int64_t *SA, int64_t *PLCP, int64_t n, int64_t threads)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0) || (threads < 0)) return -1;
    return libsais64_plcp(T, SA, PLCP, n);
}

/* reference src/libsais64.c:8378-8397 */
int64_t libsais64_plcp_gsa_omp(const uint8_t *T, const int64_t *SA, int64_t *PLCP, int64_t n, int64_t threads)
{
    if ((T == NULL) || (SA == NULL) || (PLCP == NULL) || (n < 0) || (n > 0 && T[n - 1] != 0) || (threads < 0))
        return -1;
    return libsais64_plcp_gsa(T, SA, PLCP, n);
}

/* reference src/libsais64.c:8399-8417 */
int64_t libsais64_lcp_omp(const int64_t *PLCP, const int64_t *SA, int64_t *LCP, int64_t n, int64_t threads)
{
    if ((PLCP == NULL) || (SA == NULL) || (LCP == NULL) || (n < 0) || (threads < 0)) return -1;
    return libsais64_lcp(PLCP, SA, LCP, n);
}
