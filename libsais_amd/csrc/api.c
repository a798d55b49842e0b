/*
 * api.c -- libsais 2.10.4 compatible entry points (C99 host code).
 *
 * Each function repeats the reference's argument checks and n <= 1 special cases in
 * the same order, then hands the work to the GPU engine through the internal
 * host-buffer layer (sagpu_host_*, engine.hip).  Return codes: 0 / primary index on
 * success, -1 invalid arguments, -2 when device memory cannot be obtained or the
 * device fails (the reference's allocation-failure code).
 */
#include "../../include/libsais.h"
#include "../../include/libsais64.h"

#include <stddef.h>
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
