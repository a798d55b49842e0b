/*
 * caller.c -- a C99 program that uses the library the way an existing libsais user
 * does: it includes include/libsais.h and include/libsais64.h unchanged and links
 * -lsais.  Test infrastructure (tests/test_c_caller.py runs it).
 *
 *   caller cpu   argument checks and n <= 1 cases, which never reach the device
 *                (reference src/libsais.c:7018-7121, 8363-8432; src/libsais64.c:7058-7170)
 *   caller gpu   SA, BWT, PLCP and LCP of seeded texts, compared with the oracle
 *                (oracle/oracle.h, linked as the checker)
 *
 * Prints one line per failed check and "ok <n checks>" at the end; exit status 0 when
 * every check held.
 */
#include "libsais.h"
#include "libsais64.h"
#include "oracle.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int checks, failures;

#define CHECK(cond)                                                          \
    do {                                                                     \
        checks++;                                                            \
        if (!(cond)) {                                                       \
            failures++;                                                      \
            printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);         \
        }                                                                    \
    } while (0)

static void cpu_checks(void)
{
    uint8_t T[4] = {'b', 'a', 0, 0}, U[4];
    int32_t SA[4], A[4], P[4], L[4], freq[256];
    int64_t SA64[4], A64[4], freq64[256];

    CHECK(strcmp(LIBSAIS_VERSION_STRING, "2.10.4") == 0);
    CHECK(LIBSAIS_VERSION_MAJOR == 2 && LIBSAIS_VERSION_MINOR == 10 && LIBSAIS_VERSION_PATCH == 4);

    /* -1 for invalid arguments */
    CHECK(libsais(NULL, SA, 2, 0, NULL) == -1);
    CHECK(libsais(T, NULL, 2, 0, NULL) == -1);
    CHECK(libsais(T, SA, -1, 0, NULL) == -1);
    CHECK(libsais(T, SA, 2, -1, NULL) == -1);
    CHECK(libsais_gsa(T, SA, 2, 0, NULL) == -1);            /* T[n-1] != 0 */
    CHECK(libsais_bwt(T, U, NULL, 2, 0, NULL) == -1);
    CHECK(libsais_bwt(T, NULL, A, 2, 0, NULL) == -1);
    CHECK(libsais_plcp(T, NULL, P, 2) == -1);
    CHECK(libsais_plcp_gsa(T, SA, P, 2) == -1);             /* T[n-1] != 0 */
    CHECK(libsais_lcp(P, SA, NULL, 2) == -1);
    CHECK(libsais64(T, SA64, -5, 0, NULL) == -1);
    CHECK(libsais64_bwt(T, U, A64, 2, -1, NULL) == -1);
    CHECK(libsais_ctx(NULL, T, SA, 2, 0, NULL) == -1);
    CHECK(libsais_omp(T, SA, 2, 0, NULL, -1) == -1);

    /* n == 0: 0 and freq cleared */
    memset(freq, 0x55, sizeof freq);
    CHECK(libsais(T, SA, 0, 0, freq) == 0);
    CHECK(freq[0] == 0 && freq['a'] == 0 && freq[255] == 0);
    CHECK(libsais_bwt(T, U, A, 0, 0, NULL) == 0);

    /* n == 1 */
    memset(freq, 0x55, sizeof freq);
    SA[0] = 7;
    CHECK(libsais(T, SA, 1, 0, freq) == 0 && SA[0] == 0 && freq['b'] == 1 && freq['a'] == 0);
    U[0] = 0;
    CHECK(libsais_bwt(T, U, A, 1, 0, NULL) == 1 && U[0] == 'b');
    P[0] = 9;
    CHECK(libsais_plcp(T, SA, P, 1) == 0 && P[0] == 0);
    L[0] = 9;
    CHECK(libsais_lcp(P, SA, L, 1) == 0 && L[0] == 0);
    memset(freq64, 0x55, sizeof freq64);
    CHECK(libsais64(T, SA64, 1, 0, freq64) == 0 && SA64[0] == 0 && freq64['b'] == 1 && freq64[0] == 0);

    void *ctx = libsais_create_ctx();
    CHECK(ctx != NULL);
    CHECK(libsais_ctx(ctx, T, SA, 1, 0, NULL) == 0 && SA[0] == 0);
    libsais_free_ctx(ctx);
    libsais_free_ctx(NULL);
    CHECK(libsais_create_ctx_omp(-1) == NULL);
}

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;

static uint64_t next_u64(void)
{
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void gpu_case(int32_t n, int alphabet, int period)
{
    uint8_t *T = malloc((size_t)n), *U = malloc((size_t)n), *Uo = malloc((size_t)n);
    int32_t *SA = malloc(4 * (size_t)n), *SAo = malloc(4 * (size_t)n), *A = malloc(4 * (size_t)n);
    int32_t *P = malloc(4 * (size_t)n), *Po = malloc(4 * (size_t)n), *L = malloc(4 * (size_t)n);
    int32_t *Lo = malloc(4 * (size_t)n), freq[256];
    int64_t *SA64 = malloc(8 * (size_t)n);
    static const uint8_t acgt[4] = {'A', 'C', 'G', 'T'};
    for (int32_t i = 0; i < n; i++)
        T[i] = period ? (i >= period ? T[i - period] : acgt[next_u64() % 4])
                      : (alphabet == 4 ? acgt[next_u64() % 4] : (uint8_t)(next_u64() % (uint64_t)alphabet));
    if (period) T[n / 2] ^= 1;

    CHECK(oracle_sa32(T, SAo, n, 0) == 0);
    CHECK(libsais(T, SA, n, 0, freq) == 0);
    CHECK(memcmp(SA, SAo, 4 * (size_t)n) == 0);
    int64_t f[256];
    oracle_freq(T, n, f);
    int fok = 1;
    for (int c = 0; c < 256; c++) fok &= freq[c] == (int32_t)f[c];
    CHECK(fok);

    int64_t po = oracle_bwt32(T, SAo, Uo, n);
    CHECK(libsais_bwt(T, U, A, n, 0, NULL) == (int32_t)po);
    CHECK(memcmp(U, Uo, (size_t)n) == 0);

    oracle_plcp32(T, SAo, Po, n, 0);
    CHECK(libsais_plcp(T, SA, P, n) == 0);
    CHECK(memcmp(P, Po, 4 * (size_t)n) == 0);
    oracle_lcp32(Po, SAo, Lo, n);
    CHECK(libsais_lcp(P, SA, L, n) == 0);
    CHECK(memcmp(L, Lo, 4 * (size_t)n) == 0);

    CHECK(libsais64(T, SA64, n, 0, NULL) == 0);
    int same = 1;
    for (int32_t i = 0; i < n; i++) same &= SA64[i] == (int64_t)SAo[i];
    CHECK(same);

    /* U may alias T (include/libsais.h:177) */
    memcpy(U, T, (size_t)n);
    CHECK(libsais_bwt(U, U, A, n, 0, NULL) == (int32_t)po);
    CHECK(memcmp(U, Uo, (size_t)n) == 0);

    free(T); free(U); free(Uo); free(SA); free(SAo); free(A); free(P); free(Po); free(L); free(Lo); free(SA64);
}

int main(int argc, char **argv)
{
    if (argc < 2 || (strcmp(argv[1], "cpu") && strcmp(argv[1], "gpu"))) {
        fprintf(stderr, "usage: %s cpu|gpu\n", argv[0]);
        return 2;
    }
    if (!strcmp(argv[1], "cpu")) {
        cpu_checks();
    } else {
        gpu_case(100000, 4, 0);
        gpu_case(65537, 256, 0);
        gpu_case(200003, 4, 23);
    }
    printf("%s %d checks, %d failed\n", failures ? "FAILED" : "ok", checks, failures);
    return failures ? 1 : 0;
}
