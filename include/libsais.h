/*
 * libsais.h -- drop-in header of the MI355X suffix-array engine.
 *
 * Same prototypes, version macros and return-code conventions as libsais 2.10.4
 * (include/libsais.h of the reference), for the entry points this engine provides:
 * suffix array (byte and integer alphabets), generalized suffix array, BWT (plain
 * and with auxiliary indexes), inverse BWT, PLCP (byte, GSA and integer) and LCP,
 * with the reference's context (_ctx) and OpenMP (_omp, declared when LIBSAIS_OPENMP
 * is defined, as in the reference) forms.  Contexts are tokens here (the device
 * buffers are pooled inside the engine) and the OpenMP thread count is checked but
 * the work runs on the GPU.
 *
 * All buffers are caller-owned host memory; calls are synchronous and may be made
 * from several host threads at once.  Work runs on GPU 0.
 */
#ifndef LIBSAIS_H
#define LIBSAIS_H 1

#define LIBSAIS_VERSION_MAJOR   2
#define LIBSAIS_VERSION_MINOR   10
#define LIBSAIS_VERSION_PATCH   4
#define LIBSAIS_VERSION_STRING  "2.10.4"

#ifdef _WIN32
    #ifdef LIBSAIS_SHARED
        #ifdef LIBSAIS_EXPORTS
            #define LIBSAIS_API __declspec(dllexport)
        #else
            #define LIBSAIS_API __declspec(dllimport)
        #endif
    #else
        #define LIBSAIS_API
    #endif
#else
    #define LIBSAIS_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

#include <stdint.h>

/*
 * Suffix array of T[0..n-1] into SA[0..n-1] (SA has room for n+fs entries; only the
 * first n are defined).  freq (256 entries, may be NULL) receives symbol counts.
 * Returns 0, -1 for invalid arguments, -2 when memory cannot be obtained.
 * Replaces reference include/libsais.h:84 (src/libsais.c:7018-7032).
 */
LIBSAIS_API int32_t libsais(const uint8_t * T, int32_t * SA, int32_t n, int32_t fs, int32_t * freq);

/*
 * Generalized suffix array of a 0-separated string set (T[n-1] must be 0; empty
 * strings are rejected with -1).  Every separator sorts as a distinct symbol,
 * smaller than every letter, ordered by position.
 * Replaces reference include/libsais.h:95 (src/libsais.c:7034-7048, 6886-6889).
 */
LIBSAIS_API int32_t libsais_gsa(const uint8_t * T, int32_t * SA, int32_t n, int32_t fs, int32_t * freq);

/*
 * Suffix array of an integer string (symbols in [0, k)).  T is not modified.
 * Returns 0, -1 or -2.  Replaces reference include/libsais.h:107 (src/libsais.c:7050-7063).
 */
LIBSAIS_API int32_t libsais_int(int32_t * T, int32_t * SA, int32_t n, int32_t k, int32_t fs);

/*
 * Burrows-Wheeler transform of T into U (U may alias T); A is temporary space of
 * n+fs entries whose content is unspecified on return.  Returns the primary index,
 * or -1 / -2.  Replaces reference include/libsais.h:182 (src/libsais.c:7097-7121).
 */
LIBSAIS_API int32_t libsais_bwt(const uint8_t * T, uint8_t * U, int32_t * A, int32_t n, int32_t fs, int32_t * freq);

/*
 * BWT with auxiliary indexes: U as libsais_bwt, and I[k] = 1 + the SA slot of suffix
 * k*r for k = 0..(n-1)/r (I[0] is the primary index).  r must be a power of two >= 2.
 * Returns 0, or -1 / -2.  Replaces reference include/libsais.h:196
 * (src/libsais.c:7123-7147, sampling at 4811).
 */
LIBSAIS_API int32_t libsais_bwt_aux(const uint8_t * T, uint8_t * U, int32_t * A, int32_t n, int32_t fs, int32_t * freq, int32_t r, int32_t * I);

/*
 * Inverse BWT: T holds the BWT (libsais_bwt layout) with primary index i; the original
 * string is written to U (U may alias T).  A (n+1 entries) is accepted and not used;
 * freq is accepted and not used (the counts are taken from T).  Returns 0 or -1.
 * Replaces reference include/libsais.h:289 (src/libsais.c:8030-8033).
 */
LIBSAIS_API int32_t libsais_unbwt(const uint8_t * T, uint8_t * U, int32_t * A, int32_t n, const int32_t * freq, int32_t i);

/*
 * Inverse BWT with the auxiliary indexes of libsais_bwt_aux (r = n takes I[0] as the
 * primary index).  Returns 0 or -1.
 * Replaces reference include/libsais.h:315 (src/libsais.c:8040-8056).
 */
LIBSAIS_API int32_t libsais_unbwt_aux(const uint8_t * T, uint8_t * U, int32_t * A, int32_t n, const int32_t * freq, int32_t r, const int32_t * I);

/*
 * Permuted LCP array: PLCP[i] = lcp of suffix i and the suffix preceding it in SA.
 * Returns 0 or -1.  Replaces reference include/libsais.h:368 (src/libsais.c:8363-8379).
 */
LIBSAIS_API int32_t libsais_plcp(const uint8_t * T, const int32_t * SA, int32_t * PLCP, int32_t n);

/*
 * PLCP of a generalized suffix array: matches stop at separators.
 * Replaces reference include/libsais.h:378 (src/libsais.c:8381-8397).
 */
LIBSAIS_API int32_t libsais_plcp_gsa(const uint8_t * T, const int32_t * SA, int32_t * PLCP, int32_t n);

/*
 * LCP[i] = PLCP[SA[i]] (LCP may alias SA).  Returns 0 or -1.
 * Replaces reference include/libsais.h:398 (src/libsais.c:8417-8432).
 */
LIBSAIS_API int32_t libsais_lcp(const int32_t * PLCP, const int32_t * SA, int32_t * LCP, int32_t n);

/*
 * PLCP of an integer string.  Replaces reference include/libsais.h:388
 * (src/libsais.c:8399-8415).
 */
LIBSAIS_API int32_t libsais_plcp_int(const int32_t * T, const int32_t * SA, int32_t * PLCP, int32_t n);

/*
 * Contexts (reference include/libsais.h:57, 73, 119, 131, 209, 224, 261, 277, 302, 329): a context
 * is a token, NULL on failure; the _ctx calls return -1 for a NULL context and
 * otherwise behave as the context-free calls.
 */
LIBSAIS_API void * libsais_create_ctx(void);
LIBSAIS_API void libsais_free_ctx(void * ctx);
LIBSAIS_API int32_t libsais_ctx(const void * ctx, const uint8_t * T, int32_t * SA, int32_t n, int32_t fs, int32_t * freq);
LIBSAIS_API int32_t libsais_gsa_ctx(const void * ctx, const uint8_t * T, int32_t * SA, int32_t n, int32_t fs, int32_t * freq);
LIBSAIS_API int32_t libsais_bwt_ctx(const void * ctx, const uint8_t * T, uint8_t * U, int32_t * A, int32_t n, int32_t fs, int32_t * freq);
LIBSAIS_API int32_t libsais_bwt_aux_ctx(const void * ctx, const uint8_t * T, uint8_t * U, int32_t * A, int32_t n, int32_t fs, int32_t * freq, int32_t r, int32_t * I);
LIBSAIS_API void * libsais_unbwt_create_ctx(void);
LIBSAIS_API void libsais_unbwt_free_ctx(void * ctx);
LIBSAIS_API int32_t libsais_unbwt_ctx(const void * ctx, const uint8_t * T, uint8_t * U, int32_t * A, int32_t n, const int32_t * freq, int32_t i);
LIBSAIS_API int32_t libsais_unbwt_aux_ctx(const void * ctx, const uint8_t * T, uint8_t * U, int32_t * A, int32_t n, const int32_t * freq, int32_t r, const int32_t * I);

#if defined(LIBSAIS_OPENMP)
/*
 * OpenMP forms (reference include/libsais.h:66, 144, 156, 169, 238, 253, 270, 343,
 * 357, 410, 421, 432, 443): `threads` < 0 is rejected with -1 (NULL for the context
 * constructors); otherwise the call behaves as its single-threaded form.
 */
LIBSAIS_API void * libsais_create_ctx_omp(int32_t threads);
LIBSAIS_API int32_t libsais_omp(const uint8_t * T, int32_t * SA, int32_t n, int32_t fs, int32_t * freq, int32_t threads);
LIBSAIS_API int32_t libsais_gsa_omp(const uint8_t * T, int32_t * SA, int32_t n, int32_t fs, int32_t * freq, int32_t threads);
LIBSAIS_API int32_t libsais_int_omp(int32_t * T, int32_t * SA, int32_t n, int32_t k, int32_t fs, int32_t threads);
LIBSAIS_API int32_t libsais_bwt_omp(const uint8_t * T, uint8_t * U, int32_t * A, int32_t n, int32_t fs, int32_t * freq, int32_t threads);
LIBSAIS_API int32_t libsais_bwt_aux_omp(const uint8_t * T, uint8_t * U, int32_t * A, int32_t n, int32_t fs, int32_t * freq, int32_t r, int32_t * I, int32_t threads);
LIBSAIS_API void * libsais_unbwt_create_ctx_omp(int32_t threads);
LIBSAIS_API int32_t libsais_unbwt_omp(const uint8_t * T, uint8_t * U, int32_t * A, int32_t n, const int32_t * freq, int32_t i, int32_t threads);
LIBSAIS_API int32_t libsais_unbwt_aux_omp(const uint8_t * T, uint8_t * U, int32_t * A, int32_t n, const int32_t * freq, int32_t r, const int32_t * I, int32_t threads);
LIBSAIS_API int32_t libsais_plcp_omp(const uint8_t * T, const int32_t * SA, int32_t * PLCP, int32_t n, int32_t threads);
LIBSAIS_API int32_t libsais_plcp_gsa_omp(const uint8_t * T, const int32_t * SA, int32_t * PLCP, int32_t n, int32_t threads);
LIBSAIS_API int32_t libsais_plcp_int_omp(const int32_t * T, const int32_t * SA, int32_t * PLCP, int32_t n, int32_t threads);
LIBSAIS_API int32_t libsais_lcp_omp(const int32_t * PLCP, const int32_t * SA, int32_t * LCP, int32_t n, int32_t threads);
#endif

#ifdef __cplusplus
}
#endif

#endif
