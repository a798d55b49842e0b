/*
 * libsais64.h -- 64-bit drop-in header of the MI355X suffix-array engine.
 *
 * Same prototypes and conventions as libsais 2.10.4's include/libsais64.h for the
 * suffix array, GSA, BWT (plain and auxiliary), inverse BWT, PLCP and LCP entry
 * points.  For n <= INT32_MAX the results
 * equal the 32-bit functions' results widened to int64 (as the reference delegates,
 * src/libsais64.c:7071-7083).
 */
#ifndef LIBSAIS64_H
#define LIBSAIS64_H 1

#define LIBSAIS64_VERSION_MAJOR   2
#define LIBSAIS64_VERSION_MINOR   10
#define LIBSAIS64_VERSION_PATCH   4
#define LIBSAIS64_VERSION_STRING  "2.10.4"

#ifdef _WIN32
    #ifdef LIBSAIS_SHARED
        #ifdef LIBSAIS_EXPORTS
            #define LIBSAIS64_API __declspec(dllexport)
        #else
            #define LIBSAIS64_API __declspec(dllimport)
        #endif
    #else
        #define LIBSAIS64_API
    #endif
#else
    #define LIBSAIS64_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

#include <stdint.h>

/* Replaces reference include/libsais64.h:61 (src/libsais64.c:7058-7086). */
LIBSAIS64_API int64_t libsais64(const uint8_t * T, int64_t * SA, int64_t n, int64_t fs, int64_t * freq);

/* Replaces reference include/libsais64.h:72 (src/libsais64.c:7088-7116). */
LIBSAIS64_API int64_t libsais64_gsa(const uint8_t * T, int64_t * SA, int64_t n, int64_t fs, int64_t * freq);

/* Integer string, symbols in [0, k); T is not modified.  Replaces reference
 * include/libsais64.h:84 (src/libsais64.c:7118-7131). */
LIBSAIS64_API int64_t libsais64_long(int64_t * T, int64_t * SA, int64_t n, int64_t k, int64_t fs);

/* Replaces reference include/libsais64.h:135 (src/libsais64.c:7133-7170). */
LIBSAIS64_API int64_t libsais64_bwt(const uint8_t * T, uint8_t * U, int64_t * A, int64_t n, int64_t fs, int64_t * freq);

/* Replaces reference include/libsais64.h:149 (src/libsais64.c:7172-7209). */
LIBSAIS64_API int64_t libsais64_bwt_aux(const uint8_t * T, uint8_t * U, int64_t * A, int64_t n, int64_t fs, int64_t * freq, int64_t r, int64_t * I);

/* Replaces reference include/libsais64.h:191 (src/libsais64.c:8034-8037).  n < 2^32. */
LIBSAIS64_API int64_t libsais64_unbwt(const uint8_t * T, uint8_t * U, int64_t * A, int64_t n, const int64_t * freq, int64_t i);

/* Replaces reference include/libsais64.h:204 (src/libsais64.c:8039-8064).  n < 2^32. */
LIBSAIS64_API int64_t libsais64_unbwt_aux(const uint8_t * T, uint8_t * U, int64_t * A, int64_t n, const int64_t * freq, int64_t r, const int64_t * I);

/* Replaces reference include/libsais64.h:243 (src/libsais64.c:8302-8318). */
LIBSAIS64_API int64_t libsais64_plcp(const uint8_t * T, const int64_t * SA, int64_t * PLCP, int64_t n);

/* Replaces reference include/libsais64.h:253 (src/libsais64.c:8320-8336). */
LIBSAIS64_API int64_t libsais64_plcp_gsa(const uint8_t * T, const int64_t * SA, int64_t * PLCP, int64_t n);

/* Replaces reference include/libsais64.h:263 (src/libsais64.c:8338-8353). */
LIBSAIS64_API int64_t libsais64_lcp(const int64_t * PLCP, const int64_t * SA, int64_t * LCP, int64_t n);

#if defined(LIBSAIS_OPENMP)
/* OpenMP forms (reference include/libsais64.h:97, 109, 122, 163, 178, 218, 232, 275, 286, 297):
 * `threads` < 0 is rejected with -1; otherwise as the single-threaded form. */
LIBSAIS64_API int64_t libsais64_omp(const uint8_t * T, int64_t * SA, int64_t n, int64_t fs, int64_t * freq, int64_t threads);
LIBSAIS64_API int64_t libsais64_gsa_omp(const uint8_t * T, int64_t * SA, int64_t n, int64_t fs, int64_t * freq, int64_t threads);
LIBSAIS64_API int64_t libsais64_long_omp(int64_t * T, int64_t * SA, int64_t n, int64_t k, int64_t fs, int64_t threads);
LIBSAIS64_API int64_t libsais64_bwt_omp(const uint8_t * T, uint8_t * U, int64_t * A, int64_t n, int64_t fs, int64_t * freq, int64_t threads);
LIBSAIS64_API int64_t libsais64_bwt_aux_omp(const uint8_t * T, uint8_t * U, int64_t * A, int64_t n, int64_t fs, int64_t * freq, int64_t r, int64_t * I, int64_t threads);
LIBSAIS64_API int64_t libsais64_unbwt_omp(const uint8_t * T, uint8_t * U, int64_t * A, int64_t n, const int64_t * freq, int64_t i, int64_t threads);
LIBSAIS64_API int64_t libsais64_unbwt_aux_omp(const uint8_t * T, uint8_t * U, int64_t * A, int64_t n, const int64_t * freq, int64_t r, const int64_t * I, int64_t threads);
LIBSAIS64_API int64_t libsais64_plcp_omp(const uint8_t * T, const int64_t * SA, int64_t * PLCP, int64_t n, int64_t threads);
LIBSAIS64_API int64_t libsais64_plcp_gsa_omp(const uint8_t * T, const int64_t * SA, int64_t * PLCP, int64_t n, int64_t threads);
LIBSAIS64_API int64_t libsais64_lcp_omp(const int64_t * PLCP, const int64_t * SA, int64_t * LCP, int64_t n, int64_t threads);
#endif

#ifdef __cplusplus
}
#endif

#endif
