/*
 * libsais16x64.h -- 16-bit alphabet drop-in header of the MI355X suffix-array engine.
 *
 * Same prototypes, version macros and return codes as libsais 2.10.4's
 * include/libsais16x64.h (16-bit symbols, int64 indices; n up to 2^32).
 */
#ifndef LIBSAIS16X64_H
#define LIBSAIS16X64_H 1

#define LIBSAIS16X64_VERSION_MAJOR   2
#define LIBSAIS16X64_VERSION_MINOR   10
#define LIBSAIS16X64_VERSION_PATCH   4
#define LIBSAIS16X64_VERSION_STRING  "2.10.4"

#ifdef _WIN32
    #ifdef LIBSAIS_SHARED
        #ifdef LIBSAIS_EXPORTS
            #define LIBSAIS16X64_API __declspec(dllexport)
        #else
            #define LIBSAIS16X64_API __declspec(dllimport)
        #endif
    #else
        #define LIBSAIS16X64_API
    #endif
#else
    #define LIBSAIS16X64_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

#include <stdint.h>

/* Replaces reference include/libsais16x64.h:61 (src/libsais16x64.c:7036-7064). */
LIBSAIS16X64_API int64_t libsais16x64(const uint16_t * T, int64_t * SA, int64_t n, int64_t fs, int64_t * freq);

/* Replaces reference include/libsais16x64.h:72 (src/libsais16x64.c:7066-7094). */
LIBSAIS16X64_API int64_t libsais16x64_gsa(const uint16_t * T, int64_t * SA, int64_t n, int64_t fs, int64_t * freq);

/* Replaces reference include/libsais16x64.h:84 (src/libsais16x64.c:7096-7109). */
LIBSAIS16X64_API int64_t libsais16x64_long(int64_t * T, int64_t * SA, int64_t n, int64_t k, int64_t fs);

/* Replaces reference include/libsais16x64.h:135 (src/libsais16x64.c:7111-7148). */
LIBSAIS16X64_API int64_t libsais16x64_bwt(const uint16_t * T, uint16_t * U, int64_t * A, int64_t n, int64_t fs, int64_t * freq);

/* Replaces reference include/libsais16x64.h:149 (src/libsais16x64.c:7150-7187). */
LIBSAIS16X64_API int64_t libsais16x64_bwt_aux(const uint16_t * T, uint16_t * U, int64_t * A, int64_t n, int64_t fs, int64_t * freq, int64_t r, int64_t * I);

/* Replaces reference include/libsais16x64.h:191 (src/libsais16x64.c:7784-7787). */
LIBSAIS16X64_API int64_t libsais16x64_unbwt(const uint16_t * T, uint16_t * U, int64_t * A, int64_t n, const int64_t * freq, int64_t i);

/* Replaces reference include/libsais16x64.h:204 (src/libsais16x64.c:7789-7812). */
LIBSAIS16X64_API int64_t libsais16x64_unbwt_aux(const uint16_t * T, uint16_t * U, int64_t * A, int64_t n, const int64_t * freq, int64_t r, const int64_t * I);

/* Replaces reference include/libsais16x64.h:243 (src/libsais16x64.c:8050-8066). */
LIBSAIS16X64_API int64_t libsais16x64_plcp(const uint16_t * T, const int64_t * SA, int64_t * PLCP, int64_t n);

/* Replaces reference include/libsais16x64.h:253 (src/libsais16x64.c:8068-8084). */
LIBSAIS16X64_API int64_t libsais16x64_plcp_gsa(const uint16_t * T, const int64_t * SA, int64_t * PLCP, int64_t n);

/* Replaces reference include/libsais16x64.h:263 (src/libsais16x64.c:8086-8101). */
LIBSAIS16X64_API int64_t libsais16x64_lcp(const int64_t * PLCP, const int64_t * SA, int64_t * LCP, int64_t n);

#if defined(LIBSAIS_OPENMP)
/* Replaces reference include/libsais16x64.h:97 (src/libsais16x64.c:7191-7222). */
LIBSAIS16X64_API int64_t libsais16x64_omp(const uint16_t * T, int64_t * SA, int64_t n, int64_t fs, int64_t * freq, int64_t threads);

/* Replaces reference include/libsais16x64.h:109 (src/libsais16x64.c:7224-7255). */
LIBSAIS16X64_API int64_t libsais16x64_gsa_omp(const uint16_t * T, int64_t * SA, int64_t n, int64_t fs, int64_t * freq, int64_t threads);

/* Replaces reference include/libsais16x64.h:122 (src/libsais16x64.c:7257-7273). */
LIBSAIS16X64_API int64_t libsais16x64_long_omp(int64_t * T, int64_t * SA, int64_t n, int64_t k, int64_t fs, int64_t threads);

/* Replaces reference include/libsais16x64.h:163 (src/libsais16x64.c:7275-7315). */
LIBSAIS16X64_API int64_t libsais16x64_bwt_omp(const uint16_t * T, uint16_t * U, int64_t * A, int64_t n, int64_t fs, int64_t * freq, int64_t threads);

/* Replaces reference include/libsais16x64.h:178 (src/libsais16x64.c:7317-7357). */
LIBSAIS16X64_API int64_t libsais16x64_bwt_aux_omp(const uint16_t * T, uint16_t * U, int64_t * A, int64_t n, int64_t fs, int64_t * freq, int64_t r, int64_t * I, int64_t threads);

/* Replaces reference include/libsais16x64.h:218 (src/libsais16x64.c:7816-7819). */
LIBSAIS16X64_API int64_t libsais16x64_unbwt_omp(const uint16_t * T, uint16_t * U, int64_t * A, int64_t n, const int64_t * freq, int64_t i, int64_t threads);

/* Replaces reference include/libsais16x64.h:232 (src/libsais16x64.c:7821-7847). */
LIBSAIS16X64_API int64_t libsais16x64_unbwt_aux_omp(const uint16_t * T, uint16_t * U, int64_t * A, int64_t n, const int64_t * freq, int64_t r, const int64_t * I, int64_t threads);

/* Replaces reference include/libsais16x64.h:275 (src/libsais16x64.c:8105-8124). */
LIBSAIS16X64_API int64_t libsais16x64_plcp_omp(const uint16_t * T, const int64_t * SA, int64_t * PLCP, int64_t n, int64_t threads);

/* Replaces reference include/libsais16x64.h:286 (src/libsais16x64.c:8126-8145). */
LIBSAIS16X64_API int64_t libsais16x64_plcp_gsa_omp(const uint16_t * T, const int64_t * SA, int64_t * PLCP, int64_t n, int64_t threads);

/* Replaces reference include/libsais16x64.h:297 (src/libsais16x64.c:8147-8165). */
LIBSAIS16X64_API int64_t libsais16x64_lcp_omp(const int64_t * PLCP, const int64_t * SA, int64_t * LCP, int64_t n, int64_t threads);

#endif

#ifdef __cplusplus
}
#endif

#endif
