/*
 * libsais16.h -- 16-bit alphabet drop-in header of the MI355X suffix-array engine.
 *
 * Same prototypes, version macros and return codes as libsais 2.10.4's
 * include/libsais16.h: suffix arrays (GSA, integer alphabet), BWT (plain and
 * auxiliary), inverse BWT, PLCP / LCP over 16-bit symbols, with the context and
 * OpenMP forms (the latter declared when LIBSAIS_OPENMP is defined).  freq tables
 * have 65536 entries.  Contexts are tokens; thread counts are checked and the work
 * runs on the GPU.
 */
#ifndef LIBSAIS16_H
#define LIBSAIS16_H 1

#define LIBSAIS16_VERSION_MAJOR   2
#define LIBSAIS16_VERSION_MINOR   10
#define LIBSAIS16_VERSION_PATCH   4
#define LIBSAIS16_VERSION_STRING  "2.10.4"

#ifdef _WIN32
    #ifdef LIBSAIS_SHARED
        #ifdef LIBSAIS_EXPORTS
            #define LIBSAIS16_API __declspec(dllexport)
        #else
            #define LIBSAIS16_API __declspec(dllimport)
        #endif
    #else
        #define LIBSAIS16_API
    #endif
#else
    #define LIBSAIS16_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

#include <stdint.h>

/* Replaces reference include/libsais16.h:57 (src/libsais16.c:6985-6988). */
LIBSAIS16_API void * libsais16_create_ctx(void);

/* Replaces reference include/libsais16.h:73 (src/libsais16.c:6990-6993). */
LIBSAIS16_API void libsais16_free_ctx(void * ctx);

/* Replaces reference include/libsais16.h:84 (src/libsais16.c:6995-7009). */
LIBSAIS16_API int32_t libsais16(const uint16_t * T, int32_t * SA, int32_t n, int32_t fs, int32_t * freq);

/* Replaces reference include/libsais16.h:95 (src/libsais16.c:7011-7025). */
LIBSAIS16_API int32_t libsais16_gsa(const uint16_t * T, int32_t * SA, int32_t n, int32_t fs, int32_t * freq);

/* Replaces reference include/libsais16.h:107 (src/libsais16.c:7027-7040). */
LIBSAIS16_API int32_t libsais16_int(int32_t * T, int32_t * SA, int32_t n, int32_t k, int32_t fs);

/* Replaces reference include/libsais16.h:119 (src/libsais16.c:7042-7056). */
LIBSAIS16_API int32_t libsais16_ctx(const void * ctx, const uint16_t * T, int32_t * SA, int32_t n, int32_t fs, int32_t * freq);

/* Replaces reference include/libsais16.h:131 (src/libsais16.c:7058-7072). */
LIBSAIS16_API int32_t libsais16_gsa_ctx(const void * ctx, const uint16_t * T, int32_t * SA, int32_t n, int32_t fs, int32_t * freq);

/* Replaces reference include/libsais16.h:182 (src/libsais16.c:7074-7098). */
LIBSAIS16_API int32_t libsais16_bwt(const uint16_t * T, uint16_t * U, int32_t * A, int32_t n, int32_t fs, int32_t * freq);

/* Replaces reference include/libsais16.h:196 (src/libsais16.c:7100-7123). */
LIBSAIS16_API int32_t libsais16_bwt_aux(const uint16_t * T, uint16_t * U, int32_t * A, int32_t n, int32_t fs, int32_t * freq, int32_t r, int32_t * I);

/* Replaces reference include/libsais16.h:209 (src/libsais16.c:7125-7149). */
LIBSAIS16_API int32_t libsais16_bwt_ctx(const void * ctx, const uint16_t * T, uint16_t * U, int32_t * A, int32_t n, int32_t fs, int32_t * freq);

/* Replaces reference include/libsais16.h:224 (src/libsais16.c:7151-7174). */
LIBSAIS16_API int32_t libsais16_bwt_aux_ctx(const void * ctx, const uint16_t * T, uint16_t * U, int32_t * A, int32_t n, int32_t fs, int32_t * freq, int32_t r, int32_t * I);

/* Replaces reference include/libsais16.h:261 (src/libsais16.c:7769-7772). */
LIBSAIS16_API void * libsais16_unbwt_create_ctx(void);

/* Replaces reference include/libsais16.h:277 (src/libsais16.c:7774-7777). */
LIBSAIS16_API void libsais16_unbwt_free_ctx(void * ctx);

/* Replaces reference include/libsais16.h:289 (src/libsais16.c:7779-7782). */
LIBSAIS16_API int32_t libsais16_unbwt(const uint16_t * T, uint16_t * U, int32_t * A, int32_t n, const int32_t * freq, int32_t i);

/* Replaces reference include/libsais16.h:302 (src/libsais16.c:7784-7787). */
LIBSAIS16_API int32_t libsais16_unbwt_ctx(const void * ctx, const uint16_t * T, uint16_t * U, int32_t * A, int32_t n, const int32_t * freq, int32_t i);

/* Replaces reference include/libsais16.h:315 (src/libsais16.c:7789-7805). */
LIBSAIS16_API int32_t libsais16_unbwt_aux(const uint16_t * T, uint16_t * U, int32_t * A, int32_t n, const int32_t * freq, int32_t r, const int32_t * I);

/* Replaces reference include/libsais16.h:329 (src/libsais16.c:7807-7823). */
LIBSAIS16_API int32_t libsais16_unbwt_aux_ctx(const void * ctx, const uint16_t * T, uint16_t * U, int32_t * A, int32_t n, const int32_t * freq, int32_t r, const int32_t * I);

/* Replaces reference include/libsais16.h:368 (src/libsais16.c:8064-8080). */
LIBSAIS16_API int32_t libsais16_plcp(const uint16_t * T, const int32_t * SA, int32_t * PLCP, int32_t n);

/* Replaces reference include/libsais16.h:378 (src/libsais16.c:8082-8098). */
LIBSAIS16_API int32_t libsais16_plcp_gsa(const uint16_t * T, const int32_t * SA, int32_t * PLCP, int32_t n);

/* Replaces reference include/libsais16.h:388 (src/libsais16.c:8100-8115). */
LIBSAIS16_API int32_t libsais16_lcp(const int32_t * PLCP, const int32_t * SA, int32_t * LCP, int32_t n);

#if defined(LIBSAIS_OPENMP)
/* Replaces reference include/libsais16.h:66 (src/libsais16.c:7178-7186). */
LIBSAIS16_API void * libsais16_create_ctx_omp(int32_t threads);

/* Replaces reference include/libsais16.h:144 (src/libsais16.c:7188-7205). */
LIBSAIS16_API int32_t libsais16_omp(const uint16_t * T, int32_t * SA, int32_t n, int32_t fs, int32_t * freq, int32_t threads);

/* Replaces reference include/libsais16.h:156 (src/libsais16.c:7207-7224). */
LIBSAIS16_API int32_t libsais16_gsa_omp(const uint16_t * T, int32_t * SA, int32_t n, int32_t fs, int32_t * freq, int32_t threads);

/* Replaces reference include/libsais16.h:169 (src/libsais16.c:7226-7242). */
LIBSAIS16_API int32_t libsais16_int_omp(int32_t * T, int32_t * SA, int32_t n, int32_t k, int32_t fs, int32_t threads);

/* Replaces reference include/libsais16.h:238 (src/libsais16.c:7244-7271). */
LIBSAIS16_API int32_t libsais16_bwt_omp(const uint16_t * T, uint16_t * U, int32_t * A, int32_t n, int32_t fs, int32_t * freq, int32_t threads);

/* Replaces reference include/libsais16.h:253 (src/libsais16.c:7273-7299). */
LIBSAIS16_API int32_t libsais16_bwt_aux_omp(const uint16_t * T, uint16_t * U, int32_t * A, int32_t n, int32_t fs, int32_t * freq, int32_t r, int32_t * I, int32_t threads);

/* Replaces reference include/libsais16.h:270 (src/libsais16.c:7827-7835). */
LIBSAIS16_API void * libsais16_unbwt_create_ctx_omp(int32_t threads);

/* Replaces reference include/libsais16.h:343 (src/libsais16.c:7837-7840). */
LIBSAIS16_API int32_t libsais16_unbwt_omp(const uint16_t * T, uint16_t * U, int32_t * A, int32_t n, const int32_t * freq, int32_t i, int32_t threads);

/* Replaces reference include/libsais16.h:357 (src/libsais16.c:7842-7861). */
LIBSAIS16_API int32_t libsais16_unbwt_aux_omp(const uint16_t * T, uint16_t * U, int32_t * A, int32_t n, const int32_t * freq, int32_t r, const int32_t * I, int32_t threads);

/* Replaces reference include/libsais16.h:400 (src/libsais16.c:8119-8138). */
LIBSAIS16_API int32_t libsais16_plcp_omp(const uint16_t * T, const int32_t * SA, int32_t * PLCP, int32_t n, int32_t threads);

/* Replaces reference include/libsais16.h:411 (src/libsais16.c:8140-8159). */
LIBSAIS16_API int32_t libsais16_plcp_gsa_omp(const uint16_t * T, const int32_t * SA, int32_t * PLCP, int32_t n, int32_t threads);

/* Replaces reference include/libsais16.h:422 (src/libsais16.c:8161-8179). */
LIBSAIS16_API int32_t libsais16_lcp_omp(const int32_t * PLCP, const int32_t * SA, int32_t * LCP, int32_t n, int32_t threads);

#endif

#ifdef __cplusplus
}
#endif

#endif
