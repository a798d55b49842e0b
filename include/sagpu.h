/*
 * sagpu.h -- device-level C ABI of the MI355X suffix-array engine.
 *
 * The libsais-compatible entry points (libsais.h, libsais64.h) take host buffers and
 * are synchronous.  This header is the layer beneath them: the same computations on
 * buffers that already live in device memory (HBM), ordered on the engine's stream.
 * It has no counterpart in the reference; it exists for callers that keep data
 * resident on the GPU (benchmarks, batching, the serving loop).  All pointers named
 * d_* are device pointers; h_* are host pointers.  Return codes follow libsais:
 * 0 (or the BWT primary index) on success, -1 for invalid arguments, -2 when memory
 * cannot be obtained, -3 for a device/runtime failure (sagpu_last_error() explains).
 */
#ifndef SAGPU_H
#define SAGPU_H 1

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SAGPU_FLAG_GSA 1

#ifndef SAGPU_API
#define SAGPU_API __attribute__((visibility("default")))
#endif

/* Engine lifetime.  sagpu_init is implicit on first use; it selects device 0. */
SAGPU_API int32_t     sagpu_init(void);
SAGPU_API void        sagpu_shutdown(void);
SAGPU_API const char *sagpu_last_error(void);

/* Device memory helpers (used by harnesses that do not link a GPU framework). */
SAGPU_API void   *sagpu_malloc(size_t bytes);
SAGPU_API void    sagpu_free(void *d_ptr);
SAGPU_API int32_t sagpu_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes);
SAGPU_API int32_t sagpu_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes);
SAGPU_API int32_t sagpu_memset(void *d_dst, int value, size_t bytes);
SAGPU_API int32_t sagpu_synchronize(void);

/* Caps the engine's device working memory (bytes; 0 = unlimited).  Requests beyond
 * the cap fail with -2, exactly like an allocation failure in libsais. */
SAGPU_API void    sagpu_set_memory_limit(uint64_t bytes);
/* Releases cached (pooled) device buffers. */
SAGPU_API void    sagpu_trim(void);

/*
 * Fused pipeline on a device-resident text: suffix array plus, on request, BWT,
 * PLCP and LCP, all int32, n < 2^31.  Any of d_SA, d_U, d_PLCP, d_LCP may be NULL
 * (d_SA is then kept internally).  h_freq (256 entries, may be NULL) receives the
 * symbol counts.  Returns 0, or the BWT primary index (1-based, as libsais_bwt)
 * when d_U is given.  flags: SAGPU_FLAG_GSA for generalized suffix arrays.
 */
SAGPU_API int64_t sagpu_build32(const uint8_t *d_T, int64_t n, int32_t flags, int32_t *d_SA, uint8_t *d_U,
                      int32_t *d_PLCP, int32_t *d_LCP, int64_t *h_freq);

/* PLCP from a device-resident text and suffix array (libsais_plcp semantics). */
SAGPU_API int32_t sagpu_plcp32(const uint8_t *d_T, const int32_t *d_SA, int32_t *d_PLCP, int64_t n, int32_t flags);
/* LCP[i] = PLCP[SA[i]] on device buffers (libsais_lcp semantics). */
SAGPU_API int32_t sagpu_lcp32(const int32_t *d_PLCP, const int32_t *d_SA, int32_t *d_LCP, int64_t n);

/*
 * Kernel timing and round statistics of the most recent sagpu_build32 call, for
 * benchmarks: per-kernel-family device time (ms, from events on the engine stream)
 * and the number of suffixes still unsorted entering each doubling round.
 */
typedef struct sagpu_stats {
    int32_t rounds;               /* doubling rounds executed                      */
    int32_t initial_symbols;      /* symbols packed into the initial key           */
    int32_t bits_per_symbol;
    int32_t sort_passes;          /* onesweep passes executed in total             */
    int64_t active[64];           /* active[r] = unsorted suffixes entering round r */
    int64_t sorted_elems;         /* sum over onesweep passes of elements moved     */
    double  ms_total;             /* first to last kernel of the build             */
    double  ms_onesweep;          /* all onesweep passes                            */
    int32_t onesweep_launches;
    double  ms_initial;           /* histogram + pack + initial sort + grouping    */
    double  ms_rounds;            /* doubling rounds                               */
    double  ms_bwt;
    double  ms_plcp;
    double  ms_lcp;
} sagpu_stats;

/* timing = 1 records events around kernel families (adds a little overhead). */
SAGPU_API void    sagpu_set_timing(int32_t enabled);
SAGPU_API int32_t sagpu_get_stats(sagpu_stats *out);

#ifdef __cplusplus
}
#endif

#endif
