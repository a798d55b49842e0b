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
/* Device-to-device copy by the engine's own copy kernel (16-byte vector loads and
 * stores); also the copy-bandwidth yardstick of the benchmark. */
SAGPU_API int32_t sagpu_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes);
SAGPU_API int32_t sagpu_synchronize(void);
/* Host-to-host copy split over the library's copy threads, the pool that fills and
 * drains the pinned slots of the libsais calls (threads = 0: the copy_threads option).
 * Touches no device; returns -1 for a negative thread count. */
SAGPU_API int32_t sagpu_host_memcpy(void *h_dst, const void *h_src, size_t bytes, int32_t threads);

/* Caps the device working memory of each engine context (bytes; 0 = unlimited).
 * Requests beyond the cap fail with -2, exactly like an allocation failure in libsais. */
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

/*
 * Batched generalized suffix arrays: d_T holds `count` strings, each terminated by a
 * 0 byte; string k occupies [h_offsets[k], h_offsets[k+1]) (host array of count+1
 * entries, h_offsets[0] = 0, h_offsets[count] = n, no empty strings).  For every
 * string the result equals libsais_gsa on that string alone (SA of length
 * h_offsets[k+1]-h_offsets[k], local indices) and, when d_LCP is given,
 * libsais_plcp_gsa + libsais_lcp on it; both are written into
 * d_SA[h_offsets[k] ..] and d_LCP[h_offsets[k] ..].  One pipeline run serves all
 * strings: keys carry the string id, so groups never span strings.
 */
SAGPU_API int32_t sagpu_gsa_batch32(const uint8_t *d_T, int64_t n, const int64_t *h_offsets, int32_t count,
                                    int32_t *d_SA, int32_t *d_LCP);

/*
 * Serving.  A job is one text and the outputs it asks for (host buffers the caller
 * keeps alive until the job completes):
 *   SAGPU_JOB_SA   SA[0..n)                      (libsais)
 *   SAGPU_JOB_BWT  SA[0..n), U[0..n), primary    (libsais + libsais_bwt)
 *   SAGPU_JOB_LCP  SA[0..n), LCP[0..n)           (libsais + libsais_plcp + libsais_lcp)
 * status: 1 while pending, then 0 or a libsais error code.  Times are seconds on
 * sagpu_clock()'s monotonic clock.
 */
#define SAGPU_JOB_SA  0
#define SAGPU_JOB_BWT 1
#define SAGPU_JOB_LCP 2

typedef struct sagpu_job {
    const uint8_t *T;
    int64_t n;
    int32_t kind;
    int32_t *SA;
    uint8_t *U;
    int32_t *LCP;
    int64_t primary;
    volatile int32_t status;
    int32_t batched;        /* 1 when the job ran inside a multi-job batch */
    double t_submit, t_start, t_done;
    int32_t batchable;      /* set by the server: text has no 0 byte */
    int32_t reserved_;
} sagpu_job;

/* Runs one job synchronously on the calling thread's default stream (the plain,
 * one-job-at-a-time path). */
SAGPU_API int32_t sagpu_run_job(sagpu_job *job);

/*
 * Continuous-batching server: `slots` worker streams, each with its own pooled HBM.
 * A free slot takes the oldest waiting job; if that job is small (< 1 MiB, no 0
 * bytes) it also takes the following small jobs, up to batch_bytes of text, and
 * runs them as one batched build (string-id tagged keys).  Work is admitted only
 * while the estimated device memory of everything running stays within hbm_budget
 * (0 = 90% of the device's free memory); a job that can never fit fails with -2.
 * Jobs that arrive while others run are admitted as soon as a slot and memory are
 * free, and slots run concurrently on the device.
 */
SAGPU_API void   *sagpu_server_create(int32_t slots, uint64_t hbm_budget, int64_t batch_bytes);
SAGPU_API int32_t sagpu_server_submit(void *server, sagpu_job *job);
SAGPU_API int32_t sagpu_server_wait(void *server, sagpu_job *job);
SAGPU_API int32_t sagpu_server_drain(void *server);
SAGPU_API void    sagpu_server_destroy(void *server);
SAGPU_API double  sagpu_clock(void);

/*
 * Allocation-free form of sagpu_build32.  It has no counterpart in the reference:
 * it is for callers that cannot allocate device memory during a call (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments, results and return codes are those of sagpu_build32; every working
 * buffer is carved from one static device array of SAGPU_NOALLOC_SCRATCH_BYTES
 * per GPU (held by each GPU the library runs on until the process ends), so the
 * call makes no device or pinned allocation.  n is limited to SAGPU_NOALLOC_MAX_N
 * (-1 beyond, with a message in sagpu_last_error()).  Set-up (the scratch address,
 * a stream, 64 KiB of pinned staging) happens in sagpu_init() for the current
 * device, or on the first call on a device.  All calls on one GPU share its array,
 * which keeps the last call's contents: the library serialises them (a second
 * thread waits until the first call returns), since overlapping calls would race
 * on it.
 */
#define SAGPU_NOALLOC_MAX_N 4194304
#define SAGPU_NOALLOC_SCRATCH_BYTES ((size_t)288 << 20)
SAGPU_API int64_t sagpu_build32_noalloc(const uint8_t *d_T, int64_t n, int32_t flags, int32_t *d_SA, uint8_t *d_U,
                                        int32_t *d_PLCP, int32_t *d_LCP, int64_t *h_freq);
/* PLCP from a device-resident text and suffix array (libsais_plcp semantics). */
SAGPU_API int32_t sagpu_plcp32(const uint8_t *d_T, const int32_t *d_SA, int32_t *d_PLCP, int64_t n, int32_t flags);
/* LCP[i] = PLCP[SA[i]] on device buffers (libsais_lcp semantics). */
SAGPU_API int32_t sagpu_lcp32(const int32_t *d_PLCP, const int32_t *d_SA, int32_t *d_LCP, int64_t n);

/*
 * Kernel timing and round statistics of the most recent build (of any thread), for
 * benchmarks: per-kernel-family device time (ms, from events on the engine stream)
 * and the number of suffixes still unsorted entering each doubling round.
 */
typedef struct sagpu_stats {
    int32_t rounds;               /* doubling rounds executed                      */
    int32_t initial_symbols;      /* symbols packed into the initial key           */
    int32_t bits_per_symbol;
    int32_t sort_passes;          /* onesweep passes executed in total             */
    int64_t active[64];           /* active[r] = unsorted suffixes entering round r */
    int64_t sorted_elems;         /* sum over radix passes of elements moved        */
    int64_t sorted_bytes;         /* algorithmic bytes of those passes (read+write) */
    int32_t big_rounds;           /* rounds that sorted groups too big for LDS      */
    int64_t big_elems;            /* elements those rounds sorted globally          */
    double  ms_total;             /* first to last kernel of the build             */
    double  ms_onesweep;          /* all onesweep passes                            */
    int32_t onesweep_launches;
    double  ms_initial;           /* histogram + pack + initial sort + grouping    */
    double  ms_rounds;            /* doubling rounds                               */
    double  ms_bwt;
    double  ms_plcp;
    double  ms_lcp;
    double  ms_onesweep_u32;      /* the u32-key passes among ms_onesweep           */
    int32_t onesweep_launches_u32;
    int64_t sorted_bytes_u32;     /* algorithmic bytes of the u32-key passes        */
    double  ms_plcp_pairs_pass;   /* the text-order PLCP pass over pair words      */
    int32_t plcp_pairs_launches;
    int64_t device_allocs;        /* device / pinned allocations the engine has made for
                                     its working memory in this process so far       */
    int64_t peak_bytes;           /* largest device working memory the call held at once
                                     (pool buffers, or the allocation-free scratch)   */
    /* Per-kernel device time (ms, events around each launch), launches and
     * algorithmic bytes (reads + writes at element size) of the build, timing mode
     * only.  Index: 0 radix pass u32 key + u32 value, 1 generating first radix pass,
     * 2 radix pass u64 key + u32 value, 3 k_init_emit, 4 the initial ranks' scatter,
     * 5 k_round_gather, 6 k_round_tile, 7 k_plcp_pairs, 8 k_plcp_irr_pairs,
     * 9 k_lcp_fix, 10 k_bwt_seeded; 11-15 unused. */
    double  kernel_ms[16];
    int32_t kernel_launches[16];
    int64_t kernel_bytes[16];
    int32_t round_graphs;         /* doubling-round graphs the build captured (0 when the
                                     previous build's graphs served unchanged)        */
    int32_t round_graph_launches; /* doubling rounds the build ran as graph replays   */
    int32_t big_pivot_rounds;     /* rounds whose big groups took frequent keys out of
                                     the sort (option big_pivot_min)                 */
    int64_t big_pivot_elems;      /* big-group members those rounds placed unsorted  */
} sagpu_stats;

/*
 * Stable LSD radix sort of (key, value) pairs in device memory, in place, by key
 * bits [begin_bit, end_bit).  Exposes the engine's sort primitive on its own.
 */
SAGPU_API int32_t sagpu_sort_pairs_u32(uint32_t *d_keys, uint32_t *d_vals, int64_t n, int32_t begin_bit,
                                       int32_t end_bit);
SAGPU_API int32_t sagpu_sort_pairs_u64(uint64_t *d_keys, uint32_t *d_vals, int64_t n, int32_t begin_bit,
                                       int32_t end_bit);

/*
 * Tuning knobs (process-wide):
 *   "radix_variant"  0 = onesweep (decoupled look-back), 1 = reduce-then-scan
 *   "radix_items"    keys per thread per radix tile: 8, 16 or 32 (2048-, 4096- or 8192-key
 *                    tiles; default 32)
 *   "radix_threads"  onesweep workgroup size with 32 keys per thread: 256 (8192-key tiles) or
 *                    512 (16384-key tiles, default; 32-bit keys only)
 *   "radix_bits"     digit width of the onesweep radix sort: 8 (default), 4 or 11 (4 and 11 use
 *                    256-thread tiles and the onesweep variant; the allocation-free build keeps 8)
 *   "tile_range"     list entries per workgroup of the doubling-round tile kernel: 256, 512
 *                    (default) or 1024
 *   "bwt_variant"    0 = seeded (default): the initial sort carries each suffix's BWT symbol
 *                    beside its index where both fit in 32 bits, and the text is read only for
 *                    the rows the rounds place; otherwise as 2.  1 = scatter in text order,
 *                    2 = gather T[SA[j]-1] in SA order
 *   "key_stream"     1 = byte texts (no GSA): the initial sort's keys are generated inside its
 *                    first pass from a packed code stream (default); 0 = written first
 *   "raw_round1"     1 = with key_stream, the first doubling round orders each group by the
 *                    packed key that follows (same order as its rank; default); 0 = by rank
 *   "exact_round1"   1 = with raw_round1, the first round records the exact lcp across each
 *                    split, from the two raw keys (default); 0 = a lower bound, left to the PLCP pass
 *   "plcp_hint"      1 = the fused build with LCP (texts up to 2^28) packs the exact lcp of each
 *                    suffix alone in its initial group into bits 28-31 of its ISA word (15 values
 *                    on the largest marks the initial key allows), and the PLCP pass takes it
 *                    instead of reading a pair word (default); 0 = no hints
 *   "stream_compare" 1 = with key_stream, the PLCP pass compares suffixes on the packed code
 *                    stream, 64 / bits symbols per step (default); 0 = on the bytes
 *   "host_staging"   copies between caller host memory and the device: 1 = copies of at least
 *                    4 MiB through pinned 32 MiB slots, host side split over copy threads
 *                    (default); 0 = direct hipMemcpy from pageable memory; 2 = every copy
 *                    staged in 64 KiB slots
 *   "copy_threads"   host threads for staged copies, 1-64; 0 = min(16, hardware threads) (default)
 *   "bucketed_phi"   1 = Φ of a caller's suffix array (libsais_plcp and the like) through
 *                    pairs bucketed by target, scattered bucket after bucket (default); 0 = one pass
 *   "plcp_pairs"     1 = the fused PLCP pass reads one 8-byte pair word per position, built in
 *                    SA order (default); 0 = it gathers the mark, SA[j-1] and T[k-1] separately
 *   "isa_window_log2" ISA scatter after the initial sort: (suffix, rank) pairs are bucketed by
 *                    suffix into buckets of 2^v (10..32, default 25; at most 64 buckets) and
 *                    scattered bucket after bucket
 *   "wide_from"      texts longer than this use 64-bit rank keys (default 2^31 - 1, the
 *                    largest the 32-bit keys hold; lower values exercise the wide mode)
 *   "round_graphs"   1 = doubling rounds from the second on replay a captured HIP graph per
 *                    buffer parity (parameter copy, gather, tile kernel, counter copy), updated
 *                    in place across builds (default; not in timing mode or the allocation-free
 *                    build); 0 = kernels launched one by one
 *   "windowed_scatter" 1 = texts over 2^22 write the initial ranks through a radix pass into
 *                    windows of the ISA and a scatter in which each XCD writes only its own windows
 *                    (default); 0 = one direct scatter
 *   "radix_small_n"  radix sorts of fewer keys use 2048-key tiles and run every pass without
 *                    waiting for the histograms on the host (default 2^20; 0 = never)
 *   "big_pivot_min"  rounds after the first whose big groups hold at least this many members in
 *                    all take up to 4 keys frequent in a 64-member sample of each group out of the
 *                    sort: those members become subgroups by a stable partition (default 2^20;
 *                    0 = never)
 *   "group_flags"    1 = texts of at most 2^31 suffixes let a group's members hold any slot of
 *                    the group's range as their rank (a flag in the group's list word, the rank
 *                    stored at the group's first slot + 1), so that the subgroup whose slots hold
 *                    it keeps its ranks when the group splits (default); 0 = every group ranks at
 *                    its first slot
 *   "init_rows"      rows of 64 positions per wave in the initial grouping of texts of at least
 *                    2^20 suffixes: 8 (2048-position tiles, default) or 16
 *   "perm_hist"      1 = the radix pass that groups the initial ranks into ISA windows takes its
 *                    digit histogram from the closed form of a permutation of [0, n) (default);
 *                    0 = counted
 *   "hist_windows"   1 = the initial sort's histograms, when its keys come from the code stream
 *                    with whole symbols per 8-bit digit, are derived from one histogram of the
 *                    stream's 8-bit windows (default); 0 = counted per pass
 *   "api_slots"      contexts that calls made without a libsais context share, 1-16
 *                    (default 8): that many such calls run at once, each on its own stream
 * Returns 0, or -1 for an unknown name or value.
 */
SAGPU_API int32_t sagpu_set_option(const char *name, int64_t value);

/*
 * Sort micro-benchmark: sorts n pseudo-random (u32 or u64) keys with u32 values
 * `reps` times on device-generated data and returns the mean ms per sort and per
 * radix pass.  key_bytes = 4 or 8.
 */
SAGPU_API int32_t sagpu_sort_bench(int64_t n, int32_t key_bytes, int32_t key_bits, int32_t reps, double *ms_per_sort,
                                   double *ms_per_pass);

/* timing = 1 records events around kernel families (adds a little overhead). */
SAGPU_API void    sagpu_set_timing(int32_t enabled);
SAGPU_API int32_t sagpu_get_stats(sagpu_stats *out);

#ifdef __cplusplus
}
#endif

#endif
