// engine.h -- internal interfaces between the host orchestration (engine.hip) and
// the kernel translation units.  Nothing here is part of the public C ABI.
#pragma once

#include "common.h"

#include <stddef.h>

// ------------------------------------------------------------------ scan
// Exclusive prefix sum of a u32 array in place (reduce-then-scan, 3 kernels).
struct ScanScratch {
    u32 *partials;    // [cap] per-tile sums
    u64 cap;          // tiles of 4096 elements
};
int scan_exclusive_u32(ScanScratch &sc, u32 *data, u64 count, hipStream_t s);

// ------------------------------------------------------------------ radix sort
// Keys and values of the initial sort computed from a packed code stream instead of
// read from arrays: key(p) = the kbits stream bits from bit p*bits (codes of T[p..],
// first symbol highest), value(p) = p, or (p << cb) | code of T[p-1] when cb > 0.
// pw holds the stream MSB-first in u32 words, zero past the text (pack_stream_words).
struct KeyGen {
    const u32 *pw;
    u64 words;        // pack_stream_words(n, bits)
    int bits, kbits, cb;
};
u64 pack_stream_words(u64 n, int bits);
// 1: the upfront histograms of generated keys whose digits hold whole symbols come
// from one histogram of the stream's 8-bit windows (the "hist_windows" option)
extern int g_hist_windows;
int launch_pack_stream(const uint8_t *T, u64 n, const uint8_t *lut, int bits, u32 *pw, hipStream_t s);

// Onesweep LSD radix sort, 8-bit digits, stable.  Keys are u64 or u32, values u32.
// The result ends in (k0,v0) or in (k1,v1); *result_in_alt says which.
// Kernel timing for benchmarks: when timing is on for the calling thread's engine
// context, ktime_begin / ktime_end bracket a launch with events on its stream and
// record the launch's algorithmic bytes (reads + writes at element size).  The ids
// index sagpu_stats.kernel_ms / kernel_launches / kernel_bytes.
enum KernelId {
    KI_RADIX32 = 0,   // onesweep pass, u32 key + u32 value
    KI_RADIX_GEN,     // the initial sort's first pass, keys cut from the code stream
    KI_RADIX64,       // onesweep pass, u64 key + u32 value
    KI_INIT_EMIT,     // k_init_emit
    KI_ISA_SCATTER,   // k_isa_scatter
    KI_ROUND_GATHER,  // k_round_gather
    KI_ROUND_TILE,    // k_round_tile (+ k_big_sizes)
    KI_PLCP_WORDS,    // k_plcp_pairs
    KI_PLCP_PASS,     // k_plcp_irr_pairs
    KI_LCP_FIX,       // k_lcp_fix
    KI_BWT,           // k_bwt_seeded
    KI_COUNT
};
void ktime_begin(int id);
void ktime_end(int id, u64 bytes);

// Digit histograms of a sort: counts [pass][bins] in the first RADIX_HIST_WORDS
// words, bucket offsets in the next RADIX_HIST_WORDS (passes x bins <= 6 x 2048).
constexpr u64 RADIX_HIST_WORDS = 12288;
constexpr int RADIX_MAX_PASSES = 16;   // 64-bit keys with 4-bit digits
struct RadixScratch {
    u64 *hist;        // [2 * RADIX_HIST_WORDS] global digit histograms, then offsets
    u64 *lookback;    // [tiles][bins] look-back words (epoch-tagged)
    u32 *tile_ctr;    // [RADIX_MAX_PASSES] tile tickets, one per pass
    u64 lb_words;     // capacity of lookback in words
    u32 epoch;        // advanced once per pass; lookback words of older passes are ignored
    u64 *hist_host;   // pinned host copy of the counts (RADIX_HIST_WORDS)
    int bits;         // digit width: 8 (default), 4 or 11
    // Optional hook called right before (phase 0) and after (phase 1) each pass
    // launch; the engine uses it to bracket passes with timing events.
    void (*hook)(void *user, int phase);
    void *hook_user;
    int variant;      // RADIX_ONESWEEP or RADIX_REDUCE_THEN_SCAN
    int items;        // keys per thread per tile: 8, 16 or 32 (tile = threads * items)
    int threads;      // onesweep workgroup size with 32 items: 256 or 512
    u64 small_n;      // sorts of fewer keys take the smallest tiles (items and threads ignored)
    u32 *counts;      // [counts_cap] per-tile digit counts (reduce-then-scan)
    u64 counts_cap;
    ScanScratch scan;
    // u32 sorts only: when set, the input (k0, v0) is not read; the first pass
    // generates it (and a sort whose passes are all skipped writes it to k0, v0).
    // gen_pass is 1 while that pass is bracketed by the hook.
    const KeyGen *gen;
    int gen_pass;
    // u32 sorts only: when set, the keys are a permutation of [0, n) (each value once,
    // as the suffixes of the ISA pairs), so every digit histogram has a closed form:
    // it is computed on the device without reading the keys, and the host decides
    // the constant digits without reading it back.
    int perm_keys;
};

enum { RADIX_ONESWEEP = 0, RADIX_REDUCE_THEN_SCAN = 1 };

enum { RS_THREADS = 256, RS_ITEMS = 8, RS_TILE = RS_THREADS * RS_ITEMS };  // smallest tile: sizes the scratch

int radix_sort_u64(RadixScratch &rs, u64 *k0, u32 *v0, u64 *k1, u32 *v1, u64 n, int begin_bit,
                   int end_bit, hipStream_t s, int *result_in_alt, int *passes_run);
int radix_sort_u32(RadixScratch &rs, u32 *k0, u32 *v0, u32 *k1, u32 *v1, u64 n, int begin_bit,
                   int end_bit, hipStream_t s, int *result_in_alt, int *passes_run);

// ------------------------------------------------------- rounds (rounds.hip)
// ISA words of texts of at most 2^28 suffixes may carry a PLCP hint in bits 28-31
// (the fused build with LCP: the lcp of a suffix alone in its initial group, coded
// as in PlcpHint; 0 = none).  Every ISA reader takes the rank through isa_rank; for
// longer texts nothing is packed and the word is the rank.
__host__ __device__ __forceinline__ u32 isa_rank(u32 x, u64 n) { return n <= (1ull << 28) ? (x & 0x0FFFFFFFu) : x; }
__host__ __device__ __forceinline__ u32 isa_hint(u32 x, u64 n) { return n <= (1ull << 28) ? (x >> 28) : 0u; }
// A hint h (1..15) stands for the lcp off + h - 1: off = max(0, ksym - 10) puts the
// 15 values on the largest marks an initial key allows and the exact marks of the
// first doubling round (ksym plus the few symbols two raw keys share; DNA with
// 16-symbol keys: 6..20).  on (in) asks for the hints; on (out) tells whether they
// were packed.
struct PlcpHint {
    int on = 0;
    int off = 0;
};
// LCP marks written during construction: LCP_UNKNOWN | lower bound for pairs that
// were separated by a doubling round (their lcp is resolved afterwards).
#define LCP_UNKNOWN 0x80000000u
#define LCP_BOUND_CAP 0x40000000u
enum { RT_SMALL = 256 };   // groups up to this size are sorted inside one workgroup
// texts shorter than this are grouped in 512-position tiles instead of 4096
constexpr u64 INIT_SMALL_N = 1ull << 20;

struct WsScratch {
    void *tile_agg;   // [tiles_cap] 16-byte tile aggregates, then [tiles_cap / 8192 + 2] chunk aggregates
    u64 tiles_cap;    // tiles of at least 2048 elements (512 below INIT_SMALL_N)
    u64 *total;       // [1]
    u64 *base;        // [1] reserved output base
    u32 *bcnt;        // [bcnt_cap] ISA bucket counts per tile, then their scan (initial grouping)
    u64 bcnt_cap;
    ScanScratch scan; // for bcnt
};

// The packed code stream of a byte text (KeyGen), kept for the later stages.
struct CodeStream {
    const u32 *pw;
    int bits;
};
// BWT seeding: the initial sort's values carry (suffix << cb) | code of the symbol
// before the suffix, the active lists keep that packed form, and every stage that
// places a suffix in SA slot j writes its code to codes[j] (n bytes), so the BWT is
// one sequential pass over codes (launch_bwt_seeded).  inv: code -> byte (256 bytes).
struct BwtSeed {
    int cb;
    uint8_t *codes;
    const uint8_t *inv;
};
// What the rounds need of it: cb = 0 and codes = null when the build is not seeded.
struct ActCode {
    int cb;
    uint8_t *codes;
    // Group-word flag (0x80000000 for texts of at most 2^31 suffixes, else 0).  A
    // group's members share one rank in ISA, any slot of the group's range: its
    // first slot g, or, where the group word (act_g, g | flag) carries the flag, the
    // value stored in SA[g + 1] (a group of two or more members has no suffix
    // placed in its slots yet).  When a group splits, the subgroup whose slots hold
    // that rank keeps it and the flag (storing it at its own first slot + 1), so
    // its members' ISA entries are not rewritten; the others get new ranks.  A
    // pivot subgroup (the few-key partition of big groups) takes the middle of its
    // slots, flagged: on texts of long repeats it holds nearly all of the next
    // round's group, whose surviving subgroup then keeps that rank round after
    // round.  Other new ranks are first slots, unflagged.
    u32 gf = 0;
};
// PV, PG: scratch of n entries each ((suffix, group head) pairs bucketed for the
// ISA scatter).
int group_init(WsScratch &ws, const u32 *keys, const u32 *vals, u64 n, int bits, int ksym, int gsa, u32 *SA, u32 *ISA,
               u32 *act_i, u32 *act_g, u32 *LCP, u32 *PV, u32 *PG, hipStream_t s, const BwtSeed *seed = nullptr,
               const PlcpHint *hint = nullptr);
// ISA scatter after the initial sort: buckets of 2^v suffixes (v in [10, 32]; at most
// 64 buckets, larger ones for longer texts).
void set_isa_window_log2(int v);
// log2 of the bucket size used for an n-entry permutation scatter (at most 64 buckets).
int isa_bucket_log2(u64 n);
// List entries per workgroup of the tile round kernel: 256, 512 (default) or 1024.
void set_tile_range(int v);
// Batched GSA: keys = (string id << 32) | packed symbols.
int group_init64(WsScratch &ws, const u64 *keys, const u32 *vals, u64 n, int bits, int ksym, u32 *SA, u32 *ISA,
                 u32 *act_i, u32 *act_g, u32 *LCP, u32 *PV, u32 *PG, hipStream_t s);
// First round from the code stream: key2 = packed key of i + h (h = symbols per key)
// for every group except those holding one of the suffixes shorter than h (their
// group heads in special[0..nspecial)), which keep key2 = h + ISA[i+h] / n-1-i.
struct RawKey1 {
    const u32 *pw;    // null: rank form everywhere
    int bits, kbits, nspecial;
    u32 special[32];
};
// Round 1 on raw keys: key2 is the packed `kbits`-bit code of the ksym symbols at
// i + h, so where the round splits a group, the lcp of the two adjacent subgroups
// is h + the leading symbols their keys share, exactly.  Groups in `skip` keep the
// LCP_UNKNOWN | h mark: those holding one of the last 2 ksym suffixes, whose keys
// are rank-form or padded past the text end (padding would count as shared symbols).
struct ExactLcp {
    int on = 0;
    int bits = 0, kbits = 0, nskip = 0;
    u32 skip[64] = {};
    // PLCP hints (PlcpHint) for the suffixes the first round places with an exact
    // mark: the hint bits of the rank written to ISA
    int hint = 0, hint_off = 0;
    __device__ u32 hint_bits(u32 mark) const
    {
        return hint && mark >= (u32)hint_off && mark - (u32)hint_off < 15u ? (mark - (u32)hint_off + 1u) << 28 : 0u;
    }
    __device__ bool skipped(u32 g) const
    {
        for (int t = 0; t < nskip; t++)
            if (skip[t] == g) return true;
        return false;
    }
    // symbols two different raw keys share from the top
    __device__ u32 shared(u64 a, u64 b) const
    {
        const u32 x = (u32)(a ^ b);
        return (u32)(__clz(x) - (32 - kbits)) / (u32)bits;
    }
};
// Rounds.  wide = 0: key2 is u32 (n < 2^31); wide = 1: key2 is u64 (n up to 2^32).
// The round's list length a and offset h are read from rp[0], rp[1] in device memory
// (the launch geometry is sized on the host for a list of a_grid entries, which
// must be at least rp[0]), so one captured launch serves every later round.
extern int g_tile_range;
// rows of 64 positions per wave span in the initial grouping of texts of at least
// INIT_SMALL_N suffixes (8 or 16: tiles of 2048 or 4096 positions; the "init_rows" option)
extern int g_init_rows;   // list entries per workgroup of the tile kernel (the "tile_range" option)
int round_gather(const u32 *act_i, const u32 *act_g, u64 a_grid, const u64 *rp, const u32 *ISA, u64 n, void *key2,
                 int wide, hipStream_t s, const RawKey1 *rk = nullptr, int cb = 0);
// The tile kernel's survivors go to 8 per-XCD segments of seg_i / seg_g (capacity
// 8 * round_seg_cap(a_grid) each) and are concatenated into out_i / out_g, whose
// length goes to *out_counter; seg_ctr: 128 u64.  Big-group heads go the same way
// to 8 segments of bseg_h / bseg_g (round_bigseg_cap(n) entries each, n >= every
// list length) and are concatenated into big_head / big_g / big_size, their count
// to *big_counter; boff (big_cap + 1 entries) gets their offsets in the
// concatenation of their members, *big_total the member count.
u64 round_seg_cap(u64 a);
u64 round_bigseg_cap(u64 n);
int round_tile(const u32 *act_i, const u32 *act_g, const void *key2, int wide, u64 a_grid, const u64 *rp, u32 *SA,
               u32 *ISA, u32 *out_i, u32 *out_g, u64 *out_counter, u32 *err, u32 *big_head, u32 *big_g, u32 *big_size,
               u32 *big_counter, u64 big_cap, u32 *LCP, hipStream_t s, const ActCode &ac, const ExactLcp *ex,
               u32 *seg_i, u32 *seg_g, u64 *seg_ctr, u32 *bseg_h, u32 *bseg_g, u64 bseg_cap, u64 *boff,
               u64 *big_total);
// ISA[PV[e]] = PG[e] for e < m, in the pairs' order.
int launch_isa_scatter(const u32 *PV, const u32 *PG, u64 m, u32 *ISA, hipStream_t s);
// The same for pairs grouped into windows of the target by a radix pass on bits
// [s1, s1 + 8) of PV over input bucketed by PV >> s2 (window of v: ((v >> s1) & 255)
// * nc + (v >> s2), nc buckets, 256 * nc <= 2048 windows): each XCD writes only its
// own windows.  bounds: 256 * nc + 1 u64; ctr: 256 u32.
int launch_scatter_windows(const u32 *PV, const u32 *PG, u64 m, int s1, u32 nc, int s2, u32 *ISA, u64 *bounds,
                           u32 *ctr, hipStream_t s);
// Big groups: comp = (dense group index << kb) | key2, u32 when comp32 (index and
// key2 fit 32 bits together), else u64.
int big_extract(const u32 *act_i, const void *key2, int wide, const u32 *bhead, const u64 *boff, u32 nb, u64 total,
                int kb, int comp32, void *comp, u32 *vals, hipStream_t s);
// Few-key partition of the big groups (pivots.hip): up to BP key2 values frequent
// in a sample of each group become subgroups without sorting; the rest is sorted.
enum { BP = 4 };
// keepkey of a flagged big group whose rank lies in no subgroup of its sorted members
// (key2 values are below 2^34)
constexpr u64 BIG_NOKEY = ~0ull;
struct PivRegions {
    u64 start[BP + 1];   // pivot class b: [start[b], start[b + 1]); start[BP] = pivot members
};
struct BigPivots {
    const u64 *piv = nullptr;    // [nb * BP] each group's pivot keys, ascending
    const u32 *npiv = nullptr;   // [nb]
    const u32 *eq = nullptr;     // [nb * BP] members equal to each pivot
};
int big_emit(WsScratch &ws, const void *comp, int comp32, int kb, const u32 *vals, const u32 *gslot, u64 total, u64 h,
             u32 *SA, u32 *ISA, u32 *out_i, u32 *out_g, u64 *out_counter, u32 *LCP, hipStream_t s, const ActCode &ac,
             const u32 *grank, const u64 *keepkey, const ExactLcp *ex = nullptr, const BigPivots *pv = nullptr);
// For the flagged big groups of a round without pivot classes: grank[q] = the rank
// the members hold (SA[g + 1]), keepkey[q] = the key2 of the subgroup whose slots
// hold it (group q's members are sorted[boff[q], boff[q + 1])).  Pivot rounds get
// both from big_piv_ranks.
int big_keep(const void *sorted, int comp32, const u64 *boff, int kb, u32 nb, const u32 *gslot, u32 gf,
             const u32 *SA, u32 *grank, u64 *keepkey, hipStream_t s);
u64 big_part_tiles(u64 total);
int big_pivots(const void *key2, int wide, const u32 *bhead, const u64 *boff, u32 nb, u64 *piv, u32 *npiv,
               hipStream_t s);
int big_part_count(const void *key2, int wide, const u32 *bhead, const u64 *boff, u32 nb, u64 total, const u64 *piv,
                   const u32 *npiv, u32 *ccnt, hipStream_t s);
int big_part_scatter(const u32 *act_i, const void *key2, int wide, const u32 *bhead, const u64 *boff, u32 nb,
                     u64 total, int kb, int comp32, const u64 *piv, const u32 *npiv, const u32 *cscan, u64 npivot,
                     void *comp, u32 *vals, u32 *pq, u32 *ppv, hipStream_t s);
// Rounds of at most 32 big groups count the partition per class and group, ccnt[(c *
// nb + q) * tiles + t], and write no per-member group index (pq); big_part_groups(nb)
// = nb then, else 0 (counts per class, ccnt[c * tiles + t]).
u32 big_part_groups(u32 nb);
// starts[b] (b = 0..BP) = the first member of class b after the scan of the counts
int big_part_starts(const u32 *cscan, u64 total, u32 nb, u32 *starts, hipStream_t s);
// rk, rv: nb * BP entries (the pivot subgroups' group words and ranks); grank,
// keepkey: as big_keep, for the flagged groups of a pivot round.  The pivot members
// take the next active list's entries [base, base + npivot) (base = its length
// before them; *out_counter is raised by npivot, and err flagged if it was not
// base); the partition wrote their act_i there (ppv = out_i + base).
int big_piv_ranks(const void *rest, int comp32, int kb, u64 nrest, const u32 *pq, const PivRegions &reg,
                  const u64 *piv, const u32 *npiv, const u32 *gslot, u32 gf, u32 *SA, u32 *grank, u64 *keepkey,
                  u32 nb, u32 *eq, u32 *rk, u32 *rv, u64 *out_counter, u64 base, u32 *err, const u32 *cscan,
                  u64 total, hipStream_t s);
int big_piv_emit(const u32 *pq, const PivRegions &reg, const u32 *rk, const u32 *rv, const u32 *gslot, u32 gf,
                 u64 base, u64 h, u32 *ISA, const u32 *out_i, u32 *out_g, u32 *LCP, int cb, const u32 *cscan, u32 nb,
                 u64 total, hipStream_t s);

// ------------------------------------------------------------ SA kernels
int launch_histogram(const uint8_t *T, u64 n, int gsa, u64 *hist256, u32 *gsa_bad, hipStream_t s);
// vals[p] = p, or (p << cb) | code of T[p-1] with cb > 0 (n - 1 < 2^(32 - cb)).
int launch_pack_keys32(const uint8_t *T, u64 n, const uint8_t *lut, int bits, int k, int gsa, int cb, u32 *keys,
                       u32 *vals, hipStream_t s);
// Batched GSA: sid[p] = number of 0 bytes before p (string id); keys get sid in
// their high half.  launch_string_ids needs a ScanScratch for n entries.
int launch_string_ids(const uint8_t *T, u64 n, u32 *sid, ScanScratch &sc, hipStream_t s);
int launch_pack_keys64_sid(const uint8_t *T, u64 n, const uint8_t *lut, int bits, int k, const u32 *sid, u64 *keys,
                           u32 *vals, hipStream_t s);
// Checks that string k starts at offsets[k] (sid[offsets[k]] == k and the byte
// before it is a separator); sets *bad otherwise.
int launch_check_offsets(const uint8_t *T, const u32 *sid, const int64_t *offsets, u32 count, u32 *bad, hipStream_t s);
// Batched plain-text jobs: job k's text is followed by one separator and occupies
// [off[k], off[k+1]); slot off[k] of the batched suffix array is that separator's
// suffix, slots off[k]+1 .. off[k+1]-1 are job k's suffix array.  Packed outputs
// of job k start at off[k] - k.  outLCP / outU may be null.
int launch_batch_extract(const u32 *gSA, const u32 *gLCP, const u32 *sid, const int64_t *off, const u32 *ISA,
                         const uint8_t *T, u64 N, u32 *outSA, u32 *outLCP, uint8_t *outU, hipStream_t s);
// prim[k] = rank of ISA[off[k]] - off[k]  (1-based primary index of job k's BWT);
// N = length of the batched text (selects the ISA word form, isa_rank)
int launch_batch_primary(const u32 *ISA, const int64_t *off, u32 count, u64 N, u32 *prim, hipStream_t s);
// SA_local[j] = SA[j] - offsets[sid[j]] (slot j lies in the same string block).
int launch_localize(const u32 *SA, const u32 *sid, const int64_t *offsets, u64 n, u32 *out, hipStream_t s);
// ------------------------------------------------------------ BWT / PLCP / LCP
struct LcpScratch {
    u32 *queue;       // [cap] positions whose lcp did not finish in the per-thread pass
    u32 *queue_l;     // [cap] lcp matched so far
    u32 *queue_n;     // [2]: queued positions, then those of the very-long queue
    u64 queue_cap;
    u32 *vq_pos;      // [256] very-long queue: positions still matching after the
    u64 *vq_off;      // [256] first pass's steps, their offset reached,
    u64 *vq_res;      // [256] and first mismatch
    void *tile_max;   // [tiles_cap] per-tile max for the prefix-max scan (u32 or u64), then chunk maxima
    u64 tiles_cap;
    // optional: called with 0 / 1 around the launch of the pair-word text-order pass
    void (*hook)(void *user, int phase);
    void *hook_user;
};

int launch_bwt(const uint8_t *T, const u32 *SA, const u32 *ISA, u64 n, uint8_t *U, hipStream_t s);
int launch_bwt_seeded(const uint8_t *T, const u32 *SA, const u32 *ISA, const BwtSeed &seed, u64 n, uint8_t *U,
                      hipStream_t s);
int launch_bwt16(const uint16_t *T, const u32 *SA, const u32 *ISA, u64 n, uint16_t *U, hipStream_t s);
// The same BWT in text order: U[row of suffix i] = T[i-1], reading ISA and T
// sequentially and scattering bytes.
int launch_bwt_scatter(const uint8_t *T, const u32 *ISA, u64 n, uint8_t *U, hipStream_t s);
int launch_phi(const u32 *SA, u64 n, u32 *PHI, hipStream_t s);
// The same Φ through (target, value) pairs bucketed by target (PV, PG: n entries
// each; bcur: 64 cursors, cleared here) and scattered bucket after bucket, so the
// random writes of the device stay inside one 2^wlog-entry stretch of PHI at a time.
int launch_phi_bucketed(const u32 *SA, u64 n, u32 *PHI, u32 *PV, u32 *PG, u32 *bcur, int wlog, hipStream_t s);
// PLCP from either PHI (ISA == nullptr) or ISA+SA (phi computed on the fly).
// marks (SA order, with ISA+SA only): exact lcps or LCP_UNKNOWN | lower bound from
// the construction; position i reads marks[ISA[i]] and only unknown ones compare.
// V64 (n entries) is needed when PLCP[i] + i can reach 2^32 (n > 2^32 - 1 - n/2 is
// not tracked: it is passed whenever n > INT32_MAX); otherwise PLCP is the scratch.
// pairs (n entries, byte texts with marks only): scratch for the per-slot pair words
// that replace the mark, SA[j-1] and T[k-1] gathers; U (optional) is the finished BWT.
// pw/bits (optional): the text's packed code stream, which the pair-word pass
// compares instead of the bytes.
int launch_plcp(LcpScratch &ls, const uint8_t *T, u64 n, int gsa, const u32 *PHI, const u32 *ISA,
                const u32 *SA, u32 *PLCP, u64 *V64, const u32 *marks, hipStream_t s, int sym_bytes = 1,
                u64 *pairs = nullptr, const uint8_t *U = nullptr, const u32 *pw = nullptr, int bits = 0,
                int hint_off = 0);
// LCP[j] = PLCP[SA[j]] where LCP[j] is still an unknown mark.
int launch_lcp_fix(const u32 *PLCP, const u32 *SA, u64 n, u32 *LCP, hipStream_t s);
// BWT auxiliary samples and inverse BWT (unbwt.hip).
int launch_aux_samples(const u32 *ISA, u64 n, u64 r, u64 count, u32 *out, hipStream_t s);
u64 unbwt_count_slots(u64 n);
int launch_unbwt_succ(const uint8_t *U, u64 n, u64 p, u32 *counts, ScanScratch &sc, u32 *S, hipStream_t s);
int launch_unbwt_aux_walk(const void *U, int sym_bytes, const u32 *S, u64 n, u64 p, u64 r, const u64 *rows,
                          u64 chunks, void *out, u32 *err, hipStream_t s);
// 16-bit symbols: sort keys for the stable rank, then the successor array from the
// sorted positions.
int launch_unbwt_succ16(const uint16_t *U, u64 n, u64 p, u32 *keys, u32 *vals, u32 *S, hipStream_t s);
int launch_unbwt_succ_sorted(const u32 *V, u64 n, u64 p, u32 *S, hipStream_t s);
int launch_unbwt_flags(u64 n, u32 *flags, hipStream_t s);
int launch_unbwt_sample_total(const u32 *sid, u64 n, u32 *out, hipStream_t s);
int launch_unbwt_rank(const void *U, int sym_bytes, const u32 *S, const u32 *sid, u64 n, u32 m, u32 *samp_q,
                      u32 *nxt[2], u64 *dist[2], u64 *gap, void *out, u32 *err, hipStream_t s);
int launch_lcp(const u32 *PLCP, const u32 *SA, u64 n, u32 *LCP, hipStream_t s);

// Tiny fill kernel (graph-capture safe replacement for hipMemsetAsync).
int launch_fill_u32(u32 *p, u64 count, u32 value, hipStream_t s);
int launch_fill_u64(u64 *p, u64 count, u64 value, hipStream_t s);
int launch_fill_u8(uint8_t *p, u64 count, uint8_t value, hipStream_t s);
int launch_copy(void *dst, const void *src, u64 bytes, hipStream_t s);
// Integer alphabet (libsais_int / libsais64_long): T is int32 (width 4) or int64.
int launch_pack_int(const u32 *S, u64 n, int bits, int k, int gsa, u32 *keys, u32 *vals, hipStream_t s);
// 16-bit alphabet: symbol histogram (+ GSA validity count) and dense coding.
int launch_hist16(const uint16_t *T, u64 n, int gsa, u64 *hist65536, u32 *bad, hipStream_t s);
int launch_map16(const uint16_t *T, u64 n, const u32 *lut, u32 *S, hipStream_t s);
int launch_max_sym(const void *T, int width, u64 n, u64 *out, hipStream_t s);
int launch_sym_narrow(const void *T, int width, u64 n, u32 *S, hipStream_t s);
int launch_sym_keys64(const void *T, int width, u64 n, u64 *keys, u32 *vals, hipStream_t s);
int launch_run_flags64(const u64 *keys, u64 n, u32 *flags, hipStream_t s);
int launch_dense_scatter64(const u64 *keys, const u32 *vals, const u32 *excl, u64 n, u32 *S, hipStream_t s);
