// radix.hip -- onesweep LSD radix sort for gfx950 (8-bit digits by default; 4 and 11
// bits selectable, the "radix_bits" option), wave64.
//
// One upfront kernel builds the digit histograms of every digit position at once
// (LDS-privatised, one read of the keys).  Each pass is then a single kernel:
//   * a workgroup takes a tile ticket (atomic counter), so every tile it waits on
//     belongs to a workgroup that is already running;
//   * keys are ranked inside the wave with one ballot per digit bit (peer mask of equal
//     digits, popcount below the lane) and per-wave LDS digit counters -- stable by
//     construction, no atomics on the ranking path;
//   * the per-digit tile count is published to a look-back word, and the tile's
//     global digit offsets come from decoupled look-back over preceding tiles;
//   * keys (then values) are reordered in LDS by digit and written out as runs, so
//     consecutive lanes store to consecutive addresses.
#include "engine.h"

#include <string.h>

int g_hist_windows = 1;

namespace {

constexpr int RS_WAVES = RS_THREADS / SAGPU_WAVE;
constexpr int LB_VAL_BITS = 40;
constexpr u64 LB_VAL_MASK = (1ull << LB_VAL_BITS) - 1ull;
constexpr u64 LB_AGG = 1ull, LB_PRE = 2ull;
constexpr int LB_WIN = 2;

__device__ __forceinline__ u64 lb_pack(u32 epoch, u64 flag, u64 v)
{
    return ((u64)epoch << 42) | (flag << LB_VAL_BITS) | v;
}

// Stream bits [bitpos, bitpos + nbits) (nbits <= 32) as an integer.
__device__ __forceinline__ u32 stream_bits(const u32 *__restrict__ pw, u64 bitpos, int nbits)
{
    const u64 w = bitpos >> 5;
    const u64 x = ((u64)pw[w] << 32) | (u64)pw[w + 1];
    return (u32)((x << (bitpos & 31)) >> (64 - nbits));
}
__device__ __forceinline__ u32 gen_key(const KeyGen &g, u64 p) { return stream_bits(g.pw, p * (u64)g.bits, g.kbits); }
__device__ __forceinline__ u32 gen_val(const KeyGen &g, u64 p)
{
    if (!g.cb) return (u32)p;
    const u32 prev = p ? stream_bits(g.pw, (p - 1) * (u64)g.bits, g.bits) : 0u;
    return ((u32)p << g.cb) | prev;
}

template <typename K, bool GEN, int DB = 8>
__global__ __launch_bounds__(256) void k_radix_hist(const K *__restrict__ keys, KeyGen gen, u64 n, int begin_bit,
                                                    int end_bit, int npass, u64 *__restrict__ hist)
{
    // 4 copies per digit position (one with 11-bit digits), picked by lane % 4, so
    // skewed digits (DNA) contend less on the same LDS words
    constexpr int BINS = 1 << DB;
    constexpr int C = DB > 8 ? 1 : 4;
    constexpr int NP = (8 * (int)sizeof(K) + DB - 1) / DB;
    // copy c keeps digit d at column (d + c) % BINS, so the copies of one digit sit in
    // different LDS banks
    __shared__ u32 s_hc[NP * C][BINS];
    for (int i = threadIdx.x; i < NP * C * BINS; i += blockDim.x) (&s_hc[0][0])[i] = 0;
    __syncthreads();
    const u32 copy = threadIdx.x & (C - 1);
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        K k;
        if constexpr (GEN) k = (K)gen_key(gen, i);
        else k = keys[i];
        for (int p = 0; p < npass; p++) {
            const int sh = begin_bit + DB * p;
            const u32 dm = end_bit - sh >= DB ? (u32)(BINS - 1) : ((1u << (end_bit - sh)) - 1u);
            u32 d = (u32)(k >> sh) & dm;
            if constexpr (sizeof(K) == 8) {
                // the big-group sorts of periodic and degenerate texts see long runs
                // of equal keys: a wave whose lanes share the digit adds once
                const u64 act = __ballot(1);
                const u32 d0 = __builtin_amdgcn_readfirstlane(d);
                if (__ballot(d == d0) == act) {
                    if (lane_id() == (u32)(__ffsll((long long)act) - 1)) atomicAdd(&s_hc[p * C][d0], (u32)__popcll(act));
                    continue;
                }
            }
            atomicAdd(&s_hc[p * C + copy][(d + copy) & (u32)(BINS - 1)], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < npass * BINS; i += blockDim.x) {
        const int p = i / BINS, d = i % BINS;
        u32 v = 0;
        for (int c = 0; c < C; c++) v += s_hc[p * C + c][(d + c) & (BINS - 1)];
        if (v) atomicAdd(&hist[i], (u64)v);
    }
}

// Keys generated from the code stream with whole symbols per 8-bit digit (bits in
// {1, 2, 4, 8}): digit p (from the low end) of key(i) is D(i + o_p), the s = 8 / bits
// symbols from position i + o_p, o_p = ksym - s (p + 1).  So every pass's histogram
// is the histogram H of D(j) over j in [0, n), less D(j) for j < min(o_p, n), plus
// min(o_p, n) zeros for the positions past the end (their symbols are code 0): one count per
// position instead of one per position and pass.  k_hist_windows counts H into
// hist[0, 256); k_hist_passes derives the passes' histograms from it.
__global__ __launch_bounds__(256) void k_hist_windows(KeyGen gen, u64 n, u64 *__restrict__ hist)
{
    constexpr int C = 4;   // copies picked by lane % 4, as k_radix_hist; rows of 257 words
    __shared__ u32 s_hc[C][257];   // put the copies of one window value in different banks
    for (int i = threadIdx.x; i < C * 257; i += blockDim.x) (&s_hc[0][0])[i] = 0;
    __syncthreads();
    const u32 copy = threadIdx.x & (C - 1);
    const u64 stride = (u64)gridDim.x * blockDim.x;
    // thread per stream word: the 32 / bits positions starting in word w take their
    // windows from words w and w + 1
    const u32 per = 32u / (u32)gen.bits;
    const u64 words = (n + per - 1) / per;
    for (u64 w = (u64)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) {
        const u64 x = ((u64)gen.pw[w] << 32) | (u64)gen.pw[w + 1];
        const u64 j0 = w * per;
        const u32 m = n - j0 < (u64)per ? (u32)(n - j0) : per;
        for (u32 q = 0; q < m; q++) atomicAdd(&s_hc[copy][(u32)((x << (q * (u32)gen.bits)) >> 56)], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < 256; d += blockDim.x) {
        const u32 v = s_hc[0][d] + s_hc[1][d] + s_hc[2][d] + s_hc[3][d];
        if (v) atomicAdd(&hist[d], (u64)v);
    }
}

// One workgroup of 256 threads: thread d fills bin d of every pass from H.
__global__ __launch_bounds__(256) void k_hist_passes(KeyGen gen, u64 n, int npass, u64 *__restrict__ hist)
{
    const u32 d = threadIdx.x;
    const u64 h = hist[d];
    const int s = 8 / gen.bits, ksym = gen.kbits / gen.bits;
    __syncthreads();   // every thread has read H before pass 0's bins are rewritten
    for (int p = 0; p < npass; p++) {
        const int o = ksym - s * (p + 1);
        u64 c = h;
        for (int j = 0; j < o && (u64)j < n; j++)
            if (stream_bits(gen.pw, (u64)j * (u64)gen.bits, 8) == d) c--;
        if (d == 0) c += (u64)o < n ? (u64)o : n;   // positions past the end among [o, n + o)
        hist[p * 256 + d] = c;
    }
}

// Number of v in [0, n) whose digit (v >> shift) & dmask is d (dmask = 2^b - 1):
// every full period of 2^(shift + b) values holds each digit 2^shift times, and the
// last, partial period holds d's stretch [d << shift, (d + 1) << shift) up to n.
__host__ __device__ inline u64 perm_digit_count(u64 n, int shift, u32 dmask, u32 d)
{
    const int b = __builtin_popcount(dmask);
    const u64 period = 1ull << (shift + b);
    const u64 full = n >> (shift + b), rest = n & (period - 1);
    const u64 lo = (u64)d << shift, hi = (u64)(d + 1) << shift;
    const u64 part = rest > lo ? (rest < hi ? rest : hi) - lo : 0;
    return full * (1ull << shift) + part;
}

// Digit histograms of keys that are a permutation of [0, n), from the closed form.
__global__ __launch_bounds__(256) void k_perm_hist(u64 n, int begin_bit, int end_bit, int npass, int db,
                                                   u64 *__restrict__ hist)
{
    const u32 bins = 1u << db;
    for (u32 i = threadIdx.x; i < (u32)npass * bins; i += blockDim.x) {
        const int p = (int)(i / bins), shift = begin_bit + db * p;
        const u32 dmask = end_bit - shift >= db ? bins - 1u : ((1u << (end_bit - shift)) - 1u);
        const u32 d = i % bins;
        hist[i] = d <= dmask ? perm_digit_count(n, shift, dmask, d) : 0ull;
    }
}

// Exclusive scan of each pass's histogram (bins per pass) into bucket offsets at
// hist + RADIX_HIST_WORDS.  Thread t takes bins [t * per, (t + 1) * per).
__global__ __launch_bounds__(256) void k_radix_offsets(u64 *hist, int npass, int bins)
{
    __shared__ u64 s_w[RS_WAVES];
    const int per = bins > 256 ? bins / 256 : 1;   // at most 8 (2048 bins)
    for (int p = 0; p < npass; p++) {
        const int d0 = (int)threadIdx.x * per;
        u64 loc[8];
        u64 v = 0;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            loc[q] = (q < per && d0 + q < bins) ? hist[p * bins + d0 + q] : 0ull;
            v += loc[q];
        }
        u64 inc = wave_incl_sum64(v);
        if (lane_id() == SAGPU_WAVE - 1) s_w[threadIdx.x / SAGPU_WAVE] = inc;
        __syncthreads();
        u64 carry = 0;
        for (u32 w = 0; w < threadIdx.x / SAGPU_WAVE; w++) carry += s_w[w];
        u64 run = carry + inc - v;
#pragma unroll
        for (int q = 0; q < 8; q++)
            if (q < per && d0 + q < bins) {
                hist[RADIX_HIST_WORDS + p * bins + d0 + q] = run;
                run += loc[q];
            }
        __syncthreads();
    }
}

// One radix pass over a tile of RS_TILE keys.  LOOKBACK = true: onesweep (tile
// ticket, aggregate publish, decoupled look-back).  LOOKBACK = false: the tile's
// global digit offsets were produced by an upsweep + scan (reduce-then-scan).
// GEN: the input keys and values are generated (KeyGen) instead of read.
template <typename K, bool HAS_VALS, bool LOOKBACK, int ITEMS, bool GEN = false, int NT = RS_THREADS, int DB = 8>
__global__ __launch_bounds__(NT) void k_radix_pass(const K *__restrict__ kin, const u32 *__restrict__ vin,
                                                           K *__restrict__ kout, u32 *__restrict__ vout, u64 n,
                                                           int shift, u32 dmask, const u64 *__restrict__ offsets,
                                                           u64 *__restrict__ lookback, u32 *__restrict__ tile_ctr,
                                                           u32 epoch, const u32 *__restrict__ scanned, u64 tiles,
                                                           KeyGen gen)
{
    // NT threads (256 or 512) per workgroup; BINS digits, DPT per thread in the digit
    // phase (thread t takes digits [t * DPT, (t + 1) * DPT)).
    constexpr int NW = NT / SAGPU_WAVE;
    constexpr int TILE_ = NT * ITEMS;
    constexpr int WAVE_ITEMS_ = SAGPU_WAVE * ITEMS;
    constexpr int BINS = 1 << DB;
    constexpr int DPT = BINS > NT ? BINS / NT : 1;
    // digits kept per item for the value phase: 8 bits each, 16 with 11-bit digits
    constexpr int DPB = DB > 8 ? 16 : 8, DPW = 32 / DPB;
    __shared__ u32 s_tile;
    __shared__ u32 s_wcnt[NW][BINS];
    __shared__ u32 s_dstart[BINS];
    __shared__ u64 s_gbase[BINS];
    __shared__ u32 s_wsum[NW];
    __shared__ K s_keys[TILE_];

    const u32 tid = threadIdx.x, w = tid / SAGPU_WAVE, lane = lane_id();
    if (LOOKBACK) {
        if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
    }
    for (u32 i = tid; i < (u32)NW * BINS; i += NT) (&s_wcnt[0][0])[i] = 0;
    __syncthreads();

    const u64 tile = LOOKBACK ? (u64)s_tile : (u64)blockIdx.x;
    const u64 wbase = tile * TILE_ + (u64)w * WAVE_ITEMS_;

    // Values are loaded only for the value phase, after the keys have left the
    // registers: fewer live registers per wave, more waves per SIMD.
    //
    // Positions past n (only in the last tile) take the largest key.  They rank
    // after every real key of the largest digit (position order is the tie-break),
    // so they occupy the tile's last LDS slots [tile_n, TILE_), which are never
    // written out; the only count they inflate is that digit's count in the last
    // tile's own look-back entry, which no later tile reads.
    K key[ITEMS];
    u32 rank[ITEMS];
    u32 prevc[GEN ? ITEMS / 4 : 1];   // GEN: code of T[pos-1] per item, 8 bits each
    const bool full = (tile + 1) * (u64)TILE_ <= n;
    if constexpr (GEN) {
        // The tile's stretch of the code stream (from one symbol before the tile)
        // goes to LDS with coalesced loads; keys and preceding codes are cut from
        // it.  The stretch aliases s_keys, which is first written after ranking.
        static_assert(sizeof(K) == 4, "generated keys are 32-bit");
        u32 *s_st = reinterpret_cast<u32 *>(s_keys);
        const u64 bits = (u64)gen.bits;
        const u64 bit0 = tile * (u64)TILE_ * bits;
        const u64 word0 = bit0 >= bits ? (bit0 - bits) >> 5 : 0;
        const u32 nw = (u32)(((bit0 + (u64)TILE_ * bits + 64) >> 5) - word0 + 1);
        for (u32 i = tid; i < nw; i += NT) s_st[i] = word0 + i < gen.words ? gen.pw[word0 + i] : 0u;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < ITEMS / 4; q++) prevc[q] = 0;
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            const u64 pos = wbase + (u64)j * SAGPU_WAVE + lane;
            const u64 lb = pos * bits - word0 * 32;
            const u32 lw = (u32)(lb >> 5), off = (u32)(lb & 31);
            const u64 x = ((u64)s_st[lw] << 32) | (u64)s_st[lw + 1];
            key[j] = pos < n ? (K)((x << off) >> (64 - gen.kbits)) : ~(K)0;
            if (gen.cb && pos > 0) {
                const u64 pb = lb - bits;
                const u32 pw0 = (u32)(pb >> 5), poff = (u32)(pb & 31);
                const u64 y = ((u64)s_st[pw0] << 32) | (u64)s_st[pw0 + 1];
                prevc[j / 4] |= (u32)((y << poff) >> (64 - bits)) << (8 * (j % 4));
            }
            // keep the scheduler from hoisting every item's LDS reads and 64-bit
            // temporaries (they would double the registers and halve occupancy)
            asm volatile("" : "+v"(key[j]));
            if (j % 4 == 3) __builtin_amdgcn_sched_barrier(0);
        }
    } else if (full) {
#pragma unroll
        for (int j = 0; j < ITEMS; j++) key[j] = kin[wbase + (u64)j * SAGPU_WAVE + lane];
    } else {
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            u64 pos = wbase + (u64)j * SAGPU_WAVE + lane;
            key[j] = pos < n ? kin[pos] : ~(K)0;
        }
    }

    // Wave-level stable ranking: order inside the wave is (item j, lane).
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        u32 d = (u32)(key[j] >> shift) & dmask;
        u64 m = ~0ull;
#pragma unroll
        for (int b = 0; b < DB; b++) {
            bool bit = (d >> b) & 1u;
            u64 bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        u32 below = __popcll(m & lanemask_lt());
        u32 old = s_wcnt[w][d];
        if (below == 0) s_wcnt[w][d] = old + (u32)__popcll(m);
        // rank inside the wave (< 2^16) with the digit above it: one live register
        // per item until the reorder instead of three (rank, digit, LDS address)
        rank[j] = (old + below) | (d << 16);
        // Materialise the rank here: otherwise the compiler sinks every item's rank
        // arithmetic to its first use after the look-back and keeps digit, old count
        // and peer mask of all items live (four registers per item instead of one).
        asm volatile("" : "+v"(rank[j]));
    }
    __syncthreads();

    // Per digit: wave offsets, tile count, global offset.
    u32 run[DPT];
    u32 tsum = 0;
    const bool dthread = tid * DPT < (u32)BINS;
#pragma unroll
    for (int qd = 0; qd < DPT; qd++) {
        run[qd] = 0;
        if (dthread) {
            const u32 d = tid * DPT + qd;
            u32 r = 0;
#pragma unroll
            for (int q = 0; q < NW; q++) {
                const u32 c = s_wcnt[q][d];
                s_wcnt[q][d] = r;
                r += c;
            }
            run[qd] = r;
            tsum += r;
        }
    }
    const u32 inc = wave_incl_sum(tsum);
    if (lane == SAGPU_WAVE - 1) s_wsum[w] = inc;
    if (dthread) {
        // every digit's aggregate is published before any look-back starts, so the
        // tiles after this one never wait on a digit this thread has not reached
        if (LOOKBACK) {
#pragma unroll
            for (int qd = 0; qd < DPT; qd++)
                st_release_u64(lookback + tile * BINS + tid * DPT + qd,
                               lb_pack(epoch, tile == 0 ? LB_PRE : LB_AGG, run[qd]));
        }
#pragma unroll
        for (int qd = 0; qd < DPT; qd++) {
            const u32 d = tid * DPT + qd;
            u64 gbase;
            if (LOOKBACK) {
                u64 *my = lookback + tile * BINS + d;
                u64 excl = 0;
                if (tile > 0) {
                    // Windowed look-back: the loads for up to LB_WIN predecessors are
                    // issued together so their latencies overlap; entries are consumed
                    // from the nearest one until an inclusive prefix is found.
                    long long t = (long long)tile - 1;
                    bool found = false;
                    while (!found) {
                        const int cnt = t + 1 < LB_WIN ? (int)(t + 1) : LB_WIN;
                        u64 v[LB_WIN];
#pragma unroll
                        for (int q = 0; q < LB_WIN; q++)
                            if (q < cnt) v[q] = ld_acquire_u64(lookback + (u64)(t - q) * BINS + d);
#pragma unroll
                        for (int q = 0; q < LB_WIN; q++) {
                            if (q >= cnt || found) continue;
                            while ((u32)(v[q] >> 42) != epoch || ((v[q] >> LB_VAL_BITS) & 3ull) == 0) {
                                __builtin_amdgcn_s_sleep(1);
                                v[q] = ld_acquire_u64(lookback + (u64)(t - q) * BINS + d);
                            }
                            excl += v[q] & LB_VAL_MASK;
                            if (((v[q] >> LB_VAL_BITS) & 3ull) == LB_PRE) found = true;
                        }
                        t -= cnt;
                    }
                    st_release_u64(my, lb_pack(epoch, LB_PRE, excl + run[qd]));
                }
                gbase = offsets[d] + excl;
            } else {
                gbase = scanned[(u64)d * tiles + tile];
            }
            s_gbase[d] = gbase;
        }
    }
    __syncthreads();
    if (dthread) {
        u32 carry = 0;
        for (u32 q = 0; q < w; q++) carry += s_wsum[q];
        u32 st = carry + inc - tsum;
#pragma unroll
        for (int qd = 0; qd < DPT; qd++) {
            s_dstart[tid * DPT + qd] = st;
            st += run[qd];
        }
    }
    __syncthreads();

    // Reorder keys in LDS by digit (past-the-end keys land in [tile_n, TILE_)).
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const u32 d = rank[j] >> 16;
        const u32 lp = s_dstart[d] + s_wcnt[w][d] + (rank[j] & 0xFFFFu);
        s_keys[lp] = key[j];
        rank[j] = lp;
    }
    __syncthreads();
    const u64 tbase = tile * TILE_;
    const u32 tile_n = (u32)((n - tbase) < (u64)TILE_ ? (n - tbase) : (u64)TILE_);
    // Digits of the staged keys are recomputed and kept packed in registers for the
    // value phase, so no LDS digit array is needed (one more workgroup per CU fits).
    u32 dpack[(ITEMS + DPW - 1) / DPW];
#pragma unroll
    for (int q = 0; q < (ITEMS + DPW - 1) / DPW; q++) dpack[q] = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const u32 idx = tid + (u32)j * NT;
        if (idx < tile_n) {
            const K kk = s_keys[idx];
            const u32 d = (u32)(kk >> shift) & dmask;
            dpack[j / DPW] |= d << (DPB * (j % DPW));
            kout[s_gbase[d] + (idx - s_dstart[d])] = kk;
        }
        // bound the LDS loads the scheduler hoists (each item holds four registers)
        if (j % 8 == 7) __builtin_amdgcn_sched_barrier(0);
    }
    if (HAS_VALS) {
        __syncthreads();
        u32 *s_vals = reinterpret_cast<u32 *>(s_keys);
        if constexpr (GEN) {
            // positions fit 32 bits here (values carry them); recomputed per item
            const u32 wb = (u32)wbase, nn = (u32)(n - 1);
#pragma unroll
            for (int j = 0; j < ITEMS; j++) {
                const u32 pos = wb + (u32)j * SAGPU_WAVE + lane;
                const u32 pc = (prevc[j / 4] >> (8 * (j % 4))) & 255u;
                if (pos <= nn) s_vals[rank[j]] = (pos << gen.cb) | pc;
            }
        } else if (full) {
#pragma unroll
            for (int j = 0; j < ITEMS; j++) s_vals[rank[j]] = vin[wbase + (u64)j * SAGPU_WAVE + lane];
        } else {
#pragma unroll
            for (int j = 0; j < ITEMS; j++) {
                u64 pos = wbase + (u64)j * SAGPU_WAVE + lane;
                if (pos < n) s_vals[rank[j]] = vin[pos];
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            const u32 idx = tid + (u32)j * NT;
            if (idx < tile_n) {
                const u32 d = (dpack[j / DPW] >> (DPB * (j % DPW))) & ((1u << DPB) - 1u);
                vout[s_gbase[d] + (idx - s_dstart[d])] = s_vals[idx];
            }
            if (j % 8 == 7) __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// Reduce-then-scan upsweep: per-tile digit counts, digit-major (counts[d*tiles+t]).
template <typename K, int ITEMS>
__global__ __launch_bounds__(RS_THREADS) void k_rts_upsweep(const K *__restrict__ kin, u64 n, int shift,
                                                            u32 dmask, u32 *__restrict__ counts, u64 tiles)
{
    __shared__ u32 s_h[RS_WAVES][256];
    const u32 tid = threadIdx.x, w = tid / SAGPU_WAVE;
#pragma unroll
    for (int q = 0; q < RS_WAVES; q++) s_h[q][tid] = 0;
    __syncthreads();
    const u64 base = (u64)blockIdx.x * (RS_THREADS * ITEMS);
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        u64 pos = base + (u64)j * RS_THREADS + tid;
        if (pos < n) atomicAdd(&s_h[w][(u32)(kin[pos] >> shift) & dmask], 1u);
    }
    __syncthreads();
    u32 c = 0;
#pragma unroll
    for (int q = 0; q < RS_WAVES; q++) c += s_h[q][tid];
    counts[(u64)tid * tiles + blockIdx.x] = c;
}

// A sort whose passes were all skipped still returns its input in (k0, v0).
__global__ __launch_bounds__(256) void k_gen_write(KeyGen gen, u64 n, u32 *__restrict__ keys, u32 *__restrict__ vals)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 p = (u64)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        keys[p] = gen_key(gen, p);
        vals[p] = gen_val(gen, p);
    }
}

// Word w of the MSB-first code stream: codes of the symbols overlapping stream bits
// [32w, 32w + 32); symbols at or past n are code 0.
__global__ __launch_bounds__(256) void k_pack_stream(const uint8_t *__restrict__ T, u64 n,
                                                     const uint8_t *__restrict__ lut, int bits, u32 *__restrict__ pw,
                                                     u64 words)
{
    __shared__ uint8_t s_lut[256];
    s_lut[threadIdx.x] = lut[threadIdx.x];
    __syncthreads();
    const u64 stride = (u64)gridDim.x * blockDim.x;
    if (32 % bits == 0) {
        // whole symbols per word: 32 / bits consecutive bytes, read as 4-byte groups
        const int per = 32 / bits;
        for (u64 w = (u64)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) {
            const u64 q0 = w * (u64)per;
            u32 acc = 0;
            for (int g = 0; g < per; g += 4) {
                const u64 q = q0 + g;
                u32 b4;
                if (q + 4 <= n) {
                    b4 = *reinterpret_cast<const u32 *>(T + q);   // q is a multiple of 4
                } else {
                    b4 = 0;
                    for (int t = 0; t < 4; t++)
                        if (q + t < n) b4 |= (u32)T[q + t] << (8 * t);
                }
                for (int t = 0; t < 4 && g + t < per; t++) {
                    const u32 c = q + t < n ? (u32)s_lut[(b4 >> (8 * t)) & 255u] : 0u;
                    acc = (acc << bits) | c;
                }
            }
            pw[w] = acc;
        }
        return;
    }
    for (u64 w = (u64)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) {
        const u64 b0 = w * 32, b1 = b0 + 32;
        u64 acc = 0;   // stream bits [first symbol start, ...) accumulated MSB-first
        const u64 s0 = b0 / (u64)bits, s1 = (b1 - 1) / (u64)bits;
        for (u64 q = s0; q <= s1; q++) acc = (acc << bits) | (q < n ? (u64)s_lut[T[q]] : 0ull);
        // acc holds bits [s0*bits, (s1+1)*bits); keep [b0, b1)
        const u64 end = (s1 + 1) * (u64)bits;
        pw[w] = (u32)(acc >> (end - b1));
    }
}

template <typename K, int DB>
int radix_sort_db(RadixScratch &rs, K *k0, u32 *v0, K *k1, u32 *v1, u64 n, int begin_bit, int end_bit,
                  hipStream_t s, int *result_in_alt, int *passes_run)
{
    constexpr int BINS = 1 << DB;
    const int npass = (end_bit - begin_bit + DB - 1) / DB;
    // 4- and 11-bit digits: onesweep only, 256-thread tiles
    const int variant = DB == 8 ? rs.variant : RADIX_ONESWEEP;
    // short inputs (one string of a batch, a served job) take 2048-key tiles: more
    // workgroups at once instead of a few long ones (a 16384-key tile takes about
    // 25 us, so a 20 K-key sort would run its passes on two workgroups)
    const bool small = n < rs.small_n;
    const int items = small ? 8 : rs.items == 32 ? 32 : rs.items == 8 ? 8 : 16;
    const u64 tile_keys = (u64)RS_THREADS * items;
    const u64 tiles = (n + tile_keys - 1) / tile_keys;
    if (tiles * BINS > rs.lb_words || npass > RADIX_MAX_PASSES || (u64)npass * BINS > RADIX_HIST_WORDS ||
        (variant == RADIX_REDUCE_THEN_SCAN && tiles * 256 > rs.counts_cap)) {
        sagpu_set_error("radix scratch too small", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }

    const KeyGen *gen = sizeof(K) == 4 ? rs.gen : nullptr;
    const KeyGen g0 = gen ? *gen : KeyGen{};
    if (gen && variant == RADIX_REDUCE_THEN_SCAN) {
        // the upsweep reads its keys: write the generated input out first
        if constexpr (sizeof(K) == 4) {
            hipLaunchKernelGGL(k_gen_write, dim3((unsigned)((n + 2047) / 2048 < 65535 ? (n + 2047) / 2048 : 65535)),
                               dim3(256), 0, s, g0, n, (u32 *)k0, v0);
            SAGPU_CHECK(hipGetLastError());
        }
        gen = nullptr;
    }
    // keys known to be a permutation of [0, n): closed-form histograms (u32 keys)
    const bool perm = rs.perm_keys && sizeof(K) == 4 && !gen && n < (1ull << 32);
    if (!perm) SAGPU_TRY(launch_fill_u64(rs.hist, (u64)npass * BINS, 0, s));
    {
        // 64 keys per thread; 8 for short inputs, which would otherwise leave all
        // but a few workgroups idle
        const u64 per = small ? 256 * 8 : 256 * 64;
        int blocks = (int)((n + per - 1) / per);
        if (blocks > 4096) blocks = 4096;
        // generated keys cut into whole-symbol digits: one window histogram for all passes
        const bool windows = gen && DB == 8 && begin_bit == 0 && end_bit == g0.kbits && g0.kbits % 8 == 0 &&
                             8 % g0.bits == 0 && g_hist_windows;
        if (perm) {
            hipLaunchKernelGGL(k_perm_hist, dim3(1), dim3(256), 0, s, n, begin_bit, end_bit, npass, DB, rs.hist);
        } else if (windows) {
            hipLaunchKernelGGL(k_hist_windows, dim3(blocks), dim3(256), 0, s, g0, n, rs.hist);
            hipLaunchKernelGGL(k_hist_passes, dim3(1), dim3(256), 0, s, g0, n, npass, rs.hist);
        } else if (gen)
            hipLaunchKernelGGL((k_radix_hist<K, true, DB>), dim3(blocks), dim3(256), 0, s, k0, g0, n, begin_bit,
                               end_bit, npass, rs.hist);
        else
            hipLaunchKernelGGL((k_radix_hist<K, false, DB>), dim3(blocks), dim3(256), 0, s, k0, g0, n, begin_bit,
                               end_bit, npass, rs.hist);
        hipLaunchKernelGGL(k_radix_offsets, dim3(1), dim3(256), 0, s, rs.hist, npass, BINS);
        SAGPU_CHECK(hipGetLastError());
    }
    // Histograms decide which passes can be skipped (all keys share the digit).  A
    // short input runs every pass instead: a pass over it costs less than the wait
    // for the histograms on the host.
    if (perm) {
        for (int p = 0; p < npass; p++) {
            const int shift = begin_bit + DB * p;
            const u32 dmask = end_bit - shift >= DB ? (u32)(BINS - 1) : ((1u << (end_bit - shift)) - 1u);
            for (u32 d = 0; d < (u32)BINS; d++)
                rs.hist_host[p * BINS + d] = d <= dmask ? perm_digit_count(n, shift, dmask, d) : 0ull;
        }
    } else if (!small) {
        SAGPU_CHECK(hipMemcpyAsync(rs.hist_host, rs.hist, (u64)npass * BINS * sizeof(u64), hipMemcpyDeviceToHost, s));
        SAGPU_CHECK(hipStreamSynchronize(s));
    }

    SAGPU_TRY(launch_fill_u32(rs.tile_ctr, RADIX_MAX_PASSES, 0, s));
    K *ki = k0, *ko = k1;
    u32 *vi = v0, *vo = v1;
    int alt = 0, ran = 0;
    for (int p = 0; p < npass; p++) {
        bool trivial = false;
        for (int d = 0; d < BINS && !small; d++)
            if (rs.hist_host[p * BINS + d] == n) trivial = true;
        if (trivial) continue;
        rs.epoch = (rs.epoch + 1) & ((1u << 22) - 1);
        if (rs.epoch == 0 && variant != RADIX_REDUCE_THEN_SCAN) {
            SAGPU_TRY(launch_fill_u64(rs.lookback, rs.lb_words, 0, s));
            rs.epoch = 1;
        }
        const int shift = begin_bit + DB * p;
        const u32 dmask = end_bit - shift >= DB ? (u32)(BINS - 1) : ((1u << (end_bit - shift)) - 1u);
        if constexpr (DB == 8) {
            if (variant == RADIX_REDUCE_THEN_SCAN) {
                if (items == 8)
                    hipLaunchKernelGGL((k_rts_upsweep<K, 8>), dim3((unsigned)tiles), dim3(RS_THREADS), 0, s, ki, n,
                                       shift, dmask, rs.counts, tiles);
                else if (items == 32)
                    hipLaunchKernelGGL((k_rts_upsweep<K, 32>), dim3((unsigned)tiles), dim3(RS_THREADS), 0, s, ki, n,
                                       shift, dmask, rs.counts, tiles);
                else
                    hipLaunchKernelGGL((k_rts_upsweep<K, 16>), dim3((unsigned)tiles), dim3(RS_THREADS), 0, s, ki, n,
                                       shift, dmask, rs.counts, tiles);
                SAGPU_TRY(scan_exclusive_u32(rs.scan, rs.counts, tiles * 256, s));
            }
        }
        const bool gen_now = gen && ran == 0;
        rs.gen_pass = gen_now;
        if (rs.hook) rs.hook(rs.hook_user, 0);
        const bool lb = variant != RADIX_REDUCE_THEN_SCAN;
        const u64 *offs = lb ? rs.hist + RADIX_HIST_WORDS + (u64)p * BINS : nullptr;
        u64 *lbk = lb ? rs.lookback : nullptr;
        u32 *tctr = lb ? rs.tile_ctr + p : nullptr;
        const u32 *cnts = lb ? nullptr : rs.counts;
        // onesweep passes may use 512-thread workgroups (16384-key tiles with 32
        // items: twice as long digit runs per tile, same waves per CU); 8-bit digits
        const int nt = (!small && DB == 8 && sizeof(K) == 4 && lb && rs.threads == 512 && items == 32) ? 512
                                                                                                          : RS_THREADS;
        const u64 ptiles = (n + (u64)nt * items - 1) / ((u64)nt * items);
#define RS_LAUNCH(HV, LB, IT)                                                                                     \
    hipLaunchKernelGGL((k_radix_pass<K, HV, LB, IT, false, RS_THREADS, DB>), dim3((unsigned)tiles), dim3(RS_THREADS), \
                       0, s, ki, vi, ko, vo, n, shift, dmask, offs, lbk, tctr, rs.epoch, cnts, tiles, g0)
#define RS_LAUNCH512(HV, GN)                                                                                      \
    hipLaunchKernelGGL((k_radix_pass<K, HV, true, 32, GN, 512>), dim3((unsigned)ptiles), dim3(512), 0, s, ki, vi, ko, \
                       vo, n, shift, dmask, offs, lbk, tctr, rs.epoch, cnts, ptiles, g0)
#define RS_LAUNCH_GEN(IT)                                                                                          \
    hipLaunchKernelGGL((k_radix_pass<K, true, true, IT, true, RS_THREADS, DB>), dim3((unsigned)gtiles),              \
                       dim3(RS_THREADS), 0, s, ki, vi, ko, vo, n, shift, dmask, offs, lbk, tctr, rs.epoch, cnts, gtiles, \
                       g0)
        if (nt == 512) {
            if constexpr (sizeof(K) == 4 && DB == 8) {   // 16384 u32 keys: 64 KB of LDS staging
                if (gen_now) RS_LAUNCH512(true, true);
                else if (v0) RS_LAUNCH512(true, false);
                else RS_LAUNCH512(false, false);
            }
        } else if (gen_now) {
            if constexpr (sizeof(K) == 4) {
                const int git = items;
                const u64 gtiles = (n + (u64)RS_THREADS * git - 1) / ((u64)RS_THREADS * git);
                if (git == 8) RS_LAUNCH_GEN(8); else if (git == 16) RS_LAUNCH_GEN(16); else RS_LAUNCH_GEN(32);
            }
        } else if (lb) {
            if (v0) { if (items == 32) RS_LAUNCH(true, true, 32); else if (items == 8) RS_LAUNCH(true, true, 8); else RS_LAUNCH(true, true, 16); }
            else { if (items == 32) RS_LAUNCH(false, true, 32); else if (items == 8) RS_LAUNCH(false, true, 8); else RS_LAUNCH(false, true, 16); }
        } else {
            if constexpr (DB == 8) {
                if (v0) { if (items == 32) RS_LAUNCH(true, false, 32); else if (items == 8) RS_LAUNCH(true, false, 8); else RS_LAUNCH(true, false, 16); }
                else { if (items == 32) RS_LAUNCH(false, false, 32); else if (items == 8) RS_LAUNCH(false, false, 8); else RS_LAUNCH(false, false, 16); }
            }
        }
#undef RS_LAUNCH
#undef RS_LAUNCH512
#undef RS_LAUNCH_GEN
        SAGPU_CHECK(hipGetLastError());
        if (rs.hook) rs.hook(rs.hook_user, 1);
        rs.gen_pass = 0;
        K *tk = ki; ki = ko; ko = tk;
        u32 *tv = vi; vi = vo; vo = tv;
        alt ^= 1;
        ran++;
    }
    if (gen && ran == 0) {
        if constexpr (sizeof(K) == 4) {
            hipLaunchKernelGGL(k_gen_write, dim3((unsigned)((n + 2047) / 2048 < 65535 ? (n + 2047) / 2048 : 65535)),
                               dim3(256), 0, s, g0, n, (u32 *)k0, v0);
            SAGPU_CHECK(hipGetLastError());
        }
    }
    *result_in_alt = alt;
    if (passes_run) *passes_run = ran;
    return SAGPU_OK;
}

template <typename K>
int radix_sort_impl(RadixScratch &rs, K *k0, u32 *v0, K *k1, u32 *v1, u64 n, int begin_bit, int end_bit,
                    hipStream_t s, int *result_in_alt, int *passes_run)
{
    *result_in_alt = 0;
    if (passes_run) *passes_run = 0;
    if (n <= 1 || end_bit <= begin_bit) return SAGPU_OK;
    if (rs.bits == 4) return radix_sort_db<K, 4>(rs, k0, v0, k1, v1, n, begin_bit, end_bit, s, result_in_alt, passes_run);
    if (rs.bits == 11)
        return radix_sort_db<K, 11>(rs, k0, v0, k1, v1, n, begin_bit, end_bit, s, result_in_alt, passes_run);
    return radix_sort_db<K, 8>(rs, k0, v0, k1, v1, n, begin_bit, end_bit, s, result_in_alt, passes_run);
}

} // namespace

// two words cut keys past the end; a third lets 64-bit windows start at any symbol
u64 pack_stream_words(u64 n, int bits) { return (n * (u64)bits + 31) / 32 + 3; }

int launch_pack_stream(const uint8_t *T, u64 n, const uint8_t *lut, int bits, u32 *pw, hipStream_t s)
{
    const u64 words = pack_stream_words(n, bits);
    u64 blocks = (words + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(k_pack_stream, dim3((unsigned)blocks), dim3(256), 0, s, T, n, lut, bits, pw, words);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int radix_sort_u64(RadixScratch &rs, u64 *k0, u32 *v0, u64 *k1, u32 *v1, u64 n, int begin_bit, int end_bit,
                   hipStream_t s, int *result_in_alt, int *passes_run)
{
    return radix_sort_impl<u64>(rs, k0, v0, k1, v1, n, begin_bit, end_bit, s, result_in_alt, passes_run);
}

int radix_sort_u32(RadixScratch &rs, u32 *k0, u32 *v0, u32 *k1, u32 *v1, u64 n, int begin_bit, int end_bit,
                   hipStream_t s, int *result_in_alt, int *passes_run)
{
    return radix_sort_impl<u32>(rs, k0, v0, k1, v1, n, begin_bit, end_bit, s, result_in_alt, passes_run);
}
