// rounds.hip -- grouping after the initial sort and the prefix-doubling rounds.
//
// State between rounds is the active list: act_i[k] (suffix index) and act_g[k]
// (SA slot of the first element of its group), with every group contiguous.  The
// list order across groups is free, so kernels reserve output ranges with one
// atomic per workgroup.  ISA[i] is the rank of suffix i (its group's first slot,
// the final position once it is alone).
//
// Round with offset h (all group members share their first h symbols):
//   k_round_gather   key2[k] = rank of i+h (or a length code past the end)
//   k_round_tile     groups of <= RT_SMALL members whose first element lies in a
//                    512-wide stretch of the list (tunable): ranked by key2 inside LDS,
//                    regrouped, ISA/SA written, survivors emitted.  Group extents
//                    come from a window of act_g; a big group's last list index is
//                    written to SA[g] (its SA slots are free until its suffixes are
//                    placed; the last index, not the end, so 2^32 members still fit)
//   big path         larger groups: extracted, sorted by (group, key2) with the
//                    radix sort, regrouped with a segmented scan (BigF)
// A pair of neighbours separated in a round has lcp in [h, 2h); the slot of the
// second one gets LCP_UNKNOWN | h so the LCP pass can start comparing at h.
#include "engine.h"

namespace {

// ---------------------------------------------------- warp-striped segmented scan
// A tile is split into four wave spans, each walked as R rows of 64 consecutive
// elements (R = 8 or 16: tiles of 2048 or 4096 elements; 2 for short texts), so
// every load and store of a row is coalesced.  The scan carries one sum (output index of active elements) and two
// max-scans (positions of the latest group starts).  Functors are called by every
// lane of a row (lane l holds element row_base + l), also past the end, so they can
// take neighbour values from the adjacent lanes.

struct Agg {
    u64 s;        // active count (u64: a text of 2^32 suffixes may have them all active)
    u32 m1, m2;
};

__device__ __forceinline__ u32 umax(u32 a, u32 b) { return a > b ? a : b; }

constexpr int WS_THREADS = 256, WS_ROWS = 16, WS_WAVES = 4;

template <class F, int R>
__global__ __launch_bounds__(WS_THREADS) void k_ws_reduce(F f, Agg *__restrict__ tagg)
{
    constexpr int SPAN = R * SAGPU_WAVE;
    __shared__ Agg s_w[WS_WAVES];
    const u32 w = threadIdx.x / SAGPU_WAVE, lane = lane_id();
    const u64 wbase = (u64)blockIdx.x * (WS_WAVES * SPAN) + (u64)w * SPAN;
    u64 s = 0;
    u32 m1 = 0, m2 = 0;
    for (int r = 0; r < R; r++) {
        u64 e = wbase + (u64)r * SAGPU_WAVE + lane;
        bool act;
        u32 v1, v2;
        f.scan_inputs(e, act, v1, v2);   // every lane calls (rows may shuffle); e >= n yields nothing
        s += act;
        m1 = umax(m1, v1);
        m2 = umax(m2, v2);
    }
    for (int d = 32; d >= 1; d >>= 1) {
        s += __shfl_xor(s, d, SAGPU_WAVE);
        m1 = umax(m1, __shfl_xor(m1, d, SAGPU_WAVE));
        m2 = umax(m2, __shfl_xor(m2, d, SAGPU_WAVE));
    }
    if (lane == 0) s_w[w] = Agg{s, m1, m2};
    __syncthreads();
    if (threadIdx.x == 0) {
        Agg t{0, 0, 0};
        for (int q = 0; q < WS_WAVES; q++) {
            t.s += s_w[q].s;
            t.m1 = umax(t.m1, s_w[q].m1);
            t.m2 = umax(t.m2, s_w[q].m2);
        }
        tagg[blockIdx.x] = t;
    }
}

// Exclusive scan of tile aggregates, in steps of 8192 tiles (SC_K consecutive tiles
// per thread of a 1024-thread workgroup).  A step's tiles pass through LDS: loaded and
// stored with consecutive lanes on consecutive tiles, and read by their owner thread
// from rows of SC_K + 1 (padding against bank conflicts).  Up to one step, one
// workgroup scans them (k_ws_scan_tiles); beyond, every step is a chunk of its own:
// k_ws_chunk_reduce combines each chunk, k_ws_scan_tiles scans the chunk aggregates,
// and k_ws_chunk_apply scans each chunk from its prefix, all chunks at once (the
// 131072 tiles of a 2^28-position grouping took 0.145 ms in one workgroup).
constexpr int SC_K = 8, SC_STEP = 1024 * SC_K;

// One step: tiles [b0, b0 + SC_STEP) of tagg replaced by their exclusive prefixes
// after `carry`; returns the inclusive aggregate through the step's last tile.
__device__ Agg scan_step(Agg *tagg, u64 tiles, u64 b0, Agg carry, Agg *s_t, Agg *s_w, Agg *s_end)
{
    const u32 tid = threadIdx.x, w = tid / SAGPU_WAVE;
#pragma unroll
    for (int q = 0; q < SC_K; q++) {
        const u32 j = (u32)q * 1024 + tid;   // tile b0 + j, owned by thread j / SC_K
        s_t[(j / SC_K) * (SC_K + 1) + j % SC_K] = b0 + j < tiles ? tagg[b0 + j] : Agg{0, 0, 0};
    }
    __syncthreads();
    Agg v[SC_K];
    Agg t{0, 0, 0};   // this thread's SC_K tiles combined
#pragma unroll
    for (int q = 0; q < SC_K; q++) {
        v[q] = s_t[tid * (SC_K + 1) + q];
        t.s += v[q].s;
        t.m1 = umax(t.m1, v[q].m1);
        t.m2 = umax(t.m2, v[q].m2);
    }
    const u64 is = wave_incl_sum64(t.s);
    const u32 i1 = wave_incl_max(t.m1), i2 = wave_incl_max(t.m2);
    if (lane_id() == SAGPU_WAVE - 1) s_w[w] = Agg{is, i1, i2};
    __syncthreads();
    Agg c = carry;
    for (u32 q = 0; q < w; q++) {
        c.s += s_w[q].s;
        c.m1 = umax(c.m1, s_w[q].m1);
        c.m2 = umax(c.m2, s_w[q].m2);
    }
    const u32 p1 = __shfl_up(i1, 1, SAGPU_WAVE), p2 = __shfl_up(i2, 1, SAGPU_WAVE);
    if (lane_id() > 0) {
        c.m1 = umax(c.m1, p1);
        c.m2 = umax(c.m2, p2);
    }
    c.s += is - t.s;
#pragma unroll
    for (int q = 0; q < SC_K; q++) {   // exclusive prefixes into the thread's own row
        s_t[tid * (SC_K + 1) + q] = c;
        c.s += v[q].s;
        c.m1 = umax(c.m1, v[q].m1);
        c.m2 = umax(c.m2, v[q].m2);
    }
    __syncthreads();   // every row written; s_w read
#pragma unroll
    for (int q = 0; q < SC_K; q++) {
        const u32 j = (u32)q * 1024 + tid;
        if (b0 + j < tiles) tagg[b0 + j] = s_t[(j / SC_K) * (SC_K + 1) + j % SC_K];
    }
    if (tid == 1023) *s_end = c;   // the last thread's running value covers the whole step
    __syncthreads();
    const Agg e = *s_end;
    __syncthreads();   // s_t, s_w and s_end are free for the next step
    return e;
}

__global__ __launch_bounds__(1024) void k_ws_scan_tiles(Agg *tagg, u64 tiles, u64 *total, u64 *out_counter, u64 *base)
{
    __shared__ Agg s_t[1024 * (SC_K + 1)];
    __shared__ Agg s_w[16];
    __shared__ Agg s_end;
    Agg carry{0, 0, 0};
    for (u64 b0 = 0; b0 < tiles; b0 += SC_STEP) carry = scan_step(tagg, tiles, b0, carry, s_t, s_w, &s_end);
    if (threadIdx.x == 0) {
        *total = carry.s;
        if (out_counter) *base = atomicAdd(out_counter, carry.s);
    }
}

// The aggregate of chunk blockIdx.x (tiles [b SC_STEP, (b + 1) SC_STEP)).
__global__ __launch_bounds__(1024) void k_ws_chunk_reduce(const Agg *__restrict__ tagg, u64 tiles,
                                                          Agg *__restrict__ cagg)
{
    __shared__ Agg s_w[16];
    const u64 b0 = (u64)blockIdx.x * SC_STEP;
    Agg t{0, 0, 0};
#pragma unroll
    for (int q = 0; q < SC_K; q++) {
        const u64 i = b0 + (u64)q * 1024 + threadIdx.x;
        if (i < tiles) {
            const Agg v = tagg[i];
            t.s += v.s;
            t.m1 = umax(t.m1, v.m1);
            t.m2 = umax(t.m2, v.m2);
        }
    }
    for (int d = 32; d >= 1; d >>= 1) {
        t.s += __shfl_xor(t.s, d, SAGPU_WAVE);
        t.m1 = umax(t.m1, __shfl_xor(t.m1, d, SAGPU_WAVE));
        t.m2 = umax(t.m2, __shfl_xor(t.m2, d, SAGPU_WAVE));
    }
    if (lane_id() == 0) s_w[threadIdx.x / SAGPU_WAVE] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        Agg c{0, 0, 0};
        for (int q = 0; q < 16; q++) {
            c.s += s_w[q].s;
            c.m1 = umax(c.m1, s_w[q].m1);
            c.m2 = umax(c.m2, s_w[q].m2);
        }
        cagg[blockIdx.x] = c;
    }
}

// Chunk blockIdx.x scanned from its exclusive prefix cagg[b].
__global__ __launch_bounds__(1024) void k_ws_chunk_apply(Agg *tagg, u64 tiles, const Agg *__restrict__ cagg)
{
    __shared__ Agg s_t[1024 * (SC_K + 1)];
    __shared__ Agg s_w[16];
    __shared__ Agg s_end;
    scan_step(tagg, tiles, (u64)blockIdx.x * SC_STEP, cagg[blockIdx.x], s_t, s_w, &s_end);
}

// Exclusive scan of `tiles` aggregates in ws.tile_agg (chunk aggregates after the
// tiles_cap entries); total, and the reserved base when out_counter is given.
int ws_scan(WsScratch &ws, u64 tiles, u64 *out_counter, hipStream_t s)
{
    Agg *tagg = reinterpret_cast<Agg *>(ws.tile_agg);
    if (tiles <= (u64)SC_STEP) {
        hipLaunchKernelGGL(k_ws_scan_tiles, dim3(1), dim3(1024), 0, s, tagg, tiles, ws.total, out_counter, ws.base);
    } else {
        const u64 nch = (tiles + SC_STEP - 1) / SC_STEP;
        Agg *cagg = tagg + ws.tiles_cap;
        hipLaunchKernelGGL(k_ws_chunk_reduce, dim3((unsigned)nch), dim3(1024), 0, s, (const Agg *)tagg, tiles, cagg);
        hipLaunchKernelGGL(k_ws_scan_tiles, dim3(1), dim3(1024), 0, s, cagg, nch, ws.total, out_counter, ws.base);
        hipLaunchKernelGGL(k_ws_chunk_apply, dim3((unsigned)nch), dim3(1024), 0, s, tagg, tiles, (const Agg *)cagg);
    }
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

template <class F, int R>
__global__ __launch_bounds__(WS_THREADS) void k_ws_emit(F f, const Agg *__restrict__ texcl, const u64 *__restrict__ base)
{
    constexpr int SPAN = R * SAGPU_WAVE;
    __shared__ Agg s_w[WS_WAVES];
    const u32 w = threadIdx.x / SAGPU_WAVE, lane = lane_id();
    const u64 wbase = (u64)blockIdx.x * (WS_WAVES * SPAN) + (u64)w * SPAN;
    bool act[R];
    u32 v1[R], v2[R];
    u64 s = 0;
    u32 m1 = 0, m2 = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        u64 e = wbase + (u64)r * SAGPU_WAVE + lane;
        f.scan_inputs(e, act[r], v1[r], v2[r]);
        s += act[r];
        m1 = umax(m1, v1[r]);
        m2 = umax(m2, v2[r]);
    }
    for (int d = 32; d >= 1; d >>= 1) {
        s += __shfl_xor(s, d, SAGPU_WAVE);
        m1 = umax(m1, __shfl_xor(m1, d, SAGPU_WAVE));
        m2 = umax(m2, __shfl_xor(m2, d, SAGPU_WAVE));
    }
    if (lane == 0) s_w[w] = Agg{s, m1, m2};
    __syncthreads();
    Agg c = texcl[blockIdx.x];
    for (u32 q = 0; q < w; q++) {
        c.s += s_w[q].s;
        c.m1 = umax(c.m1, s_w[q].m1);
        c.m2 = umax(c.m2, s_w[q].m2);
    }
    const u64 ob = base ? *base : 0ull;
#pragma unroll
    for (int r = 0; r < R; r++) {
        u64 e = wbase + (u64)r * SAGPU_WAVE + lane;
        u32 is = wave_incl_sum((u32)act[r]);
        u32 i1 = umax(c.m1, wave_incl_max(v1[r]));
        u32 i2 = umax(c.m2, wave_incl_max(v2[r]));
        f.emit(e, act[r], ob + c.s + is - (u64)act[r], i1, i2);   // every lane calls; e >= n writes nothing
        c.s += __shfl(is, SAGPU_WAVE - 1, SAGPU_WAVE);
        c.m1 = __shfl(i1, SAGPU_WAVE - 1, SAGPU_WAVE);
        c.m2 = __shfl(i2, SAGPU_WAVE - 1, SAGPU_WAVE);
    }
}

// ------------------------------------------------------- initial grouping
constexpr int IB_MAX = 64;   // ISA buckets
// Keys are the packed first `ksym` symbols in the low 32 bits (low ksym*bits
// bits), sorted.  With K = u64 the high half holds the string id of batched GSA,
// so groups never span two strings.
template <typename K>
struct InitF {
    const K *keys;
    const u32 *vals;
    u64 n;
    u32 sepmask;     // GSA: low symbol slot; keys with it zero contain a separator
    int bits, ksym, gsa;
    u32 *SA, *act_i, *act_g, *LCP;
    // (suffix, group head) pairs for the ISA scatter, bucketed by suffix >> wlog:
    // bucket b fills PV/PG[b << wlog, ...) (it holds exactly the suffixes of that
    // range).  bcnt[b * tiles + t] = suffixes of bucket b in tile t (k_init_reduce),
    // then its exclusive scan: every bucket before b is full, so the scan gives
    // tile t's first slot in bucket b directly, with no atomics.
    u32 *PV, *PG, *bcnt;
    int wlog;
    // 1: the head of a slot alone in its group carries its exact lcp mark as a PLCP
    // hint in bits 28-31 (mark - hint_off + 1, for marks in [hint_off, hint_off + 14];
    // texts of at most 2^28 suffixes)
    int hint, hint_off;
    u32 lowbits;   // the lowest bit of each of the ksym symbol slots of a key
    // BWT seeding (cb > 0): vals carry (suffix << cb) | code of the symbol before
    // it.  codes[e] receives that code; entries of the active list keep the packed
    // form, and the rounds write the code of every suffix they place.
    int cb;
    uint8_t *codes;
    // x / bits for 0 <= x <= 64 as (x * bmul) >> 16, bmul = 2^16 / bits + 1 (exact: the
    // error x * (bmul - 2^16 / bits) stays below 2^16 / bits for bits <= 32)
    u32 bmul;

    __device__ u32 div_bits(u32 x) const { return (x * bmul) >> 16; }
    __device__ bool sep(u32 k) const { return gsa && (k & sepmask) == 0; }
    // Group flags of element e from its key and its neighbours' keys.
    __device__ void flags(u64 e, K k, K kp, K kn, bool &act, u32 &v1) const
    {
        const bool s = sep((u32)k);
        const bool head = (e == 0) || s || kp != k;
        const bool tail = (e + 1 >= n) || s || kn != k;
        act = e < n && !(head && tail);
        v1 = e < n && head ? (u32)e : 0u;
    }
    // Leading symbols two packed keys share, stopped at a separator in GSA mode.
    __device__ u32 key_lcp(u32 k, u32 kp) const
    {
        u32 l = (u32)ksym;
        u32 x = k ^ kp;
        if (x) l = div_bits((u32)(__clz(x) - (32 - ksym * bits)));
        if (gsa) {
            // the first separator (code 0) from the top: fold each symbol's bits of ~k
            // into its lowest bit, keep those bits, take the highest
            const u32 z = ~k;
            u32 m = z;
            for (int b = 1; b < bits; b++) m &= z >> b;
            m &= lowbits;
            if (m) {
                const u32 t = (u32)(ksym - 1) - div_bits((u32)(31 - __clz(m)));
                l = t < l ? t : l;
            }
        }
        return l;
    }
    // Number of trailing code-0 symbols of a key (where end-of-text padding would be).
    __device__ u32 pad_tail(u32 k) const
    {
        const u32 kb = (u32)(ksym * bits);
        const u32 kk = kb >= 32 ? k : (k & ((1u << kb) - 1u));
        return kk ? div_bits((u32)__builtin_ctz(kk)) : (u32)ksym;
    }
    // LCP mark of slot e >= 1: the lcp when the keys decide it, else LCP_UNKNOWN |
    // lower bound.  one / onep: slot e / e-1 is alone in its group, so it keeps its
    // suffix, of length len / lenp.  A pair from two different strings (batched GSA)
    // always has a separator suffix second, whose lcp is 0 by the separator rule.
    __device__ u32 lcp_mark(u32 k, u32 kp, bool one, u64 len, bool onep, u64 lenp) const
    {
        const u32 l = key_lcp(k, kp);
        if (l >= (u32)ksym && !sep(k)) return LCP_UNKNOWN | (u32)ksym;
        if (gsa) return l;
        // A suffix shorter than ksym has code-0 padding in its key past its end; where
        // the keys agree that far, it is a prefix of the other and the lcp is its
        // length.  A slot alone in its group has its suffix's length at hand.  A slot
        // of a larger group may end up holding another member, possibly a shorter
        // one: trailing code-0 symbols of the key reaching back to l leave the pair to
        // the LCP pass with bound l.
        u64 m = l;
        if (one) m = len < m ? len : m;
        else if (pad_tail(k) + l >= (u32)ksym) return LCP_UNKNOWN | l;
        if (onep) m = lenp < m ? lenp : m;
        else if (pad_tail(kp) + l >= (u32)ksym) return LCP_UNKNOWN | l;
        return (u32)m;
    }
};

// The keys of one wave's span (R rows of 64: 16, or 2 for short texts) in
// registers, and the keys just before and after the span.  Neighbours come from
// lane shuffles, so every key is loaded once and all loads of a span are in flight
// together.
template <typename K, int R>
struct SpanKeys {
    static constexpr int SPAN = R * SAGPU_WAVE;
    K k[R];
    K before, after;

    __device__ void load(const K *keys, u64 n, u64 wbase)
    {
        const u32 lane = lane_id();
#pragma unroll
        for (int r = 0; r < R; r++) {
            const u64 e = wbase + (u64)r * SAGPU_WAVE + lane;
            k[r] = e < n ? keys[e] : K(0);
        }
        before = (wbase > 0 && wbase - 1 < n) ? keys[wbase - 1] : K(0);
        after = wbase + SPAN < n ? keys[wbase + SPAN] : K(0);
    }
    // Left and right neighbour keys of row r (r is a compile-time constant in the
    // unrolled callers).
    __device__ void neighbours(int r, K &kp, K &kn) const
    {
        const u32 lane = lane_id();
        const K up = __shfl_up(k[r], 1, SAGPU_WAVE);
        const K last = r > 0 ? __shfl(k[r - 1], SAGPU_WAVE - 1, SAGPU_WAVE) : before;
        kp = lane ? up : last;
        const K down = __shfl_down(k[r], 1, SAGPU_WAVE);
        const K first = r + 1 < R ? __shfl(k[r + 1 < R ? r + 1 : r], 0, SAGPU_WAVE) : after;
        kn = lane + 1 < SAGPU_WAVE ? down : first;
    }
};

// Initial grouping, pass 1: per tile (4 spans of R rows), the number of grouped
// elements and the position of the last group start.
template <typename K, int R>
__global__ __launch_bounds__(WS_THREADS) void k_init_reduce(InitF<K> f, Agg *__restrict__ tagg)
{
    constexpr int SPAN = R * SAGPU_WAVE;
    __shared__ Agg s_w[WS_WAVES];
    const u32 w = threadIdx.x / SAGPU_WAVE, lane = lane_id();
    const u64 wbase = (u64)blockIdx.x * (WS_WAVES * SPAN) + (u64)w * SPAN;
    __shared__ u32 s_cnt[IB_MAX];
    if (threadIdx.x < IB_MAX) s_cnt[threadIdx.x] = 0;
    SpanKeys<K, R> sk;
    sk.load(f.keys, f.n, wbase);
    u32 vv[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const u64 e = wbase + (u64)r * SAGPU_WAVE + lane;
        vv[r] = e < f.n ? f.vals[e] : 0u;
    }
    __syncthreads();
    u64 s = 0;
    u32 m1 = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const u64 e = wbase + (u64)r * SAGPU_WAVE + lane;
        K kp, kn;
        sk.neighbours(r, kp, kn);
        bool act;
        u32 v1;
        f.flags(e, sk.k[r], kp, kn, act, v1);
        s += act;
        m1 = umax(m1, v1);
        // u64 shifts: wlog may be 32 (one bucket)
        if (e < f.n) atomicAdd(&s_cnt[(u32)((u64)(vv[r] >> f.cb) >> f.wlog)], 1u);
    }
    for (int d = 32; d >= 1; d >>= 1) {
        s += __shfl_xor(s, d, SAGPU_WAVE);
        m1 = umax(m1, __shfl_xor(m1, d, SAGPU_WAVE));
    }
    if (lane == 0) s_w[w] = Agg{s, m1, 0};
    __syncthreads();
    if (threadIdx.x == 0) {
        Agg t{0, 0, 0};
        for (int q = 0; q < WS_WAVES; q++) {
            t.s += s_w[q].s;
            t.m1 = umax(t.m1, s_w[q].m1);
        }
        tagg[blockIdx.x] = t;
    }
    const u32 nb = (u32)(((f.n - 1) >> f.wlog) + 1);
    if (threadIdx.x < nb) f.bcnt[(u64)threadIdx.x * gridDim.x + blockIdx.x] = s_cnt[threadIdx.x];
}

// Initial grouping, pass 2 (after k_ws_scan_tiles): singletons into SA, grouped
// elements into the active list, LCP marks, BWT seeds, and the (suffix, group head)
// pairs into their ISA buckets.
template <typename K, int R>
__global__ __launch_bounds__(WS_THREADS) void k_init_emit(InitF<K> f, const Agg *__restrict__ texcl)
{
    constexpr int SPAN = R * SAGPU_WAVE;
    __shared__ Agg s_w[WS_WAVES];
    __shared__ u32 s_cnt[IB_MAX];
    __shared__ u64 s_base[IB_MAX];
    const u32 w = threadIdx.x / SAGPU_WAVE, lane = lane_id();
    const u64 wbase = (u64)blockIdx.x * (WS_WAVES * SPAN) + (u64)w * SPAN;
    const u64 n = f.n;
    const u32 nb = (u32)(((n - 1) >> f.wlog) + 1);
    if (threadIdx.x < IB_MAX) s_cnt[threadIdx.x] = 0;
    SpanKeys<K, R> sk;
    sk.load(f.keys, n, wbase);
    u32 vv[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const u64 e = wbase + (u64)r * SAGPU_WAVE + lane;
        vv[r] = e < n ? f.vals[e] : 0u;
    }
    __syncthreads();
    // st[r]: bit 31 grouped, bit 30 group start, low bits the rank inside the
    // workgroup's share of the element's ISA bucket
    u32 st[R];
    u64 s = 0;
    u32 m1 = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const u64 e = wbase + (u64)r * SAGPU_WAVE + lane;
        K kp, kn;
        sk.neighbours(r, kp, kn);
        bool act;
        u32 v1;
        f.flags(e, sk.k[r], kp, kn, act, v1);
        s += act;
        m1 = umax(m1, v1);
        u32 rank = 0;
        // u64 shifts: wlog may be 32 (one bucket), which a 32-bit shift cannot express
        if (e < n) rank = atomicAdd(&s_cnt[(u32)((u64)(vv[r] >> f.cb) >> f.wlog)], 1u);
        st[r] = rank | (act ? 1u << 31 : 0u) | (e < n && v1 == (u32)e && (e == 0 || v1 != 0) ? 1u << 30 : 0u);
    }
    for (int d = 32; d >= 1; d >>= 1) {
        s += __shfl_xor(s, d, SAGPU_WAVE);
        m1 = umax(m1, __shfl_xor(m1, d, SAGPU_WAVE));
    }
    if (lane == 0) s_w[w] = Agg{s, m1, 0};
    __syncthreads();
    if (threadIdx.x < nb && s_cnt[threadIdx.x]) s_base[threadIdx.x] = f.bcnt[(u64)threadIdx.x * gridDim.x + blockIdx.x];
    Agg c = texcl[blockIdx.x];
    for (u32 q = 0; q < w; q++) {
        c.s += s_w[q].s;
        c.m1 = umax(c.m1, s_w[q].m1);
    }
    __syncthreads();
    const u32 cmask = (1u << f.cb) - 1u;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const u64 e = wbase + (u64)r * SAGPU_WAVE + lane;
        const bool act = st[r] >> 31;
        const u32 is = wave_incl_sum((u32)act);
        const u32 g = umax(c.m1, wave_incl_max((st[r] >> 30) & 1u ? (u32)e : 0u));
        const u64 o = c.s + is - (u64)act;
        K kp = 0, kn;
        if (f.LCP) sk.neighbours(r, kp, kn);
        // the slot before: its value (suffix) and whether it is alone in its group
        // (unknown for the first slot of the span: taken as grouped)
        const u32 vup = __shfl_up(vv[r], 1, SAGPU_WAVE);
        const u32 vlast = r > 0 ? __shfl(vv[r > 0 ? r - 1 : 0], SAGPU_WAVE - 1, SAGPU_WAVE) : 0u;
        const u32 vp = lane ? vup : vlast;
        const u32 gup = __shfl_up(st[r] >> 31, 1, SAGPU_WAVE);
        const u32 glast = r > 0 ? __shfl(st[r > 0 ? r - 1 : 0] >> 31, SAGPU_WAVE - 1, SAGPU_WAVE) : 1u;
        const bool onep = !(lane ? gup : glast);
        if (e < n) {
            const u32 v = vv[r] >> f.cb;
            if (f.codes) f.codes[e] = (uint8_t)(vv[r] & cmask);
            u32 mark = 0;
            if (f.LCP && e > 0)
                mark = f.lcp_mark((u32)sk.k[r], (u32)kp, !act, n - (u64)(vv[r] >> f.cb), onep, n - (u64)(vp >> f.cb));
            // (an unknown mark is >= 2^31 and gets none)
            const u32 hint = (f.hint && f.LCP && !act && mark >= (u32)f.hint_off && mark - (u32)f.hint_off < 15u)
                                 ? (mark - (u32)f.hint_off + 1u) << 28
                                 : 0u;
            const u64 slot = s_base[(u32)((u64)v >> f.wlog)] + (st[r] & 0x3fffffffu);
            f.PV[slot] = v;
            f.PG[slot] = g | hint;
            // every slot is written (full lines); slots of grouped elements are
            // overwritten when their suffixes are placed
            f.SA[e] = v;
            if (act) {
                f.act_i[o] = vv[r];   // with the BWT seed the code travels in the list entry
                f.act_g[o] = g;
            }
            if (f.LCP) f.LCP[e] = mark;
        }
        c.s += __shfl(is, SAGPU_WAVE - 1, SAGPU_WAVE);
        c.m1 = __shfl(g, SAGPU_WAVE - 1, SAGPU_WAVE);
    }
}

// ISA[PV[e]] = PG[e] over the bucketed pairs, in order: the workgroups in flight at
// any time work inside one or two buckets, so the random 4-byte writes of the whole
// device land in a stretch of ISA that the memory-side cache holds and merges.
// Measured at 2^28 entries: 2^25-entry buckets (128 MiB of ISA) 4.0 ms, 2^22-2^24
// 5.4-6.5 ms, 2^26-2^27 8-10 ms, one bucket per XCD 9.8 ms.
constexpr int IW = 8;
__global__ __launch_bounds__(256) void k_isa_scatter(const u32 *__restrict__ PV, const u32 *__restrict__ PG, u64 n,
                                                     u32 *__restrict__ ISA)
{
    const u64 chunk = 256ull * IW;
    for (u64 base = (u64)blockIdx.x * chunk; base < n; base += (u64)gridDim.x * chunk) {
        u32 v[IW], g[IW];
#pragma unroll
        for (int k = 0; k < IW; k++) {
            const u64 e = base + (u64)k * 256 + threadIdx.x;
            v[k] = e < n ? PV[e] : 0u;
            g[k] = e < n ? PG[e] : 0u;
        }
#pragma unroll
        for (int k = 0; k < IW; k++)
            if (base + (u64)k * 256 + threadIdx.x < n) ISA[v[k]] = g[k];
    }
}

// ------------------------------------------------------- big-group emission
// CT = u32 when the group index and key2 fit 32 bits together (the big groups of
// degenerate texts: 16 B per element and radix pass instead of 24), else u64.
template <typename CT>
struct BigF {
    const CT *comp;      // (dense big-group index << kb) | key2, sorted
    const u32 *vals;     // list entry: suffix index, or (suffix << cb) | BWT code
    const u32 *gslot;    // group word of each big group: first SA slot | flag
    u64 n;
    u32 h;
    int kb;              // key2 width
    u32 *SA, *ISA, *out_i, *out_g, *LCP;
    ActCode ac;
    ExactLcp ex;
    BigPivots bp;        // the pivot classes taken out of the sort (bp.eq null: none)
    // flagged groups: the rank their members hold, and the key2 of the sorted
    // members' subgroup whose slots hold it (BIG_NOKEY: none of them, or a pivot's)
    const u32 *grank;
    const u64 *keepkey;

    __device__ void scan_inputs(u64 e, bool &act, u32 &v1, u32 &v2) const
    {
        if (e >= n) {
            act = false;
            v1 = v2 = 0;
            return;
        }
        const u64 c = comp[e];
        const u64 cp = e > 0 ? (u64)comp[e - 1] : ~0ull;
        const u64 cn = e + 1 < n ? (u64)comp[e + 1] : ~0ull;
        bool head = (e == 0) || c != cp;
        bool tail = (e + 1 == n) || c != cn;
        bool gch = (e == 0) || (c >> kb) != (cp >> kb);
        act = !(head && tail);
        v1 = gch ? (u32)e : 0u;
        v2 = head ? (u32)e : 0u;
    }
    __device__ void emit(u64 e, bool act, u64 o, u32 gstart, u32 hstart) const
    {
        if (e >= n) return;
        const u32 pv = vals[e];
        const u32 v = pv >> ac.cb;
        const u64 c = comp[e];
        const u32 q = (u32)(c >> kb);
        const u32 gw = gslot[q];
        const u32 g = gw & ~ac.gf;
        const u64 key = c & ((1ull << kb) - 1);
        u32 r = g + (hstart - gstart);
        if (bp.eq) {
            // members of the group's pivot classes with smaller keys rank below
            const u32 np = bp.npiv[q];
            for (u32 b = 0; b < np; b++)
                if (bp.piv[(u64)q * BP + b] < key) r += bp.eq[(u64)q * BP + b];
        }
        // the subgroup whose slots hold the group's rank keeps it (and the flag, its
        // rank then stored at its first slot + 1); new ranks here are first slots
        const bool flagged = (gw & ac.gf) != 0;
        const bool kept = flagged ? key == keepkey[q] : r == g;
        u32 hint = 0;   // PLCP hint of a suffix placed with an exact mark (ISA bits 28-31)
        if (LCP && hstart == (u32)e && r != g) {
            // a subgroup after the group's first: e - 1 is the last entry of the
            // previous subgroup (exact marks come only without pivot classes)
            u32 mark = LCP_UNKNOWN | (h < LCP_BOUND_CAP ? h : LCP_BOUND_CAP);
            if (ex.on && !ex.skipped(g) && hstart != gstart) {
                const u64 km = (1ull << kb) - 1;
                mark = h + ex.shared(c & km, (u64)comp[e - 1] & km);
                if (!act) hint = ex.hint_bits(mark);
            }
            LCP[r] = mark;
        }
        if (!kept) ISA[v] = r | hint;
        if (act) {
            out_i[o] = pv;
            out_g[o] = kept ? r | (gw & ac.gf) : r;
            if (kept && flagged && r != g && hstart == (u32)e) SA[r + 1] = grank[q];
        } else {
            SA[r] = v;
            if (ac.codes) ac.codes[r] = (uint8_t)(pv & ((1u << ac.cb) - 1u));
        }
    }
};

// Rows per wave span of the segmented scans over big-group members: 8 (2048-element
// tiles) keeps the emit pass's registers at 4 x 8 values per lane, 16 rows held
// 116 VGPRs (4 waves per SIMD).
constexpr int WS_RUN_ROWS = 8;
template <class F>
int ws_run(WsScratch &ws, const F &f, u64 n, u64 *out_counter, hipStream_t s)
{
    constexpr u64 tile = (u64)WS_WAVES * WS_RUN_ROWS * SAGPU_WAVE;
    const u64 tiles = (n + tile - 1) / tile;
    if (tiles > ws.tiles_cap) {
        sagpu_set_error("segment scratch too small", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }
    Agg *tagg = reinterpret_cast<Agg *>(ws.tile_agg);
    hipLaunchKernelGGL((k_ws_reduce<F, WS_RUN_ROWS>), dim3((unsigned)tiles), dim3(WS_THREADS), 0, s, f, tagg);
    SAGPU_TRY(ws_scan(ws, tiles, out_counter, s));
    hipLaunchKernelGGL((k_ws_emit<F, WS_RUN_ROWS>), dim3((unsigned)tiles), dim3(WS_THREADS), 0, s, f, tagg,
                       out_counter ? ws.base : (const u64 *)nullptr);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

// ------------------------------------------------------------- round kernels
// IG list entries per thread and step, all loads (the random ISA reads included)
// in flight before the stores.
constexpr int IG = 4;
template <typename KT>
__global__ __launch_bounds__(256) void k_round_gather(const u32 *__restrict__ act_i, const u32 *__restrict__ act_g,
                                                      const u64 *__restrict__ rp, const u32 *__restrict__ ISA, u64 n,
                                                      KT *__restrict__ key2, RawKey1 rk, int cb)
{
    // the round's list length and offset come from device memory (RoundParams), so
    // a captured launch serves every round
    const u64 a = rp[0], h = rp[1];
    const u64 chunk = 256ull * IG;
    for (u64 base = (u64)blockIdx.x * chunk; base < a; base += (u64)gridDim.x * chunk) {
        u32 i[IG], r[IG];
        bool raw[IG];
#pragma unroll
        for (int q = 0; q < IG; q++) {
            const u64 k = base + (u64)q * 256 + threadIdx.x;
            i[q] = k < a ? act_i[k] >> cb : 0u;
            // round 1 with a code stream: the packed key of i + h orders the group as
            // its rank does (equal keys, equal ranks), read from the 2-bits-per-symbol
            // stream instead of the 4-byte ISA; groups holding a suffix shorter than
            // h keep the rank form
            raw[q] = false;
            if (rk.pw && k < a) {
                const u32 g = act_g[k];
                raw[q] = true;
                for (int t = 0; t < rk.nspecial; t++) raw[q] = raw[q] && g != rk.special[t];
            }
        }
#pragma unroll
        for (int q = 0; q < IG; q++) {
            const u64 k = base + (u64)q * 256 + threadIdx.x;
            const u64 ih = (u64)i[q] + h;
            if (k < a && raw[q]) {
                const u64 bitpos = ih * (u64)rk.bits;
                const u64 w = bitpos >> 5;
                const u64 x = ((u64)rk.pw[w] << 32) | (u64)rk.pw[w + 1];
                r[q] = (u32)((x << (bitpos & 31)) >> (64 - rk.kbits));
            } else {
                r[q] = (k < a && ih < n) ? isa_rank(ISA[ih], n) : 0u;
            }
        }
#pragma unroll
        for (int q = 0; q < IG; q++) {
            const u64 k = base + (u64)q * 256 + threadIdx.x;
            if (k >= a) continue;
            const u64 ih = (u64)i[q] + h;
            key2[k] = raw[q] ? (KT)r[q] : ih < n ? (KT)(h + r[q]) : (KT)(n - 1 - i[q]);
        }
    }
}

constexpr int RT_THREADS = 256;
// A workgroup whose largest small group has more members ranks by sorting its range
// in LDS instead of counting over each entry's group
constexpr int RT_SORT_MIN = 96;

// RANGE = list entries whose group heads one workgroup takes (256, 512 or 1024; the
// "tile_range" option): smaller ranges use less LDS and fit more workgroups per CU.
//
// Group extents come from the workgroup's own window of act_g, [lo, lo + RANGE +
// RT_SMALL + 1): a group with its head in [lo, lo + RANGE) and at most RT_SMALL
// members ends inside it.  Groups of more members are big: their heads go to the
// big list, and the workgroup holding a big group's last entry writes that entry's
// list index to SA[g] (the group's SA slots are free until its suffixes are
// placed, and no workgroup places suffixes of a big group here), from which
// k_big_sizes takes the size after this kernel.  Big heads are reserved on the
// counter of the workgroup's XCD (seg_ctr[16 x + 8]) in that XCD's segment of
// (bseg_h, bseg_g): on texts of long repeats a round can list 10^5 big groups, and
// as many returning atomics on one address serialise for milliseconds.
template <int RANGE>
__device__ __forceinline__ void rt_big_head(u64 head, u32 g, u64 *__restrict__ seg_ctr, u32 *__restrict__ bseg_h,
                                            u32 *__restrict__ bseg_g, u64 bseg_cap, u32 *__restrict__ err)
{
    const u32 x = blockIdx.x & 7u;
    const u64 slot = atomicAdd(&seg_ctr[x * 16 + 8], 1ull);
    if (slot < bseg_cap) {
        bseg_h[(u64)x * bseg_cap + slot] = (u32)head;
        bseg_g[(u64)x * bseg_cap + slot] = g;
    } else {
        atomicOr(err, 1u);
    }
}

template <typename KT, int RANGE>
__global__ __launch_bounds__(RT_THREADS) void k_round_tile(const u32 *__restrict__ act_i, const u32 *__restrict__ act_g,
                                                           const KT *__restrict__ key2, const u64 *__restrict__ rp,
                                                           u32 *__restrict__ SA, u32 *__restrict__ ISA,
                                                           u32 *__restrict__ seg_i, u32 *__restrict__ seg_g,
                                                           u64 *__restrict__ seg_ctr, u64 seg_cap,
                                                           u32 *__restrict__ err, u32 *__restrict__ bseg_h,
                                                           u32 *__restrict__ bseg_g, u64 bseg_cap,
                                                           u32 *__restrict__ LCP, ActCode ac, ExactLcp ex)
{
    constexpr int RT_CAP = RANGE + RT_SMALL, RT_PER = RT_CAP / RT_THREADS;
    constexpr int WIN = RANGE + RT_SMALL + 1;                  // act_g window [lo, lo + WIN)
    constexpr int W_PER = (WIN + RT_THREADS - 1) / RT_THREADS; // window entries per thread
    static_assert(RT_CAP % RT_THREADS == 0, "tile capacity must split evenly");
    constexpr u32 NONE = 0xFFFFFFFFu;
    constexpr u16 NONE16 = 0xFFFFu;   // local indexes are below 2048: 16 bits each
    // act_g over the window, and over the RT_SMALL entries before it: whether the
    // group straddling lo has more than RT_SMALL members is then an LDS compare
    __shared__ u32 s_wx[RT_SMALL + WIN];
    u32 *const s_w = s_wx + RT_SMALL;   // s_w[x] = act_g[lo + x], x >= -RT_SMALL
    __shared__ u16 s_gend[RANGE];   // last local index of the group headed at local x (NONE16: past the window)
    // One block for: the keys of the range by local position (every small group of
    // this workgroup lies in [lo, lo + RT_CAP)), the local index of each position's
    // group head (NONE16: head before lo) and the survivor flags by output position;
    // the sort keys of the ranking by sorting take the whole block, and the
    // survivors' output ranks take the keys' place once the ranking is done.  The
    // list entries stay in registers of the threads that rank them, so a workgroup
    // needs about 15 B of LDS per list entry and 8 workgroups of 1024 entries fit a CU.
    constexpr u32 RAW_A = (u32)sizeof(KT) * RT_CAP + 2u * ((WIN + 1) & ~1) + RT_CAP;
    constexpr u32 RAW = RAW_A > 8u * RT_CAP ? RAW_A : 8u * RT_CAP;
    __shared__ __align__(8) uint8_t s_raw[RAW];
    KT *const s_k = reinterpret_cast<KT *>(s_raw);
    u16 *const s_hx = reinterpret_cast<u16 *>(s_raw + sizeof(KT) * RT_CAP);
    uint8_t *const s_os = s_raw + sizeof(KT) * RT_CAP + 2u * ((WIN + 1) & ~1);
    u64 *const s_o = reinterpret_cast<u64 *>(s_raw);   // sort keys
    u16 *const s_rk = reinterpret_cast<u16 *>(s_raw);  // survivors' output ranks
    __shared__ u32 s_first, s_end, s_gmax, s_wsum[4];
    __shared__ u64 s_base;

    const u64 a = rp[0];
    const u32 h = rp[1] < LCP_BOUND_CAP ? (u32)rp[1] : LCP_BOUND_CAP;   // marks carry h capped
    const u32 tid = threadIdx.x;
    const u64 lo = (u64)blockIdx.x * RANGE;
    if (lo >= a) return;
    const u64 hi = lo + RANGE < a ? lo + RANGE : a;
    const u32 wlen = (u32)((a - lo) < (u64)WIN ? (a - lo) : (u64)WIN);
    const u32 rlen = (u32)(hi - lo);
    const u32 clen = wlen < (u32)RT_CAP ? wlen : (u32)RT_CAP;

    // 1. the window, the entry before it, and the group head of every position (a
    //    max-scan over head flags: local index + 1, 0 = head before lo)
    for (u32 x = tid; x < wlen + (u32)RT_SMALL; x += RT_THREADS)
        s_wx[x] = lo + x >= (u64)RT_SMALL ? act_g[lo + x - RT_SMALL] : NONE;
    for (u32 x = tid; x < (u32)RT_CAP; x += RT_THREADS) s_os[x] = 0;
    if (tid == 0) {
        s_first = NONE;
        s_end = 0;
        s_gmax = 0;
    }
    for (u32 x = tid; x < (u32)RANGE; x += RT_THREADS) s_gend[x] = NONE16;
    __syncthreads();
    // A whole window inside one group (groups are contiguous runs of one act_g
    // value): no group ends in [lo, hi) and no small group starts there; only a
    // head at lo, of a group of at least WIN > RT_SMALL members, is recorded.  On
    // texts of long repeats most windows are of this kind.
    if (wlen == (u32)WIN && s_w[0] == s_w[WIN - 1]) {
        if (tid == 0 && (lo == 0 || s_w[-1] != s_w[0]))
            rt_big_head<RANGE>(lo, s_w[0], seg_ctr, bseg_h, bseg_g, bseg_cap, err);
        return;
    }
    {
        u32 loc[W_PER];
        u32 m = 0;
#pragma unroll
        for (int j = 0; j < W_PER; j++) {
            const u32 x = tid * W_PER + j;
            const bool head = x < wlen && (x == 0 ? (lo == 0 || s_w[-1] != s_w[0]) : s_w[x - 1] != s_w[x]);
            if (head) m = x + 1;
            loc[j] = m;
        }
        const u32 inc = wave_incl_max(m);
        if (lane_id() == SAGPU_WAVE - 1) s_wsum[tid / SAGPU_WAVE] = inc;
        __syncthreads();
        u32 c = 0;
        for (u32 q = 0; q < tid / SAGPU_WAVE; q++) c = umax(c, s_wsum[q]);
        const u32 prev = __shfl_up(inc, 1, SAGPU_WAVE);
        if (lane_id() > 0) c = umax(c, prev);
#pragma unroll
        for (int j = 0; j < W_PER; j++) {
            const u32 x = tid * W_PER + j;
            if (x < wlen) s_hx[x] = (u16)(umax(c, loc[j]) - 1u);   // 0 - 1 = NONE16
        }
    }
    __syncthreads();
    // 2. group ends: a position is its group's last entry when the next entry
    //    differs (or it is the list's last).  Ends of groups headed in [lo, hi) are
    //    recorded per head; ends in [lo, hi) of big groups get their SA marker.
    for (u32 x = tid; x < wlen; x += RT_THREADS) {
        const bool is_end = (x + 1 < wlen) ? s_w[x + 1] != s_w[x] : (lo + x + 1 == a);
        if (!is_end) continue;
        const u32 hx = s_hx[x];
        if (hx != NONE16 && hx < rlen) s_gend[hx] = (u16)x;
        if (x < rlen) {
            bool big;
            if (hx != NONE16) big = x - hx + 1 > (u32)RT_SMALL;
            else {
                // the group holding position lo has its head before lo
                const u64 e = lo + x;
                big = e >= (u64)RT_SMALL && s_w[(int)x - RT_SMALL] == s_w[x];
            }
            if (big) SA[s_w[x] & ~ac.gf] = (u32)(lo + x);
        }
    }
    __syncthreads();
    // 3. heads in [lo, hi): small groups define the tile's element range (a min and
    //    a max per wave, then one LDS atomic each), big ones go to the big list
    {
        u32 f = NONE, e = 0, gm = 0;
        for (u32 x = tid; x < rlen; x += RT_THREADS) {
            if (s_hx[x] != x) continue;
            const u32 ge = s_gend[x];
            if (ge != NONE16 && ge - x + 1 <= (u32)RT_SMALL) {
                f = x < f ? x : f;
                e = ge + 1 > e ? ge + 1 : e;
                gm = ge - x + 1 > gm ? ge - x + 1 : gm;
            } else {
                rt_big_head<RANGE>(lo + x, s_w[x], seg_ctr, bseg_h, bseg_g, bseg_cap, err);
            }
        }
        for (int d = 32; d >= 1; d >>= 1) {
            const u32 of = __shfl_xor(f, d, SAGPU_WAVE), oe = __shfl_xor(e, d, SAGPU_WAVE);
            const u32 om = __shfl_xor(gm, d, SAGPU_WAVE);
            f = of < f ? of : f;
            e = oe > e ? oe : e;
            gm = om > gm ? om : gm;
        }
        if (lane_id() == 0 && f != NONE) {
            atomicMin(&s_first, f);
            atomicMax(&s_end, e);
            atomicMax(&s_gmax, gm);
        }
    }
    __syncthreads();
    if (s_first == NONE) return;
    const u32 first = s_first, end = s_end;
    if (end < first || end > clen) {
        // inconsistent group bookkeeping: report instead of overrunning LDS
        if (tid == 0) atomicOr(err, 1u);
        return;
    }
    // the range's keys into LDS; each thread's own entries (positions first + tid +
    // j * RT_THREADS) into its registers
    u32 e_ai[RT_PER];
#pragma unroll
    for (int j = 0; j < RT_PER; j++) {
        const u32 x = first + tid + (u32)j * RT_THREADS;
        e_ai[j] = 0;
        if (x < end) {
            s_k[x] = key2[lo + x];
            e_ai[j] = act_i[lo + x];
        }
    }
    __syncthreads();

    // 4. rank inside the group by key2: lt smaller, eq equal, eqb equal before (in
    //    list order), pk the largest smaller key; local positions throughout.  The
    //    entry goes to position hx + lt + eqb, its subgroup's first slot is g + lt.
    //    Per entry of the thread: output position (NONE: none), group word for the
    //    next list, the rank written to ISA, flags (1: survives, 2: ISA written).
    u32 e_pos[RT_PER], e_word[RT_PER], e_isa[RT_PER], e_fl[RT_PER];
    auto place = [&](int j, u32 hx, u32 lt, u32 eq, u32 eqb, KT key, KT pk) {
        const u32 gw = s_w[hx];   // first slot | flag
        const u32 g = gw & ~ac.gf;
        const u32 r = g + lt;
        // the subgroup whose slots hold the group's rank keeps rank and flag (a
        // flagged group's rank is stored in SA[g + 1]; its other slots, and those of
        // the next subgroup's first slot + 1, are unplaced until this round's
        // writes, which come after a barrier); the others rank at their first slot,
        // unflagged
        const u32 gv = (gw & ac.gf) ? SA[g + 1] : g;
        const bool kept = gv - g >= lt && gv - g < lt + eq;
        u32 hint = 0;   // PLCP hint of a suffix placed with an exact mark (ISA bits 28-31)
        if (kept && (gw & ac.gf) && lt > 0 && eq > 1 && eqb == 0) SA[r + 1] = gv;
        if (LCP && lt > 0 && eqb == 0) {
            u32 mark = LCP_UNKNOWN | (h < LCP_BOUND_CAP ? h : LCP_BOUND_CAP);
            if (ex.on && !ex.skipped(g)) {
                mark = h + ex.shared((u64)key, (u64)pk);
                if (eq == 1) hint = ex.hint_bits(mark);
            }
            LCP[r] = mark;
        }
        e_pos[j] = hx + lt + eqb;
        e_word[j] = r | (kept ? (gw & ac.gf) : 0u);
        e_isa[j] = r | hint;
        e_fl[j] = (eq > 1 ? 1u : 0u) | (kept ? 0u : 2u);
    };
#pragma unroll
    for (int j = 0; j < RT_PER; j++) e_pos[j] = NONE;
    if (s_gmax <= (u32)RT_SORT_MIN) {
        // small groups: each entry counts over its group (O(size) per entry)
#pragma unroll
        for (int j = 0; j < RT_PER; j++) {
            const u32 x = first + tid + (u32)j * RT_THREADS;
            if (x >= end) continue;
            const u32 hx = s_hx[x];
            const u32 ge = s_gend[hx];
            if (ge == NONE16 || ge - hx + 1 > (u32)RT_SMALL) continue;   // big group: not ours
            const u32 gsz = ge - hx + 1;
            const KT key = s_k[x];
            u32 lt = 0, eq = 0, eqb = 0;
            KT pk = 0;
            for (u32 y = hx; y < hx + gsz; y++) {
                const KT ky = s_k[y];
                lt += ky < key;
                eq += ky == key;
                eqb += (ky == key) & (y < x);
                pk = (ky < key && ky > pk) ? ky : pk;
            }
            place(j, hx, lt, eq, eqb, key, pk);
        }
    } else {
        // groups of more than RT_SORT_MIN members: the range [first, end) sorted by
        // (group head, key2, position): 64-entry chunks sorted in registers, then
        // merged pairwise in LDS; every entry then finds its subgroup's bounds by
        // binary search inside its group's sorted stretch.  The entry at
        // sorted position p is not this thread's own: its list word is read again
        // (it was read by this workgroup just before).
        constexpr int XB = 11;   // position and head fields (RT_CAP <= 2048)
        constexpr u64 KMASK = (1ull << (64 - 2 * XB)) - 1;
        static_assert(RT_CAP <= (1 << XB), "sort fields too narrow");
        const u32 cnt = end - first;
        const u32 lane = lane_id();
        u64 comp[RT_PER];
        u32 pos[RT_PER];
#pragma unroll
        for (int j = 0; j < RT_PER; j++) {
            const u32 p = tid + (u32)j * RT_THREADS;
            // past the range: +inf, which the chunk sort moves to the chunk's end
            comp[j] = p < cnt ? ((u64)s_hx[first + p] << (64 - XB)) | ((u64)s_k[first + p] << XB) | p : ~0ull;
            pos[j] = p;
        }
        // each wave sorts its 64-entry chunks (positions j * RT_THREADS + 64 w + lane) in
        // registers, a bitonic network over the lanes
#pragma unroll
        for (u32 k = 2; k <= (u32)SAGPU_WAVE; k <<= 1) {
#pragma unroll
            for (u32 d = k >> 1; d > 0; d >>= 1) {
                const bool keep_min = ((lane & d) == 0) == ((lane & k) == 0);
#pragma unroll
                for (int j = 0; j < RT_PER; j++) {
                    const u64 o = __shfl_xor(comp[j], (int)d, SAGPU_WAVE);
                    comp[j] = keep_min ? (o < comp[j] ? o : comp[j]) : (o > comp[j] ? o : comp[j]);
                }
            }
        }
        __syncthreads();   // the keys and heads are read; the sort keys take their LDS
#pragma unroll
        for (int j = 0; j < RT_PER; j++)
            if (pos[j] < cnt) s_o[pos[j]] = comp[j];
        __syncthreads();
        // sorted runs of L entries merged pairwise: an entry moves to its offset in its
        // run plus the number of entries of the partner run below it (the composite
        // keys are distinct), found by binary search; then all entries are stored
        for (u32 L = SAGPU_WAVE; L < cnt; L <<= 1) {
            u32 np[RT_PER];
#pragma unroll
            for (int j = 0; j < RT_PER; j++) {
                np[j] = pos[j];
                if (pos[j] >= cnt) continue;
                const u32 r0 = pos[j] & ~(2u * L - 1u);
                const bool second = (pos[j] & L) != 0;
                const u32 q0 = second ? r0 : r0 + L;
                if (q0 >= cnt) continue;   // no partner run: in place
                const u32 q1 = q0 + L < cnt ? q0 + L : cnt;
                const u64 v = comp[j];
                u32 a0 = q0, a1 = q1;
                while (a0 < a1) {
                    const u32 m = (a0 + a1) >> 1;
                    if (s_o[m] < v) a0 = m + 1;
                    else a1 = m;
                }
                np[j] = pos[j] - (second ? L : 0u) + (a0 - q0);
            }
            __syncthreads();   // the searches of this level are done
#pragma unroll
            for (int j = 0; j < RT_PER; j++) {
                if (pos[j] >= cnt) continue;
                s_o[np[j]] = comp[j];
                pos[j] = np[j];
            }
            __syncthreads();
        }
        u32 rh[RT_PER], rlt[RT_PER], req[RT_PER], reqb[RT_PER];
        KT rkey[RT_PER], rpk[RT_PER];
#pragma unroll
        for (int j = 0; j < RT_PER; j++) {
            const u32 p = tid + (u32)j * RT_THREADS;
            rh[j] = NONE;
            if (p >= cnt) continue;
            const u64 c = s_o[p];
            const u32 hx = (u32)(c >> (64 - XB));
            const u32 ge = s_gend[hx];
            if (ge == NONE16 || ge - hx + 1 > (u32)RT_SMALL) continue;   // big group: not ours
            const u64 ck = c >> XB;   // (head, key)
            u32 a0 = hx - first, a1 = p;   // first entry of the subgroup: in [group start, p]
            while (a0 < a1) {
                const u32 m = (a0 + a1) >> 1;
                if ((s_o[m] >> XB) < ck) a0 = m + 1;
                else a1 = m;
            }
            u32 b0 = p + 1, b1 = ge - first + 1;   // past its last: in (p, group end]
            while (b0 < b1) {
                const u32 m = (b0 + b1) >> 1;
                if ((s_o[m] >> XB) <= ck) b0 = m + 1;
                else b1 = m;
            }
            e_ai[j] = act_i[lo + first + (u32)(c & ((1u << XB) - 1))];
            rh[j] = hx;
            rlt[j] = a0 - (hx - first);
            req[j] = b0 - a0;
            reqb[j] = p - a0;
            rkey[j] = (KT)(ck & KMASK);
            rpk[j] = rlt[j] > 0 ? (KT)((s_o[a0 - 1] >> XB) & KMASK) : (KT)0;
        }
        __syncthreads();   // the sort keys are read; the survivor flags take their LDS again
        for (u32 x = tid; x < (u32)RT_CAP; x += RT_THREADS) s_os[x] = 0;
#pragma unroll
        for (int j = 0; j < RT_PER; j++)
            if (rh[j] != NONE) place(j, rh[j], rlt[j], req[j], reqb[j], rkey[j], rpk[j]);
    }
    __syncthreads();   // every read of a flagged group's stored rank is done

    // 5. ranks and placed suffixes; survivor flags by output position
#pragma unroll
    for (int j = 0; j < RT_PER; j++) {
        if (e_pos[j] == NONE) continue;
        const u32 v = e_ai[j] >> ac.cb;   // list entries may carry the BWT code
        if (e_fl[j] & 2u) ISA[v] = e_isa[j];
        if (!(e_fl[j] & 1u)) {
            const u32 r = e_word[j] & ~ac.gf;
            SA[r] = v;
            if (ac.codes) ac.codes[r] = (uint8_t)(e_ai[j] & ((1u << ac.cb) - 1u));
        } else {
            s_os[e_pos[j]] = 1;
        }
    }
    __syncthreads();

    // 6. survivors: output ranks by position, their range reserved in this XCD's
    //    segment (one atomic on the XCD's counter).  A counter shared by all
    //    workgroups serialises them: 300 K returning atomics on one address take
    //    3.6 ms, on 8 addresses 0.57 ms; workgroups go to the XCDs round-robin
    //    (blockIdx % 8).
    u32 cnt_s[RT_PER];
    u32 tot = 0;
#pragma unroll
    for (int j = 0; j < RT_PER; j++) {
        const u32 x = first + tid * RT_PER + j;
        cnt_s[j] = (x < end && s_os[x]) ? 1u : 0u;
        tot += cnt_s[j];
    }
    const u32 inc = wave_incl_sum(tot);
    if (lane_id() == SAGPU_WAVE - 1) s_wsum[tid / SAGPU_WAVE] = inc;
    __syncthreads();
    if (tid == 0) {
        const u32 all = s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3];
        const u32 xcd = blockIdx.x & 7u;
        s_base = (u64)xcd * seg_cap + (all ? atomicAdd(&seg_ctr[xcd * 16], (u64)all) : 0ull);
    }
    {
        u32 o = inc - tot;
        for (u32 q = 0; q < tid / SAGPU_WAVE; q++) o += s_wsum[q];
#pragma unroll
        for (int j = 0; j < RT_PER; j++) {
            const u32 x = first + tid * RT_PER + j;
            if (x < end) s_rk[x] = (u16)o;
            o += cnt_s[j];
        }
    }
    __syncthreads();
    const u64 base = s_base;
#pragma unroll
    for (int j = 0; j < RT_PER; j++) {
        if (e_pos[j] == NONE || !(e_fl[j] & 1u)) continue;
        const u64 o = base + s_rk[e_pos[j]];
        seg_i[o] = e_ai[j];
        seg_g[o] = e_word[j];
    }
}

// The next active list from the 8 XCD segments of the tile kernel's output
// (segment x: [x * seg_cap, x * seg_cap + seg_ctr[16 x])), concatenated in segment
// order; *out_counter = its length (the big-group emission appends behind it).
__global__ __launch_bounds__(256) void k_seg_gather(const u32 *__restrict__ seg_i, const u32 *__restrict__ seg_g,
                                                    const u64 *__restrict__ seg_ctr, u64 seg_cap,
                                                    u32 *__restrict__ out_i, u32 *__restrict__ out_g,
                                                    u64 *__restrict__ out_counter)
{
    u64 pre[9];
    pre[0] = 0;
#pragma unroll
    for (int x = 0; x < 8; x++) pre[x + 1] = pre[x] + seg_ctr[x * 16];
    const u64 total = pre[8];
    if (blockIdx.x == 0 && threadIdx.x == 0) *out_counter = total;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += stride) {
        int x = 0;
#pragma unroll
        for (int q = 1; q < 8; q++) x += k >= pre[q];
        const u64 src = (u64)x * seg_cap + (k - pre[x]);
        out_i[k] = seg_i[src];
        out_g[k] = seg_g[src];
    }
}

// The big groups the tile kernel listed, its 8 XCD segments (segment x: [x *
// bseg_cap, x * bseg_cap + seg_ctr[16 x + 8])) concatenated in segment order, with
// their sizes: the entry holding a big group's last list index wrote it to SA[g].
// Sizes are stored minus one (a group may hold 2^32 members).  *big_counter = the
// count (entries past big_cap are not written; the host rejects such a count).
__global__ __launch_bounds__(256) void k_big_sizes(const u32 *__restrict__ bseg_h, const u32 *__restrict__ bseg_g,
                                                   const u64 *__restrict__ seg_ctr, u64 bseg_cap,
                                                   u32 *__restrict__ big_head, u32 *__restrict__ big_g,
                                                   u32 *__restrict__ big_counter, u64 big_cap,
                                                   const u32 *__restrict__ SA, u32 *__restrict__ big_size, u32 gf)
{
    u64 pre[9];
    pre[0] = 0;
#pragma unroll
    for (int x = 0; x < 8; x++) {
        const u64 c = seg_ctr[x * 16 + 8];
        pre[x + 1] = pre[x] + (c < bseg_cap ? c : bseg_cap);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *big_counter = (u32)(pre[8] < 0xFFFFFFFFull ? pre[8] : 0xFFFFFFFFull);
    const u64 nb = pre[8] < big_cap ? pre[8] : big_cap;
    for (u64 q = (u64)blockIdx.x * blockDim.x + threadIdx.x; q < nb; q += (u64)gridDim.x * blockDim.x) {
        int x = 0;
#pragma unroll
        for (int t = 1; t < 8; t++) x += q >= pre[t];
        const u64 src = (u64)x * bseg_cap + (q - pre[x]);
        const u32 hd = bseg_h[src], g = bseg_g[src];
        big_head[q] = hd;
        big_g[q] = g;
        big_size[q] = SA[g & ~gf] - hd;
    }
}

// Offsets of the big groups in the concatenation of their members, in the order
// k_big_sizes listed them (nb = *big_counter of them, at most big_cap): boff[q] =
// the members of the groups before q, boff[nb] = *total = all of them (one
// workgroup: each thread sums a contiguous chunk, then a scan of the 1024 chunk
// sums).  A listed group of at most RT_SMALL members is an inconsistency (err bit 2).
__global__ __launch_bounds__(1024) void k_big_offsets(const u32 *__restrict__ big_size,
                                                      const u32 *__restrict__ big_counter, u64 big_cap,
                                                      u64 *__restrict__ boff, u64 *__restrict__ total,
                                                      u32 *__restrict__ err)
{
    __shared__ u64 s_sum[1024 / SAGPU_WAVE];
    const u32 nb = (u32)((u64)*big_counter < big_cap ? (u64)*big_counter : big_cap);
    const u32 t = threadIdx.x;
    if (nb == 0) {   // most rounds of most texts: nothing to scan
        if (t == 0) {
            boff[0] = 0;
            *total = 0;
        }
        return;
    }
    // SC_K consecutive sizes per thread and step (coalesced reads; a Fibonacci round
    // lists 10^5 big groups)
    __shared__ u64 s_carry;
    if (t == 0) s_carry = 0;
    __syncthreads();
    const u32 w = t / SAGPU_WAVE;
    bool bad = false;
    for (u64 b0 = 0; b0 < nb; b0 += 1024 * SC_K) {
        const u64 i0 = b0 + (u64)t * SC_K;
        u64 v[SC_K];
        u64 ts = 0;
#pragma unroll
        for (int q = 0; q < SC_K; q++) {
            v[q] = i0 + q < nb ? (u64)big_size[i0 + q] + 1 : 0ull;   // sizes are stored minus one
            bad = bad || (i0 + q < nb && v[q] <= (u64)RT_SMALL);
            ts += v[q];
        }
        const u64 is = wave_incl_sum64(ts);
        if (lane_id() == SAGPU_WAVE - 1) s_sum[w] = is;
        __syncthreads();
        u64 c = s_carry + is - ts;
        for (u32 q = 0; q < w; q++) c += s_sum[q];
#pragma unroll
        for (int q = 0; q < SC_K; q++) {
            if (i0 + q < nb) boff[i0 + q] = c;
            c += v[q];
        }
        __syncthreads();
        if (t == 1023) s_carry = c;
        __syncthreads();
    }
    if (t == 0) {
        boff[nb] = s_carry;
        *total = s_carry;
    }
    if (bad) atomicOr(err, 2u);
}

// Copies the members of the big groups into the sort buffers: comp = (dense group
// index << kb) | key2, val = the list entry.
template <typename KT, typename CT>
__global__ __launch_bounds__(256) void k_big_extract(const u32 *__restrict__ act_i, const KT *__restrict__ key2,
                                                     const u32 *__restrict__ bhead, const u64 *__restrict__ boff,
                                                     u32 nb, u64 total, int kb, CT *__restrict__ comp,
                                                     u32 *__restrict__ vals)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        u32 lo = 0, hi = nb - 1;
        while (lo < hi) {
            u32 mid = (lo + hi + 1) >> 1;
            if ((u64)boff[mid] <= e) lo = mid;
            else hi = mid - 1;
        }
        u64 k = (u64)bhead[lo] + (e - boff[lo]);
        comp[e] = (CT)(((u64)lo << kb) | (u64)key2[k]);
        vals[e] = act_i[k];
    }
}

// Flagged big groups without pivot classes (one thread per group): the rank the
// members hold (stored in SA[g + 1]) and the key2 of the subgroup whose slots hold
// it, the key of the group's member at sorted position rank - g (group q's members
// are sorted[boff[q], boff[q + 1]), keyed (dense group index << kb) | key2).
template <typename CT>
__global__ __launch_bounds__(256) void k_big_keep(const CT *__restrict__ sorted, const u64 *__restrict__ boff, int kb,
                                                  u32 nb, const u32 *__restrict__ gslot, u32 gf,
                                                  const u32 *__restrict__ SA, u32 *__restrict__ grank,
                                                  u64 *__restrict__ keepkey)
{
    const u32 q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nb || !(gslot[q] & gf)) return;
    const u32 g = gslot[q] & ~gf;
    const u32 v = SA[g + 1];
    const u64 size = boff[q + 1] - boff[q];
    grank[q] = v;
    keepkey[q] = (u64)(v - g) < size ? (u64)sorted[boff[q] + (v - g)] & ((1ull << kb) - 1) : BIG_NOKEY;
}

unsigned grid_for(u64 work, unsigned per_block, unsigned cap)
{
    u64 b = (work + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

} // namespace

int g_isa_window_log2 = 25;
int g_tile_range = 512;
int g_init_rows = 8;

template <typename K>
int group_init_t(WsScratch &ws, const K *keys, const u32 *vals, u64 n, int bits, int ksym, int gsa, u32 *SA, u32 *ISA,
                 u32 *act_i, u32 *act_g, u32 *LCP, u32 *PV, u32 *PG, hipStream_t s, const BwtSeed *seed = nullptr,
                 const PlcpHint *hint = nullptr)
{
    const int cb = seed ? seed->cb : 0;
    const int wlog = isa_bucket_log2(n);
    InitF<K> f{keys, vals, n, (1u << bits) - 1u, bits, ksym, gsa, SA, act_i, act_g, LCP, PV, PG, ws.bcnt, wlog,
               hint && hint->on && n <= (1ull << 28) ? 1 : 0, hint ? hint->off : 0, 0u, cb,
               seed ? seed->codes : nullptr, 65536u / (u32)bits + 1u};
    for (int i = 0; i < ksym; i++) f.lowbits |= 1u << (bits * i);
    // short texts: 512-position tiles (2 rows per wave), so more workgroups share them
    const bool small = n < INIT_SMALL_N;
    const int rows = small ? 2 : g_init_rows;
    const u64 tile = (u64)WS_WAVES * rows * SAGPU_WAVE;
    const u64 tiles = (n + tile - 1) / tile;
    if (tiles > ws.tiles_cap) {
        sagpu_set_error("segment scratch too small", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }
    const u64 nb = ((n - 1) >> wlog) + 1;
    if (nb * tiles > ws.bcnt_cap) {
        sagpu_set_error("bucket count scratch too small", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }
    Agg *tagg = reinterpret_cast<Agg *>(ws.tile_agg);
    if (rows == 2) hipLaunchKernelGGL((k_init_reduce<K, 2>), dim3((unsigned)tiles), dim3(WS_THREADS), 0, s, f, tagg);
    else if (rows == 8) hipLaunchKernelGGL((k_init_reduce<K, 8>), dim3((unsigned)tiles), dim3(WS_THREADS), 0, s, f, tagg);
    else hipLaunchKernelGGL((k_init_reduce<K, WS_ROWS>), dim3((unsigned)tiles), dim3(WS_THREADS), 0, s, f, tagg);
    SAGPU_TRY(ws_scan(ws, tiles, (u64 *)nullptr, s));
    SAGPU_TRY(scan_exclusive_u32(ws.scan, ws.bcnt, nb * tiles, s));
    ktime_begin(KI_INIT_EMIT);
    if (rows == 2)
        hipLaunchKernelGGL((k_init_emit<K, 2>), dim3((unsigned)tiles), dim3(WS_THREADS), 0, s, f, (const Agg *)tagg);
    else if (rows == 8)
        hipLaunchKernelGGL((k_init_emit<K, 8>), dim3((unsigned)tiles), dim3(WS_THREADS), 0, s, f, (const Agg *)tagg);
    else
        hipLaunchKernelGGL((k_init_emit<K, WS_ROWS>), dim3((unsigned)tiles), dim3(WS_THREADS), 0, s, f,
                           (const Agg *)tagg);
    // key and value read; SA, the (suffix, rank) pair, LCP mark and BWT code written
    // per position (the active-list entries are not counted)
    ktime_end(KI_INIT_EMIT, n * (sizeof(K) + 4 + 4 + 8 + (LCP ? 4 : 0) + (f.codes ? 1 : 0)));
    (void)ISA;   // the caller scatters the pairs (launch_isa_scatter or the windowed form)
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_isa_scatter(const u32 *PV, const u32 *PG, u64 m, u32 *ISA, hipStream_t s)
{
    if (m == 0) return SAGPU_OK;
    hipLaunchKernelGGL(k_isa_scatter, dim3(grid_for(m, 256 * IW, 65535)), dim3(256), 0, s, PV, PG, m, ISA);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

void set_isa_window_log2(int v) { g_isa_window_log2 = v; }

int isa_bucket_log2(u64 n)
{
    // buckets of 2^wlog entries, at most IB_MAX of them
    int wlog = g_isa_window_log2;
    while ((((n - 1) >> wlog) + 1) > (u64)IB_MAX) wlog++;
    return wlog;
}
void set_tile_range(int v) { g_tile_range = v; }

int group_init(WsScratch &ws, const u32 *keys, const u32 *vals, u64 n, int bits, int ksym, int gsa, u32 *SA, u32 *ISA,
               u32 *act_i, u32 *act_g, u32 *LCP, u32 *PV, u32 *PG, hipStream_t s, const BwtSeed *seed,
               const PlcpHint *hint)
{
    return group_init_t<u32>(ws, keys, vals, n, bits, ksym, gsa, SA, ISA, act_i, act_g, LCP, PV, PG, s, seed,
                             hint);
}

int group_init64(WsScratch &ws, const u64 *keys, const u32 *vals, u64 n, int bits, int ksym, u32 *SA, u32 *ISA,
                 u32 *act_i, u32 *act_g, u32 *LCP, u32 *PV, u32 *PG, hipStream_t s)
{
    return group_init_t<u64>(ws, keys, vals, n, bits, ksym, 1, SA, ISA, act_i, act_g, LCP, PV, PG, s);
}

template <typename KT>
int round_gather_t(const u32 *act_i, const u32 *act_g, u64 a, const u64 *rp, const u32 *ISA, u64 n, KT *key2,
                   hipStream_t s, const RawKey1 &rk, int cb)
{
    hipLaunchKernelGGL(k_round_gather<KT>, dim3(grid_for(a, 256 * IG, 65535)), dim3(256), 0, s, act_i, act_g, rp, ISA,
                       n, key2, rk, cb);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

// Segment capacity of each XCD for a list of a entries: the workgroups of XCD x are
// the tiles x, x + 8, ..., each emitting at most RANGE + RT_SMALL survivors.
u64 round_seg_cap(u64 a) { return (ceil_div_u(a, (u64)g_tile_range) / 8 + 1) * (u64)(g_tile_range + RT_SMALL); }

// Big-head segment capacity of each XCD for lists of at most n entries: heads of
// groups of more than RT_SMALL = 256 members lie at least 257 entries apart, so a
// workgroup of RANGE entries holds at most RANGE / 256 of them, and an XCD's
// ceil(ceil(n / RANGE) / 8) workgroups at most n / 2048 + RANGE / 128.
u64 round_bigseg_cap(u64 n) { return n / 2048 + 16; }

template <typename KT>
int round_tile_t(const u32 *act_i, const u32 *act_g, const KT *key2, u64 a, const u64 *rp, u32 *SA, u32 *ISA,
                 u32 *out_i, u32 *out_g, u64 *out_counter, u32 *err, u32 *big_head, u32 *big_g, u32 *big_size,
                 u32 *big_counter, u64 big_cap, u32 *LCP, hipStream_t s, const ActCode &ac, const ExactLcp &ex,
                 u32 *seg_i, u32 *seg_g, u64 *seg_ctr, u32 *bseg_h, u32 *bseg_g, u64 bseg_cap, u64 *boff,
                 u64 *big_total)
{
    const u64 seg_cap = round_seg_cap(a);
    SAGPU_TRY(launch_fill_u64(seg_ctr, 8 * 16, 0, s));
#define RT_LAUNCH(R)                                                                                              \
    hipLaunchKernelGGL((k_round_tile<KT, R>), dim3(ceil_div_u(a, R)), dim3(RT_THREADS), 0, s, act_i, act_g, key2, rp, \
                       SA, ISA, seg_i, seg_g, seg_ctr, seg_cap, err, bseg_h, bseg_g, bseg_cap, LCP, ac, ex)
    // the kernel alone is timed: list entry and key read (8 + sizeof(KT) B), about
    // 16 B written per entry (ranks, placed suffixes, marks and survivors)
    ktime_begin(KI_ROUND_TILE);
    if (g_tile_range == 256) RT_LAUNCH(256);
    else if (g_tile_range == 1024) RT_LAUNCH(1024);
    else RT_LAUNCH(512);
#undef RT_LAUNCH
    ktime_end(KI_ROUND_TILE, a * (24 + sizeof(KT)));
    hipLaunchKernelGGL(k_big_sizes, dim3(grid_for(big_cap, 256, 1024)), dim3(256), 0, s, bseg_h, bseg_g,
                       (const u64 *)seg_ctr, bseg_cap, big_head, big_g, big_counter, big_cap, SA, big_size, ac.gf);
    hipLaunchKernelGGL(k_big_offsets, dim3(1), dim3(1024), 0, s, big_size, big_counter, big_cap, boff, big_total, err);
    hipLaunchKernelGGL(k_seg_gather, dim3(grid_for(a, 256 * 4, 65535)), dim3(256), 0, s, seg_i, seg_g,
                       (const u64 *)seg_ctr, seg_cap, out_i, out_g, out_counter);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

template <typename KT, typename CT>
int big_extract_t(const u32 *act_i, const KT *key2, const u32 *bhead, const u64 *boff, u32 nb, u64 total, int kb,
                  CT *comp, u32 *vals, hipStream_t s)
{
    hipLaunchKernelGGL((k_big_extract<KT, CT>), dim3(grid_for(total, 256 * 4, 65535)), dim3(256), 0, s, act_i, key2,
                       bhead, boff, nb, total, kb, comp, vals);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int round_gather(const u32 *act_i, const u32 *act_g, u64 a, const u64 *rp, const u32 *ISA, u64 n, void *key2,
                 int wide, hipStream_t s, const RawKey1 *rk, int cb)
{
    const RawKey1 r = rk ? *rk : RawKey1{};
    return wide ? round_gather_t<u64>(act_i, act_g, a, rp, ISA, n, (u64 *)key2, s, r, cb)
                : round_gather_t<u32>(act_i, act_g, a, rp, ISA, n, (u32 *)key2, s, r, cb);
}

int round_tile(const u32 *act_i, const u32 *act_g, const void *key2, int wide, u64 a, const u64 *rp, u32 *SA,
               u32 *ISA, u32 *out_i, u32 *out_g, u64 *out_counter, u32 *err, u32 *big_head, u32 *big_g, u32 *big_size,
               u32 *big_counter, u64 big_cap, u32 *LCP, hipStream_t s, const ActCode &ac, const ExactLcp *ex,
               u32 *seg_i, u32 *seg_g, u64 *seg_ctr, u32 *bseg_h, u32 *bseg_g, u64 bseg_cap, u64 *boff,
               u64 *big_total)
{
    const ExactLcp e = ex ? *ex : ExactLcp{};
    return wide ? round_tile_t<u64>(act_i, act_g, (const u64 *)key2, a, rp, SA, ISA, out_i, out_g, out_counter, err,
                                    big_head, big_g, big_size, big_counter, big_cap, LCP, s, ac, e, seg_i, seg_g,
                                    seg_ctr, bseg_h, bseg_g, bseg_cap, boff, big_total)
                : round_tile_t<u32>(act_i, act_g, (const u32 *)key2, a, rp, SA, ISA, out_i, out_g, out_counter, err,
                                    big_head, big_g, big_size, big_counter, big_cap, LCP, s, ac, e, seg_i, seg_g,
                                    seg_ctr, bseg_h, bseg_g, bseg_cap, boff, big_total);
}

int big_extract(const u32 *act_i, const void *key2, int wide, const u32 *bhead, const u64 *boff, u32 nb, u64 total,
                int kb, int comp32, void *comp, u32 *vals, hipStream_t s)
{
    if (comp32)
        return wide ? big_extract_t<u64, u32>(act_i, (const u64 *)key2, bhead, boff, nb, total, kb, (u32 *)comp, vals, s)
                    : big_extract_t<u32, u32>(act_i, (const u32 *)key2, bhead, boff, nb, total, kb, (u32 *)comp, vals, s);
    return wide ? big_extract_t<u64, u64>(act_i, (const u64 *)key2, bhead, boff, nb, total, kb, (u64 *)comp, vals, s)
                : big_extract_t<u32, u64>(act_i, (const u32 *)key2, bhead, boff, nb, total, kb, (u64 *)comp, vals, s);
}

int big_emit(WsScratch &ws, const void *comp, int comp32, int kb, const u32 *vals, const u32 *gslot, u64 total, u64 h,
             u32 *SA, u32 *ISA, u32 *out_i, u32 *out_g, u64 *out_counter, u32 *LCP, hipStream_t s, const ActCode &ac,
             const u32 *grank, const u64 *keepkey, const ExactLcp *ex, const BigPivots *pv)
{
    u32 hh = h < LCP_BOUND_CAP ? (u32)h : LCP_BOUND_CAP;
    const ExactLcp e = ex ? *ex : ExactLcp{};
    const BigPivots p = pv ? *pv : BigPivots{};
    if (ac.gf && (!grank || !keepkey)) {
        sagpu_set_error("flagged groups need their ranks and kept keys", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }
    if (comp32) {
        BigF<u32> f{(const u32 *)comp, vals, gslot, total, hh, kb, SA, ISA, out_i, out_g, LCP, ac, e, p, grank,
                    keepkey};
        return ws_run(ws, f, total, out_counter, s);
    }
    BigF<u64> f{(const u64 *)comp, vals, gslot, total, hh, kb, SA, ISA, out_i, out_g, LCP, ac, e, p, grank, keepkey};
    return ws_run(ws, f, total, out_counter, s);
}

int big_keep(const void *sorted, int comp32, const u64 *boff, int kb, u32 nb, const u32 *gslot, u32 gf,
             const u32 *SA, u32 *grank, u64 *keepkey, hipStream_t s)
{
    if (nb == 0 || !gf) return SAGPU_OK;
    const unsigned grid = (nb + 255) / 256;
    if (comp32)
        hipLaunchKernelGGL(k_big_keep<u32>, dim3(grid), dim3(256), 0, s, (const u32 *)sorted, boff, kb, nb, gslot, gf,
                           SA, grank, keepkey);
    else
        hipLaunchKernelGGL(k_big_keep<u64>, dim3(grid), dim3(256), 0, s, (const u64 *)sorted, boff, kb, nb, gslot, gf,
                           SA, grank, keepkey);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}
