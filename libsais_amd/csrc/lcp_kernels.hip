// lcp_kernels.hip -- BWT gather, Phi, PLCP (irreducible positions + prefix max), LCP.
//
// PLCP[i] = lcp(T[i..], T[Phi[i]..]) where Phi[i] is the suffix that precedes i in
// SA (Phi[SA[0]] = n, giving 0).  When T[i-1] == T[Phi[i]-1] the position is
// "reducible": Phi[i-1] = Phi[i]-1 and PLCP[i] = PLCP[i-1]-1 (Kasai's bound holds
// with equality).  So only irreducible positions are compared character by
// character; with v[i] = PLCP[i] + i for irreducible i and 0 otherwise, v is
// non-decreasing over irreducible positions and PLCP[i] = max_{j<=i} v[j] - i.
// Position 0, positions whose predecessor is suffix 0 or the sentinel row, and (in
// GSA mode) positions preceded by a separator are always irreducible.
#include "engine.h"

namespace {

unsigned grid_for(u64 work, unsigned per_block, unsigned cap)
{
    u64 b = (work + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

// BWT with the libsais layout: U[0] = T[n-1]; row j (j != j0, SA[j0] == 0) goes to
// U[j + (j < j0)].  j0 = ISA[0].
template <typename ST>
__global__ __launch_bounds__(256) void k_bwt(const ST *__restrict__ T, const u32 *__restrict__ SA,
                                             const u32 *__restrict__ ISA, u64 n, ST *__restrict__ U)
{
    const u64 j0 = isa_rank(ISA[0], n);
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        u32 s = SA[j];
        if (j == j0) U[0] = T[n - 1];
        else U[j + (j < j0)] = (s - 1u < n) ? T[s - 1] : (ST)0;
    }
}

// BWT from the seeds: every stage that placed a suffix in SA slot j wrote the code
// of its preceding symbol to codes[j], so output byte p is inv[codes[j]] with
// j = p - 1 (p <= j0) or j = p (p > j0), and U[0] = T[n-1].  One sequential pass,
// four output bytes per lane.
__global__ __launch_bounds__(256) void k_bwt_seeded(const uint8_t *__restrict__ T, const u32 *__restrict__ ISA,
                                                    const uint8_t *__restrict__ codes,
                                                    const uint8_t *__restrict__ inv, u64 n, uint8_t *__restrict__ U,
                                                    int aligned)
{
    __shared__ uint8_t s_inv[256];
    s_inv[threadIdx.x] = inv[threadIdx.x];
    __syncthreads();
    const u64 j0 = isa_rank(ISA[0], n);
    const u64 quads = (n + 3) / 4;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 q = (u64)blockIdx.x * blockDim.x + threadIdx.x; q < quads; q += stride) {
        const u64 p0 = 4 * q;
        u32 word = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const u64 p = p0 + k;
            uint8_t b = 0;
            if (p == 0) b = T[n - 1];
            else if (p < n) b = s_inv[codes[p <= j0 ? p - 1 : p]];
            word |= (u32)b << (8 * k);
        }
        if (aligned && p0 + 4 <= n) {
            *reinterpret_cast<u32 *>(U + p0) = word;
        } else {
            for (int k = 0; k < 4 && p0 + k < n; k++) U[p0 + k] = (uint8_t)(word >> (8 * k));
        }
    }
}

// Text order: suffix i (i >= 1) sits in row j = ISA[i], whose BWT symbol T[i-1] goes
// to U[j + (j < j0)]; the primary row j0 = ISA[0] is dropped and U[0] = T[n-1].
__global__ __launch_bounds__(256) void k_bwt_scatter(const uint8_t *__restrict__ T, const u32 *__restrict__ ISA,
                                                     u64 n, uint8_t *__restrict__ U)
{
    const u64 j0 = isa_rank(ISA[0], n);
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (i == 0) {
            U[0] = T[n - 1];
            continue;
        }
        const u64 j = isa_rank(ISA[i], n);
        U[j + (j < j0)] = T[i - 1];
    }
}

__global__ __launch_bounds__(256) void k_phi(const u32 *__restrict__ SA, u64 n, u32 *__restrict__ PHI)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    // Entries outside [0, n) (an invalid SA from the caller) are skipped rather than
    // written out of bounds.  The first suffix has no predecessor: PHI points it at
    // itself, a value no real predecessor takes and one that fits for n = 2^32.
    for (u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        u32 p = SA[j];
        if (p < n) PHI[p] = j == 0 ? p : SA[j - 1];
    }
}

// Φ in two passes.  k_phi_bucket: per SA slot j, the pair (SA[j], SA[j-1]) goes to
// the bucket of its target SA[j] (rank inside the workgroup's share by an LDS atomic,
// the share reserved with one atomic per bucket).  A target outside [0, n) is
// skipped, and a bucket that receives more pairs than it has entries (a caller's SA
// with repeated values) drops the excess: Φ of an invalid SA is not defined, and
// nothing is written out of bounds.
constexpr int PB_ROWS = 16, PB_TILE = 256 * PB_ROWS, PB_MAXB = 64;
__global__ __launch_bounds__(256) void k_phi_bucket(const u32 *__restrict__ SA, u64 n, int wlog,
                                                    u32 *__restrict__ PV, u32 *__restrict__ PG,
                                                    u32 *__restrict__ bcur)
{
    __shared__ u32 s_cnt[PB_MAXB];
    __shared__ u64 s_base[PB_MAXB];
    if (threadIdx.x < PB_MAXB) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const u64 base = (u64)blockIdx.x * PB_TILE;
    // valid rows in a bit mask: every u32 is a suffix index when n = 2^32, so no
    // value can stand for "skip"
    u32 p[PB_ROWS], val[PB_ROWS], rk[PB_ROWS];
    u32 valid = 0;
#pragma unroll
    for (int r = 0; r < PB_ROWS; r++) {
        const u64 j = base + (u64)r * 256 + threadIdx.x;
        p[r] = 0;
        if (j < n) {
            const u32 x = SA[j];
            if ((u64)x < n) {
                p[r] = x;
                val[r] = j == 0 ? x : SA[j - 1];   // the first suffix points at itself
                rk[r] = atomicAdd(&s_cnt[(u32)((u64)x >> wlog)], 1u);
                valid |= 1u << r;
            }
        }
    }
    __syncthreads();
    const u32 nb = (u32)(((n - 1) >> wlog) + 1);
    if (threadIdx.x < nb && s_cnt[threadIdx.x]) s_base[threadIdx.x] = atomicAdd(&bcur[threadIdx.x], s_cnt[threadIdx.x]);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PB_ROWS; r++) {
        if (!((valid >> r) & 1u)) continue;
        const u32 b = (u32)((u64)p[r] >> wlog);
        const u64 lo = (u64)b << wlog;
        const u64 cap = n - lo < (1ull << wlog) ? n - lo : (1ull << wlog);
        const u64 pos = s_base[b] + rk[r];
        if (pos < cap) {
            PV[lo + pos] = p[r];
            PG[lo + pos] = val[r];
        }
    }
}

// k_phi_scatter: PHI[PV[e]] = PG[e] over the filled part of each bucket, one sweep
// over the bucket array in order: the workgroups in flight cover a few consecutive
// entries of one bucket, so the device's random writes stay in one stretch of PHI.
__global__ __launch_bounds__(256) void k_phi_scatter(const u32 *__restrict__ PV, const u32 *__restrict__ PG,
                                                     const u32 *__restrict__ bcur, u64 n, int wlog,
                                                     u32 *__restrict__ PHI)
{
    const u64 chunk = 256ull * 8;
    for (u64 e0 = (u64)blockIdx.x * chunk; e0 < n; e0 += (u64)gridDim.x * chunk) {
        u32 v[8], g[8];
        bool ok[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const u64 e = e0 + (u64)k * 256 + threadIdx.x;
            const u64 b = e >> wlog;
            ok[k] = e < n && (e - (b << wlog)) < (u64)bcur[b];   // filled part of bucket b
            v[k] = ok[k] ? PV[e] : 0u;
            g[k] = ok[k] ? PG[e] : 0u;
        }
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (ok[k]) PHI[v[k]] = g[k];
    }
}

// 8 bytes starting at an arbitrary address; the text buffer carries >= 16 bytes of
// zero padding so both aligned words are in bounds.
__device__ __forceinline__ u64 load8u(const uint8_t *p)
{
    uintptr_t a = (uintptr_t)p;
    const u64 *q = reinterpret_cast<const u64 *>(a & ~(uintptr_t)7);
    u32 sh = (u32)(a & 7) * 8;
    u64 lo = q[0];
    if (sh == 0) return lo;
    u64 hi = q[1];
    return (lo >> sh) | (hi << (64 - sh));
}

// Lanes of x that hold a zero symbol (W-byte symbols, W = 1 or 2; for W = 4 GSA is not
// used).  Only the lowest flagged lane is exact, which is the one used.
template <int W>
__device__ __forceinline__ u64 zero_lanes(u64 x)
{
    if (W == 2) return (x - 0x0001000100010001ull) & ~x & 0x8000800080008000ull;
    return (x - 0x0101010101010101ull) & ~x & 0x8080808080808080ull;
}

// Compare up to `limit` more bytes from offset l.  Returns the new l; *done is set
// when the comparison ended (mismatch, separator, or bound).
template <int W>
__device__ __forceinline__ u64 match_bytes(const uint8_t *T, u64 i, u64 k, u64 l, u64 m, int gsa, u64 limit,
                                           bool *done)
{
    const u64 stop_at = l + limit;
    while (true) {
        if (l >= m) { *done = true; return m; }
        u64 x = load8u(T + i + l), y = load8u(T + k + l);
        u64 stop = x ^ y;
        if (gsa) stop |= zero_lanes<W>(x);
        if (stop) {
            l += (u64)(__builtin_ctzll(stop) >> 3);
            *done = true;
            return l < m ? l : m;
        }
        l += 8;
        if (l >= stop_at) { *done = (l >= m); return l < m ? l : m; }
    }
}

constexpr u64 PL_THREAD_LIMIT = 64;

// Per position: decide reducibility, compare irreducible ones up to 64 bytes, queue
// the rest.  Writes v[i] (PLCP[i] + i, or 0 for reducible positions) into PLCP.
// W = bytes per symbol: 1 for byte strings, 4 for int32 strings (libsais_plcp_int),
// whose symbols are compared as W-byte groups of the same byte comparison.
template <typename VT, int W>
__global__ __launch_bounds__(256) void k_plcp_irr(const uint8_t *__restrict__ T, u64 n, int gsa,
                                                  const u32 *__restrict__ PHI, const u32 *__restrict__ ISA,
                                                  const u32 *__restrict__ SA, VT *__restrict__ V,
                                                  u32 *__restrict__ queue, u32 *__restrict__ queue_l,
                                                  u32 *__restrict__ queue_n, u64 queue_cap,
                                                  const u32 *__restrict__ marks)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        u64 lb = 0;
        u64 k;
        if (PHI) {
            k = PHI[i];
            if (k == i) k = n;  // first suffix in SA order
        } else {
            u32 r = isa_rank(ISA[i], n);
            if (marks) {
                // construction already knows most lcps: read it where i sits in SA
                u32 p = r < n ? marks[r] : 0u;
                if (!(p & LCP_UNKNOWN)) {
                    V[i] = (VT)(p + i);
                    continue;
                }
                lb = p & ~LCP_UNKNOWN;
            }
            k = r == 0 ? n : SA[r - 1];
        }
        if (k > n) k = n;
        bool irr = (i == 0) || (k == 0) || (k >= n);
        if (!irr) {
            if (W == 1) {
                uint8_t a = T[i - 1], b = T[k - 1];
                irr = (a != b) || (gsa && a == 0);
            } else if (W == 2) {
                uint16_t a = *(const uint16_t *)(T + 2 * (i - 1)), b = *(const uint16_t *)(T + 2 * (k - 1));
                irr = (a != b) || (gsa && a == 0);
            } else {
                irr = *(const u32 *)(T + W * (i - 1)) != *(const u32 *)(T + W * (k - 1));
            }
        }
        if (!irr) { V[i] = 0; continue; }
        if (k >= n) { V[i] = (VT)i; continue; }
        u64 m = n - (i > k ? i : k); // also bounds GSA matches (they stop earlier)
        bool done = false;
        u64 l = match_bytes<W>(T, W * i, W * k, W * (lb < m ? lb : m), W * m, gsa, W * PL_THREAD_LIMIT, &done) / W;
        if (done) {
            V[i] = (VT)(l + i);
        } else {
            u32 slot = atomicAdd(queue_n, 1u);
            if (slot < queue_cap) {
                queue[slot] = (u32)i;
                queue_l[slot] = (u32)l;
                V[i] = (VT)(l + i); // provisional; overwritten by k_plcp_long
            } else {
                // queue full: finish here (correct, only slower)
                l = match_bytes<W>(T, W * i, W * k, W * l, W * m, gsa, ~0ull >> 1, &done) / W;
                V[i] = (VT)(l + i);
            }
        }
    }
}

// Per SA slot j >= 1, what the PLCP pass needs about the pair (SA[j-1], SA[j]):
//   known lcp:  the value (bit 63 clear)
//   unknown:    bit 63 | (floor(log2 bound) + 1, or 0) << 58 | T[k-1] << 32 | k,
//               k = SA[j-1]; T[k-1] is the BWT symbol of slot j-1, taken from U (the
//               BWT in output layout) when given.
// Built in SA order with sequential reads, so the text-order pass gathers one
// 8-byte word per position instead of the mark, SA[j-1] and T[k-1] separately.
// Prefix max over V (PLCP[i] + i) in tiles of MS_TILE positions.
constexpr int MS_THREADS = 256, MS_ITEMS = 16, MS_TILE = MS_THREADS * MS_ITEMS;

template <typename VT>
__device__ __forceinline__ VT vmax(VT a, VT b) { return a > b ? a : b; }

// 64 stream bits from bit `bitpos` (the stream has a zero word past every window).
__device__ __forceinline__ u64 stream_window(const u32 *__restrict__ pw, u64 bitpos)
{
    const u64 w = bitpos >> 5;
    const u32 off = (u32)(bitpos & 31);
    const u64 hi = ((u64)pw[w] << 32) | (u64)pw[w + 1];
    return off ? (hi << off) | ((u64)pw[w + 2] >> (32 - off)) : hi;
}

// lcp of suffixes i and k on the packed code stream, from l (a known lower bound)
// up to m (the shorter suffix's length): 64 / bits symbols per step.  Stops after
// `limit` symbols with *done = false, as match_bytes.
__device__ __forceinline__ u64 match_stream(const u32 *__restrict__ pw, int bits, u64 i, u64 k, u64 l, u64 m,
                                            u64 limit, bool *done)
{
    const u32 per = 64u / (u32)bits;
    const u64 keep = ~0ull << (64 - per * (u32)bits);
    const u64 stop_at = l + limit;
    while (true) {
        if (l >= m) { *done = true; return m; }
        const u64 d = (stream_window(pw, (i + l) * (u64)bits) ^ stream_window(pw, (k + l) * (u64)bits)) & keep;
        if (d) {
            l += (u64)(__clzll(d) / bits);
            *done = true;
            return l < m ? l : m;
        }
        l += per;
        if (l >= stop_at) { *done = (l >= m); return l < m ? l : m; }
    }
}

// Byte texts, pair-word form (k_plcp_pairs): per position one ISA read, one 8-byte
// pair read and the symbol before it; known lcps and reducible positions finish
// there, irreducible ones compare as k_plcp_irr.  PI positions per thread and step
// with their loads issued together.
constexpr int PI = 4;
// Each workgroup takes whole MS_TILE tiles (PI rows of 256 per step) and writes the
// tile's max of v to tmax, so the prefix max needs no separate reduce pass
// (k_plcp_long raises the entries of the positions it finishes).
template <typename VT>
__global__ __launch_bounds__(256) void k_plcp_irr_pairs(const uint8_t *__restrict__ T, u64 n, int gsa,
                                                        const u32 *__restrict__ ISA, const u64 *__restrict__ pairs,
                                                        VT *__restrict__ V, u32 *__restrict__ queue,
                                                        u32 *__restrict__ queue_l, u32 *__restrict__ queue_n,
                                                        u64 queue_cap, const u32 *__restrict__ pw, int bits,
                                                        VT *__restrict__ tmax, int hint_off, int split)
{
    // split > 1 (short texts): a tile's steps go to `split` workgroups, which raise
    // the tile's maximum (zeroed first) with atomics instead of storing it
    constexpr u64 STEP = 256ull * PI;
    static_assert(MS_TILE % STEP == 0, "tile must be whole steps");
    __shared__ VT s_m[4];
    const u64 tiles = (n + MS_TILE - 1) / MS_TILE;
    const u64 part_len = MS_TILE / (u64)split;
    for (u64 wi = blockIdx.x; wi < tiles * (u64)split; wi += gridDim.x) {
        const u64 tile = wi / (u64)split, p0 = tile * MS_TILE + (wi % (u64)split) * part_len;
        VT mx = 0;
        for (u64 base = p0; base < p0 + part_len && base < n; base += STEP) {
            u32 r[PI], hint[PI];
            u64 w[PI];
            uint8_t a[PI];
#pragma unroll
            for (int q = 0; q < PI; q++) {
                const u64 i = base + (u64)q * 256 + threadIdx.x;
                const u32 x = i < n ? ISA[i] : 0u;
                r[q] = isa_rank(x, n);
                hint[q] = isa_hint(x, n);   // 1 + PLCP[i] known from the initial grouping
                a[q] = (i > 0 && i < n) ? T[i - 1] : 0;
            }
#pragma unroll
            for (int q = 0; q < PI; q++) w[q] = (r[q] && !hint[q]) ? pairs[r[q]] : 0ull;
#pragma unroll
            for (int q = 0; q < PI; q++) {
                const u64 i = base + (u64)q * 256 + threadIdx.x;
                if (i >= n) continue;
                VT v;
                if (hint[q]) {
                    v = (VT)((u64)hint_off + hint[q] - 1 + i);
                } else if (r[q] == 0) {   // first suffix in SA order: lcp 0
                    v = (VT)i;
                } else if (!(w[q] >> 63)) {
                    v = (VT)((u64)(u32)w[q] + i);
                } else {
                    const u64 k = (u32)w[q];
                    const u32 lg = (u32)(w[q] >> 58) & 31u;
                    const u64 lb = lg ? (1ull << (lg - 1)) : 0;
                    if (i > 0 && k > 0 && a[q] == (uint8_t)(w[q] >> 32) && !(gsa && a[q] == 0)) {
                        v = 0;   // reducible: PLCP[i] = PLCP[i-1] - 1
                    } else {
                        const u64 m = n - (i > k ? i : k);
                        bool done = false;
                        // with the code stream (no GSA), 64 / bits symbols per step
                        u64 l = pw ? match_stream(pw, bits, i, k, lb < m ? lb : m, m, PL_THREAD_LIMIT, &done)
                                   : match_bytes<1>(T, i, k, lb < m ? lb : m, m, gsa, PL_THREAD_LIMIT, &done);
                        if (!done) {
                            u32 slot = atomicAdd(queue_n, 1u);
                            if (slot < queue_cap) {
                                queue[slot] = (u32)i;   // provisional value; k_plcp_long finishes it
                                queue_l[slot] = (u32)l;
                            } else {
                                l = match_bytes<1>(T, i, k, l, m, gsa, ~0ull >> 1, &done);
                            }
                        }
                        v = (VT)(l + i);
                    }
                }
                V[i] = v;
                mx = vmax(mx, v);
            }
        }
        for (int d = 32; d >= 1; d >>= 1) mx = vmax(mx, (VT)__shfl_xor(mx, d, SAGPU_WAVE));
        if (lane_id() == 0) s_m[threadIdx.x / SAGPU_WAVE] = mx;
        __syncthreads();
        if (threadIdx.x == 0) {
            const VT m = vmax(vmax(s_m[0], s_m[1]), vmax(s_m[2], s_m[3]));
            if (split == 1) tmax[tile] = m;
            else atomicMax(&tmax[tile], m);
        }
        __syncthreads();
    }
}

constexpr int PP = 8;   // slots per thread and step
__global__ __launch_bounds__(256) void k_plcp_pairs(const u32 *__restrict__ SA, const u32 *__restrict__ marks,
                                                    const uint8_t *__restrict__ U, const uint8_t *__restrict__ T,
                                                    const u32 *__restrict__ ISA, u64 n, u64 *__restrict__ pairs)
{
    const u64 j0 = isa_rank(ISA[0], n);
    const u64 chunk = 256ull * PP;
    for (u64 base = (u64)blockIdx.x * chunk; base < n; base += (u64)gridDim.x * chunk) {
        u32 m[PP], k[PP];
        uint8_t b[PP];
#pragma unroll
        for (int q = 0; q < PP; q++) {
            const u64 j = base + (u64)q * 256 + threadIdx.x;
            m[q] = (j > 0 && j < n) ? marks[j] : 0u;
        }
#pragma unroll
        for (int q = 0; q < PP; q++) {
            const u64 j = base + (u64)q * 256 + threadIdx.x;
            k[q] = (m[q] & LCP_UNKNOWN) ? SA[j - 1] : 0u;
        }
#pragma unroll
        for (int q = 0; q < PP; q++) {
            const u64 j = base + (u64)q * 256 + threadIdx.x;
            b[q] = 0;
            if ((m[q] & LCP_UNKNOWN) && k[q] != 0) {
                const u64 r = j - 1;
                b[q] = U ? U[r + (r < j0)] : T[k[q] - 1];
            }
        }
#pragma unroll
        for (int q = 0; q < PP; q++) {
            const u64 j = base + (u64)q * 256 + threadIdx.x;
            if (j >= n) continue;
            u64 w = m[q];
            if (m[q] & LCP_UNKNOWN) {
                const u32 lb = m[q] & ~LCP_UNKNOWN;
                const u32 lg = lb ? 32u - (u32)__clz(lb) : 0u;
                w = (1ull << 63) | ((u64)lg << 58) | ((u64)b[q] << 32) | (u64)k[q];
            }
            pairs[j] = w;
        }
    }
}

// One workgroup per queued position: 256 lanes x LW words of 8 bytes per step, first
// mismatch found with a workgroup min-reduction (byte offsets; W bytes per symbol).
// A position still matching after LONG_STEPS steps goes to the very-long queue
// (vq: its position and the offset reached), whose matches many workgroups compare
// together (k_plcp_vlong): a text with a few positions of lcp near n would
// otherwise keep one workgroup stepping through megabytes while the rest idle.
constexpr int LW = 8;
constexpr u64 LONG_STEP = 256ull * 8 * LW;   // bytes per workgroup step
constexpr int LONG_STEPS = 32;
constexpr u32 VQ_CAP = 256;                  // very-long queue entries
struct VLongQueue {
    u32 *n;       // [1] entries
    u32 *pos;     // [VQ_CAP] text position
    u64 *off;     // [VQ_CAP] byte offset matched so far
    u64 *res;     // [VQ_CAP] first mismatch found (atomic min), ~0 while none
};

// The smallest byte offset in [l, l + LONG_STEP) where the suffixes at i and k
// differ (or, in GSA mode, hold a separator), capped at m; ~0 if none.  Called by
// all threads of the workgroup.
template <int W>
__device__ u64 long_step(const uint8_t *__restrict__ T, int gsa, u64 i, u64 k, u64 l, u64 m, u64 *s_min)
{
    u64 first = ~0ull;
#pragma unroll
    for (int u = 0; u < LW; u++) {
        const u64 off = l + ((u64)u * 256 + threadIdx.x) * 8;
        u64 f = m;
        if (off < m) {
            const u64 x = load8u(T + W * i + off), y = load8u(T + W * k + off);
            u64 stop = x ^ y;
            if (gsa) stop |= zero_lanes<W>(x);
            f = ~0ull;
            if (stop) {
                f = off + (u64)(__builtin_ctzll(stop) >> 3);
                if (f > m) f = m;
            }
        }
        first = f < first ? f : first;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        const u64 o = __shfl_xor(first, d, SAGPU_WAVE);
        first = first < o ? first : o;
    }
    if (lane_id() == 0) s_min[threadIdx.x / SAGPU_WAVE] = first;
    __syncthreads();
    u64 r = s_min[0];
    for (int w = 1; w < 4; w++) r = r < s_min[w] ? r : s_min[w];
    __syncthreads();   // s_min is reused by the next step
    return r;
}

// The suffix that precedes suffix i in SA order (n for the first), as Φ gives it.
__device__ __forceinline__ u64 phi_of(const u32 *__restrict__ PHI, const u32 *__restrict__ ISA,
                                      const u32 *__restrict__ SA, u64 n, u64 i)
{
    u64 k;
    if (PHI) {
        k = PHI[i];
        if (k == i) k = n;
    } else {
        const u32 r = isa_rank(ISA[i], n);
        k = r == 0 ? n : SA[r - 1];
    }
    return k > n ? n : k;
}

template <typename VT, int W>
__global__ __launch_bounds__(256) void k_plcp_long(const uint8_t *__restrict__ T, u64 n, int gsa,
                                                   const u32 *__restrict__ PHI, const u32 *__restrict__ ISA,
                                                   const u32 *__restrict__ SA, VT *__restrict__ V,
                                                   const u32 *__restrict__ queue, const u32 *__restrict__ queue_l,
                                                   const u32 *__restrict__ queue_n, u64 queue_cap,
                                                   VT *__restrict__ tmax, VLongQueue vq)
{
    __shared__ u64 s_min[4];
    __shared__ u32 s_slot;
    const u32 cnt = (u32)((u64)*queue_n < queue_cap ? (u64)*queue_n : queue_cap);
    for (u32 q = blockIdx.x; q < cnt; q += gridDim.x) {
        const u64 i = queue[q];
        const u64 k = phi_of(PHI, ISA, SA, n, i);
        const u64 m = W * (n - (i > k ? i : k));
        u64 l = W * (u64)queue_l[q];
        for (int step = 0;; step++) {
            if (step == LONG_STEPS && vq.n) {
                // hand the rest of the match to the very-long pass (when it has room)
                if (threadIdx.x == 0) {
                    const u32 slot = atomicAdd(vq.n, 1u);
                    if (slot < VQ_CAP) {
                        vq.pos[slot] = (u32)i;
                        vq.off[slot] = l;
                        vq.res[slot] = ~0ull;
                    }
                    s_slot = slot;
                }
                __syncthreads();
                const u32 slot = s_slot;
                __syncthreads();
                if (slot < VQ_CAP) break;
            }
            const u64 r = long_step<W>(T, gsa, i, k, l, m, s_min);
            if (r != ~0ull) {
                if (threadIdx.x == 0) {
                    V[i] = (VT)(r / W + i);
                    // tile maxima written by the pass that queued i (pair-word form)
                    if (tmax) atomicMax(&tmax[i / MS_TILE], (VT)(r / W + i));
                }
                break;
            }
            l += LONG_STEP;
        }
    }
}

// The very-long matches: the workgroups split evenly over the queued positions;
// workgroup g of a position compares its steps g, g + G, g + 2G, ... in order until
// a mismatch at or before its step is known (atomic min of the first mismatch).
template <int W>
__global__ __launch_bounds__(256) void k_plcp_vlong(const uint8_t *__restrict__ T, u64 n, int gsa,
                                                    const u32 *__restrict__ PHI, const u32 *__restrict__ ISA,
                                                    const u32 *__restrict__ SA, VLongQueue vq)
{
    __shared__ u64 s_min[4];
    __shared__ u64 s_res;
    const u32 cnt = *vq.n < VQ_CAP ? *vq.n : VQ_CAP;
    if (cnt == 0) return;
    const u32 G = gridDim.x / cnt > 0 ? gridDim.x / cnt : 1u;
    const u32 g = blockIdx.x % G;
    for (u32 v = blockIdx.x / G; v < cnt; v += (gridDim.x / G > 0 ? gridDim.x / G : 1u)) {
        const u64 i = vq.pos[v];
        const u64 k = phi_of(PHI, ISA, SA, n, i);
        const u64 m = W * (n - (i > k ? i : k));
        for (u64 l = vq.off[v] + (u64)g * LONG_STEP;; l += (u64)G * LONG_STEP) {
            // Whether a mismatch before this step is known, read once for the whole
            // workgroup: waves reading it separately can see different values (other
            // workgroups lower it meanwhile), and a wave that left the loop would
            // leave its s_min slot stale for the others' step (uninitialised on the
            // first, so any value, 0 included, could win the min).  The second
            // barrier keeps thread 0 from rewriting s_res (next step or next
            // position) before every thread has read it.
            if (threadIdx.x == 0) s_res = *(volatile u64 *)&vq.res[v];
            __syncthreads();
            const u64 known = s_res;
            __syncthreads();
            if (known <= l) break;
            const u64 r = long_step<W>(T, gsa, i, k, l, m, s_min);
            if (r != ~0ull) {
                if (threadIdx.x == 0) atomicMin((unsigned long long *)&vq.res[v], (unsigned long long)r);
                break;
            }
        }
    }
}

// V (and the tile maxima) of the very-long matches.
template <typename VT, int W>
__global__ __launch_bounds__(256) void k_plcp_vlong_fin(VT *__restrict__ V, VT *__restrict__ tmax, VLongQueue vq)
{
    const u32 cnt = *vq.n < VQ_CAP ? *vq.n : VQ_CAP;
    const u32 v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= cnt) return;
    const u64 i = vq.pos[v];
    const VT x = (VT)(vq.res[v] / W + i);
    V[i] = x;
    if (tmax) atomicMax(&tmax[i / MS_TILE], x);
}

// Prefix max over V, then PLCP[i] = max - i.  VT = u32 works in place (V is the
// PLCP array); VT = u64 is used when PLCP[i] + i can reach 2^32.


template <typename VT>
__device__ __forceinline__ VT wave_incl_max_t(VT v)
{
#pragma unroll
    for (int d = 1; d < SAGPU_WAVE; d <<= 1) {
        VT o = __shfl_up(v, d, SAGPU_WAVE);
        if ((int)lane_id() >= d) v = vmax(v, o);
    }
    return v;
}

template <typename VT>
__global__ __launch_bounds__(MS_THREADS) void k_max_reduce(const VT *__restrict__ V, u64 n, VT *__restrict__ tmax)
{
    __shared__ VT s_w[4];
    const u64 base = (u64)blockIdx.x * MS_TILE;
    VT m = 0;
    for (int j = 0; j < MS_ITEMS; j++) {
        u64 e = base + (u64)j * MS_THREADS + threadIdx.x;
        if (e < n) m = vmax(m, V[e]);
    }
    for (int d = 32; d >= 1; d >>= 1) m = vmax(m, (VT)__shfl_xor(m, d, SAGPU_WAVE));
    if (lane_id() == 0) s_w[threadIdx.x / SAGPU_WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        VT r = s_w[0];
        for (int w = 1; w < 4; w++) r = vmax(r, s_w[w]);
        tmax[blockIdx.x] = r;
    }
}

// Exclusive max-scan of the tile maxima in steps of 8192 tiles (MX_K consecutive
// tiles per thread of a 1024-thread workgroup).  Up to one step, one workgroup scans
// them (k_max_scan_tiles); beyond, each step is a chunk: k_max_chunk_reduce takes each
// chunk's max, k_max_scan_tiles scans those, and k_max_chunk_apply scans every chunk
// from its prefix, all at once (65536 tiles of a 2^28-position text: 61 us in one
// workgroup).
constexpr int MX_K = 8, MX_STEP = 1024 * MX_K;

// One step: tiles [base, base + MX_STEP) replaced by the max of everything before
// each (from `carry`); returns the max through the step's last tile.
template <typename VT>
__device__ VT max_step(VT *tmax, u64 tiles, u64 base, VT carry, VT *s_w, VT *s_end)
{
    const u64 i0 = base + (u64)threadIdx.x * MX_K;
    VT v[MX_K];
    VT t = 0;
#pragma unroll
    for (int q = 0; q < MX_K; q++) {
        v[q] = i0 + q < tiles ? tmax[i0 + q] : (VT)0;
        t = vmax(t, v[q]);
    }
    const VT inc = wave_incl_max_t(t);
    if (lane_id() == SAGPU_WAVE - 1) s_w[threadIdx.x / SAGPU_WAVE] = inc;
    __syncthreads();
    VT c = carry;
    for (u32 q = 0; q < threadIdx.x / SAGPU_WAVE; q++) c = vmax(c, s_w[q]);
    const VT prev = __shfl_up(inc, 1, SAGPU_WAVE);
    if (lane_id() > 0) c = vmax(c, prev);
#pragma unroll
    for (int q = 0; q < MX_K; q++) {   // exclusive max of the tiles before i0 + q
        if (i0 + q < tiles) tmax[i0 + q] = c;
        c = vmax(c, v[q]);
    }
    if (threadIdx.x == 1023) *s_end = c;   // the last thread's running max covers the step
    __syncthreads();
    const VT e = *s_end;
    __syncthreads();   // s_w and s_end are free for the next step
    return e;
}

template <typename VT>
__global__ __launch_bounds__(1024) void k_max_scan_tiles(VT *tmax, u64 tiles)
{
    __shared__ VT s_w[16];
    __shared__ VT s_end;
    VT carry = 0;
    for (u64 base = 0; base < tiles; base += MX_STEP) carry = max_step(tmax, tiles, base, carry, s_w, &s_end);
}

template <typename VT>
__global__ __launch_bounds__(1024) void k_max_chunk_reduce(const VT *__restrict__ tmax, u64 tiles,
                                                           VT *__restrict__ cmax)
{
    __shared__ VT s_w[16];
    const u64 b0 = (u64)blockIdx.x * MX_STEP;
    VT t = 0;
#pragma unroll
    for (int q = 0; q < MX_K; q++) {
        const u64 i = b0 + (u64)q * 1024 + threadIdx.x;
        if (i < tiles) t = vmax(t, tmax[i]);
    }
    t = wave_incl_max_t(t);
    if (lane_id() == SAGPU_WAVE - 1) s_w[threadIdx.x / SAGPU_WAVE] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        VT c = 0;
        for (int q = 0; q < 16; q++) c = vmax(c, s_w[q]);
        cmax[blockIdx.x] = c;
    }
}

template <typename VT>
__global__ __launch_bounds__(1024) void k_max_chunk_apply(VT *tmax, u64 tiles, const VT *__restrict__ cmax)
{
    __shared__ VT s_w[16];
    __shared__ VT s_end;
    max_step(tmax, tiles, (u64)blockIdx.x * MX_STEP, cmax[blockIdx.x], s_w, &s_end);
}

// Elements 4*lane .. 4*lane+3 of a 256-element row: one 16-byte (u32) or two
// 16-byte (u64) loads per lane, or element loads (0 past n) in the last tile.
template <typename VT>
__device__ __forceinline__ void load4(const VT *V, u64 e0, u64 n, bool full, VT (&x)[4])
{
    if (full) {
        if constexpr (sizeof(VT) == 4) {
            uint4 q = *reinterpret_cast<const uint4 *>(V + e0);
            x[0] = q.x; x[1] = q.y; x[2] = q.z; x[3] = q.w;
        } else {
            ulonglong2 a = *reinterpret_cast<const ulonglong2 *>(V + e0);
            ulonglong2 b = *reinterpret_cast<const ulonglong2 *>(V + e0 + 2);
            x[0] = a.x; x[1] = a.y; x[2] = b.x; x[3] = b.y;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) x[k] = e0 + k < n ? V[e0 + k] : (VT)0;
    }
}

// Each wave owns 1024 consecutive elements of the 4096-element tile, as 4 rows of
// 256 (4 consecutive per lane), so loads and stores are 16-byte and coalesced.
template <typename VT>
__global__ __launch_bounds__(MS_THREADS) void k_max_apply(const VT *V, u64 n, const VT *__restrict__ tmax,
                                                          u32 *PLCP)
{
    constexpr int ROWS = MS_TILE / (4 * 256);   // rows of 256 per wave
    __shared__ VT s_w[MS_THREADS / SAGPU_WAVE];
    const u32 w = threadIdx.x / SAGPU_WAVE, lane = lane_id();
    const u64 wbase = (u64)blockIdx.x * MS_TILE + (u64)w * (MS_TILE / (MS_THREADS / SAGPU_WAVE));
    const bool full = wbase + (u64)ROWS * 256 <= n;
    VT v[ROWS][4], inc[ROWS];
    VT tot = 0;
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        load4(V, wbase + (u64)r * 256 + 4 * lane, n, full, v[r]);
        VT m = vmax(vmax(v[r][0], v[r][1]), vmax(v[r][2], v[r][3]));
        inc[r] = wave_incl_max_t(m);
        tot = vmax(tot, (VT)__shfl(inc[r], SAGPU_WAVE - 1, SAGPU_WAVE));
    }
    if (lane == 0) s_w[w] = tot;
    __syncthreads();
    VT c = tmax[blockIdx.x];
    for (u32 q = 0; q < w; q++) c = vmax(c, s_w[q]);
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        const u64 e0 = wbase + (u64)r * 256 + 4 * lane;
        VT prev = __shfl_up(inc[r], 1, SAGPU_WAVE);
        VT cl = lane > 0 ? vmax(c, prev) : c;
        u32 o[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            cl = vmax(cl, v[r][k]);
            o[k] = (u32)(cl - (VT)(e0 + k));
        }
        if (full) {
            *reinterpret_cast<uint4 *>(PLCP + e0) = make_uint4(o[0], o[1], o[2], o[3]);
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (e0 + k < n) PLCP[e0 + k] = o[k];
        }
        c = vmax(c, (VT)__shfl(inc[r], SAGPU_WAVE - 1, SAGPU_WAVE));
    }
}

__global__ __launch_bounds__(256) void k_lcp(const u32 *__restrict__ PLCP, const u32 *__restrict__ SA, u64 n,
                                             u32 *__restrict__ LCP)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        u32 p = SA[j];
        LCP[j] = p < n ? PLCP[p] : 0u;
    }
}

} // namespace

int launch_bwt(const uint8_t *T, const u32 *SA, const u32 *ISA, u64 n, uint8_t *U, hipStream_t s)
{
    hipLaunchKernelGGL(k_bwt<uint8_t>, dim3(grid_for(n, 256 * 8, 65535)), dim3(256), 0, s, T, SA, ISA, n, U);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_bwt_seeded(const uint8_t *T, const u32 *SA, const u32 *ISA, const BwtSeed &seed, u64 n, uint8_t *U,
                      hipStream_t s)
{
    const int aligned = ((uintptr_t)U & 3) == 0;
    (void)SA;
    hipLaunchKernelGGL(k_bwt_seeded, dim3(grid_for((n + 3) / 4, 256 * 4, 65535)), dim3(256), 0, s, T, ISA, seed.codes,
                       seed.inv, n, U, aligned);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_bwt16(const uint16_t *T, const u32 *SA, const u32 *ISA, u64 n, uint16_t *U, hipStream_t s)
{
    hipLaunchKernelGGL(k_bwt<uint16_t>, dim3(grid_for(n, 256 * 8, 65535)), dim3(256), 0, s, T, SA, ISA, n, U);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_bwt_scatter(const uint8_t *T, const u32 *ISA, u64 n, uint8_t *U, hipStream_t s)
{
    hipLaunchKernelGGL(k_bwt_scatter, dim3(grid_for(n, 256 * 8, 65535)), dim3(256), 0, s, T, ISA, n, U);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_phi_bucketed(const u32 *SA, u64 n, u32 *PHI, u32 *PV, u32 *PG, u32 *bcur, int wlog, hipStream_t s)
{
    if ((((n - 1) >> wlog) + 1) > (u64)PB_MAXB) {
        sagpu_set_error("phi: too many buckets", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }
    SAGPU_TRY(launch_fill_u32(bcur, PB_MAXB, 0, s));
    hipLaunchKernelGGL(k_phi_bucket, dim3((unsigned)((n + PB_TILE - 1) / PB_TILE)), dim3(256), 0, s, SA, n, wlog, PV,
                       PG, bcur);
    hipLaunchKernelGGL(k_phi_scatter, dim3(grid_for(n, 256 * 8, 65535)), dim3(256), 0, s, PV, PG, bcur, n, wlog, PHI);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_phi(const u32 *SA, u64 n, u32 *PHI, hipStream_t s)
{
    hipLaunchKernelGGL(k_phi, dim3(grid_for(n, 256 * 8, 65535)), dim3(256), 0, s, SA, n, PHI);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

namespace {

constexpr int LB = 8;   // entries per thread and step of the gather kernels below
__global__ __launch_bounds__(256) void k_lcp_fix(const u32 *__restrict__ PLCP, const u32 *__restrict__ SA, u64 n,
                                                 u32 *__restrict__ LCP)
{
    const u64 chunk = 256ull * LB;
    for (u64 base = (u64)blockIdx.x * chunk; base < n; base += (u64)gridDim.x * chunk) {
        u32 m[LB], p[LB], v[LB];
#pragma unroll
        for (int k = 0; k < LB; k++) {
            const u64 j = base + (u64)k * 256 + threadIdx.x;
            m[k] = j < n ? LCP[j] : 0u;
        }
#pragma unroll
        for (int k = 0; k < LB; k++) {
            const u64 j = base + (u64)k * 256 + threadIdx.x;
            p[k] = (m[k] & LCP_UNKNOWN) ? SA[j] : 0u;
        }
#pragma unroll
        for (int k = 0; k < LB; k++) v[k] = ((m[k] & LCP_UNKNOWN) && p[k] < n) ? PLCP[p[k]] : 0u;
#pragma unroll
        for (int k = 0; k < LB; k++) {
            const u64 j = base + (u64)k * 256 + threadIdx.x;
            if (m[k] & LCP_UNKNOWN) LCP[j] = v[k];
        }
    }
}

} // namespace

int launch_lcp_fix(const u32 *PLCP, const u32 *SA, u64 n, u32 *LCP, hipStream_t s)
{
    hipLaunchKernelGGL(k_lcp_fix, dim3(grid_for(n, 256 * LB, 65535)), dim3(256), 0, s, PLCP, SA, n, LCP);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

template <typename VT, int W>
int launch_plcp_t(LcpScratch &ls, const uint8_t *T, u64 n, int gsa, const u32 *PHI, const u32 *ISA, const u32 *SA,
                  u32 *PLCP, VT *V, const u32 *marks, hipStream_t s, const u64 *pairs = nullptr,
                  const u32 *pw = nullptr, int bits = 0, int hint_off = 0)
{
    const u64 tiles = (n + MS_TILE - 1) / MS_TILE;
    if (tiles > ls.tiles_cap) {
        sagpu_set_error("plcp scratch too small", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }
    VT *tmax = (VT *)ls.tile_max;
    SAGPU_TRY(launch_fill_u32(ls.queue_n, 2, 0, s));
    const bool fused_max = pairs && W == 1;   // the pair-word pass writes the tile maxima itself
    if (fused_max) {
        if constexpr (W == 1) {
            // short texts: one 1024-position step per workgroup
            const int split = n < (1ull << 20) ? (int)(MS_TILE / (256 * PI)) : 1;
            if (split > 1) SAGPU_TRY(launch_fill_u32((u32 *)tmax, tiles * sizeof(VT) / 4, 0, s));
            if (ls.hook) ls.hook(ls.hook_user, 0);
            const u64 wg = tiles * (u64)split;
            hipLaunchKernelGGL((k_plcp_irr_pairs<VT>), dim3((unsigned)(wg < 65535 ? wg : 65535)), dim3(256), 0, s, T, n,
                               gsa, ISA, pairs, V, ls.queue, ls.queue_l, ls.queue_n, ls.queue_cap,
                               gsa ? nullptr : pw, bits, tmax, hint_off, split);
            if (ls.hook) ls.hook(ls.hook_user, 1);
        }
    } else {
        hipLaunchKernelGGL((k_plcp_irr<VT, W>), dim3(grid_for(n, 256 * 8, 65535)), dim3(256), 0, s, T, n, gsa, PHI, ISA,
                           SA, V, ls.queue, ls.queue_l, ls.queue_n, ls.queue_cap, marks);
    }
    const VLongQueue vq{ls.vq_pos ? ls.queue_n + 1 : nullptr, ls.vq_pos, ls.vq_off, ls.vq_res};
    hipLaunchKernelGGL((k_plcp_long<VT, W>), dim3(4096), dim3(256), 0, s, T, n, gsa, PHI, ISA, SA, V, ls.queue, ls.queue_l,
                       ls.queue_n, ls.queue_cap, fused_max ? tmax : (VT *)nullptr, vq);
    if (vq.n) {
        hipLaunchKernelGGL((k_plcp_vlong<W>), dim3(2048), dim3(256), 0, s, T, n, gsa, PHI, ISA, SA, vq);
        hipLaunchKernelGGL((k_plcp_vlong_fin<VT, W>), dim3(VQ_CAP / 256), dim3(256), 0, s, V,
                           fused_max ? tmax : (VT *)nullptr, vq);
    }
    if (!fused_max) hipLaunchKernelGGL(k_max_reduce<VT>, dim3((unsigned)tiles), dim3(MS_THREADS), 0, s, V, n, tmax);
    if (tiles <= (u64)MX_STEP) {
        hipLaunchKernelGGL(k_max_scan_tiles<VT>, dim3(1), dim3(1024), 0, s, tmax, tiles);
    } else {
        const u64 nch = (tiles + MX_STEP - 1) / MX_STEP;
        VT *cmax = (VT *)((u64 *)ls.tile_max + ls.tiles_cap);   // chunk maxima after the tile maxima
        hipLaunchKernelGGL(k_max_chunk_reduce<VT>, dim3((unsigned)nch), dim3(1024), 0, s, (const VT *)tmax, tiles, cmax);
        hipLaunchKernelGGL(k_max_scan_tiles<VT>, dim3(1), dim3(1024), 0, s, cmax, nch);
        hipLaunchKernelGGL(k_max_chunk_apply<VT>, dim3((unsigned)nch), dim3(1024), 0, s, tmax, tiles, (const VT *)cmax);
    }
    hipLaunchKernelGGL(k_max_apply<VT>, dim3((unsigned)tiles), dim3(MS_THREADS), 0, s, V, n, tmax, PLCP);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_plcp(LcpScratch &ls, const uint8_t *T, u64 n, int gsa, const u32 *PHI, const u32 *ISA, const u32 *SA,
                u32 *PLCP, u64 *V64, const u32 *marks, hipStream_t s, int sym_bytes, u64 *pairs, const uint8_t *U,
                const u32 *pw, int bits, int hint_off)
{
    if (sym_bytes == 2) {
        if (marks || !PHI) {
            sagpu_set_error("16-bit PLCP takes the Phi form only", __FILE__, __LINE__);
            return SAGPU_EINVAL;
        }
        if (V64) return launch_plcp_t<u64, 2>(ls, T, n, gsa, PHI, ISA, SA, PLCP, V64, nullptr, s);
        return launch_plcp_t<u32, 2>(ls, T, n, gsa, PHI, ISA, SA, PLCP, PLCP, nullptr, s);
    }
    if (sym_bytes == 4) {
        if (gsa || marks || !PHI) {
            sagpu_set_error("integer PLCP takes the Phi form only", __FILE__, __LINE__);
            return SAGPU_EINVAL;
        }
        if (V64) return launch_plcp_t<u64, 4>(ls, T, n, 0, PHI, ISA, SA, PLCP, V64, nullptr, s);
        return launch_plcp_t<u32, 4>(ls, T, n, 0, PHI, ISA, SA, PLCP, PLCP, nullptr, s);
    }
    if (marks && pairs) {
        ktime_begin(KI_PLCP_WORDS);
        hipLaunchKernelGGL(k_plcp_pairs, dim3(grid_for(n, 256 * PP, 65535)), dim3(256), 0, s, SA, marks, U, T, ISA, n,
                           pairs);
        ktime_end(KI_PLCP_WORDS, 16 * n);   // mark and SA read, one pair word written
    } else {
        pairs = nullptr;
    }
    if (V64) return launch_plcp_t<u64, 1>(ls, T, n, gsa, PHI, ISA, SA, PLCP, V64, marks, s, pairs, pw, bits, hint_off);
    return launch_plcp_t<u32, 1>(ls, T, n, gsa, PHI, ISA, SA, PLCP, PLCP, marks, s, pairs, pw, bits, hint_off);
}

int launch_lcp(const u32 *PLCP, const u32 *SA, u64 n, u32 *LCP, hipStream_t s)
{
    hipLaunchKernelGGL(k_lcp, dim3(grid_for(n, 256 * 8, 65535)), dim3(256), 0, s, PLCP, SA, n, LCP);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}
