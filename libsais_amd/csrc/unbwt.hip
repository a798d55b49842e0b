// unbwt.hip -- BWT auxiliary samples and the inverse BWT.
//
// Inverse BWT.  Rows 0..n of the sorted rotations of T$: row 0 starts with $, row
// x >= 1 is SA slot x-1.  The libsais BWT layout stores the last column without the
// $ of row p (the 1-based primary index): U[q] is the last symbol of row
// x(q) = q + (q >= p).  Walking LF from row 0 reads T backwards, and
//   LF(x(q)) = 1 + #{bytes < U[q]} + #{q' < q : U[q'] == U[q]},
// a stable counting rank.  k_ub_count / scan / k_ub_succ turn that into a
// successor array S over BWT positions (S[q] = position of the next row, or
// UB_END when the next row is the $ row).
//
// Decoding, two ways:
//   - aux samples (libsais_unbwt_aux with many samples): sample k+1 gives the row
//     of suffix (k+1)*r, so chunk [k*r, (k+1)*r) is an independent walk of r
//     steps; one thread per chunk;
//   - list ranking (plain libsais_unbwt, or few samples): a hashed ~1/256 subset of
//     positions (always including position 0, the start) are samples; every sample
//     walks to the next sample (k_ub_lr_walk), pointer jumping over the samples
//     gives each one its distance to the end, and a second walk writes the text
//     (k_ub_lr_emit).  The hash keeps the gaps geometric (mean 256) whatever the
//     text; a gap longer than UB_CAP steps is reported as an error instead of
//     running on.
#include "engine.h"

namespace {

constexpr int UB_THREADS = 256, UB_ITEMS = 16, UB_TILE = UB_THREADS * UB_ITEMS;
constexpr int UB_WAVES = UB_THREADS / SAGPU_WAVE;
constexpr u32 UB_END = 0xFFFFFFFFu;
constexpr u32 UB_CAP = 1u << 16;
constexpr u32 UB_SAMPLE_MASK = 255u;

__device__ __forceinline__ bool ub_sample(u64 q)
{
    u32 h = (u32)q * 0x9E3779B1u;
    h ^= h >> 15;
    h *= 0x85EBCA77u;
    h ^= h >> 13;
    h ^= (u32)(q >> 32) * 0xC2B2AE3Du;
    return q == 0 || (h & UB_SAMPLE_MASK) == 0;
}

// Byte counts per tile, digit-major (counts[d * tiles + t]), for the stable rank.
__global__ __launch_bounds__(UB_THREADS) void k_ub_count(const uint8_t *__restrict__ U, u64 n,
                                                         u32 *__restrict__ counts, u64 tiles)
{
    __shared__ u32 s_h[UB_WAVES][256];
    const u32 tid = threadIdx.x, w = tid / SAGPU_WAVE;
#pragma unroll
    for (int q = 0; q < UB_WAVES; q++) s_h[q][tid] = 0;
    __syncthreads();
    const u64 base = (u64)blockIdx.x * UB_TILE;
#pragma unroll
    for (int j = 0; j < UB_ITEMS; j++) {
        u64 pos = base + (u64)j * UB_THREADS + tid;
        if (pos < n) atomicAdd(&s_h[w][U[pos]], 1u);
    }
    __syncthreads();
    u32 c = 0;
#pragma unroll
    for (int q = 0; q < UB_WAVES; q++) c += s_h[q][tid];
    counts[(u64)tid * tiles + blockIdx.x] = c;
}

// S[q] from the stable rank of U[q]: scanned[d * tiles + t] = bytes < d overall plus
// bytes == d in earlier tiles; inside the tile the rank is position order (wave,
// item, lane), found with a ballot peer mask as in the radix pass.
__global__ __launch_bounds__(UB_THREADS) void k_ub_succ(const uint8_t *__restrict__ U, u64 n, u64 p,
                                                        const u32 *__restrict__ scanned, u64 tiles,
                                                        u32 *__restrict__ S)
{
    __shared__ u32 s_wcnt[UB_WAVES][256];
    const u32 tid = threadIdx.x, w = tid / SAGPU_WAVE, lane = lane_id();
#pragma unroll
    for (int q = 0; q < UB_WAVES; q++) s_wcnt[q][tid] = 0;
    __syncthreads();
    const u64 tile = blockIdx.x;
    const u64 wbase = tile * UB_TILE + (u64)w * (SAGPU_WAVE * UB_ITEMS);
    u32 dig[UB_ITEMS], rank[UB_ITEMS];
#pragma unroll
    for (int j = 0; j < UB_ITEMS; j++) {
        u64 pos = wbase + (u64)j * SAGPU_WAVE + lane;
        bool ok = pos < n;
        u32 d = ok ? (u32)U[pos] : 0u;
        u64 m = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            u64 bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        u32 below = __popcll(m & lanemask_lt());
        u32 old = 0;
        if (ok) old = s_wcnt[w][d];
        if (ok && below == 0) s_wcnt[w][d] = old + (u32)__popcll(m);
        dig[j] = d;
        rank[j] = old + below;
    }
    __syncthreads();
    {
        const u32 d = tid;
        u32 run = 0;
#pragma unroll
        for (int q = 0; q < UB_WAVES; q++) {
            u32 c = s_wcnt[q][d];
            s_wcnt[q][d] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < UB_ITEMS; j++) {
        u64 pos = wbase + (u64)j * SAGPU_WAVE + lane;
        if (pos >= n) continue;
        const u32 d = dig[j];
        const u64 x = (u64)scanned[(u64)d * tiles + tile] + s_wcnt[w][d] + rank[j] + 1;  // next row
        S[pos] = x == p ? UB_END : (u32)(x - (x > p));
    }
}

// Aux decoding: one thread per chunk [c*r, min((c+1)*r, n)), walking back from the
// row of suffix (c+1)*r (row 0 for the last chunk).
template <typename ST>
__global__ __launch_bounds__(256) void k_ub_walk_aux(const ST *__restrict__ U, const u32 *__restrict__ S, u64 n,
                                                     u64 p, u64 r, const u64 *__restrict__ rows, u64 chunks,
                                                     ST *__restrict__ out, u32 *__restrict__ err)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 c = (u64)blockIdx.x * blockDim.x + threadIdx.x; c < chunks; c += stride) {
        const u64 lo = c * r, hi = lo + r < n ? lo + r : n;
        const u64 x = c + 1 < chunks ? rows[c + 1] : 0;
        if (x == p || x > n) {  // a sample that claims the $ row: inconsistent input
            atomicOr(err, 1u);
            continue;
        }
        u64 q = x - (x > p);
        for (u64 j = hi; j-- > lo;) {
            out[j] = U[q];
            if (j == lo) break;
            u32 nq = S[q];
            if (nq == UB_END) {
                atomicOr(err, 1u);
                break;
            }
            q = nq;
        }
    }
}

// flags[q] = 1 for samples (scanned afterwards into sample ids).
__global__ __launch_bounds__(256) void k_ub_flags(u64 n, u32 *__restrict__ flags)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 q = (u64)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) flags[q] = ub_sample(q) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_ub_collect(u64 n, const u32 *__restrict__ sid, u32 *__restrict__ samp_q)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 q = (u64)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
        if (ub_sample(q)) samp_q[sid[q]] = (u32)q;
}

// Each sample walks to the next sample: gap = symbols emitted from this sample up to
// the next one (or the end), nxt = the next sample's id (UB_END at the end).
__global__ __launch_bounds__(256) void k_ub_lr_walk(const u32 *__restrict__ S, const u32 *__restrict__ sid,
                                                    const u32 *__restrict__ samp_q, u32 m, u32 *__restrict__ nxt,
                                                    u64 *__restrict__ gap, u32 *__restrict__ err)
{
    for (u32 id = blockIdx.x * blockDim.x + threadIdx.x; id < m; id += gridDim.x * blockDim.x) {
        u32 c = S[samp_q[id]];
        u32 len = 1;
        while (c != UB_END && !ub_sample(c) && len < UB_CAP) {
            c = S[c];
            len++;
        }
        const bool capped = c != UB_END && !ub_sample(c);
        if (capped) atomicOr(err, 2u);  // gap over UB_CAP: the host reports an error
        nxt[id] = (c == UB_END || capped) ? UB_END : sid[c];
        gap[id] = len;
    }
}

// One pointer-jumping round: dist/nxt (in) -> dist2/nxt2 (out).
__global__ __launch_bounds__(256) void k_ub_jump(const u32 *__restrict__ nxt, const u64 *__restrict__ dist, u32 m,
                                                 u32 *__restrict__ nxt2, u64 *__restrict__ dist2)
{
    for (u32 id = blockIdx.x * blockDim.x + threadIdx.x; id < m; id += gridDim.x * blockDim.x) {
        u32 a = nxt[id];
        if (a == UB_END) {
            nxt2[id] = UB_END;
            dist2[id] = dist[id];
        } else {
            nxt2[id] = nxt[a];
            dist2[id] = dist[id] + dist[a];
        }
    }
}

// Samples on the path from position 0 (their pointer reached the end) write their
// gap: step t of the path emits out[n-1-t], and sample id starts at step
// dist[0] - dist[id].
template <typename ST>
__global__ __launch_bounds__(256) void k_ub_lr_emit(const ST *__restrict__ U, const u32 *__restrict__ S,
                                                    const u32 *__restrict__ samp_q, u32 m, const u32 *__restrict__ nxt,
                                                    const u64 *__restrict__ dist, const u64 *__restrict__ gap, u64 n,
                                                    ST *__restrict__ out)
{
    const u64 total = dist[0];
    for (u32 id = blockIdx.x * blockDim.x + threadIdx.x; id < m; id += gridDim.x * blockDim.x) {
        if (nxt[id] != UB_END || nxt[0] != UB_END) continue;  // not on the path from position 0
        const u64 t0 = total - dist[id];
        u32 q = samp_q[id];
        const u64 len = gap[id];
        for (u64 j = 0; j < len; j++) {
            const u64 t = t0 + j;
            if (t >= n) break;
            out[n - 1 - t] = U[q];
            if (j + 1 < len) q = S[q];
        }
    }
}

// Successor array from a stable sort of (symbol, position): the position sorted to
// slot k has LF row k + 1 (16-bit symbols, where per-tile counting tables do not fit).
__global__ __launch_bounds__(256) void k_ub_succ_sorted(const u32 *__restrict__ V, u64 n, u64 p, u32 *__restrict__ S)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const u64 x = k + 1;
        S[V[k]] = x == p ? UB_END : (u32)(x - (x > p));
    }
}

__global__ __launch_bounds__(256) void k_ub_keys16(const uint16_t *__restrict__ U, u64 n, u32 *__restrict__ keys,
                                                   u32 *__restrict__ vals)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 q = (u64)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        keys[q] = U[q];
        vals[q] = (u32)q;
    }
}

__global__ void k_ub_total(const u32 *__restrict__ sid, u64 n, u32 *__restrict__ out)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *out = sid[n - 1] + (ub_sample(n - 1) ? 1u : 0u);
}

__global__ __launch_bounds__(256) void k_aux_samples(const u32 *__restrict__ ISA, u64 n, u64 r, u64 count,
                                                     u32 *__restrict__ out)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) out[k] = isa_rank(ISA[k * r], n);
}

unsigned grid_of(u64 work, unsigned per_block, unsigned cap)
{
    u64 b = (work + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

} // namespace

int launch_aux_samples(const u32 *ISA, u64 n, u64 r, u64 count, u32 *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_aux_samples, dim3(grid_of(count, 256 * 4, 65535)), dim3(256), 0, s, ISA, n, r, count, out);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

u64 unbwt_count_slots(u64 n) { return 256 * ((n + UB_TILE - 1) / UB_TILE); }

int launch_unbwt_succ(const uint8_t *U, u64 n, u64 p, u32 *counts, ScanScratch &sc, u32 *S, hipStream_t s)
{
    const u64 tiles = (n + UB_TILE - 1) / UB_TILE;
    hipLaunchKernelGGL(k_ub_count, dim3((unsigned)tiles), dim3(UB_THREADS), 0, s, U, n, counts, tiles);
    SAGPU_CHECK(hipGetLastError());
    SAGPU_TRY(scan_exclusive_u32(sc, counts, tiles * 256, s));
    hipLaunchKernelGGL(k_ub_succ, dim3((unsigned)tiles), dim3(UB_THREADS), 0, s, U, n, p, counts, tiles, S);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_unbwt_aux_walk(const void *U, int sym_bytes, const u32 *S, u64 n, u64 p, u64 r, const u64 *rows,
                          u64 chunks, void *out, u32 *err, hipStream_t s)
{
    if (sym_bytes == 2)
        hipLaunchKernelGGL(k_ub_walk_aux<uint16_t>, dim3(grid_of(chunks, 256, 65535)), dim3(256), 0, s,
                           (const uint16_t *)U, S, n, p, r, rows, chunks, (uint16_t *)out, err);
    else
        hipLaunchKernelGGL(k_ub_walk_aux<uint8_t>, dim3(grid_of(chunks, 256, 65535)), dim3(256), 0, s,
                           (const uint8_t *)U, S, n, p, r, rows, chunks, (uint8_t *)out, err);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_unbwt_flags(u64 n, u32 *flags, hipStream_t s)
{
    hipLaunchKernelGGL(k_ub_flags, dim3(grid_of(n, 256 * 8, 65535)), dim3(256), 0, s, n, flags);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_unbwt_sample_total(const u32 *sid, u64 n, u32 *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_ub_total, dim3(1), dim3(64), 0, s, sid, n, out);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_unbwt_succ16(const uint16_t *U, u64 n, u64 p, u32 *keys, u32 *vals, u32 *S, hipStream_t s)
{
    (void)p;
    (void)S;
    hipLaunchKernelGGL(k_ub_keys16, dim3(grid_of(n, 256 * 8, 65535)), dim3(256), 0, s, U, n, keys, vals);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_unbwt_succ_sorted(const u32 *V, u64 n, u64 p, u32 *S, hipStream_t s)
{
    hipLaunchKernelGGL(k_ub_succ_sorted, dim3(grid_of(n, 256 * 8, 65535)), dim3(256), 0, s, V, n, p, S);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_unbwt_rank(const void *U, int sym_bytes, const u32 *S, const u32 *sid, u64 n, u32 m, u32 *samp_q,
                      u32 *nxt[2], u64 *dist[2], u64 *gap, void *out, u32 *err, hipStream_t s)
{
    hipLaunchKernelGGL(k_ub_collect, dim3(grid_of(n, 256 * 8, 65535)), dim3(256), 0, s, n, sid, samp_q);
    const unsigned g = grid_of(m, 256, 65535);
    hipLaunchKernelGGL(k_ub_lr_walk, dim3(g), dim3(256), 0, s, S, sid, samp_q, m, nxt[0], gap, err);
    SAGPU_CHECK(hipMemcpyAsync(dist[0], gap, (size_t)m * sizeof(u64), hipMemcpyDeviceToDevice, s));
    int rounds = 1;
    while ((1ull << rounds) < (u64)m + 1) rounds++;
    int cur = 0;
    for (int k = 0; k <= rounds; k++) {
        hipLaunchKernelGGL(k_ub_jump, dim3(g), dim3(256), 0, s, nxt[cur], dist[cur], m, nxt[cur ^ 1], dist[cur ^ 1]);
        cur ^= 1;
    }
    if (sym_bytes == 2)
        hipLaunchKernelGGL(k_ub_lr_emit<uint16_t>, dim3(g), dim3(256), 0, s, (const uint16_t *)U, S, samp_q, m,
                           nxt[cur], dist[cur], gap, n, (uint16_t *)out);
    else
        hipLaunchKernelGGL(k_ub_lr_emit<uint8_t>, dim3(g), dim3(256), 0, s, (const uint8_t *)U, S, samp_q, m,
                           nxt[cur], dist[cur], gap, n, (uint8_t *)out);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}
