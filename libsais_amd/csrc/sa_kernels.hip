// sa_kernels.hip -- symbol histogram, packed initial keys, doubling-round keys.
#include "engine.h"

namespace {

// 256-bin byte histogram (the libsais `freq` table) with LDS privatisation: one
// sub-histogram per wave so that the few hot bins of DNA text do not serialise
// the whole workgroup.  In GSA mode it also counts the positions that make the
// input invalid (a 0 at position 0 or a 0 directly after another 0).
__global__ __launch_bounds__(256) void k_histogram(const uint8_t *__restrict__ T, u64 n, int gsa,
                                                   u64 *__restrict__ hist, u32 *__restrict__ bad)
{
    // 16 sub-histograms: lanes of a wave spread over 4 copies (lane % 4), so a text of
    // few distinct bytes (DNA) has a quarter of the same-word atomics per step.  Rows
    // of 257 words put the copies of one byte value in different LDS banks.
    __shared__ u32 s_h[16][257];
    for (int i = threadIdx.x; i < 16 * 257; i += 256) (&s_h[0][0])[i] = 0;
    __syncthreads();
    const u32 w = (threadIdx.x / SAGPU_WAVE) * 4 + (threadIdx.x & 3);
    u32 nbad = 0;
    const u64 nvec = n / 16;
    const uint4 *T16 = reinterpret_cast<const uint4 *>(T);
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < nvec; v += stride) {
        uint4 x = T16[v];
        u32 words[4] = {x.x, x.y, x.z, x.w};
        u32 prev = (v == 0) ? 0u : T[16 * v - 1];
#pragma unroll
        for (int q = 0; q < 4; q++) {
#pragma unroll
            for (int b = 0; b < 4; b++) {
                u32 c = (words[q] >> (8 * b)) & 255u;
                atomicAdd(&s_h[w][c], 1u);
                if (gsa) nbad += (c == 0) && (prev == 0);
                prev = c;
            }
        }
    }
    for (u64 i = nvec * 16 + (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        u32 c = T[i];
        atomicAdd(&s_h[w][c], 1u);
        if (gsa) nbad += (c == 0) && (i == 0 || T[i - 1] == 0);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < 256; c += 256) {
        u32 t = 0;
        for (int q = 0; q < 16; q++) t += s_h[q][c];
        if (t) atomicAdd(&hist[c], (u64)t);
    }
    if (gsa && nbad) atomicAdd(bad, nbad);
}

constexpr int PK_THREADS = 256, PK_ITEMS = 16, PK_TILE = PK_THREADS * PK_ITEMS, PK_HALO = 64;

// Packed keys of the first k symbols of every suffix (k*bits <= 32): symbol t of
// suffix i sits in bits [bits*(k-1-t), bits*(k-t)).  Positions past the end read
// code 0, the smallest code; ties that this creates are broken by the doubling
// rounds, where a suffix shorter than the offset ranks by its length.  In GSA mode
// (separator = code 0, letters >= 1) every symbol after the first separator is
// forced to 0, so keys that contain a separator end in a zero slot.
// keys[p] = codes of T[p .. p+k) (first symbol high), vals[p] = p, or with cb > 0
// (p << cb) | code of T[p-1]: the BWT symbol of suffix p travels with it through
// the initial sort.
__global__ __launch_bounds__(PK_THREADS) void k_pack_keys32(const uint8_t *__restrict__ T, u64 n,
                                                            const uint8_t *__restrict__ lut, int bits, int k,
                                                            int gsa, int cb, u32 *__restrict__ keys,
                                                            u32 *__restrict__ vals, int items)
{
    // items <= PK_ITEMS keys per thread (short texts take fewer, so more workgroups
    // share the k-symbol loops)
    __shared__ uint8_t s_lut[256];
    __shared__ uint8_t s_c[PK_TILE + PK_HALO + 1];   // s_c[i] = code of T[base - 1 + i]
    s_lut[threadIdx.x] = lut[threadIdx.x];
    __syncthreads();
    const int tile = PK_THREADS * items;
    const u64 base = (u64)blockIdx.x * tile;
    for (int i = threadIdx.x; i < tile + PK_HALO + 1; i += PK_THREADS) {
        u64 p = base + i;   // position base - 1 + i, shifted by one
        s_c[i] = (p >= 1 && p - 1 < n) ? s_lut[T[p - 1]] : 0;
    }
    __syncthreads();
#pragma unroll 4
    for (int j = 0; j < items; j++) {
        int li = j * PK_THREADS + threadIdx.x;
        u64 p = base + li;
        if (p >= n) break;
        u32 key = 0;
        bool sep = false;
        for (int t = 0; t < k; t++) {
            u32 c = s_c[li + 1 + t];
            if (gsa) {
                if (sep) c = 0;
                sep = sep || c == 0;
            }
            key = (key << bits) | c;
        }
        keys[p] = key;
        vals[p] = cb ? ((u32)p << cb) | (u32)s_c[li] : (u32)p;
    }
}

__global__ __launch_bounds__(256) void k_zero_flags(const uint8_t *__restrict__ T, u64 n, u32 *__restrict__ f)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 p = (u64)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) f[p] = T[p] == 0;
}

__global__ __launch_bounds__(PK_THREADS) void k_pack_keys64_sid(const uint8_t *__restrict__ T, u64 n,
                                                                const uint8_t *__restrict__ lut, int bits, int k,
                                                                const u32 *__restrict__ sid, u64 *__restrict__ keys,
                                                                u32 *__restrict__ vals)
{
    __shared__ uint8_t s_lut[256];
    __shared__ uint8_t s_c[PK_TILE + PK_HALO];
    s_lut[threadIdx.x] = lut[threadIdx.x];
    __syncthreads();
    const u64 base = (u64)blockIdx.x * PK_TILE;
    for (int i = threadIdx.x; i < PK_TILE + PK_HALO; i += PK_THREADS) {
        u64 p = base + i;
        s_c[i] = p < n ? s_lut[T[p]] : 0;
    }
    __syncthreads();
#pragma unroll 4
    for (int j = 0; j < PK_ITEMS; j++) {
        int li = j * PK_THREADS + threadIdx.x;
        u64 p = base + li;
        if (p >= n) break;
        u32 key = 0;
        bool sep = false;
        for (int t = 0; t < k; t++) {
            u32 c = s_c[li + t];
            if (sep) c = 0;
            sep = sep || c == 0;
            key = (key << bits) | c;
        }
        keys[p] = ((u64)sid[p] << 32) | key;
        vals[p] = (u32)p;
    }
}

__global__ void k_check_offsets(const uint8_t *__restrict__ T, const u32 *__restrict__ sid,
                                const int64_t *__restrict__ off, u32 count, u32 *bad)
{
    for (u32 q = blockIdx.x * blockDim.x + threadIdx.x; q < count; q += gridDim.x * blockDim.x) {
        int64_t o = off[q];
        bool ok = sid[o] == q && off[q + 1] > o + 1 && T[off[q + 1] - 1] == 0 && (q == 0 || T[o - 1] == 0);
        if (!ok) atomicOr(bad, 1u);
    }
}

__global__ __launch_bounds__(256) void k_localize(const u32 *__restrict__ SA, const u32 *__restrict__ sid,
                                                  const int64_t *__restrict__ off, u64 n, u32 *__restrict__ out)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride)
        out[j] = SA[j] - (u32)off[sid[j]];
}

__global__ __launch_bounds__(256) void k_batch_extract(const u32 *__restrict__ gSA, const u32 *__restrict__ gLCP,
                                                       const u32 *__restrict__ sid, const int64_t *__restrict__ off,
                                                       const u32 *__restrict__ ISA, const uint8_t *__restrict__ T,
                                                       u64 N, u32 *__restrict__ outSA, u32 *__restrict__ outLCP,
                                                       uint8_t *__restrict__ outU)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x; j < N; j += stride) {
        const u32 k = sid[j];
        const u64 o = (u64)off[k];
        if (j == o) continue;                 // the separator's own suffix
        const u64 jj = j - o - 1;             // slot inside job k's suffix array
        const u64 po = o - k;                 // packed output offset of job k
        const u32 v = gSA[j] - (u32)o;        // suffix index inside job k
        outSA[po + jj] = v;
        if (outLCP) outLCP[po + jj] = gLCP[j];
        if (outU) {
            const u64 nk = (u64)off[k + 1] - o - 1;
            const u64 j0 = (u64)isa_rank(ISA[o], N) - o - 1;
            if (v == 0) outU[po] = T[o + nk - 1];
            else outU[po + jj + (jj < j0)] = T[o + v - 1];
        }
    }
}

__global__ void k_batch_primary(const u32 *__restrict__ ISA, const int64_t *__restrict__ off, u32 count, u64 N,
                                u32 *__restrict__ prim)
{
    for (u32 q = blockIdx.x * blockDim.x + threadIdx.x; q < count; q += gridDim.x * blockDim.x)
        prim[q] = isa_rank(ISA[off[q]], N) - (u32)off[q];
}

// Plain device copy, 16 bytes per lane per step (the copy-bandwidth yardstick).
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, u64 count)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) dst[i] = src[i];
}

__global__ void k_copy_tail(const uint8_t *__restrict__ src, uint8_t *__restrict__ dst, u64 count)
{
    for (u64 i = threadIdx.x; i < count; i += blockDim.x) dst[i] = src[i];
}

__global__ void k_fill_u32(u32 *p, u64 count, u32 v)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = v;
}

__global__ void k_fill_u8(uint8_t *p, u64 count, uint8_t v)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = v;
}

__global__ void k_fill_u64(u64 *p, u64 count, u64 v)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = v;
}

// Integer alphabet: packed keys of `k` symbols of `bits` bits each read from a u32
// symbol array (values < 2^bits).  Past the end reads 0, as for bytes.
__global__ __launch_bounds__(PK_THREADS) void k_pack_int(const u32 *__restrict__ S, u64 n, int bits, int k, int gsa,
                                                         u32 *__restrict__ keys, u32 *__restrict__ vals)
{
    __shared__ u32 s_c[PK_TILE + PK_HALO];
    const u64 base = (u64)blockIdx.x * PK_TILE;
    for (int i = threadIdx.x; i < PK_TILE + PK_HALO; i += PK_THREADS) {
        u64 p = base + i;
        s_c[i] = p < n ? S[p] : 0u;
    }
    __syncthreads();
    for (int j = 0; j < PK_ITEMS; j++) {
        int li = j * PK_THREADS + threadIdx.x;
        u64 p = base + li;
        if (p >= n) break;
        u32 key = 0;
        bool sep = false;
        for (int t = 0; t < k; t++) {
            u32 c = s_c[li + t];
            if (gsa) {  // code 0 is the separator: everything after the first one reads 0
                if (sep) c = 0;
                sep = sep || c == 0;
            }
            key = (bits >= 32 ? 0u : key << bits) | c;
        }
        keys[p] = key;
        vals[p] = (u32)p;
    }
}

// Largest symbol value (as unsigned), for the symbol width.
template <typename ST>
__global__ __launch_bounds__(256) void k_max_sym(const ST *__restrict__ T, u64 n, u64 *__restrict__ out)
{
    u64 m = 0;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        u64 v = sizeof(ST) == 4 ? (u64)(u32)T[i] : (u64)T[i];
        m = v > m ? v : m;
    }
    m = wave_incl_max64(m);
    if (lane_id() == SAGPU_WAVE - 1 && m) atomicMax((unsigned long long *)out, (unsigned long long)m);
}

// Symbols as sort keys (u32 narrowing when the values are known to fit, or u64)
// with their positions for dense ranking (vals may be null).
template <typename ST, typename K>
__global__ __launch_bounds__(256) void k_sym_keys(const ST *__restrict__ T, u64 n, K *__restrict__ keys,
                                                  u32 *__restrict__ vals)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        keys[i] = sizeof(ST) == 4 ? (K)(u32)T[i] : (K)(u64)T[i];
        if (vals) vals[i] = (u32)i;
    }
}

// After sorting (value, position): flag[e] = 1 where a new value starts.
template <typename K>
__global__ __launch_bounds__(256) void k_run_flags(const K *__restrict__ keys, u64 n, u32 *__restrict__ flags)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride)
        flags[e] = (e == 0 || keys[e] != keys[e - 1]) ? 1u : 0u;
}

// S[position] = dense rank (number of distinct smaller values), from the exclusive
// scan of the run flags.
template <typename K>
__global__ __launch_bounds__(256) void k_dense_scatter(const K *__restrict__ keys, const u32 *__restrict__ vals,
                                                       const u32 *__restrict__ excl, u64 n, u32 *__restrict__ S)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        const u32 f = (e == 0 || keys[e] != keys[e - 1]) ? 1u : 0u;
        S[vals[e]] = excl[e] + f - 1;
    }
}

// 65536-bin histogram of 16-bit symbols.  Lanes holding the same symbol are
// grouped with a ballot peer mask, so each distinct symbol of a wave costs one global
// atomic.  In GSA mode it also counts the invalid positions (a 0 first, or two
// adjacent 0s).
__global__ __launch_bounds__(256) void k_hist16(const uint16_t *__restrict__ T, u64 n, int gsa, u64 *__restrict__ hist,
                                                u32 *__restrict__ bad)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    const u64 lane = lane_id();
    u32 nbad = 0;
    for (u64 base = (u64)blockIdx.x * blockDim.x + threadIdx.x - lane; base < n; base += stride) {
        const u64 i = base + lane;
        const bool ok = i < n;
        const u32 v = ok ? T[i] : 0u;
        if (gsa && ok && v == 0 && (i == 0 || T[i - 1] == 0)) nbad++;
        u64 m = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 16; b++) {
            const bool bit = (v >> b) & 1u;
            const u64 bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        if (ok && __popcll(m & lanemask_lt()) == 0) atomicAdd((unsigned long long *)&hist[v], (unsigned long long)__popcll(m));
    }
    if (gsa && nbad) atomicAdd(bad, nbad);
}

// Dense codes of 16-bit symbols through a 65536-entry table.
__global__ __launch_bounds__(256) void k_map16(const uint16_t *__restrict__ T, u64 n, const u32 *__restrict__ lut,
                                               u32 *__restrict__ S)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) S[i] = lut[T[i]];
}

unsigned grid_for(u64 work, unsigned per_block, unsigned cap)
{
    u64 b = (work + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

} // namespace

int launch_histogram(const uint8_t *T, u64 n, int gsa, u64 *hist256, u32 *gsa_bad, hipStream_t s)
{
    hipLaunchKernelGGL(k_histogram, dim3(grid_for(n, 256 * 16 * 8, 2048)), dim3(256), 0, s, T, n, gsa, hist256,
                       gsa_bad);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_pack_keys32(const uint8_t *T, u64 n, const uint8_t *lut, int bits, int k, int gsa, int cb, u32 *keys,
                       u32 *vals, hipStream_t s)
{
    if (k > PK_HALO || k * bits > 32 || (cb && (u64)(n - 1) >> (32 - cb))) {
        sagpu_set_error("pack32: bad symbol count", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }
    const int items = n < (1u << 20) ? 1 : PK_ITEMS;
    hipLaunchKernelGGL(k_pack_keys32, dim3(ceil_div_u(n, (u64)PK_THREADS * items)), dim3(PK_THREADS), 0, s, T, n, lut,
                       bits, k, gsa, cb, keys, vals, items);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_hist16(const uint16_t *T, u64 n, int gsa, u64 *hist65536, u32 *bad, hipStream_t s)
{
    // the loop steps a whole wave at a time, so the grid is a multiple of 64 threads
    hipLaunchKernelGGL(k_hist16, dim3(grid_for(n, 256 * 16, 4096)), dim3(256), 0, s, T, n, gsa, hist65536, bad);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_map16(const uint16_t *T, u64 n, const u32 *lut, u32 *S, hipStream_t s)
{
    hipLaunchKernelGGL(k_map16, dim3(grid_for(n, 256 * 8, 65535)), dim3(256), 0, s, T, n, lut, S);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_pack_int(const u32 *S, u64 n, int bits, int k, int gsa, u32 *keys, u32 *vals, hipStream_t s)
{
    if (k > PK_HALO || k * bits > 32 || k < 1) {
        sagpu_set_error("pack_int: bad symbol count", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }
    hipLaunchKernelGGL(k_pack_int, dim3(ceil_div_u(n, PK_TILE)), dim3(PK_THREADS), 0, s, S, n, bits, k, gsa, keys,
                       vals);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_max_sym(const void *T, int width, u64 n, u64 *out, hipStream_t s)
{
    if (width == 4)
        hipLaunchKernelGGL(k_max_sym<int32_t>, dim3(grid_for(n, 256 * 8, 4096)), dim3(256), 0, s, (const int32_t *)T,
                           n, out);
    else
        hipLaunchKernelGGL(k_max_sym<int64_t>, dim3(grid_for(n, 256 * 8, 4096)), dim3(256), 0, s, (const int64_t *)T,
                           n, out);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_sym_narrow(const void *T, int width, u64 n, u32 *S, hipStream_t s)
{
    if (width == 4)
        hipLaunchKernelGGL((k_sym_keys<int32_t, u32>), dim3(grid_for(n, 256 * 8, 65535)), dim3(256), 0, s,
                           (const int32_t *)T, n, S, (u32 *)nullptr);
    else
        hipLaunchKernelGGL((k_sym_keys<int64_t, u32>), dim3(grid_for(n, 256 * 8, 65535)), dim3(256), 0, s,
                           (const int64_t *)T, n, S, (u32 *)nullptr);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_sym_keys64(const void *T, int width, u64 n, u64 *keys, u32 *vals, hipStream_t s)
{
    if (width == 4)
        hipLaunchKernelGGL((k_sym_keys<int32_t, u64>), dim3(grid_for(n, 256 * 8, 65535)), dim3(256), 0, s,
                           (const int32_t *)T, n, keys, vals);
    else
        hipLaunchKernelGGL((k_sym_keys<int64_t, u64>), dim3(grid_for(n, 256 * 8, 65535)), dim3(256), 0, s,
                           (const int64_t *)T, n, keys, vals);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_run_flags64(const u64 *keys, u64 n, u32 *flags, hipStream_t s)
{
    hipLaunchKernelGGL(k_run_flags<u64>, dim3(grid_for(n, 256 * 8, 65535)), dim3(256), 0, s, keys, n, flags);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_dense_scatter64(const u64 *keys, const u32 *vals, const u32 *excl, u64 n, u32 *S, hipStream_t s)
{
    hipLaunchKernelGGL(k_dense_scatter<u64>, dim3(grid_for(n, 256 * 8, 65535)), dim3(256), 0, s, keys, vals, excl, n,
                       S);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_string_ids(const uint8_t *T, u64 n, u32 *sid, ScanScratch &sc, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_flags, dim3(grid_for(n, 256 * 8, 65535)), dim3(256), 0, s, T, n, sid);
    SAGPU_CHECK(hipGetLastError());
    return scan_exclusive_u32(sc, sid, n, s);
}

int launch_pack_keys64_sid(const uint8_t *T, u64 n, const uint8_t *lut, int bits, int k, const u32 *sid, u64 *keys,
                           u32 *vals, hipStream_t s)
{
    if (k > PK_HALO || k * bits > 32) {
        sagpu_set_error("pack64: bad symbol count", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }
    hipLaunchKernelGGL(k_pack_keys64_sid, dim3(ceil_div_u(n, PK_TILE)), dim3(PK_THREADS), 0, s, T, n, lut, bits, k,
                       sid, keys, vals);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_check_offsets(const uint8_t *T, const u32 *sid, const int64_t *offsets, u32 count, u32 *bad, hipStream_t s)
{
    hipLaunchKernelGGL(k_check_offsets, dim3(grid_for(count, 256, 4096)), dim3(256), 0, s, T, sid, offsets, count,
                       bad);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_localize(const u32 *SA, const u32 *sid, const int64_t *offsets, u64 n, u32 *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_localize, dim3(grid_for(n, 256 * 8, 65535)), dim3(256), 0, s, SA, sid, offsets, n, out);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_batch_extract(const u32 *gSA, const u32 *gLCP, const u32 *sid, const int64_t *off, const u32 *ISA,
                         const uint8_t *T, u64 N, u32 *outSA, u32 *outLCP, uint8_t *outU, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_extract, dim3(grid_for(N, 256 * 8, 65535)), dim3(256), 0, s, gSA, gLCP, sid, off, ISA,
                       T, N, outSA, outLCP, outU);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_batch_primary(const u32 *ISA, const int64_t *off, u32 count, u64 N, u32 *prim, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_primary, dim3(grid_for(count, 256, 4096)), dim3(256), 0, s, ISA, off, count, N, prim);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_fill_u32(u32 *p, u64 count, u32 value, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(count, 256 * 4, 16384)), dim3(256), 0, s, p, count, value);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_copy(void *dst, const void *src, u64 bytes, hipStream_t s)
{
    const u64 vec = ((uintptr_t)dst % 16 == 0 && (uintptr_t)src % 16 == 0) ? bytes / 16 : 0;
    if (vec) {
        u64 blocks = (vec + 256 * 4 - 1) / (256 * 4);
        if (blocks > 8192) blocks = 8192;
        hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, s, (const uint4 *)src, (uint4 *)dst, vec);
    }
    if (bytes > vec * 16)
        hipLaunchKernelGGL(k_copy_tail, dim3(1), dim3(256), 0, s, (const uint8_t *)src + vec * 16,
                           (uint8_t *)dst + vec * 16, bytes - vec * 16);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_fill_u8(uint8_t *p, u64 count, uint8_t value, hipStream_t s)
{
    if (!count) return SAGPU_OK;
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(count, 256 * 16, 16384)), dim3(256), 0, s, p, count, value);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int launch_fill_u64(u64 *p, u64 count, u64 value, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(count, 256 * 4, 16384)), dim3(256), 0, s, p, count, value);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}
