// scan.hip -- in-place exclusive prefix sum over u32 (reduce-then-scan).
//
// Tiles of 4096 elements.  Kernel 1 sums each tile (striped, coalesced loads);
// kernel 2 scans the tile sums in one workgroup; kernel 3 stages each tile through
// LDS so that global loads and stores stay coalesced while every thread scans 16
// consecutive elements.
#include "engine.h"

namespace {

constexpr int SC_THREADS = 256, SC_ITEMS = 16, SC_TILE = SC_THREADS * SC_ITEMS;

__global__ __launch_bounds__(SC_THREADS) void k_scan_reduce(const u32 *__restrict__ in, u64 n, u32 *__restrict__ part)
{
    __shared__ u32 s_w[SC_THREADS / SAGPU_WAVE];
    const u64 base = (u64)blockIdx.x * SC_TILE;
    u32 sum = 0;
#pragma unroll
    for (int j = 0; j < SC_ITEMS; j++) {
        u64 e = base + (u64)j * SC_THREADS + threadIdx.x;
        if (e < n) sum += in[e];
    }
    for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, SAGPU_WAVE);
    if (lane_id() == 0) s_w[threadIdx.x / SAGPU_WAVE] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        u32 t = 0;
        for (int w = 0; w < SC_THREADS / SAGPU_WAVE; w++) t += s_w[w];
        part[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(1024) void k_scan_partials(u32 *part, u64 tiles)
{
    __shared__ u32 s_w[16];
    __shared__ u32 s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (u64 base = 0; base < tiles; base += 1024) {
        u64 i = base + threadIdx.x;
        u32 v = i < tiles ? part[i] : 0u;
        u32 inc = wave_incl_sum(v);
        if (lane_id() == SAGPU_WAVE - 1) s_w[threadIdx.x / SAGPU_WAVE] = inc;
        __syncthreads();
        u32 c = s_carry;
        for (u32 q = 0; q < threadIdx.x / SAGPU_WAVE; q++) c += s_w[q];
        if (i < tiles) part[i] = c + inc - v;
        __syncthreads();
        if (threadIdx.x == 1023) {
            u32 t = s_carry;
            for (int q = 0; q < 16; q++) t += s_w[q];
            s_carry = t;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_apply(u32 *__restrict__ data, u64 n, const u32 *__restrict__ part)
{
    __shared__ u32 s_d[SC_TILE + SC_TILE / 32];
    __shared__ u32 s_w[SC_THREADS / SAGPU_WAVE];
    const u64 base = (u64)blockIdx.x * SC_TILE;
    // padded index avoids bank conflicts when a thread walks 16 consecutive words
    auto pad = [](u32 x) { return x + (x >> 5); };
#pragma unroll
    for (int j = 0; j < SC_ITEMS; j++) {
        u32 x = j * SC_THREADS + threadIdx.x;
        u64 e = base + x;
        s_d[pad(x)] = e < n ? data[e] : 0u;
    }
    __syncthreads();
    u32 v[SC_ITEMS];
    u32 sum = 0;
#pragma unroll
    for (int j = 0; j < SC_ITEMS; j++) {
        v[j] = s_d[pad(threadIdx.x * SC_ITEMS + j)];
        sum += v[j];
    }
    u32 inc = wave_incl_sum(sum);
    if (lane_id() == SAGPU_WAVE - 1) s_w[threadIdx.x / SAGPU_WAVE] = inc;
    __syncthreads();
    u32 c = part[blockIdx.x];
    for (u32 q = 0; q < threadIdx.x / SAGPU_WAVE; q++) c += s_w[q];
    c += inc - sum;
#pragma unroll
    for (int j = 0; j < SC_ITEMS; j++) {
        s_d[pad(threadIdx.x * SC_ITEMS + j)] = c;
        c += v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SC_ITEMS; j++) {
        u32 x = j * SC_THREADS + threadIdx.x;
        u64 e = base + x;
        if (e < n) data[e] = s_d[pad(x)];
    }
}

} // namespace

int scan_exclusive_u32(ScanScratch &sc, u32 *data, u64 count, hipStream_t s)
{
    if (count == 0) return SAGPU_OK;
    const u64 tiles = (count + SC_TILE - 1) / SC_TILE;
    if (tiles > sc.cap) {
        sagpu_set_error("scan scratch too small", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)tiles), dim3(SC_THREADS), 0, s, data, count, sc.partials);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, sc.partials, tiles);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)tiles), dim3(SC_THREADS), 0, s, data, count, sc.partials);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}
