// seg.hip -- group segmentation after each sort (reduce-then-scan, 3 kernels).
//
// For a sorted run of keys, an element starts a group when its key differs from the
// previous one and ends a group when the next key differs.  Every element receives
// its group head (the SA slot of the group's first element: a max-scan over head
// slots, which are increasing) and its output index in the compacted list of
// still-unsorted elements (a sum-scan over "not a singleton").  Kernel 1 reduces
// each 4096-element tile, kernel 2 scans the tile partials, kernel 3 re-derives the
// flags, scans inside the tile with the carried-in partials and emits:
//   ISA[i] = group head      (for every element)
//   SA[slot] = i             (for singletons: their final position)
//   act_{i,g,p}[o] = ...     (for members of groups that still need sorting)
#include "engine.h"

namespace {

constexpr int SG_WAVES = SG_THREADS / SAGPU_WAVE;

// Initial grouping over the packed-symbol keys.
struct InitF {
    const u64 *keys;
    const u32 *vals;
    u64 n;
    u64 sepmask;
    u32 *SA, *ISA, *act_i, *act_g, *act_p;

    __device__ bool sep(u64 k) const { return sepmask != 0 && (k & sepmask) == 0; }
    __device__ void flags(u64 j, bool &head, bool &active, u32 &slot) const
    {
        u64 k = keys[j];
        bool s = sep(k);
        head = (j == 0) || s || keys[j - 1] != k;
        bool tail = (j + 1 == n) || s || keys[j + 1] != k;
        active = !(head && tail);
        slot = (u32)j;
    }
    __device__ void emit(u64 j, u32 g, bool active, u32 o) const
    {
        u32 v = vals[j];
        ISA[v] = g;
        if (active) {
            act_i[o] = v;
            act_g[o] = g;
            act_p[o] = (u32)j;
        } else {
            SA[j] = v;
        }
    }
};

// Round grouping over (old group << 32 | key2).
struct RoundF {
    const u64 *comp;
    const u32 *vals;
    const u32 *act_p;
    u64 n;
    u32 *SA, *ISA, *act_i, *act_g, *nact_p;

    __device__ void flags(u64 j, bool &head, bool &active, u32 &slot) const
    {
        u64 k = comp[j];
        head = (j == 0) || comp[j - 1] != k;
        bool tail = (j + 1 == n) || comp[j + 1] != k;
        active = !(head && tail);
        slot = act_p[j];
    }
    __device__ void emit(u64 j, u32 g, bool active, u32 o) const
    {
        u32 v = vals[j];
        ISA[v] = g;
        u32 p = act_p[j];
        if (active) {
            act_i[o] = v;
            act_g[o] = g;
            nact_p[o] = p;
        } else {
            SA[p] = v;
        }
    }
};

template <class F>
__global__ __launch_bounds__(SG_THREADS) void k_seg_reduce(F f, u32 *tile_cnt, u32 *tile_head)
{
    __shared__ u32 s_c[SG_WAVES], s_h[SG_WAVES];
    const u64 base = (u64)blockIdx.x * SG_TILE + (u64)threadIdx.x * SG_ITEMS;
    u32 cnt = 0, hmax = 0;
    for (int j = 0; j < SG_ITEMS; j++) {
        u64 e = base + j;
        if (e < f.n) {
            bool head, active;
            u32 slot;
            f.flags(e, head, active, slot);
            cnt += active;
            if (head) hmax = slot;
        }
    }
    u32 c = wave_incl_sum(cnt), h = wave_incl_max(hmax);
    const u32 w = threadIdx.x / SAGPU_WAVE;
    if (lane_id() == SAGPU_WAVE - 1) { s_c[w] = c; s_h[w] = h; }
    __syncthreads();
    if (threadIdx.x == 0) {
        u32 tc = 0, th = 0;
        for (int q = 0; q < SG_WAVES; q++) { tc += s_c[q]; th = th > s_h[q] ? th : s_h[q]; }
        tile_cnt[blockIdx.x] = tc;
        tile_head[blockIdx.x] = th;
    }
}

// Exclusive scan (sum) / exclusive scan (max) over tile partials; one workgroup.
__global__ __launch_bounds__(1024) void k_seg_scan_tiles(u32 *tile_cnt, u32 *tile_head, u64 tiles, u32 *total)
{
    __shared__ u32 s_c[16], s_h[16];
    __shared__ u32 s_carry_c, s_carry_h;
    if (threadIdx.x == 0) { s_carry_c = 0; s_carry_h = 0; }
    __syncthreads();
    for (u64 base = 0; base < tiles; base += 1024) {
        u64 i = base + threadIdx.x;
        u32 c = i < tiles ? tile_cnt[i] : 0u, h = i < tiles ? tile_head[i] : 0u;
        u32 ic = wave_incl_sum(c), ih = wave_incl_max(h);
        const u32 w = threadIdx.x / SAGPU_WAVE;
        if (lane_id() == SAGPU_WAVE - 1) { s_c[w] = ic; s_h[w] = ih; }
        __syncthreads();
        u32 cc = s_carry_c, ch = s_carry_h;
        for (u32 q = 0; q < w; q++) { cc += s_c[q]; ch = ch > s_h[q] ? ch : s_h[q]; }
        // exclusive max: max of everything strictly before i
        u32 inner = __shfl_up(ih, 1, SAGPU_WAVE);
        if (lane_id() > 0) ch = ch > inner ? ch : inner;
        if (i < tiles) {
            tile_cnt[i] = cc + ic - c;
            tile_head[i] = ch;
        }
        __syncthreads();
        if (threadIdx.x == 1023) {
            u32 tc = s_carry_c, th = s_carry_h;
            for (int q = 0; q < 16; q++) { tc += s_c[q]; th = th > s_h[q] ? th : s_h[q]; }
            s_carry_c = tc;
            s_carry_h = th;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = s_carry_c;
}

template <class F>
__global__ __launch_bounds__(SG_THREADS) void k_seg_emit(F f, const u32 *tile_cnt, const u32 *tile_head)
{
    __shared__ u32 s_c[SG_WAVES], s_h[SG_WAVES];
    const u64 base = (u64)blockIdx.x * SG_TILE + (u64)threadIdx.x * SG_ITEMS;
    bool head[SG_ITEMS], active[SG_ITEMS];
    u32 slot[SG_ITEMS];
    u32 cnt = 0, hmax = 0;
#pragma unroll
    for (int j = 0; j < SG_ITEMS; j++) {
        u64 e = base + j;
        head[j] = false;
        active[j] = false;
        slot[j] = 0;
        if (e < f.n) {
            f.flags(e, head[j], active[j], slot[j]);
            cnt += active[j];
            if (head[j]) hmax = slot[j];
        }
    }
    u32 ic = wave_incl_sum(cnt), ih = wave_incl_max(hmax);
    const u32 w = threadIdx.x / SAGPU_WAVE;
    if (lane_id() == SAGPU_WAVE - 1) { s_c[w] = ic; s_h[w] = ih; }
    __syncthreads();
    u32 cc = tile_cnt[blockIdx.x], ch = tile_head[blockIdx.x];
    for (u32 q = 0; q < w; q++) { cc += s_c[q]; ch = ch > s_h[q] ? ch : s_h[q]; }
    u32 o = cc + ic - cnt;
    u32 prev = __shfl_up(ih, 1, SAGPU_WAVE);
    u32 g = ch;
    if (lane_id() > 0) g = g > prev ? g : prev;
#pragma unroll
    for (int j = 0; j < SG_ITEMS; j++) {
        u64 e = base + j;
        if (e < f.n) {
            if (head[j]) g = slot[j];
            f.emit(e, g, active[j], o);
            o += active[j];
        }
    }
}

template <class F>
int seg_run(SegScratch &ss, const F &f, u64 n, hipStream_t s)
{
    const u64 tiles = (n + SG_TILE - 1) / SG_TILE;
    if (tiles > ss.tiles_cap) {
        sagpu_set_error("segment scratch too small", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }
    hipLaunchKernelGGL(k_seg_reduce<F>, dim3((unsigned)tiles), dim3(SG_THREADS), 0, s, f, ss.tile_cnt, ss.tile_head);
    hipLaunchKernelGGL(k_seg_scan_tiles, dim3(1), dim3(1024), 0, s, ss.tile_cnt, ss.tile_head, tiles, ss.total);
    hipLaunchKernelGGL(k_seg_emit<F>, dim3((unsigned)tiles), dim3(SG_THREADS), 0, s, f, ss.tile_cnt, ss.tile_head);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

} // namespace

int seg_init(SegScratch &ss, const u64 *keys, const u32 *vals, u64 n, u64 sepmask, u32 *SA, u32 *ISA, u32 *act_i,
             u32 *act_g, u32 *act_p, hipStream_t s)
{
    InitF f{keys, vals, n, sepmask, SA, ISA, act_i, act_g, act_p};
    return seg_run(ss, f, n, s);
}

int seg_round(SegScratch &ss, const u64 *comp, const u32 *vals, const u32 *act_p, u64 a, u32 *SA, u32 *ISA,
              u32 *nact_i, u32 *nact_g, u32 *nact_p, hipStream_t s)
{
    RoundF f{comp, vals, act_p, a, SA, ISA, nact_i, nact_g, nact_p};
    return seg_run(ss, f, a, s);
}
