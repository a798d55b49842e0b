// pivots.hip -- few-key partition of the big groups of a doubling round.
//
// On texts of long repeats (periodic texts, runs of one symbol, Fibonacci words) a
// round's big groups hold millions of members whose key2 (the rank at i + h) takes
// only one to a few values: a group of suffixes in one residue class of a period
// maps as a whole to the group of the next class.  Sorting such members by
// (group, key2) with the radix sort moves every member through every pass, round
// after round.  Here each big group samples 64 of its members' keys (one wave per
// group) and takes up to BP values that occur at least twice in the sample as
// pivots; a stable partition (count, scan, scatter over tiles of 4096 members)
// puts the members equal to pivot b of their group into class b and the others
// into the rest, and only the rest goes through the radix sort.  Every member of a
// pivot class of one group forms one new subgroup (they share the key), placed
// after the group's rest members with smaller keys and the members of the smaller
// pivots: rank = first slot of the group + #rest below + #pivot members below.
// The rest is ranked by the usual segmented scan over the sorted keys, with each
// rank raised by the members of the pivots below its key.
//
// Tie order: members of one pivot class keep their list order; members of one
// subgroup are tied, and any order of tied members is a valid list order.
#include "engine.h"

namespace {

constexpr int PT_THREADS = 256, PT_ROWS = 16, PT_SPAN = 1024, PT_TILE = 4096;
constexpr int PE_ROWS = 16;   // rows of 64 pivot members per wave in the emission
// Rounds with at most this many big groups count the partition per class and group
// (no per-member group index)
constexpr u32 PQ_GROUPS = 32;
constexpr u32 NOQ = 0xFFFFFFFFu;

__device__ __forceinline__ u32 group_of(const u64 *__restrict__ boff, u32 nb, u64 e)
{
    u32 lo = 0, hi = nb - 1;
    while (lo < hi) {
        const u32 mid = (lo + hi + 1) >> 1;
        if (boff[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// A thread's walk over increasing member indexes e of the concatenated big groups
// (group q holds [boff[q], boff[q + 1]), list entries from bhead[q]).
struct PivCursor {
    u32 q = NOQ;
    u64 beg = 0, end = 0, head = 0;
    u32 np = 0;
    u64 p[BP];

    __device__ void to(const u64 *__restrict__ boff, const u32 *__restrict__ bhead, const u64 *__restrict__ piv,
                       const u32 *__restrict__ npiv, u32 nb, u64 e)
    {
        if (q != NOQ && e >= beg && e < end) return;
        // big groups hold more than 64 members, so a step of one wave row moves
        // at most into the next group
        if (q != NOQ && e >= end && q + 1 < nb && e < boff[q + 2]) q = q + 1;
        else q = group_of(boff, nb, e);
        beg = boff[q];
        end = boff[q + 1];
        head = bhead[q];
        np = npiv[q];
#pragma unroll
        for (int b = 0; b < BP; b++) p[b] = (u32)b < np ? piv[(u64)q * BP + b] : 0ull;
    }
    // pivot index the key equals, BP for the rest
    __device__ u32 cls(u64 k) const
    {
        u32 c = BP;
#pragma unroll
        for (int b = 0; b < BP; b++)
            if ((u32)b < np && p[b] == k) c = (u32)b;
        return c;
    }
};

// One wave per big group: 64 keys sampled at even steps over the group's members;
// the values sampled at least twice, most frequent first, at most BP, ascending.
template <typename KT>
__global__ __launch_bounds__(256) void k_big_pivots(const KT *__restrict__ key2, const u32 *__restrict__ bhead,
                                                    const u64 *__restrict__ boff, u32 nb, u64 *__restrict__ piv,
                                                    u32 *__restrict__ npiv)
{
    const u32 q = blockIdx.x * (256 / SAGPU_WAVE) + threadIdx.x / SAGPU_WAVE;
    if (q >= nb) return;   // whole waves leave together
    const u32 lane = lane_id();
    const u64 size = boff[q + 1] - boff[q];
    // size > RT_SMALL >= 64: the 64 sample positions are distinct members
    const u64 k = (u64)key2[(u64)bhead[q] + (u64)lane * size / SAGPU_WAVE];
    u32 cnt = 0;
    for (int l = 0; l < SAGPU_WAVE; l++) cnt += __shfl(k, l, SAGPU_WAVE) == k;
    u64 val[BP];
    u32 np = 0;
    for (int b = 0; b < BP; b++) {
        // the most frequent remaining value; ties to the lowest lane
        u32 score = cnt >= 2 ? (cnt << 6) | (SAGPU_WAVE - 1 - lane) : 0u;
        for (int d = 32; d >= 1; d >>= 1) {
            const u32 o = __shfl_xor(score, d, SAGPU_WAVE);
            score = o > score ? o : score;
        }
        if (score == 0) break;
        const u64 kv = __shfl(k, (int)(SAGPU_WAVE - 1 - (score & 63u)), SAGPU_WAVE);
        val[np++] = kv;
        if (k == kv) cnt = 0;
    }
    if (lane == 0) {
        for (u32 i = 1; i < np; i++)   // ascending
            for (u32 j = i; j > 0 && val[j - 1] > val[j]; j--) {
                const u64 t = val[j];
                val[j] = val[j - 1];
                val[j - 1] = t;
            }
        for (u32 b = 0; b < np; b++) piv[(u64)q * BP + b] = val[b];
        npiv[q] = np;
    }
}

// Partition pass 1: members of each class per tile, class-major: ccnt[c * tiles + t].
// With gq > 0 (gq = nb groups, at most PQ_GROUPS) the counts are per class and group,
// ccnt[(c * gq + q) * tiles + t]: after the scan, the pivot subgroups' sizes and
// starts come from the counts, and no per-member group index (pq) is written.
template <typename KT>
__global__ __launch_bounds__(PT_THREADS) void k_part_count(const KT *__restrict__ key2, const u32 *__restrict__ bhead,
                                                           const u64 *__restrict__ boff, u32 nb, u64 total,
                                                           const u64 *__restrict__ piv, const u32 *__restrict__ npiv,
                                                           u32 *__restrict__ ccnt, u64 tiles, u32 gq)
{
    __shared__ u32 s_c[(BP + 1) * PQ_GROUPS];
    const u32 nc = (BP + 1) * (gq ? gq : 1u);
    for (u32 x = threadIdx.x; x < nc; x += PT_THREADS) s_c[x] = 0;
    __syncthreads();
    const u32 lane = lane_id();
    const u64 wbase = (u64)blockIdx.x * PT_TILE + (u64)(threadIdx.x / SAGPU_WAVE) * PT_SPAN;
    PivCursor cur;
    u32 cnt[BP + 1];
#pragma unroll
    for (int b = 0; b <= BP; b++) cnt[b] = 0;
    for (int r = 0; r < PT_ROWS; r++) {
        const u64 e = wbase + (u64)r * SAGPU_WAVE + lane;
        u32 c = BP + 1, q = 0;
        if (e < total) {
            cur.to(boff, bhead, piv, npiv, nb, e);
            c = cur.cls((u64)key2[cur.head + (e - cur.beg)]);
            q = cur.q;
        }
        if (gq) {
            // a row of 64 consecutive members spans at most two groups (big groups
            // hold more than 64 members)
            const u32 q0 = __shfl(q, 0, SAGPU_WAVE);
            const u64 lo = __ballot(e < total && q == q0), hi = __ballot(e < total && q != q0);
#pragma unroll
            for (int b = 0; b <= BP; b++) {
                const u64 m = __ballot(c == (u32)b);
                if (lane == 0) {
                    const u32 c0 = (u32)__popcll(m & lo), c1 = (u32)__popcll(m & hi);
                    if (c0) atomicAdd(&s_c[(u32)b * gq + q0], c0);
                    if (c1) atomicAdd(&s_c[(u32)b * gq + q0 + 1], c1);
                }
            }
        } else {
#pragma unroll
            for (int b = 0; b <= BP; b++) cnt[b] += (u32)__popcll(__ballot(c == (u32)b));
        }
    }
    if (!gq && lane == 0)
        for (int b = 0; b <= BP; b++)
            if (cnt[b]) atomicAdd(&s_c[b], cnt[b]);
    __syncthreads();
    for (u32 x = threadIdx.x; x < nc; x += PT_THREADS) ccnt[(u64)x * tiles + blockIdx.x] = s_c[x];
}

// Partition pass 2 (after the exclusive scan of ccnt): members in tile order into
// their class.  Pivot classes go to pq (dense group index; not with gq > 0) and ppv
// (the list entry) at the scanned position, ppv being the next active list's act_i
// from the range the pivot members take there; the rest to the sort buffers at
// position - npivot, keyed (dense group index << kb) | key2 as the plain extraction
// keys them.  Groups are contiguous in member order, so with per-group counts the
// tile's class-c members start at the scanned count of (c, the tile's first group).
template <typename KT, typename CT>
__global__ __launch_bounds__(PT_THREADS) void k_part_scatter(const u32 *__restrict__ act_i, const KT *__restrict__ key2,
                                                             const u32 *__restrict__ bhead,
                                                             const u64 *__restrict__ boff, u32 nb, u64 total, int kb,
                                                             const u64 *__restrict__ piv,
                                                             const u32 *__restrict__ npiv,
                                                             const u32 *__restrict__ cscan, u64 tiles, u64 npivot,
                                                             CT *__restrict__ comp, u32 *__restrict__ vals,
                                                             u32 *__restrict__ pq, u32 *__restrict__ ppv, u32 gq)
{
    __shared__ u32 s_wc[PT_THREADS / SAGPU_WAVE][BP + 1];
    const u32 lane = lane_id(), w = threadIdx.x / SAGPU_WAVE;
    const u64 wbase = (u64)blockIdx.x * PT_TILE + (u64)w * PT_SPAN;
    const u64 lt = lanemask_lt();
    PivCursor cur;
    u64 packed = 0;   // class of each row, 3 bits per row
    u32 cnt[BP + 1];
#pragma unroll
    for (int b = 0; b <= BP; b++) cnt[b] = 0;
    for (int r = 0; r < PT_ROWS; r++) {
        const u64 e = wbase + (u64)r * SAGPU_WAVE + lane;
        u32 c = BP + 1;
        if (e < total) {
            cur.to(boff, bhead, piv, npiv, nb, e);
            c = cur.cls((u64)key2[cur.head + (e - cur.beg)]);
        }
        packed |= (u64)c << (3 * r);
#pragma unroll
        for (int b = 0; b <= BP; b++) cnt[b] += (u32)__popcll(__ballot(c == (u32)b));
    }
    if (lane == 0)
#pragma unroll
        for (int b = 0; b <= BP; b++) s_wc[w][b] = cnt[b];
    __syncthreads();
    u64 pos[BP + 1];   // next output position of each class for this wave
    const u32 qt = gq ? group_of(boff, nb, (u64)blockIdx.x * PT_TILE) : 0u;   // the tile's first group
#pragma unroll
    for (int b = 0; b <= BP; b++) {
        u64 o = cscan[((u64)b * (gq ? gq : 1u) + qt) * tiles + blockIdx.x];
        for (u32 x = 0; x < w; x++) o += s_wc[x][b];
        pos[b] = o;
    }
    for (int r = 0; r < PT_ROWS; r++) {
        const u64 e = wbase + (u64)r * SAGPU_WAVE + lane;
        const u32 c = (u32)(packed >> (3 * r)) & 7u;
        u64 mine = 0;
#pragma unroll
        for (int b = 0; b <= BP; b++) {
            const u64 m = __ballot(c == (u32)b);
            if (c == (u32)b) mine = pos[b] + (u64)__popcll(m & lt);
            pos[b] += (u64)__popcll(m);
        }
        if (c > BP) continue;
        cur.to(boff, bhead, piv, npiv, nb, e);
        const u64 k = cur.head + (e - cur.beg);
        if (c < BP) {
            if (!gq) pq[mine] = cur.q;
            ppv[mine] = act_i[k];
        } else {
            const u64 o = mine - npivot;
            comp[o] = (CT)(((u64)cur.q << kb) | (u64)key2[k]);
            vals[o] = act_i[k];
        }
    }
}

// First index in sorted a[0, n) with a[i] >= v, searched by the whole wave: 64
// probes per step, so about log64(n) dependent reads instead of log2(n).
template <typename T>
__device__ u64 lower_bound_dev(const T *__restrict__ a, u64 n, u64 v)
{
    const u32 lane = lane_id();
    u64 lo = 0, hi = n;
    while (hi - lo > SAGPU_WAVE) {
        const u64 step = (hi - lo + SAGPU_WAVE - 1) / SAGPU_WAVE;
        const u64 p = lo + (u64)lane * step;
        const u64 m = __ballot(p >= hi || (u64)a[p] >= v);
        const u32 f = m ? (u32)__builtin_ctzll(m) : (u32)SAGPU_WAVE;   // first probe at or above v
        if (f == 0) return lo;
        const u64 nhi = f < (u32)SAGPU_WAVE ? lo + (u64)f * step : hi;
        lo = lo + (u64)(f - 1) * step + 1;   // the probe before holds less than v
        hi = nhi < hi ? nhi : hi;
    }
    const u64 p = lo + lane;
    const u64 m = __ballot(p >= hi || (u64)a[p] >= v);
    return m ? lo + (u64)__builtin_ctzll(m) : hi;
}

// One wave per big group (every lane computes the same values, lane 0 writes):
// members equal to each pivot (from the pivot classes, grouped by dense group
// index), the new subgroup's word and rank.  The first wave reserves the pivot
// members' entries of the next active list.
//   The group's members hold one rank v: its first slot g, or, where the group word
//   carries the flag gf, the value stored in SA[g + 1].  The subgroup whose slots
//   hold v keeps it (no ISA writes).  A pivot subgroup that does not takes the
//   middle of its slots as its rank when flags are on (flagged, stored at its first
//   slot + 1): on texts of long repeats the pivot subgroup of one round holds
//   nearly all of the next round's group, and so keeps the middle of it.
//   rk[q, b] = the subgroup's first slot | flag, its group word in the next list;
//   rv[q, b] = the rank its members take, PIV_KEEP where it keeps the group's.
//   grank[q], keepkey[q] (flagged groups) = v and the key2 of the sorted rest's
//   subgroup whose slots hold v (BIG_NOKEY if v is a pivot subgroup's).
constexpr u32 PIV_KEEP = 0xFFFFFFFFu;   // a pivot subgroup's first slot is at most n - 2
template <typename CT>
__global__ __launch_bounds__(256) void k_piv_ranks(const CT *__restrict__ rest, u64 nrest, int kb,
                                                   const u32 *__restrict__ pq, PivRegions reg,
                                                   const u64 *__restrict__ piv, const u32 *__restrict__ npiv,
                                                   const u32 *__restrict__ gslot, u32 gf, u32 *__restrict__ SA,
                                                   u32 *__restrict__ grank, u64 *__restrict__ keepkey, u32 nb,
                                                   u32 *__restrict__ eq, u32 *__restrict__ rk, u32 *__restrict__ rv,
                                                   u64 *__restrict__ out_counter, u64 base,
                                                   u32 *__restrict__ err, const u32 *__restrict__ cscan, u64 tiles,
                                                   u32 gq)
{
    // the pivot members' range of the next list starts at base (the tile kernel's
    // survivors before it; the partition wrote their act_i there already)
    if (blockIdx.x == 0 && threadIdx.x == 0 && atomicAdd(out_counter, reg.start[BP]) != base) atomicOr(err, 8u);
    const u32 q = (blockIdx.x * blockDim.x + threadIdx.x) / SAGPU_WAVE;
    if (q >= nb) return;   // whole waves leave together
    const bool w0 = lane_id() == 0;
    const u32 np = npiv[q];
    const u32 gw = gslot[q], g = gw & ~gf;
    const bool flagged = (gw & gf) != 0;
    const u32 v = flagged ? SA[g + 1] : g;   // read before any store below
    const u64 vr = (u64)(v - g);             // the rank's position among the group's slots
    const u64 g0 = (u64)q << kb;
    const u64 rest0 = lower_bound_dev(rest, nrest, g0);
    u64 below = 0;        // pivot members of smaller keys
    u64 piv_before = 0;   // pivot members whose slots all lie before position vr
    bool in_pivot = false;
    for (u32 b = 0; b < np; b++) {
        const u64 pk = piv[(u64)q * BP + b];
        const u64 lt = lower_bound_dev(rest, nrest, g0 | pk) - rest0;
        u64 c;
        if (gq) {   // per-group counts: the scanned counts of (b, q) and (b, q + 1)
            c = (u64)cscan[((u64)b * gq + q + 1) * tiles] - (u64)cscan[((u64)b * gq + q) * tiles];
        } else {
            const u64 r0 = reg.start[b], len = reg.start[b + 1] - r0;
            c = lower_bound_dev(pq + r0, len, (u64)q + 1) - lower_bound_dev(pq + r0, len, (u64)q);
        }
        if (c < 2 && w0) atomicOr(err, 4u);   // a pivot is sampled twice, so it has two members
        const u64 pos = lt + below;     // the subgroup's first position among the group's slots
        const u32 first = g + (u32)pos;
        const bool kept = vr >= pos && vr < pos + c;
        in_pivot = in_pivot || kept;
        if (pos + c <= vr) piv_before += c;
        if (w0) {
            eq[(u64)q * BP + b] = (u32)c;
            if (kept) {
                rk[(u64)q * BP + b] = first | (gw & gf);
                rv[(u64)q * BP + b] = PIV_KEEP;
                if (flagged && first != g && c >= 2) SA[first + 1] = v;
            } else {
                const u32 mid = first + (u32)((c - 1) / 2);
                rk[(u64)q * BP + b] = first | (c >= 2 ? gf : 0u);
                rv[(u64)q * BP + b] = gf && c >= 2 ? mid : first;
                if (gf && c >= 2) SA[first + 1] = mid;
            }
        }
        below += c;
    }
    if (flagged) {
        // v among the sorted rest: its position minus the pivot members before it
        const u64 rest1 = lower_bound_dev(rest, nrest, g0 + (1ull << kb));
        const u64 t = rest0 + (vr - piv_before);
        if (w0) {
            grank[q] = v;
            keepkey[q] = in_pivot || vr < piv_before || t >= rest1 ? BIG_NOKEY : (u64)rest[t] & ((1ull << kb) - 1);
        }
    }
}

// The pivot classes' members: one subgroup per (group, pivot) in the next active
// list (the reserved range at base, in class order, whose act_i the partition
// wrote): the group words, ranks written where they change, and the split's LCP
// mark at the subgroup's first slot.
__global__ __launch_bounds__(256) void k_piv_emit(const u32 *__restrict__ pq, PivRegions reg,
                                                  const u32 *__restrict__ rk, const u32 *__restrict__ rv,
                                                  const u32 *__restrict__ gslot, u32 gf, u64 base, u32 h,
                                                  u32 *__restrict__ ISA, const u32 *__restrict__ out_i,
                                                  u32 *__restrict__ out_g, u32 *__restrict__ LCP, int cb,
                                                  const u32 *__restrict__ cscan, u64 tiles, u32 gq)
{
    if (gq) {
        // per-group counts: run ri = b * gq + q of the pivot region is [R(ri), R(ri + 1))
        // with R(ri) = cscan[ri * tiles]; a lane's members only move forward
        const u64 npv = reg.start[BP];
        const u32 lane = lane_id();
        const u64 wave = ((u64)blockIdx.x * blockDim.x + threadIdx.x) / SAGPU_WAVE;
        const u64 c0 = wave * (u64)(PE_ROWS * SAGPU_WAVE);
        if (c0 >= npv) return;   // whole waves leave together
        const u32 nrun = BP * gq;
        u32 ri = 0, w = 0, r = 0, g = 0, nv = 0;
        u64 rlo = 0, rhi = 0;   // the current run
        bool have = false;
        for (int row = 0; row < PE_ROWS; row++) {
            const u64 m = c0 + (u64)row * SAGPU_WAVE + lane;
            if (m >= npv) break;
            if (!have || m >= rhi) {
                if (!have) {   // the last run starting at or before m
                    u32 lo = 0, hi = nrun - 1;
                    while (lo < hi) {
                        const u32 mid = (lo + hi + 1) >> 1;
                        if ((u64)cscan[(u64)mid * tiles] <= m) lo = mid;
                        else hi = mid - 1;
                    }
                    ri = lo;
                } else {
                    while ((u64)cscan[(u64)(ri + 1) * tiles] <= m) ri++;
                }
                have = true;
                rlo = cscan[(u64)ri * tiles];
                rhi = cscan[(u64)(ri + 1) * tiles];
                const u32 b = ri / gq, q = ri % gq;
                w = rk[(u64)q * BP + b];
                r = w & ~gf;
                g = gslot[q] & ~gf;
                nv = rv[(u64)q * BP + b];
            }
            if (nv != PIV_KEEP) ISA[out_i[base + m] >> cb] = nv;
            out_g[base + m] = w;
            if (LCP && r != g && m == rlo) LCP[r] = LCP_UNKNOWN | (h < LCP_BOUND_CAP ? h : LCP_BOUND_CAP);
        }
        return;
    }
    // each wave takes PE_ROWS rows of 64 consecutive members, a lane the same column
    // of every row: a lane's members mostly share one subgroup, whose word and rank
    // it keeps in registers, and the member before each (for the run starts) is
    // the lane to the left's, or lane 63's of the row above
    const u64 npv = reg.start[BP];
    const u32 lane = lane_id();
    const u64 wave = ((u64)blockIdx.x * blockDim.x + threadIdx.x) / SAGPU_WAVE;
    const u64 c0 = wave * (u64)(PE_ROWS * SAGPU_WAVE);
    if (c0 >= npv) return;   // whole waves leave together
    u32 cq = NOQ, cbk = 0, w = 0, r = 0, g = 0, nv = 0;
    u32 above = c0 > 0 ? pq[c0 - 1] : NOQ;   // the member before the first row's lane 0
    for (int row = 0; row < PE_ROWS; row++) {
        const u64 m = c0 + (u64)row * SAGPU_WAVE + lane;
        const bool in = m < npv;
        const u32 q = in ? pq[m] : NOQ;
        const u32 left = __shfl_up(q, 1, SAGPU_WAVE);
        const u32 qm1 = lane == 0 ? above : left;
        above = __shfl(q, SAGPU_WAVE - 1, SAGPU_WAVE);
        if (!in) continue;
        u32 b = 0;
#pragma unroll
        for (int t = 1; t < BP; t++) b += m >= reg.start[t];
        if (q != cq || b != cbk) {
            cq = q;
            cbk = b;
            w = rk[(u64)q * BP + b];
            r = w & ~gf;
            g = gslot[q] & ~gf;
            nv = rv[(u64)q * BP + b];
        }
        if (nv != PIV_KEEP) ISA[out_i[base + m] >> cb] = nv;
        out_g[base + m] = w;
        if (LCP && r != g && (m == reg.start[b] || qm1 != q))
            LCP[r] = LCP_UNKNOWN | (h < LCP_BOUND_CAP ? h : LCP_BOUND_CAP);
    }
}

} // namespace

u64 big_part_tiles(u64 total) { return (total + PT_TILE - 1) / PT_TILE; }

int big_pivots(const void *key2, int wide, const u32 *bhead, const u64 *boff, u32 nb, u64 *piv, u32 *npiv,
               hipStream_t s)
{
    if (nb == 0) return SAGPU_OK;
    const unsigned grid = (nb + 3) / 4;
    if (wide)
        hipLaunchKernelGGL(k_big_pivots<u64>, dim3(grid), dim3(256), 0, s, (const u64 *)key2, bhead, boff, nb, piv,
                           npiv);
    else
        hipLaunchKernelGGL(k_big_pivots<u32>, dim3(grid), dim3(256), 0, s, (const u32 *)key2, bhead, boff, nb, piv,
                           npiv);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

u32 big_part_groups(u32 nb) { return nb <= PQ_GROUPS ? nb : 0u; }

int big_part_count(const void *key2, int wide, const u32 *bhead, const u64 *boff, u32 nb, u64 total, const u64 *piv,
                   const u32 *npiv, u32 *ccnt, hipStream_t s)
{
    const u64 tiles = big_part_tiles(total);
    if (total == 0 || nb == 0) return SAGPU_OK;
    const u32 gq = big_part_groups(nb);
    if (wide)
        hipLaunchKernelGGL(k_part_count<u64>, dim3((unsigned)tiles), dim3(PT_THREADS), 0, s, (const u64 *)key2, bhead,
                           boff, nb, total, piv, npiv, ccnt, tiles, gq);
    else
        hipLaunchKernelGGL(k_part_count<u32>, dim3((unsigned)tiles), dim3(PT_THREADS), 0, s, (const u32 *)key2, bhead,
                           boff, nb, total, piv, npiv, ccnt, tiles, gq);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int big_part_scatter(const u32 *act_i, const void *key2, int wide, const u32 *bhead, const u64 *boff, u32 nb,
                     u64 total, int kb, int comp32, const u64 *piv, const u32 *npiv, const u32 *cscan, u64 npivot,
                     void *comp, u32 *vals, u32 *pq, u32 *ppv, hipStream_t s)
{
    const u64 tiles = big_part_tiles(total);
    if (total == 0 || nb == 0) return SAGPU_OK;
#define PT_SCATTER(KT, CT)                                                                                          \
    hipLaunchKernelGGL((k_part_scatter<KT, CT>), dim3((unsigned)tiles), dim3(PT_THREADS), 0, s, act_i,            \
                       (const KT *)key2, bhead, boff, nb, total, kb, piv, npiv, cscan, tiles, npivot, (CT *)comp, \
                       vals, pq, ppv, big_part_groups(nb))
    if (wide) {
        if (comp32) PT_SCATTER(u64, u32);
        else PT_SCATTER(u64, u64);
    } else {
        if (comp32) PT_SCATTER(u32, u32);
        else PT_SCATTER(u32, u64);
    }
#undef PT_SCATTER
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

// The class regions' starts after the exclusive scan of the class-major tile
// counts: starts[b] = ccnt[b * tiles] for b = 0..BP (one copy to the host instead of
// BP + 1).
__global__ void k_part_starts(const u32 *__restrict__ ccnt, u64 tiles, u32 gq, u32 *__restrict__ starts)
{
    if (threadIdx.x <= BP) starts[threadIdx.x] = ccnt[(u64)threadIdx.x * (gq ? gq : 1u) * tiles];
}

int big_part_starts(const u32 *cscan, u64 total, u32 nb, u32 *starts, hipStream_t s)
{
    hipLaunchKernelGGL(k_part_starts, dim3(1), dim3(64), 0, s, cscan, big_part_tiles(total), big_part_groups(nb),
                       starts);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int big_piv_ranks(const void *rest, int comp32, int kb, u64 nrest, const u32 *pq, const PivRegions &reg,
                  const u64 *piv, const u32 *npiv, const u32 *gslot, u32 gf, u32 *SA, u32 *grank, u64 *keepkey,
                  u32 nb, u32 *eq, u32 *rk, u32 *rv, u64 *out_counter, u64 base, u32 *err, const u32 *cscan,
                  u64 total, hipStream_t s)
{
    const u64 tiles = big_part_tiles(total);
    const u32 gq = big_part_groups(nb);
    if (gf && (!grank || !keepkey)) {
        sagpu_set_error("flagged groups need their ranks and kept keys", __FILE__, __LINE__);
        return SAGPU_EDEVICE;
    }
    const unsigned waves = 256 / SAGPU_WAVE;
    const unsigned grid = (nb + waves - 1) / waves > 0 ? (nb + waves - 1) / waves : 1;
    if (comp32)
        hipLaunchKernelGGL(k_piv_ranks<u32>, dim3(grid), dim3(256), 0, s, (const u32 *)rest, nrest, kb, pq, reg, piv,
                           npiv, gslot, gf, SA, grank, keepkey, nb, eq, rk, rv, out_counter, base, err, cscan, tiles, gq);
    else
        hipLaunchKernelGGL(k_piv_ranks<u64>, dim3(grid), dim3(256), 0, s, (const u64 *)rest, nrest, kb, pq, reg, piv,
                           npiv, gslot, gf, SA, grank, keepkey, nb, eq, rk, rv, out_counter, base, err, cscan, tiles, gq);
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}

int big_piv_emit(const u32 *pq, const PivRegions &reg, const u32 *rk, const u32 *rv, const u32 *gslot, u32 gf,
                 u64 base, u64 h, u32 *ISA, const u32 *out_i, u32 *out_g, u32 *LCP, int cb, const u32 *cscan, u32 nb,
                 u64 total, hipStream_t s)
{
    const u64 npv = reg.start[BP];
    if (npv == 0) return SAGPU_OK;
    const u32 hh = h < LCP_BOUND_CAP ? (u32)h : LCP_BOUND_CAP;
    const u64 per_block = 256 / SAGPU_WAVE * (u64)(PE_ROWS * SAGPU_WAVE);
    hipLaunchKernelGGL(k_piv_emit, dim3((unsigned)((npv + per_block - 1) / per_block)), dim3(256), 0, s, pq, reg, rk,
                       rv, gslot, gf, base, hh, ISA, out_i, out_g, LCP, cb, cscan, big_part_tiles(total),
                       big_part_groups(nb));
    SAGPU_CHECK(hipGetLastError());
    return SAGPU_OK;
}
