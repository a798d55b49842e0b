// common.h -- shared device helpers for the gfx950 suffix-array engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SAGPU_WAVE 64

typedef unsigned long long u64;
typedef uint32_t u32;
typedef uint16_t u16;

// Status codes used inside the engine; they map 1:1 onto the libsais return codes.
enum { SAGPU_OK = 0, SAGPU_EINVAL = -1, SAGPU_ENOMEM = -2, SAGPU_EDEVICE = -3 };

#define SAGPU_CHECK(expr)                                                      \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess) {                                                \
            sagpu_set_error(hipGetErrorString(_e), __FILE__, __LINE__);        \
            return (_e == hipErrorOutOfMemory) ? SAGPU_ENOMEM : SAGPU_EDEVICE; \
        }                                                                      \
    } while (0)

#define SAGPU_TRY(expr)                 \
    do {                                \
        int _rc = (expr);               \
        if (_rc != SAGPU_OK) return _rc; \
    } while (0)

void sagpu_set_error(const char *msg, const char *file, int line);

__device__ __forceinline__ u32 lane_id() { return __lane_id(); }

__device__ __forceinline__ u64 lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// Inclusive wave-level scan (sum) over 64 lanes.
__device__ __forceinline__ u32 wave_incl_sum(u32 v)
{
#pragma unroll
    for (int d = 1; d < SAGPU_WAVE; d <<= 1) {
        u32 o = __shfl_up(v, d, SAGPU_WAVE);
        if ((int)lane_id() >= d) v += o;
    }
    return v;
}

__device__ __forceinline__ u32 wave_incl_max(u32 v)
{
#pragma unroll
    for (int d = 1; d < SAGPU_WAVE; d <<= 1) {
        u32 o = __shfl_up(v, d, SAGPU_WAVE);
        if ((int)lane_id() >= d) v = v > o ? v : o;
    }
    return v;
}

__device__ __forceinline__ u64 wave_incl_sum64(u64 v)
{
#pragma unroll
    for (int d = 1; d < SAGPU_WAVE; d <<= 1) {
        u64 o = __shfl_up(v, d, SAGPU_WAVE);
        if ((int)lane_id() >= d) v += o;
    }
    return v;
}

__device__ __forceinline__ u64 wave_incl_max64(u64 v)
{
#pragma unroll
    for (int d = 1; d < SAGPU_WAVE; d <<= 1) {
        u64 o = __shfl_up(v, d, SAGPU_WAVE);
        if ((int)lane_id() >= d) v = v > o ? v : o;
    }
    return v;
}

// Device-scope relaxed atomics on 64-bit look-back words.
__device__ __forceinline__ u64 ld_acquire_u64(const u64 *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_release_u64(u64 *p, u64 v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

static inline unsigned ceil_div_u(u64 a, u64 b) { return (unsigned)((a + b - 1) / b); }
