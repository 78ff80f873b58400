// common.hip.h -- shared helpers for the gfx950 kernels (wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gcx.h"
#include "gcxk.h"

#define GCXK_CHECK(expr)                                                        \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess)                                                   \
            return GCX_ERR_HIP;                                                 \
    } while (0)

#define GCXK_LAUNCH_CHECK()                                                     \
    do {                                                                        \
        hipError_t e_ = hipGetLastError();                                      \
        if (e_ != hipSuccess)                                                   \
            return GCX_ERR_HIP;                                                 \
    } while (0)

static constexpr int kWave = 64;

static inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

static inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt()
{
    const int l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// OR-reduce a 64-bit value over the `width` lanes of an aligned lane group.
template <int width>
__device__ __forceinline__ uint64_t group_or64(uint64_t x)
{
#pragma unroll
    for (int off = width / 2; off > 0; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)x, off, width);
        const uint32_t hi = __shfl_xor((uint32_t)(x >> 32), off, width);
        x |= ((uint64_t)hi << 32) | lo;
    }
    return x;
}

template <int width>
__device__ __forceinline__ uint32_t group_sum32(uint32_t x)
{
#pragma unroll
    for (int off = width / 2; off > 0; off >>= 1)
        x += __shfl_xor(x, off, width);
    return x;
}

template <int width>
__device__ __forceinline__ uint32_t group_max32(uint32_t x)
{
#pragma unroll
    for (int off = width / 2; off > 0; off >>= 1) {
        const uint32_t y = __shfl_xor(x, off, width);
        x = y > x ? y : x;
    }
    return x;
}

// Order-preserving 32-bit image of a float (all non-NaN floats; -0 == +0).
__device__ __host__ __forceinline__ uint32_t float_order_key(float f)
{
    uint32_t b = __builtin_bit_cast(uint32_t, f);
    if (b == 0x80000000u)
        b = 0u;
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __host__ __forceinline__ float float_from_order_key(uint32_t k)
{
    const uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __builtin_bit_cast(float, b);
}

// Order-preserving 64-bit image of a double (non-NaN; -0 == +0).
__device__ __host__ __forceinline__ uint64_t double_order_key(double d)
{
    uint64_t b = __builtin_bit_cast(uint64_t, d);
    if (b == 0x8000000000000000ull)
        b = 0ull;
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __host__ __forceinline__ double double_from_order_key(uint64_t k)
{
    const uint64_t b = (k & 0x8000000000000000ull) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    return __builtin_bit_cast(double, b);
}

// v precedes u in the colouring order (p descending, id ascending).
template <typename K>
__device__ __forceinline__ bool precedes(K kv, uint32_t v, K ku, uint32_t u)
{
    return kv > ku || (kv == ku && v < u);
}

// SplitMix64 finaliser: the per-vertex term of the colouring digest.
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
