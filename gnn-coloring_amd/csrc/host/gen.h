/*
 * gen.h -- seeded synthetic graph generators shared by the host (C) and the
 * device (HIP) code paths. Integer-only arithmetic, so the host and device
 * versions emit bit-identical pair lists for the same arguments.
 *
 * Every generator produces 0-based vertex pairs (u, v); self loops and
 * repeats are allowed in the output and removed by the CSR builder, which
 * also symmetrises (each pair yields both directions).
 */
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define GCX_HD __host__ __device__ __forceinline__
#else
#define GCX_HD static inline
#endif

/* SplitMix64 finaliser (Steele, Lea, Flood 2014). */
GCX_HD uint64_t gcx_mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Counter-based stream: value number `idx` of sub-stream `sub` of `seed`. */
GCX_HD uint64_t gcx_hash3(uint64_t seed, uint64_t sub, uint64_t idx)
{
    return gcx_mix64(gcx_mix64(seed ^ (sub * 0xD1B54A32D192ED03ull)) + idx);
}

/* Unbiased-enough bounded draw in [0, n): multiply-shift on 32 random bits. */
GCX_HD uint32_t gcx_bounded(uint32_t r32, uint32_t n)
{
    return (uint32_t)(((uint64_t)r32 * (uint64_t)n) >> 32);
}

/* R-MAT quadrant thresholds as 32-bit integers: t0 = a, t1 = a+b, t2 = a+b+c
 * (each scaled by 2^32 and rounded down by the host before launch). */
typedef struct {
    uint32_t t0, t1, t2;
} gcx_rmat_thr;

/* Pair k of an R-MAT graph of 2^scale vertices. Level l (0 = most
 * significant bit) draws 32 bits: half of hash(seed, k, l/2). Quadrant
 * 0 = (0,0), 1 = (0,1), 2 = (1,0), 3 = (1,1) as (row bit, column bit). */
GCX_HD void gcx_rmat_pair(uint64_t seed, int scale, gcx_rmat_thr t, uint64_t k,
                          uint32_t *u, uint32_t *v)
{
    uint32_t a = 0, b = 0;
    uint64_t h = 0;
    for (int l = 0; l < scale; l++) {
        if ((l & 1) == 0)
            h = gcx_hash3(seed, k, (uint64_t)(l >> 1));
        const uint32_t r = (l & 1) ? (uint32_t)h : (uint32_t)(h >> 32);
        const uint32_t q = r < t.t0 ? 0u : r < t.t1 ? 1u : r < t.t2 ? 2u : 3u;
        a = (a << 1) | (q >> 1);
        b = (b << 1) | (q & 1u);
    }
    *u = a;
    *v = b;
}

/* Pair k of an Erdos-Renyi G(n, m) multigraph draw. */
GCX_HD void gcx_er_pair(uint64_t seed, uint32_t n, uint64_t k, uint32_t *u, uint32_t *v)
{
    const uint64_t h = gcx_hash3(seed, 0x45524552ull, k);
    *u = gcx_bounded((uint32_t)(h >> 32), n);
    *v = gcx_bounded((uint32_t)h, n);
}

/* 2D random geometric graph point i: 31-bit fixed-point coordinates. */
GCX_HD void gcx_rgg_point(uint64_t seed, uint64_t i, uint32_t *x, uint32_t *y)
{
    const uint64_t h = gcx_hash3(seed, 0x524747ull, i);
    *x = (uint32_t)(h >> 33);
    *y = (uint32_t)(h & 0x7FFFFFFFull);
}

/* Cell of a point on a cells x cells grid (row-major, y then x). */
GCX_HD uint32_t gcx_rgg_cell(uint32_t x, uint32_t y, uint32_t cells)
{
    const uint32_t cx = (uint32_t)(((uint64_t)x * cells) >> 31);
    const uint32_t cy = (uint32_t)(((uint64_t)y * cells) >> 31);
    return cy * cells + cx;
}

/* Exact integer distance test (differences < 2^31, so no overflow). */
GCX_HD int gcx_rgg_close(uint32_t xi, uint32_t yi, uint32_t xj, uint32_t yj, uint64_t r2)
{
    const uint64_t dx = xi > xj ? (uint64_t)(xi - xj) : (uint64_t)(xj - xi);
    const uint64_t dy = yi > yj ? (uint64_t)(yi - yj) : (uint64_t)(yj - yi);
    return dx * dx + dy * dy < r2;
}

#if !defined(__HIPCC__)
/* Host-side RGG parameters: grid side (cells of width >= radius) and the
 * squared radius in units of 2^-62. */
static inline uint32_t gcx_rgg_cells(double radius)
{
    double c = radius > 0.0 ? 1.0 / radius : 1.0;
    if (c < 1.0)
        c = 1.0;
    if (c > 65536.0)
        c = 65536.0;
    return (uint32_t)c;
}

static inline uint64_t gcx_rgg_r2(double radius)
{
    const double r2 = radius * radius * 4611686018427387904.0;
    return r2 >= 9.2e18 ? 9200000000000000000ull : (uint64_t)r2;
}
#endif

#undef GCX_HD
