/*
 * gen.c -- host-side seeded graph generators (pair lists). The device
 * generators in kernels/gen.hip emit the same pairs for the same arguments.
 */
#include "gcx.h"
#include "gen.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static uint32_t scale_prob(double x)
{
    if (x <= 0.0)
        return 0;
    if (x >= 1.0)
        return UINT32_MAX;
    return (uint32_t)floor(x * 4294967296.0);
}

void gcx_rmat_thresholds(double a, double b, double c, uint32_t thr[3])
{
    thr[0] = scale_prob(a);
    thr[1] = scale_prob(a + b);
    thr[2] = scale_prob(a + b + c);
}

int gcx_gen_rmat_pairs(int scale, uint32_t edge_factor, double a, double b, double c,
                       uint64_t seed, uint32_t *I, uint32_t *J)
{
    if (scale < 1 || scale > 31)
        return GCX_ERR_ARG;
    uint32_t t[3];
    gcx_rmat_thresholds(a, b, c, t);
    gcx_rmat_thr thr = {t[0], t[1], t[2]};
    const uint64_t L = (uint64_t)edge_factor << scale;
#pragma omp parallel for schedule(static)
    for (uint64_t k = 0; k < L; k++)
        gcx_rmat_pair(seed, scale, thr, k, I + k, J + k);
    return GCX_OK;
}

int gcx_gen_er_pairs(uint32_t n, uint64_t m, uint64_t seed, uint32_t *I, uint32_t *J)
{
    if (n == 0)
        return GCX_ERR_ARG;
#pragma omp parallel for schedule(static)
    for (uint64_t k = 0; k < m; k++)
        gcx_er_pair(seed, n, k, I + k, J + k);
    return GCX_OK;
}

int gcx_gen_permutation(uint32_t n, uint64_t seed, uint32_t *perm)
{
    /* Fisher-Yates driven by the counter stream (seed, "PERM", i). */
    for (uint32_t i = 0; i < n; i++)
        perm[i] = i;
    for (uint32_t i = n; i > 1; i--) {
        const uint64_t h = gcx_hash3(seed, 0x5045524Dull, i - 1);
        const uint32_t j = gcx_bounded((uint32_t)(h >> 32), i);
        const uint32_t t = perm[i - 1];
        perm[i - 1] = perm[j];
        perm[j] = t;
    }
    return GCX_OK;
}

uint64_t gcx_stencil27_pair_count(uint32_t g)
{
    /* Undirected pairs: ((3g-2)^3 - g^3) / 2. */
    const uint64_t s = 3ull * g - 2ull;
    return (s * s * s - (uint64_t)g * g * g) / 2ull;
}

int gcx_gen_stencil27_pairs(uint32_t g, uint64_t seed, uint32_t *I, uint32_t *J)
{
    /* Cell (x, y, z) -> vertex perm[x + g*(y + g*z)]; one pair for each of the
     * 13 forward offsets (dz, dy, dx) > (0, 0, 0) that stay inside the grid.
     * Pairs are written in cell order, offsets in lexicographic order. */
    if (g == 0)
        return GCX_ERR_ARG;
    const uint64_t n = (uint64_t)g * g * g;
    if (n > UINT32_MAX)
        return GCX_ERR_ARG;
    uint32_t *perm = malloc(sizeof(uint32_t) * n);
    if (!perm)
        return GCX_ERR_NOMEM;
    gcx_gen_permutation((uint32_t)n, seed, perm);
    /* Pairs emitted per z-slab are a closed form, so slabs fill in parallel. */
    uint64_t *slab_off = malloc(sizeof(uint64_t) * (g + 1));
    slab_off[0] = 0;
    for (uint32_t z = 0; z < g; z++) {
        uint64_t cnt = 0;
        for (int dz = 0; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    if (dz == 0 && (dy < 0 || (dy == 0 && dx <= 0)))
                        continue;
                    if (z + dz >= g)
                        continue;
                    const uint64_t ny = g - (dy != 0), nx = g - (dx != 0);
                    cnt += ny * nx;
                }
        slab_off[z + 1] = slab_off[z] + cnt;
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (uint32_t z = 0; z < g; z++) {
        uint64_t w = slab_off[z];
        for (uint32_t y = 0; y < g; y++)
            for (uint32_t x = 0; x < g; x++) {
                const uint32_t c = perm[x + (uint64_t)g * (y + (uint64_t)g * z)];
                for (int dz = 0; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            if (dz == 0 && (dy < 0 || (dy == 0 && dx <= 0)))
                                continue;
                            const int64_t X = (int64_t)x + dx, Y = (int64_t)y + dy, Z = (int64_t)z + dz;
                            if (X < 0 || Y < 0 || X >= g || Y >= g || Z >= g)
                                continue;
                            I[w] = c;
                            J[w] = perm[X + (uint64_t)g * (Y + (uint64_t)g * Z)];
                            w++;
                        }
            }
    }
    free(slab_off);
    free(perm);
    return GCX_OK;
}

/* ------------------------------------------------------------ 2D geometric */

int gcx_gen_rgg2d_pairs(uint32_t n, double radius, uint64_t seed, uint64_t cap,
                        uint32_t *I, uint32_t *J, uint64_t *L)
{
    /* Points: 31-bit fixed-point coordinates (gcx_rgg_point). Pairs (i, j),
     * i < j, with integer squared distance < r2 = floor(radius^2 * 2^62),
     * found through a uniform cell grid of side >= radius. Output order: by
     * i, then by neighbouring cell, then by j. */
    if (n == 0 || radius <= 0.0)
        return GCX_ERR_ARG;
    uint32_t *px = malloc(sizeof(uint32_t) * n), *py = malloc(sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; i++)
        gcx_rgg_point(seed, i, px + i, py + i);
    const uint32_t cells = gcx_rgg_cells(radius);
    const uint64_t r2 = gcx_rgg_r2(radius);
    const uint64_t nc = (uint64_t)cells * cells;
    uint32_t *start = calloc(nc + 1, sizeof(uint32_t));
    uint32_t *order = malloc(sizeof(uint32_t) * n);
    uint32_t *cell_of = malloc(sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; i++) {
        cell_of[i] = gcx_rgg_cell(px[i], py[i], cells);
        start[cell_of[i] + 1]++;
    }
    for (uint64_t c = 0; c < nc; c++)
        start[c + 1] += start[c];
    uint32_t *fill = malloc(sizeof(uint32_t) * (nc + 1));
    memcpy(fill, start, sizeof(uint32_t) * (nc + 1));
    for (uint32_t i = 0; i < n; i++)
        order[fill[cell_of[i]]++] = i;
    free(fill);
    uint64_t w = 0;
    int overflow = 0;
    for (uint32_t i = 0; i < n && !overflow; i++) {
        const uint32_t c = cell_of[i], cx = c % cells, cy = c / cells;
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int64_t X = (int64_t)cx + dx, Y = (int64_t)cy + dy;
                if (X < 0 || Y < 0 || X >= cells || Y >= cells)
                    continue;
                const uint64_t cc = (uint64_t)Y * cells + (uint64_t)X;
                for (uint32_t k = start[cc]; k < start[cc + 1]; k++) {
                    const uint32_t j = order[k];
                    if (j <= i)
                        continue;
                    if (gcx_rgg_close(px[i], py[i], px[j], py[j], r2)) {
                        if (w >= cap) {
                            overflow = 1;
                            break;
                        }
                        I[w] = i;
                        J[w] = j;
                        w++;
                    }
                }
            }
    }
    *L = w;
    free(px);
    free(py);
    free(start);
    free(order);
    free(cell_of);
    return overflow ? GCX_ERR_CAPACITY : GCX_OK;
}
