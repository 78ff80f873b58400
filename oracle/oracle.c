/*
 * oracle.c -- CPU restatement of the reference ordering + colouring path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg.
 *
 * Built with -ffp-contract=off so that every multiply-add below is exactly
 * the operation written (explicit fmaf where the reference uses FMA).
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define FEAT 16            /* fast_gnn.c:6 (DIM) */
#define LAYER_PARAMS 528   /* 32*16 weights + 16 biases, fast_gnn.c:256 */

void orc_free(void *ptr) { free(ptr); }

/* ---------------------------------------------------------------- validate */

int orc_validate(uint32_t N, const uint32_t *V, const uint32_t *E, const int *color)
{
    /* coloring.c:127-146: every colour positive, no monochromatic edge;
     * returns the largest colour (at least 1) or 0 when improper. */
    int bad = 0, top = 1;
    for (uint32_t u = 0; u < N; u++) {
        const int cu = color[u];
        if (cu <= 0)
            bad = 1;
        for (uint32_t i = V[u]; i < V[u + 1]; i++)
            if (color[E[i]] == cu)
                bad = 1;
        if (cu > top)
            top = cu;
    }
    return bad ? 0 : top;
}

/* v precedes u in the colouring order: coloring.c:72 / 249 */
static inline int precedes(const double *p, uint32_t v, uint32_t u)
{
    return p[v] > p[u] || (p[v] == p[u] && v < u);
}

/* ------------------------------------------------------------------ greedy */

void orc_greedy(uint32_t N, const uint32_t *V, const uint32_t *E, const double *p,
                int *color, uint32_t *rounds)
{
    /* coloring.c:60-125. Round-structured greedy: a vertex is coloured in the
     * round after its last predecessor; its colour is the smallest c >= 1 not
     * used by any predecessor. Equivalent to sequential greedy in the order
     * (p descending, id ascending). */
    uint32_t *pending = malloc(sizeof(uint32_t) * (N ? N : 1));
    uint32_t *cur = malloc(sizeof(uint32_t) * (N ? N : 1));
    uint32_t *nxt = malloc(sizeof(uint32_t) * (N ? N : 1));
    uint32_t maxdeg = 0, ncur = 0, depth = 0;

    for (uint32_t u = 0; u < N; u++) {
        uint32_t c = 0;
        for (uint32_t i = V[u]; i < V[u + 1]; i++)
            c += precedes(p, E[i], u);
        pending[u] = c;
        color[u] = 0;
        if (c == 0)
            cur[ncur++] = u;
        if (V[u + 1] - V[u] > maxdeg)
            maxdeg = V[u + 1] - V[u];
    }

    /* stamp[c] == u marks colour c as taken for the vertex u being coloured;
     * sized maxdeg + 2 so that the mex of maxdeg distinct colours fits. */
    uint32_t *stamp = malloc(sizeof(uint32_t) * (maxdeg + 2));
    for (uint32_t i = 0; i < maxdeg + 2; i++)
        stamp[i] = UINT32_MAX;

    while (ncur > 0) {
        uint32_t nn = 0;
        depth++;
        for (uint32_t k = 0; k < ncur; k++) {
            const uint32_t u = cur[k];
            for (uint32_t i = V[u]; i < V[u + 1]; i++) {
                const uint32_t v = E[i];
                if (precedes(p, v, u))
                    stamp[color[v]] = u;
            }
            int c = 1;
            while (stamp[c] == u)
                c++;
            color[u] = c;
            for (uint32_t i = V[u]; i < V[u + 1]; i++) {
                const uint32_t v = E[i];
                if (!precedes(p, v, u) && --pending[v] == 0)
                    nxt[nn++] = v;
            }
        }
        uint32_t *t = cur;
        cur = nxt;
        nxt = t;
        ncur = nn;
    }
    if (rounds)
        *rounds = depth;
    free(stamp);
    free(pending);
    free(cur);
    free(nxt);
}

uint64_t orc_mex_violations(uint32_t N, const uint32_t *V, const uint32_t *E,
                            const double *p, const int *color)
{
    /* Certificate of the greedy rule: color[u] == mex over predecessors. */
    uint64_t bad = 0;
#pragma omp parallel
    {
        uint32_t cap = 64;
        uint8_t *seen = calloc(cap, 1);
#pragma omp for reduction(+ : bad) schedule(dynamic, 4096)
        for (uint32_t u = 0; u < N; u++) {
            uint32_t deg = V[u + 1] - V[u];
            if (deg + 2 > cap) {
                free(seen);
                cap = deg + 2;
                seen = calloc(cap, 1);
            }
            for (uint32_t i = V[u]; i < V[u + 1]; i++) {
                uint32_t v = E[i];
                if (precedes(p, v, u) && color[v] > 0 && (uint32_t)color[v] < cap)
                    seen[color[v]] = 1;
            }
            uint32_t m = 1;
            while (m < cap && seen[m])
                m++;
            if ((uint32_t)color[u] != m)
                bad++;
            for (uint32_t i = V[u]; i < V[u + 1]; i++) {
                uint32_t v = E[i];
                if (color[v] > 0 && (uint32_t)color[v] < cap)
                    seen[color[v]] = 0;
            }
        }
        free(seen);
    }
    return bad;
}

/* -------------------------------------------------------------------- keys */

void orc_key_ff(uint32_t N, double *p)
{
    for (uint32_t i = 0; i < N; i++)
        p[i] = (double)(N - i);          /* first_fit.c:10-11 */
}

void orc_key_lf(uint32_t N, const uint32_t *V, double *p)
{
    for (uint32_t u = 0; u < N; u++)
        p[u] = (double)(V[u + 1] - V[u]); /* largest_degree_first.c:14-15 */
}

void orc_key_llf(uint32_t N, const uint32_t *V, double *p)
{
    for (uint32_t u = 0; u < N; u++) {   /* largest_degree_first.c:87-97 */
        uint32_t d = V[u + 1] - V[u];
        uint32_t bits = 0;
        while (d) {                      /* floor(log2 d) + 1 */
            bits++;
            d >>= 1;
        }
        p[u] = (double)bits;
    }
}

/* ------------------------------------------------------------------ SL/SLL */

void orc_sl(uint32_t N, const uint32_t *V, const uint32_t *E, double *p, uint32_t *rounds)
{
    /* smallest_degree_last.c:62-124. Rounds with threshold delta: every
     * remaining vertex whose round-start degree <= delta is removed with
     * p = delta; unremoved neighbours of removed vertices lose one degree per
     * removed neighbour; delta <- max(delta + 1, min remaining degree). */
    uint32_t *deg = malloc(sizeof(uint32_t) * (N ? N : 1));
    uint32_t *alive = malloc(sizeof(uint32_t) * (N ? N : 1));
    uint32_t *keep = malloc(sizeof(uint32_t) * (N ? N : 1));
    uint32_t *gone = malloc(sizeof(uint32_t) * (N ? N : 1));
    uint32_t lo = N, nalive = N, r = 0;
    for (uint32_t u = 0; u < N; u++) {
        deg[u] = V[u + 1] - V[u];
        alive[u] = u;
        p[u] = -1.0;
        if (deg[u] < lo)
            lo = deg[u];
    }
    uint32_t delta = lo;
    while (nalive > 0) {
        uint32_t nkeep = 0, ngone = 0;
        lo = N;
        r++;
        for (uint32_t k = 0; k < nalive; k++) {
            uint32_t u = alive[k];
            if (deg[u] <= delta) {
                gone[ngone++] = u;
                p[u] = (double)delta;
            } else {
                keep[nkeep++] = u;
                if (deg[u] < lo)
                    lo = deg[u];
            }
        }
        for (uint32_t k = 0; k < ngone; k++) {
            uint32_t u = gone[k];
            for (uint32_t i = V[u]; i < V[u + 1]; i++) {
                uint32_t v = E[i];
                if (p[v] >= 0.0)
                    continue;
                deg[v]--;
                if (deg[v] < lo)
                    lo = deg[v];
            }
        }
        delta = delta + 1 > lo ? delta + 1 : lo;
        uint32_t *t = alive;
        alive = keep;
        keep = t;
        nalive = nkeep;
    }
    if (rounds)
        *rounds = r;
    free(deg);
    free(alive);
    free(keep);
    free(gone);
}

void orc_sll(uint32_t N, const uint32_t *V, const uint32_t *E, double *p, uint32_t *rounds)
{
    /* smallest_degree_last.c:274-328. Threshold 2^delta, delta starts at the
     * smallest value with 2^delta >= min degree and grows by one per round. */
    uint32_t *deg = malloc(sizeof(uint32_t) * (N ? N : 1));
    uint32_t *alive = malloc(sizeof(uint32_t) * (N ? N : 1));
    uint32_t *keep = malloc(sizeof(uint32_t) * (N ? N : 1));
    uint32_t *gone = malloc(sizeof(uint32_t) * (N ? N : 1));
    uint32_t lo = N, nalive = N, r = 0;
    for (uint32_t u = 0; u < N; u++) {
        deg[u] = V[u + 1] - V[u];
        alive[u] = u;
        p[u] = -1.0;
        if (deg[u] < lo)
            lo = deg[u];
    }
    uint32_t delta = 0;
    while ((1ull << delta) < lo)
        delta++;
    while (nalive > 0) {
        uint32_t nkeep = 0, ngone = 0;
        const uint64_t thr = 1ull << delta;
        r++;
        for (uint32_t k = 0; k < nalive; k++) {
            uint32_t u = alive[k];
            if (deg[u] <= thr) {
                gone[ngone++] = u;
                p[u] = (double)delta;
            } else {
                keep[nkeep++] = u;
            }
        }
        for (uint32_t k = 0; k < ngone; k++) {
            uint32_t u = gone[k];
            for (uint32_t i = V[u]; i < V[u + 1]; i++)
                if (p[E[i]] < 0.0)
                    deg[E[i]]--;
        }
        delta++;
        uint32_t *t = alive;
        alive = keep;
        keep = t;
        nalive = nkeep;
    }
    if (rounds)
        *rounds = r;
    free(deg);
    free(alive);
    free(keep);
    free(gone);
}

/* -------------------------------------------------------------- model file */

int orc_load_model(const char *path, int *params, int *md, float **param)
{
    /* gnn.c:20-44: "params md" then params decimal floats (strtof semantics). */
    FILE *f = fopen(path, "r");
    *param = NULL;
    if (!f)
        return -1;
    if (fscanf(f, "%d %d", params, md) != 2 || *params <= 0 || *params % LAYER_PARAMS) {
        fclose(f);
        return -2;
    }
    float *w = malloc(sizeof(float) * (size_t)(*params));
    for (int i = 0; i < *params; i++) {
        if (fscanf(f, "%f", w + i) != 1) {
            free(w);
            fclose(f);
            return -2;
        }
    }
    fclose(f);
    *param = w;
    return 0;
}

/* --------------------------------------------------------------- GNN, FP32 */

static inline float relu_f(float c) { return c > 0.0f ? c : 0.0f; }

/* One dense layer on a 32-wide input row: c_j = b_j, then fmaf(W[i][j],
 * A[i], c_j) for i = 0..31 in order (fast_gnn.c:91-135), then ReLU
 * (fast_gnn.c:137-148: max(c, +0) returns +0 for c <= 0 and for NaN). */
static inline void dense_f(const float *A, const float *W, const float *b, float *h)
{
    for (int j = 0; j < FEAT; j++) {
        float c = b[j];
        for (int i = 0; i < 2 * FEAT; i++)
            c = fmaf(W[i * FEAT + j], A[i], c);
        h[j] = relu_f(c);
    }
}

static void gnn_fp32_input(const float *W, const float *b, uint32_t N, const uint32_t *V,
                           const uint32_t *E, int md, float *out)
{
    const float fmd = (float)md;
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint32_t u = 0; u < N; u++) {
        /* fast_gnn.c:49-89: neighbour means of deg/md and id/N, then self. */
        float sd = 0.0f, si = 0.0f;
        for (uint32_t i = V[u]; i < V[u + 1]; i++) {
            const uint32_t v = E[i];
            sd = sd + (float)(V[v + 1] - V[v]) / fmd;
            si = si + (float)((double)v / (double)N);
        }
        const float d = (float)(V[u + 1] - V[u]);
        if (d > 0.0f) {
            sd = sd / d;
            si = si / d;
        }
        float A[2 * FEAT];
        const float selfd = d / fmd, selfi = (float)((double)u / (double)N);
        for (int k = 0; k < 8; k++) {
            A[k] = sd;
            A[8 + k] = si;
            A[16 + k] = selfd;
            A[24 + k] = selfi;
        }
        dense_f(A, W, b, out + (size_t)u * FEAT);
    }
}

static void gnn_fp32_aggregate(uint32_t u, const uint32_t *V, const uint32_t *E,
                               const float *x, float *A)
{
    /* fast_gnn.c:13-47: feature-wise left-to-right neighbour sums, divided by
     * the degree when it is positive, then the vertex's own row. */
    float s[FEAT];
    for (int j = 0; j < FEAT; j++)
        s[j] = 0.0f;
    for (uint32_t i = V[u]; i < V[u + 1]; i++) {
        const float *r = x + (size_t)E[i] * FEAT;
        for (int j = 0; j < FEAT; j++)
            s[j] = s[j] + r[j];
    }
    const float d = (float)(V[u + 1] - V[u]);
    for (int j = 0; j < FEAT; j++) {
        A[j] = d > 0.0f ? s[j] / d : s[j];
        A[FEAT + j] = x[(size_t)u * FEAT + j];
    }
}

void orc_gnn_fp32(int params, const float *param, uint32_t N, const uint32_t *V,
                  const uint32_t *E, int md, double *p)
{
    /* fast_gnn.c:252-295: layer 0 input, layers 1..L-2 hidden, L-1 output. */
    const int L = params / LAYER_PARAMS;
    float *x = malloc(sizeof(float) * FEAT * ((size_t)N + 1));
    float *y = malloc(sizeof(float) * FEAT * ((size_t)N + 1));
    gnn_fp32_input(param, param + 2 * FEAT * FEAT, N, V, E, md, x);
    for (int l = 1; l < L; l++) {
        const float *W = param + (size_t)l * LAYER_PARAMS;
        const float *b = W + 2 * FEAT * FEAT;
        const int last = (l == L - 1);
#pragma omp parallel for schedule(dynamic, 1024)
        for (uint32_t u = 0; u < N; u++) {
            float A[2 * FEAT], h[FEAT];
            gnn_fp32_aggregate(u, V, E, x, A);
            dense_f(A, W, b, h);
            if (!last) {
                memcpy(y + (size_t)u * FEAT, h, sizeof(h));
            } else {
                /* fast_gnn.c:223-242: halves added, then the hadd tree. */
                float t[8];
                for (int k = 0; k < 8; k++)
                    t[k] = h[k] + h[k + 8];
                const float s = ((t[0] + t[1]) + (t[2] + t[3])) + ((t[4] + t[5]) + (t[6] + t[7]));
                p[u] = (double)s;
            }
        }
        float *tmp = x;
        x = y;
        y = tmp;
    }
    free(x);
    free(y);
}

/* --------------------------------------------------------------- GNN, FP64 */

void orc_gnn_fp64(int params, const float *param, uint32_t N, const uint32_t *V,
                  const uint32_t *E, int md, double *p)
{
    const int L = params / LAYER_PARAMS;
    double *x = malloc(sizeof(double) * FEAT * ((size_t)N + 1));
    double *y = malloc(sizeof(double) * FEAT * ((size_t)N + 1));
    for (int l = 0; l < L; l++) {
        const float *W = param + (size_t)l * LAYER_PARAMS;
        const float *b = W + 2 * FEAT * FEAT;
        const int last = (l == L - 1);
#pragma omp parallel for schedule(dynamic, 1024)
        for (uint32_t u = 0; u < N; u++) {
            double A[2 * FEAT];
            const double d = (double)(V[u + 1] - V[u]);
            if (l == 0) {
                double sd = 0.0, si = 0.0;
                for (uint32_t i = V[u]; i < V[u + 1]; i++) {
                    const uint32_t v = E[i];
                    sd += (double)(V[v + 1] - V[v]) / (double)md;
                    si += (double)v / (double)N;
                }
                if (d > 0.0) {
                    sd /= d;
                    si /= d;
                }
                for (int k = 0; k < 8; k++) {
                    A[k] = sd;
                    A[8 + k] = si;
                    A[16 + k] = d / (double)md;
                    A[24 + k] = (double)u / (double)N;
                }
            } else {
                double s[FEAT] = {0};
                for (uint32_t i = V[u]; i < V[u + 1]; i++)
                    for (int j = 0; j < FEAT; j++)
                        s[j] += x[(size_t)E[i] * FEAT + j];
                for (int j = 0; j < FEAT; j++) {
                    A[j] = d > 0.0 ? s[j] / d : s[j];
                    A[FEAT + j] = x[(size_t)u * FEAT + j];
                }
            }
            double h[FEAT];
            for (int j = 0; j < FEAT; j++) {
                double c = (double)b[j];
                for (int i = 0; i < 2 * FEAT; i++)
                    c += (double)W[i * FEAT + j] * A[i];
                h[j] = c > 0.0 ? c : 0.0;
            }
            if (!last) {
                memcpy(y + (size_t)u * FEAT, h, sizeof(h));
            } else {
                double s = 0.0;
                for (int j = 0; j < FEAT; j++)
                    s += h[j];
                p[u] = s;
            }
        }
        double *tmp = x;
        x = y;
        y = tmp;
    }
    free(x);
    free(y);
}

/* --------------------------------------------------------------------- CSR */

static int cmp_u32(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

void orc_csr_from_pairs(uint32_t N, uint64_t L, const uint32_t *I, const uint32_t *J,
                        uint32_t **Vout, uint32_t **Eout)
{
    /* fast_io.c:171-202: every pair contributes both directions; then each
     * row is sorted (450-455) and self loops / repeats dropped (457-477). */
    uint64_t *cnt = calloc((size_t)N + 1, sizeof(uint64_t));
    for (uint64_t k = 0; k < L; k++) {
        cnt[I[k] + 1]++;
        cnt[J[k] + 1]++;
    }
    for (uint32_t u = 0; u < N; u++)
        cnt[u + 1] += cnt[u];
    uint32_t *tmp = malloc(sizeof(uint32_t) * (cnt[N] ? cnt[N] : 1));
    uint64_t *pos = malloc(sizeof(uint64_t) * ((size_t)N + 1));
    memcpy(pos, cnt, sizeof(uint64_t) * ((size_t)N + 1));
    for (uint64_t k = 0; k < L; k++) {
        tmp[pos[I[k]]++] = J[k];
        tmp[pos[J[k]]++] = I[k];
    }
    free(pos);
    uint32_t *V = malloc(sizeof(uint32_t) * ((size_t)N + 1));
    uint64_t w = 0;
    for (uint32_t u = 0; u < N; u++) {
        uint32_t *row = tmp + cnt[u];
        const uint64_t n = cnt[u + 1] - cnt[u];
        qsort(row, n, sizeof(uint32_t), cmp_u32);
        V[u] = (uint32_t)w;
        for (uint64_t i = 0; i < n; i++) {
            if (row[i] == u || (i > 0 && row[i] == row[i - 1]))
                continue;
            tmp[w++] = row[i];
        }
    }
    V[N] = (uint32_t)w;
    free(cnt);
    *Vout = V;
    *Eout = realloc(tmp, sizeof(uint32_t) * (w ? w : 1));
}

static int has_edge(const uint32_t *V, const uint32_t *E, uint32_t a, uint32_t b)
{
    uint32_t lo = V[a], hi = V[a + 1];
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (E[mid] == b)
            return 1;
        if (E[mid] < b)
            lo = mid + 1;
        else
            hi = mid;
    }
    return 0;
}

uint32_t orc_graph_validate(uint32_t N, const uint32_t *V, const uint32_t *E)
{
    /* fast_io.c:509-540 */
    uint64_t M = 0;
    for (uint32_t u = 0; u < N; u++) {
        if (V[u + 1] < V[u])
            return 0;
        M += V[u + 1] - V[u];
        for (uint32_t i = V[u]; i < V[u + 1]; i++) {
            if (i >= V[N])
                return 0;
            uint32_t v = E[i];
            if (v >= N || v == u || (i > V[u] && v <= E[i - 1]))
                return 0;
            if (!has_edge(V, E, v, u))
                return 0;
        }
    }
    return M == V[N];
}

/* -------------------------------------------------------------- Culberson */

void orc_culberson(uint32_t N, const uint32_t *V, const uint32_t *E, int K,
                   double *p, int *color, int *colors_per_iter)
{
    /* main.c:32-45 / runner.c:25-36 */
    orc_greedy(N, V, E, p, color, NULL);
    if (colors_per_iter)
        colors_per_iter[0] = orc_validate(N, V, E, color);
    for (int k = 1; k < K; k++) {
        for (uint32_t j = 0; j < N; j++)
            p[j] = (double)color[j];
        orc_greedy(N, V, E, p, color, NULL);
        if (colors_per_iter)
            colors_per_iter[k] = orc_validate(N, V, E, color);
    }
}
