/*
 * formats.c -- the reference's other graph file formats (fast_io.h:41-122,
 * fast_io.c:239-437): adjacency graphs, PACE .gr and METIS .graph.
 *
 * Parsing is text I/O on the host. The .gr edge list goes to the GPU like a
 * .mtx pair list and comes back as the canonical CSR (rows sorted, no self
 * loops, no repeats); adjacency and METIS files already are a CSR and are
 * returned as written, as the reference does. Malformed or truncated input
 * prints "gcx: <format> parse failed" and exits with status 1 (the reference
 * reads past the end of the mapping instead).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fast_io.h"
#include "internal.h"

static void die(const char *what, int rc)
{
    fprintf(stderr, "gcx: %s failed: %s\n", what, gcx_status_string(rc));
    exit(1);
}

typedef struct {
    char *d;
    size_t n, p;
} text;

static text slurp(FILE *f, const char *what)
{
    text t = {NULL, 0, 0};
    if (fseek(f, 0, SEEK_END) != 0)
        die(what, GCX_ERR_IO);
    const long size = ftell(f);
    if (size < 0 || fseek(f, 0, SEEK_SET) != 0)
        die(what, GCX_ERR_IO);
    t.n = (size_t)size;
    t.d = malloc(t.n + 1);
    if (!t.d)
        die(what, GCX_ERR_NOMEM);
    if (t.n && fread(t.d, 1, t.n, f) != t.n)
        die(what, GCX_ERR_IO);
    t.d[t.n] = '\n';
    return t;
}

/* fast_io.c:7-15: skip non-digits, read a decimal. */
static uint32_t next_id(text *t, const char *what)
{
    while (t->p < t->n && (t->d[t->p] < '0' || t->d[t->p] > '9'))
        t->p++;
    if (t->p >= t->n)
        die(what, GCX_ERR_FORMAT);
    uint64_t v = 0;
    while (t->p < t->n && t->d[t->p] >= '0' && t->d[t->p] <= '9') {
        v = v * 10 + (uint64_t)(t->d[t->p++] - '0');
        if (v > UINT32_MAX)
            die(what, GCX_ERR_FORMAT);
    }
    return (uint32_t)v;
}

/* fast_io.c:17-22: past the next newline. */
static void skip_line(text *t)
{
    while (t->p < t->n && t->d[t->p] != '\n')
        t->p++;
    t->p++;
}

static void *xmalloc(size_t bytes, const char *what)
{
    void *p = malloc(bytes ? bytes : 1);
    if (!p)
        die(what, GCX_ERR_NOMEM);
    return p;
}

void adj_parse(FILE *f, uint32_t *N, uint32_t **V, uint32_t **E)
{
    /* fast_io.c:239-262: "AdjacencyGraph", N, M, N offsets, M targets */
    const char *what = "adjacency graph parse";
    text t = slurp(f, what);
    *N = next_id(&t, what);
    const uint32_t M = next_id(&t, what);
    *V = xmalloc(sizeof(uint32_t) * ((size_t)*N + 1), what);
    for (uint32_t j = 0; j < *N; j++)
        (*V)[j] = next_id(&t, what);
    (*V)[*N] = M;
    *E = xmalloc(sizeof(uint32_t) * (size_t)M, what);
    for (uint32_t j = 0; j < M; j++)
        (*E)[j] = next_id(&t, what);
    free(t.d);
}

void adj_store(FILE *f, uint32_t N, const uint32_t *V, const uint32_t *E)
{
    /* fast_io.c:264-275 */
    fprintf(f, "AdjacencyGraph\n");
    fprintf(f, "%d\n", N);
    fprintf(f, "%d\n", V[N]);
    for (uint32_t i = 0; i < N; i++)
        fprintf(f, "%d\n", V[i]);
    for (uint32_t i = 0; i < V[N]; i++)
        fprintf(f, "%d\n", E[i]);
}

void gr_parse(FILE *f, uint32_t *N, uint32_t **V, uint32_t **E)
{
    /* fast_io.c:283-354: 'c' comment lines, "p td N M", M 1-based edges;
     * both directions of every edge, then the canonical CSR on the GPU */
    const char *what = ".gr parse";
    text t = slurp(f, what);
    while (t.p < t.n && t.d[t.p] == 'c')
        skip_line(&t);
    *N = next_id(&t, what);
    const uint32_t M = next_id(&t, what);
    uint32_t *I = xmalloc(sizeof(uint32_t) * (size_t)M, what);
    uint32_t *J = xmalloc(sizeof(uint32_t) * (size_t)M, what);
    for (uint32_t j = 0; j < M; j++) {
        skip_line(&t);
        while (t.p < t.n && t.d[t.p] == 'c')
            skip_line(&t);
        const uint32_t x = next_id(&t, what), y = next_id(&t, what);
        if (x == 0 || y == 0 || x > *N || y > *N)
            die(what, GCX_ERR_FORMAT);
        I[j] = x - 1;
        J[j] = y - 1;
    }
    free(t.d);
    gcx_graph *g = NULL;
    int rc = gcx_graph_from_pairs(*N, M, I, J, NULL, &g);
    free(I);
    free(J);
    if (rc != GCX_OK)
        die("CSR build", rc);
    *V = xmalloc(sizeof(uint32_t) * ((size_t)*N + 1), what);
    *E = xmalloc(sizeof(uint32_t) * (size_t)gcx_graph_num_entries(g), what);
    if ((rc = gcx_graph_download(g, *V, *E)) != GCX_OK)
        die("CSR download", rc);
    gcx_graph_free(g);
}

void gr_store(FILE *f, uint32_t N, const uint32_t *V, const uint32_t *E)
{
    /* fast_io.c:356-368 */
    fprintf(f, "p td %d %d\n", N, V[N] / 2);
    for (uint32_t u = 0; u < N; u++)
        for (uint32_t i = V[u]; i < V[u + 1]; i++)
            if (E[i] > u)
                fprintf(f, "%d %d\n", u + 1, E[i] + 1);
}

void metis_parse(FILE *f, uint32_t *N, uint32_t **V, uint32_t **E, uint32_t **W)
{
    /* fast_io.c:370-419: '%' comment lines, "N M fmt", then one line per
     * vertex: [vertex weight] then 1-based neighbours [each with an edge
     * weight]; fmt 10/11 carry vertex weights, 1/11 edge weights (dropped).
     * At most 2M entries are read, as in the reference. */
    const char *what = "METIS parse";
    text t = slurp(f, what);
    while (t.p < t.n && t.d[t.p] == '%')
        skip_line(&t);
    *N = next_id(&t, what);
    const uint32_t M = next_id(&t, what);
    const uint32_t fmt = next_id(&t, what);
    const int vw = fmt == 10 || fmt == 11, ew = fmt == 1 || fmt == 11;
    const uint64_t cap = 2ull * M;
    *V = xmalloc(sizeof(uint32_t) * ((size_t)*N + 1), what);
    *E = xmalloc(sizeof(uint32_t) * (size_t)cap, what);
    *W = vw ? xmalloc(sizeof(uint32_t) * (size_t)*N, what) : NULL;
    (*V)[0] = 0;
    for (uint32_t u = 0; u < *N; u++) {
        skip_line(&t);
        while (t.p < t.n && t.d[t.p] == '%')
            skip_line(&t);
        if (vw)
            (*W)[u] = next_id(&t, what);
        uint64_t s = (*V)[u];
        while (t.p < t.n && t.d[t.p] != '\n' && s < cap) {
            const uint32_t v = next_id(&t, what);
            if (v == 0 || v > *N)
                die(what, GCX_ERR_FORMAT);
            (*E)[s++] = v - 1;
            if (ew)
                (void)next_id(&t, what);
            while (t.p < t.n && t.d[t.p] == ' ' && s < cap)
                t.p++;
        }
        (*V)[u + 1] = (uint32_t)s;
    }
    free(t.d);
}

void metis_store(FILE *f, uint32_t N, const uint32_t *V, const uint32_t *E, const uint32_t *W)
{
    /* fast_io.c:421-437 */
    fprintf(f, "%d %d %d\n", N, V[N] / 2, W == NULL ? 0 : 10);
    for (uint32_t u = 0; u < N; u++) {
        if (W != NULL)
            fprintf(f, "%d ", W[u]);
        for (uint32_t i = V[u]; i < V[u + 1]; i++)
            fprintf(f, "%d ", E[i] + 1);
        fprintf(f, "\n");
    }
}
