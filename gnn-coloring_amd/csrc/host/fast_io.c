/*
 * fast_io.c -- Matrix Market ingest and CSR utilities (fast_io.h).
 *
 * Text parsing stays on the host (mmap + OpenMP over line-aligned byte
 * ranges); the pair list then goes to the GPU, where one radix sort builds
 * the canonical CSR (csr.hip). Sorting rows and removing self loops /
 * repeats of a caller's CSR also run on the GPU.
 */
#include <hip/hip_runtime_api.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>

#include "fast_io.h"
#include "internal.h"

static void die(const char *what, int rc)
{
    fprintf(stderr, "gcx: %s failed: %s\n", what, gcx_status_string(rc));
    exit(1);
}

#define CHECK(what, expr)                                                       \
    do {                                                                        \
        int rc_ = (expr);                                                       \
        if (rc_ != GCX_OK)                                                      \
            die(what, rc_);                                                     \
    } while (0)

#define HIPCHECK(what, expr)                                                    \
    do {                                                                        \
        if ((expr) != hipSuccess)                                               \
            die(what, GCX_ERR_HIP);                                             \
    } while (0)

static int is_digit(char c) { return c >= '0' && c <= '9'; }

/* Next unsigned decimal at or after *pos (skipping any non-digits), as in
 * fast_io.c:7-15. Returns 0 when the buffer ends first. */
static int next_uint(const char *d, size_t n, size_t *pos, uint64_t *out)
{
    size_t p = *pos;
    while (p < n && !is_digit(d[p]))
        p++;
    if (p >= n)
        return 0;
    uint64_t v = 0;
    while (p < n && is_digit(d[p]))
        v = v * 10 + (uint64_t)(d[p++] - '0');
    *pos = p;
    *out = v;
    return 1;
}

static size_t line_end(const char *d, size_t n, size_t p)
{
    while (p < n && d[p] != '\n')
        p++;
    return p;
}

/* An entry line has a digit before its end and does not start with '%'. */
static int is_entry(const char *d, size_t n, size_t p)
{
    if (p >= n || d[p] == '%')
        return 0;
    for (; p < n && d[p] != '\n'; p++)
        if (is_digit(d[p]))
            return 1;
    return 0;
}

/* Parse header and L pairs (1-based) into I, J (0-based). Returns 0 on
 * success. fast_io.c:68-141 semantics: '%' lines before the header are
 * skipped, N = max(rows, cols), values after the two indices are ignored. */
static int parse_pairs(FILE *f, uint32_t *N, uint64_t *Lout, uint32_t **Iout, uint32_t **Jout)
{
    struct stat st;
    if (fstat(fileno(f), &st) != 0 || st.st_size <= 0)
        return GCX_ERR_FORMAT;
    const size_t n = (size_t)st.st_size;
    char *d = mmap(NULL, n, PROT_READ, MAP_PRIVATE, fileno(f), 0);
    if (d == MAP_FAILED)
        return GCX_ERR_IO;
    size_t p = 0;
    while (p < n && d[p] == '%')
        p = line_end(d, n, p) + 1;
    uint64_t rows = 0, cols = 0, L = 0;
    if (!next_uint(d, n, &p, &rows) || !next_uint(d, n, &p, &cols) || !next_uint(d, n, &p, &L) ||
        rows > UINT32_MAX || cols > UINT32_MAX) {
        munmap(d, n);
        return GCX_ERR_FORMAT;
    }
    *N = (uint32_t)(rows > cols ? rows : cols);
    p = line_end(d, n, p) + 1;
    const size_t body = p < n ? p : n;
    uint32_t *I = malloc(sizeof(uint32_t) * (L ? L : 1));
    uint32_t *J = malloc(sizeof(uint32_t) * (L ? L : 1));
    if (!I || !J) {
        free(I);
        free(J);
        munmap(d, n);
        return GCX_ERR_NOMEM;
    }
    const int T = omp_get_max_threads();
    size_t *start = malloc(sizeof(size_t) * (T + 1));
    uint64_t *cnt = calloc(T + 1, sizeof(uint64_t));
    int bad = 0;
    /* Line-aligned chunk starts. */
    for (int t = 0; t <= T; t++) {
        size_t s = body + (n - body) / T * t;
        if (t == T)
            s = n;
        else if (t > 0) {
            s = line_end(d, n, s - 1);
            s = s < n ? s + 1 : n;
        }
        start[t] = s;
    }
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        uint64_t c = 0;
        for (size_t q = start[t]; q < start[t + 1]; q = line_end(d, n, q) + 1)
            c += is_entry(d, n, q);
        cnt[t + 1] = c;
#pragma omp barrier
#pragma omp single
        for (int k = 0; k < T; k++)
            cnt[k + 1] += cnt[k];
        uint64_t w = cnt[t];
        for (size_t q = start[t]; q < start[t + 1] && w < L; q = line_end(d, n, q) + 1) {
            if (!is_entry(d, n, q))
                continue;
            size_t pos = q;
            uint64_t a = 0, b = 0;
            if (!next_uint(d, n, &pos, &a) || !next_uint(d, n, &pos, &b) || a == 0 || b == 0 ||
                a > *N || b > *N) {
#pragma omp atomic write
                bad = 1;
                break;
            }
            I[w] = (uint32_t)(a - 1);
            J[w] = (uint32_t)(b - 1);
            w++;
        }
    }
    const uint64_t found = cnt[T];
    free(start);
    free(cnt);
    munmap(d, n);
    if (bad || found < L) {
        free(I);
        free(J);
        return GCX_ERR_FORMAT;
    }
    *Lout = L;
    *Iout = I;
    *Jout = J;
    return GCX_OK;
}

void mtx_parse_par(FILE *f, uint32_t *N, uint32_t **V, uint32_t **E)
{
    uint32_t *I = NULL, *J = NULL;
    uint64_t L = 0;
    CHECK("matrix market parse", parse_pairs(f, N, &L, &I, &J));
    gcx_graph *g = NULL;
    CHECK("CSR build", gcx_graph_from_pairs(*N, L, I, J, NULL, &g));
    free(I);
    free(J);
    *V = malloc(sizeof(uint32_t) * (*N + 1ull));
    *E = malloc(sizeof(uint32_t) * (g->M ? g->M : 1));
    CHECK("CSR download", gcx_graph_download(g, *V, *E));
    gcx_graph_free(g);
}

void mtx_parse(FILE *f, uint32_t *N, uint32_t **V, uint32_t **E) { mtx_parse_par(f, N, V, E); }

void mtx_store(FILE *f, uint32_t N, const uint32_t *V, const uint32_t *E)
{
    /* fast_io.c:224-237, banner text kept byte for byte. */
    fprintf(f, "%%%%MatrixMarket matrix coordinate patter symmetric\n");
    fprintf(f, "%d %d %d\n", N, N, V[N] / 2);
    for (uint32_t u = 0; u < N; u++)
        for (uint32_t i = V[u]; i < V[u + 1]; i++)
            if (E[i] > u)
                fprintf(f, "%d %d\n", u + 1, E[i] + 1);
}

/* Upload a caller CSR into fresh device buffers (bounds-checked). */
static gcx_graph *upload(uint32_t N, const uint32_t *V, const uint32_t *E)
{
    gcx_graph *g = NULL;
    CHECK("graph upload", gcx_graph_from_csr(N, V, E, NULL, &g));
    return g;
}

void graph_sort_edges_par(uint32_t N, const uint32_t *V, uint32_t *E)
{
    if (N == 0 || V[N] == 0)
        return;
    gcx_graph *g = upload(N, V, E);
    CHECK("row sort", gcxk_sort_rows(N, g->V, g->E, g->M, gcx__default_stream()));
    CHECK("CSR download", gcx_graph_download(g, NULL, E));
    gcx_graph_free(g);
}

void graph_sort_edges(uint32_t N, const uint32_t *V, uint32_t *E) { graph_sort_edges_par(N, V, E); }

void graph_make_simple(uint32_t N, uint32_t *V, uint32_t **E)
{
    if (N == 0)
        return;
    gcx_graph *g = upload(N, V, *E);
    uint32_t *out = NULL;
    uint64_t M = 0;
    HIPCHECK("allocation", hipMalloc((void **)&out, 4ull * (g->M ? g->M : 1)));
    CHECK("simplify", gcxk_make_simple(N, g->V, g->E, g->M, out, &M, gcx__default_stream()));
    HIPCHECK("CSR download", hipMemcpy(V, g->V, 4ull * (N + 1ull), hipMemcpyDeviceToHost));
    if (M)
        HIPCHECK("CSR download", hipMemcpy(*E, out, 4ull * M, hipMemcpyDeviceToHost));
    (void)hipFree(out);
    gcx_graph_free(g);
    uint32_t *shrunk = realloc(*E, sizeof(uint32_t) * (M ? M : 1));
    if (shrunk)
        *E = shrunk;
}

uint32_t graph_validate_par(uint32_t N, const uint32_t *V, const uint32_t *E)
{
    if (N == 0)
        return V[0] == 0;
    gcx_graph *g = NULL;
    const int rc = gcx_graph_from_csr(N, V, E, NULL, &g);
    if (rc == GCX_ERR_ARG)
        return 0; /* indices out of range */
    CHECK("graph upload", rc);
    int valid = 0;
    CHECK("CSR check", gcx_graph_check(g, NULL, &valid));
    gcx_graph_free(g);
    return valid ? 1u : 0u;
}

uint32_t graph_validate(uint32_t N, const uint32_t *V, const uint32_t *E)
{
    return graph_validate_par(N, V, E);
}
