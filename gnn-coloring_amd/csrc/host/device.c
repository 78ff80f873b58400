/*
 * device.c -- C host layer: device objects (graphs, models, workspaces),
 * stage sequencing and the Jones-Plassmann / SL round drivers. All compute
 * runs in the HIP kernels behind gcxk.h; this file only allocates, copies,
 * launches and polls.
 */
#include <hip/hip_runtime_api.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gen.h"
#include "internal.h"

/* Rounds enqueued between two looks at the device's "done" flag. */
#define ROUND_BATCH 32u

static int g_init = 0;
static int g_init_rc = GCX_OK;
static hipStream_t g_stream = NULL;
static uint32_t g_round_grid = 0;

const char *gcx_status_string(int s)
{
    switch (s) {
    case GCX_OK: return "ok";
    case GCX_ERR_ARG: return "invalid argument";
    case GCX_ERR_NOMEM: return "out of memory";
    case GCX_ERR_HIP: return "HIP runtime error";
    case GCX_ERR_CAPACITY: return "input exceeds capacity";
    case GCX_ERR_IO: return "cannot open file";
    case GCX_ERR_FORMAT: return "malformed file";
    case GCX_ERR_NODEV: return "no gfx950 device";
    case GCX_ERR_INVALID: return "validation failed";
    case GCX_ERR_BUDGET: return "HBM budget exceeded";
    default: return "unknown status";
    }
}

const char *gcx_version(void) { return "gcx 0.1 (gfx950)"; }

int gcx__ensure_init(void)
{
    if (g_init)
        return g_init_rc;
    g_init = 1;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
        return g_init_rc = GCX_ERR_NODEV;
    if (hipSetDevice(0) != hipSuccess)
        return g_init_rc = GCX_ERR_NODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess)
        return g_init_rc = GCX_ERR_NODEV;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        fprintf(stderr, "gcx: device 0 is %s, this build targets gfx950\n", prop.gcnArchName);
        return g_init_rc = GCX_ERR_NODEV;
    }
    if (hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) != hipSuccess)
        return g_init_rc = GCX_ERR_HIP;
    /* Graphs, workspaces and temporaries come from the device's stream-ordered
     * pool; the pool keeps freed memory for reuse instead of returning it at
     * every synchronisation. */
    hipMemPool_t pool;
    uint64_t keep = UINT64_MAX;
    if (hipDeviceGetDefaultMemPool(&pool, 0) != hipSuccess ||
        hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess)
        return g_init_rc = GCX_ERR_HIP;
    g_round_grid = gcxk_round_grid();
    if (gcxk_noalloc_init() != GCX_OK)
        return g_init_rc = GCX_ERR_HIP;
    return g_init_rc = GCX_OK;
}

int gcx_init(void) { return gcx__ensure_init(); }

void *gcx__default_stream(void) { return (void *)g_stream; }

void *gcx__pick(void *stream) { return stream ? stream : (void *)g_stream; }

int gcx__dalloc(void **p, size_t bytes, void *stream)
{
    *p = NULL;
    return hipMallocAsync(p, bytes ? bytes : 16, (hipStream_t)gcx__pick(stream)) == hipSuccess
               ? GCX_OK
               : GCX_ERR_NOMEM;
}

void gcx__dfree(void *p, void *stream)
{
    if (p)
        (void)hipFreeAsync(p, (hipStream_t)gcx__pick(stream));
}

/* ------------------------------------------------------------ streams/events */

int gcx_stream_create(void **stream)
{
    TRY(gcx__ensure_init());
    hipStream_t s;
    HIPC(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return GCX_OK;
}

int gcx_stream_destroy(void *stream)
{
    if (stream)
        HIPC(hipStreamDestroy((hipStream_t)stream));
    return GCX_OK;
}

int gcx_stream_sync(void *stream)
{
    TRY(gcx__ensure_init());
    HIPC(hipStreamSynchronize((hipStream_t)gcx__pick(stream)));
    return GCX_OK;
}

int gcx_device_sync(void)
{
    TRY(gcx__ensure_init());
    HIPC(hipDeviceSynchronize());
    return GCX_OK;
}

int gcx_event_create(void **event)
{
    TRY(gcx__ensure_init());
    hipEvent_t e;
    HIPC(hipEventCreate(&e));
    *event = (void *)e;
    return GCX_OK;
}

int gcx_event_destroy(void *event)
{
    if (event)
        HIPC(hipEventDestroy((hipEvent_t)event));
    return GCX_OK;
}

int gcx_event_record(void *event, void *stream)
{
    HIPC(hipEventRecord((hipEvent_t)event, (hipStream_t)gcx__pick(stream)));
    return GCX_OK;
}

int gcx_event_elapsed_ms(void *start, void *stop, float *ms)
{
    HIPC(hipEventSynchronize((hipEvent_t)stop));
    HIPC(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return GCX_OK;
}

int gcx_mem_info(size_t *free_bytes, size_t *total_bytes)
{
    TRY(gcx__ensure_init());
    HIPC(hipMemGetInfo(free_bytes, total_bytes));
    return GCX_OK;
}

/* ------------------------------------------------------------------ profile */

typedef struct {
    int stage;
    hipEvent_t a, b;
} prof_rec;

static int g_prof_on = 0;
static prof_rec *g_prof = NULL;
static size_t g_prof_n = 0, g_prof_cap = 0;
static hipEvent_t *g_pool = NULL;
static size_t g_pool_n = 0, g_pool_cap = 0;

static hipEvent_t pool_get(void)
{
    if (g_pool_n)
        return g_pool[--g_pool_n];
    hipEvent_t e = NULL;
    if (hipEventCreate(&e) != hipSuccess)
        return NULL;
    return e;
}

static void pool_put(hipEvent_t e)
{
    if (g_pool_n == g_pool_cap) {
        size_t cap = g_pool_cap ? 2 * g_pool_cap : 64;
        hipEvent_t *p = realloc(g_pool, cap * sizeof(*p));
        if (!p) {
            (void)hipEventDestroy(e);
            return;
        }
        g_pool = p;
        g_pool_cap = cap;
    }
    g_pool[g_pool_n++] = e;
}

int gcx_calibration_copy(uint64_t bytes, float *ms)
{
    TRY(gcx__ensure_init());
    bytes &= ~(uint64_t)15;
    void *a = NULL, *b = NULL;
    hipEvent_t e0 = NULL, e1 = NULL;
    int rc = GCX_ERR_HIP;
    if (hipMalloc(&a, bytes ? bytes : 16) != hipSuccess || hipMalloc(&b, bytes ? bytes : 16) != hipSuccess ||
        hipMemsetAsync(a, 1, bytes, g_stream) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipEventRecord(e0, g_stream) != hipSuccess)
        goto done;
    if ((rc = gcxk_copy_bytes(bytes, a, b, g_stream)) != GCX_OK)
        goto done;
    rc = GCX_ERR_HIP;
    if (hipEventRecord(e1, g_stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(ms, e0, e1) != hipSuccess)
        goto done;
    rc = GCX_OK;
done:
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    if (a)
        (void)hipFree(a);
    if (b)
        (void)hipFree(b);
    return rc;
}

int gcx_profile_enable(int on)
{
    g_prof_on = on != 0;
    return GCX_OK;
}

int gcx__prof_begin(int stage, void *stream, void **token)
{
    *token = NULL;
    if (!g_prof_on)
        return GCX_OK;
    if (g_prof_n == g_prof_cap) {
        size_t cap = g_prof_cap ? 2 * g_prof_cap : 256;
        prof_rec *p = realloc(g_prof, cap * sizeof(*p));
        if (!p)
            return GCX_ERR_NOMEM;
        g_prof = p;
        g_prof_cap = cap;
    }
    prof_rec *r = &g_prof[g_prof_n];
    r->stage = stage;
    r->a = pool_get();
    r->b = pool_get();
    if (!r->a || !r->b)
        return GCX_ERR_HIP;
    HIPC(hipEventRecord(r->a, (hipStream_t)stream));
    g_prof_n++;
    *token = (void *)(uintptr_t)g_prof_n; /* index + 1 */
    return GCX_OK;
}

int gcx__prof_end(void *token, void *stream)
{
    if (!token)
        return GCX_OK;
    prof_rec *r = &g_prof[(uintptr_t)token - 1];
    HIPC(hipEventRecord(r->b, (hipStream_t)stream));
    return GCX_OK;
}

int gcx_profile_reset(void)
{
    for (size_t i = 0; i < g_prof_n; i++) {
        pool_put(g_prof[i].a);
        pool_put(g_prof[i].b);
    }
    g_prof_n = 0;
    return GCX_OK;
}

int gcx_profile_read(double ms[GCX_ST_COUNT], uint64_t launches[GCX_ST_COUNT])
{
    for (int s = 0; s < GCX_ST_COUNT; s++) {
        ms[s] = 0.0;
        launches[s] = 0;
    }
    if (g_prof_n)
        HIPC(hipDeviceSynchronize());
    for (size_t i = 0; i < g_prof_n; i++) {
        float t = 0.0f;
        HIPC(hipEventElapsedTime(&t, g_prof[i].a, g_prof[i].b));
        ms[g_prof[i].stage] += t;
        launches[g_prof[i].stage] += 1;
    }
    return gcx_profile_reset();
}

/* -------------------------------------------------------------------- graph */

static int block_get(void **p, size_t bytes, size_t *got, void *stream);
static void block_put(void *p, size_t bytes, void *stream);

uint64_t gcx_graph_footprint(uint32_t N, uint64_t M)
{
    /* V and the hub chunk table (N + 1 each), E, Er and Ed (each with
     * GCXK_ER_PAD entries of slack), hubs, the degree order and its inverse,
     * each part 256-byte aligned in one block */
    return 8ull * ((uint64_t)N + 1) + 12ull * (M + GCXK_ER_PAD) + 12ull * N + 8ull * 256;
}

/* Degree statistics, hub list (ascending ids) with its 1024-edge chunk table
 * for the JP orientation pass, and the degree order. The hub degree is the
 * caller's, else the tuner's for the graph's degree signature, else the
 * default. */
static int graph_stats(gcx_graph *g, hipStream_t s)
{
    if (!g->hchunk)
        TRY(gcx__dalloc((void **)&g->hchunk, 4ull * (g->N + 1ull), s));
    if (!g->hub_deg)
        g->hub_deg = GCXK_HUB_DEG;
    TRY(gcxk_graph_stats(g->N, g->V, g->hub_deg, g->hubs, g->hchunk, &g->nhub, &g->maxdeg,
                         &g->nhchunk, s));
    const uint32_t tuned = g->hub_fixed ? 0u : gcx__tuned_hub_degree(g->N, g->M, g->maxdeg);
    if (tuned && tuned != g->hub_deg) {
        g->hub_deg = tuned;
        TRY(gcxk_graph_stats(g->N, g->V, g->hub_deg, g->hubs, g->hchunk, &g->nhub, &g->maxdeg,
                             &g->nhchunk, s));
    }
    if (!g->perm)
        TRY(gcx__dalloc((void **)&g->perm, 4ull * (g->N ? g->N : 1), s));
    TRY(gcxk_degree_order(g->N, g->V, g->maxdeg, g->perm, s));
    if (!g->rank)
        TRY(gcx__dalloc((void **)&g->rank, 4ull * (g->N ? g->N : 1), s));
    if (!g->Er)
        TRY(gcx__dalloc((void **)&g->Er, 4ull * (g->M + GCXK_ER_PAD), s));
    if (!g->Ed)
        TRY(gcx__dalloc((void **)&g->Ed, 4ull * (g->M + GCXK_ER_PAD), s));
    TRY(gcxk_rank_columns(g->N, g->M, g->perm, g->V, g->E, g->rank, g->Er, g->Ed, s));
    /* creation returns a graph usable on any stream */
    HIPC(hipStreamSynchronize(s));
    return GCX_OK;
}

/* One allocation carved into the per-graph arrays (and V and E when with_ve
 * is set): stream-ordered allocations and frees each cost host time, which
 * adds up over the many small graphs of a serving trace. */
static int graph_block(gcx_graph *g, int with_ve, void *stream)
{
    const uint64_t N = g->N, M = g->M;
    struct {
        uint32_t **p;
        size_t bytes;
    } part[] = {
        {&g->hubs, 4 * (N ? N : 1)},     {&g->hchunk, 4 * (N + 1)},
        {&g->perm, 4 * (N ? N : 1)},     {&g->rank, 4 * (N ? N : 1)},
        {&g->Er, 4 * (M + GCXK_ER_PAD)}, {&g->Ed, 4 * (M + GCXK_ER_PAD)},
        {&g->V, 4 * (N + 1)},            {&g->E, 4 * (M + GCXK_ER_PAD)},
    };
    const size_t nparts = with_ve ? 8 : 6;
    size_t total = 0;
    for (size_t i = 0; i < nparts; i++)
        total += (part[i].bytes + 255) & ~(size_t)255;
    TRY(block_get(&g->blk, total, &g->blk_bytes, stream));
    g->blk_ve = with_ve;
    char *at = (char *)g->blk;
    for (size_t i = 0; i < nparts; i++) {
        *part[i].p = (uint32_t *)at;
        at += (part[i].bytes + 255) & ~(size_t)255;
    }
    return GCX_OK;
}

static int graph_finish(gcx_graph *g, hipStream_t s)
{
    TRY(graph_block(g, 0, s));
    return graph_stats(g, s);
}

int gcx_graph_free_on(gcx_graph *g, void *stream)
{
    if (!g)
        return GCX_OK;
    void *bufs[] = {g->V,    g->E,  g->hubs, g->hchunk, g->perm,
                    g->rank, g->Er, g->Ed,   g->gspan,  g->goff};
    /* the first 8 may live in the graph's block: V and E when blk_ve is
     * set, the next 6 whenever there is a block */
    for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); i++)
        if (!g->blk || (i >= 2 ? i >= 8 : !g->blk_ve))
            gcx__dfree(bufs[i], stream);
    block_put(g->blk, g->blk_bytes, stream);
    free(g->hgoff);
    free(g);
    return GCX_OK;
}

int gcx_graph_free(gcx_graph *g)
{
    /* any stream may still read the graph: finish all work first */
    if (g && hipDeviceSynchronize() != hipSuccess) {
        (void)gcx_graph_free_on(g, NULL);
        return GCX_ERR_HIP;
    }
    return gcx_graph_free_on(g, NULL);
}

int gcx__graph_alloc(uint32_t N, uint64_t M, void *stream, gcx_graph **out)
{
    TRY(gcx__ensure_init());
    *out = NULL;
    gcx_graph *g = calloc(1, sizeof(*g));
    if (!g)
        return GCX_ERR_NOMEM;
    g->N = N;
    g->M = M;
    if (graph_block(g, 1, stream) != GCX_OK) {
        gcx_graph_free_on(g, stream);
        return GCX_ERR_NOMEM;
    }
    *out = g;
    return GCX_OK;
}

/* Copy a host CSR into buffers sized by gcx__graph_alloc (N and V[N] may be
 * smaller than the allocation) and refresh the degree statistics. */
int gcx__graph_fill(gcx_graph *g, uint32_t N, const uint32_t *V, const uint32_t *E, void *stream)
{
    if (!g || !V || N == 0 || (V[N] && !E))
        return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    uint32_t *tmp = NULL;
    HIPC(hipMemcpyAsync(g->V, V, 4ull * (N + 1ull), hipMemcpyHostToDevice, s));
    if (V[N])
        HIPC(hipMemcpyAsync(g->E, E, 4ull * V[N], hipMemcpyHostToDevice, s));
    g->N = N;
    g->M = V[N];
    TRY(gcx__dalloc((void **)&tmp, 16, s));
    int ok = 0;
    int rc = gcxk_csr_bounds(g->N, g->M, g->V, g->E, tmp, &ok, s);
    if (rc == GCX_OK && !ok)
        rc = GCX_ERR_ARG; /* indices out of range: refuse before any kernel reads them */
    gcx__dfree(tmp, s);
    if (rc == GCX_OK)
        rc = graph_stats(g, s);
    return rc;
}

int gcx_graph_from_csr(uint32_t N, const uint32_t *V, const uint32_t *E, void *stream,
                       gcx_graph **out)
{
    TRY(gcx__ensure_init());
    *out = NULL;
    if (N == 0 || !V || (V[N] && !E))
        return GCX_ERR_ARG;
    gcx_graph *g = NULL;
    TRY(gcx__graph_alloc(N, V[N], stream, &g));
    int rc = gcx__graph_fill(g, N, V, E, stream);
    if (rc != GCX_OK) {
        gcx_graph_free_on(g, stream);
        return rc;
    }
    *out = g;
    return GCX_OK;
}

static int graph_from_device_pairs(uint32_t N, uint64_t L, uint32_t *dI, uint32_t *dJ,
                                   hipStream_t s, gcx_graph **out)
{
    gcx_graph *g = calloc(1, sizeof(*g));
    if (!g)
        return GCX_ERR_NOMEM;
    g->N = N;
    int rc = GCX_ERR_NOMEM;
    if (gcx__dalloc((void **)&g->V, 4ull * (N + 1ull), s) != GCX_OK)
        goto fail;
    if ((rc = gcxk_csr_from_pairs(N, L, dI, dJ, g->V, &g->E, &g->M, s)) != GCX_OK)
        goto fail;
    if ((rc = graph_finish(g, s)) != GCX_OK)
        goto fail;
    *out = g;
    return GCX_OK;
fail:
    (void)hipStreamSynchronize(s);
    gcx_graph_free_on(g, s);
    return rc;
}

int gcx_graph_from_pairs(uint32_t N, uint64_t L, const uint32_t *I, const uint32_t *J,
                         void *stream, gcx_graph **out)
{
    TRY(gcx__ensure_init());
    *out = NULL;
    if (N == 0 || (L && (!I || !J)))
        return GCX_ERR_ARG;
    for (uint64_t k = 0; k < L; k++)
        if (I[k] >= N || J[k] >= N)
            return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    uint32_t *dI = NULL, *dJ = NULL;
    int rc = GCX_ERR_HIP;
    if (gcx__dalloc((void **)&dI, 4ull * (L ? L : 1), s) != GCX_OK ||
        gcx__dalloc((void **)&dJ, 4ull * (L ? L : 1), s) != GCX_OK)
        goto done;
    if (L && (hipMemcpyAsync(dI, I, 4ull * L, hipMemcpyHostToDevice, s) != hipSuccess ||
              hipMemcpyAsync(dJ, J, 4ull * L, hipMemcpyHostToDevice, s) != hipSuccess))
        goto done;
    rc = graph_from_device_pairs(N, L, dI, dJ, s, out);
done:
    gcx__dfree(dI, s);
    gcx__dfree(dJ, s);
    return rc;
}

int gcx_graph_rmat(int scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                   void *stream, gcx_graph **out)
{
    TRY(gcx__ensure_init());
    *out = NULL;
    if (scale < 1 || scale > 30 || edge_factor == 0)
        return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    const uint64_t L = (uint64_t)edge_factor << scale;
    uint32_t thr[3];
    gcx_rmat_thresholds(a, b, c, thr);
    uint32_t *dI = NULL, *dJ = NULL;
    int rc = GCX_ERR_HIP;
    if (gcx__dalloc((void **)&dI, 4ull * L, s) != GCX_OK ||
        gcx__dalloc((void **)&dJ, 4ull * L, s) != GCX_OK)
        goto done;
    if ((rc = gcxk_gen_rmat(scale, L, thr[0], thr[1], thr[2], seed, dI, dJ, s)) != GCX_OK)
        goto done;
    rc = graph_from_device_pairs(1u << scale, L, dI, dJ, s, out);
done:
    gcx__dfree(dI, s);
    gcx__dfree(dJ, s);
    return rc;
}

int gcx_graph_er(uint32_t n, uint64_t m, uint64_t seed, void *stream, gcx_graph **out)
{
    TRY(gcx__ensure_init());
    *out = NULL;
    if (n == 0)
        return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    uint32_t *dI = NULL, *dJ = NULL;
    int rc = GCX_ERR_HIP;
    if (gcx__dalloc((void **)&dI, 4ull * (m ? m : 1), s) != GCX_OK ||
        gcx__dalloc((void **)&dJ, 4ull * (m ? m : 1), s) != GCX_OK)
        goto done;
    if ((rc = gcxk_gen_er(n, m, seed, dI, dJ, s)) != GCX_OK)
        goto done;
    rc = graph_from_device_pairs(n, m, dI, dJ, s, out);
done:
    gcx__dfree(dI, s);
    gcx__dfree(dJ, s);
    return rc;
}

int gcx_graph_rgg(uint32_t n, double radius, uint64_t seed, void *stream, gcx_graph **out)
{
    TRY(gcx__ensure_init());
    *out = NULL;
    if (n == 0 || !(radius > 0.0))
        return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    uint32_t *dI = NULL, *dJ = NULL;
    uint64_t L = 0;
    int rc = gcxk_gen_rgg(n, gcx_rgg_cells(radius), gcx_rgg_r2(radius), seed, &dI, &dJ, &L, s);
    if (rc == GCX_OK)
        rc = graph_from_device_pairs(n, L, dI, dJ, s, out);
    gcx__dfree(dI, s);
    gcx__dfree(dJ, s);
    return rc;
}

static int ws_fits(const gcx_workspace *ws, const gcx_graph *g);

int gcx_graph_pack(const gcx_graph *const *graphs, uint32_t count, void *stream, gcx_graph **out)
{
    TRY(gcx__ensure_init());
    *out = NULL;
    if (!graphs || count == 0)
        return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    uint64_t N = 0, M = 0;
    uint32_t *off = malloc(sizeof(uint32_t) * ((size_t)count + 1));
    if (!off)
        return GCX_ERR_NOMEM;
    for (uint32_t i = 0; i < count; i++) {
        if (!graphs[i]) {
            free(off);
            return GCX_ERR_ARG;
        }
        off[i] = (uint32_t)N;
        N += graphs[i]->N;
        M += graphs[i]->M;
    }
    off[count] = (uint32_t)N;
    if (N == 0 || N >= UINT32_MAX || M >= UINT32_MAX) {
        free(off);
        return GCX_ERR_CAPACITY;
    }
    gcx_graph *g = NULL;
    int rc = gcx__graph_alloc((uint32_t)N, M, s, &g);
    if (rc == GCX_OK)
        rc = gcx__dalloc((void **)&g->gspan, 8ull * N, s);
    if (rc == GCX_OK)
        rc = gcx__dalloc((void **)&g->goff, 4ull * ((size_t)count + 1), s);
    if (rc == GCX_OK && hipMemcpyAsync(g->goff, off, 4ull * ((size_t)count + 1),
                                       hipMemcpyHostToDevice, s) != hipSuccess)
        rc = GCX_ERR_HIP;
    if (rc == GCX_OK)
        g->hgoff = off; /* the graph keeps the host copy */
    uint64_t eoff = 0;
    for (uint32_t i = 0; rc == GCX_OK && i < count; i++) {
        const gcx_graph *h = graphs[i];
        rc = gcxk_pack_graph(h->N, h->M, h->V, h->E, off[i], (uint32_t)eoff, g->V, g->E, g->gspan,
                             s);
        eoff += h->M;
    }
    if (rc == GCX_OK)
        rc = gcxk_pack_close((uint32_t)N, (uint32_t)M, g->V, s);
    if (rc == GCX_OK) {
        g->count = count;
        rc = graph_stats(g, s); /* synchronises */
    } else {
        (void)hipStreamSynchronize(s);
    }
    if (!g || g->hgoff != off)
        free(off);
    if (rc != GCX_OK) {
        gcx_graph_free_on(g, s);
        return rc;
    }
    *out = g;
    return GCX_OK;
}

uint32_t gcx_graph_packed_count(const gcx_graph *g) { return g ? g->count : 0; }

int gcx_graph_prepare(gcx_graph *g, void *stream)
{
    TRY(gcx__ensure_init());
    if (!g)
        return GCX_ERR_ARG;
    return graph_stats(g, (hipStream_t)gcx__pick(stream));
}

int gcx_graph_set_hub_degree(gcx_graph *g, uint32_t hub_degree, void *stream)
{
    TRY(gcx__ensure_init());
    if (!g)
        return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    g->hub_fixed = hub_degree != 0;
    g->hub_deg = hub_degree ? hub_degree
                            : (gcx__tuned_hub_degree(g->N, g->M, g->maxdeg)
                                   ? gcx__tuned_hub_degree(g->N, g->M, g->maxdeg)
                                   : GCXK_HUB_DEG);
    /* the degree order does not depend on the cut: only the hub tables */
    TRY(gcxk_graph_stats(g->N, g->V, g->hub_deg, g->hubs, g->hchunk, &g->nhub, &g->maxdeg,
                         &g->nhchunk, s));
    return GCX_OK;
}

uint32_t gcx_graph_hub_degree(const gcx_graph *g) { return g ? g->hub_deg : 0; }

int gcx_packed_colors(gcx_workspace *ws, const gcx_graph *g, int *colors, uint64_t *digests,
                      void *stream)
{
    TRY(gcx__ensure_init());
    if (!ws_fits(ws, g) || !g->count || !g->goff || !colors || !digests)
        return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    const size_t n = g->count;
    /* vertex ranges of at most 64K vertices, each inside one graph */
    const uint32_t span = 1u << 16;
    size_t nr = 0;
    for (size_t k = 0; k < n; k++)
        nr += (g->hgoff[k + 1] - g->hgoff[k] + span - 1) / span;
    uint32_t *hr = malloc(sizeof(uint32_t) * 3 * (nr ? nr : 1));
    if (!hr)
        return GCX_ERR_NOMEM;
    size_t r = 0;
    for (size_t k = 0; k < n; k++)
        for (uint32_t b = g->hgoff[k]; b < g->hgoff[k + 1]; b += span, r++) {
            hr[r] = b;
            hr[nr + r] = g->hgoff[k + 1] - b < span ? g->hgoff[k + 1] : b + span;
            hr[2 * nr + r] = (uint32_t)k;
        }
    void *dbuf = NULL;
    int rc = gcx__dalloc(&dbuf, 12 * n + 12 * nr + 16, s);
    if (rc != GCX_OK) {
        free(hr);
        return rc;
    }
    uint64_t *dd = (uint64_t *)dbuf;
    int *dc = (int *)(dd + n);
    uint32_t *dr = (uint32_t *)(dc + n);
    rc = GCX_ERR_HIP;
    if (hipMemsetAsync(dbuf, 0, 12 * n, s) == hipSuccess &&
        (nr == 0 || hipMemcpyAsync(dr, hr, 12 * nr, hipMemcpyHostToDevice, s) == hipSuccess) &&
        gcxk_packed_stats((uint32_t)nr, dr, dr + nr, dr + 2 * nr, g->goff, ws->color, dc, dd, s) ==
            GCX_OK &&
        hipMemcpyAsync(colors, dc, 4 * n, hipMemcpyDeviceToHost, s) == hipSuccess &&
        hipMemcpyAsync(digests, dd, 8 * n, hipMemcpyDeviceToHost, s) == hipSuccess &&
        hipStreamSynchronize(s) == hipSuccess)
        rc = GCX_OK;
    gcx__dfree(dbuf, s);
    free(hr);
    return rc;
}

uint32_t gcx_graph_num_vertices(const gcx_graph *g) { return g ? g->N : 0; }
uint64_t gcx_graph_num_entries(const gcx_graph *g) { return g ? g->M : 0; }
uint32_t gcx_graph_max_degree(const gcx_graph *g) { return g ? g->maxdeg : 0; }
const uint32_t *gcx_graph_V_dev(const gcx_graph *g) { return g ? g->V : NULL; }
const uint32_t *gcx_graph_E_dev(const gcx_graph *g) { return g ? g->E : NULL; }

int gcx_graph_download(const gcx_graph *g, uint32_t *V, uint32_t *E)
{
    if (!g)
        return GCX_ERR_ARG;
    if (V)
        HIPC(hipMemcpy(V, g->V, 4ull * (g->N + 1ull), hipMemcpyDeviceToHost));
    if (E && g->M)
        HIPC(hipMemcpy(E, g->E, 4ull * g->M, hipMemcpyDeviceToHost));
    return GCX_OK;
}

int gcx_graph_check(const gcx_graph *g, void *stream, int *valid)
{
    TRY(gcx__ensure_init());
    if (!g || !valid)
        return GCX_ERR_ARG;
    uint32_t *tmp = NULL;
    TRY(gcx__dalloc((void **)&tmp, 16, stream));
    int rc = gcxk_graph_check(g->N, g->V, g->E, tmp, valid, gcx__pick(stream));
    gcx__dfree(tmp, stream);
    return rc;
}

/* -------------------------------------------------------------------- model */

#define LAYER_PARAMS 528

int gcx_model_from_params(int params, const float *param, int md, gcx_model **out)
{
    TRY(gcx__ensure_init());
    *out = NULL;
    if (params < 2 * LAYER_PARAMS || params % LAYER_PARAMS || !param || md <= 0)
        return GCX_ERR_FORMAT;
    gcx_model *m = calloc(1, sizeof(*m));
    if (!m)
        return GCX_ERR_NOMEM;
    m->params = params;
    m->md = md;
    m->layers = params / LAYER_PARAMS;
    m->hparam = malloc(sizeof(float) * (size_t)params);
    if (!m->hparam) {
        free(m);
        return GCX_ERR_NOMEM;
    }
    memcpy(m->hparam, param, sizeof(float) * (size_t)params);
    if (hipMalloc((void **)&m->dparam, sizeof(float) * (size_t)params) != hipSuccess ||
        hipMemcpy(m->dparam, param, sizeof(float) * (size_t)params, hipMemcpyHostToDevice) !=
            hipSuccess) {
        gcx_model_free(m);
        return GCX_ERR_HIP;
    }
    int rc = gcx__model_pack_bf16(m);
    if (rc != GCX_OK) {
        gcx_model_free(m);
        return rc;
    }
    *out = m;
    return GCX_OK;
}

int gcx_model_load(const char *path, gcx_model **out)
{
    /* gnn.c:20-44: "params md", then params decimal floats (strtof rounding). */
    *out = NULL;
    FILE *f = fopen(path, "r");
    if (!f)
        return GCX_ERR_IO;
    int params = 0, md = 0;
    if (fscanf(f, "%d %d", &params, &md) != 2 || params < 2 * LAYER_PARAMS ||
        params % LAYER_PARAMS || params > (1 << 24)) {
        fclose(f);
        return GCX_ERR_FORMAT;
    }
    float *w = malloc(sizeof(float) * (size_t)params);
    if (!w) {
        fclose(f);
        return GCX_ERR_NOMEM;
    }
    for (int i = 0; i < params; i++) {
        if (fscanf(f, "%f", w + i) != 1) {
            free(w);
            fclose(f);
            return GCX_ERR_FORMAT;
        }
    }
    fclose(f);
    int rc = gcx_model_from_params(params, w, md, out);
    free(w);
    return rc;
}

int gcx_model_free(gcx_model *m)
{
    if (!m)
        return GCX_OK;
    if (m->dparam)
        (void)hipFree(m->dparam);
    if (m->dWb)
        (void)hipFree(m->dWb);
    if (m->dWpack)
        (void)hipFree(m->dWpack);
    free(m->hparam);
    free(m);
    return GCX_OK;
}

int gcx_model_layers(const gcx_model *m) { return m ? m->layers : 0; }

/* ---------------------------------------------------------------- workspace */

/* Sort scratch of the ordered JP schedule that does not fit in the x rows. */
static size_t order_extra(uint32_t N)
{
    const size_t need = gcxk_jp_order_scratch(N, 64);
    const size_t in_x = 64ull * ((size_t)N + 1);
    return need > in_x ? need : 0;
}

uint64_t gcx_workspace_footprint(uint32_t N, uint64_t M)
{
    /* per vertex: key 8, count 4, color 4, npred 4, word offsets 4, word
     * descriptors 8, 4 lists x 4, hpos 8, rows x,y 2 x 64; per entry: E_split
     * 4; plus the state and overflow colour words, two item queues, any
     * sort scratch beyond the x rows, and the 256-byte alignment of each of
     * the 21 parts carved from one allocation */
    const uint64_t n = (uint64_t)N + 1;
    return 184ull * n + 8ull * gcxk_jp_state_words(N, M) + 4ull * M +
           4ull * gcxk_jp_ovf_words(N, M) +
           2ull * sizeof(gcxk_item) * gcxk_jp_item_capacity(N, M) + 2 * sizeof(gcxk_ctl) + 256 +
           order_extra(N) + 21ull * 256;
}

/* Pinned control-block mirrors are recycled: freeing pinned host memory
 * synchronises the whole device, which would stall other streams' jobs. A
 * block goes back with an event recorded on the freeing stream, and is
 * handed out again only after that event: the round loop keeps device ->
 * host snapshot copies into it in flight (drive_rounds), and one of them
 * could otherwise land in the next owner's block and pass it a stale flag. */
static pthread_mutex_t g_pin_lock = PTHREAD_MUTEX_INITIALIZER;
static struct {
    void *p;
    hipEvent_t ev;
} g_pin_free[64];
static int g_pin_n = 0;

static int ev_get(hipEvent_t *out);
static void ev_put(hipEvent_t e);

static void *pinned_ctl_get(void)
{
    void *p = NULL;
    hipEvent_t ev = NULL;
    pthread_mutex_lock(&g_pin_lock);
    if (g_pin_n > 0) {
        g_pin_n--;
        p = g_pin_free[g_pin_n].p;
        ev = g_pin_free[g_pin_n].ev;
    }
    pthread_mutex_unlock(&g_pin_lock);
    if (p && ev && hipEventSynchronize(ev) != hipSuccess) {
        /* the last copies into it may not have ended: never reuse it */
        ev_put(ev);
        return NULL;
    }
    ev_put(ev);
    if (!p && hipHostMalloc(&p, 4 * sizeof(gcxk_ctl), hipHostMallocDefault) != hipSuccess)
        p = NULL;
    return p;
}

static void pinned_ctl_put(void *p, void *stream)
{
    if (!p)
        return;
    hipEvent_t ev = NULL;
    if (!ev_get(&ev) || hipEventRecord(ev, (hipStream_t)gcx__pick(stream)) != hipSuccess) {
        /* cannot order the reuse after the copies: finish them, then keep it */
        ev_put(ev);
        ev = NULL;
        if (hipStreamSynchronize((hipStream_t)gcx__pick(stream)) != hipSuccess)
            return; /* leaked rather than reused while a copy may still land */
    }
    pthread_mutex_lock(&g_pin_lock);
    if (g_pin_n < 64) {
        g_pin_free[g_pin_n].p = p;
        g_pin_free[g_pin_n].ev = ev;
        g_pin_n++;
        p = NULL;
        ev = NULL;
    }
    pthread_mutex_unlock(&g_pin_lock);
    if (p) {
        if (ev)
            (void)hipEventSynchronize(ev);
        (void)hipHostFree(p);
    }
    ev_put(ev);
}

/* Side streams and events are recycled too: creating and destroying them per
 * job costs milliseconds in the serving loop. */
static hipStream_t g_aux_free[64];
static int g_aux_n = 0;
static hipEvent_t g_ev_free[256];
static int g_ev_n = 0;

static int aux_get(hipStream_t *out)
{
    pthread_mutex_lock(&g_pin_lock);
    const int have = g_aux_n > 0;
    if (have)
        *out = g_aux_free[--g_aux_n];
    pthread_mutex_unlock(&g_pin_lock);
    return have || hipStreamCreateWithFlags(out, hipStreamNonBlocking) == hipSuccess;
}

static void aux_put(hipStream_t st)
{
    if (!st)
        return;
    pthread_mutex_lock(&g_pin_lock);
    const int kept = g_aux_n < 64;
    if (kept)
        g_aux_free[g_aux_n++] = st;
    pthread_mutex_unlock(&g_pin_lock);
    if (!kept)
        (void)hipStreamDestroy(st);
}

static int ev_get(hipEvent_t *out)
{
    pthread_mutex_lock(&g_pin_lock);
    const int have = g_ev_n > 0;
    if (have)
        *out = g_ev_free[--g_ev_n];
    pthread_mutex_unlock(&g_pin_lock);
    return have || hipEventCreateWithFlags(out, hipEventDisableTiming) == hipSuccess;
}

static void ev_put(hipEvent_t e)
{
    if (!e)
        return;
    pthread_mutex_lock(&g_pin_lock);
    const int kept = g_ev_n < 256;
    if (kept)
        g_ev_free[g_ev_n++] = e;
    pthread_mutex_unlock(&g_pin_lock);
    if (!kept)
        (void)hipEventDestroy(e);
}

/* Freed workspace and graph blocks are kept for reuse: hipFreeAsync and
 * hipMallocAsync cost 50-350 us of host time each, more than a small graph's
 * whole colouring. A block is reused for a request of at least half its size
 * after the new stream waits for the event recorded when it was put back;
 * the cache holds at most BLOCK_CACHE blocks and BLOCK_CACHE_BYTES bytes, and
 * evicts the largest block first. */
#define BLOCK_CACHE 32
#define BLOCK_CACHE_BYTES (16ull << 30)
static struct {
    void *p;
    size_t bytes;
    hipEvent_t ev;
} g_bc[BLOCK_CACHE];
static int g_bc_n = 0;
static size_t g_bc_bytes = 0;
static pthread_mutex_t g_bc_lock = PTHREAD_MUTEX_INITIALIZER;

static int block_get(void **p, size_t bytes, size_t *got, void *stream)
{
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    pthread_mutex_lock(&g_bc_lock);
    int best = -1;
    for (int i = 0; i < g_bc_n; i++)
        if (g_bc[i].bytes >= bytes && g_bc[i].bytes / 2 <= bytes &&
            (best < 0 || g_bc[i].bytes < g_bc[best].bytes))
            best = i;
    if (best >= 0) {
        void *q = g_bc[best].p;
        const size_t have = g_bc[best].bytes;
        hipEvent_t ev = g_bc[best].ev;
        g_bc[best] = g_bc[--g_bc_n];
        g_bc_bytes -= have;
        pthread_mutex_unlock(&g_bc_lock);
        const hipError_t e = hipStreamWaitEvent(s, ev, 0);
        ev_put(ev);
        if (e != hipSuccess) {
            (void)hipFreeAsync(q, s);
            return GCX_ERR_HIP;
        }
        *p = q;
        *got = have;
        return GCX_OK;
    }
    pthread_mutex_unlock(&g_bc_lock);
    *got = bytes;
    return gcx__dalloc(p, bytes, stream);
}

static void block_put(void *p, size_t bytes, void *stream)
{
    if (!p)
        return;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    hipEvent_t ev = NULL;
    if (bytes > BLOCK_CACHE_BYTES / 2 || !ev_get(&ev) || hipEventRecord(ev, s) != hipSuccess) {
        ev_put(ev);
        (void)hipFreeAsync(p, s);
        return;
    }
    pthread_mutex_lock(&g_bc_lock);
    while (g_bc_n > 0 && (g_bc_n == BLOCK_CACHE || g_bc_bytes + bytes > BLOCK_CACHE_BYTES)) {
        int big = 0;
        for (int i = 1; i < g_bc_n; i++)
            if (g_bc[i].bytes > g_bc[big].bytes)
                big = i;
        /* the evicted block's last use is ordered before its event */
        (void)hipStreamWaitEvent(s, g_bc[big].ev, 0);
        (void)hipFreeAsync(g_bc[big].p, s);
        ev_put(g_bc[big].ev);
        g_bc_bytes -= g_bc[big].bytes;
        g_bc[big] = g_bc[--g_bc_n];
    }
    g_bc[g_bc_n].p = p;
    g_bc[g_bc_n].bytes = bytes;
    g_bc[g_bc_n].ev = ev;
    g_bc_n++;
    g_bc_bytes += bytes;
    pthread_mutex_unlock(&g_bc_lock);
}

int gcx_workspace_free_on(gcx_workspace *ws, void *stream)
{
    if (!ws)
        return GCX_OK;
    /* Hub kernels may still be queued on the side stream when a call failed
     * between its fork and its join: order the frees after them, so the pool
     * cannot hand these buffers to another allocation while they run. */
    if (ws->aux && ws->ev_join &&
        hipEventRecord((hipEvent_t)ws->ev_join, (hipStream_t)ws->aux) == hipSuccess)
        (void)hipStreamWaitEvent((hipStream_t)gcx__pick(stream), (hipEvent_t)ws->ev_join, 0);
    block_put(ws->arena, ws->arena_bytes, stream);
    pinned_ctl_put(ws->hctl, stream);
    aux_put((hipStream_t)ws->aux);
    void *evs[] = {ws->ev_fork, ws->ev_join, ws->ev_poll[0], ws->ev_poll[1]};
    for (size_t i = 0; i < 4; i++)
        ev_put((hipEvent_t)evs[i]);
    free(ws);
    return GCX_OK;
}

int gcx_workspace_free(gcx_workspace *ws)
{
    /* any stream may still use the workspace: finish all work first */
    if (ws && hipDeviceSynchronize() != hipSuccess) {
        (void)gcx_workspace_free_on(ws, NULL);
        return GCX_ERR_HIP;
    }
    return gcx_workspace_free_on(ws, NULL);
}

int gcx_workspace_create_on(uint32_t N, uint64_t M, void *stream, gcx_workspace **out)
{
    TRY(gcx__ensure_init());
    *out = NULL;
    if (N == 0)
        return GCX_ERR_ARG;
    gcx_workspace *ws = calloc(1, sizeof(*ws));
    if (!ws)
        return GCX_ERR_NOMEM;
    ws->capN = N;
    ws->capM = M;
    const size_t n = (size_t)N + 1;
    const size_t items = sizeof(gcxk_item) * gcxk_jp_item_capacity(N, M);
    const size_t extra = order_extra(N);
    ws->order_bytes = extra ? extra : 64 * n;
    /* One pool allocation carved into every buffer (256-byte aligned): a
     * serving job creates and frees a workspace, and each stream-ordered
     * allocation or free costs tens to hundreds of microseconds of host time. */
    struct {
        void **p;
        size_t bytes;
    } part[] = {
        {&ws->key, 8 * n},
        {(void **)&ws->count, 4 * n},
        {(void **)&ws->color, 4 * n},
        {(void **)&ws->q[0], 4 * n},
        {(void **)&ws->q[1], 4 * n},
        {(void **)&ws->h[0], 4 * n},
        {(void **)&ws->h[1], 4 * n},
        {(void **)&ws->x, 64 * n},
        {(void **)&ws->y, 64 * n},
        {(void **)&ws->pass, 256},
        {(void **)&ws->ctl, 2 * sizeof(gcxk_ctl)},
        {(void **)&ws->npred, 4 * n},
        {(void **)&ws->state, 8 * gcxk_jp_state_words(N, M)},
        {(void **)&ws->wof, 4 * n},
        {(void **)&ws->wdesc, 8 * n},
        {(void **)&ws->hpos, 8 * n},
        {(void **)&ws->ovf, 4 * gcxk_jp_ovf_words(N, M)},
        {(void **)&ws->esplit, 4 * (M ? M : 1)},
        {(void **)&ws->items[0], items},
        {(void **)&ws->items[1], items},
        {&ws->order_buf, extra},
    };
    const size_t nparts = sizeof(part) / sizeof(part[0]);
    size_t total = 0;
    for (size_t i = 0; i < nparts; i++)
        total += (part[i].bytes + 255) & ~(size_t)255;
    int ok = block_get(&ws->arena, total, &ws->arena_bytes, stream) == GCX_OK;
    if (ok) {
        char *at = (char *)ws->arena;
        for (size_t i = 0; i < nparts; i++) {
            *part[i].p = part[i].bytes ? (void *)at : NULL;
            at += (part[i].bytes + 255) & ~(size_t)255;
        }
        ok = (ws->hctl = pinned_ctl_get()) != NULL;
    }
    if (ok) {
        hipStream_t aux;
        hipEvent_t e[4];
        ok = aux_get(&aux);
        if (ok)
            ws->aux = (void *)aux;
        for (int i = 0; ok && i < 4; i++) {
            ok = ev_get(&e[i]);
            if (ok) {
                if (i == 0)
                    ws->ev_fork = e[i];
                else if (i == 1)
                    ws->ev_join = e[i];
                else
                    ws->ev_poll[i - 2] = e[i];
            }
        }
    }
    if (!ok) {
        gcx_workspace_free_on(ws, stream);
        return GCX_ERR_NOMEM;
    }
    memset(ws->hctl, 0, 4 * sizeof(gcxk_ctl));
    ws->head_max = 0; /* the head is opt-in (gcx_set_ordered_head) */
    ws->head_density = GCX_HEAD_DENSITY;
    *out = ws;
    return GCX_OK;
}

int gcx_workspace_create(uint32_t N, uint64_t M, gcx_workspace **out)
{
    TRY(gcx_workspace_create_on(N, M, NULL, out));
    /* usable on any stream once the pool allocations are complete */
    HIPC(hipStreamSynchronize((hipStream_t)gcx__default_stream()));
    return GCX_OK;
}

static int ws_fits(const gcx_workspace *ws, const gcx_graph *g)
{
    return ws && g && g->N <= ws->capN && g->M <= ws->capM;
}

/* ----------------------------------------------------------- round polling */

/* Drive a device round loop: `enqueue(r)` launches round r; the device sets
 * the flag at byte offset `flag_off` of the control block once a round sees
 * nothing left to do. Two batches stay in flight so the GPU never idles on
 * the host's check. */
typedef int (*round_fn)(void *ctx, uint32_t r, hipStream_t s);

static int drive_rounds(gcx_workspace *ws, hipStream_t s, round_fn enqueue, void *ctx,
                        size_t flag_off, uint32_t max_rounds)
{
    uint32_t r = 0;
    int k = 0;
    for (;;) {
        for (uint32_t i = 0; i < ROUND_BATCH; i++)
            TRY(enqueue(ctx, r++, s));
        gcxk_ctl *snap = ws->hctl + 2 + (k & 1);
        HIPC(hipMemcpyAsync(snap, ws->ctl, sizeof(gcxk_ctl), hipMemcpyDeviceToHost, s));
        HIPC(hipEventRecord((hipEvent_t)ws->ev_poll[k & 1], s));
        if (k > 0) {
            const int pk = (k - 1) & 1;
            HIPC(hipEventSynchronize((hipEvent_t)ws->ev_poll[pk]));
            const uint32_t flag = *(volatile uint32_t *)((char *)(ws->hctl + 2 + pk) + flag_off);
            if (flag)
                break;
        }
        if (r >= max_rounds)
            return GCX_ERR_HIP; /* the loop must end within N + 1 rounds */
        k++;
    }
    ws->last_launches = r;
    return GCX_OK;
}

/* ---------------------------------------------------------------- ordering */

typedef struct {
    gcx_workspace *ws;
    const gcx_graph *g;
    int log_variant;
    uint32_t grid;
} sl_ctx;

static int sl_round(void *p, uint32_t r, hipStream_t s)
{
    sl_ctx *c = (sl_ctx *)p;
    gcx_workspace *ws = c->ws;
    const gcx_graph *g = c->g;
    uint32_t *keyr = (uint32_t *)ws->key + (g->N + 1);
    /* long rows' chunks (gcxk_sl_chunk_capacity(M) pairs, at most M / 16 +
     * M / 2048 + 2 <= M / 2 once a row is long) in E_split, which only the
     * colouring uses; rows by rank in x, the short-row list in y (the GNN row
     * buffers) */
    TRY(gcxk_sl_scan(ws->x, ws->count, keyr, (uint32_t *)ws->key, g->perm, ws->q[r & 1],
                     ws->q[(r + 1) & 1], ws->y, ws->esplit, g->M, ws->ctl, c->log_variant, r,
                     c->grid, s));
    return gcxk_sl_dec(g->N, g->V, g->Er, ws->count, keyr, ws->y, ws->esplit, g->M, ws->ctl,
                       c->log_variant, r, s);
}

/* Persistent launches (launch_ordered, below) */
enum { LAUNCH_JP_PUSH = 0, LAUNCH_JP_PULL = 1 };

static int order_sl(gcx_workspace *ws, const gcx_graph *g, int log_variant, hipStream_t s)
{
    HIPC(hipMemsetAsync(ws->ctl, 0, sizeof(gcxk_ctl), s));
    /* degrees and alive flags by degree-order rank: count and the upper half
     * of key (key holds 2 (N + 1) words) */
    TRY(gcxk_sl_init(g->N, g->V, g->perm, ws->count, (uint32_t *)ws->key + (g->N + 1), ws->q[0],
                     ws->x, ws->ctl, s));
    sl_ctx c = {ws, g, log_variant, g_round_grid};
    TRY(drive_rounds(ws, s, sl_round, &c, offsetof(gcxk_ctl, sl_done), g->N + 2 * ROUND_BATCH + 2));
    return GCX_OK;
}

static int order_gnn_fp32(gcx_workspace *ws, const gcx_graph *g, const gcx_model *m, hipStream_t s)
{
    const int L = m->layers;
    float *bufs[2] = {ws->x, ws->y};
    const float *in = NULL; /* layer 0 reads only the graph */
    for (int l = 0; l < L; l++) {
        const int kind = l == 0 ? 0 : (l == L - 1 ? 2 : 1);
        const float *W = m->dparam + (size_t)l * LAYER_PARAMS;
        const float *b = W + 512;
        float *out = bufs[l & 1];
        void *tok;
        TRY(gcx__prof_begin(kind == 0 ? GCX_ST_GNN_IN : kind == 1 ? GCX_ST_GNN_HID : GCX_ST_GNN_OUT,
                            s, &tok));
        void *hs = gcx__fork_hubs(ws, g, 1, s);
        if (!hs)
            return GCX_ERR_HIP;
        TRY(gcx__join_hubs(ws, hs, s,
                           gcxk_gnn_fp32_layer(kind, g->N, g->V, g->E, g->Ed, g->Er, g->rank, W,
                                               b, m->md, in, out, (uint32_t *)ws->key, g->hubs,
                                               g->nhub, g->perm, g->gspan, s, hs)));
        TRY(gcx__prof_end(tok, s));
        in = out;
    }
    return GCX_OK;
}

int gcx_order(gcx_workspace *ws, const gcx_graph *g, int heuristic, const gcx_model *model,
              int precision, void *stream)
{
    TRY(gcx__ensure_init());
    if (!ws_fits(ws, g))
        return ws && g ? GCX_ERR_CAPACITY : GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    void *tok = NULL;
    int rc = GCX_OK;
    ws->curN = g->N;
    switch (heuristic) {
    case GCX_FF:
    case GCX_LF:
    case GCX_LLF:
        TRY(gcx__prof_begin(GCX_ST_KEYS, s, &tok));
        rc = heuristic == GCX_FF    ? gcxk_keys_ff(g->N, (uint32_t *)ws->key, s)
             : heuristic == GCX_LF ? gcxk_keys_lf(g->N, g->V, (uint32_t *)ws->key, s)
                                   : gcxk_keys_llf(g->N, g->V, (uint32_t *)ws->key, s);
        TRY(gcx__prof_end(tok, s));
        ws->key_kind = 0;
        ws->id_order = 0;
        break;
    case GCX_SL:
    case GCX_SLL:
        TRY(gcx__prof_begin(GCX_ST_SL, s, &tok));
        rc = order_sl(ws, g, heuristic == GCX_SLL, s);
        TRY(gcx__prof_end(tok, s));
        ws->key_kind = 0;
        ws->id_order = 0;
        break;
    case GCX_GNN:
        if (!model)
            return GCX_ERR_ARG;
        if (precision == GCX_FP32)
            rc = order_gnn_fp32(ws, g, model, s);
        else if (precision == GCX_BF16 || precision == GCX_BF16X3)
            rc = gcx__order_gnn_bf16(ws, g, model, precision == GCX_BF16X3, s);
        else
            return GCX_ERR_ARG;
        ws->key_kind = 1;
        ws->id_order = 0;
        break;
    default:
        return GCX_ERR_ARG;
    }
    if (rc == GCX_OK)
        ws->key_bits = 32;
    ws->id_order = heuristic == GCX_FF;
    return rc;
}

int gcx_set_priorities(gcx_workspace *ws, const gcx_graph *g, const double *p, void *stream)
{
    TRY(gcx__ensure_init());
    if (!ws_fits(ws, g) || !p)
        return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    double *dp = (double *)ws->x; /* 64 bytes per vertex of scratch */
    gcxk_ctl *vc = ws->ctl + 1;
    HIPC(hipMemcpyAsync(dp, p, 8ull * g->N, hipMemcpyHostToDevice, s));
    HIPC(hipMemsetAsync(vc, 0, sizeof(gcxk_ctl), s));
    TRY(gcxk_keys_scan_doubles(g->N, dp, vc, s));
    HIPC(hipMemcpyAsync(ws->hctl + 1, vc, sizeof(gcxk_ctl), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (ws->hctl[1].nan_count)
        return GCX_ERR_ARG;
    ws->curN = g->N;
    if (ws->hctl[1].nonfloat == 0) {
        TRY(gcxk_keys_from_doubles32(g->N, dp, (uint32_t *)ws->key, s));
        ws->key_bits = 32;
        ws->key_kind = 1;
        ws->id_order = 0;
    } else {
        TRY(gcxk_keys_from_doubles64(g->N, dp, (uint64_t *)ws->key, s));
        ws->key_bits = 64;
        ws->key_kind = 2;
        ws->id_order = 0;
    }
    return GCX_OK;
}

/* ------------------------------------------------------------------- colour */

typedef struct {
    gcx_workspace *ws;
    const gcx_graph *g;
    uint32_t grid;
    int mb;
} jp_ctx;

static int jp_round(void *p, uint32_t r, hipStream_t s)
{
    jp_ctx *c = (jp_ctx *)p;
    gcx_workspace *ws = c->ws;
    return gcxk_jp_round(c->g->V, ws->npred, ws->esplit, ws->state, ws->color, ws->ovf, c->mb,
                         ws->items, ws->ctl, r, c->grid, s);
}

/* The stream for hub work: the workspace's side stream, forked after what s
 * has queued, when the graph has hubs and `fork` is set; else s. NULL on a
 * HIP error. */
void *gcx__fork_hubs(gcx_workspace *ws, const gcx_graph *g, int fork, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (!fork || !g->nhub)
        return s;
    if (hipEventRecord((hipEvent_t)ws->ev_fork, s) != hipSuccess ||
        hipStreamWaitEvent((hipStream_t)ws->aux, (hipEvent_t)ws->ev_fork, 0) != hipSuccess)
        return NULL;
    return ws->aux;
}

/* Makes s wait for the hub stream hs and returns rc (or GCX_ERR_HIP when rc
 * is GCX_OK and the join fails). Every path after a successful fork calls
 * it, failed launches included: a stream-ordered free on s must come after
 * the hub kernels, or the pool could hand their buffers to another
 * allocation while they still run. */
int gcx__join_hubs(gcx_workspace *ws, void *hs, void *s, int rc)
{
    if (hs == s || !hs)
        return rc;
    if (hipEventRecord((hipEvent_t)ws->ev_join, (hipStream_t)hs) != hipSuccess ||
        hipStreamWaitEvent((hipStream_t)s, (hipEvent_t)ws->ev_join, 0) != hipSuccess)
        return rc == GCX_OK ? GCX_ERR_HIP : rc;
    return rc;
}

/* fork = 0 keeps everything on s (inside a captured graph). */
static int jp_prepare_ex(gcx_workspace *ws, const gcx_graph *g, int mb, int fork, hipStream_t s)
{
    TRY(gcxk_jp_clear(g->N, g->M, ws->ovf, g->nhub, ws->hpos, g->hubs, ws->npred, ws->state,
                      ws->color, ws->ctl, s));
    void *hs = gcx__fork_hubs(ws, g, fork, s);
    if (!hs)
        return GCX_ERR_HIP;
    const int rc = gcxk_jp_init(g->N, g->V, g->E, ws->key, ws->key_bits, mb, g->perm, ws->npred,
                                ws->esplit, ws->state, ws->color, g->hubs, g->nhub, g->hchunk,
                                g->nhchunk, ws->hpos, 0u, s, hs);
    return gcx__join_hubs(ws, hs, s, rc);
}

static int jp_prepare(gcx_workspace *ws, const gcx_graph *g, int mb, hipStream_t s)
{
    return jp_prepare_ex(ws, g, mb, 1, s);
}

static int color_rounds(gcx_workspace *ws, const gcx_graph *g, int mb, hipStream_t s)
{
    void *tok = NULL;
    TRY(gcx__prof_begin(GCX_ST_JP_INIT, s, &tok));
    TRY(jp_prepare(ws, g, mb, s));
    TRY(gcxk_jp_roots(g->N, g->V, ws->npred, ws->items[0], ws->ctl, s));
    TRY(gcx__prof_end(tok, s));
    TRY(gcx__prof_begin(GCX_ST_JP_ROUNDS, s, &tok));
    jp_ctx c = {ws, g, g_round_grid, mb};
    TRY(drive_rounds(ws, s, jp_round, &c, offsetof(gcxk_ctl, done), g->N + 2 * ROUND_BATCH + 2));
    return gcx__prof_end(tok, s);
}

/* Ordered launches. g_ordered_slots = 0: cooperative, one at a time (the
 * runtime checks residency and runs the launch alone). k >= 1: plain
 * launches that share the device with other streams' kernels, at most k at
 * once, each with at most 1/k of the resident workgroups, so their grids fit
 * together. Launch i goes to slot i mod k and waits for the previous launch
 * in that slot; after a change of k the next launch waits for all slots. The
 * lock keeps pick + wait + launch + record atomic across host threads. */
#define ORDERED_MAX_SLOTS 8
/* Largest slot count accepted. A mixed serving trace with 8 slots, 16
 * workers and GPU_MAX_HW_QUEUES=16 once ended in a GPU fault whose first
 * error was never seen; 1, 2 and 4 slots have only run cleanly. */
#define ORDERED_SLOTS_ACCEPTED 4
static hipEvent_t g_slot_done[ORDERED_MAX_SLOTS];
static int g_ordered_slots = 0;
static int g_ordered_drain = 0;
static uint32_t g_next_slot = 0;
static pthread_mutex_t g_ordered_lock = PTHREAD_MUTEX_INITIALIZER;

int gcx_set_ordered_launch(int slots)
{
    TRY(gcx__ensure_init());
    if (slots < 0 || slots > ORDERED_SLOTS_ACCEPTED)
        return GCX_ERR_ARG;
    pthread_mutex_lock(&g_ordered_lock);
    if (slots != g_ordered_slots)
        g_ordered_drain = 1;
    g_ordered_slots = slots;
    pthread_mutex_unlock(&g_ordered_lock);
    return GCX_OK;
}

int gcx_get_ordered_launch(void) { return g_ordered_slots; }

/* Enqueues the ordered kernel on s under the launch policy above: the pull
 * kernel (positions in greedy order, headers in items[0], colours by
 * position in wof, order ids in ids) or the push kernel (state words). */
static int launch_ordered(gcx_workspace *ws, const gcx_graph *g, int mb, int kind,
                          const uint32_t *ids, hipStream_t s)
{
    pthread_mutex_lock(&g_ordered_lock);
    int rc = GCX_OK;
    for (int i = 0; rc == GCX_OK && i < ORDERED_MAX_SLOTS; i++)
        if (!g_slot_done[i] &&
            hipEventCreateWithFlags(&g_slot_done[i], hipEventDisableTiming) != hipSuccess)
            rc = GCX_ERR_HIP;
    const int k = g_ordered_slots;
    const uint32_t slot = k ? g_next_slot++ % (uint32_t)k : 0u;
    for (int i = 0; rc == GCX_OK && i < ORDERED_MAX_SLOTS; i++)
        if ((g_ordered_drain || i == (int)slot) && hipStreamWaitEvent(s, g_slot_done[i], 0) != hipSuccess)
            rc = GCX_ERR_HIP;
    g_ordered_drain = 0;
    uint32_t grid = ws->ordered_grid        ? ws->ordered_grid
                    : kind == LAUNCH_JP_PULL ? gcxk_pull_grid()
                                             : gcxk_jp_ordered_grid();
    if (k) {
        /* k launches share the default grid, which leaves workgroup slots
         * free on every CU (not the occupancy capacity): two plain pull
         * launches of capacity / 2 each under 8 serving workers once stayed
         * 328 and 505 of 512 workgroups resident for the whole 2 s wait bound */
        const uint32_t dflt = kind == LAUNCH_JP_PULL ? gcxk_pull_grid() : gcxk_jp_ordered_grid();
        const uint32_t share = dflt / (uint32_t)k;
        if (grid > share)
            grid = share;
        if (grid == 0)
            rc = GCX_ERR_HIP;
    }
    /* 2 s per wait: far beyond any legitimate wait, and it still ends the kernel */
    if (rc == GCX_OK && kind == LAUNCH_JP_PULL)
        rc = gcxk_pull_colour(g->N, ws->items[0], ids, ws->esplit, ws->wof, ws->color, ws->ctl,
                              grid, 200000000ull, k == 0, s);
    else if (rc == GCX_OK)
        rc = gcxk_jp_ordered(g->V, ws->npred, ws->esplit, ws->state, ws->wdesc, ws->color, ws->ovf,
                             mb, ws->items[0], ws->ctl, grid, 200000000ull, k == 0, s);
    if (rc == GCX_OK && hipEventRecord(g_slot_done[slot], s) != hipSuccess)
        rc = GCX_ERR_HIP;
    pthread_mutex_unlock(&g_ordered_lock);
    return rc;
}

/* After an ordered launch: the control block back, and whether a wait hit
 * its bound (GCX_ERR_INVALID: the caller redoes the colouring). */
static int ordered_done(gcx_workspace *ws, hipStream_t s)
{
    HIPC(hipMemcpyAsync(ws->hctl + 2, ws->ctl, sizeof(gcxk_ctl), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    ws->last_launches = 1;
    ws->last_ordered = 1;
    return ws->hctl[2].err ? GCX_ERR_INVALID : GCX_OK;
}

/* Ordered schedule without a head (the default): orientation against the
 * workspace keys, the greedy-order items, one persistent launch. With a
 * head (gcx_set_ordered_head): the greedy order first (sort), then 32-bit rank keys
 * of that order, then the orientation against the rank keys -- which also
 * writes the predecessor positions of the first head_max vertices for the
 * head kernel -- then state words and items; the head kernel colours the
 * dense top of the order in one workgroup, and the persistent launch the
 * rest. The rank keys live in ws->count (SL scratch, free during JP). */
/* Ordered schedule with pulled colours (the default): the greedy order (sort),
 * positions rank[ids[j]] = j (ws->count), the orientation writing every
 * row's predecessor positions (hub rows on the side stream), then one
 * persistent launch. Headers live in items[0], colours by position in wof. */
static int color_pull(gcx_workspace *ws, const gcx_graph *g, int mb, hipStream_t s)
{
    void *tok = NULL;
    TRY(gcx__prof_begin(GCX_ST_JP_INIT, s, &tok));
    void *scratch = ws->order_buf ? ws->order_buf : (void *)ws->x;
    uint32_t *ids = NULL;
    TRY(gcxk_jp_sort_ids(g->N, ws->key, ws->key_bits, scratch, ws->order_bytes, &ids, s));
    TRY(gcxk_pull_rank(g->N, ids, ws->count, ws->wof, g->nhub, ws->hpos, ws->ctl, s));
    void *hs = gcx__fork_hubs(ws, g, 1, s);
    if (!hs)
        return GCX_ERR_HIP;
    TRY(gcx__join_hubs(ws, hs, s,
                       gcxk_pull_orient(g->N, g->V, g->E, ws->count, g->perm, g->hubs, g->nhub,
                                        g->hchunk, g->nhchunk, ws->hpos, ws->esplit,
                                        ws->items[0], s, hs)));
    TRY(gcx__prof_end(tok, s));
    TRY(gcx__prof_begin(GCX_ST_JP_ROUNDS, s, &tok));
    TRY(launch_ordered(ws, g, mb, LAUNCH_JP_PULL, ids, s));
    TRY(gcx__prof_end(tok, s));
    return ordered_done(ws, s);
}

static int color_ordered(gcx_workspace *ws, const gcx_graph *g, int mb, hipStream_t s)
{
    void *tok = NULL;
    TRY(gcx__prof_begin(GCX_ST_JP_INIT, s, &tok));
    void *scratch = ws->order_buf ? ws->order_buf : (void *)ws->x;
    uint32_t *cnt = (uint32_t *)ws->y, *incl = cnt + ws->capN;
    uint32_t *ids = NULL, *rk = ws->count;
    const uint32_t head_max = ws->head_max;
    if (!head_max) {
        /* no head: orientation on the workspace keys, then the greedy order */
        TRY(jp_prepare(ws, g, mb, s));
        TRY(gcxk_jp_order_items(g->N, g->V, ws->npred, ws->key, ws->key_bits, mb, scratch,
                                ws->order_bytes, cnt, incl, ws->wof, ws->wdesc, ws->state,
                                gcxk_jp_state_words(ws->capN, ws->capM), ws->items[0], ws->ctl,
                                s));
        TRY(gcx__prof_end(tok, s));
        TRY(gcx__prof_begin(GCX_ST_JP_ROUNDS, s, &tok));
        TRY(launch_ordered(ws, g, mb, LAUNCH_JP_PUSH, NULL, s));
        TRY(gcx__prof_end(tok, s));
        return ordered_done(ws, s);
    }
    TRY(gcxk_jp_clear(g->N, g->M, ws->ovf, g->nhub, ws->hpos, g->hubs, ws->npred, ws->state,
                      ws->color, ws->ctl, s));
    TRY(gcxk_jp_sort_ids(g->N, ws->key, ws->key_bits, scratch, ws->order_bytes, &ids, s));
    TRY(gcxk_jp_rank_keys(g->N, ids, rk, s));
    void *hs = gcx__fork_hubs(ws, g, 1, s);
    if (!hs)
        return GCX_ERR_HIP;
    TRY(gcx__join_hubs(ws, hs, s,
                       gcxk_jp_init(g->N, g->V, g->E, rk, 32, mb, g->perm, ws->npred, ws->esplit,
                                    ws->state, ws->color, g->hubs, g->nhub, g->hchunk, g->nhchunk,
                                    ws->hpos, head_max, s, hs)));
    TRY(gcxk_jp_order_items_ranked(g->N, g->V, ws->npred, ids, rk, head_max, ws->head_density,
                                   ws->key_bits, mb, scratch, ws->order_bytes, cnt, incl, ws->wof,
                                   ws->wdesc, ws->state, gcxk_jp_state_words(ws->capN, ws->capM),
                                   ws->items[0], ws->ctl, s));
    TRY(gcx__prof_end(tok, s));
    TRY(gcx__prof_begin(GCX_ST_JP_ROUNDS, s, &tok));
    /* 2 s per wait, as in the persistent launch */
    /* the head table lives in the second item queue (rounds schedule only):
     * 8 bytes per position, and the queue holds 8 (N + M / 64 + 2) */
    if (head_max)
        TRY(gcxk_jp_head(g->V, ws->npred, ws->esplit, ids, ws->items[1], ws->color, ws->ctl,
                         200000000ull, s));
    const int rc = launch_ordered(ws, g, mb, LAUNCH_JP_PUSH, NULL, s);
    TRY(rc);
    TRY(gcx__prof_end(tok, s));
    return ordered_done(ws, s);
}

int gcx_color(gcx_workspace *ws, const gcx_graph *g, void *stream)
{
    TRY(gcx__ensure_init());
    if (!ws_fits(ws, g) || !ws->key_bits || ws->curN != g->N)
        return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    const int mb = gcxk_jp_mask_bits(g->maxdeg);
    ws->last_fallback = 0;
    const int sched = ws->schedule == GCX_SCHED_AUTO ? gcx__tuned_schedule(g) : ws->schedule;
    ws->last_ordered = 0;
    if (sched == GCX_SCHED_ORDERED && g->N) {
        /* First-fit keys (the id order) on a graph with hubs take the push
         * form: R-MAT's hubs have ids spread over the whole range, so in id
         * order the dense core is spread over the order instead of leading it,
         * and pulled colours wait on its chains in every sweep (R-MAT 22:
         * 25.2 ms pulled, 8.5 ms pushed). Every other ordering, and first-fit
         * on graphs without hubs, is faster pulled (R-MAT 22: LF 6.0 vs 7.6,
         * SL 5.5 vs 7.2, GNN3 5.6 vs 7.8 ms; 160^3 stencil FF 2.9 vs 5.0). */
        const int push = ws->head_max || (ws->id_order && g->nhub > 0);
        const int rc = push ? color_ordered(ws, g, mb, s) : color_pull(ws, g, mb, s);
        if (rc != GCX_ERR_INVALID)
            return rc;
        ws->last_fallback = 1; /* a wait hit its bound: redo with rounds */
        ws->last_ordered = 0;
    }
    return color_rounds(ws, g, mb, s);
}

int gcx_set_ordered_head(gcx_workspace *ws, uint32_t max_positions, uint32_t min_density)
{
    if (!ws || max_positions > gcxk_jp_head_max())
        return GCX_ERR_ARG;
    ws->head_max = max_positions;
    ws->head_density = min_density;
    return GCX_OK;
}

int gcx_last_head(const gcx_workspace *ws, uint32_t *positions)
{
    if (!ws || !positions)
        return GCX_ERR_ARG;
    *positions = ws->last_ordered && !ws->last_fallback ? ws->hctl[2].head_n : 0u;
    return GCX_OK;
}

int gcx_set_schedule(gcx_workspace *ws, int schedule, uint32_t param)
{
    if (!ws || schedule < GCX_SCHED_ROUNDS || schedule > GCX_SCHED_AUTO)
        return GCX_ERR_ARG;
    ws->schedule = schedule;
    ws->ordered_grid = schedule != GCX_SCHED_ROUNDS ? param : 0;
    return GCX_OK;
}

int gcx_last_fallback(const gcx_workspace *ws, uint32_t *fell_back)
{
    if (!ws || !fell_back)
        return GCX_ERR_ARG;
    *fell_back = ws->last_fallback;
    return GCX_OK;
}

int gcx_last_rounds(const gcx_workspace *ws, uint32_t *rounds)
{
    if (!ws || !rounds)
        return GCX_ERR_ARG;
    gcxk_ctl c;
    HIPC(hipMemcpy(&c, ws->ctl, sizeof(c), hipMemcpyDeviceToHost));
    *rounds = ws->last_ordered && !ws->last_fallback ? 0 : c.rounds;
    return GCX_OK;
}

int gcx_last_launches(const gcx_workspace *ws, uint32_t *launches)
{
    if (!ws || !launches)
        return GCX_ERR_ARG;
    *launches = ws->last_launches;
    return GCX_OK;
}

/* One Culberson pass on the round schedule, enqueued on s: keys from the
 * colour classes (in the strategy's class order), orientation, roots and
 * `rounds` round launches. No host reads and no memsets, so it can be
 * captured. C bounds the colours of the input colouring. */
static int recolor_pass(gcx_workspace *ws, const gcx_graph *g, int mb, uint32_t C, int strategy,
                        uint64_t seed, uint32_t rounds, hipStream_t s)
{
    TRY(gcxk_culberson_keys(g->N, C, strategy, seed, ws->color, ws->pass, (uint32_t *)ws->y,
                            (uint32_t *)ws->key, s));
    TRY(jp_prepare_ex(ws, g, mb, 0, s));
    TRY(gcxk_jp_roots(g->N, g->V, ws->npred, ws->items[0], ws->ctl, s));
    jp_ctx c = {ws, g, g_round_grid, mb};
    for (uint32_t r = 0; r < rounds; r++)
        TRY(jp_round(&c, r, s));
    return GCX_OK;
}

static int recolor_eager(gcx_workspace *ws, const gcx_graph *g, int iterations, int strategy,
                         uint64_t seed, hipStream_t s)
{
    for (int it = 0; it < iterations; it++) {
        int C = 0;
        TRY(gcx_validate(ws, g, s, &C));
        if (C <= 0)
            return GCX_ERR_INVALID;
        TRY(gcxk_culberson_keys(g->N, (uint32_t)C, strategy, seed, ws->color, ws->pass,
                                (uint32_t *)ws->y, (uint32_t *)ws->key, s));
        ws->key_bits = 32;
        ws->key_kind = 0;
        ws->id_order = 0;
        TRY(gcx_color(ws, g, s));
    }
    return GCX_OK;
}

/* Longest captured pass: beyond this many colour classes the eager loop
 * (which polls for completion) is used. */
#define RECOLOR_GRAPH_MAX_ROUNDS 4096u

int gcx_recolor_ex(gcx_workspace *ws, const gcx_graph *g, int iterations, int strategy,
                   uint64_t seed, void *stream)
{
    /* main.c:34-43: p[j] = color[j], then JP again; repeated (strategy
     * GCX_CULB_REVERSE). Other strategies reorder the colour classes.
     *
     * A vertex's predecessors all lie in classes earlier in the order, so
     * its depth in the JP DAG is at most the number of colour classes C of
     * the input colouring, and a pass never uses more classes than it was
     * given (each vertex's class position is a valid colour again). So
     * C + 1 round launches finish every pass: the passes are captured once
     * as a HIP graph and replayed, with no host round trip between rounds
     * or passes. */
    TRY(gcx__ensure_init());
    const int eager = (strategy & GCX_CULB_NO_GRAPH) != 0;
    strategy &= ~GCX_CULB_NO_GRAPH;
    if (!ws_fits(ws, g) || iterations < 0 || ws->curN != g->N || strategy < GCX_CULB_REVERSE ||
        strategy > GCX_CULB_MIXED)
        return GCX_ERR_ARG;
    if (iterations == 0 || g->N == 0)
        return GCX_OK;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    /* the pass counter of the class orders starts at 0 for this call */
    HIPC(hipMemsetAsync(ws->pass, 0, sizeof(uint32_t), s));
    int C = 0;
    TRY(gcx_validate(ws, g, s, &C));
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIPC(hipStreamIsCapturing(s, &cap));
    /* The ordered schedule needs a sort per pass (not capturable) but runs a
     * pass as one persistent launch: with ORDERED (or AUTO choosing it) the
     * passes run one gcx_color at a time. */
    const int sched = ws->schedule == GCX_SCHED_AUTO ? gcx__tuned_schedule(g) : ws->schedule;
    if (eager || sched == GCX_SCHED_ORDERED || C <= 0 || (uint32_t)C >= RECOLOR_GRAPH_MAX_ROUNDS ||
        cap != hipStreamCaptureStatusNone)
        return recolor_eager(ws, g, iterations, strategy, seed, s);
    const int mb = gcxk_jp_mask_bits(g->maxdeg);
    ws->key_bits = 32;
    ws->key_kind = 0;
    ws->id_order = 0;
    void *tok = NULL;
    TRY(gcx__prof_begin(GCX_ST_JP_ROUNDS, s, &tok));
    hipGraph_t graph = NULL;
    hipGraphExec_t exec = NULL;
    HIPC(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = recolor_pass(ws, g, mb, (uint32_t)C, strategy, seed, (uint32_t)C + 1, s);
    const hipError_t ec = hipStreamEndCapture(s, &graph);
    if (rc == GCX_OK && ec != hipSuccess)
        rc = GCX_ERR_HIP;
    if (rc == GCX_OK && hipGraphInstantiate(&exec, graph, NULL, NULL, 0) != hipSuccess)
        rc = GCX_ERR_HIP;
    for (int it = 0; rc == GCX_OK && it < iterations; it++)
        if (hipGraphLaunch(exec, s) != hipSuccess)
            rc = GCX_ERR_HIP;
    if (rc == GCX_OK)
        rc = gcx__prof_end(tok, s);
    if (rc == GCX_OK &&
        hipMemcpyAsync(ws->hctl + 2, ws->ctl, sizeof(gcxk_ctl), hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = GCX_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess && rc == GCX_OK)
        rc = GCX_ERR_HIP;
    if (exec)
        (void)hipGraphExecDestroy(exec);
    if (graph)
        (void)hipGraphDestroy(graph);
    TRY(rc);
    /* the last pass drained its queue within the captured rounds */
    if (!ws->hctl[2].done)
        return GCX_ERR_INVALID;
    ws->last_launches = (uint32_t)C + 1;
    ws->last_fallback = 0;
    return GCX_OK;
}

int gcx_recolor(gcx_workspace *ws, const gcx_graph *g, int iterations, void *stream)
{
    return gcx_recolor_ex(ws, g, iterations, GCX_CULB_REVERSE, 0, stream);
}

int gcx_validate(gcx_workspace *ws, const gcx_graph *g, void *stream, int *max_color)
{
    TRY(gcx__ensure_init());
    if (!ws_fits(ws, g) || !max_color)
        return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    gcxk_ctl *vc = ws->ctl + 1;
    HIPC(hipMemsetAsync(vc, 0, sizeof(gcxk_ctl), s));
    void *hs = gcx__fork_hubs(ws, g, 1, s);
    if (!hs)
        return GCX_ERR_HIP;
    TRY(gcx__join_hubs(ws, hs, s,
                       gcxk_validate(g->N, g->V, g->E, ws->color, g->hubs, g->nhub, g->perm, vc,
                                     s, hs)));
    HIPC(hipMemcpyAsync(ws->hctl + 1, vc, sizeof(gcxk_ctl), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    *max_color = ws->hctl[1].bad ? 0 : (ws->hctl[1].maxcolor > 0 ? ws->hctl[1].maxcolor : 1);
    return GCX_OK;
}

int gcx_color_digest(gcx_workspace *ws, const gcx_graph *g, void *stream, uint64_t *digest)
{
    TRY(gcx__ensure_init());
    if (!ws_fits(ws, g) || !digest)
        return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    gcxk_ctl *vc = ws->ctl + 1;
    HIPC(hipMemsetAsync(vc, 0, sizeof(gcxk_ctl), s));
    TRY(gcxk_color_digest(g->N, ws->color, &vc->viol, s));
    HIPC(hipMemcpyAsync(ws->hctl + 1, vc, sizeof(gcxk_ctl), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    *digest = ws->hctl[1].viol;
    return GCX_OK;
}

int gcx_mex_check(gcx_workspace *ws, const gcx_graph *g, void *stream, uint64_t *violations)
{
    TRY(gcx__ensure_init());
    if (!ws_fits(ws, g) || !violations || !ws->key_bits)
        return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    gcxk_ctl *vc = ws->ctl + 1;
    HIPC(hipMemsetAsync(vc, 0, sizeof(gcxk_ctl), s));
    TRY(gcxk_mex_check(g->N, g->V, g->E, ws->key, ws->key_bits, ws->color, g->hubs, g->nhub,
                       g->hub_deg, vc,
                       s));
    HIPC(hipMemcpyAsync(ws->hctl + 1, vc, sizeof(gcxk_ctl), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    *violations = ws->hctl[1].viol;
    return GCX_OK;
}

int gcx_download_colors(const gcx_workspace *ws, uint32_t N, int *color, void *stream)
{
    if (!ws || N > ws->capN || !color)
        return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    HIPC(hipMemcpyAsync(color, ws->color, 4ull * N, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    return GCX_OK;
}

int gcx_download_priorities(const gcx_workspace *ws, uint32_t N, double *p, void *stream)
{
    if (!ws || N > ws->capN || !p || !ws->key_bits)
        return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    double *dp = (double *)ws->y;
    if (ws->key_bits == 32)
        TRY(gcxk_keys_to_doubles(N, (const uint32_t *)ws->key, ws->key_kind, dp, s));
    else
        TRY(gcxk_keys64_to_doubles(N, (const uint64_t *)ws->key, dp, s));
    HIPC(hipMemcpyAsync(p, dp, 8ull * N, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    return GCX_OK;
}

int gcx_upload_colors(gcx_workspace *ws, uint32_t N, const int *color, void *stream)
{
    TRY(gcx__ensure_init());
    if (!ws || N > ws->capN || !color)
        return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    HIPC(hipMemcpyAsync(ws->color, color, 4ull * N, hipMemcpyHostToDevice, s));
    HIPC(hipStreamSynchronize(s));
    return GCX_OK;
}

int gcx__order_ff_only(gcx_workspace *ws, uint32_t N, void *stream)
{
    TRY(gcx__ensure_init());
    if (!ws || N > ws->capN)
        return GCX_ERR_ARG;
    TRY(gcxk_keys_ff(N, (uint32_t *)ws->key, gcx__pick(stream)));
    ws->key_bits = 32;
    ws->key_kind = 0;
    ws->id_order = 0;
    ws->curN = N;
    return GCX_OK;
}

int *gcx_workspace_colors_dev(gcx_workspace *ws) { return ws ? ws->color : NULL; }

/* ----------------------------------------------------------------- teardown */

int gcx_shutdown(void)
{
    if (!g_init)
        return GCX_OK;
    int rc = GCX_OK;
    if (g_init_rc == GCX_OK && hipDeviceSynchronize() != hipSuccess)
        rc = GCX_ERR_HIP;
    /* cached blocks: every use is finished after the device sync */
    pthread_mutex_lock(&g_bc_lock);
    for (int i = 0; i < g_bc_n; i++) {
        if (hipFreeAsync(g_bc[i].p, g_stream) != hipSuccess)
            rc = GCX_ERR_HIP;
        (void)hipEventDestroy(g_bc[i].ev);
    }
    g_bc_n = 0;
    g_bc_bytes = 0;
    pthread_mutex_unlock(&g_bc_lock);
    if (g_stream && hipStreamSynchronize(g_stream) != hipSuccess)
        rc = GCX_ERR_HIP;
    pthread_mutex_lock(&g_pin_lock);
    for (int i = 0; i < g_pin_n; i++) {
        if (g_pin_free[i].ev)
            (void)hipEventDestroy(g_pin_free[i].ev);
        (void)hipHostFree(g_pin_free[i].p);
    }
    g_pin_n = 0;
    for (int i = 0; i < g_aux_n; i++)
        (void)hipStreamDestroy(g_aux_free[i]);
    g_aux_n = 0;
    for (int i = 0; i < g_ev_n; i++)
        (void)hipEventDestroy(g_ev_free[i]);
    g_ev_n = 0;
    pthread_mutex_unlock(&g_pin_lock);
    (void)gcx_profile_reset();
    for (size_t i = 0; i < g_pool_n; i++)
        (void)hipEventDestroy(g_pool[i]);
    free(g_pool);
    g_pool = NULL;
    g_pool_n = g_pool_cap = 0;
    free(g_prof);
    g_prof = NULL;
    g_prof_n = g_prof_cap = 0;
    pthread_mutex_lock(&g_ordered_lock);
    for (int i = 0; i < ORDERED_MAX_SLOTS; i++)
        if (g_slot_done[i]) {
            (void)hipEventDestroy(g_slot_done[i]);
            g_slot_done[i] = NULL;
        }
    g_next_slot = 0;
    g_ordered_drain = 0;
    pthread_mutex_unlock(&g_ordered_lock);
    if (g_stream)
        (void)hipStreamDestroy(g_stream);
    g_stream = NULL;
    hipMemPool_t pool;
    if (g_init_rc == GCX_OK && hipDeviceGetDefaultMemPool(&pool, 0) == hipSuccess)
        (void)hipMemPoolTrimTo(pool, 0);
    g_init = 0;
    g_init_rc = GCX_OK;
    return rc;
}
