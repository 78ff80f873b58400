/*
 * tune.c -- the autotuner: per degree distribution, the hub degree (the cut
 * between four-lanes-per-vertex and block-per-vertex processing in the GNN,
 * JP orientation, validation and mex kernels) and the JP schedule (ROUNDS,
 * one launch per DAG level, vs ORDERED, one persistent launch).
 *
 * Both choices change only the speed: colours and priorities are the same
 * under every hub degree and schedule. gcx_tune measures the candidates on a
 * graph with HIP events and records the fastest under the graph's degree
 * signature (log2 buckets of N, average degree and maximum degree); graphs
 * created later with a known signature take the recorded hub degree, and
 * workspaces on GCX_SCHED_AUTO take the recorded schedule.
 */
#include <hip/hip_runtime_api.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

#define TUNE_MAX 1024

static gcx_tuning g_table[TUNE_MAX];
static int g_count = 0;
static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;

static uint32_t bits(uint64_t x)
{
    uint32_t b = 0;
    while (x) {
        b++;
        x >>= 1;
    }
    return b;
}

uint32_t gcx__signature(uint32_t N, uint64_t M, uint32_t maxdeg)
{
    const uint64_t avg = N ? M / N : 0;
    return bits(N) | bits(avg) << 8 | bits(maxdeg) << 16;
}

uint32_t gcx_degree_signature(const gcx_graph *g)
{
    return g ? gcx__signature(g->N, g->M, g->maxdeg) : 0;
}

static int lookup(uint32_t sig, gcx_tuning *out)
{
    int found = 0;
    pthread_mutex_lock(&g_lock);
    for (int i = 0; i < g_count; i++)
        if (g_table[i].signature == sig) {
            if (out)
                *out = g_table[i];
            found = 1;
            break;
        }
    pthread_mutex_unlock(&g_lock);
    return found;
}

static void record(const gcx_tuning *t)
{
    pthread_mutex_lock(&g_lock);
    int i = 0;
    while (i < g_count && g_table[i].signature != t->signature)
        i++;
    if (i == g_count && g_count < TUNE_MAX)
        g_count++;
    if (i < TUNE_MAX)
        g_table[i] = *t;
    pthread_mutex_unlock(&g_lock);
}

uint32_t gcx__tuned_hub_degree(uint32_t N, uint64_t M, uint32_t maxdeg)
{
    gcx_tuning t;
    return lookup(gcx__signature(N, M, maxdeg), &t) ? t.hub_degree : 0;
}

int gcx__tuned_schedule(const gcx_graph *g)
{
    gcx_tuning t;
    if (lookup(gcx_degree_signature(g), &t))
        return t.schedule;
    /* untuned: one persistent launch pays off once the graph fills the device */
    return g->N >= (1u << 15) ? GCX_SCHED_ORDERED : GCX_SCHED_ROUNDS;
}

int gcx_tuning_lookup(const gcx_graph *g, gcx_tuning *out)
{
    return g ? lookup(gcx_degree_signature(g), out) : 0;
}

void gcx_tuner_reset(void)
{
    pthread_mutex_lock(&g_lock);
    g_count = 0;
    pthread_mutex_unlock(&g_lock);
}

int gcx_tuner_save(const char *path)
{
    FILE *f = fopen(path, "w");
    if (!f)
        return GCX_ERR_IO;
    pthread_mutex_lock(&g_lock);
    fprintf(f, "# signature hub_degree schedule order_ms color_ms\n");
    for (int i = 0; i < g_count; i++)
        fprintf(f, "%u %u %d %.6f %.6f\n", g_table[i].signature, g_table[i].hub_degree,
                g_table[i].schedule, g_table[i].order_ms, g_table[i].color_ms);
    pthread_mutex_unlock(&g_lock);
    return fclose(f) == 0 ? GCX_OK : GCX_ERR_IO;
}

int gcx_tuner_load(const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f)
        return GCX_ERR_IO;
    char line[256];
    int rc = GCX_OK;
    while (fgets(line, sizeof(line), f)) {
        if (line[0] == '#' || line[0] == '\n')
            continue;
        gcx_tuning t;
        if (sscanf(line, "%u %u %d %f %f", &t.signature, &t.hub_degree, &t.schedule, &t.order_ms,
                   &t.color_ms) != 5 ||
            t.hub_degree == 0 || t.schedule < GCX_SCHED_ROUNDS || t.schedule > GCX_SCHED_ORDERED) {
            rc = GCX_ERR_FORMAT;
            break;
        }
        record(&t);
    }
    fclose(f);
    return rc;
}

/* Mean of `reps` timed runs after one warm-up, in ms. */
typedef int (*tune_step)(void *ctx, hipStream_t s);

static int time_step(tune_step fn, void *ctx, hipStream_t s, int reps, float *ms)
{
    hipEvent_t a = NULL, b = NULL;
    int rc = fn(ctx, s);
    if (rc == GCX_OK && (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess))
        rc = GCX_ERR_HIP;
    if (rc == GCX_OK && hipEventRecord(a, s) != hipSuccess)
        rc = GCX_ERR_HIP;
    for (int i = 0; rc == GCX_OK && i < reps; i++)
        rc = fn(ctx, s);
    if (rc == GCX_OK && (hipEventRecord(b, s) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                         hipEventElapsedTime(ms, a, b) != hipSuccess))
        rc = GCX_ERR_HIP;
    if (rc == GCX_OK)
        *ms /= (float)reps;
    if (a)
        (void)hipEventDestroy(a);
    if (b)
        (void)hipEventDestroy(b);
    return rc;
}

typedef struct {
    gcx_workspace *ws;
    gcx_graph *g;
    const gcx_model *m;
    int precision;
} tune_ctx;

static int step_order(void *p, hipStream_t s)
{
    tune_ctx *c = (tune_ctx *)p;
    return c->m ? gcx_order(c->ws, c->g, GCX_GNN, c->m, c->precision, s)
                : gcx_order(c->ws, c->g, GCX_LF, NULL, GCX_FP32, s);
}

static int step_color(void *p, hipStream_t s)
{
    tune_ctx *c = (tune_ctx *)p;
    return gcx_color(c->ws, c->g, s);
}

static int step_all(void *p, hipStream_t s)
{
    TRY(step_order(p, s));
    return step_color(p, s);
}

int gcx_tune(gcx_graph *g, const gcx_model *m, int precision, void *stream, gcx_tuning *out)
{
    TRY(gcx__ensure_init());
    if (!g || (m && precision != GCX_FP32 && precision != GCX_BF16 && precision != GCX_BF16X3))
        return GCX_ERR_ARG;
    hipStream_t s = (hipStream_t)gcx__pick(stream);
    gcx_workspace *ws = NULL;
    TRY(gcx_workspace_create_on(g->N, g->M, s, &ws));
    tune_ctx c = {ws, g, m, precision};
    gcx_tuning best;
    memset(&best, 0, sizeof(best));
    best.signature = gcx_degree_signature(g);
    const uint32_t was = g->hub_deg;
    int rc = GCX_OK;
    /* hub degree: candidates that give distinct hub sets, timed on order +
     * colour with the current schedule (the schedule does not touch hubs) */
    const uint32_t cand[] = {256u, 1024u, 4096u, 16384u};
    float best_ms = 0.0f;
    for (size_t i = 0; rc == GCX_OK && i < sizeof(cand) / sizeof(cand[0]); i++) {
        if (i > 0 && cand[i - 1] > g->maxdeg)
            break; /* no vertex reaches this or the previous cut: same hub set */
        rc = gcx_graph_set_hub_degree(g, cand[i], s);
        if (rc == GCX_OK)
            rc = gcx_set_schedule(ws, GCX_SCHED_ROUNDS, 0);
        float ms = 0.0f;
        if (rc == GCX_OK)
            rc = time_step(step_all, &c, s, 2, &ms);
        if (rc == GCX_OK && (best.hub_degree == 0 || ms < best_ms)) {
            best_ms = ms;
            best.hub_degree = cand[i];
        }
    }
    if (rc == GCX_OK)
        rc = gcx_graph_set_hub_degree(g, best.hub_degree, s);
    if (rc == GCX_OK)
        rc = time_step(step_order, &c, s, 2, &best.order_ms);
    /* schedule: the same priorities coloured both ways */
    float ms_sched[2] = {0.0f, 0.0f};
    for (int sc = GCX_SCHED_ROUNDS; rc == GCX_OK && sc <= GCX_SCHED_ORDERED; sc++) {
        rc = gcx_set_schedule(ws, sc, 0);
        if (rc == GCX_OK)
            rc = time_step(step_color, &c, s, 2, &ms_sched[sc]);
    }
    if (rc == GCX_OK) {
        best.schedule = ms_sched[GCX_SCHED_ORDERED] < ms_sched[GCX_SCHED_ROUNDS] ? GCX_SCHED_ORDERED
                                                                                : GCX_SCHED_ROUNDS;
        best.color_ms = ms_sched[best.schedule];
        record(&best);
        if (out)
            *out = best;
    } else if (was) {
        (void)gcx_graph_set_hub_degree(g, was, s);
    }
    (void)gcx_workspace_free_on(ws, s);
    if (hipStreamSynchronize(s) != hipSuccess && rc == GCX_OK)
        rc = GCX_ERR_HIP;
    return rc;
}
