/*
 * serve.c -- request loop over the device-resident path (gcx_serve.h).
 *
 * The submitting thread queues each request at its arrival time; worker
 * threads take the queue head when its footprint fits the remaining budget
 * (first come, first served: a large head job waits for memory rather than
 * being overtaken) -- except express workers (gcx_server_set_express), which
 * take the earliest queued job whose footprint is at most the express size,
 * so small jobs are not held behind multi-million-vertex ones -- and run
 * generate -> order -> colour -> validate -> digest
 * on their own stream, then free the job's memory in stream order so no call
 * synchronises the whole device. Jobs on different workers overlap on the
 * GPU; ordered JP launches are serialised by the library itself.
 */
#define _GNU_SOURCE
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "gcx_serve.h"
#include "internal.h"

typedef struct worker {
    struct gcx_server *srv;
    int id;
    hipStream_t stream;
    pthread_t thread;
} worker;

struct gcx_server {
    int nworkers;
    worker *w;
    uint64_t budget;
    gcx_model *model[5]; /* by layer count 2..4 */
    const gcx_graph **resident;
    uint32_t nresident, resident_cap;
    int max_batch; /* continuous batching of resident-graph jobs (<= 1: off) */
    int express_workers;    /* workers 0 .. express_workers-1 take small jobs only */
    uint64_t express_bytes; /* ... those with a footprint of at most this */
    pthread_mutex_t lock;
    pthread_cond_t cv;      /* queue / memory changed */
    pthread_cond_t done_cv; /* a job finished */
    int stop;
    /* current run */
    const gcx_request *req;
    gcx_reply *rep;
    size_t *queue, head, tail, done, n;
    unsigned char *taken; /* taken[k]: queue entry k was admitted (express workers skip ahead) */
    uint64_t in_use, peak;
    struct timespec t0;
};

static double since(const struct timespec *t0)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)(t.tv_sec - t0->tv_sec) + 1e-9 * (double)(t.tv_nsec - t0->tv_nsec);
}

static int rmat_scale(uint32_t n)
{
    int s = 1;
    while (s < 31 && (1u << s) < n)
        s++;
    return s;
}

static double rgg_radius(const gcx_request *r)
{
    return r->param > 0.0 ? r->param : sqrt(12.0 / (M_PI * (double)(r->n ? r->n : 1)));
}

/* Generated pairs (upper bound for RGG) and vertex count. */
static uint64_t request_pairs(const gcx_request *r, uint32_t *N)
{
    switch (r->generator) {
    case GCX_GEN_RMAT: {
        const int sc = rmat_scale(r->n);
        *N = 1u << sc;
        const uint64_t ef = r->param >= 1.0 ? (uint64_t)r->param : 16u;
        return ef << sc;
    }
    case GCX_GEN_ER:
        *N = r->n;
        return (uint64_t)llround((double)r->n * (r->param > 0.0 ? r->param : 16.0) / 2.0);
    default: {
        *N = r->n;
        const double rad = rgg_radius(r);
        const double pairs = 0.5 * (double)r->n * (double)r->n * M_PI * rad * rad;
        return (uint64_t)(1.25 * pairs) + 1024u;
    }
    }
}

uint64_t gcx_request_footprint(const gcx_request *r)
{
    if (r->generator == GCX_GEN_RESIDENT)
        return 0;
    uint32_t N = 0;
    const uint64_t L = request_pairs(r, &N);
    const uint64_t M = 2 * L;
    /* build: pairs (8L), two key arrays (32L), flags + positions (16L), the
     * radix-sort scratch and E (4M) */
    const uint64_t build = 8 * L + 32 * L + 16 * L + 8 * L + 4 * M + (64u << 20);
    const uint64_t ws = gcx_workspace_footprint(N, M);
    return gcx_graph_footprint(N, M) + (build > ws ? build : ws);
}

static int make_graph(const gcx_request *r, hipStream_t s, gcx_graph **g)
{
    switch (r->generator) {
    case GCX_GEN_RMAT:
        return gcx_graph_rmat(rmat_scale(r->n), r->param >= 1.0 ? (uint32_t)r->param : 16u, 0.57,
                              0.19, 0.19, r->seed, s, g);
    case GCX_GEN_ER: {
        uint32_t N;
        const uint64_t m = request_pairs(r, &N);
        return gcx_graph_er(r->n, m, r->seed, s, g);
    }
    case GCX_GEN_RGG:
        return gcx_graph_rgg(r->n, rgg_radius(r), r->seed, s, g);
    default:
        return GCX_ERR_ARG;
    }
}

static const gcx_graph *resident_graph(const gcx_server *S, const gcx_request *r)
{
    return (r->generator == GCX_GEN_RESIDENT && r->n < S->nresident) ? S->resident[r->n] : NULL;
}

static uint64_t job_footprint(const gcx_server *S, const gcx_request *r)
{
    if (r->generator != GCX_GEN_RESIDENT)
        return gcx_request_footprint(r);
    const gcx_graph *g = resident_graph(S, r);
    return g ? gcx_workspace_footprint(g->N, g->M) : 0;
}

static void run_job(worker *w, const gcx_request *r, gcx_reply *out)
{
    gcx_server *S = w->srv;
    hipStream_t s = w->stream;
    const gcx_graph *res = resident_graph(S, r);
    gcx_graph *g = (gcx_graph *)res;
    gcx_workspace *ws = NULL;
    int rc = res ? GCX_OK : r->generator == GCX_GEN_RESIDENT ? GCX_ERR_ARG : make_graph(r, s, &g);
    if (rc == GCX_OK)
        rc = gcx_workspace_create_on(g->N, g->M, s, &ws);
    if (rc == GCX_OK)
        rc = gcx_set_schedule(ws, r->schedule, 0);
    if (rc == GCX_OK) {
        const gcx_model *m = NULL;
        if (r->heuristic == GCX_GNN)
            m = (r->layers >= 2 && r->layers <= 4) ? S->model[r->layers] : NULL;
        rc = (r->heuristic == GCX_GNN && !m) ? GCX_ERR_ARG
                                             : gcx_order(ws, g, r->heuristic, m, r->precision, s);
    }
    if (rc == GCX_OK)
        rc = gcx_color(ws, g, s);
    int colors = 0;
    uint64_t digest = 0;
    if (rc == GCX_OK)
        rc = gcx_validate(ws, g, s, &colors);
    if (rc == GCX_OK)
        rc = gcx_color_digest(ws, g, s, &digest);
    out->status = rc;
    out->N = g ? g->N : 0;
    out->M = g ? g->M : 0;
    out->colors = colors;
    out->digest = digest;
    out->worker = w->id;
    (void)gcx_workspace_free_on(ws, s);
    if (!res)
        (void)gcx_graph_free_on(g, s);
}

/* Jobs that can share one packed pass: resident graphs, same ordering. */
static int batchable(const gcx_server *S, const gcx_request *a, const gcx_request *b)
{
    return resident_graph(S, a) && resident_graph(S, b) && a->heuristic == b->heuristic &&
           (a->heuristic != GCX_GNN || (a->layers == b->layers && a->precision == b->precision)) &&
           a->schedule == b->schedule;
}

/* HBM of one packed pass over graphs with nv vertices and ne entries in
 * total: the pack and its workspace. */
static uint64_t pack_footprint(uint64_t nv, uint64_t ne)
{
    return gcx_graph_footprint((uint32_t)nv, ne) + gcx_workspace_footprint((uint32_t)nv, ne);
}

#define MAX_BATCH 64

/* One pass over the packed graphs of reqs[idx[0..n)]. */
static void run_batch(worker *w, const size_t *idx, int n)
{
    gcx_server *S = w->srv;
    hipStream_t s = w->stream;
    const gcx_graph *gs[MAX_BATCH];
    for (int k = 0; k < n; k++)
        gs[k] = resident_graph(S, &S->req[idx[k]]);
    const gcx_request *r = &S->req[idx[0]];
    gcx_graph *pack = NULL;
    gcx_workspace *ws = NULL;
    int colors[MAX_BATCH];
    uint64_t digests[MAX_BATCH];
    int rc = gcx_graph_pack(gs, (uint32_t)n, s, &pack);
    if (rc == GCX_OK)
        rc = gcx_workspace_create_on(pack->N, pack->M, s, &ws);
    if (rc == GCX_OK)
        rc = gcx_set_schedule(ws, r->schedule, 0);
    if (rc == GCX_OK) {
        const gcx_model *m = NULL;
        if (r->heuristic == GCX_GNN)
            m = (r->layers >= 2 && r->layers <= 4) ? S->model[r->layers] : NULL;
        rc = (r->heuristic == GCX_GNN && !m) ? GCX_ERR_ARG
                                             : gcx_order(ws, pack, r->heuristic, m, r->precision, s);
    }
    if (rc == GCX_OK)
        rc = gcx_color(ws, pack, s);
    int proper = 0;
    if (rc == GCX_OK)
        rc = gcx_validate(ws, pack, s, &proper);
    if (rc == GCX_OK)
        rc = gcx_packed_colors(ws, pack, colors, digests, s);
    for (int k = 0; k < n; k++) {
        gcx_reply *out = &S->rep[idx[k]];
        out->status = rc;
        out->batch = n;
        out->N = gs[k]->N;
        out->M = gs[k]->M;
        out->colors = rc == GCX_OK && proper ? colors[k] : 0;
        out->digest = rc == GCX_OK ? digests[k] : 0;
        out->worker = w->id;
    }
    (void)gcx_workspace_free_on(ws, s);
    (void)gcx_graph_free_on(pack, s);
}

/* Queue position of the job worker w may admit now, or -1: the head when it
 * fits the budget; for an express worker, the earliest job of at most the
 * express size that fits together with the head. Express admissions thus
 * never hold the head back for memory: while the head waits for memory,
 * the queue is first come, first served for every worker. */
static long pick_job(gcx_server *S, const worker *w)
{
    while (S->head < S->tail && S->taken[S->head])
        S->head++;
    if (S->head >= S->tail)
        return -1;
    const uint64_t head_fp = S->rep[S->queue[S->head]].footprint;
    if (w->id >= S->express_workers)
        return S->in_use + head_fp <= S->budget ? (long)S->head : -1;
    for (size_t k = S->head; k < S->tail; k++) {
        const uint64_t fp = S->rep[S->queue[k]].footprint;
        const uint64_t need = S->in_use + fp + (k == S->head ? 0 : head_fp);
        if (!S->taken[k] && fp <= S->express_bytes && need <= S->budget)
            return (long)k;
    }
    return -1;
}

static void *worker_main(void *arg)
{
    worker *w = (worker *)arg;
    gcx_server *S = w->srv;
    pthread_mutex_lock(&S->lock);
    for (;;) {
        long k0 = -1;
        while (!S->stop && (k0 = pick_job(S, w)) < 0)
            pthread_cond_wait(&S->cv, &S->lock);
        if (S->stop)
            break;
        S->taken[k0] = 1;
        const size_t i = S->queue[k0];
        gcx_reply *rp = &S->rep[i];
        size_t idx[MAX_BATCH];
        int n = 1;
        idx[0] = i;
        uint64_t reserve = rp->footprint;
        if (S->max_batch > 1 && w->id >= S->express_workers && resident_graph(S, &S->req[i])) {
            /* take the compatible jobs queued right behind it while the
             * packed pass fits the budget */
            const gcx_graph *g0 = resident_graph(S, &S->req[i]);
            uint64_t nv = g0->N, ne = g0->M;
            for (size_t k = (size_t)k0 + 1; n < S->max_batch && n < MAX_BATCH && k < S->tail;
                 k++) {
                if (S->taken[k])
                    continue;
                const size_t j = S->queue[k];
                const gcx_graph *gj = resident_graph(S, &S->req[j]);
                if (!batchable(S, &S->req[i], &S->req[j]) || nv + gj->N >= UINT32_MAX ||
                    ne + gj->M >= UINT32_MAX ||
                    S->in_use + pack_footprint(nv + gj->N, ne + gj->M) > S->budget)
                    break;
                nv += gj->N;
                ne += gj->M;
                idx[n++] = j;
                S->taken[k] = 1;
            }
            if (n > 1)
                reserve = pack_footprint(nv, ne);
        }
        S->in_use += reserve;
        if (S->in_use > S->peak)
            S->peak = S->in_use;
        const double t_admit = since(&S->t0);
        for (int k = 0; k < n; k++) {
            S->rep[idx[k]].t_admit = t_admit;
            if (n > 1)
                S->rep[idx[k]].footprint = reserve;
        }
        pthread_mutex_unlock(&S->lock);
        if (n > 1) {
            run_batch(w, idx, n);
        } else {
            run_job(w, &S->req[i], rp);
            rp->batch = 1;
        }
        /* the job's frees are ordered on this stream; wait for them so the
         * budget never counts memory the pool has not got back */
        (void)hipStreamSynchronize(w->stream);
        pthread_mutex_lock(&S->lock);
        S->in_use -= reserve;
        const double t_done = since(&S->t0);
        for (int k = 0; k < n; k++)
            S->rep[idx[k]].t_done = t_done;
        S->done += (size_t)n;
        pthread_cond_broadcast(&S->cv);
        pthread_cond_broadcast(&S->done_cv);
    }
    pthread_mutex_unlock(&S->lock);
    return NULL;
}

int gcx_server_create(int workers, uint64_t hbm_budget, const char *model_dir, gcx_server **out)
{
    *out = NULL;
    int rc = gcx_init();
    if (rc != GCX_OK)
        return rc;
    if (workers < 1 || workers > 64)
        return GCX_ERR_ARG;
    gcx_server *S = calloc(1, sizeof(*S));
    if (!S)
        return GCX_ERR_NOMEM;
    S->w = calloc((size_t)workers, sizeof(worker));
    if (!S->w) {
        free(S);
        return GCX_ERR_NOMEM;
    }
    pthread_mutex_init(&S->lock, NULL);
    pthread_cond_init(&S->cv, NULL);
    pthread_cond_init(&S->done_cv, NULL);
    if (!hbm_budget) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            gcx_server_destroy(S);
            return GCX_ERR_HIP;
        }
        hbm_budget = (uint64_t)fr / 4 * 3;
    }
    S->budget = hbm_budget;
    for (int l = 2; l <= 4; l++) {
        char path[4096];
        snprintf(path, sizeof(path), "%s/m%d.gnn", model_dir ? model_dir : "models", l);
        if ((rc = gcx_model_load(path, &S->model[l])) != GCX_OK) {
            gcx_server_destroy(S);
            return rc;
        }
    }
    for (int i = 0; i < workers; i++) {
        worker *w = &S->w[i];
        w->srv = S;
        w->id = i;
        if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess ||
            pthread_create(&w->thread, NULL, worker_main, w) != 0) {
            w->stream = NULL;
            gcx_server_destroy(S);
            return GCX_ERR_HIP;
        }
        S->nworkers = i + 1;
    }
    *out = S;
    return GCX_OK;
}

int gcx_server_run(gcx_server *S, const gcx_request *reqs, size_t n, gcx_reply *replies)
{
    if (!S || (n && (!reqs || !replies)))
        return GCX_ERR_ARG;
    for (size_t i = 1; i < n; i++)
        if (reqs[i].arrival < reqs[i - 1].arrival)
            return GCX_ERR_ARG;
    size_t *queue = malloc(sizeof(size_t) * (n ? n : 1));
    unsigned char *taken = calloc(n ? n : 1, 1);
    if (!queue || !taken) {
        free(queue);
        free(taken);
        return GCX_ERR_NOMEM;
    }
    memset(replies, 0, sizeof(gcx_reply) * n);
    for (size_t i = 0; i < n; i++)
        replies[i].footprint = job_footprint(S, &reqs[i]);
    pthread_mutex_lock(&S->lock);
    S->req = reqs;
    S->rep = replies;
    S->queue = queue;
    S->taken = taken;
    S->head = S->tail = S->done = 0;
    S->n = n;
    S->in_use = S->peak = 0;
    clock_gettime(CLOCK_MONOTONIC, &S->t0);
    pthread_mutex_unlock(&S->lock);
    for (size_t i = 0; i < n; i++) {
        /* sleep to within 200 us of the arrival, then spin */
        for (;;) {
            const double wait = reqs[i].arrival - since(&S->t0);
            if (wait <= 0.0)
                break;
            if (wait > 2e-4) {
                struct timespec ts = {0, (long)((wait - 1e-4) * 1e9)};
                nanosleep(&ts, NULL);
            }
        }
        pthread_mutex_lock(&S->lock);
        gcx_reply *rp = &replies[i];
        rp->t_arrival = since(&S->t0);
        if (rp->footprint > S->budget) {
            rp->status = GCX_ERR_BUDGET;
            rp->t_admit = rp->t_done = rp->t_arrival;
            rp->worker = -1;
            S->done++;
        } else {
            S->queue[S->tail++] = i;
            pthread_cond_broadcast(&S->cv);
        }
        pthread_mutex_unlock(&S->lock);
    }
    pthread_mutex_lock(&S->lock);
    while (S->done < n)
        pthread_cond_wait(&S->done_cv, &S->lock);
    S->queue = NULL;
    S->taken = NULL;
    S->req = NULL;
    S->rep = NULL;
    pthread_mutex_unlock(&S->lock);
    free(queue);
    free(taken);
    return GCX_OK;
}

int gcx_server_add_graph(gcx_server *S, const gcx_graph *g, uint32_t *id)
{
    if (!S || !g || !id)
        return GCX_ERR_ARG;
    const uint64_t bytes = gcx_graph_footprint(g->N, g->M);
    pthread_mutex_lock(&S->lock);
    int rc = GCX_OK;
    if (S->req)
        rc = GCX_ERR_ARG; /* a run is active */
    else if (bytes > S->budget)
        rc = GCX_ERR_BUDGET;
    else if (S->nresident == S->resident_cap) {
        const uint32_t cap = S->resident_cap ? 2 * S->resident_cap : 8;
        const gcx_graph **p = realloc(S->resident, sizeof(*p) * cap);
        if (!p)
            rc = GCX_ERR_NOMEM;
        else {
            S->resident = p;
            S->resident_cap = cap;
        }
    }
    if (rc == GCX_OK) {
        S->budget -= bytes;
        *id = S->nresident;
        S->resident[S->nresident++] = g;
    }
    pthread_mutex_unlock(&S->lock);
    return rc;
}

uint64_t gcx_server_peak_reserved(const gcx_server *S) { return S ? S->peak : 0; }

int gcx_server_set_batching(gcx_server *S, int max_batch)
{
    if (!S || max_batch < 0 || max_batch > MAX_BATCH)
        return GCX_ERR_ARG;
    pthread_mutex_lock(&S->lock);
    const int active = S->req != NULL;
    if (!active)
        S->max_batch = max_batch;
    pthread_mutex_unlock(&S->lock);
    return active ? GCX_ERR_ARG : GCX_OK;
}

int gcx_server_set_express(gcx_server *S, int workers, uint64_t max_footprint)
{
    if (!S || workers < 0 || workers >= S->nworkers)
        return GCX_ERR_ARG;
    pthread_mutex_lock(&S->lock);
    const int active = S->req != NULL;
    if (!active) {
        S->express_workers = workers;
        S->express_bytes = max_footprint;
    }
    pthread_mutex_unlock(&S->lock);
    return active ? GCX_ERR_ARG : GCX_OK;
}

int gcx_server_destroy(gcx_server *S)
{
    if (!S)
        return GCX_OK;
    pthread_mutex_lock(&S->lock);
    S->stop = 1;
    pthread_cond_broadcast(&S->cv);
    pthread_mutex_unlock(&S->lock);
    for (int i = 0; i < S->nworkers; i++)
        pthread_join(S->w[i].thread, NULL);
    for (int i = 0; i < S->nworkers; i++)
        if (S->w[i].stream)
            (void)hipStreamDestroy(S->w[i].stream);
    for (int l = 2; l <= 4; l++)
        gcx_model_free(S->model[l]);
    pthread_cond_destroy(&S->cv);
    pthread_cond_destroy(&S->done_cv);
    pthread_mutex_destroy(&S->lock);
    free(S->resident);
    free(S->w);
    free(S);
    return GCX_OK;
}
