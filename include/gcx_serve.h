/*
 * gcx_serve.h -- a request loop for colouring jobs on one MI355X.
 *
 * Jobs of different graph sizes and heuristics arrive over time; a fixed set
 * of worker threads, each with its own HIP stream, runs them concurrently on
 * the device. A job is admitted when it reaches the head of the arrival-order
 * queue and its HBM footprint (gcx_request_footprint) fits in what the
 * budget has left; memory freed by finished jobs admits the next ones. Graph
 * construction, priorities, colouring and validation all stay on the device;
 * only the colour count and a digest of the colouring come back. Models are
 * loaded once and shared.
 *
 * No counterpart in the reference (its runner colours one host graph at a
 * time); this is the serving extension of gcx.h.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "gcx.h"

#ifdef __cplusplus
extern "C" {
#endif

enum gcx_generator {
    GCX_GEN_RMAT = 0, /* R-MAT, a/b/c = .57/.19/.19, n rounded up to 2^scale */
    GCX_GEN_ER = 1,   /* Erdos-Renyi G(n, m), m = n * degree / 2 pairs */
    GCX_GEN_RGG = 2,  /* 2D random geometric graph in the unit square */
    GCX_GEN_RESIDENT = 3 /* a graph registered with gcx_server_add_graph;
                            n is its id, param and seed are unused */
};

typedef struct {
    int generator;    /* GCX_GEN_* */
    uint32_t n;       /* vertices */
    double param;     /* RMAT: edge factor; ER: average degree; RGG: radius
                         (0 = sqrt(12 / (pi n)), about 12 neighbours) */
    uint64_t seed;
    int heuristic;    /* GCX_FF .. GCX_GNN */
    int layers;       /* GNN layers 2..4 (models m2/m3/m4) */
    int precision;    /* GCX_FP32 / GCX_BF16 (GNN only) */
    int schedule;     /* GCX_SCHED_ROUNDS / GCX_SCHED_ORDERED */
    double arrival;   /* seconds after the start of gcx_server_run */
} gcx_request;

typedef struct {
    int status;         /* GCX_OK, GCX_ERR_BUDGET (never fits), or an error */
    int batch;          /* jobs in the pass that served this one (1 = alone) */
    uint32_t N;
    uint64_t M;         /* CSR entries of the generated graph */
    int colors;         /* validated colour count (0 = invalid) */
    uint64_t digest;    /* gcx_color_digest of the colouring */
    uint64_t footprint; /* HBM bytes reserved at admission */
    int worker;
    double t_arrival, t_admit, t_done; /* seconds since the start of the run */
} gcx_reply;

typedef struct gcx_server gcx_server;

/* workers >= 1 threads/streams; hbm_budget in bytes (0 = 3/4 of free HBM);
 * model_dir holds m2.gnn, m3.gnn, m4.gnn (NULL = "models"). */
int gcx_server_create(int workers, uint64_t hbm_budget, const char *model_dir, gcx_server **out);
/* Runs n requests sorted by arrival time: each is queued at its arrival
 * time and replies[i] is filled when request i is done. Blocks until all
 * are done. */
int gcx_server_run(gcx_server *s, const gcx_request *reqs, size_t n, gcx_reply *replies);
int gcx_server_destroy(gcx_server *s);
/* Registers a device graph that requests name by id (GCX_GEN_RESIDENT):
 * jobs on it read the resident CSR and only allocate a workspace. The caller
 * keeps ownership and frees the graph after the server. Its bytes are taken
 * from the budget for the server's lifetime. Not while a run is active. */
int gcx_server_add_graph(gcx_server *s, const gcx_graph *g, uint32_t *id);
/* HBM bytes a request may occupy at its peak (graph build or colouring);
 * for GCX_GEN_RESIDENT the server counts the graph's workspace instead and
 * this returns 0. */
uint64_t gcx_request_footprint(const gcx_request *r);
/* Peak HBM bytes reserved by admitted jobs during the last run. */
uint64_t gcx_server_peak_reserved(const gcx_server *s);
/* Continuous batching (0 or 1 = off, the default): a worker that admits a
 * request on a resident graph also takes up to max_batch - 1 more queued
 * right behind it with the same heuristic, layers, precision and schedule
 * (as budget allows), packs their graphs into one block-diagonal graph
 * (gcx_graph_pack) and runs them as one ordering + colouring pass. Replies
 * are those of the jobs alone (per-graph colour count and digest). Not while
 * a run is active. */
int gcx_server_set_batching(gcx_server *s, int max_batch);
/* Express workers (0 = off, the default): workers 0 .. workers - 1 admit
 * only jobs whose footprint (gcx_request_footprint, or the workspace of a
 * resident-graph job) is at most max_footprint bytes, taking the earliest
 * such job in the queue even when larger jobs arrived before it; the other
 * workers keep first-come-first-served order. Small jobs then do not wait
 * behind multi-million-vertex builds for a worker. workers must be below the
 * server's worker count. Not while a run is active. */
int gcx_server_set_express(gcx_server *s, int workers, uint64_t max_footprint);

#ifdef __cplusplus
}
#endif
