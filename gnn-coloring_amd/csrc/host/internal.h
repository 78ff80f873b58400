/*
 * internal.h -- private structures of the C host layer.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "gcx.h"
#include "gcxk.h"

struct gcx_graph {
    uint32_t N;
    uint64_t M;
    uint32_t *V;      /* device, N+1 */
    uint32_t *E;      /* device, M */
    uint32_t *hubs;   /* device, nhub entries (degree >= hub_deg) */
    uint32_t hub_deg; /* block-per-vertex cut (0 until the stats run) */
    int hub_fixed;    /* hub_deg set by the caller: the tuner leaves it */
    uint32_t nhub;
    uint32_t maxdeg;
    uint32_t *hchunk; /* device, nhub+1: prefix of 1024-edge chunks per hub */
    uint32_t nhchunk;
    uint32_t *perm;   /* device, N: ids by degree descending (hubs first) */
    uint32_t *rank;   /* device, N: rank[perm[i]] = i */
    uint32_t *Er;     /* device, M + pad: rank[E[j]] (GNN rows live at degree-order positions) */
    uint32_t *Ed;     /* device, M + pad: deg(E[j]) (the GNN input layer's degree feature) */
    uint32_t *gspan;  /* device, 2N, or NULL: (first vertex, count) of each vertex's packed graph */
    uint32_t *goff;   /* device, count + 1, or NULL: first vertex of each packed graph */
    uint32_t *hgoff;  /* host copy of goff (malloc'ed), or NULL */
    uint32_t count;   /* packed graphs (0 = a single graph) */
    void *blk;        /* one device allocation holding hubs, hchunk, perm, rank, Er, Ed */
    int blk_ve;       /* ... and V and E */
    size_t blk_bytes; /* size of blk */
};

struct gcx_model {
    int params, md, layers;
    float *dparam; /* device copy of the weights */
    float *hparam; /* host copy */
    uint16_t *dWb;    /* bf16 weights, 1024 per layer: W[k][j] k-major, hi then lo */
    uint16_t *dWpack; /* bf16 MFMA B operands, 1024 per layer, lane order, hi then lo */
};

struct gcx_workspace {
    uint32_t capN;
    uint64_t capM;
    void *key;        /* 8 * capN bytes: key32 or key64 view */
    int key_bits;     /* 32 / 64 for the current priorities, 0 = none */
    int key_kind;     /* 32-bit keys: 0 unsigned integer, 1 float order */
    int id_order;     /* the keys are first-fit's: the greedy order is the id order */
    uint32_t *count;  /* SL: remaining degrees */
    int *color;
    uint32_t *q[2];   /* SL alive lists */
    uint32_t *h[2];   /* SL removed list */
    uint32_t *npred;  /* JP: predecessors per vertex */
    uint64_t *state;  /* JP state words (gcxk_jp_state_words; layouts in jp.hip) */
    uint32_t *wof;    /* JP (ordered): first state word of each vertex, N + 1 */
    uint64_t *wdesc;  /* JP (ordered): state word descriptor of each vertex, N */
    void *arena;      /* the one device allocation all of the above are carved from */
    size_t arena_bytes;
    uint32_t *ovf;    /* JP: colour-set overflow words (colours > mb) */
    uint32_t *esplit; /* JP: rows split into predecessors | successors (capM) */
    gcxk_item *items[2]; /* JP: round queues; items[0] = greedy-order items */
    uint32_t *hpos;   /* JP: per-hub E_split positions (2 per vertex slot) */
    float *x, *y;     /* GNN rows, 16 floats (or 16 bf16) per vertex */
    gcxk_ctl *ctl;    /* device control block */
    gcxk_ctl *hctl;   /* pinned host mirror */
    void *aux;        /* side stream for hub kernels */
    void *ev_fork, *ev_join;
    void *ev_poll[2];
    uint32_t last_rounds;
    uint32_t last_sl_rounds;
    uint32_t last_launches; /* round kernels enqueued by the last round loop */
    int schedule;           /* GCX_SCHED_ROUNDS or GCX_SCHED_ORDERED */
    uint32_t ordered_grid;  /* workgroups of the ordered launch (0 = all resident) */
    uint32_t last_fallback; /* 1 when the last ordered launch hit its wait bound */
    int last_ordered;       /* the last gcx_color ran the ordered schedule */
    uint32_t *pass;         /* Culberson pass counter (device word; survives gcx_color) */
    void *order_buf;        /* sort scratch when it does not fit in x */
    size_t order_bytes;     /* bytes of sort scratch available */
    uint32_t curN;    /* vertex count of the graph the contents belong to */
    uint32_t head_max;     /* ordered schedule: longest head prefix (0 = no head) */
    uint32_t head_density; /* ... and the mean predecessor count that makes a block dense */
};

/* Default head density: a 1024-position block of the greedy order belongs to
 * the head when its vertices have at least this many predecessors on
 * average (R-MAT scale 22, GNN3: ~460-1000 in the first 32,768 positions;
 * 27-point stencils and Erdos-Renyi graphs stay far below it). */
#define GCX_HEAD_DENSITY 64u

/* Return GCX_ERR_HIP when a HIP call fails / the status of a failing call. */
#define HIPC(expr)                                                              \
    do {                                                                        \
        if ((expr) != hipSuccess)                                               \
            return GCX_ERR_HIP;                                                 \
    } while (0)

#define TRY(expr)                                                               \
    do {                                                                        \
        int rc_ = (expr);                                                       \
        if (rc_ != GCX_OK)                                                      \
            return rc_;                                                         \
    } while (0)

/* tune.c */
uint32_t gcx__signature(uint32_t N, uint64_t M, uint32_t maxdeg);
/* Recorded hub degree for this degree signature, 0 if none. */
uint32_t gcx__tuned_hub_degree(uint32_t N, uint64_t M, uint32_t maxdeg);
/* Recorded schedule for the graph's signature, else the size rule. */
int gcx__tuned_schedule(const gcx_graph *g);

/* device.c */
int gcx__graph_alloc(uint32_t N, uint64_t M, void *stream, gcx_graph **out);
/* Stream-ordered device memory from the device pool (stream NULL = the
 * library's stream); freeing never synchronises the device. */
int gcx__dalloc(void **p, size_t bytes, void *stream);
void gcx__dfree(void *p, void *stream);
int gcx__graph_fill(gcx_graph *g, uint32_t N, const uint32_t *V, const uint32_t *E, void *stream);
/* FF keys (N - id) for a workspace without a graph (first_fit.c:8-19). */
int gcx__order_ff_only(gcx_workspace *ws, uint32_t N, void *stream);
/* bf16.c */
int gcx__model_pack_bf16(gcx_model *m);
/* split = 0: GCX_BF16, 1: GCX_BF16X3. */
int gcx__order_gnn_bf16(gcx_workspace *ws, const gcx_graph *g, const gcx_model *m, int split,
                        void *stream);
int gcx__ensure_init(void);
void *gcx__default_stream(void);
void *gcx__pick(void *stream);
int gcx__prof_begin(int stage, void *stream, void **token);
/* Hub work on the workspace's side stream: fork returns the stream to put
 * it on (s itself when the graph has no hubs or fork is 0; NULL on a HIP
 * error); join makes s wait for it and passes rc through. Every path after
 * a successful fork joins, failed launches included. */
void *gcx__fork_hubs(gcx_workspace *ws, const gcx_graph *g, int fork, void *s);
int gcx__join_hubs(gcx_workspace *ws, void *hs, void *s, int rc);
int gcx__prof_end(void *token, void *stream);
