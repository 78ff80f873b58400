/*
 * gcx.h -- device-resident extension API of the MI355X colouring library.
 *
 * The reference-compatible entry points (coloring.h, fast_gnn.h, gnn.h,
 * runner.h, the ordering headers, fast_io.h) take host arrays and copy them
 * to the GPU on every call. This header adds what those signatures cannot
 * express: graphs, weights and scratch that stay resident in HBM across
 * calls, explicit HIP streams (passed as void *), status codes, BF16 GNN
 * inference and batched colouring. It has no counterpart in the reference.
 *
 * All functions return GCX_OK (0) or a negative gcx_status. Streams are
 * hipStream_t values passed as void * (NULL = the library's own stream).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum gcx_status {
    GCX_OK = 0,
    GCX_ERR_ARG = -1,      /* invalid argument (size, NaN priority, ...) */
    GCX_ERR_NOMEM = -2,    /* host or device allocation failed */
    GCX_ERR_HIP = -3,      /* a HIP runtime call failed */
    GCX_ERR_CAPACITY = -4, /* input larger than the workspace / scratch */
    GCX_ERR_IO = -5,       /* file cannot be opened */
    GCX_ERR_FORMAT = -6,   /* malformed or truncated file */
    GCX_ERR_NODEV = -7,    /* no usable gfx950 device */
    GCX_ERR_INVALID = -8,  /* a colouring or CSR failed validation */
    GCX_ERR_BUDGET = -9    /* admission rejected: HBM budget exceeded */
};

/* Ordering heuristics (names follow the reference's runner labels). */
enum gcx_heuristic {
    GCX_FF = 0,   /* first fit:            p = N - id          */
    GCX_LF = 1,   /* largest degree first: p = deg             */
    GCX_LLF = 2,  /* largest log degree:   p = bitlength(deg)  */
    GCX_SL = 3,   /* smallest degree last: peel level          */
    GCX_SLL = 4,  /* smallest log degree last                  */
    GCX_GNN = 5   /* GraphSAGE-16 priorities from a gcx_model  */
};

enum gcx_precision {
    GCX_FP32 = 0,  /* bit-exact with the reference's FP32 GNN */
    GCX_BF16 = 1,  /* BF16 rows + MFMA layer transforms, FP32 accumulate */
    GCX_BF16X3 = 2 /* FP32 rows; operands and weights split into bf16 hi + lo,
                      three MFMAs per product: within a few FP32 ulps of FP32 */
};

typedef struct gcx_graph gcx_graph;
typedef struct gcx_model gcx_model;
typedef struct gcx_workspace gcx_workspace;

const char *gcx_status_string(int status);

/* Initialise device 0 (must be gfx950). Called implicitly by every entry
 * point; returns GCX_ERR_NODEV when no usable GPU is present. */
int gcx_init(void);
/* Library version string and the gfx target it was built for. */
const char *gcx_version(void);
/* Waits for all device work, then releases what the library keeps for
 * reuse: cached freed blocks, pinned control blocks, side streams, events,
 * the ordered-launch slot events and the library's stream. Call it after
 * freeing the graphs, workspaces, models and servers, before exit; a later
 * call of any entry point initialises the library again. */
int gcx_shutdown(void);

/* ---- streams and timing (thin wrappers so callers need no HIP headers) */
int gcx_stream_create(void **stream);
int gcx_stream_destroy(void *stream);
int gcx_stream_sync(void *stream);
int gcx_device_sync(void);
int gcx_event_create(void **event);
int gcx_event_destroy(void *event);
int gcx_event_record(void *event, void *stream);
int gcx_event_elapsed_ms(void *start, void *stop, float *ms);
/* Free and total HBM in bytes. */
int gcx_mem_info(size_t *free_bytes, size_t *total_bytes);

/* ---- device-resident graphs (canonical CSR: symmetric, sorted, simple) ----
 * Graphs and workspaces live in the device's stream-ordered memory pool. */
int gcx_graph_from_csr(uint32_t N, const uint32_t *V, const uint32_t *E, void *stream,
                       gcx_graph **out);
/* Host pair list (0-based) -> canonical CSR built on the device. */
int gcx_graph_from_pairs(uint32_t N, uint64_t L, const uint32_t *I, const uint32_t *J,
                         void *stream, gcx_graph **out);
/* Pairs generated on the device (same pairs as the host generators). */
int gcx_graph_rmat(int scale, uint32_t edge_factor, double a, double b, double c,
                   uint64_t seed, void *stream, gcx_graph **out);
int gcx_graph_er(uint32_t n, uint64_t m, uint64_t seed, void *stream, gcx_graph **out);
/* 2D random geometric graph: n points in the unit square, an edge for every
 * pair closer than `radius` (same pairs as gcx_gen_rgg2d_pairs). */
int gcx_graph_rgg(uint32_t n, double radius, uint64_t seed, void *stream, gcx_graph **out);
/* Free after finishing all device work (like hipFree). */
int gcx_graph_free(gcx_graph *g);
/* Stream-ordered free: the memory returns to the device pool once the work
 * already enqueued on `stream` is done; every other use of g must be ordered
 * before it. Never synchronises the device. */
int gcx_graph_free_on(gcx_graph *g, void *stream);
uint32_t gcx_graph_num_vertices(const gcx_graph *g);
uint64_t gcx_graph_num_entries(const gcx_graph *g);
uint32_t gcx_graph_max_degree(const gcx_graph *g);
int gcx_graph_download(const gcx_graph *g, uint32_t *V, uint32_t *E);
/* Device CSR check (sorted, simple, symmetric, V[N] == M): 1 valid, 0 not. */
int gcx_graph_check(const gcx_graph *g, void *stream, int *valid);
/* ---- autotuner (no reference counterpart) ----
 * Per degree distribution: the hub degree, i.e. the cut between four lanes
 * per vertex (16 vertices per wave) and one workgroup per vertex in the GNN,
 * JP orientation and validation kernels, and the JP schedule. Neither changes
 * results. A graph's degree signature = log2 buckets of N, the average and
 * the maximum degree. */
typedef struct {
    uint32_t signature;
    uint32_t hub_degree; /* block-per-vertex from this degree up */
    int schedule;        /* GCX_SCHED_ROUNDS / GCX_SCHED_ORDERED */
    float order_ms;      /* measured ordering time at the chosen hub degree */
    float color_ms;      /* measured colouring time under the chosen schedule */
} gcx_tuning;
uint32_t gcx_degree_signature(const gcx_graph *g);
/* Measures hub degrees 256..16384 (order + colour) and both schedules on g
 * (GNN with model m in `precision`, or LF keys when m is NULL), records the
 * fastest under g's signature and applies the hub degree to g. Graphs
 * created later with the same signature take the recorded hub degree. */
int gcx_tune(gcx_graph *g, const gcx_model *m, int precision, void *stream, gcx_tuning *out);
/* 1 and *out filled when g's signature has a record, else 0. */
int gcx_tuning_lookup(const gcx_graph *g, gcx_tuning *out);
void gcx_tuner_reset(void);
/* Text file of records: "signature hub_degree schedule order_ms color_ms". */
int gcx_tuner_save(const char *path);
int gcx_tuner_load(const char *path);
/* Fix g's hub degree (0 = the tuned value or the default); rebuilds the hub
 * tables on `stream` and waits for it. Not while other streams use g
 * (gcx_tune calls it too). */
int gcx_graph_set_hub_degree(gcx_graph *g, uint32_t hub_degree, void *stream);
uint32_t gcx_graph_hub_degree(const gcx_graph *g);

/* Pack graphs into one block-diagonal graph (graph i's vertices follow graph
 * i-1's). Ordering + colouring the pack gives every graph exactly the
 * priorities and colours it gets alone (the GNN keeps graph-local id
 * features), in one pass over the device: many small graphs become one
 * launch sequence and one persistent JP kernel. */
int gcx_graph_pack(const gcx_graph *const *graphs, uint32_t count, void *stream, gcx_graph **out);
/* Redo the per-graph preparation every constructor does once: degree
 * statistics, the hub list and its chunk table, the degree order, its
 * inverse and the columns remapped to it (Er). Results do not change; the
 * call exists to time that one-time cost. Synchronises `stream`. */
int gcx_graph_prepare(gcx_graph *g, void *stream);
/* Number of graphs packed into g (0 for a plain graph). */
uint32_t gcx_graph_packed_count(const gcx_graph *g);
/* Bytes of HBM a graph with N vertices and M entries occupies. */
uint64_t gcx_graph_footprint(uint32_t N, uint64_t M);

/* ---- GNN weights ("params md" + params floats, layers of 32x16 + 16) */
int gcx_model_load(const char *path, gcx_model **out);
int gcx_model_from_params(int params, const float *param, int md, gcx_model **out);
int gcx_model_free(gcx_model *m);
int gcx_model_layers(const gcx_model *m);

/* ---- workspace: scratch for ordering + colouring of graphs up to (N, M) */
int gcx_workspace_create(uint32_t N, uint64_t M, gcx_workspace **out);
/* Allocation ordered on `stream` (usable there at once, elsewhere after it). */
int gcx_workspace_create_on(uint32_t N, uint64_t M, void *stream, gcx_workspace **out);
int gcx_workspace_free(gcx_workspace *ws);
/* Stream-ordered free, as gcx_graph_free_on. */
int gcx_workspace_free_on(gcx_workspace *ws, void *stream);
uint64_t gcx_workspace_footprint(uint32_t N, uint64_t M);

/* Compute priorities for g into the workspace (model/precision used by
 * GCX_GNN only). */
int gcx_order(gcx_workspace *ws, const gcx_graph *g, int heuristic, const gcx_model *model,
              int precision, void *stream);
/* Load caller priorities (host doubles, N of them) into the workspace. */
int gcx_set_priorities(gcx_workspace *ws, const gcx_graph *g, const double *p, void *stream);
/* Jones-Plassmann colouring with the workspace's priorities (p descending,
 * id ascending); colours 1..C land in the workspace. */
int gcx_color(gcx_workspace *ws, const gcx_graph *g, void *stream);
/* JP schedule; all give the same colours.
 *   ROUNDS:  one launch per DAG level over a compacted frontier.
 *   ORDERED: work items sorted in greedy order, then one persistent launch
 *            of `param` workgroups (0 = 6 per CU; at most as many as are
 *            resident at once) in which each wave takes a fixed stride of
 *            the items and waits on their predecessor counts. Concurrent
 *            ordered launches follow gcx_set_ordered_launch. If a wait
 *            exceeds its bound the call redoes the colouring with ROUNDS.
 *   AUTO:    the schedule the autotuner recorded for the graph's degree
 *            signature (gcx_tune), else ORDERED from 2^15 vertices up. */
enum gcx_schedule { GCX_SCHED_ROUNDS = 0, GCX_SCHED_ORDERED = 1, GCX_SCHED_AUTO = 2 };
int gcx_set_schedule(gcx_workspace *ws, int schedule, uint32_t param);
/* How ORDERED launches share the device (process-wide). 0: cooperative
 * launches, one at a time, each running alone on the device. k = 1..4:
 * plain launches that overlap other streams' kernels, at most k at once,
 * each with at most 1/k of the resident workgroups. Colours are the same
 * under every setting. */
int gcx_set_ordered_launch(int slots);
int gcx_get_ordered_launch(void);
/* The ORDERED schedule's head (no reference counterpart): before the
 * persistent launch, one workgroup colours the dense top of the greedy order
 * with the colours in LDS -- on skewed graphs most of the DAG's depth sits
 * there. It covers the leading 1024-position blocks whose vertices have at
 * least min_density predecessors on average, at most max_positions (<=
 * 32768; 0 turns the head off). Colours never change. Off by default: on
 * R-MAT scale 22 (GNN3) it is still slower than the persistent launch alone
 * (an LDS hop in the head costs 0.2-0.8 us against ~2 us for a device-scope
 * hop, but folding each vertex's ~500 predecessor colours has not been
 * taken off the chain). Suggested setting: 32768 positions, density 64. */
int gcx_set_ordered_head(gcx_workspace *ws, uint32_t max_positions, uint32_t min_density);
/* Greedy-order positions the head coloured in the last gcx_color (0: no
 * head, or not the ORDERED schedule). */
int gcx_last_head(const gcx_workspace *ws, uint32_t *positions);
/* 1 when the last gcx_color fell back from ORDERED to ROUNDS. */
int gcx_last_fallback(const gcx_workspace *ws, uint32_t *fell_back);
/* Culberson iterated greedy: `iterations` more passes, each with p = colour
 * (main.c:34-43); see gcx_recolor_ex for how passes run. */
int gcx_recolor(gcx_workspace *ws, const gcx_graph *g, int iterations, void *stream);
/* Iterated greedy with other colour-class orders (no reference counterpart;
 * every order keeps the colour count non-increasing):
 *   REVERSE        higher colours first (= gcx_recolor)
 *   LARGEST_FIRST  classes by size descending, ties by colour descending
 *   RANDOM         seeded random class order, a new order every pass
 *                  (pass i of a call uses seed + i)
 *   MIXED          pass i uses REVERSE, LARGEST_FIRST, RANDOM for i mod 3
 * Uses the workspace's GNN row buffer as scratch. Under the ROUNDS schedule
 * each pass is a captured graph replay; under ORDERED (or AUTO choosing it)
 * each pass is one gcx_color call (a sort plus one persistent launch).
 * OR-ing GCX_CULB_NO_GRAPH into the strategy forces one gcx_color call per
 * pass under any schedule. */
enum gcx_culberson {
    GCX_CULB_REVERSE = 0,
    GCX_CULB_LARGEST_FIRST = 1,
    GCX_CULB_RANDOM = 2,
    GCX_CULB_MIXED = 3,
    GCX_CULB_NO_GRAPH = 0x100
};
int gcx_recolor_ex(gcx_workspace *ws, const gcx_graph *g, int iterations, int strategy,
                   uint64_t seed, void *stream);
/* Proper-colouring check of the workspace colours: *max_color = C or 0. */
int gcx_validate(gcx_workspace *ws, const gcx_graph *g, void *stream, int *max_color);
/* Per packed graph of g: its colour count (largest colour) and the digest
 * gcx_color_digest gives for that graph coloured alone (count entries). */
int gcx_packed_colors(gcx_workspace *ws, const gcx_graph *g, int *colors, uint64_t *digests,
                      void *stream);
/* Order-independent 64-bit digest of the workspace colours: the sum over u
 * of a hash of (u, colour[u]); equal colourings give equal digests. */
int gcx_color_digest(gcx_workspace *ws, const gcx_graph *g, void *stream, uint64_t *digest);
/* Greedy-rule certificate: number of vertices whose colour is not the mex of
 * their predecessors' colours under the workspace priorities. */
int gcx_mex_check(gcx_workspace *ws, const gcx_graph *g, void *stream, uint64_t *violations);

int gcx_download_colors(const gcx_workspace *ws, uint32_t N, int *color, void *stream);
/* Load caller colours (host ints, N of them) into the workspace. */
int gcx_upload_colors(gcx_workspace *ws, uint32_t N, const int *color, void *stream);
/* Priorities as doubles (the reference's p). */
int gcx_download_priorities(const gcx_workspace *ws, uint32_t N, double *p, void *stream);
/* Number of JP rounds (DAG depth) of the last gcx_color call under ROUNDS
 * (0 under ORDERED, which has no rounds). */
int gcx_last_rounds(const gcx_workspace *ws, uint32_t *rounds);
/* Kernels the last colouring loop enqueued (ROUNDS: rounds + the empty
 * tail; ORDERED: 1). */
int gcx_last_launches(const gcx_workspace *ws, uint32_t *launches);

/* Device pointers for callers that bring their own kernels. */
int *gcx_workspace_colors_dev(gcx_workspace *ws);
const uint32_t *gcx_graph_V_dev(const gcx_graph *g);
const uint32_t *gcx_graph_E_dev(const gcx_graph *g);

/* ---- per-stage kernel timing (HIP events around each launch) ---- */
enum gcx_stage {
    GCX_ST_KEYS = 0,
    GCX_ST_SL = 1,
    GCX_ST_GNN_IN = 2,
    GCX_ST_GNN_HID = 3,
    GCX_ST_GNN_OUT = 4,
    GCX_ST_JP_INIT = 5,
    GCX_ST_JP_ROUNDS = 6,
    GCX_ST_COUNT = 7
};
/* Copy `bytes` between two fresh HBM buffers with a plain streaming kernel
 * (for calibrating hardware byte counters); *ms = kernel time. */
int gcx_calibration_copy(uint64_t bytes, float *ms);
int gcx_profile_enable(int on);
int gcx_profile_reset(void);
/* Accumulated milliseconds and launch count per stage (synchronises). */
int gcx_profile_read(double ms[GCX_ST_COUNT], uint64_t launches[GCX_ST_COUNT]);

/* ---- host generators (pair lists; same pairs as the device versions) ---- */
void gcx_rmat_thresholds(double a, double b, double c, uint32_t thr[3]);
int gcx_gen_rmat_pairs(int scale, uint32_t edge_factor, double a, double b, double c,
                       uint64_t seed, uint32_t *I, uint32_t *J);
int gcx_gen_er_pairs(uint32_t n, uint64_t m, uint64_t seed, uint32_t *I, uint32_t *J);
int gcx_gen_permutation(uint32_t n, uint64_t seed, uint32_t *perm);
uint64_t gcx_stencil27_pair_count(uint32_t g);
int gcx_gen_stencil27_pairs(uint32_t g, uint64_t seed, uint32_t *I, uint32_t *J);
int gcx_gen_rgg2d_pairs(uint32_t n, double radius, uint64_t seed, uint64_t cap, uint32_t *I,
                        uint32_t *J, uint64_t *L);

#ifdef __cplusplus
}
#endif
