/*
 * gcxk.h -- the thin C ABI between the C host layer and the HIP kernels.
 * Every pointer here is a device pointer; `stream` is a hipStream_t.
 * Every launcher returns GCX_OK or GCX_ERR_HIP (launch failure).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Vertices with degree >= GCXK_HUB_DEG are processed one per workgroup. */
#define GCXK_HUB_DEG 1024u

/* Control block shared by the JP / SL / validate kernels (device memory). */
typedef struct {
    uint32_t qn[3];         /* small-vertex frontier sizes, slot = round % 3 */
    uint32_t hn[3];         /* hub frontier sizes, slot = round % 3 */
    uint32_t rounds;        /* non-empty JP rounds executed */
    uint32_t done;          /* set when a JP round found an empty frontier */
    uint32_t bad;           /* validation: violations found */
    int32_t maxcolor;       /* validation: largest colour */
    uint64_t viol;          /* mex certificate violations */
    uint32_t nan_count;     /* priorities that are NaN (rejected) */
    uint32_t nonfloat;      /* priorities not exactly representable as float */
    uint32_t sl_alive[3];   /* SL: alive-list length, slot = round % 3 */
    uint32_t sl_gone[3];    /* SL: removed-this-round list length */
    uint32_t sl_mininv[3];  /* SL: UINT32_MAX - min remaining degree */
    uint32_t sl_delta[3];   /* SL: threshold of the round */
    uint32_t sl_rounds;
    uint32_t sl_done;
} gcxk_ctl;

/* ---- graph helpers (csr.hip) ---- */
/* Canonical CSR from pairs already on the device. Allocates *dE (hipMalloc,
 * caller frees) and fills dV (N+1 entries). Temporaries are freed inside. */
int gcxk_csr_from_pairs(uint32_t N, uint64_t L, const uint32_t *dI, const uint32_t *dJ,
                        uint32_t *dV, uint32_t **dE, uint64_t *M, void *stream);
/* Hub list: vertices with degree >= GCXK_HUB_DEG, ascending id. Writes the
 * count to *nhub (host) and the ids to dhub (capacity N). Also returns max
 * degree. Uses dtmp (>= 8 bytes device scratch). */
int gcxk_graph_stats(uint32_t N, const uint32_t *dV, uint32_t *dhub, uint32_t *dtmp,
                     uint32_t *nhub, uint32_t *maxdeg, void *stream);
/* CSR validity (sorted, simple, symmetric): *valid = 1 or 0. */
int gcxk_graph_check(uint32_t N, const uint32_t *dV, const uint32_t *dE, uint32_t *dtmp,
                     int *valid, void *stream);
/* Index bounds every kernel relies on: V[0] = 0, V non-decreasing,
 * V[N] = M, all E < N. *ok = 1 when they hold. */
int gcxk_csr_bounds(uint32_t N, uint64_t M, const uint32_t *dV, const uint32_t *dE,
                    uint32_t *dtmp, int *ok, void *stream);
/* Sort every row ascending in place (repeats kept); synchronises. */
int gcxk_sort_rows(uint32_t N, const uint32_t *dV, uint32_t *dE, uint64_t M, void *stream);
/* Drop self loops and repeats from sorted rows: dV rewritten, survivors
 * compacted into dE_out (capacity M); *M_out = new entry count. */
int gcxk_make_simple(uint32_t N, uint32_t *dV, const uint32_t *dE, uint64_t M, uint32_t *dE_out,
                     uint64_t *M_out, void *stream);
/* Device pair generators. */
int gcxk_gen_rmat(int scale, uint64_t L, uint32_t t0, uint32_t t1, uint32_t t2, uint64_t seed,
                  uint32_t *dI, uint32_t *dJ, void *stream);
int gcxk_gen_er(uint32_t n, uint64_t m, uint64_t seed, uint32_t *dI, uint32_t *dJ, void *stream);

/* ---- ordering keys (keys.hip); key32: larger = earlier ---- */
int gcxk_keys_ff(uint32_t N, uint32_t *key, void *stream);
int gcxk_keys_lf(uint32_t N, const uint32_t *dV, uint32_t *key, void *stream);
int gcxk_keys_llf(uint32_t N, const uint32_t *dV, uint32_t *key, void *stream);
int gcxk_keys_from_colors(uint32_t N, const int *color, uint32_t *key, void *stream);
/* Host doubles (already on device) -> key32 when every value is exactly a
 * float, else key64; counts NaNs and non-floats in ctl. */
int gcxk_keys_scan_doubles(uint32_t N, const double *p, gcxk_ctl *ctl, void *stream);
int gcxk_keys_from_doubles32(uint32_t N, const double *p, uint32_t *key, void *stream);
int gcxk_keys_from_doubles64(uint32_t N, const double *p, uint64_t *key, void *stream);
/* key32 -> doubles: kind 0 = unsigned integer, 1 = float order key. */
int gcxk_keys_to_doubles(uint32_t N, const uint32_t *key, int kind, double *p, void *stream);
int gcxk_keys64_to_doubles(uint32_t N, const uint64_t *key, double *p, void *stream);

/* ---- SL / SLL peeling (sl.hip) ---- */
/* key[u] = UINT32_MAX marks "not yet removed"; a removed vertex's key is the
 * round threshold (delta) at its removal. ctl must be zeroed. */
int gcxk_sl_init(uint32_t N, const uint32_t *dV, uint32_t *deg, uint32_t *key, uint32_t *alive0,
                 gcxk_ctl *ctl, void *stream);
/* Round r, phase 1: remove alive vertices with degree <= threshold. */
int gcxk_sl_scan(uint32_t *deg, uint32_t *key, const uint32_t *alive_in, uint32_t *alive_out,
                 uint32_t *gone, gcxk_ctl *ctl, int log_variant, uint32_t r, uint32_t grid,
                 void *stream);
/* Round r, phase 2: decrement unremoved neighbours of removed vertices. */
int gcxk_sl_dec(const uint32_t *dV, const uint32_t *dE, uint32_t *deg, const uint32_t *key,
                const uint32_t *gone, gcxk_ctl *ctl, uint32_t r, uint32_t grid, void *stream);

/* ---- GraphSAGE-16 (gnn.hip) ---- */
/* FP32, bit-exact with the reference op order. kind 0 = input layer (x
 * unused), 1 = hidden (x -> out), 2 = output (x -> key, float order keys).
 * Rows are 16 floats. Hub rows run on hub_stream, the rest on stream; the
 * caller joins the two streams before the next layer. */
int gcxk_gnn_fp32_layer(int kind, uint32_t N, const uint32_t *dV, const uint32_t *dE,
                        const float *W, const float *b, int md, const float *x, float *out,
                        uint32_t *key, const uint32_t *hubs, uint32_t nhub, void *stream,
                        void *hub_stream);
/* BF16 rows (16 x bf16 = 32 B), FP32 aggregation, MFMA layer transform.
 * Wpack: the layer's two MFMA B operands in lane order (512 bf16); Wb: the
 * layer's weights W[k][j] in bf16 (512); b: FP32 biases (16). */
int gcxk_gnn_bf16_layer(int kind, uint32_t N, const uint32_t *dV, const uint32_t *dE,
                        const uint16_t *Wpack, const uint16_t *Wb, const float *b, int md,
                        const uint16_t *x, uint16_t *out, uint32_t *key, const uint32_t *hubs,
                        uint32_t nhub, void *stream, void *hub_stream);

/* ---- Jones-Plassmann (jp.hip) ---- */
/* Work item: (vertex, chunk of 64 successor edges). Stored as two uint32. */
typedef struct gcxk_item {
    uint32_t u, k;
} gcxk_item;

/* Overflow colour-set words for N vertices / M entries (colours > 64). */
static inline uint64_t gcxk_jp_ovf_words(uint32_t N, uint64_t M)
{
    return M / 32u + (uint64_t)N + 2u;
}
/* Item queue capacity: every vertex contributes max(1, ceil(nsucc / 64)). */
static inline uint64_t gcxk_jp_item_capacity(uint32_t N, uint64_t M)
{
    return (uint64_t)N + M / 64u + 2u;
}
/* Hub chunks for the orientation pass: 4096 edges each. */
#define GCXK_HUB_CHUNK_EDGES 4096u

/* Zero colours, colour sets (mask64: N words, ovf: gcxk_jp_ovf_words), the
 * hubs' counters and per-hub positions (hpos: 2 per hub), and the control
 * block. A plain kernel, so the sequence can be captured in a graph. */
int gcxk_jp_clear(uint32_t N, uint64_t M, int *color, uint64_t *mask64, uint32_t *ovf,
                  const uint32_t *hubs, uint32_t nhub, uint32_t *count, uint32_t *npred,
                  uint32_t *hpos, gcxk_ctl *ctl, void *stream);
/* Orientation: key_bits = 32 or 64. E_split rows = predecessors then
 * successors; count = npred = predecessor counts; roots queued as items in
 * slot 0. hchunk: per-hub prefix of 4096-edge chunks (nhub + 1 entries),
 * nhchunk = total chunks. */
int gcxk_jp_init(uint32_t N, const uint32_t *dV, const uint32_t *dE, const void *key,
                 int key_bits, uint32_t *count, uint32_t *npred, uint32_t *esplit,
                 gcxk_item *q0, const uint32_t *hubs, uint32_t nhub, const uint32_t *hchunk,
                 uint32_t nhchunk, uint32_t *hpos, gcxk_ctl *ctl, void *stream);
/* One JP round r: colours the items of slot r, queues the next round's. */
int gcxk_jp_round(const uint32_t *dV, const uint32_t *npred, const uint32_t *esplit,
                  uint32_t *count, int *color, uint64_t *mask64, uint32_t *ovf,
                  gcxk_item *const q[2], gcxk_ctl *ctl, uint32_t r, uint32_t grid, void *stream);
/* Degrees of the listed vertices (for the hub chunk table). */
int gcxk_gather_degrees(const uint32_t *dV, const uint32_t *list, uint32_t n, uint32_t *out,
                        void *stream);
/* Proper-colouring check: ctl->bad, ctl->maxcolor (ctl fields zeroed by caller). */
int gcxk_validate(uint32_t N, const uint32_t *dV, const uint32_t *dE, const int *color,
                  const uint32_t *hubs, uint32_t nhub, gcxk_ctl *ctl, void *stream);
/* Greedy-rule certificate: ctl->viol. */
int gcxk_mex_check(uint32_t N, const uint32_t *dV, const uint32_t *dE, const void *key,
                   int key_bits, const int *color, gcxk_ctl *ctl, void *stream);

/* Number of workgroups the round kernel should use (fills the GPU). */
uint32_t gcxk_round_grid(void);

#ifdef __cplusplus
}
#endif
