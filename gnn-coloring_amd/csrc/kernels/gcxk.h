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

/* Default hub degree: vertices with degree >= a graph's hub degree are
 * processed one per workgroup (block-per-vertex), the others four lanes per
 * vertex (16 vertices per wave). Each graph carries its own value. */
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
    uint32_t sl_gone[3];    /* SL: removed-this-round short-row list length */
    uint32_t sl_mininv[3];  /* SL: UINT32_MAX - min remaining degree */
    uint32_t sl_delta[3];   /* SL: threshold of the round */
    uint32_t sl_rounds;
    uint32_t sl_done;
    uint32_t total_items;   /* JP (ordered): work items in greedy order */
    uint32_t err;           /* JP (ordered): a wave hit its wait limit */
    uint32_t head_n;        /* JP (ordered): greedy-order positions the head kernel colours */
    uint32_t sl_chunks[3];  /* SL: removed-this-round long-row chunk list length */
    uint32_t sl_supers[3];  /* SL: removed-this-round super-chunk list length */
} gcxk_ctl;

/* ---- graph helpers (csr.hip) ---- */
/* Canonical CSR from pairs already on the device. Allocates *dE (hipMalloc,
 * caller frees) and fills dV (N+1 entries). Temporaries are freed inside. */
int gcxk_csr_from_pairs(uint32_t N, uint64_t L, const uint32_t *dI, const uint32_t *dJ,
                        uint32_t *dV, uint32_t **dE, uint64_t *M, void *stream);
/* Hub list: vertices with degree >= hub_deg in ascending id order into
 * dhub (capacity N), their 1024-edge chunk prefix into dchunk (capacity N+1,
 * dchunk[nhub] = total chunks); counts and the maximum degree to the host.
 * One host synchronisation. */
int gcxk_graph_stats(uint32_t N, const uint32_t *dV, uint32_t hub_deg, uint32_t *dhub,
                     uint32_t *dchunk, uint32_t *nhub, uint32_t *maxdeg, uint32_t *nchunk,
                     void *stream);
/* Degree order: perm = vertex ids by degree descending, id ascending (so the
 * nhub hubs come first). Stream-ordered temporaries; no synchronisation. */
/* rank[perm[i]] = i, Er[j] = rank[E[j]] (the columns as positions in the
 * degree order, so that rows stored in degree order put the hubs' rows
 * together: they take most neighbour reads on skewed graphs) and
 * Ed[j] = deg(E[j]) (the GNN input layer's degree feature, read in CSR order
 * instead of gathered). Er and Ed hold M + GCXK_ER_PAD entries, the padding
 * written as 0; E must also be allocated with GCXK_ER_PAD entries of slack
 * (the input layer reads whole 16-byte words). */
#define GCXK_ER_PAD 8u
int gcxk_rank_columns(uint32_t N, uint64_t M, const uint32_t *perm, const uint32_t *dV,
                      const uint32_t *dE, uint32_t *rank, uint32_t *Er, uint32_t *Ed,
                      void *stream);
int gcxk_degree_order(uint32_t N, const uint32_t *dV, uint32_t maxdeg, uint32_t *perm,
                      void *stream);
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
/* 2D random geometric graph: allocates *dI, *dJ (caller frees), *L pairs. */
int gcxk_gen_rgg(uint32_t n, uint32_t cells, uint64_t r2, uint64_t seed, uint32_t **dI,
                 uint32_t **dJ, uint64_t *L, void *stream);

/* ---- packed graphs (batch.hip) ---- */
/* Copy one graph (n vertices, m entries) into a pack at vertex offset off and
 * entry offset eoff (Vout rows, Eout columns shifted by off) and write its
 * vertices' graph span gspan[2(off + i)] = off, gspan[2(off + i) + 1] = n
 * (the GNN's id feature is graph-local: (float)((double)(v - off) / n)). */
int gcxk_pack_graph(uint32_t n, uint64_t m, const uint32_t *dV, const uint32_t *dE, uint32_t off,
                    uint32_t eoff, uint32_t *Vout, uint32_t *Eout, uint32_t *gspan,
                    void *stream);
/* Vout[N] = M. */
int gcxk_pack_close(uint32_t N, uint32_t M, uint32_t *Vout, void *stream);
/* Per packed graph (goff: count + 1 first vertices): largest colour into
 * maxc and the colouring digest over graph-local ids into digest (both
 * zeroed by the caller). The vertices come as nranges ranges
 * [rbeg[i], rend[i]) each inside graph rgid[i]; one workgroup per range. */
int gcxk_packed_stats(uint32_t nranges, const uint32_t *rbeg, const uint32_t *rend,
                      const uint32_t *rgid, const uint32_t *goff, const int *color, int *maxc,
                      uint64_t *digest, void *stream);

/* ---- ordering keys (keys.hip); key32: larger = earlier ---- */
int gcxk_keys_ff(uint32_t N, uint32_t *key, void *stream);
int gcxk_keys_lf(uint32_t N, const uint32_t *dV, uint32_t *key, void *stream);
int gcxk_keys_llf(uint32_t N, const uint32_t *dV, uint32_t *key, void *stream);
int gcxk_keys_from_colors(uint32_t N, const int *color, uint32_t *key, void *stream);
/* Culberson pass keys for strategy 0 reverse (key = colour), 1 largest class
 * first, 2 seeded random class order, 3 mixed (pass i: i mod 3). C = number
 * of colour classes (colours 1..C); pass = the pass counter (one device word,
 * advanced once per call on the device); scratch holds 2C + 2 words (class
 * sizes and class keys). */
int gcxk_culberson_keys(uint32_t N, uint32_t C, int strategy, uint64_t seed, const int *color,
                        uint32_t *pass, uint32_t *scratch, uint32_t *key, void *stream);
/* Host doubles (already on device) -> key32 when every value is exactly a
 * float, else key64; counts NaNs and non-floats in ctl. */
int gcxk_keys_scan_doubles(uint32_t N, const double *p, gcxk_ctl *ctl, void *stream);
int gcxk_keys_from_doubles32(uint32_t N, const double *p, uint32_t *key, void *stream);
int gcxk_keys_from_doubles64(uint32_t N, const double *p, uint64_t *key, void *stream);
/* key32 -> doubles: kind 0 = unsigned integer, 1 = float order key. */
int gcxk_keys_to_doubles(uint32_t N, const uint32_t *key, int kind, double *p, void *stream);
int gcxk_keys64_to_doubles(uint32_t N, const uint64_t *key, double *p, void *stream);

/* Streaming copy of `bytes` (multiple of 16) for byte-counter calibration. */
int gcxk_copy_bytes(uint64_t bytes, const void *src, void *dst, void *stream);

/* ---- SL / SLL peeling (sl.hip) ---- */
/* The peel's state is kept at degree-order ranks (perm / rank): degr and
 * keyr (N entries each) are the degrees and alive flags by rank (keyr[a] =
 * UINT32_MAX: alive, else the threshold at removal), the alive lists hold
 * ranks; key (by vertex id) receives each vertex's threshold (delta) at its
 * removal. ctl must be zeroed. */
int gcxk_sl_init(uint32_t N, const uint32_t *dV, const uint32_t *perm, uint32_t *degr,
                 uint32_t *keyr, uint32_t *alive0, void *rrow, gcxk_ctl *ctl, void *stream);
/* Round r, phase 1: remove alive vertices with degree <= threshold. rrow (N
 * pairs, from gcxk_sl_init): each rank's (first entry, degree). gone (N
 * pairs) receives (first entry, length) of the removed rows with 1 .. 16
 * entries; chunks (capacity gcxk_sl_chunk_capacity(M) pairs, M = CSR
 * entries) the CSR ranges [begin, end) of at most 64 entries that cover the
 * rows of 17 .. 4096 entries, followed by the ranges of at most 4096 entries
 * that cover the longer rows. */
int gcxk_sl_scan(const void *rrow, const uint32_t *degr, uint32_t *keyr, uint32_t *key,
                 const uint32_t *perm, const uint32_t *alive_in, uint32_t *alive_out, void *gone,
                 void *chunks, uint64_t M, gcxk_ctl *ctl, int log_variant, uint32_t r,
                 uint32_t grid, void *stream);
uint64_t gcxk_sl_chunk_capacity(uint64_t M);
/* Round r, phase 2: decrement unremoved neighbours of removed vertices,
 * reading neighbour ranks from Er = rank[E]. Decrements to the first 32768
 * ranks are summed per workgroup in LDS and applied once per rank. */
int gcxk_sl_dec(uint32_t N, const uint32_t *dV, const uint32_t *Er, uint32_t *degr,
                const uint32_t *keyr, const void *gone, const void *chunks, uint64_t M,
                gcxk_ctl *ctl, int log_variant, uint32_t r, void *stream);

/* ---- GraphSAGE-16 (gnn.hip) ---- */
/* FP32, bit-exact with the reference op order. kind 0 = input layer (reads
 * the graph only: neighbour ids E and their degrees Ed -> out), 1 = hidden
 * (x -> out), 2 = output (x -> key, float order keys).
 * Rows are 16 floats, stored by degree-order position (row of vertex v =
 * rank[v]; neighbours read through Er, gcxk_rank_columns). Hub rows run on
 * hub_stream, the rest on stream in the degree order perm (non-hubs are
 * perm[nhub .. N)); the caller joins the two streams before the next layer.
 * Priorities (key) stay indexed by vertex id. gspan (input layer only): 2N
 * values, (first vertex, vertex count) of each vertex's packed graph, or
 * NULL for a single graph (id feature (float)((double)v / N)). E and Ed must
 * carry GCXK_ER_PAD readable entries past M. */
int gcxk_gnn_fp32_layer(int kind, uint32_t N, const uint32_t *dV, const uint32_t *dE,
                        const uint32_t *Ed, const uint32_t *Er, const uint32_t *rank,
                        const float *W, const float *b, int md, const float *x, float *out,
                        uint32_t *key, const uint32_t *hubs, uint32_t nhub, const uint32_t *perm,
                        const uint32_t *gspan, void *stream, void *hub_stream);
/* MFMA layer transform with BF16 operands and FP32 accumulation.
 * split_mode 0: bf16 rows (16 x bf16 = 32 B), operands rounded to bf16.
 * split_mode 1: FP32 rows, operands and weights split into bf16 hi + lo,
 * three MFMAs per product (near-FP32 results).
 * Graph arrays and rows as in the FP32 layer. Wpack: the layer's two
 * MFMA B operands in lane order, hi parts (512 bf16) then lo parts (512);
 * b: FP32 biases (16). */
int gcxk_gnn_bf16_layer(int kind, int split_mode, uint32_t N, const uint32_t *dV,
                        const uint32_t *dE, const uint32_t *Ed, const uint32_t *Er,
                        const uint32_t *rank, const uint16_t *Wpack, const float *b, int md,
                        const void *x, void *out, uint32_t *key, const uint32_t *hubs,
                        uint32_t nhub, const uint32_t *perm, const uint32_t *gspan, void *stream,
                        void *hub_stream);

/* ---- Jones-Plassmann (jp.hip) ---- */
/* Work item: (vertex, chunk of 64 successor edges). Stored as two uint32. */
typedef struct gcxk_item {
    uint32_t u, k;
} gcxk_item;

/* Overflow colour-set words for N vertices / M entries (colours past the
 * state word's colour bits, which are at least 32). */
static inline uint64_t gcxk_jp_ovf_words(uint32_t N, uint64_t M)
{
    return M / 32u + (uint64_t)N + 2u;
}
/* State words: the round schedule uses one per vertex; the ordered schedule
 * gives vertex v W(v) = min(8, ceil((npred(v)+1)/mb)) <= 1 + npred(v)/32, and
 * the predecessor counts sum to M/2. */
static inline uint64_t gcxk_jp_state_words(uint32_t N, uint64_t M)
{
    /* counting fields: ceil((np + 2) / 8) <= np / 8 + 9 / 8 words per vertex,
     * bit sets fewer; the predecessor counts sum to at most M (M / 2 for a
     * symmetric CSR) */
    return (uint64_t)N + N / 8u + M / 8u + 64u;
}
/* Item queue capacity: every vertex contributes max(1, ceil(nsucc / 64)). */
static inline uint64_t gcxk_jp_item_capacity(uint32_t N, uint64_t M)
{
    return (uint64_t)N + M / 64u + 2u;
}
/* Hub chunks for the orientation pass: 1024 edges each (R-MAT 22 orientation stage: 4096 -> 1.75 ms, 2048 -> 1.68, 1024 -> 1.64, 512 -> 1.66). */
#define GCXK_HUB_CHUNK_EDGES 1024u

/* Colour bits in the per-vertex JP state word: 64 - bitlength(maxdeg); the
 * rest of the word counts pending predecessors. */
int gcxk_jp_mask_bits(uint32_t maxdeg);
/* Zero the overflow colour words (gcxk_jp_ovf_words), the per-hub positions
 * (hpos: 2 per hub), the hubs' colour / npred / state and the control block.
 * A plain kernel, so the sequence can be captured in a graph. */
int gcxk_jp_clear(uint32_t N, uint64_t M, uint32_t *ovf, uint32_t nhub, uint32_t *hpos,
                  const uint32_t *hubs, uint32_t *npred, uint64_t *state, int *color,
                  gcxk_ctl *ctl, void *stream);
/* Orientation: key_bits = 32 or 64, mb = gcxk_jp_mask_bits(maxdeg), perm =
 * the graph's degree order (hubs first). E_split[V[u] + npred[u] ..
 * V[u+1]) = u's successors (the predecessor part is left unwritten: nothing
 * reads it); npred = predecessor counts; state = npred << mb; color = 1 for
 * degree-0 vertices, else 0. hchunk: per-hub prefix of 1024-edge chunks
 * (nhub + 1 entries), nhchunk = total chunks. The hub chunks run on
 * hub_stream (may equal stream), after gcxk_jp_clear; the caller joins. */
int gcxk_jp_init(uint32_t N, const uint32_t *dV, const uint32_t *dE, const void *key,
                 int key_bits, int mb, const uint32_t *perm, uint32_t *npred, uint32_t *esplit,
                 uint64_t *state, int *color, const uint32_t *hubs, uint32_t nhub,
                 const uint32_t *hchunk, uint32_t nhchunk, uint32_t *hpos, uint32_t head_max,
                 void *stream, void *hub_stream);
/* head_max > 0 (with 32-bit rank keys from gcxk_jp_rank_keys only): every
 * vertex whose greedy-order position N - 1 - key is below head_max also
 * gets its predecessors' positions in E_split[V[u] .. V[u] + npred[u]), for
 * the head kernel. */
/* Greedy order alone: ids (inside scratch, sized by gcxk_jp_order_scratch)
 * = the vertices sorted by key descending, id ascending. */
int gcxk_jp_sort_ids(uint32_t N, const void *key, int key_bits, void *scratch,
                     size_t scratch_bytes, uint32_t **ids, void *stream);
/* rk[ids[i]] = N - 1 - i: 32-bit keys of the same order with no ties. */
int gcxk_jp_rank_keys(uint32_t N, const uint32_t *ids, uint32_t *rk, void *stream);
/* Longest prefix the head kernel may colour (positions). */
uint32_t gcxk_jp_head_max(void);
/* Ordered schedule after gcxk_jp_sort_ids + rank keys + orientation with
 * head_max: ctl->head_n = the dense head of the order (leading 1024-position
 * blocks whose mean predecessor count is at least head_density, at most
 * head_max positions; 0 = no head), then state words, descriptors (head
 * vertices flagged) and items as gcxk_jp_order_items. key_bits: the keys
 * the ids were sorted by (sizes the scratch layout). */
int gcxk_jp_order_items_ranked(uint32_t N, const uint32_t *dV, const uint32_t *npred,
                               const uint32_t *ids, const uint32_t *rk, uint32_t head_max,
                               uint32_t head_density, int key_bits, int mb, void *scratch,
                               size_t scratch_bytes, uint32_t *cnt, uint32_t *incl, uint32_t *wof,
                               uint64_t *wdesc, uint64_t *words, uint64_t words_cap,
                               gcxk_item *items, gcxk_ctl *ctl, void *stream);
/* Head kernel: one workgroup colours greedy-order positions 0 .. head_n
 * (LDS colours); run before gcxk_jp_ordered, which takes those colours as
 * given. hdr: 8 bytes of scratch per head position (gcxk_jp_head_max() at
 * most). A wait past spin_ticks sets ctl->err (the ordered launch then does
 * nothing and the caller redoes the colouring). */
int gcxk_jp_head(const uint32_t *dV, const uint32_t *npred, const uint32_t *esplit,
                 const uint32_t *ids, void *hdr, int *color, gcxk_ctl *ctl, uint64_t spin_ticks,
                 void *stream);
/* Round schedule: queue the roots (vertices with successors and no
 * predecessors) as round-0 items in q0 / ctl->qn[0]. */
int gcxk_jp_roots(uint32_t N, const uint32_t *dV, const uint32_t *npred, gcxk_item *q0,
                  gcxk_ctl *ctl, void *stream);
/* One JP round r: colours the items of slot r, queues the next round's. */
int gcxk_jp_round(const uint32_t *dV, const uint32_t *npred, const uint32_t *esplit,
                  uint64_t *state, int *color, uint32_t *ovf, int mb, gcxk_item *const q[2],
                  gcxk_ctl *ctl, uint32_t r, uint32_t grid, void *stream);
/* Ordered schedule, step 1 (after gcxk_jp_init): the state word offsets
 * wof (N + 1 entries; vertex v owns words[wof[v] .. wof[v+1]), zeroed,
 * words_cap = gcxk_jp_state_words), the per-vertex word descriptors wdesc
 * (N entries) and the items of every vertex in greedy
 * order (p descending, id ascending) into `items`, their number into
 * ctl->total_items. scratch: gcxk_jp_order_scratch(N, key_bits) bytes; cnt:
 * N words, incl: N + 1 words. */
size_t gcxk_jp_order_scratch(uint32_t N, int key_bits);
int gcxk_jp_order_items(uint32_t N, const uint32_t *dV, const uint32_t *npred, const void *key,
                        int key_bits, int mb, void *scratch, size_t scratch_bytes, uint32_t *cnt,
                        uint32_t *incl, uint32_t *wof, uint64_t *wdesc, uint64_t *words,
                        uint64_t words_cap, gcxk_item *items, gcxk_ctl *ctl, void *stream);
/* Default workgroups of the ordered kernel: 6 per CU, at most what is
 * resident at once (0 on error). */
uint32_t gcxk_jp_ordered_grid(void);
/* Ordered schedule, step 2: one launch of `grid` workgroups (0 =
 * gcxk_jp_ordered_grid(); larger values are capped at the resident maximum);
 * wave w colours items w, w + W, ... waiting on each item's state word.
 * spin_ticks bounds every wait (100 MHz ticks); exceeding it sets ctl->err
 * and ends the kernel. cooperative = 1 launches cooperatively (the runtime
 * checks residency and runs it alone); 0 launches it as a plain kernel that
 * shares the device with other streams (the caller keeps the grids of
 * concurrent ordered launches within the resident maximum). */
int gcxk_jp_ordered(const uint32_t *dV, const uint32_t *npred, const uint32_t *esplit,
                    uint64_t *words, const uint64_t *wdesc, int *color, uint32_t *ovf, int mb,
                    const gcxk_item *items, gcxk_ctl *ctl, uint32_t grid, uint64_t spin_ticks,
                    int cooperative, void *stream);
/* Workgroups of the ordered kernel resident on the device at once (0 on
 * error). */
uint32_t gcxk_jp_ordered_cap(void);

/* ---- ordered schedule with pulled colours (jp_pull.hip) ----
 * Step 1, after gcxk_jp_sort_ids: zero ctl and hcnt[0 .. nhub), then
 * rank[ids[j]] = j (greedy-order positions) and pcol[j] = 0. */
int gcxk_pull_rank(uint32_t N, const uint32_t *ids, uint32_t *rank, uint32_t *pcol, uint32_t nhub,
                   uint32_t *hcnt, gcxk_ctl *ctl, void *stream);
/* Step 2: every row's predecessor positions to its front in esplit, hdr[j] =
 * (V[ids[j]], predecessor count) (8 bytes per position). Hub
 * rows (chunk table hchunk) run on hub_stream (may equal stream; the caller
 * forks it after step 1 and joins it before step 3). */
int gcxk_pull_orient(uint32_t N, const uint32_t *dV, const uint32_t *dE, const uint32_t *rank,
                     const uint32_t *perm, const uint32_t *hubs, uint32_t nhub,
                     const uint32_t *hchunk, uint32_t nhchunk, uint32_t *hcnt, uint32_t *esplit,
                     void *hdr, void *stream, void *hub_stream);
/* Default workgroups of the pull kernel (6 per CU, at most the resident
 * maximum; 0 on error) and the resident maximum. */
uint32_t gcxk_pull_grid(void);
uint32_t gcxk_pull_cap(void);
/* Step 3: one persistent launch colouring positions in greedy order (color[]
 * by vertex id, pcol by position). grid, spin_ticks and cooperative as for
 * gcxk_jp_ordered; a wait past spin_ticks sets ctl->err. */
int gcxk_pull_colour(uint32_t N, const void *hdr, const uint32_t *ids, const uint32_t *esplit,
                     uint32_t *pcol, int *color, gcxk_ctl *ctl, uint32_t grid, uint64_t spin_ticks,
                     int cooperative, void *stream);
/* Degrees of the listed vertices (for the hub chunk table). */
int gcxk_gather_degrees(const uint32_t *dV, const uint32_t *list, uint32_t n, uint32_t *out,
                        void *stream);
/* Proper-colouring check: ctl->bad, ctl->maxcolor (ctl fields zeroed by
 * caller). Hub rows run on hub_stream (may equal stream); the caller joins. */
int gcxk_validate(uint32_t N, const uint32_t *dV, const uint32_t *dE, const int *color,
                  const uint32_t *hubs, uint32_t nhub, const uint32_t *perm, gcxk_ctl *ctl,
                  void *stream, void *hub_stream);
/* Greedy-rule certificate: ctl->viol. */
int gcxk_mex_check(uint32_t N, const uint32_t *dV, const uint32_t *dE, const void *key,
                   int key_bits, const int *color, const uint32_t *hubs, uint32_t nhub,
                   uint32_t hub_deg, gcxk_ctl *ctl, void *stream);

/* Order-independent digest of color[0..N): adds sum_u mix(u, color[u]) to
 * *out (device, zeroed by the caller). */
int gcxk_color_digest(uint32_t N, const int *color, uint64_t *out, void *stream);
/* Number of workgroups the round kernel should use (fills the GPU). */
uint32_t gcxk_round_grid(void);

/* The no-allocation entry point's device scratch (noalloc.hip): looks up its
 * address on the current device; called by the library's initialisation. */
int gcxk_noalloc_init(void);

#ifdef __cplusplus
}
#endif
