/*
 * oracle.h -- CPU restatement of the GNN-COLORING ordering + colouring path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP path is
 * compared against. Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it; the product library never links or calls it.
 *
 * Every function cites the reference file:line whose semantics it restates
 * (paths relative to the KennethLangedal/GNN-COLORING tree).
 *
 * Pinning (how each function is checked, see tests/test_oracle_*.py):
 *   orc_greedy / orc_validate / orc_mex_violations : golden colours from the
 *       compiled reference, brute-force sequential greedy on tiny graphs,
 *       hand-derived known answers (path, star, clique, odd cycle, bipartite).
 *   orc_key_* : closed forms + reference goldens.
 *   orc_sl / orc_sll : hand-derived known answers (P5, star, clique, delta
 *       jump) + reference goldens.
 *   orc_gnn_fp32 : bit-exact against the compiled reference's
 *       cgnn_graph_sage_16_par on the golden graphs.
 *   orc_gnn_fp64 : parity unpinned against the reference (the reference has
 *       no FP64 path); checked against orc_gnn_fp32 within FP32 rounding and
 *       against a hand-computed single-vertex case.
 *   orc_csr_from_pairs : scipy COO->CSR canonical form + reference mtx goldens.
 *   orc_load_model : the reference's own loader (fscanf "%f") on models/.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* coloring.c:127-146 */
int orc_validate(uint32_t N, const uint32_t *V, const uint32_t *E, const int *color);

/* coloring.c:60-125 (greedy_internal); *rounds receives the DAG depth. */
void orc_greedy(uint32_t N, const uint32_t *V, const uint32_t *E, const double *p,
                int *color, uint32_t *rounds);

/* Number of vertices u with color[u] != mex{color[v] : v pred of u}. */
uint64_t orc_mex_violations(uint32_t N, const uint32_t *V, const uint32_t *E,
                            const double *p, const int *color);

/* first_fit.c:8-12 */
void orc_key_ff(uint32_t N, double *p);
/* largest_degree_first.c:12-16 */
void orc_key_lf(uint32_t N, const uint32_t *V, double *p);
/* largest_degree_first.c:87-97 */
void orc_key_llf(uint32_t N, const uint32_t *V, double *p);
/* smallest_degree_last.c:62-124; *rounds receives the peel round count. */
void orc_sl(uint32_t N, const uint32_t *V, const uint32_t *E, double *p, uint32_t *rounds);
/* smallest_degree_last.c:274-328 */
void orc_sll(uint32_t N, const uint32_t *V, const uint32_t *E, double *p, uint32_t *rounds);

/* gnn.c:20-44: returns 0 on success, -1 if the file cannot be opened, -2 if it
 * is truncated or malformed. *param is malloc'ed (caller frees). */
int orc_load_model(const char *path, int *params, int *md, float **param);

/* fast_gnn.c:252-295 (op order of message_kernel_input / message_kernel /
 * dense_kernel / dense_kernel_output). p must hold N doubles. */
void orc_gnn_fp32(int params, const float *param, uint32_t N, const uint32_t *V,
                  const uint32_t *E, int md, double *p);
/* Same mathematics evaluated in double precision. */
void orc_gnn_fp64(int params, const float *param, uint32_t N, const uint32_t *V,
                  const uint32_t *E, int md, double *p);

/* fast_io.c:171-202 + 450-455 + 457-477: 0-based pairs -> canonical CSR
 * (symmetric, rows sorted, no self loops, no duplicates). V has N+1 entries;
 * E is malloc'ed with V[N] entries (caller frees both). */
void orc_csr_from_pairs(uint32_t N, uint64_t L, const uint32_t *I, const uint32_t *J,
                        uint32_t **V, uint32_t **E);

/* fast_io.c:509-540: 1 if sorted, simple and symmetric, else 0. */
uint32_t orc_graph_validate(uint32_t N, const uint32_t *V, const uint32_t *E);

/* main.c:32-45: JP, then K-1 times p[j] = color[j] and JP again.
 * colors_per_iter (K entries) receives the colour count after each pass. */
void orc_culberson(uint32_t N, const uint32_t *V, const uint32_t *E, int K,
                   double *p, int *color, int *colors_per_iter);

void orc_free(void *ptr);

#ifdef __cplusplus
}
#endif
