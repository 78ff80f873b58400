/*
 * gnn.h -- GNN ordering glue (drop-in for the reference's include/gnn.h:4-7).
 * Weights are read from "<dir>/m<l>.gnn" where <dir> is $GCX_MODEL_DIR when
 * set and "models" (relative to the working directory, as in gnn.c:25)
 * otherwise. A missing or malformed model prints a message and exits with
 * status 1, as the reference does (gnn.c:27-32).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces gnn.c:85-91: p[0..N) = GraphSAGE priorities of model m<l>.gnn. */
void gnn_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, double *p, int l);

/* Replaces gnn.c:93-104: timed GNN<l> + colouring runner. */
void gnn_runner(uint32_t N, const uint32_t *V, const uint32_t *E, int l,
                int it, int k, int t_tot, int ntc, const int *nt);

#ifdef __cplusplus
}
#endif
