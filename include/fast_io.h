/*
 * fast_io.h -- Matrix Market I/O and CSR utilities (drop-in for the
 * reference's include/fast_io.h:35-39 and 128-148).
 *
 * mtx_parse / mtx_parse_par read "M N L" and L 1-based index pairs (values
 * and '%' comment lines ignored, N = max(M, N), fast_io.c:68-141) and build
 * the CSR on the GPU. The result is already canonical (rows sorted, no self
 * loops, no repeats), which is what the reference reaches after
 * graph_sort_edges + graph_make_simple; those two calls are then no-ops.
 * The adjacency, .gr and METIS formats (fast_io.h:41-122) are read on the
 * host; .gr also comes back canonical (the reference leaves both directions
 * of every edge, repeats included, in input order).
 */
#pragma once
#include <stdio.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replace fast_io.c:204-222. *V (N+1) and *E (V[N]) are malloc'ed. */
void mtx_parse(FILE *f, uint32_t *N, uint32_t **V, uint32_t **E);
void mtx_parse_par(FILE *f, uint32_t *N, uint32_t **V, uint32_t **E);
/* Replaces fast_io.c:224-237: writes the u < v entries, 1-based. */
void mtx_store(FILE *f, uint32_t N, const uint32_t *V, const uint32_t *E);

/* Replace fast_io.c:239-275 (fast_io.h:69-71): "AdjacencyGraph", N, M, the
 * N row offsets and M targets, 0-based; returned as written. */
void adj_parse(FILE *f, uint32_t *N, uint32_t **V, uint32_t **E);
void adj_store(FILE *f, uint32_t N, const uint32_t *V, const uint32_t *E);
/* Replace fast_io.c:283-368 (fast_io.h:99-101): PACE .gr ("c" comments,
 * "p td N M", M 1-based edges). gr_parse returns the canonical CSR. */
void gr_parse(FILE *f, uint32_t *N, uint32_t **V, uint32_t **E);
void gr_store(FILE *f, uint32_t N, const uint32_t *V, const uint32_t *E);
/* Replace fast_io.c:370-437 (fast_io.h:120-122): METIS .graph ("%" comments,
 * "N M fmt", one line of 1-based neighbours per vertex; fmt 10/11 put a
 * vertex weight first, returned in *W, else *W = NULL; fmt 1/11 give every
 * neighbour an edge weight, dropped). Returned as written. */
void metis_parse(FILE *f, uint32_t *N, uint32_t **V, uint32_t **E, uint32_t **W);
void metis_store(FILE *f, uint32_t N, const uint32_t *V, const uint32_t *E, const uint32_t *W);

/* Replace fast_io.c:444-455: sort every row ascending (repeats kept). */
void graph_sort_edges(uint32_t N, const uint32_t *V, uint32_t *E);
void graph_sort_edges_par(uint32_t N, const uint32_t *V, uint32_t *E);
/* Replaces fast_io.c:457-477: drop self loops and repeats from sorted rows,
 * shift V, shrink *E with realloc. */
void graph_make_simple(uint32_t N, uint32_t *V, uint32_t **E);
/* Replace fast_io.c:479-540: 1 if sorted, simple and symmetric, else 0. */
uint32_t graph_validate(uint32_t N, const uint32_t *V, const uint32_t *E);
uint32_t graph_validate_par(uint32_t N, const uint32_t *V, const uint32_t *E);

#ifdef __cplusplus
}
#endif
