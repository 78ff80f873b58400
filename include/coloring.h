/*
 * coloring.h -- colouring engine entry points (drop-in for the reference's
 * include/coloring.h:6-29). Every call runs on the GPU: the host arrays are
 * copied to HBM, Jones-Plassmann colours the graph in the order
 * (p descending, vertex id ascending), and the colours come back.
 * The sequential names (greedy_*) produce the same colours as the parallel
 * ones, as in the reference (coloring.c:60-125 vs 223-330).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces coloring.c:7-16. Returns a malloc'ed int[N] (caller frees). */
int *greedy_coloring(uint32_t N, const uint32_t *V, const uint32_t *E, const double *p);
/* Replaces coloring.c:26-41: opaque handle sized for N vertices, M entries. */
void *greedy_setup(uint32_t N, uint32_t M);
/* Replaces coloring.c:43-58. */
void greedy_cleanup(void *gd);
/* Replaces coloring.c:60-125: colours 1..C into color[0..N). */
void greedy_internal(uint32_t N, const uint32_t *V, const uint32_t *E,
                     void *gd, const double *p, int *color);
/* Replaces coloring.c:127-146: largest colour if proper (all > 0), else 0. */
int validate_coloring(uint32_t N, const uint32_t *V, const uint32_t *E, const int *color);

/* Replaces coloring.c:148-157. */
int *jones_plassmann_coloring(uint32_t N, const uint32_t *V, const uint32_t *E, const double *p);
/* Replaces coloring.c:168-197 (N vertices, M = V[N] entries). */
void *jones_plassmann_setup(uint32_t N, uint32_t M);
/* Replaces coloring.c:199-221. */
void jones_plassmann_cleanup(void *jp);
/* Replaces coloring.c:223-330. */
void jones_plassmann_internal(uint32_t N, const uint32_t *V, const uint32_t *E,
                              void *jp, const double *p, int *color);
/* Replaces coloring.c:332-351. */
int validate_coloring_par(uint32_t N, const uint32_t *V, const uint32_t *E, const int *color);

/* No counterpart in the reference. For callers that cannot allocate (code
 * that must not call hipMalloc, such as tight serving loops or harnesses that
 * forbid allocation): the colours of jones_plassmann_internal, with no
 * setup handle and no allocation of any kind. Every device buffer of the
 * call lives in one 128 MiB device array that the library holds for the
 * process's lifetime. Limits: N <= GCX_NOALLOC_MAX_N vertices and
 * V[N] <= GCX_NOALLOC_MAX_M entries; larger inputs return GCX_ERR_CAPACITY
 * (-4) with a message on stderr. Returns 0 (GCX_OK) or a negative gcx.h
 * status: GCX_ERR_ARG for out-of-range indices or a NaN priority.
 * Synchronous; calls are serialised inside the process (they share the
 * array). */
#define GCX_NOALLOC_SCRATCH_MB 128
#define GCX_NOALLOC_MAX_N (1u << 18)
#define GCX_NOALLOC_MAX_M (1u << 23)
int jones_plassmann_internal_noalloc(uint32_t N, const uint32_t *V, const uint32_t *E,
                                     const double *p, int *color);

#ifdef __cplusplus
}
#endif
