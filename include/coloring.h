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

#ifdef __cplusplus
}
#endif
