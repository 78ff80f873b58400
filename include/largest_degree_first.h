/*
 * largest_degree_first.h -- LF (p = degree) and LLF (p = bit length of the
 * degree) heuristics; drop-in for the reference's
 * include/largest_degree_first.h:4-20.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces largest_degree_first.c:12-16. */
void largest_degree_first_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, double *p);
/* Replace largest_degree_first.c:36-47 and 59-70. */
int *largest_degree_first_coloring(uint32_t N, const uint32_t *V, const uint32_t *E);
int *largest_degree_first_coloring_par(uint32_t N, const uint32_t *V, const uint32_t *E);
/* Replaces largest_degree_first.c:72-81. */
void largest_degree_first_runner(uint32_t N, const uint32_t *V, const uint32_t *E,
                                 int it, int k, int t_tot, int ntc, const int *nt);

/* Replaces largest_degree_first.c:87-97. */
void largest_log_degree_first_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, double *p);
/* Replace largest_degree_first.c:106-117 and 140-151. */
int *largest_log_degree_first_coloring(uint32_t N, const uint32_t *V, const uint32_t *E);
int *largest_log_degree_first_coloring_par(uint32_t N, const uint32_t *V, const uint32_t *E);
/* Replaces largest_degree_first.c:153-162. */
void largest_log_degree_first_runner(uint32_t N, const uint32_t *V, const uint32_t *E,
                                     int it, int k, int t_tot, int ntc, const int *nt);

#ifdef __cplusplus
}
#endif
