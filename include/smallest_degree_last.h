/*
 * smallest_degree_last.h -- SL and SLL peeling heuristics (drop-in for the
 * reference's include/smallest_degree_last.h:4-20). p[u] is the peel
 * threshold at which u was removed.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces smallest_degree_last.c:126-131 (rounds of 62-124). */
void smallest_degree_last_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, double *p);
/* Replace smallest_degree_last.c:140-153 and 248-261. */
int *smallest_degree_last_coloring(uint32_t N, const uint32_t *V, const uint32_t *E);
int *smallest_degree_last_coloring_par(uint32_t N, const uint32_t *V, const uint32_t *E);
/* Replaces smallest_degree_last.c:263-272. */
void smallest_degree_last_runner(uint32_t N, const uint32_t *V, const uint32_t *E,
                                 int it, int k, int t_tot, int ntc, const int *nt);

/* Replaces smallest_degree_last.c:330-335 (rounds of 274-328). */
void smallest_log_degree_last_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, double *p);
/* Replace smallest_degree_last.c:344-357 and 445-458. */
int *smallest_log_degree_last_coloring(uint32_t N, const uint32_t *V, const uint32_t *E);
int *smallest_log_degree_last_coloring_par(uint32_t N, const uint32_t *V, const uint32_t *E);
/* Replaces smallest_degree_last.c:460-469. */
void smallest_log_degree_last_runner(uint32_t N, const uint32_t *V, const uint32_t *E,
                                     int it, int k, int t_tot, int ntc, const int *nt);

#ifdef __cplusplus
}
#endif
