/*
 * first_fit.h -- natural-order heuristic (drop-in for the reference's
 * include/first_fit.h:4-13): p[i] = N - i.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces first_fit.c:8-12. */
void first_fit_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, double *p);
/* Replaces first_fit.c:14-19. */
void first_fit_ordering_par(uint32_t N, const uint32_t *V, const uint32_t *E, double *p);
/* Replace first_fit.c:21-37: malloc'ed int[N] colours. */
int *first_fit_coloring(uint32_t N, const uint32_t *V, const uint32_t *E);
int *first_fit_coloring_par(uint32_t N, const uint32_t *V, const uint32_t *E);
/* Replaces first_fit.c:89-98. */
void first_fit_runner(uint32_t N, const uint32_t *V, const uint32_t *E,
                      int it, int k, int t_tot, int ntc, const int *nt);

#ifdef __cplusplus
}
#endif
