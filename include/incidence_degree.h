/*
 * incidence_degree.h -- incidence-degree (ID) and saturation-degree (SD, SDA)
 * orderings (drop-in for the reference's include/incidence_degree.h:4-8).
 *
 * Each writes p[0..N) = N, N-1, ... in the order a sequential heap walk
 * picks the vertices (SDA gives one value to each batch of equal keys), so
 * the colouring that follows is greedy in that order. The walk has no
 * parallel form with the same output, so these three run on the host CPU;
 * the colouring of the priorities they produce runs on the GPU. Memory:
 * about 29 bytes per vertex, plus 8 bytes per CSR entry for SD and SDA.
 * A failed allocation prints a message and exits with status 1.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces incidence_degree.c:107-170: the next vertex has the most ordered
 * neighbours, then the largest degree, then the largest id. */
void incidence_degree_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, double *p);

/* Replaces incidence_degree.c:172-257: DSATUR order; the next vertex has the
 * most distinct colours among its coloured neighbours, then the largest
 * degree, then the largest id. */
void saturation_degree_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, double *p);

/* Replaces incidence_degree.c:259-360: SD started from every vertex of
 * maximum degree in its component, picking all heap-top vertices with equal
 * (saturation, degree) at once; they share one priority. */
void saturation_degree_ordering_alt(uint32_t N, const uint32_t *V, const uint32_t *E, double *p);

#ifdef __cplusplus
}
#endif
