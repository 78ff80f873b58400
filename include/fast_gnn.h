/*
 * fast_gnn.h -- GraphSAGE-16 priorities (drop-in for the reference's
 * include/fast_gnn.h:13-19). Inputs are degree/md and id/N features; the
 * output p[u] is the sum of the last layer's 16 features. Runs on the GPU
 * with the reference's FP32 operation order, so p is bit-identical.
 * x and y are accepted for signature compatibility; the GPU keeps its rows
 * in HBM and does not write them. p needs N entries (the reference asks for
 * N + 8 and writes junk past N; this library writes exactly N).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces fast_gnn.c:252-295. */
void cgnn_graph_sage_16(int params, const float *param,
                        uint32_t N, const uint32_t *V, const uint32_t *E, int md,
                        float *x, float *y, double *p);

/* Replaces fast_gnn.c:297-347. */
void cgnn_graph_sage_16_par(int params, const float *param,
                            uint32_t N, const uint32_t *V, const uint32_t *E, int md,
                            float *x, float *y, double *p);

#ifdef __cplusplus
}
#endif
