/*
 * bf16.c -- BF16 GraphSAGE ordering (rows stored as bf16, FP32 neighbour
 * aggregation, layer transforms on MFMA with FP32 accumulation).
 */
#include "internal.h"

int gcx__order_gnn_bf16(gcx_workspace *ws, const gcx_graph *g, const gcx_model *m, void *stream)
{
    (void)ws;
    (void)g;
    (void)m;
    (void)stream;
    return GCX_ERR_ARG;
}
