/*
 * bf16.c -- BF16 GraphSAGE ordering: weight packing for the MFMA operand
 * layout and the layer sequence (kernels in gnn_bf16.hip).
 */
#include <hip/hip_runtime_api.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

#define LAYER_PARAMS 528

/* Round to nearest even (Inf/NaN kept). */
static uint16_t to_bf16(float f)
{
    uint32_t b;
    memcpy(&b, &f, 4);
    if ((b & 0x7F800000u) == 0x7F800000u)
        return (uint16_t)((b >> 16) | ((b & 0xFFFFu) ? 0x40u : 0u));
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}

static float from_bf16(uint16_t h)
{
    const uint32_t b = (uint32_t)h << 16;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

/* Per layer (1024 values each): Wb = W[k][j] in bf16, k-major, hi parts
 * (512) then lo parts (512, bf16(W - hi)); and the two MFMA B operands in
 * lane order, hi then lo: pack[p*512 + m*256 + 4*lane + e] = part p of
 * W[16m + 4*(lane/16) + e][lane%16] (v_mfma_f32_16x16x16_bf16: B[k][j] lives
 * in lane j + 16*(k/4), element k%4). */
int gcx__model_pack_bf16(gcx_model *m)
{
    const int L = m->layers;
    const size_t per = 1024;
    uint16_t *wb = malloc(sizeof(uint16_t) * per * (size_t)L);
    uint16_t *pk = malloc(sizeof(uint16_t) * per * (size_t)L);
    if (!wb || !pk) {
        free(wb);
        free(pk);
        return GCX_ERR_NOMEM;
    }
    for (int l = 0; l < L; l++) {
        const float *W = m->hparam + (size_t)l * LAYER_PARAMS;
        uint16_t *hi = wb + per * l, *lo = hi + 512;
        for (int k = 0; k < 512; k++) {
            hi[k] = to_bf16(W[k]);
            lo[k] = to_bf16(W[k] - from_bf16(hi[k]));
        }
        for (int p = 0; p < 2; p++)
            for (int mm = 0; mm < 2; mm++)
                for (int lane = 0; lane < 64; lane++)
                    for (int e = 0; e < 4; e++) {
                        const int k = 16 * mm + 4 * (lane / 16) + e, j = lane % 16;
                        pk[per * l + 512 * p + 256 * mm + 4 * lane + e] = hi[512 * p + 16 * k + j];
                    }
    }
    int rc = GCX_OK;
    const size_t bytes = sizeof(uint16_t) * per * (size_t)L;
    if (hipMalloc((void **)&m->dWb, bytes) != hipSuccess ||
        hipMalloc((void **)&m->dWpack, bytes) != hipSuccess ||
        hipMemcpy(m->dWb, wb, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->dWpack, pk, bytes, hipMemcpyHostToDevice) != hipSuccess)
        rc = GCX_ERR_HIP;
    free(wb);
    free(pk);
    return rc;
}

int gcx__order_gnn_bf16(gcx_workspace *ws, const gcx_graph *g, const gcx_model *m, int split,
                        void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    const int L = m->layers;
    void *bufs[2] = {ws->x, ws->y};
    const void *in = NULL; /* layer 0 reads only the graph */
    for (int l = 0; l < L; l++) {
        const int kind = l == 0 ? 0 : (l == L - 1 ? 2 : 1);
        const float *b = m->dparam + (size_t)l * LAYER_PARAMS + 512;
        void *out = bufs[l & 1];
        void *tok;
        int rc = gcx__prof_begin(kind == 0 ? GCX_ST_GNN_IN : kind == 1 ? GCX_ST_GNN_HID : GCX_ST_GNN_OUT,
                                 s, &tok);
        if (rc != GCX_OK)
            return rc;
        void *hs = gcx__fork_hubs(ws, g, 1, s);
        if (!hs)
            return GCX_ERR_HIP;
        rc = gcxk_gnn_bf16_layer(kind, split, g->N, g->V, g->E, g->Ed, g->Er, g->rank,
                                 m->dWpack + 1024 * (size_t)l, b, m->md, in, out,
                                 (uint32_t *)ws->key, g->hubs, g->nhub, g->perm, g->gspan, s, hs);
        if ((rc = gcx__join_hubs(ws, hs, s, rc)) != GCX_OK)
            return rc;
        if ((rc = gcx__prof_end(tok, s)) != GCX_OK)
            return rc;
        in = out;
    }
    return GCX_OK;
}
