// gnn.hip -- GraphSAGE-16 inference (mean aggregation, concat self, ReLU),
// FP32, reproducing the reference's operation order bit for bit
// (fast_gnn.c:13-243):
//   layer 0  A = [mean fdeg(v) x8, mean fid(v) x8, fdeg(u) x8, fid(u) x8]
//            fdeg(v) = (float)deg(v) / (float)md, fid(v) = (float)((double)v / N)
//   hidden   A = [mean over neighbours of x[v][0..15], x[u][0..15]]
//   dense    c_j = b_j; c_j = fmaf(W[i][j], A[i], c_j) for i = 0..31; h = c > 0 ? c : +0
//   output   t_k = h_k + h_{k+8}; p = ((t0+t1)+(t2+t3)) + ((t4+t5)+(t6+t7))
// Neighbour sums run left to right in CSR order from +0 and are divided by
// the degree when it is positive.
//
// Rows live at degree-order positions (row of v = rank[v], neighbours read
// through the graph's Er = rank[E]): the hubs' rows, which take most
// neighbour reads on skewed graphs, sit together in a few cache lines.
//
// Mapping: vertices below the hub degree run in degree order (the graph's
// perm), 4 lanes per vertex and 16 vertices per wave (see sage_quads); the
// sequential dependency is only on the ordered adds, the row loads are
// issued ahead. Hubs (degree >= the graph's hub degree) get a 256-thread
// workgroup that stages 256 neighbour rows per step into LDS (feature-major,
// double-buffered, the next step's loads in flight during the sum) while one
// 16-lane group performs the ordered sum (hub_sum.hip.h); the hub kernel runs
// on a second stream next to the group kernel.
//
// This file is compiled with -ffp-contract=off: the only fused operations
// are the explicit fmaf calls of the dense step.
#include <type_traits>

#include "common.hip.h"
#include "hub_sum.hip.h"
#include "input_sum.hip.h"

namespace {

constexpr int kBlock = 256;
constexpr int F = 16;
// Threads (= rows per staged chunk) of the hub kernel. Measured on R-MAT
// scale 22 (GNN3 ordering): 256 -> 4.43 ms, 512 -> 4.87 ms, 1024 -> 6.08 ms;
// more, smaller workgroups keep more of the hubs' add chains running.
constexpr int kHubBlock = 256;

enum LayerKind { kInput = 0, kHidden = 1, kOutput = 2 };

__device__ __forceinline__ float relu(float c) { return c > 0.0f ? c : 0.0f; }

// Dense step for lane j of a 16-lane group; A[i] for i < 16 is a_lo of lane
// i, for i >= 16 a_hi of lane i - 16.
__device__ __forceinline__ float dense_group(const float (&w)[2 * F], float bias, float a_lo,
                                             float a_hi)
{
    float c = bias;
#pragma unroll
    for (int i = 0; i < F; i++)
        c = fmaf(w[i], __shfl(a_lo, i, F), c);
#pragma unroll
    for (int i = 0; i < F; i++)
        c = fmaf(w[F + i], __shfl(a_hi, i, F), c);
    return relu(c);
}

// Layer-0 dense step: A = [sd x8, si x8, selfd x8, selfi x8].
__device__ __forceinline__ float dense_input(const float (&w)[2 * F], float bias, float sd, float si,
                                             float selfd, float selfi)
{
    float c = bias;
#pragma unroll
    for (int i = 0; i < 8; i++)
        c = fmaf(w[i], sd, c);
#pragma unroll
    for (int i = 8; i < 16; i++)
        c = fmaf(w[i], si, c);
#pragma unroll
    for (int i = 16; i < 24; i++)
        c = fmaf(w[i], selfd, c);
#pragma unroll
    for (int i = 24; i < 32; i++)
        c = fmaf(w[i], selfi, c);
    return relu(c);
}

// Output reduction over lanes 0..7 of the group; valid in lane 0.
__device__ __forceinline__ float output_tree(float h)
{
    const float t = h + __shfl_down(h, 8, F);
    const float s1 = t + __shfl_xor(t, 1, F);
    const float s2 = s1 + __shfl_xor(s1, 2, F);
    return s2 + __shfl_xor(s2, 4, F);
}

__device__ __forceinline__ void load_weights(const float *__restrict__ W, int j, float (&w)[2 * F])
{
#pragma unroll
    for (int i = 0; i < 2 * F; i++)
        w[i] = W[i * F + j];
}

// ------------------------------------------------------------- quad kernel

// Non-hub vertices in degree order (perm[p0 .. p1)), 4 lanes per vertex:
// lane q owns features 4q .. 4q+3 (one 16-byte slice of each 64-byte row), so
// one load instruction fetches 16 neighbour rows and a wave's 16 vertices
// have near-equal degrees. Each lane keeps four independent ordered sums.
// The weights sit in LDS as float4 W[i][4q .. 4q+3]; the dense step takes
// A[i] from the owning lane of the quad.
constexpr int kQuads = kBlock / 4; // vertices per workgroup step
constexpr int kUnroll = 8;         // neighbour rows in flight per lane (16: slower)

template <int kind>
__global__ __launch_bounds__(kBlock) void sage_quads(const uint32_t *__restrict__ V,
                                                     const uint32_t *__restrict__ Er,
                                                     const float *__restrict__ W,
                                                     const float *__restrict__ b,
                                                     const float *__restrict__ x,
                                                     float *__restrict__ out,
                                                     uint32_t *__restrict__ key,
                                                     const uint32_t *__restrict__ perm, uint32_t p0,
                                                     uint32_t p1)
{
    static_assert(kind != kInput, "layer 0 runs in sage_lanes_input");
    __shared__ float4 Ws[2 * F][4];
    __shared__ float4 bs[4];
    for (int t = threadIdx.x; t < 2 * F * F; t += kBlock)
        reinterpret_cast<float *>(Ws)[t] = W[t];
    if (threadIdx.x < F)
        reinterpret_cast<float *>(bs)[threadIdx.x] = b[threadIdx.x];
    __syncthreads();
    const int q = threadIdx.x & 3;
    const float4 *X = reinterpret_cast<const float4 *>(x);
    // whole quads leave the loop together, so the width-4 shuffles stay valid
    for (uint32_t g = p0 + blockIdx.x * kQuads + threadIdx.x / 4; g < p1;
         g += gridDim.x * kQuads) {
        const uint32_t u = perm[g];
        const uint32_t beg = V[u], end = V[u + 1];
        const float fdeg_f = (float)(end - beg);
        const float4 bb = bs[q];
        float c[4] = {bb.x, bb.y, bb.z, bb.w};
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
        uint32_t i = beg;
        for (; i + kUnroll <= end; i += kUnroll) {
            uint32_t v[kUnroll];
#pragma unroll
            for (int t = 0; t < kUnroll; t++)
                v[t] = Er[i + t];
            float4 r[kUnroll];
#pragma unroll
            for (int t = 0; t < kUnroll; t++)
                r[t] = X[(size_t)v[t] * 4 + q];
#pragma unroll
            for (int t = 0; t < kUnroll; t++) {
                s0 = s0 + r[t].x;
                s1 = s1 + r[t].y;
                s2 = s2 + r[t].z;
                s3 = s3 + r[t].w;
            }
        }
        if (i < end) { // the last < kUnroll rows, loads in flight together
            uint32_t v[kUnroll];
#pragma unroll
            for (int t = 0; t < kUnroll; t++)
                v[t] = i + t < end ? Er[i + t] : 0u;
            float4 r[kUnroll];
#pragma unroll
            for (int t = 0; t < kUnroll; t++)
                r[t] = i + t < end ? X[(size_t)v[t] * 4 + q] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
            for (int t = 0; t < kUnroll; t++) {
                if (i + t < end) {
                    s0 = s0 + r[t].x;
                    s1 = s1 + r[t].y;
                    s2 = s2 + r[t].z;
                    s3 = s3 + r[t].w;
                }
            }
        }
        if (fdeg_f > 0.0f) {
            s0 = s0 / fdeg_f;
            s1 = s1 / fdeg_f;
            s2 = s2 / fdeg_f;
            s3 = s3 / fdeg_f;
        }
        const float4 self = X[(size_t)g * 4 + q];
        const float ag[4] = {s0, s1, s2, s3};
        const float sf[4] = {self.x, self.y, self.z, self.w};
#pragma unroll
        for (int i = 0; i < F; i++) {
            const float a = __shfl(ag[i % 4], i / 4, 4);
            const float4 w = Ws[i][q];
            c[0] = fmaf(w.x, a, c[0]);
            c[1] = fmaf(w.y, a, c[1]);
            c[2] = fmaf(w.z, a, c[2]);
            c[3] = fmaf(w.w, a, c[3]);
        }
#pragma unroll
        for (int i = 0; i < F; i++) {
            const float a = __shfl(sf[i % 4], i / 4, 4);
            const float4 w = Ws[F + i][q];
            c[0] = fmaf(w.x, a, c[0]);
            c[1] = fmaf(w.y, a, c[1]);
            c[2] = fmaf(w.z, a, c[2]);
            c[3] = fmaf(w.w, a, c[3]);
        }
        float h[4];
#pragma unroll
        for (int k = 0; k < 4; k++)
            h[k] = relu(c[k]);
        if (kind == kOutput) {
            // t_k = h_k + h_{k+8}: lanes 0,1 hold h_0..7, lanes 2,3 hold h_8..15
            float t[4];
#pragma unroll
            for (int k = 0; k < 4; k++)
                t[k] = h[k] + __shfl(h[k], (q & 1) + 2, 4);
            const float half = (t[0] + t[1]) + (t[2] + t[3]);
            const float p = half + __shfl(half, 1, 4);
            if (q == 0)
                key[u] = float_order_key(p);
        } else {
            reinterpret_cast<float4 *>(out)[(size_t)g * 4 + q] = make_float4(h[0], h[1], h[2], h[3]);
        }
    }
}

// ------------------------------------------------------- input lane kernel

constexpr int kLaneSteps = 1; // blocks of 8 neighbours per load step (2 need more registers than pay)

// Layer 0 for the vertices below the hub degree (perm[p0 .. p1), heaviest
// first): one lane per vertex, so a wave holds 64 vertices of near-equal
// degree and each lane keeps its own two ordered sums (insum::lane_sums,
// streaming E and Ed). The dense step runs in the lane:
// c_j = fmaf(W[i][j], A[i], c_j) for i = 0..31, with the weights read from
// LDS as broadcast float4s. One workgroup per 256 positions: the dispatcher
// hands out the light workgroups as the heavy ones finish.
template <bool kRecip>
__global__ __launch_bounds__(kBlock) void sage_lanes_input(
    uint32_t N, const uint32_t *__restrict__ V, const uint32_t *__restrict__ E,
    const uint32_t *__restrict__ Ed, const float *__restrict__ W, const float *__restrict__ b,
    float fmd, float *__restrict__ out, const uint32_t *__restrict__ perm, uint32_t p0,
    uint32_t p1, const uint2 *__restrict__ gspan)
{
    __shared__ float4 Ws[2 * F][4];
    __shared__ float4 bs[4];
    for (int t = threadIdx.x; t < 2 * F * F; t += kBlock)
        reinterpret_cast<float *>(Ws)[t] = W[t];
    if (threadIdx.x < F)
        reinterpret_cast<float *>(bs)[threadIdx.x] = b[threadIdx.x];
    __syncthreads();
    const uint32_t g = p0 + blockIdx.x * kBlock + threadIdx.x;
    if (g >= p1)
        return;
    const uint32_t u = perm[g];
    const uint32_t beg = V[u], end = V[u + 1];
    const uint2 sp = insum::span_of(u, N, gspan);
    const float2 s = insum::lane_sums<kLaneSteps, kRecip>(beg, end, E, Ed, fmd, sp);
    const float fdeg_f = (float)(end - beg);
    float sd = s.x, si = s.y;
    if (fdeg_f > 0.0f) {
        sd = sd / fdeg_f;
        si = si / fdeg_f;
    }
    const float selfd = insum::deg_feature(end - beg, fmd), selfi = insum::id_feature(u, sp);
    float c[F];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float4 bb = bs[q];
        c[4 * q] = bb.x;
        c[4 * q + 1] = bb.y;
        c[4 * q + 2] = bb.z;
        c[4 * q + 3] = bb.w;
    }
#pragma unroll
    for (int i = 0; i < 2 * F; i++) {
        const float a = i < 8 ? sd : i < 16 ? si : i < 24 ? selfd : selfi;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float4 w = Ws[i][q];
            c[4 * q] = fmaf(w.x, a, c[4 * q]);
            c[4 * q + 1] = fmaf(w.y, a, c[4 * q + 1]);
            c[4 * q + 2] = fmaf(w.z, a, c[4 * q + 2]);
            c[4 * q + 3] = fmaf(w.w, a, c[4 * q + 3]);
        }
    }
    float4 *row = reinterpret_cast<float4 *>(out + (size_t)g * F);
#pragma unroll
    for (int q = 0; q < 4; q++)
        row[q] = make_float4(relu(c[4 * q]), relu(c[4 * q + 1]), relu(c[4 * q + 2]),
                             relu(c[4 * q + 3]));
}

// -------------------------------------------------------------- hub kernel

// Neighbours per loader thread per chunk of the input layer's hub sums:
// 192 loader threads x 8 = 1536 per chunk (measured on R-MAT scale 22 against
// 4 and 16).
constexpr int kHubInRows = 8;

template <int kind>
__global__ __launch_bounds__(kHubBlock) void sage_hubs(GnnCsr gr, const float *__restrict__ W,
                                                       const float *__restrict__ b, float fmd,
                                                       const float *__restrict__ x,
                                                       float *__restrict__ out,
                                                       uint32_t *__restrict__ key)
{
    // the input layer stages 2 features per neighbour, the others 16 (each
    // kind reserves only its own stage: the 33 KB row stage held the input
    // layer's hub workgroups to fewer per CU)
    using Stage = std::conditional_t<kind == kInput, insum::PairStage<kHubBlock, kHubInRows>,
                                     hubsum::Stage<kHubBlock>>;
    __shared__ __attribute__((aligned(16))) Stage st;
    const uint32_t u = gr.perm[blockIdx.x]; // hubs by degree, largest first
    const uint32_t beg = gr.V[u], end = gr.V[u + 1];
    const uint32_t deg = end - beg;
    const int tid = threadIdx.x;
    const int j = tid & (F - 1);
    // Layer 0: lane 0 sums fdeg, lane 1 sums fid. Other layers: lane j sums
    // feature j.
    float acc;
    uint2 sp = make_uint2(0u, gr.N);
    if constexpr (kind == kInput) {
        sp = insum::span_of(u, gr.N, gr.gspan);
        acc = insum::pair_sums(st, beg, end, gr.E, gr.Ed, fmd, sp);
    } else {
        acc = hubsum::ordered_sum<kHubBlock, false>(st, beg, end, gr.Er,
                                                    [&](uint32_t v, float4 *row) {
            const float4 *src = reinterpret_cast<const float4 *>(x + (size_t)v * F);
#pragma unroll
            for (int q = 0; q < 4; q++)
                row[q] = src[q];
        });
    }
    if (tid >= kWave)
        return;
    // Wave 0 finishes; lanes 16..63 mirror lanes 0..15 harmlessly.
    float w[2 * F];
    load_weights(W, j, w);
    const float bias = b[j];
    const float fdeg_f = (float)deg;
    const uint32_t row = gr.rank[u];
    float h;
    if (kind == kInput) {
        float sd = __shfl(acc, 0, F), si = __shfl(acc, 1, F);
        if (fdeg_f > 0.0f) {
            sd = sd / fdeg_f;
            si = si / fdeg_f;
        }
        h = dense_input(w, bias, sd, si, insum::deg_feature(deg, fmd), insum::id_feature(u, sp));
    } else {
        float a = acc;
        if (fdeg_f > 0.0f)
            a = a / fdeg_f;
        h = dense_group(w, bias, a, x[(size_t)row * F + j]);
    }
    if (tid < F) {
        if (kind == kOutput) {
            const float p = output_tree(h);
            if (tid == 0)
                key[u] = float_order_key(p);
        } else {
            out[(size_t)row * F + j] = h;
        }
    }
}

// Workgroups of the quad kernel: enough to fill the device several times
// over, looping over the vertices.
constexpr uint32_t kQuadGrid = 2048;

template <int kind>
int launch_layer(const GnnCsr &gr, const float *W, const float *b, float fmd, bool recip_ok,
                 const float *x, float *out, uint32_t *key, void *stream, void *hub_stream)
{
    if (gr.N == 0)
        return GCX_OK;
    // Hubs go first on their own stream so that they overlap the quads.
    if (gr.nhub)
        sage_hubs<kind><<<gr.nhub, kHubBlock, 0, as_stream(hub_stream)>>>(gr, W, b, fmd, x, out,
                                                                          key);
    const uint32_t rest = gr.N - gr.nhub;
    if constexpr (kind == kInput) {
        const uint32_t grid = div_up(rest, (uint32_t)kBlock);
        if (rest && recip_ok)
            sage_lanes_input<true><<<grid, kBlock, 0, as_stream(stream)>>>(
                gr.N, gr.V, gr.E, gr.Ed, W, b, fmd, out, gr.perm, gr.nhub, gr.N, gr.gspan);
        else if (rest)
            sage_lanes_input<false><<<grid, kBlock, 0, as_stream(stream)>>>(
                gr.N, gr.V, gr.E, gr.Ed, W, b, fmd, out, gr.perm, gr.nhub, gr.N, gr.gspan);
    } else if (rest) {
        uint32_t grid = div_up(rest, kQuads);
        grid = grid < kQuadGrid ? grid : kQuadGrid;
        sage_quads<kind><<<grid, kBlock, 0, as_stream(stream)>>>(gr.V, gr.Er, W, b, x, out, key,
                                                                 gr.perm, gr.nhub, gr.N);
    }
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

} // namespace

extern "C" int gcxk_gnn_fp32_layer(int kind, uint32_t N, const uint32_t *dV, const uint32_t *dE,
                                   const uint32_t *Ed, const uint32_t *Er, const uint32_t *rank,
                                   const float *W, const float *b, int md, const float *x,
                                   float *out, uint32_t *key, const uint32_t *hubs, uint32_t nhub,
                                   const uint32_t *perm, const uint32_t *gspan, void *stream,
                                   void *hub_stream)
{
    const GnnCsr gr = {dV, dE, Ed, Er, rank, perm, hubs, reinterpret_cast<const uint2 *>(gspan), N,
                    nhub};
    switch (kind) {
    case 0:
        return launch_layer<kInput>(gr, W, b, (float)md, insum::recip_md_ok(md), nullptr, out,
                                    nullptr, stream, hub_stream);
    case 1:
        return launch_layer<kHidden>(gr, W, b, 0.0f, false, x, out, nullptr, stream, hub_stream);
    case 2:
        return launch_layer<kOutput>(gr, W, b, 0.0f, false, x, nullptr, key, stream, hub_stream);
    default:
        return GCX_ERR_ARG;
    }
}
