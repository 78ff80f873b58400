// gnn_bf16.hip -- GraphSAGE-16 with MFMA layer transforms (BF16 operands,
// FP32 accumulation), in two precisions:
//
//   BF16   rows stored as bf16 (32 B per vertex); neighbour means accumulated
//          in FP32; the dense step multiplies the bf16-rounded operands
//          [mean | self] (16 vertices x 32) by the bf16-rounded weights
//          (32 x 16) on v_mfma_f32_16x16x16_bf16 with an FP32 bias; ReLU in
//          FP32; the next layer's rows are rounded to bf16; the output layer
//          sums the 16 FP32 features into the priority.
//   SPLIT  rows stored in FP32 and summed in CSR order as in the FP32 path;
//          every operand a and weight w is split into bf16 parts
//          a = a_hi + a_lo, w = w_hi + w_lo, and the dense step runs three
//          MFMAs per K half: a_hi w_hi + a_hi w_lo + a_lo w_hi (the a_lo w_lo
//          term is below 2^-16 relative). Priorities are within a few FP32
//          ulps of the FP32 path's; no step needs FP64.
//
// Rows live at degree-order positions (row of v = rank[v]; neighbours read
// through Er = rank[E]), as in the FP32 kernel.
//
// Mapping (as in the FP32 kernel): vertices below the hub degree in the
// graph's degree order, 4 lanes per vertex, 16 vertices per wave; lane q of a
// vertex loads slice q (features 4q .. 4q+3) of every neighbour row, 8 rows
// in flight, and keeps four ordered FP32 sums. The wave's 16 [mean | self]
// rows are then moved into the MFMA A layout with lane permutes (no LDS) and
// multiplied by the weights held in registers in the B layout. Hubs (degree
// >= the graph's hub degree) are handled by one workgroup each on the side
// stream: the same CSR-order sums and the same MFMA sequence, so results do
// not depend on the hub degree; both kernels only read the previous layer and
// write disjoint rows.
#include "common.hip.h"
#include "hub_sum.hip.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int F = 16;
constexpr int kBlock = 256;
constexpr int kUnroll = 8;        // neighbour rows in flight per lane
constexpr int kGrid = 2048;       // persistent workgroups of the quad kernel

enum { kInput = 0, kHidden = 1, kOutput = 2 };
enum { kBf16 = 0, kSplit = 1 };

__device__ __forceinline__ uint16_t to_bf16(float f)
{
    uint32_t b = __builtin_bit_cast(uint32_t, f);
    if ((b & 0x7F800000u) == 0x7F800000u) // Inf / NaN: keep, quiet NaNs
        return (uint16_t)((b >> 16) | ((b & 0xFFFFu) ? 0x40u : 0u));
    b += 0x7FFFu + ((b >> 16) & 1u); // round to nearest even
    return (uint16_t)(b >> 16);
}

__device__ __forceinline__ float from_bf16(uint32_t h) { return __builtin_bit_cast(float, h << 16); }

__device__ __forceinline__ float fid(uint32_t v, uint32_t N) { return (float)((double)v / (double)N); }

// Id feature: v / N, or the caller's per-vertex value (graph-local for batches).
__device__ __forceinline__ float fid_of(uint32_t v, uint32_t N, const float *__restrict__ fidv)
{
    return fidv ? fidv[v] : fid(v, N);
}

// Slice q (features 4q .. 4q+3) of row r: 8 bytes of bf16 or 16 bytes of FP32.
template <int mode>
__device__ __forceinline__ float4 load_slice(const void *__restrict__ x, uint32_t r, int q)
{
    if (mode == kBf16) {
        const uint2 w = reinterpret_cast<const uint2 *>(x)[(size_t)r * 4 + q];
        return make_float4(from_bf16(w.x & 0xFFFFu), from_bf16(w.x >> 16), from_bf16(w.y & 0xFFFFu),
                           from_bf16(w.y >> 16));
    } else {
        return reinterpret_cast<const float4 *>(x)[(size_t)r * 4 + q];
    }
}

// bf16 parts of a: hi = bf16(a); lo = bf16(a - hi) (exact difference in FP32).
__device__ __forceinline__ void split(float a, short &hi, short &lo)
{
    const uint16_t h = to_bf16(a);
    hi = (short)h;
    lo = (short)to_bf16(a - from_bf16(h));
}

template <int kind, int mode>
__global__ __launch_bounds__(kBlock) void sage_mfma_quads(uint32_t N, const uint32_t *__restrict__ V,
                                                          const uint32_t *__restrict__ E,
                                                          const uint32_t *__restrict__ Er,
                                                          const uint16_t *__restrict__ Wpack,
                                                          const float *__restrict__ bias, float fmd,
                                                          const void *__restrict__ x,
                                                          void *__restrict__ out,
                                                          uint32_t *__restrict__ key,
                                                          const uint32_t *__restrict__ perm,
                                                          uint32_t p0, uint32_t p1,
                                                          const float *__restrict__ fidv)
{
    const int lane = lane_id();
    const int q = lane & 3;       // quad lane: features 4q .. 4q+3 of vertex lane / 4
    const int mi = lane & 15;     // MFMA lane: A row mi / C column mi
    const int mq = lane >> 4;     // MFMA lane: A k-slice mq / C rows 4mq .. 4mq+3
    const int src = 4 * mi + mq;  // quad lane holding row mi, slice mq
    // B operands: pack[m*256 + 4*lane + e] = W[16m + 4*(lane/16) + e][lane%16]
    const bf16x4 bh0 = *reinterpret_cast<const bf16x4 *>(Wpack + 4 * lane);
    const bf16x4 bh1 = *reinterpret_cast<const bf16x4 *>(Wpack + 256 + 4 * lane);
    bf16x4 bl0 = {0, 0, 0, 0}, bl1 = {0, 0, 0, 0};
    if (mode == kSplit) {
        bl0 = *reinterpret_cast<const bf16x4 *>(Wpack + 512 + 4 * lane);
        bl1 = *reinterpret_cast<const bf16x4 *>(Wpack + 768 + 4 * lane);
    }
    const float bj = bias[mi];
    const uint32_t wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const uint32_t nwave = gridDim.x * (kBlock / kWave);
    // the trip count is wave-uniform: the MFMAs below run with every lane
    for (uint32_t base = p0 + wave * 16u; base < p1; base += nwave * 16u) {
        const uint32_t pos = base + (uint32_t)(lane >> 2);
        const bool valid = pos < p1;
        const uint32_t u = valid ? perm[pos] : 0u;
        uint32_t beg = 0, end = 0;
        if (valid) {
            beg = V[u];
            end = V[u + 1];
        }
        const float fdeg = (float)(end - beg);
        float m[4], sf[4];
        if (kind == kInput) {
            // lane q computes the features of neighbours i0 + 4t + q, then
            // every lane adds all 16 in CSR order (the hub kernel's order)
            float sd = 0.0f, si = 0.0f;
            for (uint32_t i0 = beg; i0 < end; i0 += 16) {
                float fd[4], fi[4];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const uint32_t i = i0 + 4 * t + q;
                    fd[t] = 0.0f;
                    fi[t] = 0.0f;
                    if (i < end) {
                        const float2 r = reinterpret_cast<const float2 *>(x)[Er[i]];
                        fd[t] = r.x;
                        fi[t] = r.y;
                    }
                }
                const uint32_t nb = end - i0 < 16u ? end - i0 : 16u;
#pragma unroll
                for (int t = 0; t < 4; t++)
#pragma unroll
                    for (int m2 = 0; m2 < 4; m2++) {
                        const float a = __shfl(fd[t], m2, 4), e = __shfl(fi[t], m2, 4);
                        if ((uint32_t)(4 * t + m2) < nb) {
                            sd = sd + a;
                            si = si + e;
                        }
                    }
            }
            if (fdeg > 0.0f) {
                sd = sd / fdeg;
                si = si / fdeg;
            }
            // k = 4q .. 4q+3 of [sd x8, si x8] and of [fdeg x8, fid x8]
            const float a = q < 2 ? sd : si;
            const float b = valid ? (q < 2 ? fdeg / fmd : fid_of(u, N, fidv)) : 0.0f;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                m[e] = a;
                sf[e] = b;
            }
        } else {
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
                    r[t] = load_slice<mode>(x, v[t], q);
#pragma unroll
                for (int t = 0; t < kUnroll; t++) {
                    s0 = s0 + r[t].x;
                    s1 = s1 + r[t].y;
                    s2 = s2 + r[t].z;
                    s3 = s3 + r[t].w;
                }
            }
            for (; i < end; i++) {
                const float4 r = load_slice<mode>(x, Er[i], q);
                s0 = s0 + r.x;
                s1 = s1 + r.y;
                s2 = s2 + r.z;
                s3 = s3 + r.w;
            }
            if (fdeg > 0.0f) {
                s0 = s0 / fdeg;
                s1 = s1 / fdeg;
                s2 = s2 / fdeg;
                s3 = s3 / fdeg;
            }
            m[0] = s0;
            m[1] = s1;
            m[2] = s2;
            m[3] = s3;
            const float4 self = valid ? load_slice<mode>(x, pos, q) : make_float4(0.f, 0.f, 0.f, 0.f);
            sf[0] = self.x;
            sf[1] = self.y;
            sf[2] = self.z;
            sf[3] = self.w;
        }
        // A operands: lane (mi, mq) takes slice mq of vertex row mi
        bf16x4 ah0, ah1, al0, al1;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const float a = __shfl(m[e], src), b = __shfl(sf[e], src);
            if (mode == kSplit) {
                short h0, l0, h1, l1;
                split(a, h0, l0);
                split(b, h1, l1);
                ah0[e] = h0;
                al0[e] = l0;
                ah1[e] = h1;
                al1[e] = l1;
            } else {
                ah0[e] = (short)to_bf16(a);
                ah1[e] = (short)to_bf16(b);
            }
        }
        f32x4 c = {bj, bj, bj, bj};
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah0, bh0, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah1, bh1, c, 0, 0, 0);
        if (mode == kSplit) {
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah0, bl0, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah1, bl1, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al0, bh0, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al1, bh1, c, 0, 0, 0);
        }
        // c[e] = H[vertex row 4mq + e][feature mi]
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int r = 4 * mq + e;
            const uint32_t v = __shfl(u, 4 * r);
            const bool ok = base + (uint32_t)r < p1;
            float h = c[e] > 0.0f ? c[e] : 0.0f;
            if (kind == kOutput) {
                h += __shfl_xor(h, 1, 16);
                h += __shfl_xor(h, 2, 16);
                h += __shfl_xor(h, 4, 16);
                h += __shfl_xor(h, 8, 16);
                if (mi == 0 && ok)
                    key[v] = float_order_key(h);
            } else if (ok) {
                const size_t row = base + (uint32_t)r; // degree-order position
                if (mode == kBf16)
                    reinterpret_cast<uint16_t *>(out)[row * F + mi] = to_bf16(h);
                else
                    reinterpret_cast<float *>(out)[row * F + mi] = h;
            }
        }
    }
}

// One workgroup per hub. Neighbour rows are staged into LDS feature-major,
// 512 per step with the next step's loads in flight (hub_sum.hip.h), and
// lanes 0..15 of wave 0 sum them in CSR order -- the order of the quad
// kernel's sums, so the mean is the same bits whichever kernel handles the
// vertex. Wave 0 then runs the same MFMA
// sequence with the hub as row 0 of the A tile (the other rows zero): each
// output row of an MFMA depends only on its own A row, so the priorities and
// rows do not depend on the hub degree.
// Threads (= rows per staged chunk) of the hub kernel. Measured on R-MAT
// scale 22 (GNN3 ordering, BF16): 256 -> 4.91 ms, 512 -> 4.63 ms,
// 1024 -> 5.80 ms.
constexpr int kHubChunk = 512;

template <int kind, int mode>
__global__ __launch_bounds__(kHubChunk) void sage_mfma_hubs(uint32_t N, const uint32_t *__restrict__ V,
                                                            const uint32_t *__restrict__ E,
                                                            const uint32_t *__restrict__ Er,
                                                            const uint32_t *__restrict__ rank,
                                                            const uint16_t *__restrict__ Wpack,
                                                            const float *__restrict__ bias, float fmd,
                                                            const void *__restrict__ x,
                                                            void *__restrict__ out,
                                                            uint32_t *__restrict__ key,
                                                            const uint32_t *__restrict__ hubs,
                                                            const float *__restrict__ fidv)
{
    __shared__ __attribute__((aligned(16))) hubsum::Stage<kHubChunk> st;
    const uint32_t u = hubs[blockIdx.x];
    const uint32_t beg = V[u], end = V[u + 1];
    const uint32_t deg = end - beg;
    const int tid = threadIdx.x;
    // layer 0: lane 0 sums fdeg, lane 1 fid; else lane j feature j
    float acc;
    if (kind == kInput) {
        acc = hubsum::ordered_sum<kHubChunk, true>(st, beg, end, Er, [&](uint32_t v, float4 *row) {
            const float2 r = reinterpret_cast<const float2 *>(x)[v];
            row[0] = make_float4(r.x, r.y, 0.0f, 0.0f);
        });
    } else {
        acc = hubsum::ordered_sum<kHubChunk, false>(st, beg, end, Er, [&](uint32_t v, float4 *row) {
#pragma unroll
            for (int q = 0; q < 4; q++)
                row[q] = load_slice<mode>(x, v, q);
        });
    }
    if (tid >= kWave)
        return;
    const int lane = tid;
    const int mi = lane & 15, mq = lane >> 4;
    const float fdeg = (float)deg;
    const uint32_t self_row = rank[u];
    // mean features (lane j holds feature j; layer 0: lanes 0, 1)
    float mean = acc;
    if (fdeg > 0.0f)
        mean = mean / fdeg;
    float a[4], b[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int k = 4 * mq + e;
        if (kind == kInput) {
            const float sd = __shfl(mean, 0), si = __shfl(mean, 1);
            a[e] = mq < 2 ? sd : si;
            b[e] = mq < 2 ? fdeg / fmd : fid_of(u, N, fidv);
        } else {
            a[e] = __shfl(mean, k);
            const float4 r4 = load_slice<mode>(x, self_row, mq);
            b[e] = e == 0 ? r4.x : e == 1 ? r4.y : e == 2 ? r4.z : r4.w;
        }
    }
    bf16x4 ah0 = {0, 0, 0, 0}, ah1 = {0, 0, 0, 0}, al0 = {0, 0, 0, 0}, al1 = {0, 0, 0, 0};
    if (mi == 0) { // A row 0 = the hub; rows 1..15 stay zero
#pragma unroll
        for (int e = 0; e < 4; e++) {
            if (mode == kSplit) {
                short h0, l0, h1, l1;
                split(a[e], h0, l0);
                split(b[e], h1, l1);
                ah0[e] = h0;
                al0[e] = l0;
                ah1[e] = h1;
                al1[e] = l1;
            } else {
                ah0[e] = (short)to_bf16(a[e]);
                ah1[e] = (short)to_bf16(b[e]);
            }
        }
    }
    const bf16x4 bh0 = *reinterpret_cast<const bf16x4 *>(Wpack + 4 * lane);
    const bf16x4 bh1 = *reinterpret_cast<const bf16x4 *>(Wpack + 256 + 4 * lane);
    const float bj = bias[mi];
    f32x4 c = {bj, bj, bj, bj};
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah0, bh0, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah1, bh1, c, 0, 0, 0);
    if (mode == kSplit) {
        const bf16x4 bl0 = *reinterpret_cast<const bf16x4 *>(Wpack + 512 + 4 * lane);
        const bf16x4 bl1 = *reinterpret_cast<const bf16x4 *>(Wpack + 768 + 4 * lane);
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah0, bl0, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah1, bl1, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al0, bh0, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al1, bh1, c, 0, 0, 0);
    }
    // row 0 of C: lanes 0..15 (mq = 0), element 0
    float h = c[0] > 0.0f ? c[0] : 0.0f;
    if (kind == kOutput) {
        h += __shfl_xor(h, 1, 16);
        h += __shfl_xor(h, 2, 16);
        h += __shfl_xor(h, 4, 16);
        h += __shfl_xor(h, 8, 16);
        if (lane == 0)
            key[u] = float_order_key(h);
    } else if (lane < F) {
        if (mode == kBf16)
            reinterpret_cast<uint16_t *>(out)[(size_t)self_row * F + mi] = to_bf16(h);
        else
            reinterpret_cast<float *>(out)[(size_t)self_row * F + mi] = h;
    }
}

template <int kind, int mode>
int launch(uint32_t N, const uint32_t *V, const uint32_t *E, const uint32_t *Er, const uint32_t *rank,
           const uint16_t *Wpack, const float *b, float fmd, const void *x, void *out,
           uint32_t *key, const uint32_t *hubs, uint32_t nhub, const uint32_t *perm,
           const float *fidv, void *stream, void *hub_stream)
{
    if (N == 0)
        return GCX_OK;
    if (nhub)
        sage_mfma_hubs<kind, mode><<<nhub, kHubChunk, 0, as_stream(hub_stream)>>>(
            N, V, E, Er, rank, Wpack, b, fmd, x, out, key, hubs, fidv);
    if (N > nhub) {
        const uint32_t waves = div_up(N - nhub, 16u);
        const uint32_t need = div_up(waves, (uint32_t)(kBlock / kWave));
        sage_mfma_quads<kind, mode><<<need < (uint32_t)kGrid ? need : (uint32_t)kGrid, kBlock, 0,
                                      as_stream(stream)>>>(N, V, E, Er, Wpack, b, fmd, x, out,
                                                           key, perm, nhub, N, fidv);
    }
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

template <int mode>
int layer(int kind, uint32_t N, const uint32_t *dV, const uint32_t *dE, const uint32_t *Er,
          const uint32_t *rank, const uint16_t *Wpack, const float *b, float fmd, const void *x,
          void *out, uint32_t *key, const uint32_t *hubs, uint32_t nhub, const uint32_t *perm,
          const float *fidv, void *stream, void *hub_stream)
{
    switch (kind) {
    case 0:
        return launch<kInput, mode>(N, dV, dE, Er, rank, Wpack, b, fmd, x, out, nullptr,
                                    hubs, nhub, perm, fidv, stream, hub_stream);
    case 1:
        return launch<kHidden, mode>(N, dV, dE, Er, rank, Wpack, b, fmd, x, out, nullptr, hubs,
                                     nhub, perm, nullptr, stream, hub_stream);
    case 2:
        return launch<kOutput, mode>(N, dV, dE, Er, rank, Wpack, b, fmd, x, nullptr, key, hubs,
                                     nhub, perm, nullptr, stream, hub_stream);
    default:
        return GCX_ERR_ARG;
    }
}

} // namespace

extern "C" int gcxk_gnn_bf16_layer(int kind, int split_mode, uint32_t N, const uint32_t *dV,
                                   const uint32_t *dE, const uint32_t *Er, const uint32_t *rank,
                                   const uint16_t *Wpack, const float *b, int md, const void *x,
                                   void *out, uint32_t *key, const uint32_t *hubs, uint32_t nhub,
                                   const uint32_t *perm, const float *fidv, void *stream,
                                   void *hub_stream)
{
    const float fmd = (float)md;
    return split_mode ? layer<kSplit>(kind, N, dV, dE, Er, rank, Wpack, b, fmd, x, out, key, hubs,
                                      nhub, perm, fidv, stream, hub_stream)
                      : layer<kBf16>(kind, N, dV, dE, Er, rank, Wpack, b, fmd, x, out, key, hubs,
                                     nhub, perm, fidv, stream, hub_stream);
}
