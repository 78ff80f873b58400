// gnn_bf16.hip -- GraphSAGE-16 with MFMA layer transforms (BF16 operands,
// FP32 accumulation), in two precisions:
//
//   BF16   rows stored as bf16 (32 B per vertex); neighbour means accumulated
//          in FP32 (blocked order, below); the dense step multiplies the
//          bf16-rounded operands
//          [mean | self] (16 vertices x 32) by the bf16-rounded weights
//          (32 x 16) on v_mfma_f32_16x16x16_bf16 with an FP32 bias; ReLU in
//          FP32; the next layer's rows are rounded to bf16; the output layer
//          sums the 16 FP32 features into the priority.
//   SPLIT  rows stored in FP32 and summed in FP32 (blocked order, below);
//          every operand a and weight w is split into bf16 parts
//          a = a_hi + a_lo, w = w_hi + w_lo, and the dense step runs three
//          MFMAs per K half: a_hi w_hi + a_hi w_lo + a_lo w_hi (the a_lo w_lo
//          term is below 2^-16 relative). Priorities are within a few FP32
//          ulps of the FP32 path's; no step needs FP64.
//
// Rows live at degree-order positions (row of v = rank[v]; neighbours read
// through Er = rank[E]), as in the FP32 kernel.
//
// Neighbour sums of the hidden and output layers are blocked: blocks of 32
// consecutive CSR entries, each summed in order, block sums added in order
// (hubsum::kSumBlock). The quad and hub kernels use the same order, so
// results do not depend on the hub degree, and a hub's dependent add chain
// is deg / 32 long (the FP32 path keeps the reference's single chain, which
// made the largest hub, 162,724 entries on R-MAT 22, the layer's critical
// path). The input layer keeps the FP32 path's sums.
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
#include <type_traits>

#include "common.hip.h"
#include "hub_sum.hip.h"
#include "input_sum.hip.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int F = 16;
constexpr int kBlock = 256;
constexpr int kUnroll = 8;        // neighbour rows in flight per lane
constexpr int kGrid = 2048;       // persistent workgroups of the quad kernel
constexpr int kLaneSteps = 1;     // input layer: blocks of 8 neighbours per load step
constexpr int kHubInRows = 8;     // input layer hubs: neighbours per loader thread per chunk

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
__global__ __launch_bounds__(kBlock) void sage_mfma_quads(const uint32_t *__restrict__ V,
                                                          const uint32_t *__restrict__ Er,
                                                          const uint16_t *__restrict__ Wpack,
                                                          const float *__restrict__ bias,
                                                          const void *__restrict__ x,
                                                          void *__restrict__ out,
                                                          uint32_t *__restrict__ key,
                                                          const uint32_t *__restrict__ perm,
                                                          uint32_t p0, uint32_t p1)
{
    static_assert(kind != kInput, "layer 0 runs in sage_mfma_lanes_input");
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
        // blocked sums (hub_sum.hip.h): blocks of kSumBlock entries in
        // order, the block sums added in order -- as the hub kernel does
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
        for (uint32_t b0 = beg; b0 < end; b0 += hubsum::kSumBlock) {
            const uint32_t be = b0 + hubsum::kSumBlock < end ? b0 + hubsum::kSumBlock : end;
            float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f, t3 = 0.0f;
            uint32_t i = b0;
            for (; i + kUnroll <= be; i += kUnroll) {
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
                    t0 = t0 + r[t].x;
                    t1 = t1 + r[t].y;
                    t2 = t2 + r[t].z;
                    t3 = t3 + r[t].w;
                }
            }
            if (i < be) { // the last < kUnroll rows, loads in flight together
                uint32_t v[kUnroll];
#pragma unroll
                for (int t = 0; t < kUnroll; t++)
                    v[t] = i + t < be ? Er[i + t] : 0u;
                float4 r[kUnroll];
#pragma unroll
                for (int t = 0; t < kUnroll; t++)
                    r[t] = i + t < be ? load_slice<mode>(x, v[t], q)
                                      : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
                for (int t = 0; t < kUnroll; t++) {
                    if (i + t < be) {
                        t0 = t0 + r[t].x;
                        t1 = t1 + r[t].y;
                        t2 = t2 + r[t].z;
                        t3 = t3 + r[t].w;
                    }
                }
            }
            s0 = s0 + t0;
            s1 = s1 + t1;
            s2 = s2 + t2;
            s3 = s3 + t3;
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

// Layer 0 for the vertices below the hub degree: one lane per vertex, 64
// vertices per wave (insum::lane_sums, the FP32 path's sums), then four
// 16-vertex MFMA tiles. A row mi of tile t is the vertex of lane 16t + mi:
// K half 0 is [sd x8, si x8], K half 1 [fdeg/md x8, fid x8], so lane
// (mi, mq) takes sd or si (fdeg/md or fid) of that vertex by mq < 2.
template <int mode, bool kRecip>
__global__ __launch_bounds__(kBlock) void sage_mfma_lanes_input(GnnCsr gr,
                                                                const uint16_t *__restrict__ Wpack,
                                                                const float *__restrict__ bias,
                                                                float fmd, void *__restrict__ out)
{
    const uint32_t p0 = gr.nhub, p1 = gr.N;
    const int lane = lane_id();
    const int mi = lane & 15, mq = lane >> 4;
    const uint32_t wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const uint32_t base = p0 + wave * (uint32_t)kWave;
    if (base >= p1) // wave-uniform: the MFMAs below run with every lane
        return;
    const bf16x4 bh0 = *reinterpret_cast<const bf16x4 *>(Wpack + 4 * lane);
    const bf16x4 bh1 = *reinterpret_cast<const bf16x4 *>(Wpack + 256 + 4 * lane);
    bf16x4 bl0 = {0, 0, 0, 0}, bl1 = {0, 0, 0, 0};
    if (mode == kSplit) {
        bl0 = *reinterpret_cast<const bf16x4 *>(Wpack + 512 + 4 * lane);
        bl1 = *reinterpret_cast<const bf16x4 *>(Wpack + 768 + 4 * lane);
    }
    const float bj = bias[mi];
    const uint32_t pos = base + (uint32_t)lane;
    const bool valid = pos < p1;
    const uint32_t u = valid ? gr.perm[pos] : 0u;
    uint32_t beg = 0, end = 0;
    if (valid) {
        beg = gr.V[u];
        end = gr.V[u + 1];
    }
    const uint2 sp = insum::span_of(u, gr.N, gr.gspan);
    const float2 s = insum::lane_sums<kLaneSteps, kRecip>(beg, end, gr.E, gr.Ed, fmd, sp);
    const float fdeg = (float)(end - beg);
    float sd = s.x, si = s.y;
    if (fdeg > 0.0f) {
        sd = sd / fdeg;
        si = si / fdeg;
    }
    const float selfd = insum::deg_feature(end - beg, fmd);
    const float selfi = valid ? insum::id_feature(u, sp) : 0.0f;
#pragma unroll
    for (int t = 0; t < kWave / 16; t++) {
        const int src = 16 * t + mi;
        const float sdt = __shfl(sd, src), sit = __shfl(si, src);
        const float fdt = __shfl(selfd, src), fit = __shfl(selfi, src);
        const float a = mq < 2 ? sdt : sit, b = mq < 2 ? fdt : fit;
        bf16x4 ah0, ah1, al0, al1;
        if (mode == kSplit) {
            short h0, l0, h1, l1;
            split(a, h0, l0);
            split(b, h1, l1);
            ah0 = bf16x4{h0, h0, h0, h0};
            al0 = bf16x4{l0, l0, l0, l0};
            ah1 = bf16x4{h1, h1, h1, h1};
            al1 = bf16x4{l1, l1, l1, l1};
        } else {
            const short h0 = (short)to_bf16(a), h1 = (short)to_bf16(b);
            ah0 = bf16x4{h0, h0, h0, h0};
            ah1 = bf16x4{h1, h1, h1, h1};
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
        // c[e] = H[tile row 4mq + e][feature mi]
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const size_t row = base + (uint32_t)(16 * t + 4 * mq + e); // degree-order position
            if (row < p1) {
                const float h = c[e] > 0.0f ? c[e] : 0.0f;
                if (mode == kBf16)
                    reinterpret_cast<uint16_t *>(out)[row * F + mi] = to_bf16(h);
                else
                    reinterpret_cast<float *>(out)[row * F + mi] = h;
            }
        }
    }
}

// One workgroup per hub. Neighbour rows are staged into LDS feature-major,
// 512 per step with the next step's loads in flight (hub_sum.hip.h); 256
// threads sum the chunk's 16 blocks x 16 features and lanes 0..15 of wave 0
// add the block sums in order -- the quad kernel's order, so the mean is the
// same bits whichever kernel handles the vertex. Wave 0 then runs the same MFMA
// sequence with the hub as row 0 of the A tile (the other rows zero): each
// output row of an MFMA depends only on its own A row, so the priorities and
// rows do not depend on the hub degree.
// Threads (= rows per staged chunk) of the hub kernel. Hidden and output
// layers, measured on R-MAT scale 22 (GNN3 ordering, BF16): 256 -> 4.91 ms,
// 512 -> 4.63 ms, 1024 -> 5.80 ms. The input layer stages two features per
// neighbour, 8 neighbours per loader thread, in 256-thread workgroups, as
// the FP32 input layer does (the hidden layers' 512 threads and 4
// neighbours per loader made it 0.81 ms against FP32's 0.69).
constexpr int kHubChunk = 512;

template <int kind>
constexpr int hub_threads()
{
    return kind == kInput ? 256 : kHubChunk;
}

template <int kind, int mode>
__global__ __launch_bounds__(hub_threads<kind>()) void sage_mfma_hubs(GnnCsr gr,
                                                            const uint16_t *__restrict__ Wpack,
                                                            const float *__restrict__ bias,
                                                            float fmd, const void *__restrict__ x,
                                                            void *__restrict__ out,
                                                            uint32_t *__restrict__ key)
{
    // the input layer stages 2 features per neighbour, the others 16
    constexpr int T = hub_threads<kind>();
    using Stage = std::conditional_t<kind == kInput, insum::PairStage<T, kHubInRows>,
                                     hubsum::BlockedStage<T>>;
    __shared__ __attribute__((aligned(16))) Stage st;
    const uint32_t u = gr.perm[blockIdx.x]; // hubs by degree, largest first
    const uint32_t beg = gr.V[u], end = gr.V[u + 1];
    const uint32_t deg = end - beg;
    const int tid = threadIdx.x;
    // layer 0: lane 0 sums fdeg, lane 1 fid; else lane j feature j
    float acc;
    uint2 sp = make_uint2(0u, gr.N);
    if constexpr (kind == kInput) {
        sp = insum::span_of(u, gr.N, gr.gspan);
        acc = insum::pair_sums(st, beg, end, gr.E, gr.Ed, fmd, sp);
    } else {
        acc = hubsum::ordered_sum<T, false, true>(st, beg, end, gr.Er,
                                                    [&](uint32_t v, float4 *row) {
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
    const uint32_t self_row = gr.rank[u];
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
            b[e] = mq < 2 ? insum::deg_feature(deg, fmd) : insum::id_feature(u, sp);
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
int launch(const GnnCsr &gr, const uint16_t *Wpack, const float *b, float fmd, bool recip_ok,
           const void *x, void *out, uint32_t *key, void *stream, void *hub_stream)
{
    if (gr.N == 0)
        return GCX_OK;
    if (gr.nhub)
        sage_mfma_hubs<kind, mode><<<gr.nhub, hub_threads<kind>(), 0, as_stream(hub_stream)>>>(
            gr, Wpack, b, fmd, x, out, key);
    if constexpr (kind == kInput) {
        if (gr.N > gr.nhub) {
            const uint32_t waves = div_up(gr.N - gr.nhub, (uint32_t)kWave);
            const uint32_t grid = div_up(waves, (uint32_t)(kBlock / kWave));
            if (recip_ok)
                sage_mfma_lanes_input<mode, true>
                    <<<grid, kBlock, 0, as_stream(stream)>>>(gr, Wpack, b, fmd, out);
            else
                sage_mfma_lanes_input<mode, false>
                    <<<grid, kBlock, 0, as_stream(stream)>>>(gr, Wpack, b, fmd, out);
        }
    } else if (gr.N > gr.nhub) {
        const uint32_t waves = div_up(gr.N - gr.nhub, 16u);
        const uint32_t need = div_up(waves, (uint32_t)(kBlock / kWave));
        sage_mfma_quads<kind, mode><<<need < (uint32_t)kGrid ? need : (uint32_t)kGrid, kBlock, 0,
                                      as_stream(stream)>>>(gr.V, gr.Er, Wpack, b, x, out, key,
                                                           gr.perm, gr.nhub, gr.N);
    }
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

template <int mode>
int layer(int kind, const GnnCsr &gr, const uint16_t *Wpack, const float *b, int md,
          const void *x, void *out, uint32_t *key, void *stream, void *hub_stream)
{
    const float fmd = (float)md;
    const bool ok = insum::recip_md_ok(md);
    switch (kind) {
    case 0:
        return launch<kInput, mode>(gr, Wpack, b, fmd, ok, nullptr, out, nullptr, stream,
                                    hub_stream);
    case 1:
        return launch<kHidden, mode>(gr, Wpack, b, fmd, ok, x, out, nullptr, stream, hub_stream);
    case 2:
        return launch<kOutput, mode>(gr, Wpack, b, fmd, ok, x, nullptr, key, stream, hub_stream);
    default:
        return GCX_ERR_ARG;
    }
}

} // namespace

extern "C" int gcxk_gnn_bf16_layer(int kind, int split_mode, uint32_t N, const uint32_t *dV,
                                   const uint32_t *dE, const uint32_t *Ed, const uint32_t *Er,
                                   const uint32_t *rank, const uint16_t *Wpack, const float *b,
                                   int md, const void *x, void *out, uint32_t *key,
                                   const uint32_t *hubs, uint32_t nhub, const uint32_t *perm,
                                   const uint32_t *gspan, void *stream, void *hub_stream)
{
    const GnnCsr gr = {dV, dE, Ed, Er, rank, perm, hubs, reinterpret_cast<const uint2 *>(gspan), N,
                       nhub};
    return split_mode ? layer<kSplit>(kind, gr, Wpack, b, md, x, out, key, stream, hub_stream)
                      : layer<kBf16>(kind, gr, Wpack, b, md, x, out, key, stream, hub_stream);
}
