// gnn_bf16.hip -- GraphSAGE-16 with BF16 rows and MFMA layer transforms.
//
// Precision contract: neighbour rows are stored as bf16 (32 B per vertex);
// neighbour means are accumulated in FP32; the dense step multiplies
// bf16-rounded operands [mean | self] (16 vertices x 32) by the bf16-rounded
// weights (32 x 16) on v_mfma_f32_16x16x16_bf16 with FP32 accumulation and an
// FP32 bias; ReLU in FP32; the next layer's rows are rounded to bf16; the
// output layer sums the 16 FP32 features into the priority.
//
// Mapping: a 256-thread workgroup owns 64 vertices that are consecutive in
// the graph's degree order (perm, hubs excluded), so their gathers take
// near-equal time. Phase 1:
// four lanes per vertex stride its neighbours, each loading whole 32-byte
// rows (two 16-byte loads) into 16 FP32 partial sums, then a 2-step xor
// reduction combines the four partials. The bf16 operand tile [64 x 32]
// goes to LDS. Phase 2: each wave multiplies its 16 vertices with two MFMAs
// (K = 0..15 means, K = 16..31 self) and stores bf16 rows (or priorities).
// Hubs (degree >= GCXK_HUB_DEG) are skipped by the tile kernel and handled
// by one workgroup each on the side stream (parallel row reduction, then the
// same bf16 arithmetic on VALU); both kernels only read the previous layer
// and write disjoint rows.
#include "common.hip.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int F = 16;
constexpr int kBlock = 256;
constexpr int kTile = 64;       // vertices per workgroup
constexpr int kSub = 4;         // lanes per vertex in the gather phase
constexpr int kHubBlock = 256;

enum { kInput = 0, kHidden = 1, kOutput = 2 };

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

// Accumulate the 16 bf16 features of row r (two 16-byte loads).
__device__ __forceinline__ void add_row(const uint16_t *__restrict__ x, uint32_t r, float (&s)[F])
{
    const uint4 *p = reinterpret_cast<const uint4 *>(x + (size_t)r * F);
    const uint4 a = p[0], b = p[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int k = 0; k < 8; k++) {
        s[2 * k] += from_bf16(w[k] & 0xFFFFu);
        s[2 * k + 1] += from_bf16(w[k] >> 16);
    }
}

template <int kind>
__global__ __launch_bounds__(kBlock) void sage_bf16_tiles(uint32_t N, const uint32_t *__restrict__ V,
                                                          const uint32_t *__restrict__ E,
                                                          const uint16_t *__restrict__ Wpack,
                                                          const float *__restrict__ bias, float fmd,
                                                          const uint16_t *__restrict__ x,
                                                          uint16_t *__restrict__ out,
                                                          uint32_t *__restrict__ key,
                                                          const uint32_t *__restrict__ perm,
                                                          uint32_t p0, const float *__restrict__ fidv)
{
    // Operand tile: 64 vertices x 32 bf16, padded to 36 to spread LDS banks.
    __shared__ uint16_t A[kTile][36];
    __shared__ uint32_t tu[kTile]; // vertex of each tile row, N = none
    const int t = threadIdx.x;
    const int vi = t / kSub, sub = t % kSub;
    const uint32_t pos = p0 + blockIdx.x * kTile + vi;
    const bool skip = pos >= N;
    const uint32_t u = skip ? 0u : perm[pos];
    uint32_t beg = 0, end = 0;
    if (!skip) {
        beg = V[u];
        end = V[u + 1];
    }
    if (sub == 0)
        tu[vi] = skip ? N : u;
    const uint32_t deg = end - beg;
    const float fdeg = (float)deg;
    if (kind == kInput) {
        float sd = 0.0f, si = 0.0f;
        if (!skip)
            for (uint32_t i = beg + sub; i < end; i += kSub) {
                const uint32_t v = E[i];
                sd += (float)(V[v + 1] - V[v]) / fmd;
                si += fid_of(v, N, fidv);
            }
        sd += __shfl_xor(sd, 1);
        si += __shfl_xor(si, 1);
        sd += __shfl_xor(sd, 2);
        si += __shfl_xor(si, 2);
        if (fdeg > 0.0f) {
            sd /= fdeg;
            si /= fdeg;
        }
        // k = 4*sub .. 4*sub+3 of [sd x8, si x8] and of [fdeg x8, fid x8]
        const uint16_t lo = to_bf16(sub < 2 ? sd : si);
        const uint16_t hi =
            to_bf16(sub < 2 ? (skip ? 0.0f : fdeg / fmd) : (skip ? 0.0f : fid_of(u, N, fidv)));
#pragma unroll
        for (int e = 0; e < 4; e++) {
            A[vi][4 * sub + e] = lo;
            A[vi][16 + 4 * sub + e] = hi;
        }
    } else {
        float s[F];
#pragma unroll
        for (int k = 0; k < F; k++)
            s[k] = 0.0f;
        if (!skip)
            for (uint32_t i = beg + sub; i < end; i += kSub)
                add_row(x, E[i], s);
#pragma unroll
        for (int k = 0; k < F; k++) {
            s[k] += __shfl_xor(s[k], 1);
            s[k] += __shfl_xor(s[k], 2);
        }
        const uint16_t *self = x + (size_t)(skip ? 0 : u) * F;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int k = 4 * sub + e;
            const float m = fdeg > 0.0f ? s[k] / fdeg : s[k];
            A[vi][k] = to_bf16(m);
            A[vi][16 + k] = skip ? (uint16_t)0 : self[k];
        }
    }
    __syncthreads();
    // Phase 2: wave w multiplies vertices 16w..16w+15.
    const int lane = lane_id(), w = t / kWave;
    const int i = lane & 15, q = lane >> 4;
    const int row = 16 * w + i;
    bf16x4 a0, a1;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        a0[e] = (short)A[row][4 * q + e];
        a1[e] = (short)A[row][16 + 4 * q + e];
    }
    const bf16x4 b0 = *reinterpret_cast<const bf16x4 *>(Wpack + 4 * lane);
    const bf16x4 b1 = *reinterpret_cast<const bf16x4 *>(Wpack + 256 + 4 * lane);
    const float bj = bias[i];
    f32x4 c = {bj, bj, bj, bj};
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b0, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b1, c, 0, 0, 0);
    // c[e] = H[tile row 16w + 4q + e][feature i]
    const int rbase = 16 * w + 4 * q;
    if (kind == kOutput) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            float h = c[e] > 0.0f ? c[e] : 0.0f;
            h += __shfl_xor(h, 1, 16);
            h += __shfl_xor(h, 2, 16);
            h += __shfl_xor(h, 4, 16);
            h += __shfl_xor(h, 8, 16);
            const uint32_t v = tu[rbase + e];
            if (i == 0 && v < N)
                key[v] = float_order_key(h);
        }
    } else {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const uint32_t v = tu[rbase + e];
            if (v < N)
                out[(size_t)v * F + i] = to_bf16(c[e] > 0.0f ? c[e] : 0.0f);
        }
    }
}

template <int kind>
__global__ __launch_bounds__(kHubBlock) void sage_bf16_hubs(uint32_t N, const uint32_t *__restrict__ V,
                                                            const uint32_t *__restrict__ E,
                                                            const uint16_t *__restrict__ Wb,
                                                            const float *__restrict__ bias, float fmd,
                                                            const uint16_t *__restrict__ x,
                                                            uint16_t *__restrict__ out,
                                                            uint32_t *__restrict__ key,
                                                            const uint32_t *__restrict__ hubs,
                                                            const float *__restrict__ fidv)
{
    __shared__ float part[kHubBlock][F + 1];
    __shared__ float mean[F];
    const uint32_t u = hubs[blockIdx.x];
    const uint32_t beg = V[u], end = V[u + 1];
    const float fdeg = (float)(end - beg);
    const int t = threadIdx.x;
    float s[F];
#pragma unroll
    for (int k = 0; k < F; k++)
        s[k] = 0.0f;
    if (kind == kInput) {
        for (uint32_t i = beg + t; i < end; i += kHubBlock) {
            const uint32_t v = E[i];
            s[0] += (float)(V[v + 1] - V[v]) / fmd;
            s[1] += fid_of(v, N, fidv);
        }
    } else {
        for (uint32_t i = beg + t; i < end; i += kHubBlock)
            add_row(x, E[i], s);
    }
#pragma unroll
    for (int k = 0; k < F; k++)
        part[t][k] = s[k];
    __syncthreads();
    if (t < F) {
        float acc = 0.0f;
        for (int r = 0; r < kHubBlock; r++) // fixed order: deterministic
            acc += part[r][t];
        mean[t] = fdeg > 0.0f ? acc / fdeg : acc;
    }
    __syncthreads();
    if (t >= F)
        return;
    float a[2 * F];
    if (kind == kInput) {
        const float sd = from_bf16(to_bf16(mean[0])), si = from_bf16(to_bf16(mean[1]));
        const float d = from_bf16(to_bf16(fdeg / fmd)),
                    id = from_bf16(to_bf16(fid_of(u, N, fidv)));
        for (int k = 0; k < 8; k++) {
            a[k] = sd;
            a[8 + k] = si;
            a[16 + k] = d;
            a[24 + k] = id;
        }
    } else {
        for (int k = 0; k < F; k++) {
            a[k] = from_bf16(to_bf16(mean[k]));
            a[F + k] = from_bf16(x[(size_t)u * F + k]);
        }
    }
    float c = bias[t];
    for (int k = 0; k < 2 * F; k++)
        c += from_bf16(Wb[k * F + t]) * a[k];
    const float h = c > 0.0f ? c : 0.0f;
    if (kind == kOutput) {
        float sum = h;
        sum += __shfl_xor(sum, 1, 16);
        sum += __shfl_xor(sum, 2, 16);
        sum += __shfl_xor(sum, 4, 16);
        sum += __shfl_xor(sum, 8, 16);
        if (t == 0)
            key[u] = float_order_key(sum);
    } else {
        out[(size_t)u * F + t] = to_bf16(h);
    }
}

template <int kind>
int launch(uint32_t N, const uint32_t *V, const uint32_t *E, const uint16_t *Wpack,
           const uint16_t *Wb, const float *b, float fmd, const uint16_t *x, uint16_t *out,
           uint32_t *key, const uint32_t *hubs, uint32_t nhub, const uint32_t *perm,
           const float *fidv, void *stream, void *hub_stream)
{
    if (N == 0)
        return GCX_OK;
    if (nhub)
        sage_bf16_hubs<kind><<<nhub, kHubBlock, 0, as_stream(hub_stream)>>>(N, V, E, Wb, b, fmd, x,
                                                                            out, key, hubs, fidv);
    if (N > nhub)
        sage_bf16_tiles<kind><<<div_up(N - nhub, kTile), kBlock, 0, as_stream(stream)>>>(
            N, V, E, Wpack, b, fmd, x, out, key, perm, nhub, fidv);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

} // namespace

extern "C" int gcxk_gnn_bf16_layer(int kind, uint32_t N, const uint32_t *dV, const uint32_t *dE,
                                   const uint16_t *Wpack, const uint16_t *Wb, const float *b,
                                   int md, const uint16_t *x, uint16_t *out, uint32_t *key,
                                   const uint32_t *hubs, uint32_t nhub, const uint32_t *perm,
                                   const float *fidv, void *stream, void *hub_stream)
{
    const float fmd = (float)md;
    switch (kind) {
    case 0:
        return launch<kInput>(N, dV, dE, Wpack, Wb, b, fmd, nullptr, out, nullptr, hubs, nhub,
                              perm, fidv, stream, hub_stream);
    case 1:
        return launch<kHidden>(N, dV, dE, Wpack, Wb, b, fmd, x, out, nullptr, hubs, nhub, perm,
                               nullptr, stream, hub_stream);
    case 2:
        return launch<kOutput>(N, dV, dE, Wpack, Wb, b, fmd, x, nullptr, key, hubs, nhub, perm,
                               nullptr, stream, hub_stream);
    default:
        return GCX_ERR_ARG;
    }
}
