// input_sum.hip.h -- the input layer's ordered neighbour sums.
//
// Layer 0 of the reference's GraphSAGE (src/fast_gnn.c:51-89) sums, for each
// vertex, the degree feature (float)deg(v) / (float)md and the id feature
// (float)((double)v / N) of its neighbours left to right in CSR order in FP32,
// starting from +0. Here both features come from sequential per-edge reads:
// the neighbour id E[j] and its degree Ed[j] (written with the graph), so a
// neighbour costs 8 streamed bytes and two divisions by per-vertex constants
// (with the reciprocals refined once, see Recip32 / Recip64) instead of a
// random gather. In a pack of graphs the id feature is graph-local,
// (v - off) / n for the (first vertex, vertex count) of the vertex's own graph;
// an edge never leaves its graph, so the source vertex's span serves all its
// neighbours. Two mappings:
//
//   lane_sums  one lane per vertex (vertices below the hub degree); the
//              lane walks its list in aligned steps of 8U entries, so E and
//              Ed arrive as 16-byte loads covering whole 32U-byte spans, all
//              of a step's loads issued before its adds.
//   pair_sums  one workgroup per hub; each thread computes R features per
//              chunk of T * R rows into two LDS columns, lanes 0 and 1 of
//              wave 0 add the columns in order while the next chunk's loads
//              are in flight.
//
// Both add the same terms in the same order, so a vertex's sums do not depend
// on which mapping (i.e. on the hub degree) handled it.
#pragma once
#include "common.hip.h"
#include "hub_sum.hip.h"

// Graph arrays of the GNN layer kernels (device pointers): CSR, per-edge
// neighbour degrees Ed and degree-order columns Er, the degree order and its
// inverse, the hubs, and for packs of graphs each vertex's graph span.
struct GnnCsr {
    const uint32_t *V, *E, *Ed, *Er, *rank, *perm, *hubs;
    const uint2 *gspan;
    uint32_t N, nhub;
};

namespace insum {

// Entries after the last edge that E and Ed carry (readable; Ed's are 0):
// lane_sums reads whole 16-byte words, at most 3 entries past a list's end.
static_assert(GCXK_ER_PAD >= 3u, "lane_sums reads up to 3 entries past the last edge");

// (first vertex, vertex count) of the graph holding u: its packed graph's, or
// (0, N) for a single graph.
__device__ __forceinline__ uint2 span_of(uint32_t u, uint32_t N, const uint2 *__restrict__ gspan)
{
    return gspan ? gspan[u] : make_uint2(0u, N);
}

__device__ __forceinline__ float deg_feature(uint32_t deg, float fmd) { return (float)deg / fmd; }

__device__ __forceinline__ float id_feature(uint32_t v, uint2 sp)
{
    return (float)((double)(v - sp.x) / (double)sp.y);
}

// Division by a divisor shared by many numerators. The compiler's IEEE
// division runs v_div_scale on both operands, refines v_rcp of the scaled
// divisor by Newton steps, forms the quotient and its correction with FMAs
// (v_div_fmas) and passes the result through v_div_fixup. For a divisor in
// [1, 2^32] and a numerator that is 0 or in [1, 2^32] no scaling applies and
// the fixup returns its input, so running the same steps with the
// divisor's reciprocal refined once gives the same quotient bit for bit, at
// 3 (FP64) or 5 (FP32) operations per numerator instead of about 12.
struct Recip64 {
    double b, y;
};

__device__ __forceinline__ Recip64 recip64(double b)
{
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    return {b, y};
}

__device__ __forceinline__ double div64(double a, const Recip64 &d)
{
    const double q = a * d.y;
    const double r = __builtin_fma(-d.b, q, a);
    return __builtin_fma(r, d.y, q);
}

struct Recip32 {
    float b, y;
};

__device__ __forceinline__ Recip32 recip32(float b)
{
    float y = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y, 1.0f);
    y = __builtin_fmaf(e, y, y);
    return {b, y};
}

__device__ __forceinline__ float div32(float a, const Recip32 &d)
{
    float q = a * d.y;
    float r = __builtin_fmaf(-d.b, q, a);
    q = __builtin_fmaf(r, d.y, q);
    r = __builtin_fmaf(-d.b, q, a);
    return __builtin_fmaf(r, d.y, q);
}

// True when the model's md makes (float)md a divisor the shared-reciprocal
// steps cover (md >= 1; an int is below 2^32).
static inline bool recip_md_ok(int md) { return md >= 1; }

// The two features of a neighbour: deg_feature(deg, md) and
// id_feature(v, sp) bit for bit; with kRecip through the shared reciprocals.
template <bool kRecip>
struct Features {
    Recip32 md;
    Recip64 n;
    float fmd;
    uint2 sp;
    __device__ __forceinline__ Features(float fmd_, uint2 sp_) : fmd(fmd_), sp(sp_)
    {
        if (kRecip) {
            md = recip32(fmd_);
            n = recip64((double)sp_.y);
        }
    }
    __device__ __forceinline__ float deg(uint32_t d) const
    {
        return kRecip ? div32((float)d, md) : deg_feature(d, fmd);
    }
    __device__ __forceinline__ float id(uint32_t v) const
    {
        return kRecip ? (float)div64((double)(v - sp.x), n) : id_feature(v, sp);
    }
};

__device__ __forceinline__ uint32_t word(const uint4 &w, int e)
{
    return e == 0 ? w.x : e == 1 ? w.y : e == 2 ? w.z : w.w;
}

// Ordered sums (x: degree features, y: id features) over the neighbours
// E[i], i in [beg, end), of a vertex whose graph spans sp, by one lane. Each
// step takes the 8U entries of an aligned 32U-byte span: 2U 16-byte loads of
// E and of Ed (clamped to the 16-byte word holding the last entry), then the
// adds. Entries outside [beg, end) add +0: the features are
// non-negative and the sums start at +0, never become -0, and s + (+0) = s
// for every other s, so the result is the plain ordered sum bit for bit. The
// mask is applied to the bits (a select on a loaded value lets the compiler
// turn the load into a branch around it).
template <int U, bool kRecip>
__device__ __forceinline__ float2 lane_sums(uint32_t beg, uint32_t end,
                                            const uint32_t *__restrict__ E,
                                            const uint32_t *__restrict__ Ed, float fmd, uint2 sp)
{
    float sd = 0.0f, si = 0.0f;
    if (beg >= end)
        return make_float2(sd, si);
    const Features<kRecip> ft(fmd, sp);
    const uint4 *PE = reinterpret_cast<const uint4 *>(E);
    const uint4 *PD = reinterpret_cast<const uint4 *>(Ed);
    const uint32_t qmax = (end - 1u) / 4u; // 16-byte word of the last entry
    for (uint32_t i0 = beg & ~(8u * U - 1u); i0 < end; i0 += 8u * U) {
        uint4 ie[2 * U], id[2 * U];
#pragma unroll
        for (int j = 0; j < 2 * U; j++) {
            uint32_t q = i0 / 4u + (uint32_t)j;
            q = q < qmax ? q : qmax;
            ie[j] = PE[q];
            id[j] = PD[q];
        }
#pragma unroll
        for (int t = 0; t < 8 * U; t++) {
            const uint32_t i = i0 + (uint32_t)t;
            const uint32_t keep = (i >= beg && i < end) ? ~0u : 0u;
            const float fd = ft.deg(word(id[t / 4], t % 4));
            const float fi = ft.id(word(ie[t / 4], t % 4));
            sd = sd + __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, fd) & keep);
            si = si + __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, fi) & keep);
        }
    }
    return make_float2(sd, si);
}

// LDS of pair_sums: two chunk buffers of (T - 64) * R rows per column.
template <int T, int R>
struct PairStage {
    static constexpr int kRows = (T - kWave) * R;
    float col[2][2][kRows + 4]; // [buffer][0: degree, 1: id][row]
};

// Ordered sums of the features of E[i], i in [beg, end), by a workgroup of
// T threads; returns the degree sum in lane 0 and the id sum in lane 1 of
// wave 0 (other threads return 0). Wave 0 only adds; the other T - 64
// threads load and compute the features. Chunk c holds rows
// beg + c*C + r*(T-64) + (tid-64) for r < R, C = (T-64)*R (coalesced loads).
// In step c lanes 0 and 1 add chunk c from LDS while the loaders issue chunk
// c+2's loads and write chunk c+1's features (loaded during step c-1) into
// the other buffer, so a load has a whole step to arrive. Must be called by
// every thread (it synchronises).
//
// The loaded values live across the step boundary in two register sets used
// in turn (steps go in pairs); an empty asm with a memory clobber after each
// set's loads keeps the compiler from folding the loop-carried loads into
// one load at the top of the step, which would undo the prefetch.
template <int T, int R>
__device__ __forceinline__ float pair_sums(PairStage<T, R> &st, uint32_t beg, uint32_t end,
                                           const uint32_t *__restrict__ E,
                                           const uint32_t *__restrict__ Ed, float fmd, uint2 sp)
{
    static_assert(T > kWave, "pair_sums needs loader waves besides wave 0");
    constexpr int L = T - kWave;
    constexpr uint32_t C = (uint32_t)PairStage<T, R>::kRows;
    const int tid = threadIdx.x;
    const bool adder = tid < kWave; // wave-uniform
    const int lt = tid - kWave;     // loader thread
    const uint32_t deg = end - beg;
    const uint32_t nchunk = (deg + C - 1u) / C;
    float acc = 0.0f;
    if (nchunk == 0)
        return acc;
    auto pos = [&](uint32_t c, int r) {
        return (uint64_t)beg + (uint64_t)c * C + (uint64_t)(r * L + lt);
    };
    struct Raw {
        uint32_t v[R], d[R];
    };
    // clamped to the last row, so the loads are unconditional
    auto fetch = [&](uint32_t c, Raw &x) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint64_t i = pos(c, r);
            const uint64_t k = i < end ? i : end - 1;
            x.v[r] = E[k];
            x.d[r] = Ed[k];
        }
        asm volatile("" ::: "memory");
    };
    auto put = [&](uint32_t c, int s, const Raw &x) {
#pragma unroll
        for (int r = 0; r < R; r++)
            if (pos(c, r) < end) {
                st.col[s][0][r * L + lt] = deg_feature(x.d[r], fmd);
                st.col[s][1][r * L + lt] = id_feature(x.v[r], sp);
            }
    };
    auto add = [&](uint32_t c) {
        if (tid < 2) {
            const uint32_t left = deg - c * C;
            acc = hubsum::ordered_column_sum(st.col[c & 1][tid], left < C ? left : C, acc);
        }
    };
    Raw a, b; // a: chunk c+1 in even steps, b: in odd steps
    if (!adder) {
        fetch(0, b);
        fetch(1, a);
        put(0, 0, b);
    }
    __syncthreads();
    for (uint32_t c = 0; c < nchunk; c += 2) {
        if (adder) {
            add(c);
        } else if (c + 1 < nchunk) {
            fetch(c + 2, b);
            put(c + 1, 1, a);
        }
        __syncthreads();
        if (c + 1 >= nchunk)
            break;
        if (adder) {
            add(c + 1);
        } else if (c + 2 < nchunk) {
            fetch(c + 3, a);
            put(c + 2, 0, b);
        }
        __syncthreads();
    }
    return acc;
}

} // namespace insum
