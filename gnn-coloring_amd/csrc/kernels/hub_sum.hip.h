// hub_sum.hip.h -- ordered neighbour sums of one hub vertex per workgroup.
//
// The reference sums a vertex's neighbour features left to right in FP32
// (src/fast_gnn.c:13-47, 49-89), so a hub's sum is one dependent add chain
// per feature and cannot be split across lanes without changing the bits.
// What can be overlapped is the gather that feeds the chain: the workgroup
// stages chunks of kRows neighbour rows feature-major in LDS, one row per
// thread, and 16 lanes of wave 0 (lane j: feature j) walk the staged chunk
// in order. The next chunk's rows are in flight while the current chunk is
// summed; two LDS chunk buffers alternate, one summed while the other is
// filled.
#pragma once
#include "common.hip.h"

namespace hubsum {

constexpr int kF = 16;
constexpr int kPad = 4; // LDS row padding: the 16 summing lanes' 16-byte reads spread over banks

// Two chunk buffers of kRows rows each; kRows = threads per workgroup.
template <int kRows>
struct Stage {
    float v[2][kF][kRows + kPad];
};

// Sums the first `rows` entries of one staged feature column, in order,
// into acc: 32 rows of LDS reads are issued ahead of their 32 adds.
__device__ __forceinline__ float ordered_column_sum(const float *col, uint32_t rows, float acc)
{
    uint32_t r = 0;
    for (; r + 32 <= rows; r += 32) {
        float4 q[8];
#pragma unroll
        for (int t = 0; t < 8; t++)
            q[t] = *reinterpret_cast<const float4 *>(col + r + 4 * t);
#pragma unroll
        for (int t = 0; t < 8; t++) {
            acc = acc + q[t].x;
            acc = acc + q[t].y;
            acc = acc + q[t].z;
            acc = acc + q[t].w;
        }
    }
    for (; r < rows; r++)
        acc = acc + col[r];
    return acc;
}

// Ordered sums of the rows x[Er[beg..end)] (as returned by `load`) for the
// calling workgroup of kRows threads (rows per chunk). `pair` = true: rows are two features
// (load fills row[0].x, .y) and lanes 0, 1 sum them; else 16 features
// (row[0..3]) and lane j < 16 sums feature j. Returns the sum in lanes
// 0..15 (0..1) of wave 0; other threads return 0. Must be called by every
// thread of the workgroup (it synchronises).
//
// Step c: issue the row loads of chunk c+1 (its index arrived during step
// c-1) and the index load of chunk c+2, sum chunk c from LDS, then write
// chunk c+1's rows into the other LDS buffer and synchronise. The summing
// wave waits for at most one memory round trip per chunk, overlapped with
// its own sum.
template <int kRows, bool pair, typename Load>
__device__ __forceinline__ float ordered_sum(Stage<kRows> &st, uint32_t beg, uint32_t end,
                                             const uint32_t *__restrict__ Er, Load load)
{
    constexpr int Q = pair ? 1 : 4;
    const int tid = threadIdx.x;
    const uint32_t deg = end - beg;
    const uint32_t nchunk = (deg + kRows - 1) / kRows;
    const bool summer = tid < (pair ? 2 : kF);
    float acc = 0.0f;
    if (nchunk == 0)
        return acc;

    // 64-bit positions: the index load looks two chunks past the end. Index
    // loads are unconditional, clamped to the last row (deg >= 1 here), so
    // no branch hides them from the compiler's in-order wait counting.
    auto pos = [&](uint32_t c) { return (uint64_t)beg + (uint64_t)c * kRows + (uint64_t)tid; };
    auto fetch_idx = [&](uint32_t c) {
        const uint64_t i = pos(c);
        return Er[i < end ? i : end - 1];
    };
    auto put = [&](uint32_t c, int s, const float4 *src) {
        if (pos(c) < end) {
            if (pair) {
                st.v[s][0][tid] = src[0].x;
                st.v[s][1][tid] = src[0].y;
            } else {
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    st.v[s][4 * q + 0][tid] = src[q].x;
                    st.v[s][4 * q + 1][tid] = src[q].y;
                    st.v[s][4 * q + 2][tid] = src[q].z;
                    st.v[s][4 * q + 3][tid] = src[q].w;
                }
            }
        }
    };

    float4 row[Q];
    load(fetch_idx(0), row);
    uint32_t idx = fetch_idx(1);
    put(0, 0, row);
    __syncthreads();
    for (uint32_t c = 0; c < nchunk; c++) {
        const int s = c & 1;
        const bool more = c + 1 < nchunk;
        if (more) {
            load(idx, row);
            idx = fetch_idx(c + 2);
        }
        if (summer) {
            const uint32_t left = deg - c * kRows;
            acc = ordered_column_sum(st.v[s][tid], left < (uint32_t)kRows ? left : kRows, acc);
        }
        if (more)
            put(c + 1, s ^ 1, row);
        __syncthreads();
    }
    return acc;
}

} // namespace hubsum
