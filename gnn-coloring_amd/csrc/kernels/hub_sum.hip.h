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

// Blocked sums (the BF16 / BF16X3 layers): a row's neighbours are summed in
// blocks of kSumBlock consecutive CSR entries, each block in order from 0,
// and the block sums are added in order: acc = ((0 + B0) + B1) + ... The
// quad kernel sums a vertex's row the same way, so the bits do not depend on
// which kernel handles a vertex, and a hub's add chain is deg / kSumBlock
// long instead of deg.
constexpr int kSumBlock = 32;

template <int kRows>
struct BlockedStage {
    float v[2][kF][kRows + kPad];
    float bs[2][kRows / kSumBlock][kF]; // a chunk's block sums, by parity
};

// Sums the first `rows` entries of one staged feature column, in order,
// into acc. Blocks of 32 rows alternate between two register sets so that
// the LDS reads of the next block are in flight during the adds of the
// current one; scheduling fences keep the compiler from moving the reads
// back behind the adds (which would put every block's LDS latency on the
// add chain).
__device__ __forceinline__ float ordered_column_sum(const float *col, uint32_t rows, float acc)
{
    auto rd = [&](float4(&q)[8], uint32_t r) {
#pragma unroll
        for (int t = 0; t < 8; t++)
            q[t] = *reinterpret_cast<const float4 *>(col + r + 4 * t);
    };
    auto add = [&](const float4(&q)[8]) {
#pragma unroll
        for (int t = 0; t < 8; t++) {
            acc = acc + q[t].x;
            acc = acc + q[t].y;
            acc = acc + q[t].z;
            acc = acc + q[t].w;
        }
    };
    uint32_t r = 0;
    if (rows >= 32) {
        float4 A[8], B[8];
        rd(A, 0);
        // A holds rows [r-32, r), read but not yet added
        for (r = 32; r + 64 <= rows; r += 64) {
            rd(B, r);
            __builtin_amdgcn_sched_barrier(0);
            add(A);
            __builtin_amdgcn_sched_barrier(0);
            rd(A, r + 32);
            __builtin_amdgcn_sched_barrier(0);
            add(B);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (r + 32 <= rows) {
            rd(B, r);
            __builtin_amdgcn_sched_barrier(0);
            add(A);
            add(B);
            r += 32;
        } else {
            add(A);
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
// Step c: issue the row loads of chunk c+1 (its index was loaded two steps
// earlier) and the index load of chunk c+3, sum chunk c from LDS, then write
// chunk c+1's rows into the other LDS buffer and synchronise. The summing
// wave waits for at most one memory round trip per chunk, overlapped with
// its own sum.
// Sum of column f's rows [r0, r1) of one staged chunk, in order from 0.
__device__ __forceinline__ float block_column_sum(const float *col, uint32_t r0, uint32_t r1)
{
    float acc = 0.0f;
    uint32_t r = r0;
    if (r1 - r0 == (uint32_t)kSumBlock) {
        float4 q[kSumBlock / 4];
#pragma unroll
        for (int t = 0; t < kSumBlock / 4; t++)
            q[t] = *reinterpret_cast<const float4 *>(col + r0 + 4 * t);
#pragma unroll
        for (int t = 0; t < kSumBlock / 4; t++) {
            acc = acc + q[t].x;
            acc = acc + q[t].y;
            acc = acc + q[t].z;
            acc = acc + q[t].w;
        }
        return acc;
    }
    for (; r < r1; r++)
        acc = acc + col[r];
    return acc;
}

template <int kRows, bool pair, bool blocked = false, typename StageT, typename Load>
__device__ __forceinline__ float ordered_sum(StageT &st, uint32_t beg, uint32_t end,
                                             const uint32_t *__restrict__ Er, Load load)
{
    static_assert(!blocked || (!pair && kRows % kSumBlock == 0 && kRows >= kF * 1),
                  "blocked sums: 16 features, whole blocks per chunk");
    constexpr uint32_t kBlocks = blocked ? kRows / kSumBlock : 1;
    constexpr int Q = pair ? 1 : 4;
    const int tid = threadIdx.x;
    const uint32_t deg = end - beg;
    const uint32_t nchunk = (deg + kRows - 1) / kRows;
    const bool summer = tid < (pair ? 2 : kF);
    float acc = 0.0f;
    if (nchunk == 0)
        return acc;

    // 64-bit positions: the index loads look three chunks past the end. Index
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

    // Index loads run two chunks ahead of the row loads (ia: the next odd
    // or even chunk, ib: the one after): under a saturated HBM a load takes
    // longer than one chunk's sum. Steps go in pairs so the two index
    // registers keep fixed roles.
    float4 row[Q];
    load(fetch_idx(0), row);
    uint32_t ia = fetch_idx(1), ib = fetch_idx(2);
    put(0, 0, row);
    __syncthreads();
    // The loads of the step after the last chunk are made anyway (clamped,
    // in bounds) and not staged: a branch around them would cost the exact
    // wait counts.
    auto step = [&](uint32_t c, uint32_t &inext) {
        const int s = c & 1;
        load(inext, row);
        inext = fetch_idx(c + 3);
        const uint32_t left = deg - c * kRows;
        const uint32_t rows = left < (uint32_t)kRows ? left : kRows;
        if constexpr (blocked) {
            // thread (block b, feature f) sums chunk c's block b of column f;
            // the summers add chunk c-1's block sums (full chunk: kBlocks),
            // written before the previous step's barrier
            const uint32_t b = (uint32_t)tid / kF, fe = (uint32_t)tid % kF;
            if (b < kBlocks && b * kSumBlock < rows) {
                const uint32_t r1 = (b + 1) * kSumBlock < rows ? (b + 1) * kSumBlock : rows;
                st.bs[s][b][fe] = block_column_sum(st.v[s][fe], b * kSumBlock, r1);
            }
            if (summer && c > 0) {
#pragma unroll
                for (uint32_t q = 0; q < kBlocks; q++)
                    acc = acc + st.bs[s ^ 1][q][tid];
            }
        } else if (summer) {
            acc = ordered_column_sum(st.v[s][tid], rows, acc);
        }
        put(c + 1, s ^ 1, row); // no-op past the last row
        __syncthreads();
    };
    for (uint32_t c = 0; c < nchunk; c += 2) {
        step(c, ia); // chunk c+1's index is in ia, chunk c+3's goes there
        if (c + 1 < nchunk)
            step(c + 1, ib);
    }
    if constexpr (blocked) {
        // the last chunk's block sums (visible after the last step's barrier)
        if (summer) {
            const uint32_t c = nchunk - 1, left = deg - c * kRows;
            const uint32_t nb = (left + kSumBlock - 1) / kSumBlock;
            for (uint32_t q = 0; q < nb; q++)
                acc = acc + st.bs[c & 1][q][tid];
        }
    }
    return acc;
}

} // namespace hubsum
