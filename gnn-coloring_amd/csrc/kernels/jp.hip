// jp.hip -- Jones-Plassmann colouring, proper-colouring check and the
// greedy-rule certificate for gfx950.
//
// Semantics (reference coloring.c:223-330): v precedes u iff
// p[v] > p[u] || (p[v] == p[u] && v < u); each vertex takes the smallest
// colour >= 1 unused by its predecessors. The result equals sequential greedy
// in (p descending, id ascending) order, so it does not depend on how the
// work below is scheduled.
//
// Orientation (init): npred[u] = number of predecessors, and u's successors
// fill E_split[V[u] + npred[u] .. V[u+1]) (the successor part of the split
// that coloring.c:241-264 builds). The predecessor part is never read --
// colours arrive through the state words -- so it is not written.
//
// Per-vertex state word: state[v] = (pending predecessors << mb) | colour
// bits, bit c-1 set when a predecessor took colour c (1 <= c <= mb), where
// mb = 64 - bitlength(max degree) leaves just enough high bits for the count.
// Colours mb+1 .. deg+1 live in per-vertex overflow words at word offset
// V[v]/32 + v. When u takes colour c <= mb it ORs bit c-1 into state[v] of
// each successor and then subtracts 1 << mb from the same word: both are
// atomics on one location, so the value a consumer reads with count zero
// already holds every predecessor's bit, and the consumer's single atomic
// load yields both "ready" and the colour set. For c > mb the overflow OR is
// waited for (s_waitcnt) before the decrement.
//
// A vertex with s successors is ceil(s/64) work items (at least one); one
// wave handles one item, i.e. up to 64 successor edges, so a hub's
// successors are spread over many waves. Every item of u recomputes the
// (identical) mex; item 0 stores the colour. Vertices of degree 0 take
// colour 1 at init and have no items.
//
// Two schedules, same colours:
//   rounds  -- one launch per DAG level; the frontier is an item queue; the
//              state word above (its returned old value tells a producer that
//              it released a successor).
//   ordered -- items sorted in greedy order; one persistent launch whose
//              wave w takes items w, w + W, w + 2W, ... and waits on each
//              item's state. The earliest unfinished item is always ready
//              and its wave is at it, so the launch completes when all W
//              waves are resident; a wall-clock bound on every wait ends the
//              kernel with ctl->err set if they are not. Here each vertex has
//              up to 8 state words, each with its own report counter (see
//              jp_ordered_kernel), so colours up to 8*mb need no overflow
//              words and no wait between a report's OR and its count.
#include "common.hip.h"

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kGroup = 16;      // lanes per vertex in whole-graph passes
constexpr uint32_t kChunk = 64; // successor edges per work item
constexpr int kRootBlock = 1024;

__device__ __forceinline__ uint32_t ovf_base(uint32_t Vv, uint32_t v) { return Vv / 32u + v; }

// Overflow words of a vertex of degree d whose state words hold colours
// 1 .. cw: colours cw+1 .. d+1.
__device__ __forceinline__ uint32_t ovf_words(uint32_t d, uint32_t cw)
{
    return d + 1u > cw ? (d + 1u - cw + 31u) / 32u : 0u;
}

__device__ __forceinline__ uint32_t n_items(uint32_t nsucc)
{
    return nsucc ? (nsucc + kChunk - 1u) / kChunk : 1u;
}

__device__ __forceinline__ uint64_t load_state(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t load_word(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Smallest colour > cw not in u's overflow words. Wave uniform (all lanes call
// with the same u). mex <= npred + 1 <= deg + 1, so a free bit exists.
__device__ __forceinline__ uint32_t mex_overflow(uint32_t cw, uint32_t u, uint32_t Vu, uint32_t deg,
                                                 const uint32_t *__restrict__ ovf)
{
    const uint32_t base = ovf_base(Vu, u), words = ovf_words(deg, cw);
    const int lane = lane_id();
    for (uint32_t wb = 0; wb < words; wb += kWave) {
        const uint32_t wi = wb + (uint32_t)lane;
        const uint32_t w = wi < words ? load_word(&ovf[base + wi]) : 0xFFFFFFFFu;
        const uint64_t b = __ballot(w != 0xFFFFFFFFu);
        if (b) {
            const int fl = __ffsll((long long)b) - 1;
            const uint32_t wf = __shfl(w, fl);
            return cw + 1u + 32u * (wb + (uint32_t)fl) + (uint32_t)__builtin_ctz(~wf);
        }
    }
    return 0u; // unreachable for a consistent state
}

// Smallest colour not in u's set (one state word s with colours 1..mb).
__device__ __forceinline__ uint32_t mex_of(uint64_t s, int mb, uint32_t u, uint32_t Vu,
                                           uint32_t deg, const uint32_t *__restrict__ ovf)
{
    const uint64_t free = ~s & ((1ull << mb) - 1ull);
    if (free)
        return (uint32_t)__builtin_ctzll(free) + 1u;
    return mex_overflow((uint32_t)mb, u, Vu, deg, ovf);
}

// Record colour c > cw in successor v's overflow words (colours past v's
// degree + 1 can never be v's mex and are dropped).
__device__ __forceinline__ void or_overflow(const uint32_t *__restrict__ V, uint32_t *ovf,
                                            uint32_t v, uint32_t c, uint32_t cw)
{
    const uint32_t Vv = V[v], b = c - cw - 1u;
    if (b < 32u * ovf_words(V[v + 1] - Vv, cw))
        atomicOr(&ovf[ovf_base(Vv, v) + b / 32u], 1u << (b % 32u));
}

// Wave-wide: lanes with `want` append items (v, 0..cnt-1) to q. All lanes call.
__device__ __forceinline__ void push_items(bool want, uint32_t v, uint32_t cnt, uint2 *q,
                                           uint32_t *n)
{
    if (__ballot(want) == 0)
        return;
    const int lane = lane_id();
    const uint32_t mine = want ? cnt : 0u;
    uint32_t incl = mine;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off);
        if (lane >= off)
            incl += y;
    }
    const uint32_t total = __shfl(incl, kWave - 1);
    uint32_t base = 0;
    if (lane == kWave - 1)
        base = atomicAdd(n, total);
    base = __shfl(base, kWave - 1);
    const uint32_t start = base + incl - mine;
    for (uint32_t k = 0; k < mine; k++)
        q[start + k] = make_uint2(v, k);
}

// Zero the overflow words, the per-hub positions and the control block.
// (State words and predecessor counts are written by the orientation pass.)
// A plain kernel, so the sequence can be captured in a graph.
__global__ __launch_bounds__(kBlock) void jp_clear(uint64_t nwords, uint32_t *__restrict__ ovf,
                                                   uint32_t nhub, uint32_t *__restrict__ hpos,
                                                   const uint32_t *__restrict__ hubs,
                                                   uint32_t *__restrict__ npred,
                                                   uint64_t *__restrict__ state,
                                                   int *__restrict__ color, gcxk_ctl *ctl)
{
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    const uint64_t t0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    for (uint64_t i = t0; i < nwords; i += stride)
        ovf[i] = 0u;
    for (uint64_t h = t0; h < 2ull * nhub; h += stride)
        hpos[h] = 0u;
    // the hubs' counts start at zero: jp_init_hub_chunks adds into them,
    // possibly while jp_init_quads runs on another stream
    for (uint64_t h = t0; h < nhub; h += stride) {
        const uint32_t u = hubs[h];
        color[u] = 0;
        npred[u] = 0u;
        state[u] = 0ull;
    }
    if (t0 < sizeof(gcxk_ctl) / 4)
        reinterpret_cast<uint32_t *>(ctl)[t0] = 0u;
}

// Orientation in degree order (perm), 4 lanes per vertex so a wave's 16
// vertices have near-equal degrees: successors go to the back of the row in
// E_split. Also writes every non-hub vertex's colour slot (1 for degree 0,
// else 0), npred and state[u] = npred << mb. The hubs (perm[0 .. nhub)) are
// left alone: jp_clear zeroed them and jp_init_hub_chunks places their
// successors and adds their counts.
constexpr int kQuad = 4;
// 64 edges in flight per vertex (16 per lane; 62 / 82 VGPRs for 32- / 64-bit
// keys). R-MAT 22 first-fit colouring (push form) with 4 / 16 / 32 steps:
// 8.42 / 8.38 / 8.38 ms
constexpr int kQuadSteps = 16;

template <typename K>
__global__ __launch_bounds__(kBlock) void jp_init_quads(uint32_t N, const uint32_t *__restrict__ V,
                                                        const uint32_t *__restrict__ E,
                                                        const K *__restrict__ key,
                                                        const uint32_t *__restrict__ perm,
                                                        uint32_t nhub,
                                                        uint32_t *__restrict__ npred,
                                                        uint32_t *__restrict__ esplit,
                                                        uint64_t *__restrict__ state,
                                                        int *__restrict__ color, int mb,
                                                        uint32_t head_max)
{
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint32_t pi = (uint32_t)(t / kQuad);
    const int gl = (int)(t % kQuad);
    const int gsh = (lane_id() / kQuad) * kQuad; // first lane of the quad
    if (pi >= N) // whole quads leave together
        return;
    if (pi < nhub)
        return;
    const uint32_t u = perm[pi];
    const uint32_t beg = V[u], end = V[u + 1];
    uint32_t np = 0, ns = 0;
    if (end > beg) {
        const K ku = key[u];
        // rank keys: the predecessors of a head-prefix vertex go to the front
        // of its row as greedy-order positions (N - 1 - key)
        const bool emit = head_max != 0u && N - 1u - (uint32_t)ku < head_max;
        const uint32_t below = (1u << gl) - 1u;
        for (uint32_t i0 = beg; i0 < end; i0 += kQuad * kQuadSteps) {
            uint32_t v[kQuadSteps];
            K kv[kQuadSteps];
            bool in[kQuadSteps], pr[kQuadSteps];
#pragma unroll
            for (int s = 0; s < kQuadSteps; s++) {
                const uint32_t i = i0 + kQuad * s + gl;
                in[s] = i < end;
                v[s] = in[s] ? E[i] : 0u;
            }
#pragma unroll
            for (int s = 0; s < kQuadSteps; s++) {
                kv[s] = in[s] ? key[v[s]] : K(0);
                pr[s] = in[s] && precedes(kv[s], v[s], ku, u);
            }
#pragma unroll
            for (int s = 0; s < kQuadSteps; s++) {
                const uint32_t gp = (uint32_t)(__ballot(pr[s]) >> gsh) & 0xFu;
                const uint32_t gs = (uint32_t)(__ballot(in[s] && !pr[s]) >> gsh) & 0xFu;
                if (in[s] && !pr[s])
                    esplit[end - 1u - ns - __popc(gs & below)] = v[s];
                if (emit && pr[s])
                    esplit[beg + np + __popc(gp & below)] = N - 1u - (uint32_t)kv[s];
                np += __popc(gp);
                ns += __popc(gs);
            }
        }
    }
    if (gl == 0) {
        color[u] = end == beg ? 1 : 0;
        npred[u] = np;
        state[u] = (uint64_t)np << mb;
    }
}

// Hubs: chunk b of hub h covers edges [V[u] + b*kHubEdges, ...); successor
// positions in E_split come from a per-hub counter (filled from the back of
// the row; their order is irrelevant) and predecessors are only counted.
constexpr uint32_t kHubEdges = GCXK_HUB_CHUNK_EDGES;

constexpr int kHubPer = (int)(kHubEdges / kBlock); // entries per thread of a chunk

template <typename K>
__global__ __launch_bounds__(kBlock) void jp_init_hub_chunks(const uint32_t *__restrict__ V,
                                                             const uint32_t *__restrict__ E,
                                                             const K *__restrict__ key,
                                                             uint32_t *__restrict__ npred,
                                                             uint64_t *__restrict__ state,
                                                             uint32_t *__restrict__ esplit,
                                                             const uint32_t *__restrict__ hubs,
                                                             const uint32_t *__restrict__ hchunk,
                                                             uint32_t nhub,
                                                             uint32_t *__restrict__ hpos, int mb,
                                                             uint32_t N, uint32_t head_max)
{
    __shared__ uint32_t s_cnt[3][kWavesPerBlock]; // [1] successors, [2] predecessors
    __shared__ uint32_t s_base, s_pbase;
    // hub of this block: last h with hchunk[h] <= blockIdx.x
    uint32_t lo = 0, hi = nhub;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (hchunk[mid] <= blockIdx.x)
            lo = mid;
        else
            hi = mid;
    }
    const uint32_t h = lo, u = hubs[h];
    const uint32_t beg = V[u], end = V[u + 1];
    const uint32_t cb = beg + (blockIdx.x - hchunk[h]) * kHubEdges;
    const uint32_t ce = cb + kHubEdges < end ? cb + kHubEdges : end;
    const K ku = key[u];
    const int lane = lane_id(), w = threadIdx.x / kWave;
    // all of the chunk's loads in flight, then one reservation per block
    uint32_t v[kHubPer];
    K kv[kHubPer];
    bool in[kHubPer], sc[kHubPer];
#pragma unroll
    for (int q = 0; q < kHubPer; q++) {
        const uint32_t i = cb + (uint32_t)q * kBlock + threadIdx.x;
        in[q] = i < ce;
        v[q] = in[q] ? E[i] : 0u;
    }
    uint32_t np = 0, ns = 0;
#pragma unroll
    for (int q = 0; q < kHubPer; q++) {
        kv[q] = in[q] ? key[v[q]] : K(0);
        sc[q] = in[q] && !precedes(kv[q], v[q], ku, u);
        np += (in[q] && !sc[q]) ? 1u : 0u;
        ns += sc[q] ? 1u : 0u;
    }
    // rank keys: a head-prefix hub also writes its predecessors' greedy-order
    // positions (N - 1 - key) to the front of its row
    const bool emit = head_max != 0u && N - 1u - (uint32_t)ku < head_max;
    uint32_t is = ns, ip = np; // inclusive scans of the counts over the wave
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t a = __shfl_up(is, off), b = __shfl_up(ip, off);
        if (lane >= off) {
            is += a;
            ip += b;
        }
    }
    if (lane == kWave - 1) {
        s_cnt[1][w] = is;
        s_cnt[2][w] = ip;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tp = 0, ts = 0;
        for (int q = 0; q < kWavesPerBlock; q++) {
            const uint32_t b = s_cnt[1][q], c = s_cnt[2][q];
            s_cnt[1][q] = ts;
            s_cnt[2][q] = tp;
            ts += b;
            tp += c;
        }
        s_base = ts ? atomicAdd(&hpos[2 * h + 1], ts) : 0u;
        s_pbase = (tp && emit) ? atomicAdd(&hpos[2 * h], tp) : 0u;
        if (tp) {
            atomicAdd(&npred[u], tp);
            atomicAdd((unsigned long long *)&state[u], (unsigned long long)tp << mb);
        }
    }
    __syncthreads();
    // successors fill the row from its back, predecessor positions (head
    // prefix only) from its front; the order within each part is irrelevant
    uint32_t pos = s_base + s_cnt[1][w] + is - ns;
#pragma unroll
    for (int q = 0; q < kHubPer; q++)
        if (sc[q])
            esplit[end - 1u - pos++] = v[q];
    if (emit) {
        uint32_t pp = s_pbase + s_cnt[2][w] + ip - np;
#pragma unroll
        for (int q = 0; q < kHubPer; q++)
            if (in[q] && !sc[q])
                esplit[beg + pp++] = N - 1u - (uint32_t)kv[q];
    }
}

// ------------------------------------------------------------ round schedule

// Round-0 queue: every vertex with successors and no predecessors. One queue
// reservation per workgroup.
__global__ __launch_bounds__(kRootBlock) void jp_collect_roots(uint32_t N,
                                                               const uint32_t *__restrict__ V,
                                                               const uint32_t *__restrict__ npred,
                                                               uint2 *__restrict__ q0,
                                                               gcxk_ctl *ctl)
{
    __shared__ uint32_t wsum[kRootBlock / kWave];
    __shared__ uint32_t bbase;
    const uint32_t u = blockIdx.x * kRootBlock + threadIdx.x;
    const int lane = lane_id(), w = threadIdx.x / kWave;
    uint32_t cnt = 0;
    if (u < N) {
        const uint32_t d = V[u + 1] - V[u], np = npred[u];
        if (d > 0 && np == 0)
            cnt = n_items(d);
    }
    uint32_t incl = cnt;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off);
        if (lane >= off)
            incl += y;
    }
    if (lane == kWave - 1)
        wsum[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int i = 0; i < kRootBlock / kWave; i++) {
            const uint32_t x = wsum[i];
            wsum[i] = tot;
            tot += x;
        }
        bbase = tot ? atomicAdd(&ctl->qn[0], tot) : 0u;
    }
    __syncthreads();
    const uint32_t start = bbase + wsum[w] + incl - cnt;
    for (uint32_t k = 0; k < cnt; k++)
        q0[start + k] = make_uint2(u, k);
}

__global__ __launch_bounds__(kBlock) void jp_round_kernel(const uint32_t *__restrict__ V,
                                                          const uint32_t *__restrict__ npred,
                                                          const uint32_t *__restrict__ esplit,
                                                          uint64_t *__restrict__ state,
                                                          int *__restrict__ color,
                                                          uint32_t *__restrict__ ovf,
                                                          const uint2 *__restrict__ qin,
                                                          uint2 *__restrict__ qout, gcxk_ctl *ctl,
                                                          uint32_t r, int mb)
{
    const uint32_t si = r % 3, so = (r + 1) % 3, sr = (r + 2) % 3;
    const uint32_t n = ctl->qn[si];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // Slot sr held round r-1's input; round r+1 appends to it.
        ctl->qn[sr] = 0;
        if (n != 0)
            ctl->rounds = r + 1;
        else
            ctl->done = 1;
    }
    if (n == 0)
        return;
    const int lane = lane_id();
    const uint64_t one = 1ull << mb;
    const uint32_t w0 = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    for (uint32_t idx = w0; idx < n; idx += nw) {
        const uint2 it = qin[idx];
        const uint32_t u = it.x, k = it.y;
        const uint32_t Vu = V[u], Vu1 = V[u + 1];
        const uint32_t np = npred[u];
        // u's count reached zero in an earlier launch: a plain load is current
        const uint32_t c = mex_of(state[u], mb, u, Vu, Vu1 - Vu, ovf);
        if (k == 0 && lane == 0)
            color[u] = (int)c;
        const uint32_t i = Vu + np + k * kChunk + (uint32_t)lane;
        const bool has = i < Vu1;
        const uint32_t v = has ? esplit[i] : 0u;
        bool rel = false;
        if (c <= (uint32_t)mb) {
            if (has) {
                atomicOr((unsigned long long *)&state[v], 1ull << (c - 1u));
                const uint64_t old = atomicAdd((unsigned long long *)&state[v], 0ull - one);
                rel = (old >> mb) == 1ull;
            }
        } else {
            if (has)
                or_overflow(V, ovf, v, c, mb);
            __builtin_amdgcn_s_waitcnt(0);
            if (has)
                rel = (atomicAdd((unsigned long long *)&state[v], 0ull - one) >> mb) == 1ull;
        }
        uint32_t cnt = 0;
        if (rel)
            cnt = n_items(V[v + 1] - V[v] - npred[v]);
        push_items(rel, v, cnt, qout, &ctl->qn[so]);
    }
}

// ----------------------------------------------------------- ordered schedule

// Sort input: complemented keys (so an ascending stable sort yields p
// descending, id ascending) and the ids.
template <typename K>
__global__ __launch_bounds__(kBlock) void jp_order_keys(uint32_t N, const K *__restrict__ key,
                                                        K *__restrict__ nkey,
                                                        uint32_t *__restrict__ ids)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N) {
        nkey[i] = ~key[i];
        ids[i] = i;
    }
}

// Items per sorted position (0 for degree-0 vertices).
__global__ __launch_bounds__(kBlock) void jp_item_counts(uint32_t N, const uint32_t *__restrict__ V,
                                                         const uint32_t *__restrict__ npred,
                                                         const uint32_t *__restrict__ ids,
                                                         uint32_t *__restrict__ cnt)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N) {
        const uint32_t u = ids[i];
        const uint32_t d = V[u + 1] - V[u];
        cnt[i] = d ? n_items(d - npred[u]) : 0u;
    }
}

// Items in greedy order; incl = inclusive prefix of the counts.
__global__ __launch_bounds__(kBlock) void jp_expand(uint32_t N, const uint32_t *__restrict__ ids,
                                                    const uint32_t *__restrict__ incl,
                                                    uint2 *__restrict__ items, gcxk_ctl *ctl)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N)
        return;
    const uint32_t end = incl[i], beg = i ? incl[i - 1] : 0u;
    const uint32_t u = ids[i];
    for (uint32_t j = beg; j < end; j++)
        items[j] = make_uint2(u, j - beg);
    if (i == N - 1)
        ctl->total_items = end;
}

// State words of the ordered schedule. A vertex with np predecessors uses one
// of two layouts (W words each, wn[N] = 0 closes the scan):
//  counting fields (np <= 255): fields of B bits (B = 4 for np <= 15, else
//    8), F = 64 / B per word; field f counts the predecessors with colour
//    f + 1, the last field of the last word those with any colour past the
//    fields. W = ceil((np + 2) / F), so every colour v can take (mex <=
//    np + 1) has its own field, and no field can exceed np < 2^B;
//  bit sets (np > 255): word i = (reports << mb) | bits of colours
//    i*mb+1 .. (i+1)*mb, W = min(8, ceil((np + 1) / mb)).
constexpr uint32_t kMaxWords = 8;
constexpr uint32_t kFieldMaxPred = 255;

__device__ __forceinline__ uint32_t field_bits(uint32_t np)
{
    return np <= 15u ? 4u : (np <= kFieldMaxPred ? 8u : 0u);
}

__device__ __forceinline__ uint32_t state_words(uint32_t np, int mb)
{
    const uint32_t B = field_bits(np);
    if (B)
        return (np + 2u + 64u / B - 1u) / (64u / B);
    const uint32_t w = (np + 1u + (uint32_t)mb - 1u) / (uint32_t)mb;
    return w < kMaxWords ? w : kMaxWords;
}

__global__ __launch_bounds__(kBlock) void jp_word_counts(uint32_t N, const uint32_t *__restrict__ npred,
                                                         int mb, uint32_t *__restrict__ wn)
{
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    if (v < N)
        wn[v] = state_words(npred[v], mb);
    else if (v == N)
        wn[v] = 0u;
}

// Per-vertex descriptor read by the ordered kernel (one 8-byte load instead
// of two offsets): first word (bits 0-39), word count (40-47), field width
// B (48-55; 0 = bit-set words), and bit 56 set when the head kernel colours
// the vertex (rank keys rk given and its greedy-order position N - 1 - rk[v]
// below ctl->head_n).
constexpr int kHeadFlagBit = 56;

__global__ __launch_bounds__(kBlock) void jp_word_desc(uint32_t N, const uint32_t *__restrict__ npred,
                                                       const uint32_t *__restrict__ wof,
                                                       const uint32_t *__restrict__ rk,
                                                       const gcxk_ctl *__restrict__ ctl,
                                                       uint64_t *__restrict__ wdesc)
{
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    if (v < N) {
        const uint64_t n = wof[v + 1] - wof[v];
        const bool head = rk && N - 1u - rk[v] < ctl->head_n;
        wdesc[v] = (uint64_t)wof[v] | (n << 40) | ((uint64_t)field_bits(npred[v]) << 48) |
                   ((uint64_t)head << kHeadFlagBit);
    }
}

// Rank keys: rk[ids[i]] = N - 1 - i for the vertices in greedy order (ids), so
// larger = earlier and no two are equal.
__global__ __launch_bounds__(kBlock) void jp_rank_keys(uint32_t N, const uint32_t *__restrict__ ids,
                                                       uint32_t *__restrict__ rk)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N)
        rk[ids[i]] = N - 1u - i;
}

// ------------------------------------------------------------- head kernel
//
// The top of the greedy order of a skewed graph is a dense core whose
// vertices form most of the DAG's depth: on R-MAT scale 22 (GNN3 order) the
// first 32,768 positions hold 2,516 of the 2,741 levels of the longest
// chain, each of their vertices has at most ~1000 predecessors, and every
// predecessor of a vertex in such a prefix lies in the prefix too. One
// workgroup colours that prefix with the colours in LDS: a hop of the chain
// is then an LDS store seen by an LDS poll instead of a device-scope atomic
// seen by a device-scope poll. The rest of the graph follows in the ordered
// kernel, which takes the head's colours as given.
// Waits read the real-time clock (for their 2 s bound) only from their 64th
// poll on: s_memrealtime is a scalar memory read, and the next LDS or
// scalar access waits for it (one shared counter), which put the clock's
// latency on every hop of the chain when the first poll read it.
constexpr uint32_t kClockPoll = 64;

constexpr int kHeadBlock = 1024;
constexpr int kHeadWaves = kHeadBlock / kWave;
constexpr uint32_t kHeadMax = 32768;          // prefix positions (32-bit colours in LDS)
constexpr uint32_t kHeadWin = 4096;           // colours per mex window
constexpr uint32_t kHeadWinWords = kHeadWin / 32;
constexpr int kHeadSlots = 16;                // predecessor positions per lane and batch
constexpr uint32_t kHeadBatch = kHeadSlots * kWave;
constexpr uint32_t kHeadBlockPos = 1024;      // positions per density block
constexpr uint32_t kHeadWaveWords = kHeadWinWords + kWave; // bitmap + a word per lane
constexpr size_t kHeadLds = (size_t)kHeadWaves * kHeadWaveWords * 4 + 4 * kHeadMax + 4;

// Head extent: the leading 1024-position blocks of the greedy order whose
// mean predecessor count is at least min_density (a dense core), at most
// head_max positions; ctl->head_n = their total (0: no head).
__global__ __launch_bounds__(kHeadBlock) void jp_head_extent(uint32_t N, const uint32_t *__restrict__ ids,
                                                             const uint32_t *__restrict__ npred,
                                                             uint32_t head_max,
                                                             uint32_t min_density, gcxk_ctl *ctl)
{
    __shared__ uint32_t part[kHeadWaves];
    __shared__ uint32_t s_go;
    const uint32_t lim = N < head_max ? N : head_max;
    const int w = threadIdx.x / kWave;
    uint32_t P = 0;
    for (uint32_t b = 0; b < lim; b += kHeadBlockPos) {
        const uint32_t i = b + threadIdx.x;
        uint32_t x = i < lim ? npred[ids[i]] : 0u;
        x = group_sum32<kWave>(x);
        if (lane_id() == 0)
            part[w] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t sum = 0;
            for (int q = 0; q < kHeadWaves; q++)
                sum += part[q];
            const uint32_t n = lim - b < kHeadBlockPos ? lim - b : kHeadBlockPos;
            s_go = sum >= (uint64_t)min_density * n;
        }
        __syncthreads();
        if (!s_go)
            break;
        P = b + kHeadBlockPos < lim ? b + kHeadBlockPos : lim;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        ctl->head_n = P;
}

// Head table: hdr[i] = (first entry of the row, predecessor count) of the
// vertex at greedy-order position i < ctl->head_n, so the head kernel reads
// one record per position instead of ids -> V and npred.
__global__ __launch_bounds__(kBlock) void jp_head_table(const uint32_t *__restrict__ V,
                                                        const uint32_t *__restrict__ npred,
                                                        const uint32_t *__restrict__ ids,
                                                        const gcxk_ctl *__restrict__ ctl,
                                                        uint2 *__restrict__ hdr)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < ctl->head_n) {
        const uint32_t u = ids[i];
        hdr[i] = make_uint2(V[u], npred[u]);
    }
}

// A head position's work: its vertex, row start, predecessor count and the
// first batch of its predecessor positions (registers).
struct head_vtx {
    uint32_t u, base, np;
    uint32_t pos[kHeadSlots];
};

struct head_hdr {
    uint32_t u, base, np;
};

__device__ __forceinline__ head_hdr head_fetch_hdr(uint32_t i, uint32_t P,
                                                   const uint32_t *__restrict__ ids,
                                                   const uint2 *__restrict__ hdr)
{
    head_hdr h = {0u, 0u, 0u};
    if (i < P) {
        const uint2 r = hdr[i];
        h.u = ids[i];
        h.base = r.x;
        h.np = r.y;
    }
    return h;
}

__device__ __forceinline__ void head_take(head_vtx &x, const head_hdr &h)
{
    x.u = h.u;
    x.base = h.base;
    x.np = h.np;
}

// Loads are unconditional (slots past the list re-read its last entry, and a
// vertex without predecessors reads entry 0), so the number of loads in
// flight is the same on every path and the compiler can wait for exactly
// the older ones; slots past the list are masked out by the caller.
__device__ __forceinline__ void head_fetch_list(head_vtx &x, uint32_t b0,
                                                const uint32_t *__restrict__ esplit, int lane)
{
    const uint32_t last = x.np ? x.base + x.np - 1u : 0u;
#pragma unroll
    for (int t = 0; t < kHeadSlots; t++) {
        const uint32_t j = x.base + b0 + (uint32_t)(kWave * t + lane);
        x.pos[t] = esplit[j < last ? j : last];
    }
}

// Colour one head position from its predecessors' LDS colours; x.pos holds
// the first batch. Returns false when a wait hit its bound (ctl->err set).
// LDS colours: relaxed workgroup-scope atomics keep them LDS accesses
// (ds_read / ds_write) that other waves' stores become visible to.
__device__ __forceinline__ uint32_t lds_get(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Generic colouring of a head position: any number of predecessors (in
// batches of 1024) and colours past the first window (4096 colours per pass
// over the, by then complete, list). x.pos holds the first batch. Returns
// the colour, or 0 when a wait hit its bound (ctl->err set).
__device__ __noinline__ uint32_t head_colour_any(head_vtx &x, uint32_t *bm, uint32_t *col,
                                                 const uint32_t *__restrict__ esplit,
                                                 gcxk_ctl *ctl, uint64_t spin_ticks, int lane)
{
    uint32_t c = x.np ? 0u : 1u;
    for (uint32_t win = 0; c == 0; win++) {
        const uint32_t lo = win * kHeadWin; // window: colours lo + 1 .. lo + kHeadWin
        for (uint32_t b0 = 0; b0 < x.np; b0 += kHeadBatch) {
            if (b0 || win)
                head_fetch_list(x, b0, esplit, lane);
            uint32_t need = 0;
#pragma unroll
            for (int t = 0; t < kHeadSlots; t++)
                need |= b0 + (uint32_t)(kWave * t + lane) < x.np ? 1u << t : 0u;
            uint32_t polls = 0;
            uint64_t t0 = 0;
            for (;;) {
                uint32_t cv[kHeadSlots];
#pragma unroll
                for (int t = 0; t < kHeadSlots; t++)
                    cv[t] = lds_get(&col[x.pos[t]]);
#pragma unroll
                for (int t = 0; t < kHeadSlots; t++) {
                    if ((need & (1u << t)) && cv[t]) {
                        need &= ~(1u << t);
                        if (cv[t] > lo && cv[t] <= lo + kHeadWin)
                            atomicOr(&bm[(cv[t] - lo - 1u) / 32u], 1u << ((cv[t] - lo - 1u) % 32u));
                    }
                }
                if (__ballot(need != 0u) == 0ull)
                    break;
                if (polls == kClockPoll)
                    t0 = __builtin_amdgcn_s_memrealtime();
                __builtin_amdgcn_s_sleep(1);
                if ((++polls & 255u) == 0u &&
                    (__builtin_amdgcn_s_memrealtime() - t0 > spin_ticks ||
                     __hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                    if (lane == 0)
                        __hip_atomic_store(&ctl->err, 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    return 0u;
                }
            }
        }
        const uint32_t w0 = bm[lane], w1 = bm[lane + kWave];
        const uint64_t f0 = __ballot(w0 != 0xFFFFFFFFu), f1 = __ballot(w1 != 0xFFFFFFFFu);
        if (f0) {
            const int fl = __ffsll((long long)f0) - 1;
            c = lo + 32u * (uint32_t)fl +
                (uint32_t)__builtin_ctz(~__builtin_amdgcn_readlane(w0, fl)) + 1u;
        } else if (f1) {
            const int fl = __ffsll((long long)f1) - 1;
            c = lo + 32u * (uint32_t)(kWave + fl) +
                (uint32_t)__builtin_ctz(~__builtin_amdgcn_readlane(w1, fl)) + 1u;
        }
        bm[lane] = 0u;
        bm[lane + kWave] = 0u;
    }
    return c;
}

// Smallest colour of the window (1 .. 4096) whose bit is clear in the wave's
// bitmap words (w0: words 0-63, w1: words 64-127, one of each per lane); 0
// when all are set. Marks it set in w0 / w1 of its lane.
__device__ __forceinline__ uint32_t head_take_free(uint32_t &w0, uint32_t &w1, int lane)
{
    const uint64_t f0 = __ballot(w0 != 0xFFFFFFFFu), f1 = __ballot(w1 != 0xFFFFFFFFu);
    const uint64_t f = f0 ? f0 : f1;
    if (!f)
        return 0u;
    const int fl = __ffsll((long long)f) - 1;
    const uint32_t word = f0 ? __builtin_amdgcn_readlane(w0, fl) : __builtin_amdgcn_readlane(w1, fl);
    const uint32_t b = (uint32_t)__builtin_ctz(~word);
    if (lane == fl) {
        if (f0)
            w0 |= 1u << b;
        else
            w1 |= 1u << b;
    }
    return 32u * (uint32_t)(f0 ? fl : kWave + fl) + b + 1u;
}

// Wait until `*p >= target` (LDS word); false when the wait hit its bound.
__device__ __forceinline__ bool head_wait_ge(const uint32_t *p, uint32_t target, uint32_t &polls,
                                             uint64_t &t0, gcxk_ctl *ctl, uint64_t spin_ticks,
                                             int lane, uint32_t &got)
{
    while ((got = lds_get(p)) < target) {
        __builtin_amdgcn_s_sleep(1);
        if (++polls == kClockPoll)
            t0 = __builtin_amdgcn_s_memrealtime();
        if ((polls & 255u) == 0u &&
            (__builtin_amdgcn_s_memrealtime() - t0 > spin_ticks ||
             __hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
            if (lane == 0)
                __hip_atomic_store(&ctl->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
    }
    return true;
}

// Colour of one head position, or 0 when a wait hit its bound. The
// watermark *wm is the smallest position not yet coloured. Common case (at
// most 1024 predecessors, colours below 4096): with L the predecessor
// latest in the greedy order, the wave waits until wm >= L -- then every
// other predecessor is coloured -- folds those colours into its LDS bitmap
// in one gather, finds the two smallest colours absent from it, m0 < m1,
// and waits for L's colour cl: the colour is cl == m0 ? m1 : m0. Only that
// last wait is on the chain. Everything else goes to head_colour_any.
__device__ __forceinline__ uint32_t head_colour(head_vtx &x, uint32_t *bm, uint32_t *col,
                                                const uint32_t *wm,
                                                const uint32_t *__restrict__ esplit,
                                                gcxk_ctl *ctl, uint64_t spin_ticks, int lane)
{
    if (x.np == 0u)
        return 1u;
    if (x.np > kHeadBatch)
        return head_colour_any(x, bm, col, esplit, ctl, spin_ticks, lane);
    uint32_t latest = 0;
#pragma unroll
    for (int t = 0; t < kHeadSlots; t++) {
        const bool in = (uint32_t)(kWave * t + lane) < x.np;
        latest = in && x.pos[t] > latest ? x.pos[t] : latest;
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const uint32_t y = __shfl_xor(latest, off);
        latest = y > latest ? y : latest;
    }
    uint32_t polls = 0, got = 0;
    uint64_t t0 = 0;
    if (!head_wait_ge(wm, latest, polls, t0, ctl, spin_ticks, lane, got))
        return 0u;
    // every predecessor but `latest` is coloured: fold them (branch-free;
    // lanes with nothing to add OR 0 into a word of their own)
    uint32_t cv[kHeadSlots];
#pragma unroll
    for (int t = 0; t < kHeadSlots; t++)
        cv[t] = lds_get(&col[x.pos[t]]);
#pragma unroll
    for (int t = 0; t < kHeadSlots; t++) {
        const bool ok = (uint32_t)(kWave * t + lane) < x.np && x.pos[t] != latest;
        const bool in_win = ok && cv[t] <= kHeadWin;
        const uint32_t widx = in_win ? (cv[t] - 1u) / 32u : kHeadWinWords + (uint32_t)lane;
        atomicOr(&bm[widx], in_win ? 1u << ((cv[t] - 1u) % 32u) : 0u);
    }
    uint32_t w0 = bm[lane], w1 = bm[lane + kWave];
    const uint32_t m0 = head_take_free(w0, w1, lane);
    const uint32_t m1 = head_take_free(w0, w1, lane);
    if (m1 == 0u) // colours past the window: the generic path (the bitmap
                  // keeps what was folded; folding it again changes nothing)
        return head_colour_any(x, bm, col, esplit, ctl, spin_ticks, lane);
    bm[lane] = 0u;
    bm[lane + kWave] = 0u;
    uint32_t cl = 0;
    if (!head_wait_ge(&col[latest], 1u, polls, t0, ctl, spin_ticks, lane, cl))
        return 0u;
    return cl == m0 ? m1 : m0;
}

// Record colour c of position i and advance the watermark past it and any
// positions already coloured after it. A wave that finds the watermark at
// its own position moves it with a compare-and-swap, then re-checks the next
// position: a wave colouring that position meanwhile either sees the moved
// watermark (and advances it itself) or is seen by the re-check, so the
// watermark never stalls behind a coloured position. Lane 0 only.
__device__ __forceinline__ void head_finish(uint32_t i, uint32_t c, uint32_t u, uint32_t P,
                                            uint32_t *col, uint32_t *wm, int *__restrict__ color)
{
    __hip_atomic_store(&col[i], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    color[u] = (int)c;
    uint32_t k = i;
    while (lds_get(wm) == k) {
        uint32_t n = k + 1u;
        while (n < P && lds_get(&col[n]) != 0u)
            n++;
        uint32_t expect = k;
        if (!__hip_atomic_compare_exchange_strong(wm, &expect, n, __ATOMIC_RELAXED,
                                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
            break;
        if (n >= P || lds_get(&col[n]) == 0u)
            break; // a wave colouring n after this sees wm == n
        k = n;     // n was coloured meanwhile: keep going from it
    }
}

// One workgroup of 16 waves; wave w colours positions w, w + 16, ... of the
// prefix in order. A position's predecessors are the greedy-order positions
// the orientation pass wrote to the front of its row (esplit), at most 1024
// per batch in registers. The wave loads the batch and waits -- on one LDS
// word, the colour of the predecessor latest in the greedy order -- so that
// memory latency falls into the wait; once that colour is set it gathers
// all the predecessors' colours into its LDS bitmap (polling again for any
// still missing) and takes the first clear bit: the mex. Colours past a
// window of 4096 take further passes over the (by then complete) list. The
// next position's header is loaded one position ahead.
//
// The first unfinished position has all its predecessors finished and its
// wave is at it, so the kernel completes; every wait is bounded all the
// same.
__global__ __launch_bounds__(kHeadBlock) void jp_head_kernel(const uint32_t *__restrict__ esplit,
                                                             const uint32_t *__restrict__ ids,
                                                             const uint2 *__restrict__ hdr,
                                                             int *__restrict__ color, gcxk_ctl *ctl,
                                                             uint64_t spin_ticks)
{
    extern __shared__ uint32_t head_lds[];
    const int w = threadIdx.x / kWave, lane = lane_id();
    uint32_t *bm = head_lds + (size_t)w * kHeadWaveWords;
    uint32_t *col = head_lds + (size_t)kHeadWaves * kHeadWaveWords;
    uint32_t *wm = col + kHeadMax; // watermark: smallest position not yet coloured
    const uint32_t P = ctl->head_n;
    for (uint32_t i = threadIdx.x; i < P; i += kHeadBlock)
        col[i] = 0u;
    for (uint32_t i = threadIdx.x; i < (uint32_t)kHeadWaves * kHeadWaveWords; i += kHeadBlock)
        head_lds[i] = 0u;
    if (threadIdx.x == 0)
        *wm = 0u;
    __syncthreads();
    head_vtx x;
    head_hdr next = head_fetch_hdr((uint32_t)w, P, ids, hdr);
    for (uint32_t i = (uint32_t)w; i < P; i += kHeadWaves) {
        head_take(x, next);
        next = head_fetch_hdr(i + kHeadWaves, P, ids, hdr);
        head_fetch_list(x, 0, esplit, lane);
        const uint32_t c = head_colour(x, bm, col, wm, esplit, ctl, spin_ticks, lane);
        if (c == 0u)
            return;
        if (lane == 0)
            head_finish(i, c, x.u, P, col, wm, color);
    }
}

// Zero the wof[N] state words (grid-stride; the count is read on the device).
__global__ __launch_bounds__(kBlock) void jp_zero_words(const uint32_t *__restrict__ wof, uint32_t N,
                                                        uint64_t cap, uint64_t *__restrict__ st)
{
    uint64_t n = wof[N];
    n = n < cap ? n : cap;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * kBlock)
        st[i] = 0ull;
}

// Ordered schedule. Vertex v owns state words st[off .. off + W) in the
// layout its descriptor names (jp_word_counts). Every report from a
// predecessor with colour c is one atomic on one word:
//  counting fields: add 1 to the field of c (or to the last field when c is
//    past the fields); v's reports are the sum of its fields;
//  bit sets: OR c's bit, then add 1 << mb, both on the word that holds c (one
//    location: no wait between them); a colour past v's words can never be
//    v's mex, so it is only counted (add to word 0), except for vertices with
//    the full 8 words, whose higher colours go to overflow words (waited for,
//    then counted in word 0).
// v is ready when its reports sum to npred[v]: reports only grow and each one
// sits in the word it marks, so loads whose reports sum to npred have seen
// every mark. Lanes 0..W-1 poll the words in one load instruction. The
// mex is the first zero field, or the first clear bit.
//
// Waits sleep between polls by how many reports are still missing: a wave
// one or two reports from ready polls every 64 clocks, a wave that still
// waits for many sleeps 1024. Thousands of waiting waves otherwise load the
// memory system enough to slow the critical chain (2.5% faster than a
// backoff by poll count on R-MAT 22).
__device__ __forceinline__ void poll_sleep(uint32_t missing)
{
    if (missing <= 2)
        __builtin_amdgcn_s_sleep(1);
    else if (missing <= 8)
        __builtin_amdgcn_s_sleep(4);
    else
        __builtin_amdgcn_s_sleep(16);
}

__device__ __forceinline__ uint32_t field_sum(uint64_t w, uint32_t B)
{
    // per-word sums never exceed npred <= 255, so the byte sum has no carry
    if (B == 4u)
        w = (w & 0x0F0F0F0F0F0F0F0Full) + ((w >> 4) & 0x0F0F0F0F0F0F0F0Full);
    return (uint32_t)((w * 0x0101010101010101ull) >> 56);
}

// The lowest bit of every zero field of w; the last field of the last word
// (colours past the fields) never counts as free.
__device__ __forceinline__ uint64_t zero_fields(uint64_t w, uint32_t B, bool last)
{
    uint64_t t, low;
    if (B == 4u) {
        t = w | (w >> 2);
        t |= t >> 1;
        low = 0x1111111111111111ull;
    } else {
        t = w | (w >> 4);
        t |= t >> 2;
        t |= t >> 1;
        low = 0x0101010101010101ull;
    }
    uint64_t z = ~t & low;
    if (last)
        z &= ~(1ull << (64u - B));
    return z;
}

__global__ __launch_bounds__(kBlock) void jp_ordered_kernel(const uint32_t *__restrict__ V,
                                                            const uint32_t *__restrict__ npred,
                                                            const uint32_t *__restrict__ esplit,
                                                            uint64_t *__restrict__ st,
                                                            const uint64_t *__restrict__ wdesc,
                                                            int *__restrict__ color,
                                                            uint32_t *__restrict__ ovf,
                                                            const uint2 *__restrict__ items,
                                                            gcxk_ctl *ctl, int mb,
                                                            uint64_t spin_ticks)
{
    const uint32_t n = ctl->err ? 0u : ctl->total_items; // a failed head ends the launch
    const int lane = lane_id();
    const uint64_t one = 1ull << mb, low = one - 1ull;
    const uint32_t cap = kMaxWords * (uint32_t)mb;
    const uint32_t w0 = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    for (uint32_t j = w0; j < n; j += nw) {
        const uint2 it = items[j];
        const uint32_t u = it.x, k = it.y;
        const uint32_t Vu = V[u], Vu1 = V[u + 1];
        const uint32_t np = npred[u];
        const uint64_t du = wdesc[u];
        const uint32_t wo = (uint32_t)du, wn = (uint32_t)(du >> 40) & 0xFFu;
        const uint32_t Bu = (uint32_t)(du >> 48) & 0xFFu;
        const uint32_t i = Vu + np + k * kChunk + (uint32_t)lane;
        const bool has = i < Vu1;
        const uint32_t v = has ? esplit[i] : 0u;
        const uint64_t dv = has ? wdesc[v] : 0ull;
        const uint32_t vo = (uint32_t)dv, vn = (uint32_t)(dv >> 40) & 0xFFu;
        const uint32_t Bv = (uint32_t)(dv >> 48) & 0xFFu;
        // head-coloured successors get no reports (nothing waits on them)
        const bool report = has && !((dv >> kHeadFlagBit) & 1ull);
        const bool mine = (uint32_t)lane < wn;
        uint64_t w = 0;
        uint32_t polls = 0;
        uint64_t t0 = 0;
        // a head-coloured vertex (coloured by the head kernel, which ran
        // before this launch) only reports its colour to its successors
        const bool head_u = (du >> kHeadFlagBit) & 1ull;
        for (; !head_u;) {
            w = mine ? load_state(&st[wo + lane]) : 0ull;
            uint32_t r = mine ? (Bu ? field_sum(w, Bu) : (uint32_t)(w >> mb)) : 0u;
            r += __shfl_xor(r, 1);
            r += __shfl_xor(r, 2);
            r += __shfl_xor(r, 4);
            r += __shfl_xor(r, 8);
            if (wn > 16u) { // wave-uniform
                r += __shfl_xor(r, 16);
                r += __shfl_xor(r, 32);
            }
            r = __shfl(r, 0);
            if (r == np)
                break;
            if (polls == kClockPoll)
                t0 = __builtin_amdgcn_s_memrealtime();
            poll_sleep(np - r);
            if ((++polls & 255u) == 0u &&
                (__builtin_amdgcn_s_memrealtime() - t0 > spin_ticks || load_word(&ctl->err))) {
                if (lane == 0)
                    __hip_atomic_store(&ctl->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return;
            }
        }
        // mex: the first zero field, or the first word with a clear bit (all
        // 8 words full: overflow)
        uint32_t c;
        if (head_u) {
            c = (uint32_t)color[u];
        } else if (Bu) {
            const uint64_t z = mine ? zero_fields(w, Bu, (uint32_t)lane == wn - 1u) : 0ull;
            const uint64_t bz = __ballot(z != 0ull);
            if (!bz) { // cannot happen: np + 1 colour fields, at most np counted
                if (lane == 0)
                    __hip_atomic_store(&ctl->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return;
            }
            const int fl = __ffsll((long long)bz) - 1;
            const uint32_t lo = __shfl((uint32_t)z, fl), hi = __shfl((uint32_t)(z >> 32), fl);
            c = (uint32_t)fl * (64u / Bu) +
                (uint32_t)__builtin_ctzll(((uint64_t)hi << 32) | lo) / Bu + 1u;
        } else {
            const uint64_t fr = mine ? (~w & low) : 0ull;
            const uint64_t bw = __ballot(fr != 0ull);
            if (bw) {
                const int fl = __ffsll((long long)bw) - 1;
                const uint32_t lo = __shfl((uint32_t)fr, fl), hi = __shfl((uint32_t)(fr >> 32), fl);
                c = (uint32_t)fl * (uint32_t)mb +
                    (uint32_t)__builtin_ctzll(((uint64_t)hi << 32) | lo) + 1u;
            } else {
                c = mex_overflow(cap, u, Vu, Vu1 - Vu, ovf);
            }
        }
        if (k == 0 && lane == 0 && !head_u)
            color[u] = (int)c;
        bool late = false;
        if (report) {
            if (Bv) {
                const uint32_t F = 64u / Bv, top = vn * F - 1u;
                const uint32_t fi = c - 1u < top ? c - 1u : top;
                atomicAdd((unsigned long long *)&st[vo + fi / F], 1ull << (Bv * (fi % F)));
            } else {
                const uint32_t wi = (c - 1u) / (uint32_t)mb;
                if (wi < vn) {
                    unsigned long long *p = (unsigned long long *)&st[vo + wi];
                    atomicOr(p, 1ull << (c - 1u - wi * (uint32_t)mb));
                    atomicAdd(p, (unsigned long long)one);
                } else if (vn == kMaxWords) {
                    or_overflow(V, ovf, v, c, cap);
                    late = true;
                } else {
                    atomicAdd((unsigned long long *)&st[vo], (unsigned long long)one);
                }
            }
        }
        if (__ballot(late)) {
            __builtin_amdgcn_s_waitcnt(0);
            if (late)
                atomicAdd((unsigned long long *)&st[vo], (unsigned long long)one);
        }
    }
}

// ------------------------------------------------------------- validation

// Vertices in degree order (perm), 4 lanes each, grid-stride; one atomic per
// workgroup. Hub rows (perm[0 .. nhub)) only get their own colour checked
// here; their edges are validate_hubs' work.
__global__ __launch_bounds__(kBlock) void validate_quads(uint32_t N, const uint32_t *__restrict__ V,
                                                         const uint32_t *__restrict__ E,
                                                         const int *__restrict__ color,
                                                         const uint32_t *__restrict__ perm,
                                                         uint32_t nhub, gcxk_ctl *ctl)
{
    __shared__ uint32_t sbad[kWavesPerBlock], stop[kWavesPerBlock];
    const uint32_t gl = threadIdx.x % kQuad;
    const uint32_t qpb = kBlock / kQuad;
    uint32_t bad = 0, top = 0;
    for (uint64_t p = (uint64_t)blockIdx.x * qpb + threadIdx.x / kQuad; p < N;
         p += (uint64_t)gridDim.x * qpb) {
        const uint32_t u = perm[p];
        const int cu = color[u];
        if (p >= nhub) {
            const uint32_t beg = V[u], end = V[u + 1];
            for (uint32_t i = beg + gl; i < end; i += kQuad)
                bad += (color[E[i]] == cu) ? 1u : 0u;
        }
        if (gl == 0) {
            bad += cu <= 0 ? 1u : 0u;
            top = cu > 0 && (uint32_t)cu > top ? (uint32_t)cu : top;
        }
    }
    bad = group_sum32<kWave>(bad);
    top = group_max32<kWave>(top);
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        sbad[w] = bad;
        stop[w] = top;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t b = 0, t = 0;
        for (int i = 0; i < kWavesPerBlock; i++) {
            b += sbad[i];
            t = stop[i] > t ? stop[i] : t;
        }
        if (b)
            atomicAdd(&ctl->bad, b);
        if (t)
            atomicMax(&ctl->maxcolor, (int)t);
    }
}

__global__ __launch_bounds__(kBlock) void validate_hubs(const uint32_t *__restrict__ V,
                                                        const uint32_t *__restrict__ E,
                                                        const int *__restrict__ color,
                                                        const uint32_t *__restrict__ hubs,
                                                        gcxk_ctl *ctl)
{
    const uint32_t u = hubs[blockIdx.x];
    const uint32_t beg = V[u], end = V[u + 1];
    const int cu = color[u];
    uint32_t bad = 0;
    for (uint32_t i = beg + threadIdx.x; i < end; i += kBlock)
        bad += (color[E[i]] == cu) ? 1u : 0u;
    bad = group_sum32<kWave>(bad);
    if (lane_id() == 0 && bad)
        atomicAdd(&ctl->bad, bad);
}

// ------------------------------------------------------------- certificate

// mex over the predecessors' colours, one 64-colour window per pass. 16 lanes
// per non-hub vertex; hubs take a whole workgroup each.
template <typename K>
__global__ __launch_bounds__(kBlock) void mex_check_groups(uint32_t N, const uint32_t *__restrict__ V,
                                                           const uint32_t *__restrict__ E,
                                                           const K *__restrict__ key,
                                                           const int *__restrict__ color,
                                                           uint32_t hub_deg, gcxk_ctl *ctl)
{
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint32_t u = (uint32_t)(t / kGroup);
    const int gl = (int)(t % kGroup);
    if (u >= N) // whole groups leave together
        return;
    const uint32_t beg = V[u], end = V[u + 1];
    if (end - beg >= hub_deg)
        return;
    const K ku = key[u];
    uint32_t base = 0;
    uint64_t mask;
    for (;;) {
        mask = 0;
        for (uint32_t i = beg + gl; i < end; i += kGroup) {
            const uint32_t v = E[i];
            if (precedes(key[v], v, ku, u)) {
                const uint32_t c = (uint32_t)color[v];
                if (c > base && c <= base + 64)
                    mask |= 1ull << (c - base - 1);
            }
        }
        mask = group_or64<kGroup>(mask);
        if (mask != ~0ull)
            break;
        base += 64;
    }
    const uint32_t mex = base + (uint32_t)__builtin_ctzll(~mask) + 1u;
    if (gl == 0 && (uint32_t)color[u] != mex)
        atomicAdd((unsigned long long *)&ctl->viol, 1ull);
}

template <typename K>
__global__ __launch_bounds__(kBlock) void mex_check_hubs(const uint32_t *__restrict__ V,
                                                         const uint32_t *__restrict__ E,
                                                         const K *__restrict__ key,
                                                         const int *__restrict__ color,
                                                         const uint32_t *__restrict__ hubs,
                                                         gcxk_ctl *ctl)
{
    __shared__ unsigned long long smask;
    const uint32_t u = hubs[blockIdx.x];
    const uint32_t beg = V[u], end = V[u + 1];
    const K ku = key[u];
    uint32_t base = 0;
    uint64_t mask;
    for (;;) {
        if (threadIdx.x == 0)
            smask = 0ull;
        __syncthreads();
        mask = 0;
        for (uint32_t i = beg + threadIdx.x; i < end; i += kBlock) {
            const uint32_t v = E[i];
            if (precedes(key[v], v, ku, u)) {
                const uint32_t c = (uint32_t)color[v];
                if (c > base && c <= base + 64)
                    mask |= 1ull << (c - base - 1);
            }
        }
        mask = group_or64<kWave>(mask);
        if (lane_id() == 0 && mask)
            atomicOr(&smask, (unsigned long long)mask);
        __syncthreads();
        mask = smask;
        __syncthreads();
        if (mask != ~0ull)
            break;
        base += 64;
    }
    const uint32_t mex = base + (uint32_t)__builtin_ctzll(~mask) + 1u;
    if (threadIdx.x == 0 && (uint32_t)color[u] != mex)
        atomicAdd((unsigned long long *)&ctl->viol, 1ull);
}

// Workgroups of the ordered kernel: default, and how many fit on the device
// at once.
uint32_t g_ordered_grid = 0, g_ordered_cap = 0;

// Byte offset of the sort / scan temporary storage inside the scratch:
// after two key arrays and two id arrays, rounded up to 256 bytes.
size_t order_tmp_offset(uint32_t N, size_t key_size)
{
    return (2 * key_size * N + 8ull * N + 255) & ~(size_t)255;
}

template <typename K>
int sort_order(uint32_t N, const void *key, void *scratch, size_t scratch_bytes, uint32_t **ids,
               hipStream_t s)
{
    // scratch: nkey[2][N] (K), id[2][N], then rocprim temporary storage
    char *p = static_cast<char *>(scratch);
    K *k0 = reinterpret_cast<K *>(p);
    K *k1 = k0 + N;
    uint32_t *i0 = reinterpret_cast<uint32_t *>(k1 + N);
    uint32_t *i1 = i0 + N;
    const size_t used = order_tmp_offset(N, sizeof(K));
    char *tmp = p + used;
    if (used > scratch_bytes)
        return GCX_ERR_CAPACITY;
    size_t tmp_bytes = scratch_bytes - used;
    jp_order_keys<K><<<div_up(N, kBlock), kBlock, 0, s>>>(N, static_cast<const K *>(key), k0, i0);
    GCXK_LAUNCH_CHECK();
    rocprim::double_buffer<K> kb(k0, k1);
    rocprim::double_buffer<uint32_t> vb(i0, i1);
    size_t need = 0;
    GCXK_CHECK(rocprim::radix_sort_pairs(nullptr, need, kb, vb, N, 0, (unsigned)(8 * sizeof(K)), s));
    if (need > tmp_bytes)
        return GCX_ERR_CAPACITY;
    GCXK_CHECK(rocprim::radix_sort_pairs(tmp, need, kb, vb, N, 0, (unsigned)(8 * sizeof(K)), s));
    *ids = vb.current();
    return GCX_OK;
}

// Order-independent digest of a colouring: sum over u of mix(u, colour[u]).
__global__ __launch_bounds__(kBlock) void color_digest_kernel(uint32_t N, const int *__restrict__ color,
                                                              unsigned long long *out)
{
    __shared__ unsigned long long part[kWavesPerBlock];
    unsigned long long h = 0;
    for (uint64_t u = (uint64_t)blockIdx.x * kBlock + threadIdx.x; u < N;
         u += (uint64_t)gridDim.x * kBlock)
        h += mix64((u << 32) | (uint32_t)color[u]);
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)h, off), hi = __shfl_xor((uint32_t)(h >> 32), off);
        h += ((uint64_t)hi << 32) | lo;
    }
    if (lane_id() == 0)
        part[threadIdx.x / kWave] = h;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < kWavesPerBlock; i++)
            t += part[i];
        atomicAdd(out, t);
    }
}

} // namespace

extern "C" uint32_t gcxk_round_grid(void) { return 1024u; }

extern "C" int gcxk_color_digest(uint32_t N, const int *color, uint64_t *out, void *stream)
{
    if (N == 0)
        return GCX_OK;
    uint32_t grid = div_up(N, kBlock);
    grid = grid < 1024u ? grid : 1024u;
    color_digest_kernel<<<grid, kBlock, 0, as_stream(stream)>>>(
        N, color, reinterpret_cast<unsigned long long *>(out));
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_jp_mask_bits(uint32_t maxdeg)
{
    const int cb = maxdeg ? 32 - __builtin_clz(maxdeg) : 1;
    return 64 - cb;
}

extern "C" int gcxk_jp_clear(uint32_t N, uint64_t M, uint32_t *ovf, uint32_t nhub, uint32_t *hpos,
                             const uint32_t *hubs, uint32_t *npred, uint64_t *state, int *color,
                             gcxk_ctl *ctl, void *stream)
{
    const uint64_t nwords = gcxk_jp_ovf_words(N, M);
    uint64_t n = nwords > 2ull * nhub ? nwords : 2ull * nhub;
    if (n < 64)
        n = 64;
    uint32_t grid = div_up(n, kBlock);
    if (grid > 4096)
        grid = 4096;
    jp_clear<<<grid, kBlock, 0, as_stream(stream)>>>(nwords, ovf, nhub, hpos, hubs, npred, state,
                                                     color, ctl);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_jp_init(uint32_t N, const uint32_t *dV, const uint32_t *dE, const void *key,
                            int key_bits, int mb, const uint32_t *perm, uint32_t *npred,
                            uint32_t *esplit, uint64_t *state, int *color, const uint32_t *hubs,
                            uint32_t nhub, const uint32_t *hchunk, uint32_t nhchunk,
                            uint32_t *hpos, uint32_t head_max, void *stream, void *hub_stream)
{
    if (N == 0)
        return GCX_OK;
    if (head_max && key_bits != 32)
        return GCX_ERR_ARG; // predecessor positions come from rank keys
    hipStream_t s = as_stream(stream), hs = as_stream(hub_stream);
    const uint32_t grid = div_up((uint64_t)N * kQuad, kBlock);
    if (key_bits == 32) {
        jp_init_quads<uint32_t><<<grid, kBlock, 0, s>>>(N, dV, dE, (const uint32_t *)key, perm,
                                                        nhub, npred, esplit, state, color, mb,
                                                        head_max);
        if (nhub)
            jp_init_hub_chunks<uint32_t><<<nhchunk, kBlock, 0, hs>>>(
                dV, dE, (const uint32_t *)key, npred, state, esplit, hubs, hchunk, nhub, hpos, mb,
                N, head_max);
    } else {
        jp_init_quads<uint64_t><<<grid, kBlock, 0, s>>>(N, dV, dE, (const uint64_t *)key, perm,
                                                        nhub, npred, esplit, state, color, mb, 0u);
        if (nhub)
            jp_init_hub_chunks<uint64_t><<<nhchunk, kBlock, 0, hs>>>(
                dV, dE, (const uint64_t *)key, npred, state, esplit, hubs, hchunk, nhub, hpos, mb,
                N, 0u);
    }
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_jp_roots(uint32_t N, const uint32_t *dV, const uint32_t *npred, gcxk_item *q0,
                             gcxk_ctl *ctl, void *stream)
{
    if (N == 0)
        return GCX_OK;
    jp_collect_roots<<<div_up(N, kRootBlock), kRootBlock, 0, as_stream(stream)>>>(
        N, dV, npred, reinterpret_cast<uint2 *>(q0), ctl);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_jp_round(const uint32_t *dV, const uint32_t *npred, const uint32_t *esplit,
                             uint64_t *state, int *color, uint32_t *ovf, int mb,
                             gcxk_item *const q[2], gcxk_ctl *ctl, uint32_t r, uint32_t grid,
                             void *stream)
{
    jp_round_kernel<<<grid, kBlock, 0, as_stream(stream)>>>(
        dV, npred, esplit, state, color, ovf, reinterpret_cast<const uint2 *>(q[r & 1]),
        reinterpret_cast<uint2 *>(q[(r + 1) & 1]), ctl, r, mb);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" size_t gcxk_jp_order_scratch(uint32_t N, int key_bits)
{
    const size_t ks = key_bits == 32 ? 4 : 8;
    size_t need = 0;
    if (N) {
        if (key_bits == 32) {
            rocprim::double_buffer<uint32_t> kb(nullptr, nullptr);
            rocprim::double_buffer<uint32_t> vb(nullptr, nullptr);
            if (rocprim::radix_sort_pairs(nullptr, need, kb, vb, N, 0, 32u) != hipSuccess)
                return 0;
        } else {
            rocprim::double_buffer<uint64_t> kb(nullptr, nullptr);
            rocprim::double_buffer<uint32_t> vb(nullptr, nullptr);
            if (rocprim::radix_sort_pairs(nullptr, need, kb, vb, N, 0, 64u) != hipSuccess)
                return 0;
        }
    }
    size_t scan = 0, scan2 = 0;
    if (N && (rocprim::inclusive_scan(nullptr, scan, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                      N, rocprim::plus<uint32_t>()) != hipSuccess ||
              rocprim::exclusive_scan(nullptr, scan2, (const uint32_t *)nullptr,
                                      (uint32_t *)nullptr, 0u, (size_t)N + 1,
                                      rocprim::plus<uint32_t>()) != hipSuccess))
        return 0;
    need = need > scan ? need : scan;
    need = need > scan2 ? need : scan2;
    return order_tmp_offset(N, ks) + need;
}

// State word offsets, descriptors (head flags when rk is given) and zeroed
// words; incl (N + 1) is used for the per-vertex word counts.
static int order_words(uint32_t N, const uint32_t *npred, const uint32_t *rk, int mb, char *tmp,
                       size_t tmp_bytes, uint32_t *incl, uint32_t *wof, uint64_t *wdesc,
                       uint64_t *words, uint64_t words_cap, gcxk_ctl *ctl, hipStream_t s)
{
    size_t need = 0;
    jp_word_counts<<<div_up((uint64_t)N + 1, kBlock), kBlock, 0, s>>>(N, npred, mb, incl);
    GCXK_LAUNCH_CHECK();
    GCXK_CHECK(rocprim::exclusive_scan(nullptr, need, incl, wof, 0u, (size_t)N + 1,
                                       rocprim::plus<uint32_t>(), s));
    if (need > tmp_bytes)
        return GCX_ERR_CAPACITY;
    GCXK_CHECK(rocprim::exclusive_scan(tmp, need, incl, wof, 0u, (size_t)N + 1,
                                       rocprim::plus<uint32_t>(), s));
    jp_word_desc<<<div_up(N, kBlock), kBlock, 0, s>>>(N, npred, wof, rk, ctl, wdesc);
    GCXK_LAUNCH_CHECK();
    uint32_t zg = div_up((uint64_t)N + 1, kBlock);
    zg = zg < 2048u ? zg : 2048u;
    jp_zero_words<<<zg, kBlock, 0, s>>>(wof, N, words_cap, words);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

// Work items of the vertices in greedy order (ids): counts into cnt, their
// inclusive scan into incl, then the items and ctl->total_items.
static int order_items_from(uint32_t N, const uint32_t *dV, const uint32_t *npred,
                            const uint32_t *ids, char *tmp, size_t tmp_bytes, uint32_t *cnt,
                            uint32_t *incl, gcxk_item *items, gcxk_ctl *ctl, hipStream_t s)
{
    size_t need = 0;
    jp_item_counts<<<div_up(N, kBlock), kBlock, 0, s>>>(N, dV, npred, ids, cnt);
    GCXK_LAUNCH_CHECK();
    GCXK_CHECK(rocprim::inclusive_scan(nullptr, need, cnt, incl, N, rocprim::plus<uint32_t>(), s));
    if (need > tmp_bytes)
        return GCX_ERR_CAPACITY;
    GCXK_CHECK(rocprim::inclusive_scan(tmp, need, cnt, incl, N, rocprim::plus<uint32_t>(), s));
    jp_expand<<<div_up(N, kBlock), kBlock, 0, s>>>(N, ids, incl, reinterpret_cast<uint2 *>(items),
                                                   ctl);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_jp_order_items(uint32_t N, const uint32_t *dV, const uint32_t *npred,
                                   const void *key, int key_bits, int mb, void *scratch,
                                   size_t scratch_bytes, uint32_t *cnt, uint32_t *incl,
                                   uint32_t *wof, uint64_t *wdesc, uint64_t *words,
                                   uint64_t words_cap, gcxk_item *items, gcxk_ctl *ctl,
                                   void *stream)
{
    if (N == 0)
        return GCX_OK;
    hipStream_t s = as_stream(stream);
    const size_t ks = key_bits == 32 ? 4 : 8;
    if (order_tmp_offset(N, ks) > scratch_bytes)
        return GCX_ERR_CAPACITY;
    char *tmp = static_cast<char *>(scratch) + order_tmp_offset(N, ks);
    const size_t tmp_bytes = scratch_bytes - order_tmp_offset(N, ks);
    int rc = order_words(N, npred, nullptr, mb, tmp, tmp_bytes, incl, wof, wdesc, words, words_cap,
                         ctl, s);
    if (rc != GCX_OK)
        return rc;
    uint32_t *ids = nullptr;
    rc = key_bits == 32 ? sort_order<uint32_t>(N, key, scratch, scratch_bytes, &ids, s)
                        : sort_order<uint64_t>(N, key, scratch, scratch_bytes, &ids, s);
    if (rc != GCX_OK)
        return rc;
    // the sort's temporary storage is free again: reuse it for the scan
    return order_items_from(N, dV, npred, ids, tmp, tmp_bytes, cnt, incl, items, ctl, s);
}

extern "C" int gcxk_jp_sort_ids(uint32_t N, const void *key, int key_bits, void *scratch,
                                size_t scratch_bytes, uint32_t **ids, void *stream)
{
    *ids = nullptr;
    if (N == 0)
        return GCX_OK;
    hipStream_t s = as_stream(stream);
    return key_bits == 32 ? sort_order<uint32_t>(N, key, scratch, scratch_bytes, ids, s)
                          : sort_order<uint64_t>(N, key, scratch, scratch_bytes, ids, s);
}

extern "C" int gcxk_jp_rank_keys(uint32_t N, const uint32_t *ids, uint32_t *rk, void *stream)
{
    if (N == 0)
        return GCX_OK;
    jp_rank_keys<<<div_up(N, kBlock), kBlock, 0, as_stream(stream)>>>(N, ids, rk);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_jp_order_items_ranked(uint32_t N, const uint32_t *dV, const uint32_t *npred,
                                          const uint32_t *ids, const uint32_t *rk,
                                          uint32_t head_max, uint32_t head_density, int key_bits,
                                          int mb, void *scratch, size_t scratch_bytes,
                                          uint32_t *cnt, uint32_t *incl, uint32_t *wof,
                                          uint64_t *wdesc, uint64_t *words, uint64_t words_cap,
                                          gcxk_item *items, gcxk_ctl *ctl, void *stream)
{
    if (N == 0)
        return GCX_OK;
    hipStream_t s = as_stream(stream);
    const size_t ks = key_bits == 32 ? 4 : 8;
    if (order_tmp_offset(N, ks) > scratch_bytes)
        return GCX_ERR_CAPACITY;
    char *tmp = static_cast<char *>(scratch) + order_tmp_offset(N, ks);
    const size_t tmp_bytes = scratch_bytes - order_tmp_offset(N, ks);
    if (head_max > kHeadMax)
        head_max = kHeadMax;
    if (head_max) {
        jp_head_extent<<<1, kHeadBlock, 0, s>>>(N, ids, npred, head_max, head_density, ctl);
        GCXK_LAUNCH_CHECK();
    }
    const int rc = order_words(N, npred, head_max ? rk : nullptr, mb, tmp, tmp_bytes, incl, wof,
                               wdesc, words, words_cap, ctl, s);
    if (rc != GCX_OK)
        return rc;
    return order_items_from(N, dV, npred, ids, tmp, tmp_bytes, cnt, incl, items, ctl, s);
}

extern "C" uint32_t gcxk_jp_head_max(void) { return kHeadMax; }

extern "C" int gcxk_jp_head(const uint32_t *dV, const uint32_t *npred, const uint32_t *esplit,
                            const uint32_t *ids, void *hdr, int *color, gcxk_ctl *ctl,
                            uint64_t spin_ticks, void *stream)
{
    static int lds_set = 0; // the LDS limit attribute is per function, set once
    if (!lds_set) {
        GCXK_CHECK(hipFuncSetAttribute((const void *)jp_head_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kHeadLds));
        lds_set = 1;
    }
    hipStream_t s = as_stream(stream);
    uint2 *h = static_cast<uint2 *>(hdr);
    jp_head_table<<<div_up(kHeadMax, kBlock), kBlock, 0, s>>>(dV, npred, ids, ctl, h);
    GCXK_LAUNCH_CHECK();
    jp_head_kernel<<<1, kHeadBlock, kHeadLds, s>>>(esplit, ids, h, color, ctl, spin_ticks);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" uint32_t gcxk_jp_ordered_grid(void)
{
    if (!g_ordered_grid) {
        int per_cu = 0, dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, jp_ordered_kernel, kBlock, 0) !=
                hipSuccess ||
            per_cu <= 0)
            return 0;
        // 6 workgroups per CU (8 fit). Alone on the device, orientation + JP
        // on the 160^3 stencil take 5.33 / 4.98 ms at 6 / 8 per CU and R-MAT
        // 22 is flat (7.65 / 7.64 ms); but a launch that needs every resident
        // slot cannot become resident while other streams' kernels hold some:
        // with 8 serving workers, cooperative launches of 8 per CU fell from
        // 69 to 0.6 jobs/s. 6 per CU leaves room for them.
        const uint32_t cap = (uint32_t)per_cu * (uint32_t)prop.multiProcessorCount;
        const uint32_t want = 6u * (uint32_t)prop.multiProcessorCount;
        g_ordered_grid = want < cap ? want : cap;
        g_ordered_cap = cap;
    }
    return g_ordered_grid;
}

extern "C" uint32_t gcxk_jp_ordered_cap(void) { return gcxk_jp_ordered_grid() ? g_ordered_cap : 0; }

extern "C" int gcxk_jp_ordered(const uint32_t *dV, const uint32_t *npred, const uint32_t *esplit,
                               uint64_t *words, const uint64_t *wdesc, int *color, uint32_t *ovf,
                               int mb, const gcxk_item *items, gcxk_ctl *ctl, uint32_t grid,
                               uint64_t spin_ticks, int cooperative, void *stream)
{
    const uint32_t dflt = gcxk_jp_ordered_grid();
    if (!dflt)
        return GCX_ERR_HIP;
    if (grid == 0)
        grid = dflt;
    if (grid > g_ordered_cap)
        grid = g_ordered_cap;
    const uint2 *it = reinterpret_cast<const uint2 *>(items);
    void *args[] = {(void *)&dV,  (void *)&npred, (void *)&esplit, (void *)&words,
                    (void *)&wdesc, (void *)&color, (void *)&ovf,  (void *)&it,
                    (void *)&ctl, (void *)&mb,    (void *)&spin_ticks};
    if (cooperative) {
        GCXK_CHECK(hipLaunchCooperativeKernel((const void *)jp_ordered_kernel, dim3(grid),
                                              dim3(kBlock), args, 0, as_stream(stream)));
    } else {
        // grid <= the resident capacity: once the kernels of other streams
        // (none of which wait on anything) drain, every workgroup is resident
        GCXK_CHECK(hipLaunchKernel((const void *)jp_ordered_kernel, dim3(grid), dim3(kBlock), args,
                                   0, as_stream(stream)));
    }
    return GCX_OK;
}

extern "C" int gcxk_validate(uint32_t N, const uint32_t *dV, const uint32_t *dE, const int *color,
                             const uint32_t *hubs, uint32_t nhub, const uint32_t *perm,
                             gcxk_ctl *ctl, void *stream, void *hub_stream)
{
    if (N == 0)
        return GCX_OK;
    uint32_t grid = div_up((uint64_t)N * kQuad, kBlock);
    if (grid > 2048)
        grid = 2048;
    validate_quads<<<grid, kBlock, 0, as_stream(stream)>>>(N, dV, dE, color, perm, nhub, ctl);
    if (nhub)
        validate_hubs<<<nhub, kBlock, 0, as_stream(hub_stream)>>>(dV, dE, color, hubs, ctl);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_mex_check(uint32_t N, const uint32_t *dV, const uint32_t *dE, const void *key,
                              int key_bits, const int *color, const uint32_t *hubs, uint32_t nhub,
                              uint32_t hub_deg, gcxk_ctl *ctl, void *stream)
{
    if (N == 0)
        return GCX_OK;
    hipStream_t s = as_stream(stream);
    const uint32_t grid = div_up((uint64_t)N * kGroup, kBlock);
    if (key_bits == 32) {
        mex_check_groups<uint32_t><<<grid, kBlock, 0, s>>>(N, dV, dE, (const uint32_t *)key, color,
                                                           hub_deg, ctl);
        if (nhub)
            mex_check_hubs<uint32_t><<<nhub, kBlock, 0, s>>>(dV, dE, (const uint32_t *)key, color,
                                                             hubs, ctl);
    } else {
        mex_check_groups<uint64_t><<<grid, kBlock, 0, s>>>(N, dV, dE, (const uint64_t *)key, color,
                                                           hub_deg, ctl);
        if (nhub)
            mex_check_hubs<uint64_t><<<nhub, kBlock, 0, s>>>(dV, dE, (const uint64_t *)key, color,
                                                             hubs, ctl);
    }
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}
