// jp_pull.hip -- the ordered Jones-Plassmann schedule with pulled colours.
//
// Same result as the reference's jones_plassmann_internal (coloring.c:223-330)
// and as sequential greedy in (p descending, id ascending) order: every vertex
// takes the smallest colour >= 1 that none of its predecessors holds.
//
// Layout, all in greedy-order positions (position j = the j-th vertex of the
// greedy order, ids[j]):
//   rank[v]  = position of v (rank[ids[j]] = j);
//   E_split  = each row's predecessor POSITIONS at its front,
//              E_split[V[u] .. V[u] + np(u)); the rest of the row is unused;
//   hdr[j]   = (V[ids[j]], np(ids[j]));
//   pcol[j]  = colour of position j, 0 while it has none.
// Predecessors of a vertex are exactly the earlier positions among its
// neighbours, so v precedes u iff rank[v] < rank[u]. (Gathering the
// positions through the degree-order columns Er instead of E, which puts the
// hubs' words together, was 2% slower on R-MAT 22.)
//
// One persistent launch: wave w colours positions in increasing order (see
// jp_pull_kernel for which). A position gathers its predecessors' colours
// from pcol and waits for the missing ones; nothing is pushed to successors,
// so a hop of the DAG's longest chain is one colour store seen by one load
// (no atomics). While more than kFast predecessors are missing the wave
// polls only the latest of them in greedy order (one word), re-gathering the
// others when it is coloured or every kRegather polls (re-gathering only
// when it is coloured was 1.6x slower on R-MAT 22: the latest predecessor
// in greedy order is often not the last one coloured). Once at most kFast
// are missing, the nmiss + 1 smallest colours absent from the predecessors
// seen so far are the only candidates: the colour is the smallest candidate
// that none of the missing predecessors takes, so the last wait is on those
// few words alone.
//
// The earliest uncoloured position has all its predecessors coloured and its
// wave is at it (all of the wave's earlier positions are coloured), so the
// launch completes once every wave is resident; every wait is also bounded by
// the wall clock (ctl->err, and the caller colours with rounds instead).
#include "common.hip.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kQuad = 4;
// 128 edges in flight per vertex (32 per lane, 86 VGPRs): the waves of the
// longest non-hub rows set the pass's length. R-MAT 22 orientation with 4 /
// 8 / 16 / 32 steps: 1.43 / 1.35 / 1.30 / 1.27 ms (64 steps: 161 VGPRs)
constexpr int kQuadSteps = 32;
constexpr uint32_t kHubEdges = GCXK_HUB_CHUNK_EDGES;
constexpr int kHubPer = (int)(kHubEdges / kBlock);
constexpr int kSlots = 16;                  // predecessor positions per lane (1024 per wave)
constexpr uint32_t kBatch = kSlots * kWave; // largest list kept in registers
constexpr uint32_t kWin = 4096;             // colours per bitmap window
constexpr uint32_t kWinWords = kWin / 32;
constexpr uint32_t kWaveWords = kWinWords + kWave; // bitmap + a dummy word per lane
constexpr uint32_t kFast = 16;              // missing predecessors polled word by word
constexpr uint32_t kRegather = 8;           // latest-missing polls between re-gathers
constexpr uint32_t kClockPoll = 64;         // polls before the wall clock is read
constexpr uint32_t kSingleRounds = 16;      // positions per wave coloured one at a time first
constexpr uint32_t kChunkPos = 4;           // consecutive positions per chunk after that
constexpr uint32_t kQuarter = kWave / kChunkPos; // lanes per position in a chunk's groups
constexpr uint32_t kSmall = 64;             // most predecessors a group position may have

__device__ __forceinline__ uint32_t load_col(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Orders the wave's LDS accesses (lanes of one wave share the bitmap).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int W>
__device__ __forceinline__ uint32_t wave_sum(uint32_t x)
{
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1)
        x += __shfl_xor(x, off);
    return x;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t x)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const uint32_t y = __shfl_xor(x, off);
        x = y > x ? y : x;
    }
    return x;
}

// ---------------------------------------------------------------- prepare

__global__ __launch_bounds__(kBlock) void pull_clear(uint32_t nhub, uint32_t *__restrict__ hcnt,
                                                     gcxk_ctl *ctl)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t < nhub)
        hcnt[t] = 0u;
    if (t < sizeof(gcxk_ctl) / 4)
        reinterpret_cast<uint32_t *>(ctl)[t] = 0u;
}

// rank[ids[i]] = i, and position i's colour slot cleared (in order here,
// rather than at a random position per vertex in the orientation pass).
__global__ __launch_bounds__(kBlock) void pull_rank(uint32_t N, const uint32_t *__restrict__ ids,
                                                    uint32_t *__restrict__ rank,
                                                    uint32_t *__restrict__ pcol)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N) {
        rank[ids[i]] = i;
        pcol[i] = 0u;
    }
}

// Non-hub vertices in degree order (perm[nhub .. N)), 4 lanes per vertex so a
// wave's 16 vertices have near-equal degrees: predecessor positions to the
// front of the row, and the header at the vertex's position.
__global__ __launch_bounds__(kBlock) void pull_orient_quads(uint32_t N, const uint32_t *__restrict__ V,
                                                            const uint32_t *__restrict__ E,
                                                            const uint32_t *__restrict__ rank,
                                                            const uint32_t *__restrict__ perm,
                                                            uint32_t nhub,
                                                            uint32_t *__restrict__ esplit,
                                                            uint2 *__restrict__ hdr)
{
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint32_t pi = (uint32_t)(t / kQuad) + nhub;
    const int gl = (int)(t % kQuad);
    const int gsh = (lane_id() / kQuad) * kQuad; // first lane of the quad
    if (pi >= N) // whole quads leave together
        return;
    const uint32_t u = perm[pi];
    const uint32_t beg = V[u], end = V[u + 1];
    const uint32_t ru = rank[u];
    const uint32_t below = (1u << gl) - 1u;
    uint32_t np = 0;
    for (uint32_t i0 = beg; i0 < end; i0 += kQuad * kQuadSteps) {
        uint32_t v[kQuadSteps], rv[kQuadSteps];
        bool in[kQuadSteps];
#pragma unroll
        for (int s = 0; s < kQuadSteps; s++) {
            const uint32_t i = i0 + kQuad * s + gl;
            in[s] = i < end;
            v[s] = in[s] ? E[i] : 0u;
        }
#pragma unroll
        for (int s = 0; s < kQuadSteps; s++)
            rv[s] = in[s] ? rank[v[s]] : 0u;
#pragma unroll
        for (int s = 0; s < kQuadSteps; s++) {
            const bool pr = in[s] && rv[s] < ru;
            const uint32_t gp = (uint32_t)(__ballot(pr) >> gsh) & 0xFu;
            if (pr)
                esplit[beg + np + __popc(gp & below)] = rv[s];
            np += __popc(gp);
        }
    }
    if (gl == 0)
        hdr[ru] = make_uint2(beg, np);
}

// Hubs: chunk b of hub h covers entries [V[u] + b*kHubEdges, ...); one
// position reservation per chunk in hcnt[h] (the order of a row's
// predecessors is irrelevant), whose final value is the hub's count.
__global__ __launch_bounds__(kBlock) void pull_orient_hub_chunks(const uint32_t *__restrict__ V,
                                                                 const uint32_t *__restrict__ E,
                                                                 const uint32_t *__restrict__ rank,
                                                                 uint32_t *__restrict__ esplit,
                                                                 const uint32_t *__restrict__ hubs,
                                                                 const uint32_t *__restrict__ hchunk,
                                                                 uint32_t nhub,
                                                                 uint32_t *__restrict__ hcnt)
{
    __shared__ uint32_t s_cnt[kWavesPerBlock];
    __shared__ uint32_t s_base;
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
    const uint32_t ru = rank[u];
    const int lane = lane_id(), w = threadIdx.x / kWave;
    uint32_t v[kHubPer], rv[kHubPer];
    bool pr[kHubPer];
#pragma unroll
    for (int q = 0; q < kHubPer; q++) {
        const uint32_t i = cb + (uint32_t)q * kBlock + threadIdx.x;
        v[q] = i < ce ? E[i] : 0u;
    }
    uint32_t np = 0;
#pragma unroll
    for (int q = 0; q < kHubPer; q++) {
        const uint32_t i = cb + (uint32_t)q * kBlock + threadIdx.x;
        rv[q] = i < ce ? rank[v[q]] : 0u;
        pr[q] = i < ce && rv[q] < ru;
        np += pr[q] ? 1u : 0u;
    }
    uint32_t ip = np; // inclusive scan over the wave
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t a = __shfl_up(ip, off);
        if (lane >= off)
            ip += a;
    }
    if (lane == kWave - 1)
        s_cnt[w] = ip;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tp = 0;
        for (int q = 0; q < kWavesPerBlock; q++) {
            const uint32_t c = s_cnt[q];
            s_cnt[q] = tp;
            tp += c;
        }
        s_base = tp ? atomicAdd(&hcnt[h], tp) : 0u;
    }
    __syncthreads();
    uint32_t pos = s_base + s_cnt[w] + ip - np;
#pragma unroll
    for (int q = 0; q < kHubPer; q++)
        if (pr[q])
            esplit[beg + pos++] = rv[q];
}

__global__ __launch_bounds__(kBlock) void pull_hub_headers(const uint32_t *__restrict__ V,
                                                           const uint32_t *__restrict__ rank,
                                                           const uint32_t *__restrict__ hubs,
                                                           uint32_t nhub,
                                                           const uint32_t *__restrict__ hcnt,
                                                           uint2 *__restrict__ hdr)
{
    const uint32_t h = blockIdx.x * kBlock + threadIdx.x;
    if (h < nhub) {
        const uint32_t u = hubs[h], ru = rank[u];
        hdr[ru] = make_uint2(V[u], hcnt[h]);
    }
}

// ------------------------------------------------------------ colouring

// Smallest colour of the window (1 .. kWin) whose bit is clear in the wave's
// bitmap words (w0: words 0-63, w1: words 64-127, one of each per lane); 0
// when all are set. Marks it set in w0 / w1 of its lane.
__device__ __forceinline__ uint32_t take_free(uint32_t &w0, uint32_t &w1, int lane)
{
    const uint64_t f0 = __ballot(w0 != 0xFFFFFFFFu), f1 = __ballot(w1 != 0xFFFFFFFFu);
    const uint64_t f = f0 ? f0 : f1;
    if (!f)
        return 0u;
    const int fl = __ffsll((long long)f) - 1;
    const uint32_t word = f0 ? __shfl(w0, fl) : __shfl(w1, fl);
    const uint32_t b = (uint32_t)__builtin_ctz(~word);
    if (lane == fl) {
        if (f0)
            w0 |= 1u << b;
        else
            w1 |= 1u << b;
    }
    return 32u * (uint32_t)(f0 ? fl : kWave + fl) + b + 1u;
}

// Marks colour c (window colours lo+1 .. lo+kWin) in the wave's bitmap;
// lanes with nothing to mark OR 0 into a word of their own.
__device__ __forceinline__ void mark(uint32_t *bm, uint32_t c, uint32_t lo, bool on, int lane)
{
    const bool in = on && c > lo && c - lo <= kWin;
    const uint32_t b = c - lo - 1u;
    atomicOr(&bm[in ? b / 32u : kWinWords + (uint32_t)lane], in ? 1u << (b % 32u) : 0u);
}

// Bounded wait bookkeeping: false (and ctl->err set) once the wall clock
// passed its bound or another wave already failed.
__device__ __forceinline__ bool wait_ok(uint32_t &polls, uint64_t &t0, gcxk_ctl *ctl,
                                        uint64_t spin_ticks, int lane)
{
    if (++polls == kClockPoll)
        t0 = __builtin_amdgcn_s_memrealtime();
    if ((polls & 255u) == 0u &&
        (__builtin_amdgcn_s_memrealtime() - t0 > spin_ticks ||
         __hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        if (lane == 0)
            __hip_atomic_store(&ctl->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return false;
    }
    return true;
}

__device__ __forceinline__ void far_sleep(uint32_t nmiss)
{
    if (nmiss > 64u)
        __builtin_amdgcn_s_sleep(32);
    else if (nmiss > 16u)
        __builtin_amdgcn_s_sleep(16);
    else
        __builtin_amdgcn_s_sleep(4);
}

// Colour of a position with 1 <= np <= 64*S predecessors at esplit[beg ..),
// or 0 when a wait hit its bound.
template <int S>
__device__ __forceinline__ uint32_t pull_list(uint32_t beg, uint32_t np, uint32_t *bm,
                                              const uint32_t *__restrict__ esplit,
                                              const uint32_t *pcol, gcxk_ctl *ctl,
                                              uint64_t spin_ticks, int lane)
{
    uint32_t p[S], cv[S];
    const uint32_t last = beg + np - 1u;
    // unconditional loads (slots past the list re-read its last entry), so
    // the same number is in flight on every path
#pragma unroll
    for (int s = 0; s < S; s++) {
        const uint32_t i = beg + (uint32_t)(kWave * s + lane);
        p[s] = esplit[i < last ? i : last];
    }
#pragma unroll
    for (int s = 0; s < S; s++)
        cv[s] = load_col(&pcol[p[s]]);
    bm[lane] = 0u;
    bm[lane + kWave] = 0u;
    wave_sync();
    uint32_t miss = 0;
#pragma unroll
    for (int s = 0; s < S; s++) {
        const bool valid = (uint32_t)(kWave * s + lane) < np;
        miss |= (valid && cv[s] == 0u) ? 1u << s : 0u;
        mark(bm, cv[s], 0u, valid && cv[s] != 0u, lane);
    }
    uint32_t polls = 0;
    uint64_t t0 = 0;
    uint32_t nmiss;
    for (;;) {
        nmiss = wave_sum<kWave>((uint32_t)__popc(miss));
        if (nmiss <= kFast)
            break;
        // many missing: poll the latest of them in greedy order
        uint32_t L = 0;
#pragma unroll
        for (int s = 0; s < S; s++)
            L = ((miss >> s) & 1u) && p[s] > L ? p[s] : L;
        L = wave_max(L);
        for (uint32_t q = 0; q < kRegather; q++) {
            const uint32_t cl = lane == 0 ? load_col(&pcol[L]) : 0u;
            if (__shfl(cl, 0))
                break;
            far_sleep(nmiss);
            if (!wait_ok(polls, t0, ctl, spin_ticks, lane))
                return 0u;
        }
#pragma unroll
        for (int s = 0; s < S; s++)
            if ((miss >> s) & 1u)
                cv[s] = load_col(&pcol[p[s]]);
#pragma unroll
        for (int s = 0; s < S; s++) {
            const bool now = ((miss >> s) & 1u) && cv[s] != 0u;
            miss &= now ? ~(1u << s) : ~0u;
            mark(bm, cv[s], 0u, now, lane);
        }
    }
    wave_sync();
    uint32_t w0 = bm[lane], w1 = bm[lane + kWave];
    if (nmiss == 0u)
        return take_free(w0, w1, lane);
    // candidates: the nmiss + 1 smallest colours absent so far (all within
    // the window: np + 1 <= 1025 < kWin)
    // lane k holds the k-th candidate
    uint32_t cand = 0;
    for (uint32_t k = 0; k <= nmiss; k++) {
        const uint32_t c = take_free(w0, w1, lane);
        cand = (uint32_t)lane == k ? c : cand;
    }
    // lane k < nmiss polls the k-th missing predecessor
    uint32_t mine = 0, k = 0;
#pragma unroll
    for (int s = 0; s < S; s++) {
        uint64_t b = __ballot((miss >> s) & 1u);
        while (b) {
            const int l = __ffsll((long long)b) - 1;
            const uint32_t x = __shfl(p[s], l);
            mine = (uint32_t)lane == k ? x : mine;
            k++;
            b &= b - 1ull;
        }
    }
    const bool poller = (uint32_t)lane < nmiss;
    uint32_t got = 0;
    for (;;) {
        if (poller && got == 0u)
            got = load_col(&pcol[mine]);
        if (__ballot(poller && got == 0u) == 0ull)
            break;
        __builtin_amdgcn_s_sleep(1);
        if (!wait_ok(polls, t0, ctl, spin_ticks, lane))
            return 0u;
    }
    for (uint32_t q = 0; q <= nmiss; q++) {
        const uint32_t c = __builtin_amdgcn_readlane(cand, q);
        if (__ballot(poller && got == c) == 0ull)
            return c;
    }
    return 0u; // unreachable: nmiss colours exclude at most nmiss candidates
}

// Any number of predecessors: batches of kBatch, each waited for in full, per
// colour window of kWin (a further window only when the first is full, which
// needs np >= kWin). 0 when a wait hit its bound.
__device__ __noinline__ uint32_t pull_any(uint32_t beg, uint32_t np, uint32_t *bm,
                                          const uint32_t *__restrict__ esplit, const uint32_t *pcol,
                                          gcxk_ctl *ctl, uint64_t spin_ticks, int lane)
{
    uint32_t polls = 0;
    uint64_t t0 = 0;
    const uint32_t last = beg + np - 1u;
    for (uint32_t lo = 0;; lo += kWin) {
        bm[lane] = 0u;
        bm[lane + kWave] = 0u;
        wave_sync();
        for (uint32_t b0 = 0; b0 < np; b0 += kBatch) {
            uint32_t p[kSlots], cv[kSlots];
#pragma unroll
            for (int s = 0; s < kSlots; s++) {
                const uint32_t i = beg + b0 + (uint32_t)(kWave * s + lane);
                p[s] = esplit[i < last ? i : last];
            }
            uint32_t miss = 0;
#pragma unroll
            for (int s = 0; s < kSlots; s++)
                miss |= b0 + (uint32_t)(kWave * s + lane) < np ? 1u << s : 0u;
            // gather, then while any is missing poll the latest missing one
            // in greedy order and re-gather when it is coloured or every
            // kRegather polls (as pull_list's slow phase)
            for (;;) {
#pragma unroll
                for (int s = 0; s < kSlots; s++)
                    if ((miss >> s) & 1u)
                        cv[s] = load_col(&pcol[p[s]]);
#pragma unroll
                for (int s = 0; s < kSlots; s++) {
                    const bool now = ((miss >> s) & 1u) && cv[s] != 0u;
                    miss &= now ? ~(1u << s) : ~0u;
                    mark(bm, cv[s], lo, now, lane);
                }
                const uint32_t nmiss = wave_sum<kWave>((uint32_t)__popc(miss));
                if (nmiss == 0u)
                    break;
                uint32_t L = 0;
#pragma unroll
                for (int s = 0; s < kSlots; s++)
                    L = ((miss >> s) & 1u) && p[s] > L ? p[s] : L;
                L = wave_max(L);
                for (uint32_t q = 0; q < kRegather; q++) {
                    const uint32_t cl = lane == 0 ? load_col(&pcol[L]) : 0u;
                    if (__shfl(cl, 0))
                        break;
                    far_sleep(nmiss);
                    if (!wait_ok(polls, t0, ctl, spin_ticks, lane))
                        return 0u;
                }
            }
        }
        wave_sync();
        uint32_t w0 = bm[lane], w1 = bm[lane + kWave];
        const uint32_t c = take_free(w0, w1, lane);
        if (c)
            return lo + c;
    }
}

// Colour of one position with a whole wave (any predecessor count); 0 when
// a wait hit its bound.
__device__ __forceinline__ uint32_t pull_one(uint32_t beg, uint32_t np, uint32_t *bm,
                                             const uint32_t *__restrict__ esplit,
                                             const uint32_t *pcol, gcxk_ctl *ctl,
                                             uint64_t spin_ticks, int lane)
{
    if (np == 0u)
        return 1u;
    if (np <= kWave)
        return pull_list<1>(beg, np, bm, esplit, pcol, ctl, spin_ticks, lane);
    if (np <= 2u * kWave)
        return pull_list<2>(beg, np, bm, esplit, pcol, ctl, spin_ticks, lane);
    if (np <= 4u * kWave)
        return pull_list<4>(beg, np, bm, esplit, pcol, ctl, spin_ticks, lane);
    if (np <= 8u * kWave)
        return pull_list<8>(beg, np, bm, esplit, pcol, ctl, spin_ticks, lane);
    if (np <= kBatch)
        return pull_list<kSlots>(beg, np, bm, esplit, pcol, ctl, spin_ticks, lane);
    return pull_any(beg, np, bm, esplit, pcol, ctl, spin_ticks, lane);
}

// Up to four consecutive positions at once, 16 lanes each (groups with
// active = false take no part), every one with at most kSmall = 64
// predecessors: lane g of a group loads predecessors g, g + 16, g + 32,
// g + 48 and waits for their colours; the mex (<= 65) is the first clear bit
// of the group's OR of 1 << (colour - 1), or 65 when colours 1..64 are all
// taken. A group stores its colour as soon as it has it (a later position of
// the chunk may wait for it). False when a wait hit its bound.
__device__ __forceinline__ bool pull_quarter(bool active, uint32_t jq, uint32_t beg, uint32_t np,
                                             uint32_t u, const uint32_t *__restrict__ esplit,
                                             uint32_t *pcol, int *__restrict__ color,
                                             gcxk_ctl *ctl, uint64_t spin_ticks, int lane)
{
    constexpr int S = (int)(kSmall / kQuarter);
    const uint32_t g = (uint32_t)lane & (kQuarter - 1u);
    const uint32_t sh = (uint32_t)lane & ~(kQuarter - 1u);
    uint32_t pp[S], cv[S];
    uint32_t in = 0;
#pragma unroll
    for (int t = 0; t < S; t++)
        in |= active && g + kQuarter * (uint32_t)t < np ? 1u << t : 0u;
#pragma unroll
    for (int t = 0; t < S; t++)
        pp[t] = (in >> t) & 1u ? esplit[beg + g + kQuarter * (uint32_t)t] : 0u;
#pragma unroll
    for (int t = 0; t < S; t++)
        cv[t] = (in >> t) & 1u ? load_col(&pcol[pp[t]]) : 0u;
    bool stored = !active;
    uint32_t polls = 0;
    uint64_t t0 = 0;
    for (;;) {
        uint32_t miss = 0;
        uint64_t m = 0;
#pragma unroll
        for (int t = 0; t < S; t++) {
            const bool on = (in >> t) & 1u;
            miss |= on && cv[t] == 0u ? 1u << t : 0u;
            m |= on && cv[t] != 0u && cv[t] <= 64u ? 1ull << (cv[t] - 1u) : 0ull;
        }
        const uint32_t gm = (uint32_t)(__ballot(miss != 0u) >> sh) & ((1u << kQuarter) - 1u);
#pragma unroll
        for (int off = (int)kQuarter / 2; off > 0; off >>= 1) {
            const uint32_t lo = __shfl_xor((uint32_t)m, off), hi = __shfl_xor((uint32_t)(m >> 32), off);
            m |= ((uint64_t)hi << 32) | lo;
        }
        if (gm == 0u && !stored) {
            const uint32_t c = ~m ? (uint32_t)__builtin_ctzll(~m) + 1u : 65u;
            if (g == 0u) {
                __hip_atomic_store(&pcol[jq], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                color[u] = (int)c;
            }
            stored = true;
        }
        if (__ballot(!stored) == 0ull)
            return true;
        __builtin_amdgcn_s_sleep(1);
        if (!wait_ok(polls, t0, ctl, spin_ticks, lane))
            return false;
#pragma unroll
        for (int t = 0; t < S; t++)
            if ((miss >> t) & 1u)
                cv[t] = load_col(&pcol[pp[t]]);
    }
}

// Positions [0, T), T = min(N, kSingleRounds * W): one position per wave at a
// time, wave w taking w, w + W, ... Positions [T, N): chunks of kChunkPos
// consecutive positions, wave w taking chunks w, w + W, ... (runs of 16 to
// 256 consecutive positions per wave were 1.3-20x slower on R-MAT 22: a
// wave's run waits on the runs before it); within a chunk
// the positions with at most kSmall predecessors are coloured by 16-lane
// groups at once and the others by the whole wave, in order. Each wave meets
// its positions in increasing order either way.
__global__ __launch_bounds__(kBlock, 4) void jp_pull_kernel(uint32_t N, const uint2 *__restrict__ hdr,
                                                         const uint32_t *__restrict__ ids,
                                                         const uint32_t *__restrict__ esplit,
                                                         uint32_t *pcol, int *__restrict__ color,
                                                         gcxk_ctl *ctl, uint64_t spin_ticks)
{
    __shared__ uint32_t s_bm[kWavesPerBlock][kWaveWords];
    const int lane = lane_id(), w = threadIdx.x / kWave;
    uint32_t *bm = s_bm[w];
    const uint32_t n = ctl->err ? 0u : N;
    // Wave ids run across workgroups first: consecutive positions go to
    // consecutive workgroups, which the dispatcher places on the XCDs in turn
    // (R-MAT 22: 4.41 -> 4.17 ms against four consecutive positions per
    // workgroup; keeping runs of consecutive positions on one XCD: 5.62 ms).
    const uint32_t w0 = (uint32_t)w * gridDim.x + blockIdx.x;
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    const uint32_t T = (uint64_t)kSingleRounds * nw < n ? kSingleRounds * nw : n;
    uint2 h = w0 < T ? hdr[w0] : make_uint2(0u, 0u);
    uint32_t u = w0 < T ? ids[w0] : 0u;
    for (uint32_t j = w0; j < T; j += nw) {
        // the next position's header and id are loaded a position ahead
        const uint32_t jn = j + nw;
        const uint2 hn = jn < T ? hdr[jn] : make_uint2(0u, 0u);
        const uint32_t un = jn < T ? ids[jn] : 0u;
        const uint32_t c = pull_one(h.x, h.y, bm, esplit, pcol, ctl, spin_ticks, lane);
        if (c == 0u)
            return; // a wait hit its bound: ctl->err is set
        if (lane == 0) {
            __hip_atomic_store(&pcol[j], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            color[u] = (int)c;
        }
        h = hn;
        u = un;
    }
    // chunks: group q = lane / kQuarter holds position c0 + q
    const uint32_t q = (uint32_t)lane / kQuarter;
    const uint32_t stride = kChunkPos * nw;
    uint32_t c0 = T + kChunkPos * w0;
    uint32_t jq = c0 + q;
    h = c0 < n && jq < n ? hdr[jq] : make_uint2(0u, 0u);
    u = c0 < n && jq < n ? ids[jq] : 0u;
    for (; c0 < n; c0 += stride) {
        jq = c0 + q;
        const uint32_t cn = c0 + stride, jn = cn + q;
        const uint2 hn = cn < n && jn < n ? hdr[jn] : make_uint2(0u, 0u);
        const uint32_t un = cn < n && jn < n ? ids[jn] : 0u;
        // large positions (more than kSmall predecessors) take the whole
        // wave one at a time; the small ones between them go by groups, so
        // every group phase only holds positions before the next large one
        const uint64_t lm = __ballot(jq < n && h.y > kSmall);
        uint32_t large = 0; // bit q: position c0 + q is large
#pragma unroll
        for (uint32_t k = 0; k < kChunkPos; k++)
            large |= (uint32_t)((lm >> (k * kQuarter)) & 1ull) << k;
        for (uint32_t k = 0; k < kChunkPos && c0 + k < n;) {
            const uint32_t rest = large >> k;
            const uint32_t f = rest ? k + (uint32_t)__builtin_ctz(rest) : kChunkPos;
            if (f > k) { // small positions k .. f - 1
                const bool act = q >= k && q < f && jq < n;
                if (!pull_quarter(act, jq, h.x, h.y, u, esplit, pcol, color, ctl, spin_ticks,
                                  lane))
                    return;
            }
            if (f < kChunkPos && c0 + f < n) {
                const uint32_t src = f * kQuarter;
                const uint32_t beg = __shfl(h.x, src), np = __shfl(h.y, src),
                               uf = __shfl(u, src);
                const uint32_t c = pull_one(beg, np, bm, esplit, pcol, ctl, spin_ticks, lane);
                if (c == 0u)
                    return;
                if (lane == 0) {
                    __hip_atomic_store(&pcol[c0 + f], c, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    color[uf] = (int)c;
                }
            }
            k = f + 1u;
        }
        h = hn;
        u = un;
    }
}

uint32_t g_pull_grid = 0, g_pull_cap = 0;

} // namespace

extern "C" int gcxk_pull_rank(uint32_t N, const uint32_t *ids, uint32_t *rank, uint32_t *pcol,
                              uint32_t nhub, uint32_t *hcnt, gcxk_ctl *ctl, void *stream)
{
    hipStream_t s = as_stream(stream);
    const uint32_t words = (uint32_t)(sizeof(gcxk_ctl) / 4);
    pull_clear<<<div_up(nhub > words ? nhub : words, kBlock), kBlock, 0, s>>>(nhub, hcnt, ctl);
    GCXK_LAUNCH_CHECK();
    if (N) {
        pull_rank<<<div_up(N, kBlock), kBlock, 0, s>>>(N, ids, rank, pcol);
        GCXK_LAUNCH_CHECK();
    }
    return GCX_OK;
}

extern "C" int gcxk_pull_orient(uint32_t N, const uint32_t *dV, const uint32_t *dE,
                                const uint32_t *rank, const uint32_t *perm, const uint32_t *hubs,
                                uint32_t nhub, const uint32_t *hchunk, uint32_t nhchunk,
                                uint32_t *hcnt, uint32_t *esplit, void *hdr, void *stream,
                                void *hub_stream)
{
    if (N == 0)
        return GCX_OK;
    uint2 *h = static_cast<uint2 *>(hdr);
    const uint64_t quads = (uint64_t)(N - nhub) * kQuad;
    if (quads) {
        pull_orient_quads<<<div_up(quads, kBlock), kBlock, 0, as_stream(stream)>>>(
            N, dV, dE, rank, perm, nhub, esplit, h);
        GCXK_LAUNCH_CHECK();
    }
    if (nhub) {
        hipStream_t hs = as_stream(hub_stream);
        if (nhchunk) {
            pull_orient_hub_chunks<<<nhchunk, kBlock, 0, hs>>>(dV, dE, rank, esplit, hubs, hchunk,
                                                               nhub, hcnt);
            GCXK_LAUNCH_CHECK();
        }
        pull_hub_headers<<<div_up(nhub, kBlock), kBlock, 0, hs>>>(dV, rank, hubs, nhub, hcnt, h);
        GCXK_LAUNCH_CHECK();
    }
    return GCX_OK;
}

extern "C" uint32_t gcxk_pull_grid(void)
{
    if (!g_pull_grid) {
        int per_cu = 0, dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, jp_pull_kernel, kBlock, 0) !=
                hipSuccess ||
            per_cu <= 0)
            return 0;
        // one resident workgroup per CU fewer than fit (3 of 4 at 128
        // VGPRs): a cooperative launch that needs every resident slot cannot
        // become resident while other streams' kernels hold some (8 serving
        // workers: 0.8 jobs/s on R-MAT 22 at 4 per CU)
        const uint32_t cus = (uint32_t)prop.multiProcessorCount;
        const uint32_t cap = (uint32_t)per_cu * cus;
        g_pull_grid = per_cu > 1 ? cap - cus : cap;
        g_pull_cap = cap;
    }
    return g_pull_grid;
}

extern "C" uint32_t gcxk_pull_cap(void) { return gcxk_pull_grid() ? g_pull_cap : 0; }

extern "C" int gcxk_pull_colour(uint32_t N, const void *hdr, const uint32_t *ids,
                                const uint32_t *esplit, uint32_t *pcol, int *color, gcxk_ctl *ctl,
                                uint32_t grid, uint64_t spin_ticks, int cooperative, void *stream)
{
    if (N == 0)
        return GCX_OK;
    const uint32_t dflt = gcxk_pull_grid();
    if (!dflt)
        return GCX_ERR_HIP;
    if (grid == 0)
        grid = dflt;
    if (grid > g_pull_cap)
        grid = g_pull_cap;
    const uint2 *h = static_cast<const uint2 *>(hdr);
    void *args[] = {(void *)&N,    (void *)&h,     (void *)&ids, (void *)&esplit,
                    (void *)&pcol, (void *)&color, (void *)&ctl, (void *)&spin_ticks};
    if (cooperative) {
        GCXK_CHECK(hipLaunchCooperativeKernel((const void *)jp_pull_kernel, dim3(grid),
                                              dim3(kBlock), args, 0, as_stream(stream)));
    } else {
        // grid <= the resident capacity: once the kernels of other streams
        // (none of which wait on anything) drain, every workgroup is resident
        GCXK_CHECK(hipLaunchKernel((const void *)jp_pull_kernel, dim3(grid), dim3(kBlock), args, 0,
                                   as_stream(stream)));
    }
    return GCX_OK;
}
