// jp.hip -- Jones-Plassmann colouring, proper-colouring check and the
// greedy-rule certificate for gfx950.
//
// Semantics (reference coloring.c:223-330): v precedes u iff
// p[v] > p[u] || (p[v] == p[u] && v < u); each vertex takes the smallest
// colour >= 1 unused by its predecessors. The result equals sequential greedy
// in (p descending, id ascending) order, so it is independent of how rounds
// are scheduled here.
//
// Round structure: a frontier of ready vertices (all predecessors coloured)
// is an independent set, so within a round a neighbour is a predecessor iff
// it already has a colour, and a successor iff its colour is still 0. The
// round kernel therefore needs no priority gathers: it reads E and colours,
// takes the mex of the non-zero colours, and decrements the successors'
// pending-predecessor counts; a count reaching zero queues that vertex for
// the next round. Frontier vertices are coloured one per wave (64 lanes
// stride the row); hubs (degree >= GCXK_HUB_DEG) one per workgroup with an
// LDS colour bitmap.
#include "common.hip.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kGroup = 16;             // lanes per vertex in whole-graph passes
constexpr int kHubWords = 1024;        // LDS bitmap: 32K colours per window

// Wave-aggregated append of `v` (where `want`) to queue q with counter *n.
__device__ __forceinline__ void wave_push(bool want, uint32_t v, uint32_t *q, uint32_t *n)
{
    const uint64_t b = __ballot(want);
    if (b == 0)
        return;
    const int leader = __ffsll((long long)b) - 1;
    uint32_t base = 0;
    if (lane_id() == leader)
        base = atomicAdd(n, (uint32_t)__popcll(b));
    base = __shfl(base, leader);
    if (want)
        q[base + __popcll(b & lanemask_lt())] = v;
}

// Decrement v's pending count; return true when this call released it.
__device__ __forceinline__ bool release(uint32_t *count, uint32_t v)
{
    return atomicSub(count + v, 1u) == 1u;
}

template <typename K>
__global__ __launch_bounds__(kBlock) void jp_init_groups(uint32_t N, const uint32_t *__restrict__ V,
                                                         const uint32_t *__restrict__ E,
                                                         const K *__restrict__ key,
                                                         uint32_t *__restrict__ count,
                                                         int *__restrict__ color,
                                                         uint32_t *__restrict__ q0, gcxk_ctl *ctl)
{
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint32_t u = (uint32_t)(t / kGroup);
    const int gl = (int)(t % kGroup);
    const bool active = u < N;
    uint32_t beg = 0, end = 0;
    if (active) {
        beg = V[u];
        end = V[u + 1];
    }
    const bool hub = (end - beg) >= GCXK_HUB_DEG;
    uint32_t cnt = 0;
    if (active && !hub) {
        const K ku = key[u];
        for (uint32_t i = beg + gl; i < end; i += kGroup) {
            const uint32_t v = E[i];
            cnt += precedes(key[v], v, ku, u) ? 1u : 0u;
        }
    }
    cnt = group_sum32<kGroup>(cnt);
    const bool lead = active && !hub && gl == 0;
    if (lead) {
        count[u] = cnt;
        color[u] = 0;
    }
    wave_push(lead && cnt == 0, u, q0, &ctl->qn[0]);
}

template <typename K>
__global__ __launch_bounds__(kBlock) void jp_init_hubs(const uint32_t *__restrict__ V,
                                                       const uint32_t *__restrict__ E,
                                                       const K *__restrict__ key,
                                                       uint32_t *__restrict__ count,
                                                       int *__restrict__ color,
                                                       uint32_t *__restrict__ h0,
                                                       const uint32_t *__restrict__ hubs,
                                                       gcxk_ctl *ctl)
{
    __shared__ uint32_t part[kWavesPerBlock];
    const uint32_t u = hubs[blockIdx.x];
    const uint32_t beg = V[u], end = V[u + 1];
    const K ku = key[u];
    uint32_t cnt = 0;
    for (uint32_t i = beg + threadIdx.x; i < end; i += kBlock) {
        const uint32_t v = E[i];
        cnt += precedes(key[v], v, ku, u) ? 1u : 0u;
    }
    cnt = group_sum32<kWave>(cnt);
    if (lane_id() == 0)
        part[threadIdx.x / kWave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int w = 0; w < kWavesPerBlock; w++)
            s += part[w];
        count[u] = s;
        color[u] = 0;
        if (s == 0)
            h0[atomicAdd(&ctl->hn[0], 1u)] = u;
    }
}

// Colour one frontier vertex with the whole wave. All lanes call this.
__device__ void colour_by_wave(uint32_t u, const uint32_t *__restrict__ V,
                               const uint32_t *__restrict__ E, uint32_t *__restrict__ count,
                               int *__restrict__ color, uint32_t *qout, uint32_t *qn,
                               uint32_t *hout, uint32_t *hn)
{
    const uint32_t beg = V[u], end = V[u + 1];
    const int lane = lane_id();
    uint64_t mask = 0;
    // Pass 1: colours 1..64 into the mask, successors released.
    for (uint32_t i0 = beg; i0 < end; i0 += kWave) {
        const uint32_t i = i0 + lane;
        bool small = false, big = false;
        uint32_t v = 0;
        if (i < end) {
            v = E[i];
            const int c = color[v];
            if (c == 0) {
                if (release(count, v)) {
                    const bool is_hub = (V[v + 1] - V[v]) >= GCXK_HUB_DEG;
                    big = is_hub;
                    small = !is_hub;
                }
            } else if (c <= 64) {
                mask |= 1ull << (c - 1);
            }
        }
        wave_push(small, v, qout, qn);
        wave_push(big, v, hout, hn);
    }
    mask = group_or64<kWave>(mask);
    uint32_t base = 0;
    // Further windows only when colours 1..64 are all taken.
    while (mask == ~0ull) {
        base += 64;
        mask = 0;
        for (uint32_t i = beg + lane; i < end; i += kWave) {
            const uint32_t c = (uint32_t)color[E[i]];
            if (c > base && c <= base + 64)
                mask |= 1ull << (c - base - 1);
        }
        mask = group_or64<kWave>(mask);
    }
    if (lane == 0)
        color[u] = (int)(base + (uint32_t)__builtin_ctzll(~mask) + 1u);
}

// Colour one hub with the whole workgroup (LDS bitmap window of 32K colours).
__device__ void colour_by_block(uint32_t u, const uint32_t *__restrict__ V,
                                const uint32_t *__restrict__ E, uint32_t *__restrict__ count,
                                int *__restrict__ color, uint32_t *qout, uint32_t *qn,
                                uint32_t *hout, uint32_t *hn, uint32_t *bitmap, uint32_t *found)
{
    const uint32_t beg = V[u], end = V[u + 1];
    uint32_t base = 0;
    bool first = true;
    for (;;) {
        for (int w = threadIdx.x; w < kHubWords; w += kBlock)
            bitmap[w] = 0;
        if (threadIdx.x == 0)
            *found = UINT32_MAX;
        __syncthreads();
        for (uint32_t i0 = beg; i0 < end; i0 += kBlock) {
            const uint32_t i = i0 + threadIdx.x;
            bool small = false, big = false;
            uint32_t v = 0;
            if (i < end) {
                v = E[i];
                const uint32_t c = (uint32_t)color[v];
                if (c == 0) {
                    if (first && release(count, v)) {
                        const bool is_hub = (V[v + 1] - V[v]) >= GCXK_HUB_DEG;
                        big = is_hub;
                        small = !is_hub;
                    }
                } else if (c > base && c - base - 1 < 32u * kHubWords) {
                    const uint32_t b = c - base - 1;
                    atomicOr(bitmap + (b >> 5), 1u << (b & 31));
                }
            }
            wave_push(small, v, qout, qn);
            wave_push(big, v, hout, hn);
        }
        __syncthreads();
        for (int w = threadIdx.x; w < kHubWords; w += kBlock)
            if (bitmap[w] != 0xFFFFFFFFu)
                atomicMin(found, (uint32_t)w);
        __syncthreads();
        const uint32_t w = *found;
        if (w != UINT32_MAX) {
            if (threadIdx.x == 0)
                color[u] = (int)(base + 32u * w + (uint32_t)__builtin_ctz(~bitmap[w]) + 1u);
            __syncthreads();
            return;
        }
        base += 32u * kHubWords;
        first = false;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void jp_round_kernel(const uint32_t *__restrict__ V,
                                                          const uint32_t *__restrict__ E,
                                                          uint32_t *__restrict__ count,
                                                          int *__restrict__ color,
                                                          const uint32_t *__restrict__ qin,
                                                          uint32_t *__restrict__ qout,
                                                          const uint32_t *__restrict__ hin,
                                                          uint32_t *__restrict__ hout,
                                                          gcxk_ctl *ctl, uint32_t r)
{
    __shared__ uint32_t bitmap[kHubWords];
    __shared__ uint32_t found;
    const uint32_t si = r % 3, so = (r + 1) % 3, sr = (r + 2) % 3;
    const uint32_t ns = ctl->qn[si], nh = ctl->hn[si];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // Slot sr is the input of round r-1 and the output of round r+1.
        ctl->qn[sr] = 0;
        ctl->hn[sr] = 0;
        if (ns + nh != 0)
            ctl->rounds = r + 1;
        else
            ctl->done = 1;
    }
    if (ns + nh == 0)
        return;
    for (uint32_t h = blockIdx.x; h < nh; h += gridDim.x)
        colour_by_block(hin[h], V, E, count, color, qout, &ctl->qn[so], hout, &ctl->hn[so], bitmap,
                        &found);
    const uint32_t w0 = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    for (uint32_t k = w0; k < ns; k += nw)
        colour_by_wave(qin[k], V, E, count, color, qout, &ctl->qn[so], hout, &ctl->hn[so]);
}

// ---------------------------------------------------------------- validate

__global__ __launch_bounds__(kBlock) void validate_groups(uint32_t N, const uint32_t *__restrict__ V,
                                                          const uint32_t *__restrict__ E,
                                                          const int *__restrict__ color,
                                                          gcxk_ctl *ctl)
{
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint32_t u = (uint32_t)(t / kGroup);
    const int gl = (int)(t % kGroup);
    uint32_t bad = 0, top = 0;
    if (u < N) {
        const uint32_t beg = V[u], end = V[u + 1];
        const int cu = color[u];
        if (end - beg < GCXK_HUB_DEG) {
            for (uint32_t i = beg + gl; i < end; i += kGroup)
                bad += (color[E[i]] == cu) ? 1u : 0u;
        }
        if (gl == 0) {
            bad += cu <= 0 ? 1u : 0u;
            top = cu > 0 ? (uint32_t)cu : 0u;
        }
    }
    bad = group_sum32<kWave>(bad);
    top = group_max32<kWave>(top);
    if (lane_id() == 0) {
        if (bad)
            atomicAdd(&ctl->bad, bad);
        atomicMax(&ctl->maxcolor, (int)top);
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

template <typename K>
__global__ __launch_bounds__(kBlock) void mex_check_kernel(uint32_t N, const uint32_t *__restrict__ V,
                                                           const uint32_t *__restrict__ E,
                                                           const K *__restrict__ key,
                                                           const int *__restrict__ color,
                                                           gcxk_ctl *ctl)
{
    const uint32_t u = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (u >= N)
        return;
    const int lane = lane_id();
    const uint32_t beg = V[u], end = V[u + 1];
    const K ku = key[u];
    uint32_t base = 0;
    uint64_t mask;
    for (;;) {
        mask = 0;
        for (uint32_t i = beg + lane; i < end; i += kWave) {
            const uint32_t v = E[i];
            if (precedes(key[v], v, ku, u)) {
                const uint32_t c = (uint32_t)color[v];
                if (c > base && c <= base + 64)
                    mask |= 1ull << (c - base - 1);
            }
        }
        mask = group_or64<kWave>(mask);
        if (mask != ~0ull)
            break;
        base += 64;
    }
    const uint32_t mex = base + (uint32_t)__builtin_ctzll(~mask) + 1u;
    if (lane == 0 && (uint32_t)color[u] != mex)
        atomicAdd((unsigned long long *)&ctl->viol, 1ull);
}

} // namespace

extern "C" uint32_t gcxk_round_grid(void) { return 1024u; }

extern "C" int gcxk_jp_init(uint32_t N, const uint32_t *dV, const uint32_t *dE, const void *key,
                            int key_bits, uint32_t *count, int *color, uint32_t *q0, uint32_t *h0,
                            const uint32_t *hubs, uint32_t nhub, gcxk_ctl *ctl, void *stream)
{
    if (N == 0)
        return GCX_OK;
    const uint32_t grid = div_up((uint64_t)N * kGroup, kBlock);
    if (key_bits == 32) {
        jp_init_groups<uint32_t><<<grid, kBlock, 0, as_stream(stream)>>>(
            N, dV, dE, (const uint32_t *)key, count, color, q0, ctl);
        if (nhub)
            jp_init_hubs<uint32_t><<<nhub, kBlock, 0, as_stream(stream)>>>(
                dV, dE, (const uint32_t *)key, count, color, h0, hubs, ctl);
    } else {
        jp_init_groups<uint64_t><<<grid, kBlock, 0, as_stream(stream)>>>(
            N, dV, dE, (const uint64_t *)key, count, color, q0, ctl);
        if (nhub)
            jp_init_hubs<uint64_t><<<nhub, kBlock, 0, as_stream(stream)>>>(
                dV, dE, (const uint64_t *)key, count, color, h0, hubs, ctl);
    }
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_jp_round(uint32_t N, const uint32_t *dV, const uint32_t *dE, uint32_t *count,
                             int *color, uint32_t *const q[2], uint32_t *const h[2], gcxk_ctl *ctl,
                             uint32_t r, uint32_t grid, void *stream)
{
    (void)N;
    jp_round_kernel<<<grid, kBlock, 0, as_stream(stream)>>>(dV, dE, count, color, q[r & 1],
                                                            q[(r + 1) & 1], h[r & 1],
                                                            h[(r + 1) & 1], ctl, r);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_validate(uint32_t N, const uint32_t *dV, const uint32_t *dE, const int *color,
                             const uint32_t *hubs, uint32_t nhub, gcxk_ctl *ctl, void *stream)
{
    if (N == 0)
        return GCX_OK;
    validate_groups<<<div_up((uint64_t)N * kGroup, kBlock), kBlock, 0, as_stream(stream)>>>(
        N, dV, dE, color, ctl);
    if (nhub)
        validate_hubs<<<nhub, kBlock, 0, as_stream(stream)>>>(dV, dE, color, hubs, ctl);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_mex_check(uint32_t N, const uint32_t *dV, const uint32_t *dE, const void *key,
                              int key_bits, const int *color, gcxk_ctl *ctl, void *stream)
{
    if (N == 0)
        return GCX_OK;
    const uint32_t grid = div_up(N, kWavesPerBlock);
    if (key_bits == 32)
        mex_check_kernel<uint32_t><<<grid, kBlock, 0, as_stream(stream)>>>(
            N, dV, dE, (const uint32_t *)key, color, ctl);
    else
        mex_check_kernel<uint64_t><<<grid, kBlock, 0, as_stream(stream)>>>(
            N, dV, dE, (const uint64_t *)key, color, ctl);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}
