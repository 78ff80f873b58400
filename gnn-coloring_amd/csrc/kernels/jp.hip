// jp.hip -- Jones-Plassmann colouring, proper-colouring check and the
// greedy-rule certificate for gfx950.
//
// Semantics (reference coloring.c:223-330): v precedes u iff
// p[v] > p[u] || (p[v] == p[u] && v < u); each vertex takes the smallest
// colour >= 1 unused by its predecessors. The result equals sequential greedy
// in (p descending, id ascending) order, so it does not depend on how the
// rounds below are scheduled.
//
// Orientation (init): every row is split into predecessors first and
// successors last (E_split, as coloring.c:241-264 does); count[u] = npred[u]
// = number of predecessors; vertices with none are the first frontier.
//
// Push-based rounds: when u takes colour c it ORs bit c-1 into the colour set
// of each successor v and decrements count[v]; the successor's colour set is
// therefore complete when its count reaches zero, and its mex costs one
// 64-bit word (colours 1..64) or a short scan of its overflow words (colours
// 65..deg+1, a per-vertex region at word offset V[v]/32 + v). A ready vertex
// becomes ceil(nsucc/64) work items (at least one); one wave handles one
// item, i.e. up to 64 successor edges, so a hub's successors are spread over
// many waves of the same round. Every item of u recomputes the (identical)
// mex; item 0 stores the colour.
#include "common.hip.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kGroup = 16;             // lanes per vertex in whole-graph passes
constexpr uint32_t kChunk = 64;        // successor edges per work item

__device__ __forceinline__ uint32_t ovf_base(uint32_t Vv, uint32_t v) { return Vv / 32u + v; }

// Overflow words of a vertex of degree d: colours 65 .. d+1.
__device__ __forceinline__ uint32_t ovf_words(uint32_t d)
{
    return d >= 64u ? (d + 1u - 64u + 31u) / 32u : 0u;
}

__device__ __forceinline__ uint32_t n_items(uint32_t nsucc)
{
    return nsucc ? (nsucc + kChunk - 1u) / kChunk : 1u;
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

// Zero colours, colour sets, hub counters and the control block.
__global__ __launch_bounds__(kBlock) void jp_clear(uint32_t N, uint64_t nwords, int *__restrict__ color,
                                                   uint64_t *__restrict__ mask64,
                                                   uint32_t *__restrict__ ovf,
                                                   const uint32_t *__restrict__ hubs, uint32_t nhub,
                                                   uint32_t *__restrict__ count,
                                                   uint32_t *__restrict__ npred,
                                                   uint32_t *__restrict__ hpos, gcxk_ctl *ctl)
{
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    const uint64_t t0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    for (uint64_t i = t0; i < N; i += stride) {
        color[i] = 0;
        mask64[i] = 0ull;
    }
    for (uint64_t i = t0; i < nwords; i += stride)
        ovf[i] = 0u;
    for (uint64_t h = t0; h < nhub; h += stride) {
        count[hubs[h]] = 0u;
        npred[hubs[h]] = 0u;
        hpos[2 * h] = 0u;
        hpos[2 * h + 1] = 0u;
    }
    if (t0 < sizeof(gcxk_ctl) / 4)
        reinterpret_cast<uint32_t *>(ctl)[t0] = 0u;
}

template <typename K>
__global__ __launch_bounds__(kBlock) void jp_init_groups(uint32_t N, const uint32_t *__restrict__ V,
                                                         const uint32_t *__restrict__ E,
                                                         const K *__restrict__ key,
                                                         uint32_t *__restrict__ count,
                                                         uint32_t *__restrict__ npred,
                                                         uint32_t *__restrict__ esplit,
                                                         uint2 *__restrict__ q0, gcxk_ctl *ctl)
{
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint32_t u = (uint32_t)(t / kGroup);
    const int gl = (int)(t % kGroup);
    const int gw = lane_id() / kGroup;
    const bool active = u < N;
    uint32_t beg = 0, end = 0;
    if (active) {
        beg = V[u];
        end = V[u + 1];
    }
    const bool hub = (end - beg) >= GCXK_HUB_DEG;
    uint32_t np = 0, ns = 0;
    if (active && !hub) {
        const K ku = key[u];
        const uint32_t below = (1u << gl) - 1u;
        for (uint32_t i0 = beg; i0 < end; i0 += kGroup) {
            const uint32_t i = i0 + gl;
            const bool in = i < end;
            bool pr = false;
            uint32_t v = 0;
            if (in) {
                v = E[i];
                pr = precedes(key[v], v, ku, u);
            }
            const uint32_t gp = (uint32_t)(__ballot(in && pr) >> (kGroup * gw)) & 0xFFFFu;
            const uint32_t gs = (uint32_t)(__ballot(in && !pr) >> (kGroup * gw)) & 0xFFFFu;
            if (in) {
                if (pr)
                    esplit[beg + np + __popc(gp & below)] = v;
                else
                    esplit[end - 1u - ns - __popc(gs & below)] = v;
            }
            np += __popc(gp);
            ns += __popc(gs);
        }
    }
    const bool lead = active && !hub && gl == 0;
    if (lead) {
        count[u] = np;
        npred[u] = np;
    }
    push_items(lead && np == 0, u, n_items(end - beg - np), q0, &ctl->qn[0]);
}

// Hubs: chunk b of hub h covers edges [V[u] + b*kHubEdges, ...); positions in
// E_split come from per-hub counters (predecessors from the front,
// successors from the back); the order inside each part is irrelevant.
constexpr uint32_t kHubEdges = 4096;

template <typename K>
__global__ __launch_bounds__(kBlock) void jp_init_hub_chunks(const uint32_t *__restrict__ V,
                                                             const uint32_t *__restrict__ E,
                                                             const K *__restrict__ key,
                                                             uint32_t *__restrict__ count,
                                                             uint32_t *__restrict__ npred,
                                                             uint32_t *__restrict__ esplit,
                                                             const uint32_t *__restrict__ hubs,
                                                             const uint32_t *__restrict__ hchunk,
                                                             uint32_t nhub,
                                                             uint32_t *__restrict__ hpos)
{
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
    const int lane = lane_id();
    uint32_t npb = 0;
    for (uint32_t i0 = cb + (threadIdx.x - lane); i0 < ce; i0 += kBlock) {
        const uint32_t i = i0 + lane;
        const bool in = i < ce;
        bool pr = false;
        uint32_t v = 0;
        if (in) {
            v = E[i];
            pr = precedes(key[v], v, ku, u);
        }
        const uint64_t bp = __ballot(in && pr), bs = __ballot(in && !pr);
        uint32_t pbase = 0, sbase = 0;
        if (lane == 0) {
            if (bp)
                pbase = atomicAdd(&hpos[2 * h], (uint32_t)__popcll(bp));
            if (bs)
                sbase = atomicAdd(&hpos[2 * h + 1], (uint32_t)__popcll(bs));
        }
        pbase = __shfl(pbase, 0);
        sbase = __shfl(sbase, 0);
        if (in) {
            if (pr)
                esplit[beg + pbase + __popcll(bp & lanemask_lt())] = v;
            else
                esplit[end - 1u - sbase - __popcll(bs & lanemask_lt())] = v;
        }
        npb += (uint32_t)__popcll(bp);
    }
    if (lane == 0 && npb) {
        atomicAdd(&count[u], npb);
        atomicAdd(&npred[u], npb);
    }
}

__global__ __launch_bounds__(kBlock) void jp_hub_roots(const uint32_t *__restrict__ V,
                                                       const uint32_t *__restrict__ npred,
                                                       const uint32_t *__restrict__ hubs,
                                                       uint32_t nhub, uint2 *__restrict__ q0,
                                                       gcxk_ctl *ctl)
{
    const uint32_t h = blockIdx.x * kBlock + threadIdx.x;
    bool root = false;
    uint32_t u = 0, cnt = 0;
    if (h < nhub) {
        u = hubs[h];
        root = npred[u] == 0;
        cnt = n_items(V[u + 1] - V[u] - npred[u]);
    }
    push_items(root, u, cnt, q0, &ctl->qn[0]);
}

__global__ __launch_bounds__(kBlock) void jp_round_kernel(const uint32_t *__restrict__ V,
                                                          const uint32_t *__restrict__ npred,
                                                          const uint32_t *__restrict__ esplit,
                                                          uint32_t *__restrict__ count,
                                                          int *__restrict__ color,
                                                          uint64_t *__restrict__ mask64,
                                                          uint32_t *__restrict__ ovf,
                                                          const uint2 *__restrict__ qin,
                                                          uint2 *__restrict__ qout, gcxk_ctl *ctl,
                                                          uint32_t r)
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
    const uint32_t w0 = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    for (uint32_t idx = w0; idx < n; idx += nw) {
        const uint2 it = qin[idx];
        const uint32_t u = it.x, k = it.y;
        const uint32_t Vu = V[u], Vu1 = V[u + 1];
        const uint32_t np = npred[u];
        const uint64_t m = mask64[u];
        uint32_t c;
        if (m != ~0ull) {
            c = (uint32_t)__builtin_ctzll(~m) + 1u;
        } else {
            // mex <= npred + 1 <= deg + 1, so a free bit lies inside the region
            const uint32_t base = ovf_base(Vu, u), words = ovf_words(Vu1 - Vu);
            c = 0;
            for (uint32_t wb = 0; wb < words; wb += kWave) {
                const uint32_t wi = wb + lane;
                const uint32_t w = wi < words ? ovf[base + wi] : 0xFFFFFFFFu;
                const uint64_t b = __ballot(w != 0xFFFFFFFFu);
                if (b) {
                    const int fl = __ffsll((long long)b) - 1;
                    const uint32_t wf = __shfl(w, fl);
                    c = 65u + 32u * (wb + (uint32_t)fl) + (uint32_t)__builtin_ctz(~wf);
                    break;
                }
            }
        }
        if (k == 0 && lane == 0)
            color[u] = (int)c;
        const uint32_t i = Vu + np + k * kChunk + (uint32_t)lane;
        bool rel = false;
        uint32_t v = 0;
        if (i < Vu1) {
            v = esplit[i];
            if (c <= 64u) {
                atomicOr((unsigned long long *)&mask64[v], 1ull << (c - 1u));
            } else {
                const uint32_t Vv = V[v], b = c - 65u;
                if (b < 32u * ovf_words(V[v + 1] - Vv))
                    atomicOr(&ovf[ovf_base(Vv, v) + b / 32u], 1u << (b % 32u));
            }
            rel = atomicSub(&count[v], 1u) == 1u;
        }
        uint32_t cnt = 0;
        if (rel)
            cnt = n_items(V[v + 1] - V[v] - npred[v]);
        push_items(rel, v, cnt, qout, &ctl->qn[so]);
    }
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

extern "C" int gcxk_jp_clear(uint32_t N, uint64_t M, int *color, uint64_t *mask64, uint32_t *ovf,
                             const uint32_t *hubs, uint32_t nhub, uint32_t *count,
                             uint32_t *npred, uint32_t *hpos, gcxk_ctl *ctl, void *stream)
{
    const uint64_t nwords = gcxk_jp_ovf_words(N, M);
    uint64_t n = nwords > N ? nwords : N;
    if (n < 64)
        n = 64;
    uint32_t grid = div_up(n, kBlock);
    if (grid > 4096)
        grid = 4096;
    jp_clear<<<grid, kBlock, 0, as_stream(stream)>>>(N, nwords, color, mask64, ovf, hubs, nhub,
                                                     count, npred, hpos, ctl);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_jp_init(uint32_t N, const uint32_t *dV, const uint32_t *dE, const void *key,
                            int key_bits, uint32_t *count, uint32_t *npred, uint32_t *esplit,
                            gcxk_item *q0i, const uint32_t *hubs, uint32_t nhub,
                            const uint32_t *hchunk, uint32_t nhchunk, uint32_t *hpos,
                            gcxk_ctl *ctl, void *stream)
{
    if (N == 0)
        return GCX_OK;
    uint2 *q0 = reinterpret_cast<uint2 *>(q0i);
    hipStream_t s = as_stream(stream);
    const uint32_t grid = div_up((uint64_t)N * kGroup, kBlock);
    if (key_bits == 32) {
        jp_init_groups<uint32_t><<<grid, kBlock, 0, s>>>(N, dV, dE, (const uint32_t *)key, count,
                                                         npred, esplit, q0, ctl);
        if (nhub)
            jp_init_hub_chunks<uint32_t><<<nhchunk, kBlock, 0, s>>>(
                dV, dE, (const uint32_t *)key, count, npred, esplit, hubs, hchunk, nhub, hpos);
    } else {
        jp_init_groups<uint64_t><<<grid, kBlock, 0, s>>>(N, dV, dE, (const uint64_t *)key, count,
                                                         npred, esplit, q0, ctl);
        if (nhub)
            jp_init_hub_chunks<uint64_t><<<nhchunk, kBlock, 0, s>>>(
                dV, dE, (const uint64_t *)key, count, npred, esplit, hubs, hchunk, nhub, hpos);
    }
    if (nhub)
        jp_hub_roots<<<div_up(nhub, kBlock), kBlock, 0, s>>>(dV, npred, hubs, nhub, q0, ctl);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_jp_round(const uint32_t *dV, const uint32_t *npred, const uint32_t *esplit,
                             uint32_t *count, int *color, uint64_t *mask64, uint32_t *ovf,
                             gcxk_item *const q[2], gcxk_ctl *ctl, uint32_t r, uint32_t grid,
                             void *stream)
{
    jp_round_kernel<<<grid, kBlock, 0, as_stream(stream)>>>(
        dV, npred, esplit, count, color, mask64, ovf, reinterpret_cast<const uint2 *>(q[r & 1]),
        reinterpret_cast<uint2 *>(q[(r + 1) & 1]), ctl, r);
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
