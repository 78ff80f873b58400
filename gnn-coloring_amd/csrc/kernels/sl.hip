// sl.hip -- smallest-degree-last (SL) and smallest-log-degree-last (SLL)
// peeling, reference smallest_degree_last.c:155-246 and 359-443.
//
// Round r with threshold T_r (SL: T = delta, SLL: T = 2^delta):
//   scan: every alive u with round-start deg[u] <= T is removed (key = delta);
//         the others stay alive, and their round-start degrees feed the min;
//   dec:  every unremoved neighbour of a removed vertex loses one degree per
//         removed neighbour; the new degrees feed the min;
//   next: SL delta <- max(delta + 1, min), SLL delta <- delta + 1.
// Initial delta: SL = min degree; SLL = smallest d with 2^d >= min degree.
// The min is kept inverted (UINT32_MAX - m) so that a zeroed slot is the
// identity of atomicMax. Slots rotate by round % 3 so that each kernel resets
// only a slot no kernel of the current round touches.
#include "common.hip.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kAlive = UINT32_MAX;
constexpr int kDecSteps = 4;
constexpr int kScanPer = 8;

__device__ __forceinline__ void wave_min_inv(uint32_t m, uint32_t *slot)
{
    // m = UINT32_MAX - value (0 = no contribution)
    m = group_max32<kWave>(m);
    if (lane_id() == 0 && m)
        atomicMax(slot, m);
}

__global__ __launch_bounds__(kBlock) void sl_init_kernel(uint32_t N, const uint32_t *__restrict__ V,
                                                         uint32_t *__restrict__ deg,
                                                         uint32_t *__restrict__ key,
                                                         uint32_t *__restrict__ alive0,
                                                         gcxk_ctl *ctl)
{
    const uint32_t u = blockIdx.x * kBlock + threadIdx.x;
    uint32_t inv = 0;
    if (u < N) {
        const uint32_t d = V[u + 1] - V[u];
        deg[u] = d;
        key[u] = kAlive;
        alive0[u] = u;
        inv = UINT32_MAX - d;
    }
    if (u == 0) {
        ctl->sl_alive[0] = N;
    }
    // Min degree goes to slot 2 (the slot of "round -1").
    wave_min_inv(inv, &ctl->sl_mininv[2]);
}

__device__ __forceinline__ uint32_t round_delta(const gcxk_ctl *ctl, int log_variant, uint32_t r)
{
    const uint32_t pm = (r + 2) % 3; // slot of round r-1
    const uint32_t mn = UINT32_MAX - ctl->sl_mininv[pm];
    if (r == 0) {
        if (!log_variant)
            return mn;
        uint32_t d = 0;
        while (d < 32 && (1ull << d) < mn)
            d++;
        return d;
    }
    const uint32_t prev = ctl->sl_delta[pm];
    if (log_variant)
        return prev + 1;
    return prev + 1 > mn ? prev + 1 : mn;
}

__global__ __launch_bounds__(kBlock) void sl_scan_kernel(uint32_t *__restrict__ deg,
                                                         uint32_t *__restrict__ key,
                                                         const uint32_t *__restrict__ alive_in,
                                                         uint32_t *__restrict__ alive_out,
                                                         uint32_t *__restrict__ gone, gcxk_ctl *ctl,
                                                         int log_variant, uint32_t r)
{
    const uint32_t s = r % 3, s1 = (r + 1) % 3, s2 = (r + 2) % 3;
    const uint32_t n = ctl->sl_alive[s];
    const uint32_t delta = round_delta(ctl, log_variant, r);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->sl_delta[s] = delta;
        // Slot s1 is next round's: its gone list and min start empty.
        ctl->sl_gone[s1] = 0;
        ctl->sl_mininv[s1] = 0;
        // Slot s2 held round r-1's input; it becomes round r+1's input.
        ctl->sl_alive[s2] = 0;
        if (n)
            ctl->sl_rounds = r + 1;
        else
            ctl->sl_done = 1;
    }
    if (n == 0)
        return;
    const uint64_t thr = log_variant ? (delta >= 32 ? (1ull << 32) : (1ull << delta)) : delta;
    // Tiles of kBlock * kScanPer entries: one reservation per list per tile
    // (all blocks append to the same two counters, whose atomics serialise).
    __shared__ uint32_t s_cnt[2][kWavesPerBlock];
    __shared__ uint32_t s_base[2];
    const uint32_t tile = kBlock * kScanPer;
    const uint32_t ntiles = (n + tile - 1) / tile;
    const int lane = lane_id(), w = threadIdx.x / kWave;
    uint32_t inv = 0;
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) { // block-uniform trip count
        uint32_t u[kScanPer];
        bool rm[kScanPer], keep[kScanPer];
        uint32_t cr = 0, ck = 0;
#pragma unroll
        for (int q = 0; q < kScanPer; q++) {
            const uint32_t i = t * tile + (uint32_t)q * kBlock + threadIdx.x;
            rm[q] = keep[q] = false;
            u[q] = i < n ? alive_in[i] : 0u;
            if (i < n) {
                const uint32_t d = deg[u[q]];
                if ((uint64_t)d <= thr) {
                    rm[q] = true;
                    key[u[q]] = delta;
                } else {
                    keep[q] = true;
                    const uint32_t m = UINT32_MAX - d;
                    inv = m > inv ? m : inv;
                }
            }
            cr += rm[q] ? 1u : 0u;
            ck += keep[q] ? 1u : 0u;
        }
        uint32_t ir = cr, ik = ck; // inclusive scans over the wave
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t a = __shfl_up(ir, off), b = __shfl_up(ik, off);
            if (lane >= off) {
                ir += a;
                ik += b;
            }
        }
        if (lane == kWave - 1) {
            s_cnt[0][w] = ir;
            s_cnt[1][w] = ik;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tr = 0, tk = 0;
            for (int q = 0; q < kWavesPerBlock; q++) {
                const uint32_t a = s_cnt[0][q], b = s_cnt[1][q];
                s_cnt[0][q] = tr;
                s_cnt[1][q] = tk;
                tr += a;
                tk += b;
            }
            s_base[0] = tr ? atomicAdd(&ctl->sl_gone[s], tr) : 0u;
            s_base[1] = tk ? atomicAdd(&ctl->sl_alive[s1], tk) : 0u;
        }
        __syncthreads();
        uint32_t pr = s_base[0] + s_cnt[0][w] + ir - cr;
        uint32_t pk = s_base[1] + s_cnt[1][w] + ik - ck;
#pragma unroll
        for (int q = 0; q < kScanPer; q++) {
            if (rm[q])
                gone[pr++] = u[q];
            else if (keep[q])
                alive_out[pk++] = u[q];
        }
        __syncthreads(); // s_cnt and s_base are rewritten by the next tile
    }
    wave_min_inv(inv, &ctl->sl_mininv[s]);
}

// In rounds that remove at least kAggMinGone vertices, decrements to the
// kHubSlots highest-degree vertices (hub_bit set, slot = degree-order rank)
// are summed in LDS and applied once per block: in R-MAT graphs most removed
// vertices hang off the same few hubs, whose counters would otherwise take
// thousands of serialised atomics per round.
constexpr uint32_t kHubSlots = 1024;
constexpr uint32_t kAggMinGone = 1u << 16;

__device__ __forceinline__ bool is_hub(const uint32_t *__restrict__ hub_bit, uint32_t v)
{
    return (hub_bit[v >> 5] >> (v & 31u)) & 1u;
}

// Decrement the alive neighbours among entries beg + t0, beg + t0 + stride,
// ... (< end) of a removed vertex's row, STEPS entries per thread in flight
// (loads, then atomics, then uses); folds the new degrees into inv.
template <int STEPS>
__device__ __forceinline__ uint32_t dec_row(uint32_t beg, uint32_t end, uint32_t t0, uint32_t stride,
                                            const uint32_t *__restrict__ E,
                                            uint32_t *__restrict__ deg,
                                            const uint32_t *__restrict__ key,
                                            const uint32_t *__restrict__ hub_bit,
                                            const uint32_t *__restrict__ rank, uint32_t *s_hub,
                                            uint32_t inv)
{
    for (uint32_t i0 = beg + t0; i0 < end; i0 += stride * STEPS) {
        uint32_t v[STEPS], d[STEPS];
        bool live[STEPS];
#pragma unroll
        for (int t = 0; t < STEPS; t++) {
            const uint32_t i = i0 + (uint32_t)t * stride;
            live[t] = i < end;
            v[t] = live[t] ? E[i] : 0u;
        }
#pragma unroll
        for (int t = 0; t < STEPS; t++) {
            live[t] = live[t] && key[v[t]] == kAlive;
            if (live[t] && hub_bit && is_hub(hub_bit, v[t])) {
                atomicAdd(&s_hub[rank[v[t]]], 1u);
                live[t] = false;
            }
        }
#pragma unroll
        for (int t = 0; t < STEPS; t++)
            d[t] = live[t] ? atomicSub(deg + v[t], 1u) : 0u;
#pragma unroll
        for (int t = 0; t < STEPS; t++) {
            const uint32_t m = live[t] ? UINT32_MAX - (d[t] - 1u) : 0u;
            inv = m > inv ? m : inv;
        }
    }
    return inv;
}

// In rounds that remove many vertices, rows of at most kLaneRow entries are
// decremented by one thread, rows of at most kBigRow by one wave, longer
// rows by the whole block (one wave per removed vertex made R-MAT 22's first
// rounds, 1.8 M isolated and 0.55 M degree-1 vertices, take 0.6-0.7 ms each);
// in the others every removed vertex gets a wave (a thread per vertex left
// most of the grid idle in rounds that remove a few thousand long rows).
constexpr uint32_t kLaneRow = 16;
constexpr uint32_t kBigRow = 2048;

__global__ __launch_bounds__(kBlock) void sl_dec_kernel(const uint32_t *__restrict__ V,
                                                        const uint32_t *__restrict__ E,
                                                        uint32_t *__restrict__ deg,
                                                        const uint32_t *__restrict__ key,
                                                        const uint32_t *__restrict__ gone,
                                                        const uint32_t *__restrict__ hub_bit,
                                                        const uint32_t *__restrict__ rank,
                                                        const uint32_t *__restrict__ perm,
                                                        uint32_t nslot, gcxk_ctl *ctl, uint32_t r)
{
    __shared__ uint32_t s_beg[kBlock], s_end[kBlock];
    __shared__ uint32_t s_nrow;
    __shared__ uint32_t s_hub[kHubSlots];
    const uint32_t s = r % 3;
    const uint32_t n = ctl->sl_gone[s];
    if (n == 0)
        return;
    // only rounds that remove many vertices pile up on the hubs' counters;
    // in the others the LDS clear and flush cost more than they save
    if (n < kAggMinGone)
        nslot = 0;
    if (nslot == 0)
        hub_bit = nullptr;
    for (uint32_t i = threadIdx.x; i < nslot; i += kBlock)
        s_hub[i] = 0u;
    const int w = threadIdx.x / kWave;
    uint32_t inv = 0;
    if (n >= gridDim.x * (kBlock / 4u)) {
        // many removed vertices (early rounds: mostly degree 0 or 1): a
        // thread per removed vertex; rows of at most kLaneRow entries are
        // walked by their own thread, longer ones are listed in LDS and taken
        // by a wave each, rows past kBigRow by the whole block
        for (uint32_t g = blockIdx.x * kBlock; g < n; g += gridDim.x * kBlock) {
            if (threadIdx.x == 0)
                s_nrow = 0u;
            __syncthreads();
            const uint32_t k = g + threadIdx.x;
            uint32_t beg = 0, end = 0;
            if (k < n) {
                const uint32_t u = gone[k];
                beg = V[u];
                end = V[u + 1];
            }
            if (end - beg <= kLaneRow) {
                inv = dec_row<kDecSteps>(beg, end, 0u, 1u, E, deg, key, hub_bit, rank, s_hub, inv);
            } else {
                const uint32_t slot = atomicAdd(&s_nrow, 1u);
                s_beg[slot] = beg;
                s_end[slot] = end;
            }
            __syncthreads();
            const uint32_t m = s_nrow;
            for (uint32_t j = (uint32_t)w; j < m; j += kWavesPerBlock)
                if (s_end[j] - s_beg[j] <= kBigRow)
                    inv = dec_row<kDecSteps>(s_beg[j], s_end[j], (uint32_t)lane_id(), kWave, E, deg,
                                             key, hub_bit, rank, s_hub, inv);
            for (uint32_t j = 0; j < m; j++)
                if (s_end[j] - s_beg[j] > kBigRow) // block-uniform
                    inv = dec_row<kDecSteps>(s_beg[j], s_end[j], threadIdx.x, kBlock, E, deg, key,
                                             hub_bit, rank, s_hub, inv);
            __syncthreads(); // s_nrow, s_beg and s_end are rewritten by the next group
        }
    } else {
        // few removed vertices (late rounds: long rows): a wave per removed
        // vertex, so that every wave of the grid has a row; rows past kBigRow
        // by the whole block after the group's short rows
        __syncthreads(); // s_hub cleared
        for (uint32_t g = blockIdx.x * kWavesPerBlock; g < n; g += gridDim.x * kWavesPerBlock) {
            const uint32_t k = g + (uint32_t)w;
            uint32_t beg = 0, end = 0;
            if (k < n) {
                const uint32_t u = gone[k];
                beg = V[u];
                end = V[u + 1];
            }
            const bool big = end - beg > kBigRow;
            if (!big)
                inv = dec_row<kDecSteps>(beg, end, (uint32_t)lane_id(), kWave, E, deg, key, hub_bit,
                                         rank, s_hub, inv);
            if (lane_id() == 0) {
                s_beg[w] = big ? beg : 0u;
                s_end[w] = big ? end : 0u;
            }
            __syncthreads();
            for (int q = 0; q < kWavesPerBlock; q++)
                if (s_end[q]) // block-uniform
                    inv = dec_row<kDecSteps>(s_beg[q], s_end[q], threadIdx.x, kBlock, E, deg, key,
                                             hub_bit, rank, s_hub, inv);
            __syncthreads(); // s_beg / s_end are rewritten by the next group
        }
    }
    for (uint32_t i = threadIdx.x; i < nslot; i += kBlock) {
        const uint32_t c = s_hub[i];
        if (c) {
            const uint32_t d = atomicSub(deg + perm[i], c) - c;
            const uint32_t m = UINT32_MAX - d;
            inv = m > inv ? m : inv;
        }
    }
    wave_min_inv(inv, &ctl->sl_mininv[s]);
}

// hub_bit: one bit per vertex, set for perm[0 .. nslot).
__global__ __launch_bounds__(kBlock) void sl_hub_bits(uint32_t words, const uint32_t *__restrict__ perm,
                                                      uint32_t nslot, uint32_t *__restrict__ hub_bit,
                                                      int phase)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (phase == 0) {
        if (i < words)
            hub_bit[i] = 0u;
    } else if (i < nslot) {
        const uint32_t v = perm[i];
        atomicOr(&hub_bit[v >> 5], 1u << (v & 31u));
    }
}

} // namespace

extern "C" int gcxk_sl_init(uint32_t N, const uint32_t *dV, uint32_t *deg, uint32_t *key,
                            uint32_t *alive0, gcxk_ctl *ctl, void *stream)
{
    sl_init_kernel<<<div_up(N ? N : 1, kBlock), kBlock, 0, as_stream(stream)>>>(N, dV, deg, key,
                                                                                 alive0, ctl);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_sl_scan(uint32_t *deg, uint32_t *key, const uint32_t *alive_in,
                            uint32_t *alive_out, uint32_t *gone, gcxk_ctl *ctl, int log_variant,
                            uint32_t r, uint32_t grid, void *stream)
{
    sl_scan_kernel<<<grid, kBlock, 0, as_stream(stream)>>>(deg, key, alive_in, alive_out, gone, ctl,
                                                           log_variant, r);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_sl_dec(const uint32_t *dV, const uint32_t *dE, uint32_t *deg,
                           const uint32_t *key, const uint32_t *gone, const uint32_t *hub_bit,
                           const uint32_t *rank, const uint32_t *perm, uint32_t nslot,
                           gcxk_ctl *ctl, uint32_t r, uint32_t grid, void *stream)
{
    if (nslot > kHubSlots)
        nslot = kHubSlots;
    if (!hub_bit)
        nslot = 0;
    sl_dec_kernel<<<grid, kBlock, 0, as_stream(stream)>>>(dV, dE, deg, key, gone, hub_bit, rank,
                                                          perm, nslot, ctl, r);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" uint32_t gcxk_sl_hub_slots(uint32_t N) { return N < kHubSlots ? N : kHubSlots; }

extern "C" int gcxk_sl_hub_bits(uint32_t N, const uint32_t *perm, uint32_t *hub_bit, void *stream)
{
    const uint32_t words = N / 32u + 1u, nslot = gcxk_sl_hub_slots(N);
    sl_hub_bits<<<div_up(words, kBlock), kBlock, 0, as_stream(stream)>>>(words, perm, nslot,
                                                                         hub_bit, 0);
    GCXK_LAUNCH_CHECK();
    if (nslot) {
        sl_hub_bits<<<div_up(nslot, kBlock), kBlock, 0, as_stream(stream)>>>(words, perm, nslot,
                                                                             hub_bit, 1);
        GCXK_LAUNCH_CHECK();
    }
    return GCX_OK;
}
