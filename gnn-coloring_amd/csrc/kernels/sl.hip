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

__device__ __forceinline__ void wave_append(bool want, uint32_t v, uint32_t *q, uint32_t *n)
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
    const uint32_t stride = gridDim.x * kBlock;
    // Uniform trip count per wave so that the ballots see every lane.
    const uint32_t n_pad = (n + kWave - 1) / kWave * kWave;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i - lane_id() < n_pad; i += stride) {
        bool rm = false, keep = false;
        uint32_t u = 0, inv = 0;
        if (i < n) {
            u = alive_in[i];
            const uint32_t d = deg[u];
            if ((uint64_t)d <= thr) {
                rm = true;
                key[u] = delta;
            } else {
                keep = true;
                inv = UINT32_MAX - d;
            }
        }
        wave_append(rm, u, gone, &ctl->sl_gone[s]);
        wave_append(keep, u, alive_out, &ctl->sl_alive[s1]);
        wave_min_inv(inv, &ctl->sl_mininv[s]);
    }
}

__global__ __launch_bounds__(kBlock) void sl_dec_kernel(const uint32_t *__restrict__ V,
                                                        const uint32_t *__restrict__ E,
                                                        uint32_t *__restrict__ deg,
                                                        const uint32_t *__restrict__ key,
                                                        const uint32_t *__restrict__ gone,
                                                        gcxk_ctl *ctl, uint32_t r)
{
    const uint32_t s = r % 3;
    const uint32_t n = ctl->sl_gone[s];
    const uint32_t w0 = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    const uint32_t nw = gridDim.x * kWavesPerBlock;
    uint32_t inv = 0;
    for (uint32_t k = w0; k < n; k += nw) {
        const uint32_t u = gone[k];
        const uint32_t beg = V[u], end = V[u + 1];
        for (uint32_t i = beg + lane_id(); i < end; i += kWave) {
            const uint32_t v = E[i];
            if (key[v] == kAlive) {
                const uint32_t d = atomicSub(deg + v, 1u) - 1u;
                const uint32_t m = UINT32_MAX - d;
                inv = m > inv ? m : inv;
            }
        }
    }
    wave_min_inv(inv, &ctl->sl_mininv[s]);
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
                           const uint32_t *key, const uint32_t *gone, gcxk_ctl *ctl, uint32_t r,
                           uint32_t grid, void *stream)
{
    sl_dec_kernel<<<grid, kBlock, 0, as_stream(stream)>>>(dV, dE, deg, key, gone, ctl, r);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}
