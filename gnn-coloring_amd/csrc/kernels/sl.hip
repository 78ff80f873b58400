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
//
// Rank space: the peel keeps its per-vertex state (degrees, alive flags,
// alive lists) at degree-order ranks (perm / rank: degree descending, ties
// by id) and reads neighbours from Er = rank[E], so the highest-degree
// vertices are the first ranks. They take most decrements: on R-MAT 22,
// 78% of SL's 61.5 M decrements go to the 32768 first ranks, and in the
// heavy late rounds all of them (round 221: 4.7 M decrements onto 16.7 K
// vertices). Every workgroup sums its decrements to the first kSumSlots
// ranks in LDS and applies each sum with one atomic, adjacent ranks from
// adjacent lanes; only decrements to later ranks are device atomics. On
// MI355X, device-scope atomics scattered over a few thousand hot addresses
// run at about 4.5 G/s; summed in LDS and flushed in rank order the same
// decrements run at about 170 G/s.
#include "common.hip.h"

namespace {

constexpr int kBlock = 256;   // init and scan
constexpr int kDBlock = 1024; // decrements: one workgroup per CU or per two
constexpr int kDWaves = kDBlock / kWave;
constexpr uint32_t kAlive = UINT32_MAX;
constexpr int kDecSteps = 4;
constexpr int kScanPer = 8;
// LDS decrement sums: two ranks per 32-bit word (16-bit halves), 64 KB (a
// 128 KB table of 65536 ranks was 0.8 ms slower on R-MAT 22 SL at one
// workgroup per CU). A workgroup's sum for one rank is at most the number
// of removed rows it handles (one entry per row and neighbour in a simple
// graph); the sums are used only while that bound stays below 2^16, so the
// low half never carries into the high one.
constexpr uint32_t kSumWords = 16384;
constexpr uint32_t kSumSlots = 2 * kSumWords;
constexpr uint32_t kMinSumSlots = 1024;
constexpr uint64_t kMaxRowsPerGroup = 65535;
constexpr uint32_t kLaneRow = 16;   // rows up to this length: a thread each
constexpr uint32_t kChunk = kWave;  // longer rows: chunks of a lane per entry
constexpr int kChunkSteps = 4;
constexpr uint32_t kOwnChunks = 4;  // chunks of a row its own thread lists
// Rows of more than kSuperChunks chunks (removed hubs, up to hundreds of
// thousands of entries) are listed as super chunks of kSuperChunks * kChunk
// entries, which the decrement pass cuts into chunks: writing a hub's
// thousands of chunk descriptors fell to one wave of the scan and made the
// late SL rounds' scans 5-10x longer than their median.
constexpr uint32_t kSuperChunks = 64;
constexpr uint32_t kSuperLen = kSuperChunks * kChunk;

// Folds m = UINT32_MAX - value (0 = no contribution) into *slot with
// atomicMax, one atomic per workgroup of T threads (one per wave put
// thousands of atomics per round on one word). Every thread of the
// workgroup calls it.
template <int T>
__device__ __forceinline__ void block_min_inv(uint32_t m, uint32_t *slot)
{
    __shared__ uint32_t s_m[T / kWave];
    m = group_max32<kWave>(m);
    if (lane_id() == 0)
        s_m[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t b = 0;
        for (int w = 0; w < T / kWave; w++)
            b = s_m[w] > b ? s_m[w] : b;
        if (b)
            atomicMax(slot, b);
    }
}

__device__ __forceinline__ void fold_min(uint32_t d, uint32_t &inv)
{
    const uint32_t m = UINT32_MAX - d;
    inv = m > inv ? m : inv;
}

// rrow[a] = (first entry, degree) of rank a's row, read by the scan in one
// load instead of perm -> V.
__global__ __launch_bounds__(kBlock) void sl_init_kernel(uint32_t N, const uint32_t *__restrict__ V,
                                                         const uint32_t *__restrict__ perm,
                                                         uint32_t *__restrict__ degr,
                                                         uint32_t *__restrict__ keyr,
                                                         uint32_t *__restrict__ alive0,
                                                         uint2 *__restrict__ rrow, gcxk_ctl *ctl)
{
    const uint32_t a = blockIdx.x * kBlock + threadIdx.x;
    if (a < N) {
        const uint32_t u = perm[a];
        const uint32_t b = V[u], d = V[u + 1] - b;
        degr[a] = d;
        rrow[a] = make_uint2(b, d);
        keyr[a] = kAlive;
        alive0[a] = a;
        // the last rank has the smallest degree: it goes to slot 2 (the
        // slot of "round -1")
        if (a == N - 1u)
            ctl->sl_mininv[2] = UINT32_MAX - d;
    }
    if (a == 0) {
        ctl->sl_alive[0] = N;
    }
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

// alive_in / alive_out hold ranks. A removed vertex's threshold goes to key
// (by id); its row (first entry, length) goes to gone when it has 1 ..
// kLaneRow entries, else to chunks as ranges of at most kChunk entries, or,
// past kSuperChunks chunks, to supers as ranges of at most kSuperLen, so
// that the decrement pass has only equal work items and no workgroup-wide
// steps.
__global__ __launch_bounds__(kBlock) void sl_scan_kernel(const uint2 *__restrict__ rrow,
                                                         const uint32_t *__restrict__ degr,
                                                         uint32_t *__restrict__ keyr,
                                                         uint32_t *__restrict__ key,
                                                         const uint32_t *__restrict__ perm,
                                                         const uint32_t *__restrict__ alive_in,
                                                         uint32_t *__restrict__ alive_out,
                                                         uint2 *__restrict__ gone,
                                                         uint2 *__restrict__ chunks,
                                                         uint2 *__restrict__ supers, gcxk_ctl *ctl,
                                                         int log_variant, uint32_t r)
{
    const uint32_t s = r % 3, s1 = (r + 1) % 3, s2 = (r + 2) % 3;
    const uint32_t n = ctl->sl_alive[s];
    const uint32_t delta = round_delta(ctl, log_variant, r);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->sl_delta[s] = delta;
        // Slot s1 is next round's: its work lists and min start empty.
        ctl->sl_gone[s1] = 0;
        ctl->sl_chunks[s1] = 0;
        ctl->sl_supers[s1] = 0;
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
    // (all blocks append to the same counters, whose atomics serialise).
    constexpr int kW = kBlock / kWave;
    __shared__ uint32_t s_cnt[4][kW];
    __shared__ uint32_t s_base[4];
    const uint32_t tile = kBlock * kScanPer;
    const uint32_t ntiles = (n + tile - 1) / tile;
    const int lane = lane_id(), w = threadIdx.x / kWave;
    uint32_t inv = 0;
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) { // block-uniform trip count
        uint32_t a[kScanPer], u[kScanPer], beg[kScanPer], len[kScanPer];
        bool rm[kScanPer], keep[kScanPer];
#pragma unroll
        for (int q = 0; q < kScanPer; q++) {
            const uint32_t i = t * tile + (uint32_t)q * kBlock + threadIdx.x;
            a[q] = i < n ? alive_in[i] : 0u;
        }
#pragma unroll
        for (int q = 0; q < kScanPer; q++) {
            const uint32_t i = t * tile + (uint32_t)q * kBlock + threadIdx.x;
            rm[q] = keep[q] = false;
            if (i < n) {
                const uint32_t d = degr[a[q]];
                if ((uint64_t)d <= thr) {
                    rm[q] = true;
                    keyr[a[q]] = delta;
                } else {
                    keep[q] = true;
                    fold_min(d, inv);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kScanPer; q++) {
            const uint2 rw = rm[q] ? rrow[a[q]] : make_uint2(0u, 0u);
            u[q] = rm[q] ? perm[a[q]] : 0u; // for key only
            beg[q] = rw.x;
            len[q] = rw.y;
        }
        uint32_t cr = 0, ck = 0, cc = 0, cs = 0;
#pragma unroll
        for (int q = 0; q < kScanPer; q++) {
            const uint32_t nch = len[q] > kLaneRow ? (len[q] + kChunk - 1u) / kChunk : 0u;
            cr += (len[q] >= 1u && len[q] <= kLaneRow) ? 1u : 0u;
            ck += keep[q] ? 1u : 0u;
            cc += nch <= kSuperChunks ? nch : 0u;
            cs += nch > kSuperChunks ? (len[q] + kSuperLen - 1u) / kSuperLen : 0u;
        }
        uint32_t ir = cr, ik = ck, ic = cc, is = cs; // inclusive scans over the wave
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t x = __shfl_up(ir, off), y = __shfl_up(ik, off), z = __shfl_up(ic, off),
                           v = __shfl_up(is, off);
            if (lane >= off) {
                ir += x;
                ik += y;
                ic += z;
                is += v;
            }
        }
        if (lane == kWave - 1) {
            s_cnt[0][w] = ir;
            s_cnt[1][w] = ik;
            s_cnt[2][w] = ic;
            s_cnt[3][w] = is;
        }
        __syncthreads();
        if (threadIdx.x < 4) {
            uint32_t tot = 0;
            for (int q = 0; q < kW; q++) {
                const uint32_t x = s_cnt[threadIdx.x][q];
                s_cnt[threadIdx.x][q] = tot;
                tot += x;
            }
            uint32_t *ctr = threadIdx.x == 0   ? &ctl->sl_gone[s]
                            : threadIdx.x == 1 ? &ctl->sl_alive[s1]
                            : threadIdx.x == 2 ? &ctl->sl_chunks[s]
                                               : &ctl->sl_supers[s];
            s_base[threadIdx.x] = tot ? atomicAdd(ctr, tot) : 0u;
        }
        __syncthreads();
        uint32_t pr = s_base[0] + s_cnt[0][w] + ir - cr;
        uint32_t pk = s_base[1] + s_cnt[1][w] + ik - ck;
        uint32_t pc = s_base[2] + s_cnt[2][w] + ic - cc;
        uint32_t ps = s_base[3] + s_cnt[3][w] + is - cs;
#pragma unroll
        for (int q = 0; q < kScanPer; q++) {
            const uint32_t nch = len[q] > kLaneRow ? (len[q] + kChunk - 1u) / kChunk : 0u;
            if (rm[q]) {
                key[u[q]] = delta;
                if (nch && nch <= kOwnChunks) {
                    const uint32_t end = beg[q] + len[q];
                    for (uint32_t b = beg[q]; b < end; b += kChunk)
                        chunks[pc++] = make_uint2(b, end - b > kChunk ? b + kChunk : end);
                } else if (nch == 0 && len[q]) {
                    gone[pr++] = make_uint2(beg[q], len[q]);
                }
            } else if (keep[q]) {
                alive_out[pk++] = a[q];
            }
            // rows of more than kOwnChunks chunks are written by the whole
            // wave, one row at a time: as chunks up to kSuperChunks, longer
            // rows as super chunks
            uint64_t big = __ballot(rm[q] && nch > kOwnChunks);
            const bool sup = nch > kSuperChunks;
            const uint32_t nunit = sup ? (len[q] + kSuperLen - 1u) / kSuperLen : nch;
            const uint32_t p_row = sup ? ps : pc;
            if (rm[q] && nch > kOwnChunks) {
                if (sup)
                    ps += nunit;
                else
                    pc += nunit;
            }
            while (big) {
                const int src = __builtin_ctzll(big);
                big &= big - 1;
                const uint32_t b0 = __shfl(beg[q], src), e0 = b0 + __shfl(len[q], src);
                const uint32_t p0 = __shfl(p_row, src), n0 = __shfl(nunit, src);
                const bool s0 = __shfl((int)sup, src) != 0;
                uint2 *dst = s0 ? supers : chunks;
                const uint32_t unit = s0 ? kSuperLen : kChunk;
                for (uint32_t j = (uint32_t)lane; j < n0; j += kWave) {
                    const uint32_t b = b0 + j * unit;
                    dst[p0 + j] = make_uint2(b, e0 - b > unit ? b + unit : e0);
                }
            }
        }
        __syncthreads(); // s_cnt and s_base are rewritten by the next tile
    }
    block_min_inv<kBlock>(inv, &ctl->sl_mininv[s]);
}

// Decrement the alive neighbours among entries beg + t0, beg + t0 + stride,
// ... (< end) of a removed vertex's row, STEPS entries per thread in flight
// (column loads, then flag loads, then the decrements): ranks below nsum
// into the workgroup's LDS sums, the others with a device atomic whose new
// degree is folded into inv.
template <int STEPS>
__device__ __forceinline__ uint32_t dec_row(uint32_t beg, uint32_t end, uint32_t t0, uint32_t stride,
                                            const uint32_t *__restrict__ Er,
                                            uint32_t *__restrict__ degr,
                                            const uint32_t *__restrict__ keyr, uint32_t nsum,
                                            uint32_t *s_sum, uint32_t inv)
{
    for (uint32_t i0 = beg + t0; i0 < end; i0 += stride * STEPS) {
        uint32_t t[STEPS], kv[STEPS];
#pragma unroll
        for (int q = 0; q < STEPS; q++) {
            const uint32_t i = i0 + (uint32_t)q * stride;
            t[q] = i < end ? Er[i] : 0u;
        }
#pragma unroll
        for (int q = 0; q < STEPS; q++) {
            const uint32_t i = i0 + (uint32_t)q * stride;
            kv[q] = i < end ? keyr[t[q]] : 0u;
        }
#pragma unroll
        for (int q = 0; q < STEPS; q++) {
            if (kv[q] != kAlive)
                continue;
            if (t[q] < nsum)
                atomicAdd(&s_sum[t[q] >> 1], 1u << ((t[q] & 1u) * 16u));
            else
                fold_min(atomicSub(degr + t[q], 1u) - 1u, inv);
        }
    }
    return inv;
}

// Short rows (at most kLaneRow entries): a thread each; long rows: a wave
// per chunk of kChunk entries (a lane per entry), kChunkSteps chunks in
// flight. Super chunk x stands for chunks nc + x * kSuperChunks .. + 63 of
// the chunk sequence (empty past its end). No workgroup-wide steps until the
// LDS sums are flushed.
__global__ __launch_bounds__(kDBlock) void sl_dec_kernel(const uint32_t *__restrict__ V,
                                                         const uint32_t *__restrict__ Er,
                                                         uint32_t *__restrict__ degr,
                                                         const uint32_t *__restrict__ keyr,
                                                         const uint2 *__restrict__ gone,
                                                         const uint2 *__restrict__ chunks,
                                                         const uint2 *__restrict__ supers,
                                                         uint32_t nsum, gcxk_ctl *ctl, uint32_t r)
{
    __shared__ uint32_t s_sum[kSumWords];
    const uint32_t s = r % 3;
    const uint32_t n = ctl->sl_gone[s], nc = ctl->sl_chunks[s];
    const uint32_t nu = nc + ctl->sl_supers[s] * kSuperChunks; // chunks in all
    // nothing left alive (the last round): no degree to lower
    if ((n == 0 && nu == 0) || ctl->sl_alive[(r + 1) % 3] == 0)
        return;
    // LDS sums over the first nsum ranks, nsum sized to the round: the
    // power of two from kMinSumSlots that covers the entries a workgroup
    // takes. Every workgroup clears and flushes its table once per round,
    // which light rounds do not repay. R-MAT 22 with fixed tables of 8192 /
    // 16384 / 32768 ranks: SL 12.6 / 13.0 / 13.8 ms, SLL (heavy rounds)
    // 3.85 / 3.60 / 3.53; sized per round (4x / 1x the entries per
    // workgroup, from 8192; 1x from 4096; 1x from 1024): SL 12.9 / 12.7 /
    // 12.5 / 12.3, SLL 3.53 - 3.54
    {
        const uint64_t per_wg = ((uint64_t)n * (kLaneRow / 2u) + (uint64_t)nu * kChunk) / gridDim.x;
        uint32_t want = kMinSumSlots;
        while (want < nsum && (uint64_t)want < per_wg)
            want <<= 1;
        nsum = want < nsum ? want : nsum;
    }
    const uint64_t rows_per = kDBlock, chunks_per = (uint64_t)kDWaves * kChunkSteps;
    const uint64_t rows = (n + gridDim.x * rows_per - 1) / (gridDim.x * rows_per) * rows_per;
    const uint64_t chs = (nu + gridDim.x * chunks_per - 1) / (gridDim.x * chunks_per) * chunks_per;
    if (rows + chs > kMaxRowsPerGroup)
        nsum = 0; // a 16-bit sum could overflow
    const uint32_t words = (nsum + 1u) / 2u;
    for (uint32_t i = threadIdx.x; i < words; i += kDBlock)
        s_sum[i] = 0u;
    __syncthreads();
    uint32_t inv = 0;
    for (uint32_t k = blockIdx.x * kDBlock + threadIdx.x; k < n; k += gridDim.x * kDBlock) {
        const uint2 row = gone[k];
        inv = dec_row<kDecSteps>(row.x, row.x + row.y, 0u, 1u, Er, degr, keyr, nsum, s_sum, inv);
    }
    const uint32_t nw = gridDim.x * kDWaves;
    const uint32_t lane = (uint32_t)lane_id();
    for (uint32_t c0 = blockIdx.x * kDWaves + threadIdx.x / kWave; c0 < nu; c0 += nw * kChunkSteps) {
        uint32_t t[kChunkSteps], kv[kChunkSteps];
        bool in[kChunkSteps];
#pragma unroll
        for (int q = 0; q < kChunkSteps; q++) {
            const uint32_t c = c0 + (uint32_t)q * nw; // wave-uniform
            uint2 ch = make_uint2(0u, 0u);
            if (c < nc) {
                ch = chunks[c];
            } else if (c < nu) {
                const uint32_t x = c - nc;
                const uint2 sp = supers[x / kSuperChunks];
                const uint32_t b = sp.x + (x % kSuperChunks) * kChunk;
                if (b < sp.y) // the last super chunk of a row can end early
                    ch = make_uint2(b, sp.y - b > kChunk ? b + kChunk : sp.y);
            }
            in[q] = ch.x + lane < ch.y;
            t[q] = in[q] ? Er[ch.x + lane] : 0u;
        }
#pragma unroll
        for (int q = 0; q < kChunkSteps; q++)
            kv[q] = in[q] ? keyr[t[q]] : 0u;
#pragma unroll
        for (int q = 0; q < kChunkSteps; q++) {
            if (kv[q] != kAlive)
                continue;
            if (t[q] < nsum)
                atomicAdd(&s_sum[t[q] >> 1], 1u << ((t[q] & 1u) * 16u));
            else
                fold_min(atomicSub(degr + t[q], 1u) - 1u, inv);
        }
    }
    __syncthreads();
    // one atomic per rank with a non-zero sum; a wave covers 128 consecutive
    // ranks
    for (uint32_t i = threadIdx.x; i < words; i += kDBlock) {
        const uint32_t c = s_sum[i];
        const uint32_t lo = c & 0xFFFFu, hi = c >> 16;
        if (lo)
            fold_min(atomicSub(degr + 2u * i, lo) - lo, inv);
        if (hi)
            fold_min(atomicSub(degr + 2u * i + 1u, hi) - hi, inv);
    }
    block_min_inv<kDBlock>(inv, &ctl->sl_mininv[s]);
}

uint32_t g_dec_grid = 0;

} // namespace

extern "C" int gcxk_sl_init(uint32_t N, const uint32_t *dV, const uint32_t *perm, uint32_t *degr,
                            uint32_t *keyr, uint32_t *alive0, void *rrow, gcxk_ctl *ctl,
                            void *stream)
{
    sl_init_kernel<<<div_up(N ? N : 1, kBlock), kBlock, 0, as_stream(stream)>>>(
        N, dV, perm, degr, keyr, alive0, static_cast<uint2 *>(rrow), ctl);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

namespace {

// A row of kLaneRow < len <= kSuperLen entries takes ceil(len / kChunk) <=
// len / 16 chunks (kChunk = 64 = 4 * kLaneRow): at most M / 16 over all
// rows. A longer row takes ceil(len / kSuperLen) <= len / 2048 super chunks
// (len > kSuperLen = 4096): at most M / 2048. The super list follows the
// chunk list.
uint64_t chunk_list_capacity(uint64_t M) { return M / kLaneRow + 1; }
uint64_t super_list_capacity(uint64_t M) { return M / (kSuperLen / 2u) + 1; }

} // namespace

extern "C" uint64_t gcxk_sl_chunk_capacity(uint64_t M)
{
    return chunk_list_capacity(M) + super_list_capacity(M);
}

extern "C" int gcxk_sl_scan(const void *rrow, const uint32_t *degr, uint32_t *keyr,
                            uint32_t *key, const uint32_t *perm, const uint32_t *alive_in,
                            uint32_t *alive_out, void *gone, void *chunks, uint64_t M,
                            gcxk_ctl *ctl, int log_variant, uint32_t r, uint32_t grid,
                            void *stream)
{
    uint2 *ch = static_cast<uint2 *>(chunks);
    sl_scan_kernel<<<grid, kBlock, 0, as_stream(stream)>>>(
        static_cast<const uint2 *>(rrow), degr, keyr, key, perm, alive_in, alive_out,
        static_cast<uint2 *>(gone), ch, ch + chunk_list_capacity(M), ctl, log_variant, r);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_sl_dec(uint32_t N, const uint32_t *dV, const uint32_t *Er, uint32_t *degr,
                           const uint32_t *keyr, const void *gone, const void *chunks, uint64_t M,
                           gcxk_ctl *ctl, int log_variant, uint32_t r, void *stream)
{
    if (!g_dec_grid) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
            return GCX_ERR_HIP;
        g_dec_grid = (uint32_t)prop.multiProcessorCount;
    }
    // Fewer workgroups sum more decrements per LDS table (fewer flushed
    // atomics) but keep fewer loads in flight. R-MAT 22 SL (thin degree
    // layers, 0.5-1.2 M decrements per middle round): 2 per CU 17.9 ms, 1 per
    // CU 15.4, 1 per 2 CUs 14.0, 1 per 4 CUs 14.4. SLL (layers of doubling
    // width, up to several M decrements a round): 3.9, 3.5, 3.8, 5.1 ms.
    // The 27-point stencil (160^3) is within 0.1 ms of its best at either.
    const uint32_t grid = log_variant ? g_dec_grid : (g_dec_grid + 1) / 2;
    const uint32_t nsum = N < kSumSlots ? N : kSumSlots;
    const uint2 *ch = static_cast<const uint2 *>(chunks);
    sl_dec_kernel<<<grid, kDBlock, 0, as_stream(stream)>>>(
        dV, Er, degr, keyr, static_cast<const uint2 *>(gone), ch, ch + chunk_list_capacity(M),
        nsum, ctl, r);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}
