// csr.hip -- device-side canonical CSR construction and graph checks.
//
// Reference path (fast_io.c:171-202, 450-455, 457-477): atomic degree count,
// serial prefix sum, atomic scatter of both directions, per-row qsort, serial
// removal of self loops and repeats. Here the same canonical CSR comes from
// one radix sort: every pair (u, v) becomes the keys (u << b | v) and
// (v << b | u) with b = bit length of N-1, self loops become the all-ones key
// that sorts last; after sorting, the first occurrence of each key is kept,
// its row is counted, and the row counts are scanned into V. Row order is
// ascending v because the key's low bits are v.
#include "common.hip.h"
#include "gen.h"

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void expand_pairs(uint64_t L, const uint32_t *__restrict__ I,
                                                       const uint32_t *__restrict__ J, int b,
                                                       uint64_t *__restrict__ keys)
{
    const uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= L)
        return;
    const uint64_t u = I[k], v = J[k];
    const uint64_t none = (b == 32) ? ~0ull : ((1ull << (2 * b)) - 1ull);
    if (u == v) {
        keys[2 * k] = none;
        keys[2 * k + 1] = none;
    } else {
        keys[2 * k] = (u << b) | v;
        keys[2 * k + 1] = (v << b) | u;
    }
}

// Keep the first of each run of equal keys; self loops carry the all-ones
// key, which sorts last and is dropped.
__global__ __launch_bounds__(kBlock) void mark_unique(uint64_t n, const uint64_t *__restrict__ keys,
                                                      int b, uint32_t *__restrict__ flag)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n)
        return;
    const uint64_t none = (b == 32) ? ~0ull : ((1ull << (2 * b)) - 1ull);
    const uint64_t k = keys[i];
    flag[i] = (k != none && (i == 0 || keys[i - 1] != k)) ? 1u : 0u;
}

// Columns and row offsets of the kept entries (keys sorted, pos = exclusive
// scan of flag, total = kept count). Kept entry i lands at pos[i] and opens
// every row after the previous distinct key's row up to its own; the last
// one also closes the rows after it. No atomics: each V[u] has one writer.
__global__ __launch_bounds__(kBlock) void scatter_rows(uint64_t n, const uint64_t *__restrict__ keys,
                                                       const uint32_t *__restrict__ flag,
                                                       const uint32_t *__restrict__ pos, int b,
                                                       uint32_t N, uint32_t total,
                                                       uint32_t *__restrict__ V,
                                                       uint32_t *__restrict__ E)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n || !flag[i])
        return;
    const uint64_t mask = (b == 32) ? 0xFFFFFFFFull : ((1ull << b) - 1ull);
    const uint64_t k = keys[i];
    const uint32_t r = (uint32_t)(k >> b), j = pos[i];
    E[j] = (uint32_t)(k & mask);
    // keys[i-1] < keys[i] (i is kept), so it is a real entry of an earlier
    // or the same row
    const uint32_t r0 = i == 0 ? 0u : (uint32_t)(keys[i - 1] >> b) + 1u;
    for (uint32_t u = r0; u <= r; u++)
        V[u] = j;
    if (j + 1u == total)
        for (uint32_t u = r + 1u; u <= N; u++)
            V[u] = total;
}


// Hub flags and hub chunk counts per vertex (entry N = 0 closes the scans),
// and the maximum degree (one atomic per workgroup).
__global__ __launch_bounds__(kBlock) void hub_flags(uint32_t N, const uint32_t *__restrict__ V,
                                                    uint32_t hub_deg, uint32_t *__restrict__ flag,
                                                    uint32_t *__restrict__ chunks,
                                                    uint32_t *__restrict__ maxdeg)
{
    __shared__ uint32_t wmax[kBlock / kWave];
    const uint32_t u = blockIdx.x * kBlock + threadIdx.x;
    uint32_t d = 0;
    if (u < N) {
        d = V[u + 1] - V[u];
        const bool hub = d >= hub_deg;
        flag[u] = hub ? 1u : 0u;
        chunks[u] = hub ? (d + GCXK_HUB_CHUNK_EDGES - 1u) / GCXK_HUB_CHUNK_EDGES : 0u;
    } else if (u == N) {
        flag[u] = 0u;
        chunks[u] = 0u;
    }
    d = group_max32<kWave>(d);
    if (lane_id() == 0)
        wmax[threadIdx.x / kWave] = d;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = 0;
        for (int i = 0; i < kBlock / kWave; i++)
            m = wmax[i] > m ? wmax[i] : m;
        if (m)
            atomicMax(maxdeg, m);
    }
}

// Hub list in ascending id order and its chunk prefix (hchunk[nhub] = total).
__global__ __launch_bounds__(kBlock) void hub_scatter(uint32_t N, const uint32_t *__restrict__ flag,
                                                      const uint32_t *__restrict__ pos,
                                                      const uint32_t *__restrict__ cpre,
                                                      uint32_t *__restrict__ hubs,
                                                      uint32_t *__restrict__ hchunk)
{
    const uint32_t u = blockIdx.x * kBlock + threadIdx.x;
    if (u < N && flag[u]) {
        hubs[pos[u]] = u;
        hchunk[pos[u]] = cpre[u];
    } else if (u == N) {
        hchunk[pos[N]] = cpre[N];
    }
}

// Binary search for x in row `row`; a malformed row (decreasing or past the
// end of E) reads nothing and counts as "absent".
__device__ bool row_contains(uint32_t N, const uint32_t *__restrict__ V,
                             const uint32_t *__restrict__ E, uint32_t row, uint32_t x)
{
    uint32_t lo = V[row], hi = V[row + 1];
    if (hi < lo || hi > V[N])
        return false;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        const uint32_t y = E[mid];
        if (y == x)
            return true;
        if (y < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return false;
}

// fast_io.c:509-540: V monotone, entries in range, strictly ascending rows
// (no repeats), no self loops, every edge present in both directions.
__global__ __launch_bounds__(kBlock) void check_csr(uint32_t N, const uint32_t *__restrict__ V,
                                                    const uint32_t *__restrict__ E,
                                                    uint32_t *__restrict__ bad)
{
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint32_t u = (uint32_t)(t / kWave);
    if (u >= N)
        return;
    const uint32_t beg = V[u], end = V[u + 1];
    uint32_t err = 0;
    if (end < beg || end > V[N]) {
        err = 1;
    } else {
        for (uint32_t i = beg + lane_id(); i < end; i += kWave) {
            const uint32_t v = E[i];
            if (v >= N || v == u || (i > beg && v <= E[i - 1]) || !row_contains(N, V, E, v, u))
                err = 1;
        }
    }
    const uint64_t any = __ballot(err != 0);
    if (lane_id() == 0 && any)
        atomicAdd(bad, 1u);
}

__global__ __launch_bounds__(kBlock) void gen_rmat_kernel(int scale, uint64_t L, gcx_rmat_thr thr,
                                                          uint64_t seed, uint32_t *__restrict__ I,
                                                          uint32_t *__restrict__ J)
{
    const uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k < L)
        gcx_rmat_pair(seed, scale, thr, k, I + k, J + k);
}

__global__ __launch_bounds__(kBlock) void gen_er_kernel(uint32_t n, uint64_t m, uint64_t seed,
                                                        uint32_t *__restrict__ I,
                                                        uint32_t *__restrict__ J)
{
    const uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k < m)
        gcx_er_pair(seed, n, k, I + k, J + k);
}

// Bounds check that every kernel indexing by V and E relies on: V[0] = 0,
// V non-decreasing, V[N] = M and every entry of E below N.
__global__ __launch_bounds__(kBlock) void csr_bounds(uint32_t N, uint64_t M,
                                                     const uint32_t *__restrict__ V,
                                                     const uint32_t *__restrict__ E,
                                                     uint32_t *__restrict__ bad)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t err = 0;
    if (i < N)
        err |= V[i + 1] < V[i] ? 1u : 0u;
    if (i == 0)
        err |= (V[0] != 0 || (uint64_t)V[N] != M) ? 1u : 0u;
    if (i < M)
        err |= E[i] >= N ? 1u : 0u;
    const uint64_t any = __ballot(err != 0);
    if (lane_id() == 0 && any)
        atomicAdd(bad, 1u);
}

// One key per CSR entry: (row << b) | column, for sorting rows in place.
__global__ __launch_bounds__(kBlock) void row_keys(uint32_t N, const uint32_t *__restrict__ V,
                                                   const uint32_t *__restrict__ E, int b,
                                                   uint64_t *__restrict__ keys)
{
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint32_t u = (uint32_t)(t / kWave);
    if (u >= N)
        return;
    for (uint32_t i = V[u] + lane_id(); i < V[u + 1]; i += kWave)
        keys[i] = ((uint64_t)u << b) | E[i];
}

__global__ __launch_bounds__(kBlock) void keys_low(uint64_t M, const uint64_t *__restrict__ keys,
                                                   int b, uint32_t *__restrict__ E)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t mask = (b == 32) ? 0xFFFFFFFFull : ((1ull << b) - 1ull);
    if (i < M)
        E[i] = (uint32_t)(keys[i] & mask);
}

// graph_make_simple (fast_io.c:457-477): in a row sorted ascending, entry j
// survives unless it is the row's own vertex or not larger than entry j-1.
__global__ __launch_bounds__(kBlock) void simple_flags(uint32_t N, const uint32_t *__restrict__ V,
                                                       const uint32_t *__restrict__ E,
                                                       uint32_t *__restrict__ flag,
                                                       uint32_t *__restrict__ deg)
{
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint32_t u = (uint32_t)(t / kWave);
    if (u >= N)
        return;
    const uint32_t beg = V[u], end = V[u + 1];
    uint32_t kept = 0;
    for (uint32_t i = beg + lane_id(); i < end; i += kWave) {
        const uint32_t v = E[i];
        const bool keep = v != u && !(i > beg && v <= E[i - 1]);
        flag[i] = keep ? 1u : 0u;
        kept += keep ? 1u : 0u;
    }
    kept = group_sum32<kWave>(kept);
    if (lane_id() == 0)
        deg[u] = kept;
}

__global__ __launch_bounds__(kBlock) void compact_entries(uint64_t M, const uint32_t *__restrict__ E,
                                                          const uint32_t *__restrict__ flag,
                                                          const uint32_t *__restrict__ pos,
                                                          uint32_t *__restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < M && flag[i])
        out[pos[i]] = E[i];
}

__global__ __launch_bounds__(kBlock) void gather_degrees(const uint32_t *__restrict__ V,
                                                         const uint32_t *__restrict__ list, uint32_t n,
                                                         uint32_t *__restrict__ out)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n)
        out[i] = V[list[i] + 1] - V[list[i]];
}

// ------------------------------------------------ random geometric (2D) pairs

__global__ __launch_bounds__(kBlock) void rgg_cells(uint32_t n, uint64_t seed, uint32_t cells,
                                                    uint32_t *__restrict__ px, uint32_t *__restrict__ py,
                                                    uint32_t *__restrict__ cell_of,
                                                    uint32_t *__restrict__ cellcnt)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n)
        return;
    uint32_t x, y;
    gcx_rgg_point(seed, i, &x, &y);
    px[i] = x;
    py[i] = y;
    const uint32_t c = gcx_rgg_cell(x, y, cells);
    cell_of[i] = c;
    atomicAdd(cellcnt + c, 1u);
}

__global__ __launch_bounds__(kBlock) void rgg_fill(uint32_t n, const uint32_t *__restrict__ cell_of,
                                                   uint32_t *__restrict__ fill,
                                                   uint32_t *__restrict__ order)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n)
        order[atomicAdd(fill + cell_of[i], 1u)] = i;
}

// Pass 0 counts, pass 1 writes the pairs (i, j), j > i, of point i.
template <int pass>
__global__ __launch_bounds__(kBlock) void rgg_pairs(uint32_t n, uint32_t cells, uint64_t r2,
                                                    const uint32_t *__restrict__ px,
                                                    const uint32_t *__restrict__ py,
                                                    const uint32_t *__restrict__ cell_of,
                                                    const uint32_t *__restrict__ cellstart,
                                                    const uint32_t *__restrict__ order,
                                                    uint32_t *__restrict__ cnt,
                                                    const uint64_t *__restrict__ off,
                                                    uint32_t *__restrict__ I, uint32_t *__restrict__ J)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t xi = px[i], yi = py[i], c = cell_of[i];
    const int64_t cx = c % cells, cy = c / cells;
    uint32_t k = 0;
    uint64_t w = pass ? off[i] : 0;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            const int64_t X = cx + dx, Y = cy + dy;
            if (X < 0 || Y < 0 || X >= cells || Y >= cells)
                continue;
            const uint32_t cc = (uint32_t)(Y * cells + X);
            for (uint32_t s = cellstart[cc]; s < cellstart[cc + 1]; s++) {
                const uint32_t j = order[s];
                if (j > i && gcx_rgg_close(xi, yi, px[j], py[j], r2)) {
                    if (pass) {
                        I[w] = i;
                        J[w] = j;
                        w++;
                    } else {
                        k++;
                    }
                }
            }
        }
    if (!pass)
        cnt[i] = k;
}

int bit_length(uint32_t x)
{
    int b = 0;
    while (x) {
        b++;
        x >>= 1;
    }
    return b;
}

// Stream-ordered temporary from the device memory pool: freeing it never
// synchronises the device, so work on other streams keeps running.
struct DevBuf {
    void *p = nullptr;
    hipStream_t s = nullptr;
    ~DevBuf()
    {
        if (p)
            (void)hipFreeAsync(p, s);
    }
    hipError_t alloc(size_t bytes, hipStream_t st)
    {
        s = st;
        return hipMallocAsync(&p, bytes ? bytes : 8, st);
    }
};

// Several temporaries carved from one DevBuf (256-byte aligned parts): each
// stream-ordered allocation and free costs host time, which adds up over the
// many small graphs of a serving trace.
struct DevParts {
    DevBuf buf;
    size_t total = 0;
    size_t add(size_t bytes)
    {
        const size_t at = total;
        total += (bytes + 255) & ~(size_t)255;
        return at;
    }
    hipError_t alloc(hipStream_t st) { return buf.alloc(total, st); }
    template <typename T> T *at(size_t off) { return (T *)((char *)buf.p + off); }
};

// Sort keys for the degree order: maxdeg - deg (so ascending = degree
// descending) and the vertex ids.
__global__ __launch_bounds__(kBlock) void degree_keys(uint32_t N, const uint32_t *__restrict__ V,
                                                      uint32_t maxdeg, uint32_t *__restrict__ k,
                                                      uint32_t *__restrict__ id)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N) {
        k[i] = maxdeg - (V[i + 1] - V[i]);
        id[i] = i;
    }
}

} // namespace

extern "C" int gcxk_csr_from_pairs(uint32_t N, uint64_t L, const uint32_t *dI, const uint32_t *dJ,
                                   uint32_t *dV, uint32_t **dE, uint64_t *M, void *stream)
{
    hipStream_t s = as_stream(stream);
    *dE = nullptr;
    *M = 0;
    if (N == 0)
        return GCX_ERR_ARG;
    const int b = bit_length(N - 1) ? bit_length(N - 1) : 1;
    const uint64_t n = 2 * L;
    DevParts tmp;
    uint64_t *k_out = nullptr;
    uint32_t *flag = nullptr, *pos = nullptr;
    uint32_t total = 0;
    if (L) {
        size_t tb = 0, tb2 = 0;
        GCXK_CHECK(rocprim::radix_sort_keys(nullptr, tb, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                            (size_t)n, 0, 2 * b, s));
        GCXK_CHECK(rocprim::exclusive_scan(nullptr, tb2, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                           0u, (size_t)n, rocprim::plus<uint32_t>(), s));
        const size_t o_in = tmp.add(n * sizeof(uint64_t)), o_out = tmp.add(n * sizeof(uint64_t));
        const size_t o_flag = tmp.add(n * sizeof(uint32_t)), o_pos = tmp.add(n * sizeof(uint32_t));
        const size_t o_t = tmp.add(tb ? tb : 1), o_t2 = tmp.add(tb2 ? tb2 : 1);
        GCXK_CHECK(tmp.alloc(s));
        uint64_t *k_in = tmp.at<uint64_t>(o_in);
        k_out = tmp.at<uint64_t>(o_out);
        flag = tmp.at<uint32_t>(o_flag);
        pos = tmp.at<uint32_t>(o_pos);
        expand_pairs<<<div_up(L, kBlock), kBlock, 0, s>>>(L, dI, dJ, b, k_in);
        GCXK_LAUNCH_CHECK();
        GCXK_CHECK(rocprim::radix_sort_keys(tmp.at<void>(o_t), tb, k_in, k_out, (size_t)n, 0, 2 * b,
                                            s));
        mark_unique<<<div_up(n, kBlock), kBlock, 0, s>>>(n, k_out, b, flag);
        GCXK_LAUNCH_CHECK();
        GCXK_CHECK(rocprim::exclusive_scan(tmp.at<void>(o_t2), tb2, flag, pos, 0u, (size_t)n,
                                           rocprim::plus<uint32_t>(), s));
        uint32_t last[2] = {0, 0};
        GCXK_CHECK(hipMemcpyAsync(&last[0], pos + (n - 1), 4, hipMemcpyDeviceToHost, s));
        GCXK_CHECK(hipMemcpyAsync(&last[1], flag + (n - 1), 4, hipMemcpyDeviceToHost, s));
        GCXK_CHECK(hipStreamSynchronize(s));
        total = last[0] + last[1];
    }
    uint32_t *E = nullptr;
    // GCXK_ER_PAD entries of slack: the GNN input layer reads whole 16-byte words
    GCXK_CHECK(hipMallocAsync((void **)&E, (total + GCXK_ER_PAD) * sizeof(uint32_t), s));
    hipError_t e = hipSuccess;
    if (total) {
        scatter_rows<<<div_up(n, kBlock), kBlock, 0, s>>>(n, k_out, flag, pos, b, N, total, dV, E);
        e = hipGetLastError();
    } else {
        e = hipMemsetAsync(dV, 0, ((size_t)N + 1) * sizeof(uint32_t), s);
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        (void)hipFreeAsync(E, s);
        return GCX_ERR_HIP;
    }
    *dE = E;
    *M = total;
    return GCX_OK;
}

// rank[v] = i for v = perm[i] (position of each vertex in the degree order),
// and rd[v] = (rank[v], deg(v)): the two per-vertex values the column pass
// gathers, in one 8-byte word.
__global__ __launch_bounds__(kBlock) void rank_deg(uint32_t N, const uint32_t *__restrict__ perm,
                                                   const uint32_t *__restrict__ V,
                                                   uint32_t *__restrict__ rank,
                                                   uint2 *__restrict__ rd)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N) {
        const uint32_t v = perm[i];
        rank[v] = i;
        rd[v] = make_uint2(i, V[v + 1] - V[v]);
    }
}

// Er[j] = rank[E[j]] (the CSR's columns as degree-order positions) and
// Ed[j] = deg(E[j]), then GCXK_ER_PAD zeros in both.
__global__ __launch_bounds__(kBlock) void rank_cols(uint64_t M, const uint32_t *__restrict__ E,
                                                    const uint2 *__restrict__ rd,
                                                    uint32_t *__restrict__ Er,
                                                    uint32_t *__restrict__ Ed)
{
    const uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j < M) {
        const uint2 x = rd[E[j]];
        Er[j] = x.x;
        Ed[j] = x.y;
    } else if (j < M + GCXK_ER_PAD) {
        Er[j] = 0u;
        Ed[j] = 0u;
    }
}

extern "C" int gcxk_rank_columns(uint32_t N, uint64_t M, const uint32_t *perm,
                                 const uint32_t *dV, const uint32_t *dE, uint32_t *rank,
                                 uint32_t *Er, uint32_t *Ed, void *stream)
{
    hipStream_t s = as_stream(stream);
    DevBuf rd;
    GCXK_CHECK(rd.alloc(8ull * N, s));
    if (N)
        rank_deg<<<div_up(N, kBlock), kBlock, 0, s>>>(N, perm, dV, rank,
                                                      static_cast<uint2 *>(rd.p));
    rank_cols<<<(unsigned)div_up(M + GCXK_ER_PAD, (uint64_t)kBlock), kBlock, 0, s>>>(
        M, dE, static_cast<const uint2 *>(rd.p), Er, Ed);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_degree_order(uint32_t N, const uint32_t *dV, uint32_t maxdeg, uint32_t *perm,
                                 void *stream)
{
    if (N == 0)
        return GCX_OK;
    hipStream_t s = as_stream(stream);
    const unsigned bits = maxdeg ? 32u - (unsigned)__builtin_clz(maxdeg) : 1u;
    size_t tb = 0;
    GCXK_CHECK(rocprim::radix_sort_pairs(nullptr, tb, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                         (uint32_t *)nullptr, (uint32_t *)nullptr, N, 0u, bits, s));
    DevParts tmp;
    const size_t o0 = tmp.add(4ull * N), o1 = tmp.add(4ull * N), oid = tmp.add(4ull * N);
    const size_t ot = tmp.add(tb ? tb : 1);
    GCXK_CHECK(tmp.alloc(s));
    uint32_t *k0 = tmp.at<uint32_t>(o0), *k1 = tmp.at<uint32_t>(o1), *id = tmp.at<uint32_t>(oid);
    degree_keys<<<div_up(N, kBlock), kBlock, 0, s>>>(N, dV, maxdeg, k0, id);
    GCXK_LAUNCH_CHECK();
    GCXK_CHECK(rocprim::radix_sort_pairs(tmp.at<void>(ot), tb, k0, k1, id, perm, N, 0u, bits, s));
    return GCX_OK;
}

extern "C" int gcxk_graph_stats(uint32_t N, const uint32_t *dV, uint32_t hub_deg, uint32_t *dhub,
                                uint32_t *dchunk, uint32_t *nhub, uint32_t *maxdeg, uint32_t *nchunk,
                                void *stream)
{
    hipStream_t s = as_stream(stream);
    *nhub = *maxdeg = *nchunk = 0;
    const size_t n1 = (size_t)N + 1;
    size_t tb = 0;
    GCXK_CHECK(rocprim::exclusive_scan(nullptr, tb, (uint32_t *)nullptr, (uint32_t *)nullptr, 0u,
                                       n1, rocprim::plus<uint32_t>(), s));
    DevParts tmp;
    const size_t of = tmp.add(4 * n1), oc = tmp.add(4 * n1), op = tmp.add(4 * n1),
                 ocp = tmp.add(4 * n1), om = tmp.add(4), ot = tmp.add(tb ? tb : 1);
    GCXK_CHECK(tmp.alloc(s));
    uint32_t *f = tmp.at<uint32_t>(of), *c = tmp.at<uint32_t>(oc), *ps = tmp.at<uint32_t>(op),
             *cp = tmp.at<uint32_t>(ocp), *mx = tmp.at<uint32_t>(om);
    GCXK_CHECK(hipMemsetAsync(mx, 0, 4, s));
    hub_flags<<<div_up(n1, kBlock), kBlock, 0, s>>>(N, dV, hub_deg, f, c, mx);
    GCXK_LAUNCH_CHECK();
    GCXK_CHECK(rocprim::exclusive_scan(tmp.at<void>(ot), tb, f, ps, 0u, n1,
                                       rocprim::plus<uint32_t>(), s));
    GCXK_CHECK(rocprim::exclusive_scan(tmp.at<void>(ot), tb, c, cp, 0u, n1,
                                       rocprim::plus<uint32_t>(), s));
    hub_scatter<<<div_up(n1, kBlock), kBlock, 0, s>>>(N, f, ps, cp, dhub, dchunk);
    GCXK_LAUNCH_CHECK();
    uint32_t h[3] = {0, 0, 0};
    GCXK_CHECK(hipMemcpyAsync(&h[0], ps + N, 4, hipMemcpyDeviceToHost, s));
    GCXK_CHECK(hipMemcpyAsync(&h[1], mx, 4, hipMemcpyDeviceToHost, s));
    GCXK_CHECK(hipMemcpyAsync(&h[2], cp + N, 4, hipMemcpyDeviceToHost, s));
    GCXK_CHECK(hipStreamSynchronize(s));
    *nhub = h[0];
    *maxdeg = h[1];
    *nchunk = h[2];
    return GCX_OK;
}

extern "C" int gcxk_graph_check(uint32_t N, const uint32_t *dV, const uint32_t *dE, uint32_t *dtmp,
                                int *valid, void *stream)
{
    hipStream_t s = as_stream(stream);
    GCXK_CHECK(hipMemsetAsync(dtmp, 0, sizeof(uint32_t), s));
    if (N)
        check_csr<<<div_up((uint64_t)N * kWave, kBlock), kBlock, 0, s>>>(N, dV, dE, dtmp);
    GCXK_LAUNCH_CHECK();
    uint32_t bad = 0;
    GCXK_CHECK(hipMemcpyAsync(&bad, dtmp, sizeof(bad), hipMemcpyDeviceToHost, s));
    GCXK_CHECK(hipStreamSynchronize(s));
    *valid = bad == 0;
    return GCX_OK;
}

extern "C" int gcxk_gen_rmat(int scale, uint64_t L, uint32_t t0, uint32_t t1, uint32_t t2,
                             uint64_t seed, uint32_t *dI, uint32_t *dJ, void *stream)
{
    if (L == 0)
        return GCX_OK;
    gcx_rmat_thr thr = {t0, t1, t2};
    gen_rmat_kernel<<<div_up(L, kBlock), kBlock, 0, as_stream(stream)>>>(scale, L, thr, seed, dI,
                                                                         dJ);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_gen_er(uint32_t n, uint64_t m, uint64_t seed, uint32_t *dI, uint32_t *dJ,
                           void *stream)
{
    if (m == 0)
        return GCX_OK;
    gen_er_kernel<<<div_up(m, kBlock), kBlock, 0, as_stream(stream)>>>(n, m, seed, dI, dJ);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_csr_bounds(uint32_t N, uint64_t M, const uint32_t *dV, const uint32_t *dE,
                               uint32_t *dtmp, int *ok, void *stream)
{
    hipStream_t s = as_stream(stream);
    GCXK_CHECK(hipMemsetAsync(dtmp, 0, sizeof(uint32_t), s));
    const uint64_t n = (uint64_t)N > M ? (uint64_t)N : M;
    csr_bounds<<<div_up(n ? n : 1, kBlock), kBlock, 0, s>>>(N, M, dV, dE, dtmp);
    GCXK_LAUNCH_CHECK();
    uint32_t bad = 0;
    GCXK_CHECK(hipMemcpyAsync(&bad, dtmp, sizeof(bad), hipMemcpyDeviceToHost, s));
    GCXK_CHECK(hipStreamSynchronize(s));
    *ok = bad == 0;
    return GCX_OK;
}

extern "C" int gcxk_sort_rows(uint32_t N, const uint32_t *dV, uint32_t *dE, uint64_t M,
                              void *stream)
{
    hipStream_t s = as_stream(stream);
    if (M == 0 || N == 0)
        return GCX_OK;
    const int b = bit_length(N - 1) ? bit_length(N - 1) : 1;
    DevBuf keys, keys2, temp;
    GCXK_CHECK(keys.alloc(M * sizeof(uint64_t), s));
    GCXK_CHECK(keys2.alloc(M * sizeof(uint64_t), s));
    row_keys<<<div_up((uint64_t)N * kWave, kBlock), kBlock, 0, s>>>(N, dV, dE, b,
                                                                    (uint64_t *)keys.p);
    GCXK_LAUNCH_CHECK();
    size_t tb = 0;
    GCXK_CHECK(rocprim::radix_sort_keys(nullptr, tb, (uint64_t *)keys.p, (uint64_t *)keys2.p,
                                        (size_t)M, 0, 2 * b, s));
    GCXK_CHECK(temp.alloc(tb, s));
    GCXK_CHECK(rocprim::radix_sort_keys(temp.p, tb, (uint64_t *)keys.p, (uint64_t *)keys2.p,
                                        (size_t)M, 0, 2 * b, s));
    keys_low<<<div_up(M, kBlock), kBlock, 0, s>>>(M, (uint64_t *)keys2.p, b, dE);
    GCXK_LAUNCH_CHECK();
    GCXK_CHECK(hipStreamSynchronize(s));
    return GCX_OK;
}

extern "C" int gcxk_make_simple(uint32_t N, uint32_t *dV, const uint32_t *dE, uint64_t M,
                                uint32_t *dE_out, uint64_t *M_out, void *stream)
{
    hipStream_t s = as_stream(stream);
    DevBuf flag, pos, deg, temp, temp2;
    GCXK_CHECK(flag.alloc(M * sizeof(uint32_t), s));
    GCXK_CHECK(pos.alloc(M * sizeof(uint32_t), s));
    GCXK_CHECK(deg.alloc(((size_t)N + 1) * sizeof(uint32_t), s));
    GCXK_CHECK(hipMemsetAsync(deg.p, 0, ((size_t)N + 1) * sizeof(uint32_t), s));
    if (N)
        simple_flags<<<div_up((uint64_t)N * kWave, kBlock), kBlock, 0, s>>>(
            N, dV, dE, (uint32_t *)flag.p, (uint32_t *)deg.p);
    GCXK_LAUNCH_CHECK();
    if (M) {
        size_t tb = 0;
        GCXK_CHECK(rocprim::exclusive_scan(nullptr, tb, (uint32_t *)flag.p, (uint32_t *)pos.p, 0u,
                                           (size_t)M, rocprim::plus<uint32_t>(), s));
        GCXK_CHECK(temp.alloc(tb, s));
        GCXK_CHECK(rocprim::exclusive_scan(temp.p, tb, (uint32_t *)flag.p, (uint32_t *)pos.p, 0u,
                                           (size_t)M, rocprim::plus<uint32_t>(), s));
        compact_entries<<<div_up(M, kBlock), kBlock, 0, s>>>(M, dE, (uint32_t *)flag.p,
                                                             (uint32_t *)pos.p, dE_out);
        GCXK_LAUNCH_CHECK();
    }
    size_t tb2 = 0;
    GCXK_CHECK(rocprim::exclusive_scan(nullptr, tb2, (uint32_t *)deg.p, dV, 0u, (size_t)N + 1,
                                       rocprim::plus<uint32_t>(), s));
    GCXK_CHECK(temp2.alloc(tb2, s));
    GCXK_CHECK(rocprim::exclusive_scan(temp2.p, tb2, (uint32_t *)deg.p, dV, 0u, (size_t)N + 1,
                                       rocprim::plus<uint32_t>(), s));
    uint32_t m32 = 0;
    GCXK_CHECK(hipMemcpyAsync(&m32, dV + N, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    GCXK_CHECK(hipStreamSynchronize(s));
    *M_out = m32;
    return GCX_OK;
}

extern "C" int gcxk_gather_degrees(const uint32_t *dV, const uint32_t *list, uint32_t n,
                                   uint32_t *out, void *stream)
{
    if (n == 0)
        return GCX_OK;
    gather_degrees<<<div_up(n, kBlock), kBlock, 0, as_stream(stream)>>>(dV, list, n, out);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_gen_rgg(uint32_t n, uint32_t cells, uint64_t r2, uint64_t seed, uint32_t **dI,
                            uint32_t **dJ, uint64_t *L, void *stream)
{
    hipStream_t s = as_stream(stream);
    *dI = *dJ = nullptr;
    *L = 0;
    if (n == 0 || cells == 0)
        return GCX_ERR_ARG;
    const uint64_t nc = (uint64_t)cells * cells;
    if (nc >= UINT32_MAX)
        return GCX_ERR_ARG;
    DevBuf px, py, cellof, cellcnt, cellstart, fill, order, cnt, off, t1, t2;
    GCXK_CHECK(px.alloc(4ull * n, s));
    GCXK_CHECK(py.alloc(4ull * n, s));
    GCXK_CHECK(cellof.alloc(4ull * n, s));
    GCXK_CHECK(order.alloc(4ull * n, s));
    GCXK_CHECK(cellcnt.alloc(4ull * (nc + 1), s));
    GCXK_CHECK(cellstart.alloc(4ull * (nc + 1), s));
    GCXK_CHECK(fill.alloc(4ull * (nc + 1), s));
    GCXK_CHECK(cnt.alloc(4ull * (n + 1ull), s));
    GCXK_CHECK(off.alloc(8ull * (n + 1ull), s));
    GCXK_CHECK(hipMemsetAsync(cellcnt.p, 0, 4ull * (nc + 1), s));
    GCXK_CHECK(hipMemsetAsync(cnt.p, 0, 4ull * (n + 1ull), s));
    uint32_t *P = (uint32_t *)px.p, *Q = (uint32_t *)py.p, *CO = (uint32_t *)cellof.p;
    rgg_cells<<<div_up(n, kBlock), kBlock, 0, s>>>(n, seed, cells, P, Q, CO, (uint32_t *)cellcnt.p);
    GCXK_LAUNCH_CHECK();
    size_t tb = 0;
    GCXK_CHECK(rocprim::exclusive_scan(nullptr, tb, (uint32_t *)cellcnt.p, (uint32_t *)cellstart.p,
                                       0u, (size_t)nc + 1, rocprim::plus<uint32_t>(), s));
    GCXK_CHECK(t1.alloc(tb, s));
    GCXK_CHECK(rocprim::exclusive_scan(t1.p, tb, (uint32_t *)cellcnt.p, (uint32_t *)cellstart.p, 0u,
                                       (size_t)nc + 1, rocprim::plus<uint32_t>(), s));
    GCXK_CHECK(hipMemcpyAsync(fill.p, cellstart.p, 4ull * (nc + 1), hipMemcpyDeviceToDevice, s));
    rgg_fill<<<div_up(n, kBlock), kBlock, 0, s>>>(n, CO, (uint32_t *)fill.p, (uint32_t *)order.p);
    rgg_pairs<0><<<div_up(n, kBlock), kBlock, 0, s>>>(n, cells, r2, P, Q, CO,
                                                     (const uint32_t *)cellstart.p,
                                                     (const uint32_t *)order.p, (uint32_t *)cnt.p,
                                                     nullptr, nullptr, nullptr);
    GCXK_LAUNCH_CHECK();
    size_t tb2 = 0;
    GCXK_CHECK(rocprim::exclusive_scan(nullptr, tb2, (uint32_t *)cnt.p, (uint64_t *)off.p, 0ull,
                                       (size_t)n + 1, rocprim::plus<uint64_t>(), s));
    GCXK_CHECK(t2.alloc(tb2, s));
    GCXK_CHECK(rocprim::exclusive_scan(t2.p, tb2, (uint32_t *)cnt.p, (uint64_t *)off.p, 0ull,
                                       (size_t)n + 1, rocprim::plus<uint64_t>(), s));
    uint64_t total = 0;
    GCXK_CHECK(hipMemcpyAsync(&total, (uint64_t *)off.p + n, 8, hipMemcpyDeviceToHost, s));
    GCXK_CHECK(hipStreamSynchronize(s));
    uint32_t *I = nullptr, *J = nullptr;
    if (hipMallocAsync((void **)&I, 4ull * (total ? total : 1), s) != hipSuccess ||
        hipMallocAsync((void **)&J, 4ull * (total ? total : 1), s) != hipSuccess) {
        if (I)
            (void)hipFreeAsync(I, s);
        return GCX_ERR_NOMEM;
    }
    rgg_pairs<1><<<div_up(n, kBlock), kBlock, 0, s>>>(n, cells, r2, P, Q, CO,
                                                     (const uint32_t *)cellstart.p,
                                                     (const uint32_t *)order.p, nullptr,
                                                     (const uint64_t *)off.p, I, J);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        (void)hipFreeAsync(I, s);
        (void)hipFreeAsync(J, s);
        return GCX_ERR_HIP;
    }
    *dI = I;
    *dJ = J;
    *L = total;
    return GCX_OK;
}
