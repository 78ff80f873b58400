// batch.hip -- many graphs packed into one block-diagonal graph.
//
// Packed graph g occupies vertices [off_g, off_g + n_g) and entries
// [eoff_g, eoff_g + m_g). It is the union of the graphs with shifted ids, so
// its CSR is canonical when each graph's is, ids keep their order inside a
// graph (JP's id tie-break is unchanged), and no edge joins two graphs: JP on
// the pack gives every graph exactly the colours it gets alone. The GNN's
// id feature must stay graph-local, (float)((double)(v - off_g) / n_g), so
// the pack carries each vertex's graph span gspan[v] = (off_g, n_g).
#include "common.hip.h"

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void pack_rows(uint32_t n, const uint32_t *__restrict__ V,
                                                    uint32_t off, uint32_t eoff,
                                                    uint32_t *__restrict__ Vout,
                                                    uint2 *__restrict__ gspan)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        Vout[off + i] = eoff + V[i];
        gspan[off + i] = make_uint2(off, n);
    }
}

__global__ __launch_bounds__(kBlock) void pack_cols(uint64_t m, const uint32_t *__restrict__ E,
                                                    uint32_t off, uint32_t *__restrict__ Eout)
{
    const uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j < m)
        Eout[j] = off + E[j];
}

__global__ void pack_close(uint32_t N, uint32_t M, uint32_t *__restrict__ Vout)
{
    Vout[N] = M;
}

// Per packed graph: the largest colour and the colouring digest over
// graph-local ids (the digest gcx_color_digest gives for the graph alone).
// Block b reduces the vertex range [beg[b], end[b]) of graph gid[b] and
// adds one partial per counter, so a large graph costs one atomic per
// range, not per vertex.
__global__ __launch_bounds__(kBlock) void packed_stats(const uint32_t *__restrict__ rbeg,
                                                       const uint32_t *__restrict__ rend,
                                                       const uint32_t *__restrict__ rgid,
                                                       const uint32_t *__restrict__ goff,
                                                       const int *__restrict__ color,
                                                       int *__restrict__ maxc,
                                                       unsigned long long *__restrict__ digest)
{
    __shared__ int smax[kBlock / kWave];
    __shared__ unsigned long long ssum[kBlock / kWave];
    const uint32_t beg = rbeg[blockIdx.x], end = rend[blockIdx.x], g = rgid[blockIdx.x];
    const uint32_t base = goff[g];
    int m = 0;
    uint64_t h = 0;
    for (uint32_t v = beg + threadIdx.x; v < end; v += kBlock) {
        const int c = color[v];
        m = c > m ? c : m;
        h += mix64(((uint64_t)(v - base) << 32) | (uint32_t)c);
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const int mo = __shfl_xor(m, off);
        m = mo > m ? mo : m;
        const uint32_t lo = __shfl_xor((uint32_t)h, off), hi = __shfl_xor((uint32_t)(h >> 32), off);
        h += ((uint64_t)hi << 32) | lo;
    }
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        smax[w] = m;
        ssum[w] = h;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int bm = 0;
        unsigned long long bs = 0;
        for (int i = 0; i < kBlock / kWave; i++) {
            bm = smax[i] > bm ? smax[i] : bm;
            bs += ssum[i];
        }
        atomicMax(&maxc[g], bm);
        atomicAdd(&digest[g], bs);
    }
}

} // namespace

extern "C" int gcxk_pack_graph(uint32_t n, uint64_t m, const uint32_t *dV, const uint32_t *dE,
                               uint32_t off, uint32_t eoff, uint32_t *Vout, uint32_t *Eout,
                               uint32_t *gspan, void *stream)
{
    hipStream_t s = as_stream(stream);
    if (n)
        pack_rows<<<div_up(n, kBlock), kBlock, 0, s>>>(n, dV, off, eoff, Vout,
                                                       reinterpret_cast<uint2 *>(gspan));
    if (m)
        pack_cols<<<div_up(m, kBlock), kBlock, 0, s>>>(m, dE, off, Eout + eoff);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_pack_close(uint32_t N, uint32_t M, uint32_t *Vout, void *stream)
{
    pack_close<<<1, 1, 0, as_stream(stream)>>>(N, M, Vout);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_packed_stats(uint32_t nranges, const uint32_t *rbeg, const uint32_t *rend,
                                 const uint32_t *rgid, const uint32_t *goff, const int *color,
                                 int *maxc, uint64_t *digest, void *stream)
{
    if (nranges == 0)
        return GCX_OK;
    packed_stats<<<nranges, kBlock, 0, as_stream(stream)>>>(
        rbeg, rend, rgid, goff, color, maxc, reinterpret_cast<unsigned long long *>(digest));
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}
