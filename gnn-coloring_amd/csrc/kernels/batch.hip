// batch.hip -- many graphs packed into one block-diagonal graph.
//
// Packed graph g occupies vertices [off_g, off_g + n_g) and entries
// [eoff_g, eoff_g + m_g). It is the union of the graphs with shifted ids, so
// its CSR is canonical when each graph's is, ids keep their order inside a
// graph (JP's id tie-break is unchanged), and no edge joins two graphs: JP on
// the pack gives every graph exactly the colours it gets alone. The GNN's
// id feature must stay graph-local, so the pack carries fid[v] =
// (float)((double)(v - off_g) / n_g).
#include "common.hip.h"

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void pack_rows(uint32_t n, const uint32_t *__restrict__ V,
                                                    uint32_t off, uint32_t eoff,
                                                    uint32_t *__restrict__ Vout,
                                                    float *__restrict__ fid)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        Vout[off + i] = eoff + V[i];
        fid[off + i] = (float)((double)i / (double)n);
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
__global__ __launch_bounds__(kBlock) void packed_stats(uint32_t N, uint32_t count,
                                                       const uint32_t *__restrict__ goff,
                                                       const int *__restrict__ color,
                                                       int *__restrict__ maxc,
                                                       unsigned long long *__restrict__ digest)
{
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= N)
        return;
    uint32_t lo = 0, hi = count; // last g with goff[g] <= v
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (goff[mid] <= v)
            lo = mid;
        else
            hi = mid;
    }
    const int c = color[v];
    atomicMax(&maxc[lo], c);
    atomicAdd(&digest[lo], (unsigned long long)mix64(((uint64_t)(v - goff[lo]) << 32) | (uint32_t)c));
}

} // namespace

extern "C" int gcxk_pack_graph(uint32_t n, uint64_t m, const uint32_t *dV, const uint32_t *dE,
                               uint32_t off, uint32_t eoff, uint32_t *Vout, uint32_t *Eout,
                               float *fid, void *stream)
{
    hipStream_t s = as_stream(stream);
    if (n)
        pack_rows<<<div_up(n, kBlock), kBlock, 0, s>>>(n, dV, off, eoff, Vout, fid);
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

extern "C" int gcxk_packed_stats(uint32_t N, uint32_t count, const uint32_t *goff, const int *color,
                                 int *maxc, uint64_t *digest, void *stream)
{
    if (N == 0)
        return GCX_OK;
    packed_stats<<<div_up(N, kBlock), kBlock, 0, as_stream(stream)>>>(
        N, count, goff, color, maxc, reinterpret_cast<unsigned long long *>(digest));
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}
