// noalloc.hip -- Jones-Plassmann colouring behind the reference's
// jones_plassmann_internal interface (host V, E, p in; host colours out)
// without any allocation: every device buffer of the call -- the staged
// graph and priorities, keys, predecessor counts, the split rows, state
// words, work items, sort scratch and the control block -- is carved from
// one file-scope device array, `scratch`. The kernels are the library's own
// (orientation, greedy-order items, the persistent ordered kernel, and the
// round kernels as the fallback); they take pointers into the array.
//
// The array's device address comes from hipGetSymbolAddress, once per
// device, when the library initialises (gcxk_noalloc_init, called by
// gcx_init with the device current). In host code the array's name is only a
// stand-in for that address; kernels and copies always get the pointer.
//
// Calls are synchronous and serialised by a process-wide lock, so two calls
// of one process never use the array at once. The array keeps the last
// call's contents: everything a call reads is written or cleared by the same
// call first.
#include <pthread.h>
#include <stdio.h>

#include "coloring.h"
#include "common.hip.h"

extern "C" int gcx__ensure_init(void);
extern "C" void *gcx__default_stream(void);

namespace {

constexpr size_t kScratchBytes = size_t(GCX_NOALLOC_SCRATCH_MB) << 20;
constexpr int kMaxDevices = 64;
constexpr uint32_t kRoundBatch = 32;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Per-device address of `scratch` (hipGetSymbolAddress at initialisation).
void *g_scratch_ptr[kMaxDevices];
pthread_mutex_t g_noalloc_lock = PTHREAD_MUTEX_INITIALIZER;

__global__ void iota_kernel(uint32_t n, uint32_t *__restrict__ out)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n)
        out[i] = i;
}

// Byte offsets of every buffer for a graph of N vertices, M entries and
// key_bits-bit keys; `total` = bytes used.
struct layout {
    size_t V, E, p, key, color, npred, esplit, state, wof, wdesc, ovf, items0, items1, perm, cnt,
        incl, sort, ctl, tmp, total;
    size_t sort_bytes;
    uint64_t state_words;
};

size_t carve(size_t *at, size_t bytes)
{
    const size_t off = *at;
    *at += (bytes + 255) & ~size_t(255);
    return off;
}

layout make_layout(uint32_t N, uint64_t M, int key_bits)
{
    layout L{};
    const size_t n = size_t(N) + 1;
    const size_t items = sizeof(gcxk_item) * gcxk_jp_item_capacity(N, M);
    size_t at = 0;
    L.state_words = gcxk_jp_state_words(N, M);
    L.sort_bytes = gcxk_jp_order_scratch(N, key_bits);
    L.V = carve(&at, 4 * n);
    L.E = carve(&at, 4 * (M ? M : 1));
    L.p = carve(&at, 8 * n);
    L.key = carve(&at, 8 * n);
    L.color = carve(&at, 4 * n);
    L.npred = carve(&at, 4 * n);
    L.esplit = carve(&at, 4 * (M ? M : 1));
    L.state = carve(&at, 8 * L.state_words);
    L.wof = carve(&at, 4 * n);
    L.wdesc = carve(&at, 8 * n);
    L.ovf = carve(&at, 4 * gcxk_jp_ovf_words(N, M));
    L.items0 = carve(&at, items);
    L.items1 = carve(&at, items);
    L.perm = carve(&at, 4 * n);
    L.cnt = carve(&at, 4 * n);
    L.incl = carve(&at, 4 * n);
    L.sort = carve(&at, L.sort_bytes);
    L.ctl = carve(&at, sizeof(gcxk_ctl));
    L.tmp = carve(&at, 16);
    L.total = at;
    return L;
}

template <typename T> T *at(void *base, size_t off)
{
    return reinterpret_cast<T *>(static_cast<unsigned char *>(base) + off);
}

int read_ctl(const gcxk_ctl *dctl, gcxk_ctl *h, hipStream_t s)
{
    GCXK_CHECK(hipMemcpyAsync(h, dctl, sizeof(*h), hipMemcpyDeviceToHost, s));
    GCXK_CHECK(hipStreamSynchronize(s));
    return GCX_OK;
}

// Rounds schedule over the orientation already in the scratch (the fallback
// when an ordered wait hits its bound).
int color_by_rounds(uint32_t N, uint64_t M, void *b, const layout &L, int key_bits, int mb,
                    hipStream_t s)
{
    gcxk_ctl *ctl = at<gcxk_ctl>(b, L.ctl);
    const uint32_t *V = at<uint32_t>(b, L.V);
    int rc = gcxk_jp_clear(N, M, at<uint32_t>(b, L.ovf), 0, nullptr, nullptr,
                           at<uint32_t>(b, L.npred), at<uint64_t>(b, L.state),
                           at<int>(b, L.color), ctl, s);
    if (rc == GCX_OK)
        rc = gcxk_jp_init(N, V, at<uint32_t>(b, L.E), at<void>(b, L.key), key_bits, mb,
                          at<uint32_t>(b, L.perm), at<uint32_t>(b, L.npred),
                          at<uint32_t>(b, L.esplit), at<uint64_t>(b, L.state),
                          at<int>(b, L.color), nullptr, 0, nullptr, 0, nullptr, 0u, s, s);
    gcxk_item *q[2] = {at<gcxk_item>(b, L.items0), at<gcxk_item>(b, L.items1)};
    if (rc == GCX_OK)
        rc = gcxk_jp_roots(N, V, at<uint32_t>(b, L.npred), q[0], ctl, s);
    // every round either colours a non-empty frontier or sets `done`; a DAG
    // over N vertices has at most N levels
    const uint32_t max_rounds = N + 2 * kRoundBatch + 2;
    for (uint32_t r = 0; rc == GCX_OK;) {
        for (uint32_t i = 0; rc == GCX_OK && i < kRoundBatch; i++, r++)
            rc = gcxk_jp_round(V, at<uint32_t>(b, L.npred), at<uint32_t>(b, L.esplit),
                               at<uint64_t>(b, L.state), at<int>(b, L.color),
                               at<uint32_t>(b, L.ovf), mb, q, ctl, r, gcxk_round_grid(), s);
        gcxk_ctl h;
        if (rc == GCX_OK)
            rc = read_ctl(ctl, &h, s);
        if (rc != GCX_OK || h.done)
            break;
        if (r >= max_rounds)
            rc = GCX_ERR_HIP;
    }
    return rc;
}

int color_noalloc(uint32_t N, const uint32_t *V, const uint32_t *E, const double *p, int *color)
{
    const uint64_t M = V[N];
    if (N > GCX_NOALLOC_MAX_N || M > GCX_NOALLOC_MAX_M) {
        fprintf(stderr,
                "gcx: jones_plassmann_internal_noalloc: N=%u, M=%llu exceed its limits "
                "(N <= %u, M <= %u)\n",
                N, (unsigned long long)M, (unsigned)GCX_NOALLOC_MAX_N,
                (unsigned)GCX_NOALLOC_MAX_M);
        return GCX_ERR_CAPACITY;
    }
    // sized for 64-bit keys, whose sort needs the most room (32-bit keys use
    // a prefix of the same sort region); the limits above always fit
    const layout L = make_layout(N, M, 64);
    if (L.sort_bytes == 0 || L.total > kScratchBytes) {
        fprintf(stderr,
                "gcx: jones_plassmann_internal_noalloc: N=%u, M=%llu need %zu bytes, more than "
                "its %zu-byte device scratch\n",
                N, (unsigned long long)M, L.total, kScratchBytes);
        return GCX_ERR_CAPACITY;
    }
    int dev = 0;
    GCXK_CHECK(hipGetDevice(&dev));
    void *b = (dev >= 0 && dev < kMaxDevices) ? g_scratch_ptr[dev] : nullptr;
    if (!b)
        return GCX_ERR_NODEV;
    hipStream_t s = as_stream(gcx__default_stream());
    uint32_t maxdeg = 0;
    for (uint32_t u = 0; u < N; u++) {
        if (V[u + 1] < V[u])
            return GCX_ERR_ARG;
        const uint32_t d = V[u + 1] - V[u];
        maxdeg = d > maxdeg ? d : maxdeg;
    }
    // stage the inputs, then check every index before any kernel follows it
    gcxk_ctl *ctl = at<gcxk_ctl>(b, L.ctl);
    GCXK_CHECK(hipMemcpyAsync(at<void>(b, L.V), V, 4ull * (N + 1ull), hipMemcpyHostToDevice, s));
    if (M)
        GCXK_CHECK(hipMemcpyAsync(at<void>(b, L.E), E, 4ull * M, hipMemcpyHostToDevice, s));
    GCXK_CHECK(hipMemcpyAsync(at<void>(b, L.p), p, 8ull * N, hipMemcpyHostToDevice, s));
    int ok = 0;
    int rc = gcxk_csr_bounds(N, M, at<uint32_t>(b, L.V), at<uint32_t>(b, L.E),
                             at<uint32_t>(b, L.tmp), &ok, s);
    if (rc != GCX_OK)
        return rc;
    if (!ok)
        return GCX_ERR_ARG;
    // priorities -> keys: 32-bit when every value is exactly a float
    GCXK_CHECK(hipMemsetAsync(ctl, 0, sizeof(gcxk_ctl), s));
    if ((rc = gcxk_keys_scan_doubles(N, at<double>(b, L.p), ctl, s)) != GCX_OK)
        return rc;
    gcxk_ctl h;
    if ((rc = read_ctl(ctl, &h, s)) != GCX_OK)
        return rc;
    if (h.nan_count)
        return GCX_ERR_ARG;
    const int key_bits = h.nonfloat ? 64 : 32;
    if (key_bits == 32) {
        rc = gcxk_keys_from_doubles32(N, at<double>(b, L.p), at<uint32_t>(b, L.key), s);
    } else {
        rc = gcxk_keys_from_doubles64(N, at<double>(b, L.p), at<uint64_t>(b, L.key), s);
    }
    if (rc != GCX_OK)
        return rc;
    const int mb = gcxk_jp_mask_bits(maxdeg);
    iota_kernel<<<div_up(N, 256), 256, 0, s>>>(N, at<uint32_t>(b, L.perm));
    GCXK_LAUNCH_CHECK();
    // orientation (no hub split: every vertex four lanes, in id order)
    rc = gcxk_jp_clear(N, M, at<uint32_t>(b, L.ovf), 0, nullptr, nullptr, at<uint32_t>(b, L.npred),
                       at<uint64_t>(b, L.state), at<int>(b, L.color), ctl, s);
    if (rc == GCX_OK)
        rc = gcxk_jp_init(N, at<uint32_t>(b, L.V), at<uint32_t>(b, L.E), at<void>(b, L.key),
                          key_bits, mb, at<uint32_t>(b, L.perm), at<uint32_t>(b, L.npred),
                          at<uint32_t>(b, L.esplit), at<uint64_t>(b, L.state),
                          at<int>(b, L.color), nullptr, 0, nullptr, 0, nullptr, 0u, s, s);
    if (rc == GCX_OK)
        rc = gcxk_jp_order_items(N, at<uint32_t>(b, L.V), at<uint32_t>(b, L.npred),
                                 at<void>(b, L.key), key_bits, mb, at<void>(b, L.sort),
                                 L.sort_bytes, at<uint32_t>(b, L.cnt), at<uint32_t>(b, L.incl),
                                 at<uint32_t>(b, L.wof), at<uint64_t>(b, L.wdesc),
                                 at<uint64_t>(b, L.state), L.state_words,
                                 at<gcxk_item>(b, L.items0), ctl, s);
    // one cooperative persistent launch; every wait is bounded (2 s)
    if (rc == GCX_OK)
        rc = gcxk_jp_ordered(at<uint32_t>(b, L.V), at<uint32_t>(b, L.npred),
                             at<uint32_t>(b, L.esplit), at<uint64_t>(b, L.state),
                             at<uint64_t>(b, L.wdesc), at<int>(b, L.color),
                             at<uint32_t>(b, L.ovf), mb, at<gcxk_item>(b, L.items0), ctl, 0,
                             200000000ull, 1, s);
    if (rc == GCX_OK)
        rc = read_ctl(ctl, &h, s);
    if (rc == GCX_OK && h.err)
        rc = color_by_rounds(N, M, b, L, key_bits, mb, s);
    if (rc != GCX_OK)
        return rc;
    GCXK_CHECK(hipMemcpyAsync(color, at<int>(b, L.color), 4ull * N, hipMemcpyDeviceToHost, s));
    GCXK_CHECK(hipStreamSynchronize(s));
    return GCX_OK;
}

} // namespace

// Called by the library's initialisation with the device current.
extern "C" int gcxk_noalloc_init(void)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return GCX_ERR_NODEV;
    if (g_scratch_ptr[dev])
        return GCX_OK;
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
        return GCX_ERR_HIP;
    g_scratch_ptr[dev] = ptr;
    return GCX_OK;
}

extern "C" int jones_plassmann_internal_noalloc(uint32_t N, const uint32_t *V, const uint32_t *E,
                                                const double *p, int *color)
{
    int rc = gcx__ensure_init();
    if (rc != GCX_OK)
        return rc;
    if (N == 0)
        return GCX_OK;
    if (!V || !p || !color || (V[N] && !E))
        return GCX_ERR_ARG;
    pthread_mutex_lock(&g_noalloc_lock);
    rc = color_noalloc(N, V, E, p, color);
    pthread_mutex_unlock(&g_noalloc_lock);
    return rc;
}
