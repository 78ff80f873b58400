// keys.hip -- elementwise ordering keys (32-bit; larger = coloured earlier).
//   FF  first_fit.c:8-19              key = N - id
//   LF  largest_degree_first.c:12-16  key = deg
//   LLF largest_degree_first.c:87-97  key = bit length of deg (0 for deg 0)
// plus conversions between the reference's double priorities and keys.
#include "common.hip.h"

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void keys_ff(uint32_t N, uint32_t *__restrict__ key)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N)
        key[i] = N - i;
}

__global__ __launch_bounds__(kBlock) void keys_lf(uint32_t N, const uint32_t *__restrict__ V,
                                                  uint32_t *__restrict__ key)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N)
        key[i] = V[i + 1] - V[i];
}

__global__ __launch_bounds__(kBlock) void keys_llf(uint32_t N, const uint32_t *__restrict__ V,
                                                   uint32_t *__restrict__ key)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N) {
        const uint32_t d = V[i + 1] - V[i];
        key[i] = d ? 32u - (uint32_t)__builtin_clz(d) : 0u;
    }
}

__global__ __launch_bounds__(kBlock) void keys_from_colors(uint32_t N, const int *__restrict__ c,
                                                           uint32_t *__restrict__ key)
{
    // Culberson step (main.c:38-40): p = colour; colours are >= 1.
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N)
        key[i] = (uint32_t)c[i];
}

// ---- Culberson class orders (iterated greedy with reordered colour classes)
//
// A pass colours the vertices class by class in some order of the previous
// colour classes; every order gives at most as many colours. key[v] =
// ckey[color[v]], larger = earlier:
//   reverse        ckey[c] = c (the reference's p = colour, main.c:38-40)
//   largest first  classes by size descending, ties by colour descending
//   random         ckey[c] = (hi32(mix64((seed + pass) ^ c)) << b) | c with
//                  b = 12 (b = bits of C when C >= 4096)
//   mixed          pass i uses reverse, largest first, random for i mod 3
// The pass counter lives in device memory and advances once per pass, so a
// captured pass replays with a new order each time.

__global__ __launch_bounds__(kBlock) void class_clear(uint32_t C, uint32_t *__restrict__ cnt)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i <= C)
        cnt[i] = 0u;
}

constexpr uint32_t kClassLds = 4096;

__global__ __launch_bounds__(kBlock) void class_count(uint32_t N, uint32_t C, const int *__restrict__ color,
                                                      uint32_t *__restrict__ cnt)
{
    __shared__ uint32_t h[kClassLds];
    const bool lds = C < kClassLds;
    if (lds)
        for (uint32_t i = threadIdx.x; i <= C; i += kBlock)
            h[i] = 0u;
    __syncthreads();
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < N; v += gridDim.x * kBlock) {
        const uint32_t c = (uint32_t)color[v];
        if (c <= C) {
            if (lds)
                atomicAdd(&h[c], 1u);
            else
                atomicAdd(&cnt[c], 1u);
        }
    }
    __syncthreads();
    if (lds)
        for (uint32_t i = threadIdx.x; i <= C; i += kBlock)
            if (h[i])
                atomicAdd(&cnt[i], h[i]);
}

// One workgroup: ckey for classes 1..C, then the pass counter advances.
__global__ __launch_bounds__(1024) void class_keys(uint32_t C, int strategy, uint64_t seed,
                                                   const uint32_t *__restrict__ cnt,
                                                   uint32_t *__restrict__ ckey, uint32_t *pass)
{
    const uint32_t it = *pass;
    const int st = strategy == 3 ? (int)(it % 3u) : strategy;
    const uint32_t b = C < 4096u ? 12u : 32u - (uint32_t)__builtin_clz(C);
    for (uint32_t c = threadIdx.x + 1; c <= C; c += 1024) {
        uint32_t k = c;
        if (st == 1) {
            // rank among classes: larger first, then higher colour first
            const uint32_t sc = cnt[c];
            uint32_t r = 0;
            for (uint32_t d = 1; d <= C; d++) {
                const uint32_t sd = cnt[d];
                r += (sd > sc || (sd == sc && d > c)) ? 1u : 0u;
            }
            k = C - r;
        } else if (st == 2) {
            const uint64_t h = mix64((seed + it) ^ (uint64_t)c);
            k = b >= 32u ? c : (((uint32_t)(h >> 32)) << b) | c;
        }
        ckey[c] = k;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        *pass = it + 1u;
}

__global__ __launch_bounds__(kBlock) void keys_from_classes(uint32_t N, uint32_t C, const int *__restrict__ color,
                                                            const uint32_t *__restrict__ ckey,
                                                            uint32_t *__restrict__ key)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N) {
        const uint32_t c = (uint32_t)color[i];
        key[i] = c <= C ? ckey[c] : c;
    }
}

__global__ __launch_bounds__(kBlock) void scan_doubles(uint32_t N, const double *__restrict__ p,
                                                       gcxk_ctl *ctl)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    bool nan = false, nonfloat = false;
    if (i < N) {
        const double d = p[i];
        nan = d != d;
        nonfloat = !nan && (double)(float)d != d;
    }
    const uint64_t bn = __ballot(nan), bf = __ballot(nonfloat);
    if (lane_id() == 0) {
        if (bn)
            atomicAdd(&ctl->nan_count, (uint32_t)__popcll(bn));
        if (bf)
            atomicAdd(&ctl->nonfloat, (uint32_t)__popcll(bf));
    }
}

__global__ __launch_bounds__(kBlock) void doubles_to_key32(uint32_t N, const double *__restrict__ p,
                                                           uint32_t *__restrict__ key)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N)
        key[i] = float_order_key((float)p[i]);
}

__global__ __launch_bounds__(kBlock) void doubles_to_key64(uint32_t N, const double *__restrict__ p,
                                                           uint64_t *__restrict__ key)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N)
        key[i] = double_order_key(p[i]);
}

__global__ __launch_bounds__(kBlock) void key32_to_doubles(uint32_t N, const uint32_t *__restrict__ key,
                                                           int kind, double *__restrict__ p)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N)
        p[i] = kind == 0 ? (double)key[i] : (double)float_from_order_key(key[i]);
}

__global__ __launch_bounds__(kBlock) void key64_to_doubles(uint32_t N, const uint64_t *__restrict__ key,
                                                           double *__restrict__ p)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N)
        p[i] = double_from_order_key(key[i]);
}

// Plain streaming copy, used to calibrate HBM byte counters.
__global__ __launch_bounds__(kBlock) void copy_u32x4(uint64_t n4, const uint4 *__restrict__ src,
                                                     uint4 *__restrict__ dst)
{
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n4; i += stride)
        dst[i] = src[i];
}

} // namespace

extern "C" int gcxk_copy_bytes(uint64_t bytes, const void *src, void *dst, void *stream)
{
    const uint64_t n4 = bytes / 16u;
    if (n4 == 0)
        return GCX_OK;
    copy_u32x4<<<8192, kBlock, 0, as_stream(stream)>>>(n4, (const uint4 *)src, (uint4 *)dst);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

#define ELEMENTWISE(kernel, ...)                                                \
    do {                                                                        \
        if (N == 0)                                                             \
            return GCX_OK;                                                      \
        kernel<<<div_up(N, kBlock), kBlock, 0, as_stream(stream)>>>(__VA_ARGS__); \
        GCXK_LAUNCH_CHECK();                                                    \
        return GCX_OK;                                                          \
    } while (0)

extern "C" int gcxk_keys_ff(uint32_t N, uint32_t *key, void *stream) { ELEMENTWISE(keys_ff, N, key); }

extern "C" int gcxk_keys_lf(uint32_t N, const uint32_t *dV, uint32_t *key, void *stream)
{
    ELEMENTWISE(keys_lf, N, dV, key);
}

extern "C" int gcxk_keys_llf(uint32_t N, const uint32_t *dV, uint32_t *key, void *stream)
{
    ELEMENTWISE(keys_llf, N, dV, key);
}

extern "C" int gcxk_keys_from_colors(uint32_t N, const int *color, uint32_t *key, void *stream)
{
    ELEMENTWISE(keys_from_colors, N, color, key);
}

extern "C" int gcxk_culberson_keys(uint32_t N, uint32_t C, int strategy, uint64_t seed,
                                   const int *color, uint32_t *pass, uint32_t *scratch,
                                   uint32_t *key, void *stream)
{
    if (N == 0)
        return GCX_OK;
    if (strategy == 0)
        return gcxk_keys_from_colors(N, color, key, stream);
    if (strategy < 0 || strategy > 3)
        return GCX_ERR_ARG;
    hipStream_t s = as_stream(stream);
    uint32_t *cnt = scratch, *ckey = cnt + C + 1;
    class_clear<<<div_up((uint64_t)C + 1, kBlock), kBlock, 0, s>>>(C, cnt);
    const uint32_t grid = div_up(N, kBlock) < 1024u ? div_up(N, kBlock) : 1024u;
    class_count<<<grid, kBlock, 0, s>>>(N, C, color, cnt);
    class_keys<<<1, 1024, 0, s>>>(C, strategy, seed, cnt, ckey, pass);
    keys_from_classes<<<div_up(N, kBlock), kBlock, 0, s>>>(N, C, color, ckey, key);
    GCXK_LAUNCH_CHECK();
    return GCX_OK;
}

extern "C" int gcxk_keys_scan_doubles(uint32_t N, const double *p, gcxk_ctl *ctl, void *stream)
{
    ELEMENTWISE(scan_doubles, N, p, ctl);
}

extern "C" int gcxk_keys_from_doubles32(uint32_t N, const double *p, uint32_t *key, void *stream)
{
    ELEMENTWISE(doubles_to_key32, N, p, key);
}

extern "C" int gcxk_keys_from_doubles64(uint32_t N, const double *p, uint64_t *key, void *stream)
{
    ELEMENTWISE(doubles_to_key64, N, p, key);
}

extern "C" int gcxk_keys_to_doubles(uint32_t N, const uint32_t *key, int kind, double *p,
                                    void *stream)
{
    ELEMENTWISE(key32_to_doubles, N, key, kind, p);
}

extern "C" int gcxk_keys64_to_doubles(uint32_t N, const uint64_t *key, double *p, void *stream)
{
    ELEMENTWISE(key64_to_doubles, N, key, p);
}
