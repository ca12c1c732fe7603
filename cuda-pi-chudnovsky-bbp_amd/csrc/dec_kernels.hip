// ============================================================================
//  dec_kernels.hip -- device building blocks of divide-and-conquer decimal
//  conversion (the reference converts with repeated division by 10^9 over the
//  whole number, src/bignum_div.cpp:239-267): batched right shifts of
//  magnitudes (the quotient estimates q = x V >> s of every node of a level)
//  and the decimal digits of the leaves.
// ============================================================================
#include "pdh_internal.h"
#include <algorithm>

namespace pdh {

namespace {

constexpr int SH_NT = 256;
constexpr int SH_K = SHR_TILE / SH_NT;
static_assert(SH_NT * SH_K == SHR_TILE, "tile size");

// Right shift (out[j] = limbs j + w and j + w + 1 of a joined and shifted
// right by s) or left shift (out[j] = limbs j - w and j - w - 1 joined and
// shifted left by s), bits = 32 w + s. The normalized length follows from
// a's top limb alone, which the first thread of the node reads: right, the
// top candidate is la - w - 1, nonzero unless a's top limb is below 2^s (and
// then limb la - w - 2 holds that limb's bits); left, la + w, nonzero iff the
// top limb has bits above 32 - s. nodes == nullptr: one node in `one`.
__global__ __launch_bounds__(SH_NT) void k_shift_batch(const ShiftNode* __restrict__ nodes,
                                                       const uint32_t* __restrict__ tile_node, ShiftNode one) {
    const uint32_t t = blockIdx.x;
    const ShiftNode N = nodes ? nodes[tile_node[t]] : one;
    const uint64_t la = *N.a_len;
    const uint64_t w = N.bits >> 5;
    const uint32_t s = (uint32_t)(N.bits & 31);
    const uint64_t base = (uint64_t)(t - N.tile0) * SHR_TILE;
    for (int k = 0; k < SH_K; ++k) {
        const uint64_t j = base + (uint64_t)k * SH_NT + threadIdx.x;
        if (j >= N.out_cap) break;
        uint32_t v;
        if (N.left) {
            const uint32_t hi = j >= w && j - w < la ? N.a[j - w] : 0u;
            const uint32_t lo = j >= w + 1 && j - w - 1 < la ? N.a[j - w - 1] : 0u;
            v = s ? (hi << s) | (lo >> (32 - s)) : hi;
        } else {
            const uint64_t i = j + w;
            const uint32_t lo = i < la ? N.a[i] : 0u;
            const uint32_t hi = i + 1 < la ? N.a[i + 1] : 0u;
            v = s ? (lo >> s) | (hi << (32 - s)) : lo;
        }
        N.out[j] = v;
    }
    if (t == N.tile0 && threadIdx.x == 0) {
        uint64_t len = 0;
        if (la) {
            const uint32_t top = N.a[la - 1];
            if (N.left) len = la + w + ((s && (top >> (32 - s))) ? 1 : 0);
            else if (la > w) len = la - w - ((top >> s) ? 0 : 1);
        }
        *N.out_len = len < N.out_cap ? len : N.out_cap;      // a result that does not fit is truncated
    }
}

// One thread per value: nine digits per round from the right end, each
// round one pass of division by 10^9 from the top limb down (the value is
// its own scratch).
__global__ __launch_bounds__(64) void k_decimal_leaves(uint32_t* __restrict__ x, uint64_t stride,
                                                       const uint64_t* __restrict__ x_len, uint64_t n,
                                                       uint32_t width, char* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    uint32_t* v = x + i * stride;
    uint64_t len = x_len[i];
    char* o = out + i * width;
    for (uint32_t pos = width; pos >= 9; pos -= 9) {
        uint64_t r = 0;
        for (uint64_t k = len; k-- > 0;) {
            const uint64_t cur = (r << 32) | v[k];
            const uint64_t q = cur / 1000000000u;
            v[k] = (uint32_t)q;
            r = cur - q * 1000000000u;
        }
        while (len && v[len - 1] == 0) --len;
        uint32_t r32 = (uint32_t)r;
        for (int d = 1; d <= 9; ++d) {
            o[pos - d] = (char)('0' + r32 % 10u);
            r32 /= 10u;
        }
    }
}

// The same with the values staged in LDS (limb-major, thread t's limb k at
// k * LV_NT + t): the division passes then wait on LDS, not on global memory.
constexpr int LV_NT = 32;
constexpr uint64_t LV_MAX_STRIDE = 480;      // 480 x 32 x 4 B = 60 KiB of LDS

__global__ __launch_bounds__(LV_NT) void k_decimal_leaves_lds(const uint32_t* __restrict__ x, uint64_t stride,
                                                              const uint64_t* __restrict__ x_len, uint64_t n,
                                                              uint32_t width, char* __restrict__ out) {
    extern __shared__ uint32_t lv[];
    const uint64_t i = (uint64_t)blockIdx.x * LV_NT + threadIdx.x;
    if (i >= n) return;
    uint64_t len = x_len[i];
    const uint32_t* src = x + i * stride;
    for (uint64_t k = 0; k < len; ++k) lv[k * LV_NT + threadIdx.x] = src[k];
    char* o = out + i * width;
    for (uint32_t pos = width; pos >= 9; pos -= 9) {
        uint64_t r = 0;
        for (uint64_t k = len; k-- > 0;) {
            uint32_t* cell = &lv[k * LV_NT + threadIdx.x];
            const uint64_t cur = (r << 32) | *cell;
            const uint64_t q = cur / 1000000000u;
            *cell = (uint32_t)q;
            r = cur - q * 1000000000u;
        }
        while (len && lv[(len - 1) * LV_NT + threadIdx.x] == 0) --len;
        uint32_t r32 = (uint32_t)r;
        for (int d = 1; d <= 9; ++d) {
            o[pos - d] = (char)('0' + r32 % 10u);
            r32 /= 10u;
        }
    }
}

__global__ void k_fill_zero(uint32_t* __restrict__ p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

} // namespace

hipError_t launch_fill_zero(uint32_t* p, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((n + 1023) / 1024, 4096);
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
    g_launches.fetch_add(1, std::memory_order_relaxed);
    return hipGetLastError();
}

hipError_t launch_shift_batch(const ShiftNode* d_nodes, const uint32_t* d_tile_node, uint32_t ntiles, hipStream_t s,
                              const ShiftNode* one) {
    if (ntiles == 0) return hipSuccess;
    const ShiftNode single = one ? *one : ShiftNode{};
    if (one) { d_nodes = nullptr; d_tile_node = nullptr; }
    hipEvent_t t0 = timing_begin(s);
    hipLaunchKernelGGL(k_shift_batch, dim3(ntiles), dim3(SH_NT), 0, s, d_nodes, d_tile_node, single);
    g_launches.fetch_add(1, std::memory_order_relaxed);
    hipError_t e = hipGetLastError();
    timing_end("k_shift_batch", t0, s);
    return e;
}

hipError_t launch_decimal_leaves(uint32_t* x, uint64_t stride, const uint64_t* x_len, uint64_t n, uint32_t width,
                                 char* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipEvent_t t0 = timing_begin(s);
    if (stride <= LV_MAX_STRIDE)
        hipLaunchKernelGGL(k_decimal_leaves_lds, dim3((unsigned)((n + LV_NT - 1) / LV_NT)), dim3(LV_NT),
                           (size_t)stride * LV_NT * 4, s, x, stride, x_len, n, width, out);
    else
        hipLaunchKernelGGL(k_decimal_leaves, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, x, stride, x_len, n,
                           width, out);
    g_launches.fetch_add(1, std::memory_order_relaxed);
    hipError_t e = hipGetLastError();
    timing_end("k_decimal_leaves", t0, s);
    return e;
}

} // namespace pdh
