// ============================================================================
//  bbp_kernels.hip -- BBP hex-digit extraction for gfx950.
//
//  Replaces the reference's k_bbp / bbp_hex_digits_cuda (src/bbp_cuda.cu:
//  30-70) built on include/bbp_kernel.h:70-137.
//
//  Bit-exactness. The digit depends on the FP64 operation sequence, which is
//  kept exactly: per series (j, d) the head adds (double)r/(double)m for
//  k = 0..d in ascending order with sum -= floor(sum) after each add
//  (bbp_kernel.h:103-108), then the same tail loop (:112-120), then
//  4*S1 - 2*S4 - 1*S5 - 1*S6 left to right, floor, the negative fix and
//  (int)(16*s) (:128-137). This file is compiled with -ffp-contract=off so
//  none of those operations is fused.
//
//  The integer r = 16^(d-k) mod m is unique, so it may be computed by any
//  exact method. Here it is left-to-right square-and-multiply on FP64 values
//  (v_fma_f64 issues at full rate on gfx950, 64-bit integer modulo does not
//  exist in hardware): for |r| < 16m, m < 2^31,
//      p = r*r (rounded), e = fma(r, r, -p) (exact error, r*r = p + e),
//      q = rint(p / m),   r' = fma(-q, m, p) + e
//  where fma(-q, m, p) is exact because p - q*m is an integer below 2^31, so
//  r' = r*r - q*m exactly and |r'| < m. A set exponent bit multiplies by 16
//  without reduction; the next squaring reduces. The result is brought into
//  [0, m) at the end.
//
//  Layout: 4 lanes per position (one per j in {1,4,5,6}); lane 4p combines
//  the four series through cross-lane shuffles in the reference's order.
// ============================================================================
#include "pdh_internal.h"
#include <cmath>

namespace pdh {

__device__ __forceinline__ double modpow16_exact(uint64_t e, double m, double minv) {
    double r = 1.0;
    if (e) {
        const int top = 63 - __clzll((long long)e);
        for (int bit = top; bit >= 0; --bit) {
            const double p = r * r;
            const double err = fma(r, r, -p);
            const double q = __builtin_rint(p * minv);
            r = fma(-q, m, p) + err;
            if ((e >> bit) & 1) r = r * 16.0;
        }
    }
    const double q = floor(r * minv);
    r = fma(-q, m, r);
    if (r < 0.0) r += m;
    if (r >= m) r -= m;
    return r;
}

// {16^d S_j}: bbp_kernel.h:99-122 with the integer part done exactly above.
__device__ double bbp_series_dev(uint32_t j, uint64_t d) {
    double sum = 0.0;
    for (uint64_t k = 0; k <= d; ++k) {
        const uint64_t mi = 8ull * k + j;
        const double m = (double)mi;
        const double r = (mi == 1) ? 0.0 : modpow16_exact(d - k, m, 1.0 / m);
        sum += r / m;
        sum -= floor(sum);
    }
    double p16 = 1.0 / 16.0;
    uint64_t k = d + 1;
    while (p16 > 1e-18) {
        const uint64_t mi = 8ull * k + j;
        sum += p16 / (double)mi;
        p16 /= 16.0;
        ++k;
    }
    sum -= floor(sum);
    return sum;
}

__global__ __launch_bounds__(256) void k_bbp(const uint64_t* __restrict__ pos, uint64_t start,
                                             size_t count, uint8_t* __restrict__ out) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t pi = gid >> 2;
    const int q = (int)(gid & 3);
    const bool active = pi < count;
    uint64_t position = active ? (pos ? pos[pi] : start + pi) : 1;
    // Positions outside [1, 2^28] (possible only through the device entry
    // point) are not computed: their value is 0xFF.
    const bool valid = position >= 1 && position <= ((uint64_t)1 << 28);
    if (!valid) position = 1;
    const uint32_t j = q == 0 ? 1u : (uint32_t)(q + 3);   // 1, 4, 5, 6
    const double S = bbp_series_dev(j, position - 1);
    const int lane = threadIdx.x & 63;
    const int g = lane & ~3;
    const double s1 = __shfl(S, g + 0);
    const double s4 = __shfl(S, g + 1);
    const double s5 = __shfl(S, g + 2);
    const double s6 = __shfl(S, g + 3);
    if (active && q == 0 && !valid) out[pi] = 0xFF;
    if (active && q == 0 && valid) {
        double s = 4.0 * s1 - 2.0 * s4 - 1.0 * s5 - 1.0 * s6;
        s -= floor(s);
        if (s < 0.0) s += 1.0;
        out[pi] = (uint8_t)(int)(16.0 * s);
    }
}

hipError_t launch_bbp(const uint64_t* d_pos, uint64_t start, size_t count, uint8_t* d_out,
                      hipStream_t s) {
    if (count == 0) return hipSuccess;
    const size_t lanes = count * 4;
    const dim3 grid((unsigned)((lanes + 255) / 256));
    hipEvent_t t0 = timing_begin(s);
    hipLaunchKernelGGL(k_bbp, grid, dim3(256), 0, s, d_pos, start, count, d_out);
    g_launches.fetch_add(1, std::memory_order_relaxed);
    hipError_t e = hipGetLastError();
    timing_end("k_bbp", t0, s);
    return e;
}

} // namespace pdh
