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
//  exact method: below m = 2^30 by integer Montgomery arithmetic (see
//  modpow16_mont), above by left-to-right square-and-multiply on FP64 values
//  (v_fma_f64 issues at full rate on gfx950, 64-bit integer modulo does not
//  exist in hardware), exact for every m < 2^48 -- the moduli of positions
//  up to 2^45, where the reference's bbp_mulmod (include/bbp_kernel.h:70-83)
//  is exact -- with residues kept in |r| <= 0.54 m (see modpow16_exact).
//
//  Layout: 4 lanes per position (one per j in {1,4,5,6}); lane 4p combines
//  the four series through cross-lane shuffles in the reference's order.
// ============================================================================
#include "pdh_internal.h"
#include <algorithm>
#include <cmath>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

namespace pdh {

// 1/m to ~2^-80 relative: the hardware estimate (v_rcp_f64) refined by two
// Newton steps. The quotient estimates below only need 2^-40.
__device__ __forceinline__ double recip_f64(double m) {
    double r = __builtin_amdgcn_rcp(m);
    double e = fma(-m, r, 1.0);
    r = fma(r, e, r);
    e = fma(-m, r, 1.0);
    return fma(r, e, r);
}

// 16^e mod m for 1 < m < 2^48 on FP64 values, with signed residues |r| <=
// 0.54 m between steps. A square: p = r*r (rounded), e = fma(r, r, -p) (the
// exact error, r^2 = p + e, |e| <= 2^43 for r^2 < 2^96), q = rint(p / m), off
// by at most 0.54 from r^2 / m (r^2 / m <= 0.3 m < 2^47 and 1/m is correct to
// about 2^-52); fma(-q, m, p) is the exact p - q m, an integer below 0.54 m +
// 2^43 < 2^53, and adding e gives r^2 - q m exactly, again |r'| <= 0.54 m. A
// set exponent bit multiplies by 16 (exact) and subtracts rint(16 r / m) m
// the same way. The top three exponent bits start as the exact power 2^(4t).
// The result is brought into [0, m) at the end.
__device__ __forceinline__ double modpow16_exact(uint64_t e, double m, double minv) {
    double r = 1.0;
    if (e) {
        int bit = 63 - __clzll((long long)e);
        if (bit >= 2) {
            const int t = (int)(e >> (bit - 2));          // t in [4, 7]: 16^t = 2^(4t) <= 2^28
            r = ldexp(1.0, 4 * t);
            r = fma(-__builtin_rint(r * minv), m, r);
            bit -= 3;
        }
        for (; bit >= 0; --bit) {
            const double p = r * r;
            const double err = fma(r, r, -p);
            const double q = __builtin_rint(p * minv);
            r = fma(-q, m, p) + err;
            if ((e >> bit) & 1) {
                const double v = r * 16.0;
                r = fma(-__builtin_rint(v * minv), m, v);
            }
        }
    }
    const double q = floor(r * minv);
    r = fma(-q, m, r);
    if (r < 0.0) r += m;
    if (r >= m) r -= m;
    return r;
}

// The same residue by integer Montgomery arithmetic (R = 2^32) for 1 < m <
// 2^30: m = 2^s mo with mo odd (s <= 2 for m = 8k + j, j in {1, 4, 5, 6}),
// and for e >= 1, 16^e = 2^s 2^E with E = 4e - s, so 16^e mod m = 2^s (2^E mod
// mo). 2^E mod mo runs a left-to-right square-and-double ladder on lazy
// residues in [0, 2 mo): a square r^2 < 4 mo^2 < 2^32 mo reduces to [0, 2 mo)
// without a final correction, a double goes to [0, 4 mo) and one unsigned
// min brings it back. Seven VALU per exponent bit against fifteen for the
// FP64 ladder. minv = 1/m (double) gives the Montgomery one, 2^32 mod mo.
__device__ __forceinline__ uint32_t modpow16_mont(uint64_t e, uint32_t m, double minv) {
    if (e == 0) return 1u;
    const int s = __builtin_ctz(m);
    const uint32_t mo = m >> s;
    const uint32_t E = (uint32_t)(4 * e) - (uint32_t)s;
    uint32_t inv = mo;                                  // mo * mo = 1 mod 8
    for (int i = 0; i < 4; ++i) inv *= 2u - mo * inv;   // 3 -> 48 correct bits
    const uint32_t ninv = 0u - inv;
    // 2^32 mod mo from a floor estimate of 2^32 / mo (within one)
    const uint64_t q = (uint64_t)(4294967296.0 * (minv * (double)(1u << s)));
    int64_t rr = (int64_t)(((uint64_t)1 << 32) - q * mo);
    if (rr < 0) rr += mo;
    if (rr >= (int64_t)mo) rr -= mo;
    uint32_t r = (uint32_t)rr;
    const uint32_t two_mo = mo << 1;
    for (int b = 31 - __clz(E); b >= 0; --b) {
        const uint64_t t = (uint64_t)r * r;
        const uint32_t mq = (uint32_t)t * ninv;
        r = (uint32_t)(((uint64_t)mq * mo + t) >> 32);
        r <<= (E >> b) & 1u;
        const uint32_t d = r - two_mo;
        r = d < r ? d : r;
    }
    // out of Montgomery form: in [0, mo], and mo only when mo = 1
    const uint32_t mq = r * ninv;
    uint32_t y = (uint32_t)(((uint64_t)mq * mo + r) >> 32);
    if (y >= mo) y -= mo;
    return y << s;
}

// 16^e mod m for 1 < m < 2^48: Montgomery below 2^30, the FP64 ladder above.
__device__ __forceinline__ double modpow16(uint64_t e, uint64_t mi, double m, double minv) {
    return mi < ((uint64_t)1 << 30) ? (double)modpow16_mont(e, (uint32_t)mi, minv) : modpow16_exact(e, m, minv);
}

// The largest position the kernels compute: every modulus 8k + j with
// k <= position - 1 stays below 2^48.
constexpr uint64_t BBP_POS_LIMIT = (uint64_t)1 << 45;
__device__ __forceinline__ bool bbp_pos_ok(uint64_t p) { return p >= 1 && p <= BBP_POS_LIMIT; }

// {16^d S_j}: bbp_kernel.h:99-122 with the integer part done exactly above.
__device__ double bbp_series_dev(uint32_t j, uint64_t d) {
    double sum = 0.0;
    for (uint64_t k = 0; k <= d; ++k) {
        const uint64_t mi = 8ull * k + j;
        const double m = (double)mi;
        const double r = (mi == 1) ? 0.0 : modpow16(d - k, mi, m, recip_f64(m));
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

// A work queue of groups of 16 positions (one wavefront: 4 lanes each),
// handed out longest first (positions are ascending, so from the last group
// down): series lengths d + 1 vary widely, and a static split leaves some
// SIMDs with much more work than others. Every wavefront loops until the
// queue (a counter zeroed before the launch) runs past the last group.
__global__ __launch_bounds__(256) void k_bbp(const uint64_t* __restrict__ pos, uint64_t start,
                                             size_t count, uint8_t* __restrict__ out,
                                             unsigned long long* __restrict__ next) {
    const size_t ngroups = (count + 15) / 16;
    const int lane = threadIdx.x & 63;
    const int q = lane & 3;
    const int g = lane & ~3;
    for (;;) {
        unsigned long long idx = 0;
        if (lane == 0) idx = atomicAdd(next, 1ull);
        idx = __shfl(idx, 0);
        if (idx >= ngroups) break;                       // the same on every lane
        const size_t pi = (ngroups - 1 - idx) * 16 + (size_t)(lane >> 2);
        const bool active = pi < count;
        uint64_t position = active ? (pos ? pos[pi] : start + pi) : 1;
        // Positions outside [1, 2^45] (possible only through the device entry
        // point) are not computed: their value is 0xFF.
        const bool valid = bbp_pos_ok(position);
        if (!valid) position = 1;
        const uint32_t j = q == 0 ? 1u : (uint32_t)(q + 3);   // 1, 4, 5, 6
        const double S = bbp_series_dev(j, position - 1);
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
}

// ---- few positions ------------------------------------------------------------------
// With few positions, one lane per series leaves most of the GPU idle while
// each lane walks d + 1 terms, so the terms of every series are computed in
// parallel (one lane per term). The reference's sum is an ordered FP64 fold,
// which is inherently sequential; two paths handle it.
//
// Fixed-point path (default). Each head term r/m is also taken exactly to 64
// fractional bits, F = floor(r 2^64 / m), and summed with wrapping u64 adds
// (mod 2^64 is mod 1), in any order: an atomic reduction. The 14 tail terms
// 16^-i/m (the reference's loop runs while p16 > 1e-18, i.e. i = 1..14) are
// floor(2^(64-4i)/m). T = 4F1 - 2F4 - F5 - F6 (mod 2^64) then lies within
// Mu(d) of 2^64 times the reference's FP64 value s (mod 2^64), where, in
// units of 2^-64:
//   head, per term: |fl(r/m) - r/m| <= 2^-54 and the add into [0, 2) rounds
//     by <= 2^-53; sum -= floor(sum) is exact: 3 * 2^10 per term, d+1 terms;
//   tail: 14 adds (<= 2^-53) and divides (<= 2^-58): below 15 * 2^11;
//   combination 4 s1 - 2 s4 - s5 - s6: three roundings below 4, <= 3 * 2^12;
//   so |2^64 s - 2^64 s_exact| <= 8 (3072 (d+1) + 30720) + 12288, and the
//   floors of the fixed point lose below d+15 units per series, at most
//   4 (d+15) after the combination: Mu(d) = 24584 (d+1) + 262144 bounds all.
// When T is farther than 2 Mu(d) from every multiple of 2^60, the
// reference's digit floor(16 frac(s)) equals T >> 60 (both values lie
// strictly inside the same sixteenth of the circle). Otherwise the position
// is marked undecided (0xFE) and recomputed by the ordered path below. With
// d = 10^8 that happens with probability about 4e-6 per position.
//
// Ordered path (fallback, or forced by pdh_bbp_set_exact). The head sum is
// split into chunks of CH terms: k_bbp_terms computes each chunk's terms
// (double)r/(double)m in parallel, k_bbp_fold then adds them to the series'
// running sum in ascending k with sum -= floor(sum) after each add -- the
// reference's operation sequence, unchanged.
struct BbpSeries { uint64_t d; double sum; uint32_t j; uint32_t valid; };

__global__ __launch_bounds__(256) void k_bbp_init(const uint64_t* __restrict__ pos, uint64_t start, uint32_t stride,
                                                  size_t count, BbpSeries* __restrict__ st, uint64_t* __restrict__ acc) {
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= 4 * count) return;
    const size_t pi = s >> 2;
    const int q = (int)(s & 3);
    const uint64_t position = pos ? pos[pi] : start + pi * stride;
    const bool valid = bbp_pos_ok(position);
    st[s] = BbpSeries{valid ? position - 1 : 0, 0.0, q == 0 ? 1u : (uint32_t)(q + 3), valid ? 1u : 0u};
    if (acc) acc[s] = 0;
}

// floor(r 2^32 / m) and the remainder, for integers 0 <= r < m < 2^48; c32 =
// 2^32 / m to ~2^-50 relative, so the estimate (below 2^32) is within 2^-18
// of the true quotient and one correction step suffices. r 2^32 - q m is
// formed mod 2^64: its true value lies in (-m, 2m), so the wrapped
// difference read as signed is exact.
__device__ __forceinline__ uint32_t div_step(uint64_t r, uint64_t m, double c32, uint64_t& rem) {
    uint32_t q = (uint32_t)((double)r * c32);
    int64_t t = (int64_t)((r << 32) - (uint64_t)q * m);
    if (t < 0) { --q; t += (int64_t)m; } else if (t >= (int64_t)m) { ++q; t -= (int64_t)m; }
    rem = (uint64_t)t;
    return q;
}

// floor(r 2^64 / m) for integers 0 <= r < m < 2^48.
__device__ __forceinline__ uint64_t frac64(uint64_t r, uint64_t m, double minv) {
    const double c32 = minv * 4294967296.0;
    uint64_t rem;
    const uint32_t hi = div_step(r, m, c32, rem);
    const uint32_t lo = div_step(rem, m, c32, rem);
    return ((uint64_t)hi << 32) | lo;
}

// r a 2^-32 mod mo, lazy: for r, a < 2^32 with r a < 2^32 mo the result is
// below 2 mo.
__device__ __forceinline__ uint32_t mont_mul_lazy(uint32_t r, uint32_t a, uint32_t mo, uint32_t ninv) {
    const uint64_t t = (uint64_t)r * a;
    const uint32_t mq = (uint32_t)t * ninv;
    return (uint32_t)(((uint64_t)mq * mo + t) >> 32);
}

// floor((16^e mod m) 2^64 / m) for e >= 1 and 1 < m < 2^30, with no division.
// With m = 2^s mo (mo odd) and r = 16^e mod m = 2^s r_o, r_o = 2^(4e - s) mod
// mo, the fraction is F = floor(r_o 2^64 / mo), taken as two 32-bit halves:
// r_o 2^32 = F_hi mo + rem1 and rem1 2^32 = F_lo mo + rem2, so F = F_hi 2^32 +
// F_lo with rem1 = 2^(4e - s + 32) mod mo and rem2 = 2^(4e - s + 64) mod mo.
// Each half is below 2^32 and exact, so F_hi = -rem1 mo^-1 and F_lo = -rem2
// mo^-1 mod 2^32. Both remainders come from one square-and-double ladder on
// E = 4e - s + 32 run in Montgomery form (R = 2^32): it ends holding
// 2^E R = 2^(4e - s + 64) mod mo, which is rem2 up to one subtraction, and its
// Montgomery reduction is 2^E mod mo = rem1.
//
// The ladder is modpow16_mont's, for callers whose first active lane holds
// the wavefront's largest exponent (lanes taking consecutive terms k of one
// series: e = d - k falls with the lane). Its bit loop runs on that lane's
// bit count, read once into a scalar register: no per-lane counter or exit
// test, and the bit is one v_bfe. A lane with a shorter exponent squares the
// Montgomery one through its leading zero bits, which keeps its residue. The
// set-up is trimmed too: the inverse starts from (3 mo) xor 2, right to five
// bits for odd mo, so three Newton steps reach 40; and the Montgomery one
// 2^32 mod mo is taken in the lazy range [0, 2 mo) from a 32-bit quotient
// estimate q within one of floor(2^32 / mo): 2^32 - q mo, formed mod 2^32,
// lies in [-mo, 2 mo), and one add of mo fixes the negative case. The
// indices fit 32 bits: positions are at most 2^28, so m = 8k + j < 2^31.
// FOLD (every modulus below 2^29): each bit's doubling goes into its square,
// r <- (r * (r << bit)) 2^-32 (mont_mul_lazy): with r < 2 mo the product is
// below 8 mo^2 < 2^32 mo, so r stays below 2 mo -- five VALU per bit (the bit,
// the shift, the product and its reduction) against seven.
template <bool FOLD>
__device__ __forceinline__ uint64_t frac64_mont_wave(uint32_t e, uint32_t m) {
    const int s = __builtin_ctz(m);
    const uint32_t mo = m >> s;
    const uint32_t E = 4u * e - (uint32_t)s + 32u;
    uint32_t inv = (3u * mo) ^ 2u;
    for (int i = 0; i < 3; ++i) inv *= 2u - mo * inv;
    const uint32_t ninv = 0u - inv;
    // q within one of floor(2^32 / mo): a single-precision reciprocal is off by
    // at most 2^-22.4 relative, under half a unit while 2^32 / mo < 2^21;
    // smaller moduli (the first terms of a series) divide exactly
    const uint32_t q = mo >= 2048u ? (uint32_t)(4294967296.0f * __builtin_amdgcn_rcpf((float)mo))
                                   : (uint32_t)(0x100000000ull / mo);
    uint32_t r = 0u - q * mo;
    if ((int32_t)r < 0) r += mo;
    const uint32_t two_mo = mo << 1;
    const uint32_t Emax = __builtin_amdgcn_readfirstlane(E);
    for (int b = 31 - __clz(Emax); b >= 0; --b) {
        if constexpr (FOLD) {
            r = mont_mul_lazy(r, r << ((E >> b) & 1u), mo, ninv);
        } else {
            r = mont_mul_lazy(r, r, mo, ninv) << ((E >> b) & 1u);
            const uint32_t d = r - two_mo;
            r = d < r ? d : r;
        }
    }
    const uint32_t rem2 = r >= mo ? r - mo : r;
    const uint32_t mq = r * ninv;
    uint32_t rem1 = (uint32_t)(((uint64_t)mq * mo + r) >> 32);
    if (rem1 >= mo) rem1 -= mo;
    return ((uint64_t)(rem1 * ninv) << 32) | (uint64_t)(rem2 * ninv);
}

// Virtual block v: series v / bps, terms [kb * 256 * KT, (kb + 1) * 256 * KT).
// Lane t takes k = base + i * 256 + t, so a wavefront's exponents d - k are
// consecutive and its square-and-multiply loops have the same length.
// WIDE (some position above 2^28, so moduli up to 2^48 and 64-bit term
// indices): every term takes the FP64 ladder, which costs about what the
// Montgomery one does per exponent bit.
template <bool WIDE>
__global__ __launch_bounds__(256) void k_bbp_fx_terms(const BbpSeries* __restrict__ st, unsigned bps, int KT,
                                                      uint64_t* __restrict__ acc) {
    const unsigned s = blockIdx.x / bps;
    const uint64_t kb = blockIdx.x - (uint64_t)s * bps;
    const BbpSeries S = st[s];
    const uint64_t base64 = kb * 256u * (uint64_t)KT;
    if (!S.valid || base64 > S.d) return;
    if constexpr (WIDE) {
        uint64_t local = 0;
        for (int i = 0; i < KT; ++i) {
            const uint64_t k = base64 + (uint64_t)i * 256u + threadIdx.x;
            if (k > S.d) break;
            const uint64_t mi = 8ull * k + S.j;
            if (mi == 1) continue;                    // 16^d mod 1 = 0
            const double m = (double)mi;
            const double minv = recip_f64(m);
            const uint64_t r = k == S.d ? 1u : (uint64_t)modpow16_exact(S.d - k, m, minv);
            local += frac64(r, mi, minv);
        }
        for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o);
        __shared__ uint64_t wpart[4];
        if ((threadIdx.x & 63) == 0) wpart[threadIdx.x >> 6] = local;
        __syncthreads();
        if (threadIdx.x == 0)
            atomicAdd((unsigned long long*)&acc[s], (unsigned long long)(wpart[0] + wpart[1] + wpart[2] + wpart[3]));
        return;
    }
    // positions are at most 2^28 (the entry points check), so d, k < 2^28 and
    // m = 8k + j < 2^31: the term indices run in 32 bits
    const uint32_t d = (uint32_t)S.d, base = (uint32_t)base64;
    uint64_t local = 0;
    for (int i = 0; i < KT; ++i) {
        const uint32_t k = base + (uint32_t)i * 256u + threadIdx.x;
        if (k > d) break;
        // every modulus of this step, 8k + j <= 8 (kb 256 KT + 256 i + 255) + 7,
        // is below 2^29 (the FOLD ladder) when that k is at most 2^26 - 2
        const bool fold = base + (uint32_t)i * 256u + 255u <= (1u << 26) - 2u;
        const uint32_t mi = 8u * k + S.j;
        if (mi == 1) continue;                        // 16^d mod 1 = 0
        if (mi < (1u << 30) && k < d) {
            local += fold ? frac64_mont_wave<true>(d - k, mi) : frac64_mont_wave<false>(d - k, mi);
        } else {                                      // 16^0 = 1, or m >= 2^30
            const double m = (double)mi;
            const double minv = recip_f64(m);
            const uint32_t r = k == d ? 1u : (uint32_t)modpow16_exact(d - k, m, minv);
            local += frac64(r, mi, minv);
        }
    }
    for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o);
    __shared__ uint64_t part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = local;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicAdd((unsigned long long*)&acc[s], (unsigned long long)(part[0] + part[1] + part[2] + part[3]));
}

// The same fraction when the whole wavefront shares the exponent e (lanes
// hold different positions, so different moduli) and s (one series j per
// wavefront): E = 4e - s + 32 is then wave-uniform, its bits are scalar, and
// a zero bit skips the doubling by a scalar branch -- three VALU for a zero
// bit and six for a one, against seven for every bit of the per-lane ladder.
// Otherwise step for step frac64_mont_wave.
__device__ __forceinline__ uint32_t mont_sq_lazy(uint32_t r, uint32_t mo, uint32_t ninv) {
    const uint64_t t = (uint64_t)r * r;
    const uint32_t mq = (uint32_t)t * ninv;
    return (uint32_t)(((uint64_t)mq * mo + t) >> 32);
}

__device__ __forceinline__ uint32_t dbl_lazy(uint32_t r, uint32_t two_mo) {
    // an empty volatile statement keeps the caller's test a scalar branch:
    // if-converted (both values computed, one selected) a bit would cost
    // eight VALU
    __asm__ volatile("" ::: "memory");
    r <<= 1;
    const uint32_t d = r - two_mo;
    return d < r ? d : r;
}

// FOLD: every modulus of the wavefront has mo < 2^29. Each bit's doubling
// then goes into its square, r <- (r * (r << b)) 2^-32: with r < 2 mo the
// product is below 8 mo^2 < 2^32 mo, so the result stays below 2 mo, at four
// VALU per bit (shift by the scalar bit, multiply, Montgomery reduction) and
// no branch.
template <bool FOLD>
__device__ __forceinline__ uint64_t frac64_mont_uniform(uint32_t E, uint32_t m, uint32_t s) {
    const uint32_t mo = m >> s;
    uint32_t inv = (3u * mo) ^ 2u;
    for (int i = 0; i < 3; ++i) inv *= 2u - mo * inv;
    const uint32_t ninv = 0u - inv;
    const uint32_t q = mo >= 2048u ? (uint32_t)(4294967296.0f * __builtin_amdgcn_rcpf((float)mo))
                                   : (uint32_t)(0x100000000ull / mo);
    uint32_t r = 0u - q * mo;
    if ((int32_t)r < 0) r += mo;
    const uint32_t two_mo = mo << 1;
    // E >= 34. Its leading one squares the Montgomery one (no change) and
    // doubles it; the n bits below it, left-aligned in x, go four per trip
    // (fewer scalar instructions per bit), the last n mod 4 one per trip.
    int n = 31 - __clz(E);
    uint32_t x = E << (32 - n);
    if constexpr (FOLD) {
        r <<= 1;                                          // the leading one, reduced to [0, 2 mo)
        const uint32_t d = r - two_mo;
        r = d < r ? d : r;
        for (; n >= 4; n -= 4, x <<= 4) {
            r = mont_mul_lazy(r, r << (x >> 31), mo, ninv);
            r = mont_mul_lazy(r, r << ((x >> 30) & 1u), mo, ninv);
            r = mont_mul_lazy(r, r << ((x >> 29) & 1u), mo, ninv);
            r = mont_mul_lazy(r, r << ((x >> 28) & 1u), mo, ninv);
        }
        for (; n > 0; --n, x <<= 1) r = mont_mul_lazy(r, r << (x >> 31), mo, ninv);
    } else {
        r = dbl_lazy(r, two_mo);
        for (; n >= 4; n -= 4, x <<= 4) {
            r = mont_sq_lazy(r, mo, ninv);
            if ((int32_t)x < 0) r = dbl_lazy(r, two_mo);
            r = mont_sq_lazy(r, mo, ninv);
            if ((int32_t)(x << 1) < 0) r = dbl_lazy(r, two_mo);
            r = mont_sq_lazy(r, mo, ninv);
            if ((int32_t)(x << 2) < 0) r = dbl_lazy(r, two_mo);
            r = mont_sq_lazy(r, mo, ninv);
            if ((int32_t)(x << 3) < 0) r = dbl_lazy(r, two_mo);
        }
        for (; n > 0; --n, x <<= 1) {
            r = mont_sq_lazy(r, mo, ninv);
            if ((int32_t)x < 0) r = dbl_lazy(r, two_mo);
        }
    }
    const uint32_t rem2 = r >= mo ? r - mo : r;
    const uint32_t mq = r * ninv;
    uint32_t rem1 = (uint32_t)(((uint64_t)mq * mo + r) >> 32);
    if (rem1 >= mo) rem1 -= mo;
    return ((uint64_t)(rem1 * ninv) << 32) | (uint64_t)(rem2 * ninv);
}

// Positions in lanes: block = 64 consecutive positions (in the order given)
// x 4 series, one wavefront per series j; block v covers group v / chunks and
// the exponents e of chunk v % chunks, [c EK, (c + 1) EK). For each e every
// lane takes its own position's term k = d - e (none when e > d), so the
// wavefront's exponent is uniform. Each lane sums its terms in a register and
// adds them to its series' accumulator once. Suited to positions that lie
// close together within a group (a lane idles for e above its d).
__global__ __launch_bounds__(256) void k_bbp_fx_lanes(const BbpSeries* __restrict__ st, size_t count,
                                                      uint32_t chunks, uint32_t EK, uint64_t* __restrict__ acc) {
    const uint32_t g = blockIdx.x / chunks;
    const uint32_t c = blockIdx.x - g * chunks;
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    const size_t pi = (size_t)g * 64 + lane;
    const bool in = pi < count;
    const BbpSeries S = in ? st[4 * pi + w] : BbpSeries{0, 0.0, 1, 0};
    const bool act = in && S.valid;
    // positions are at most 2^28, so d, e, k < 2^28 and m = 8k + j < 2^31
    const uint32_t d = act ? (uint32_t)S.d : 0;
    uint32_t dm = d;
    for (int o = 32; o > 0; o >>= 1) dm = max(dm, (uint32_t)__shfl_xor((int)dm, o));
    dm = __builtin_amdgcn_readfirstlane(dm);
    const uint32_t e0 = c * EK;
    if (e0 > dm) return;                                  // the same in all four wavefronts
    const uint32_t e1 = min(e0 + EK - 1u, dm);
    const uint32_t j = w == 0 ? 1u : w + 3u;              // 1, 4, 5, 6
    const uint32_t sj = j == 4 ? 2u : (j == 6 ? 1u : 0u); // 2-adic valuation of 8k + j
    uint64_t local = 0;
    // At exponent e every modulus of the wavefront is 8k + j <= 8 (dm - e) + 6:
    // below 2^29 (the FOLD ladder) from e_fold = dm - (2^26 - 2) on. The two
    // ranges of e are wave-uniform, so each loop runs one ladder.
    const uint32_t e_fold = dm > (1u << 26) - 2u ? dm - ((1u << 26) - 2u) : 0u;
    const uint32_t e_mid = min(max(e_fold, e0), e1 + 1u);
    for (uint32_t e = e0; e < e_mid; ++e) {
        if (!act || e > d) continue;
        const uint32_t k = d - e;
        const uint32_t mi = 8u * k + j;
        if (mi == 1) continue;                            // 16^d mod 1 = 0
        if (mi < (1u << 30) && e > 0) {
            local += frac64_mont_uniform<false>(4u * e - sj + 32u, mi, sj);
        } else {                                          // 16^0 = 1, or m >= 2^30
            const double m = (double)mi;
            const double minv = recip_f64(m);
            const uint32_t r = e == 0 ? 1u : (uint32_t)modpow16_exact(e, m, minv);
            local += frac64(r, mi, minv);
        }
    }
    for (uint32_t e = e_mid; e <= e1; ++e) {
        if (!act || e > d) continue;
        const uint32_t k = d - e;
        const uint32_t mi = 8u * k + j;
        if (mi == 1) continue;                            // 16^d mod 1 = 0
        if (e > 0) {
            local += frac64_mont_uniform<true>(4u * e - sj + 32u, mi, sj);
        } else {                                          // 16^0 = 1
            const double m = (double)mi;
            local += frac64(1u, mi, recip_f64(m));
        }
    }
    if (act) atomicAdd((unsigned long long*)&acc[4 * pi + w], (unsigned long long)local);
}

// ---- moduli in lanes -----------------------------------------------------------------
// A wavefront holds 64 consecutive terms k = k0 + lane of one series j, that
// is 64 fixed moduli m = 8k + j, and steps through the positions of its
// group of 64 in ascending order. With w the Montgomery form of
// 16^e 2^-s mod mo (e = d - k, m = 2^s mo), w is rem1 = 2^(4e - s + 32) mod mo
// of frac64_mont_wave, and Montgomery multiplying it by the form of 2^32
// (16^8) gives rem2; the fraction then costs that product, two unsigned
// minima and two multiplies (nine VALU). Going from position d to the next,
// d', multiplies w by 16^(d' - d): the gap is wave-uniform, so its two low
// nibbles index the lane's tables of 16^n and 16^(16 n) (n < 16) -- two
// products below a gap of 2^8 -- and its higher part takes a uniform power
// of 16^256. Each lane's own ladder (the long one, e ~ d) runs once, at the
// first position it reaches, and the set-up once per 64 terms. The term
// with e = 0 (k = d) is the one integer 16^0 = 1 that the 2^-s form does
// not give, and goes through frac64. Moduli stay below 2^30 (positions
// up to 2^27 - 1), so every lazy product of residues below 2 mo is below
// 4 mo^2 < 2^32 mo and reduces to [0, 2 mo).
constexpr uint32_t BBP_MODULI_DMAX = (uint32_t)BBP_MODULI_DMAX_HOST;   // largest d = position - 1 this layout takes
// Residues (blocks of 64 terms) per lane. Two share each position step's
// scalar work and staging but take 152 VGPRs (three waves per SIMD): C4
// chunk 32 took 2.94 s against 2.61 s with one (5.30 s held to four waves,
// where the tables spill).
constexpr uint32_t BBP_MODULI_NR = 1;

__device__ __forceinline__ uint32_t dbl_red(uint32_t r, uint32_t two_mo) {
    r <<= 1;
    const uint32_t d = r - two_mo;
    return d < r ? d : r;
}

// x / 2 mod mo for x in [0, 2 mo), mo odd: the result is below 3 mo / 2.
__device__ __forceinline__ uint32_t half_mod(uint32_t x, uint32_t mo) { return (x & 1u) ? (x + mo) >> 1 : x >> 1; }

// Montgomery form of 2^E for a per-lane E whose largest value sits in the
// first active lane (lanes with a shorter E square the form of one through
// their leading zero bits, which keeps it).
template <bool FOLD>
__device__ __forceinline__ uint32_t pow2_lane(uint32_t E, uint32_t r1, uint32_t mo, uint32_t ninv, uint32_t two_mo) {
    const uint32_t Emax = __builtin_amdgcn_readfirstlane(E);
    uint32_t r = r1;
    for (int b = 31 - __clz(Emax); b >= 0; --b) {
        if constexpr (FOLD) {
            r = mont_mul_lazy(r, r << ((E >> b) & 1u), mo, ninv);
        } else {
            r = mont_mul_lazy(r, r, mo, ninv) << ((E >> b) & 1u);
            const uint32_t d = r - two_mo;
            r = d < r ? d : r;
        }
    }
    return r;
}

__device__ __forceinline__ void bbp_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum the batch of eight staged slots over the 64 moduli lanes: lane (r, part)
// reads slot r at lanes part + 8 t (rows of 72 u64 put each half-wavefront's
// 32 reads in 64 distinct banks), three shuffles finish the sum, and sorted
// position 8 b + r (batch b) is kept by lane 8 r + b.
__device__ __forceinline__ void bbp_moduli_flush(const uint64_t* __restrict__ buf, uint32_t batch, uint32_t lane,
                                                 uint64_t& acc_l) {
    bbp_wave_sync();
    const uint32_t r = lane >> 3, part = lane & 7u;
    uint64_t sum = 0;
#pragma unroll
    for (int t = 0; t < 8; ++t) sum += buf[r * 72u + part + 8u * t];
    sum += __shfl_xor(sum, 1);
    sum += __shfl_xor(sum, 2);
    sum += __shfl_xor(sum, 4);
    if (part == batch) acc_l += sum;
    bbp_wave_sync();
}

// w *= form of 16^gp for a wave-uniform gap gp: gp = n0 + 16 n1 + 256 r
// takes the lane's entries U0[n0] (form of 16^n0) and U1[n1] (16^(16 n1)),
// read by the uniform index (entry 0 is the form of one: no test, most gaps
// have both nibbles), and V^r (V = form of 16^256) by a uniform
// square-and-multiply.
__device__ __forceinline__ uint32_t bbp_gap_mul(uint32_t w, uint32_t gp, const uint32_t* U0, const uint32_t* U1,
                                                uint32_t V, uint32_t mo, uint32_t ninv) {
    const uint32_t n0 = gp & 15u, n1 = (gp >> 4) & 15u, r = gp >> 8;
    w = mont_mul_lazy(w, U0[n0], mo, ninv);
    w = mont_mul_lazy(w, U1[n1], mo, ninv);
    if (r) {
        uint32_t p = V;
        for (int b = 30 - __clz(r); b >= 0; --b) {
            p = mont_mul_lazy(p, p, mo, ninv);
            if ((r >> b) & 1u) p = mont_mul_lazy(p, V, mo, ninv);
        }
        w = mont_mul_lazy(w, p, mo, ninv);
    }
    return w;
}

// F = floor((16^e mod m) 2^64 / m) for e >= 1 from w (0 for w = 0). For
// x < 2 mo, min(x, x - mo) (unsigned) is x mod mo: two VALU.
__device__ __forceinline__ uint64_t bbp_moduli_frac(uint32_t w, uint32_t T3, uint32_t mo, uint32_t ninv) {
    const uint32_t rem1 = min(w, w - mo);
    const uint32_t t = mont_mul_lazy(w, T3, mo, ninv);
    const uint32_t rem2 = min(t, t - mo);
    return ((uint64_t)(rem1 * ninv) << 32) | (uint64_t)(rem2 * ninv);
}

// One modulus of a lane: its set-up and running residue. Its tables U0[n]
// (form of 16^n) and U1[n] (16^(16 n)) are plain local arrays beside it: an
// array member read at a run-time index would keep the whole struct in
// scratch memory.
struct ModLane {
    uint32_t k, mi, mo, ninv, r1, two_mo, V, T8, h, w;
    bool live, started;
};

__device__ __forceinline__ void mod_lane_init(ModLane& L, uint32_t* U0, uint32_t* U1, uint32_t k, uint32_t j,
                                              uint32_t sj, uint32_t dmax) {
    L.k = k;
    L.mi = 8u * k + j;
    L.live = k <= dmax && L.mi != 1u;
    const uint32_t mo = L.live ? L.mi >> sj : 3u;
    L.mo = mo;
    uint32_t inv = (3u * mo) ^ 2u;
    for (int i = 0; i < 3; ++i) inv *= 2u - mo * inv;
    L.ninv = 0u - inv;
    const uint32_t q = mo >= 2048u ? (uint32_t)(4294967296.0f * __builtin_amdgcn_rcpf((float)mo))
                                   : (uint32_t)(0x100000000ull / mo);
    uint32_t r1 = 0u - q * mo;                       // 2^32 mod mo in [0, 2 mo)
    if ((int32_t)r1 < 0) r1 += mo;
    L.r1 = r1;
    L.two_mo = mo << 1;
    U0[0] = r1;
    U0[1] = dbl_red(dbl_red(dbl_red(dbl_red(r1, L.two_mo), L.two_mo), L.two_mo), L.two_mo);
#pragma unroll
    for (int t = 2; t < 16; ++t) U0[t] = mont_mul_lazy(U0[t - 1], U0[1], mo, L.ninv);
    U1[0] = r1;
    U1[1] = mont_mul_lazy(U0[15], U0[1], mo, L.ninv);
#pragma unroll
    for (int t = 2; t < 16; ++t) U1[t] = mont_mul_lazy(U1[t - 1], U1[1], mo, L.ninv);
    L.V = mont_mul_lazy(U1[15], U1[1], mo, L.ninv);  // 16^256
    L.T8 = U0[8];                                    // 16^8 = 2^32: rem2 from rem1
    uint32_t h = r1;                                 // form of 2^-s
    if (sj >= 1u) h = half_mod(h, mo);
    if (sj >= 2u) h = half_mod(h, mo);
    L.h = h;
    L.w = 0;
    L.started = false;
}

// w = form of 16^(d - k) 2^-s for a lane reaching position d >= k.
template <bool FOLD>
__device__ __forceinline__ void mod_lane_start(ModLane& L, uint32_t d) {
    const uint32_t e = d - L.k;
    L.w = e == 0u ? L.h : mont_mul_lazy(pow2_lane<FOLD>(4u * e, L.r1, L.mo, L.ninv, L.two_mo), L.h, L.mo, L.ninv);
    L.started = true;
}

// NR blocks of 64 terms at once, k0 .. k0 + 64 NR - 1 (lane l holds
// k0 + l + 64 r, r < NR), of a wavefront over its sorted positions
// key[0 .. nv) (key held by lane i), with FOLD for every modulus below 2^29.
// The NR residues of a lane share each position step (the gap, its scalar
// work, the staging and the sum), and their fractions are added before the
// staging. Positions up to the largest live k (the triangle, only in blocks
// that reach into the group) start residues one by one and hold the e = 0
// terms; from the first position above it every live residue has started
// and e >= 1, and a residue with no terms keeps w = 0 (so F = 0) through the
// gap products. Each position's sum over the lanes goes to the lane
// p(l) = 8 (l & 7) + (l >> 3) holds (acc_l).
template <bool FOLD, int NR>
__device__ __forceinline__ void bbp_moduli_kblock(uint32_t k0, uint32_t j, uint32_t sj, uint32_t key, uint32_t nv,
                                                  uint32_t dmax, uint32_t lane, uint64_t* __restrict__ buf,
                                                  uint64_t& acc_l) {
    static_assert(NR == 1 || NR == 2, "one or two residues per lane");
    // two named lanes (not an array of them), so that each table stays a
    // register array read at a uniform index
    ModLane La, Lb;
    uint32_t U0a[16], U1a[16], U0b[16], U1b[16];
    auto each = [&](auto&& f) {
        f(La, U0a, U1a);
        if constexpr (NR == 2) f(Lb, U0b, U1b);
    };
    mod_lane_init(La, U0a, U1a, k0 + lane, j, sj, dmax);
    if constexpr (NR == 2) mod_lane_init(Lb, U0b, U1b, k0 + lane + 64u, j, sj, dmax);
    // the first position these blocks reach (the smallest d >= k0), and the
    // first above their largest live k
    const uint32_t kmax = min(k0 + 64u * NR - 1u, dmax);
    const uint32_t first = (uint32_t)__builtin_ctzll(__ballot(lane < nv && key >= k0));
    const uint64_t above = __ballot(lane < nv && key > kmax);
    const uint32_t i1 = above ? (uint32_t)__builtin_ctzll(above) : nv;
    // slots of the first batch before it hold zero
    for (uint32_t sl = first & ~7u; sl < first; ++sl) buf[(sl & 7u) * 72u + lane] = 0;
    uint32_t dprev = 0;
    uint32_t i = first;
    for (; i < i1; ++i) {                            // the triangle
        const uint32_t di = (uint32_t)__builtin_amdgcn_readlane((int)key, (int)i);
        uint64_t F = 0;
        each([&](ModLane& M, const uint32_t* U0, const uint32_t* U1) {
            if (i > first && di != dprev) M.w = bbp_gap_mul(M.w, di - dprev, U0, U1, M.V, M.mo, M.ninv);
            if (M.live && !M.started && M.k <= di) mod_lane_start<FOLD>(M, di);
            if (M.started)
                F += M.k == di ? frac64(1u, M.mi, recip_f64((double)M.mi)) : bbp_moduli_frac(M.w, M.T8, M.mo, M.ninv);
        });
        dprev = di;
        buf[(i & 7u) * 72u + lane] = F;
        if ((i & 7u) == 7u) bbp_moduli_flush(buf, i >> 3, lane, acc_l);
    }
    if (i < nv) {
        // position i1: the residues started in the triangle take the gap,
        // the others start here (e = d - k >= 1)
        const uint32_t di = (uint32_t)__builtin_amdgcn_readlane((int)key, (int)i);
        each([&](ModLane& M, const uint32_t* U0, const uint32_t* U1) {
            if (i > first && di != dprev) M.w = bbp_gap_mul(M.w, di - dprev, U0, U1, M.V, M.mo, M.ninv);
            if (M.live && !M.started) mod_lane_start<FOLD>(M, di);
        });
        dprev = di;
        for (;;) {
            uint64_t F = 0;
            each([&](ModLane& M, const uint32_t*, const uint32_t*) { F += bbp_moduli_frac(M.w, M.T8, M.mo, M.ninv); });
            buf[(i & 7u) * 72u + lane] = F;
            if ((i & 7u) == 7u) bbp_moduli_flush(buf, i >> 3, lane, acc_l);
            if (++i >= nv) break;
            const uint32_t dn = (uint32_t)__builtin_amdgcn_readlane((int)key, (int)i);
            each([&](ModLane& M, const uint32_t* U0, const uint32_t* U1) {
                M.w = bbp_gap_mul(M.w, dn - dprev, U0, U1, M.V, M.mo, M.ninv);
            });
            dprev = dn;
        }
    }
    // the last batch, when it is partial
    if (nv & 7u) {
        for (uint32_t z = nv & 7u; z < 8u; ++z) buf[z * 72u + lane] = 0;
        bbp_moduli_flush(buf, (nv - 1u) >> 3, lane, acc_l);
    }
}

// Block = group g of 64 consecutive positions (in the order given) x 4
// series, one wavefront per series j; chunk c of the block takes the
// k-blocks [c KB, (c + 1) KB). Each wavefront sorts its group's positions
// in registers (bitonic, ascending, invalid ones last) and adds each
// position's sum over its k-blocks with one atomic.
__global__ __launch_bounds__(256) void k_bbp_fx_moduli(const BbpSeries* __restrict__ st, size_t count,
                                                       uint32_t chunks, uint32_t KB, uint64_t* __restrict__ acc) {
    __shared__ uint64_t sbuf[4][8 * 72];
    const uint32_t g = blockIdx.x / chunks;
    const uint32_t c = blockIdx.x - g * chunks;
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    const size_t pi = (size_t)g * 64 + lane;
    const bool in = pi < count;
    const BbpSeries S = in ? st[4 * pi + w] : BbpSeries{0, 0.0, 1, 0};
    const bool act = in && S.valid;
    uint32_t key = act ? (uint32_t)S.d : 0xFFFFFFFFu;    // d <= 2^27 - 2 here
    uint32_t idx = lane;
    for (uint32_t kk = 2; kk <= 64; kk <<= 1) {
        for (uint32_t jj = kk >> 1; jj > 0; jj >>= 1) {
            const uint32_t ok = (uint32_t)__shfl_xor((int)key, (int)jj), oi = (uint32_t)__shfl_xor((int)idx, (int)jj);
            const bool up = (lane & kk) == 0, low = (lane & jj) == 0;
            const bool less = ok < key || (ok == key && oi < idx);
            const bool more = ok > key || (ok == key && oi > idx);
            if (low == up ? less : more) { key = ok; idx = oi; }
        }
    }
    const uint32_t nv = (uint32_t)__popcll(__ballot(act));
    if (nv == 0) return;                                  // the same in all four wavefronts
    const uint32_t dmax = (uint32_t)__builtin_amdgcn_readlane((int)key, (int)(nv - 1));
    const uint32_t j = w == 0 ? 1u : w + 3u;              // 1, 4, 5, 6
    const uint32_t sj = j == 4 ? 2u : (j == 6 ? 1u : 0u);
    uint64_t* buf = sbuf[w];
    uint64_t acc_l = 0;
    // KB is a multiple of BBP_MODULI_NR: each step takes NR blocks of the chunk
    const uint32_t kb_end = dmax / 64u + 1u;
    for (uint32_t kb = c * KB; kb < min((c + 1u) * KB, kb_end); kb += BBP_MODULI_NR) {
        const uint32_t k0 = kb * 64u;
        // the largest modulus, 8 min(k0 + 64 NR - 1, dmax) + 6, below 2^29
        if (min(k0 + 64u * BBP_MODULI_NR - 1u, dmax) <= (1u << 26) - 2u)
            bbp_moduli_kblock<true, BBP_MODULI_NR>(k0, j, sj, key, nv, dmax, lane, buf, acc_l);
        else
            bbp_moduli_kblock<false, BBP_MODULI_NR>(k0, j, sj, key, nv, dmax, lane, buf, acc_l);
    }
    const uint32_t p = 8u * (lane & 7u) + (lane >> 3);
    const uint32_t orig = (uint32_t)__shfl((int)idx, (int)p);
    if (p < nv && acc_l) atomicAdd((unsigned long long*)&acc[4 * ((size_t)g * 64 + orig) + w], (unsigned long long)acc_l);
}

// Per group of 64 consecutive entries of pos: out[0] = max(d) (atomic max),
// out[1] += sum of d + 1 over valid positions, out[2] += 64 (max d of the
// group + 1): out[1] / out[2] is the share of lanes k_bbp_fx_lanes keeps busy.
__global__ void k_bbp_group_scan(const uint64_t* __restrict__ pos, size_t count,
                                 unsigned long long* __restrict__ out) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g * 64 >= count) return;
    unsigned long long mx = 0, sum = 0;
    bool any = false;
    const size_t end = min(count, g * 64 + 64);
    for (size_t i = g * 64; i < end; ++i) {
        const uint64_t p = pos[i];
        if (bbp_pos_ok(p)) {
            any = true;
            sum += p;
            if (p - 1 > mx) mx = p - 1;
        }
    }
    if (!any) return;
    atomicMax(&out[0], mx);
    atomicAdd(&out[1], sum);
    atomicAdd(&out[2], 64ull * (mx + 1));
}

// T = 4 F1 - 2 F4 - F5 - F6 (mod 2^64) of one position: its four head sums
// plus the 14 tail terms floor(2^(64-4i) / m).
__device__ __forceinline__ uint64_t bbp_fx_combine(const BbpSeries* P, const uint64_t* a, uint64_t d) {
    uint64_t F[4];
    for (int q = 0; q < 4; ++q) {
        uint64_t f = a[q];
        for (int i = 1; i <= 14; ++i) f += ((uint64_t)1 << (64 - 4 * i)) / (8ull * (d + i) + P[q].j);
        F[q] = f;
    }
    return 4 * F[0] - 2 * F[1] - F[2] - F[3];
}

__device__ __forceinline__ uint64_t bbp_mu(uint64_t d) { return 24584ull * (d + 1) + 262144ull; }

// Distance of a 64-bit fraction from the nearest digit boundary (a multiple of 2^60).
__device__ __forceinline__ uint64_t bbp_edge_dist(uint64_t T) {
    const uint64_t cell = (uint64_t)1 << 60;
    const uint64_t x = T & (cell - 1);
    return x < cell - x ? x : cell - x;
}

__global__ __launch_bounds__(256) void k_bbp_fx_finish(const BbpSeries* __restrict__ st, size_t count,
                                                       const uint64_t* __restrict__ acc,
                                                       uint8_t* __restrict__ out, int odd_undecided) {
    const size_t pi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= count) return;
    const BbpSeries* P = st + 4 * pi;
    if (!P[0].valid) { out[pi] = 0xFF; return; }
    const uint64_t d = P[0].d;
    const uint64_t T = bbp_fx_combine(P, acc + 4 * pi, d);
    const bool decided = bbp_edge_dist(T) > 2 * bbp_mu(d) && !(odd_undecided && (pi & 1));
    out[pi] = decided ? (uint8_t)(T >> 60) : (uint8_t)0xFE;
}

// Digit blocks of a contiguous range. Anchor t sits at position start + G t
// (d = position - 1) and gives the digits of the G positions from there on.
// With V(d) = frac(16^d pi), V(d + i) = frac(16^i V(d)), so 16^i T mod 2^64
// approximates 2^64 V(d + i) within 16^i e1, where e1 = 4 (d + 15) + 2
// bounds |T - 2^64 V(d)|: the fixed-point floors lose under 4 (d + 15) units
// after the combination and the series' terms beyond the 14 tail terms add
// under 2 (8 sum_{i >= 15} 16^(16 - i) / (8 (d + i)) < 18 / (d + 15)). The
// reference's FP64 value at d + i is within Mu(d + i) of the same 2^64
// V(d + i) (Mu already holds the floors' 4 (d + 15) and these 2 units beside
// its rounding terms). So when 16^i T mod 2^64 is farther than 16^i e1 +
// Mu(d + i) from every digit boundary, the reference's digit at d + i is its
// top four bits. Digit 0 keeps the per-position rule (2 Mu(d)). The host picks
// G so that 16^(G-1) e1 stays below 2^46: e1 <= 2^31 and Mu < 2^43 keep every
// margin below 2^60, with no overflow.
__global__ __launch_bounds__(256) void k_bbp_fx_block(const BbpSeries* __restrict__ st, size_t anchors, uint32_t G,
                                                      size_t nout, const uint64_t* __restrict__ acc,
                                                      uint8_t* __restrict__ out, int odd_undecided) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= anchors) return;
    const BbpSeries* P = st + 4 * t;
    const uint64_t d = P[0].d;
    const uint64_t T = P[0].valid ? bbp_fx_combine(P, acc + 4 * t, d) : 0;
    const uint64_t e1 = 4 * (d + 15) + 2;
    for (uint32_t i = 0; i < G; ++i) {
        const size_t idx = t * G + i;
        if (idx >= nout) break;
        if (!P[0].valid) { out[idx] = 0xFF; continue; }
        const uint64_t Ti = T << (4 * i);
        const uint64_t margin = i == 0 ? 2 * bbp_mu(d) : (e1 << (4 * i)) + bbp_mu(d + i);
        const bool decided = bbp_edge_dist(Ti) > margin && !(odd_undecided && (idx & 1));
        out[idx] = decided ? (uint8_t)(Ti >> 60) : (uint8_t)0xFE;
    }
}

__global__ __launch_bounds__(256) void k_bbp_terms(const BbpSeries* __restrict__ st, size_t nseries, uint64_t k0,
                                                   size_t lanes, double* __restrict__ buf) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= lanes) return;
    const size_t i = g / nseries, s = g - i * nseries;        // term-major: buf[i * nseries + s]
    const BbpSeries S = st[s];
    const uint64_t k = k0 + i;
    if (!S.valid || k > S.d) return;
    const uint64_t mi = 8ull * k + S.j;
    const double m = (double)mi;
    const double r = (mi == 1) ? 0.0 : modpow16(S.d - k, mi, m, recip_f64(m));
    buf[g] = r / m;
}

__global__ __launch_bounds__(64) void k_bbp_fold(BbpSeries* __restrict__ st, size_t nseries, uint64_t k0,
                                                 uint64_t ch, const double* __restrict__ buf) {
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseries) return;
    BbpSeries S = st[s];
    if (!S.valid || k0 > S.d) return;
    const uint64_t n = (S.d + 1 - k0) < ch ? (S.d + 1 - k0) : ch;
    const double* __restrict__ t = buf + s;
    double sum = S.sum;
    uint64_t i = 0;
    for (; i + 8 <= n; i += 8) {                      // eight loads in flight, then the ordered adds
        double v[8];
        for (int u = 0; u < 8; ++u) v[u] = t[(i + u) * nseries];
        for (int u = 0; u < 8; ++u) { sum += v[u]; sum -= floor(sum); }
    }
    for (; i < n; ++i) { sum += t[i * nseries]; sum -= floor(sum); }
    st[s].sum = sum;
}

__global__ __launch_bounds__(256) void k_bbp_finish(const BbpSeries* __restrict__ st, size_t count,
                                                    uint8_t* __restrict__ out) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t pi = gid >> 2;
    const bool active = pi < count;
    double S = 0.0;
    BbpSeries me{0, 0.0, 1, 0};
    if (active) {
        me = st[gid];
        double sum = me.sum;
        double p16 = 1.0 / 16.0;                 // tail, bbp_kernel.h:112-120
        uint64_t k = me.d + 1;
        while (p16 > 1e-18) {
            sum += p16 / (double)(8ull * k + me.j);
            p16 /= 16.0;
            ++k;
        }
        sum -= floor(sum);
        S = sum;
    }
    const int lane = threadIdx.x & 63;
    const int g = lane & ~3;
    const double s1 = __shfl(S, g + 0);
    const double s4 = __shfl(S, g + 1);
    const double s5 = __shfl(S, g + 2);
    const double s6 = __shfl(S, g + 3);
    if (active && (gid & 3) == 0) {
        if (!me.valid) { out[pi] = 0xFF; return; }
        double s = 4.0 * s1 - 2.0 * s4 - 1.0 * s5 - 1.0 * s6;
        s -= floor(s);
        if (s < 0.0) s += 1.0;
        out[pi] = (uint8_t)(int)(16.0 * s);
    }
}

static int chunk_log(size_t S) {
    int log_ch = 15;
    while (log_ch > 8 && (S << log_ch) > ((size_t)1 << 24)) --log_ch;
    return log_ch;
}

size_t bbp_few_ws_bytes(size_t count) {
    const size_t S = 4 * count;
    const size_t head = (S * sizeof(BbpSeries) + 255) & ~(size_t)255;
    const size_t acc = (S * 8 + 255) & ~(size_t)255;
    const size_t buf = (S << chunk_log(S)) * sizeof(double);
    return head + (acc > buf ? acc : buf);
}

#define BBP_LAUNCH(name, s, ...) do { hipEvent_t _t0 = timing_begin(s); hipLaunchKernelGGL(__VA_ARGS__); \
    g_launches.fetch_add(1, std::memory_order_relaxed); hipError_t _e = hipGetLastError(); \
    timing_end(name, _t0, s); if (_e != hipSuccess) return _e; } while (0)

// dmax: the largest d = position - 1 among valid positions (known to the host).
hipError_t launch_bbp_fx(const uint64_t* d_pos, uint64_t start, size_t count, uint64_t dmax, uint8_t* d_out,
                         void* ws, int odd_undecided, hipStream_t s, uint32_t block, size_t nout, int tk) {
    const size_t S = 4 * count;
    BbpSeries* st = (BbpSeries*)ws;
    uint64_t* acc = (uint64_t*)((char*)ws + ((S * sizeof(BbpSeries) + 255) & ~(size_t)255));
    const uint32_t stride = d_pos ? 1u : block;
    BBP_LAUNCH("k_bbp_init", s, k_bbp_init, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, s, d_pos, start, stride,
               count, st, acc);
    // terms per lane: at least 32, more while the grid passes 2^17 blocks (and
    // a series has more than one): few long blocks beat many short ones, most
    // of them empty for the shorter series (measured on 1..65,536: 2^23-block
    // cap 102 ms, 2^18 52 ms; 2^17 also halves 1..32,768 to 12.6 ms)
    int KT = 32;
    uint64_t bps = (dmax + 1 + 256ull * KT - 1) / (256ull * KT);
    while (bps > 1 && bps * S > ((uint64_t)1 << 17)) { KT *= 2; bps = (dmax + 1 + 256ull * KT - 1) / (256ull * KT); }
    if (tk == BBP_TK_MODULI && dmax <= BBP_MODULI_DMAX) {
        // about 2^20 blocks: each wavefront takes KB consecutive blocks of 64
        // terms, its set-up and sort shared by their 64 x 64 terms
        const uint64_t groups = (count + 63) / 64;
        const uint64_t kblocks = dmax / 64 + 1;
        const uint64_t want = std::max<uint64_t>(((uint64_t)1 << 20) / groups, 1);
        const uint64_t chunks = std::min<uint64_t>(want, kblocks);
        const uint64_t KB = ((kblocks + chunks - 1) / chunks + BBP_MODULI_NR - 1) / BBP_MODULI_NR * BBP_MODULI_NR;
        BBP_LAUNCH("k_bbp_fx_moduli", s, k_bbp_fx_moduli, dim3((unsigned)(groups * chunks)), dim3(256), 0, s,
                   (const BbpSeries*)st, count, (uint32_t)chunks, (uint32_t)KB, acc);
    } else if (tk != BBP_TK_TERMS && dmax < ((uint64_t)1 << 28)) {          // 32-bit term indices
        // about 2^20 blocks, but at least 64 exponents per block (measured on
        // 1..65,536: 2^14 blocks of 4,096 exponents 49 ms, 2^18 of 256 29 ms,
        // 2^20 of 64 28 ms; every 7th position to 65,536: 10 exponents per
        // block 8.5 ms, 37-147 4.2 ms)
        const uint64_t groups = (count + 63) / 64;
        const uint64_t want = std::max<uint64_t>(((uint64_t)1 << 20) / groups, 1);
        const uint64_t EK = std::max<uint64_t>((dmax + want) / want, 64);
        const uint64_t chunks = (dmax + EK) / EK;
        BBP_LAUNCH("k_bbp_fx_lanes", s, k_bbp_fx_lanes, dim3((unsigned)(groups * chunks)), dim3(256), 0, s,
                   (const BbpSeries*)st, count, (uint32_t)chunks, (uint32_t)EK, acc);
    } else if (dmax < ((uint64_t)1 << 28)) {
        BBP_LAUNCH("k_bbp_fx_terms", s, k_bbp_fx_terms<false>, dim3((unsigned)(bps * S)), dim3(256), 0, s,
                   (const BbpSeries*)st, (unsigned)bps, KT, acc);
    } else {
        BBP_LAUNCH("k_bbp_fx_terms_wide", s, k_bbp_fx_terms<true>, dim3((unsigned)(bps * S)), dim3(256), 0, s,
                   (const BbpSeries*)st, (unsigned)bps, KT, acc);
    }
    if (d_pos || block <= 1)
        BBP_LAUNCH("k_bbp_fx_finish", s, k_bbp_fx_finish, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s,
                   (const BbpSeries*)st, count, (const uint64_t*)acc, d_out, odd_undecided);
    else
        BBP_LAUNCH("k_bbp_fx_block", s, k_bbp_fx_block, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s,
                   (const BbpSeries*)st, count, block, nout, (const uint64_t*)acc, d_out, odd_undecided);
    return hipSuccess;
}

hipError_t launch_bbp_chunked(const uint64_t* d_pos, uint64_t start, size_t count, uint64_t dmax, uint8_t* d_out,
                              void* ws, hipStream_t s) {
    const size_t S = 4 * count;
    const int log_ch = chunk_log(S);
    BbpSeries* st = (BbpSeries*)ws;
    double* buf = (double*)((char*)ws + ((S * sizeof(BbpSeries) + 255) & ~(size_t)255));
    BBP_LAUNCH("k_bbp_init", s, k_bbp_init, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, s, d_pos, start, 1u,
               count, st, (uint64_t*)nullptr);
    const uint64_t ch = (uint64_t)1 << log_ch;
    const size_t lanes = S << log_ch;
    for (uint64_t k0 = 0; k0 <= dmax; k0 += ch) {
        BBP_LAUNCH("k_bbp_terms", s, k_bbp_terms, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, s,
                   (const BbpSeries*)st, S, k0, lanes, buf);
        BBP_LAUNCH("k_bbp_fold", s, k_bbp_fold, dim3((unsigned)((S + 63) / 64)), dim3(64), 0, s, st, S, k0, ch,
                   (const double*)buf);
    }
    BBP_LAUNCH("k_bbp_finish", s, k_bbp_finish, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, s,
               (const BbpSeries*)st, count, d_out);
    return hipSuccess;
}

// Positions and their indices sorted by the low 27 bits of the position
// (every position the moduli kernel takes lies below 2^27; others land
// anywhere, which changes the cost and never a digit). With temp null,
// *temp_bytes receives the scratch size rocPRIM's radix sort needs.
hipError_t bbp_sort_positions(const uint64_t* d_pos, size_t n, uint64_t* d_sorted, uint32_t* d_perm, void* temp,
                              size_t* temp_bytes, hipStream_t s) {
    return rocprim::radix_sort_pairs(temp, *temp_bytes, d_pos, d_sorted, rocprim::counting_iterator<uint32_t>(0u),
                                     d_perm, (unsigned)n, 0u, 27u, s);
}

__global__ __launch_bounds__(256) void k_bbp_scatter(const uint8_t* __restrict__ in, const uint32_t* __restrict__ perm,
                                                     size_t n, uint8_t* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[perm[i]] = in[i];
}

hipError_t launch_bbp_scatter(const uint8_t* d_in, const uint32_t* d_perm, size_t n, uint8_t* d_out, hipStream_t s) {
    BBP_LAUNCH("k_bbp_scatter", s, k_bbp_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_in, d_perm, n,
               d_out);
    return hipSuccess;
}

hipError_t launch_bbp_group_scan(const uint64_t* d_pos, size_t count, uint64_t* d_out3, hipStream_t s) {
    const size_t groups = (count + 63) / 64;
    BBP_LAUNCH("k_bbp_group_scan", s, k_bbp_group_scan, dim3((unsigned)((groups + 255) / 256)), dim3(256), 0, s,
               d_pos, count, (unsigned long long*)d_out3);
    return hipSuccess;
}


// ---- integer multiply issue-rate probe ------------------------------------------------
// The ceiling the BBP throughput is compared with: eight independent chains of
// t = lo32(t) * x + t (one v_mad_u64_u32 each) per lane over a fixed grid,
// every access in bounds (one u64 out per thread).
__global__ __launch_bounds__(256) void k_probe_mad(uint64_t* __restrict__ out, uint32_t seed, int iters) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t x = (seed ^ (uint32_t)gid) | 1u;
    uint64_t t[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) t[c] = gid * 8 + c;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < 8; ++c) t[c] = (uint64_t)(uint32_t)t[c] * x + t[c];
    }
    uint64_t acc = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) acc += t[c];
    out[gid] = acc;
}

hipError_t probe_mad_rate(double* lane_ops_per_s) {
    constexpr unsigned blocks = 8192, threads = 256;
    constexpr int iters = 4096;
    uint64_t* d = nullptr;
    hipError_t e = hipMalloc(&d, (size_t)blocks * threads * 8);
    if (e != hipSuccess) return e;
    hipEvent_t a = nullptr, b = nullptr;
    float ms = 0.f;
    if ((e = hipEventCreate(&a)) == hipSuccess && (e = hipEventCreate(&b)) == hipSuccess) {
        hipLaunchKernelGGL(k_probe_mad, dim3(blocks), dim3(threads), 0, 0, d, 7u, iters);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(a, 0);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_probe_mad, dim3(blocks), dim3(threads), 0, 0, d, 11u, iters);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(b, 0);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipFree(d);
    if (e == hipSuccess) *lane_ops_per_s = (double)blocks * threads * iters * 8 / (ms * 1e-3);
    return e;
}

hipError_t launch_bbp(const uint64_t* d_pos, uint64_t start, size_t count, uint8_t* d_out, uint64_t* d_next,
                      hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipError_t ez = launch_zero_len(d_next, s);
    if (ez != hipSuccess) return ez;
    // at most as many wavefronts as groups, and no more than the GPU holds
    const size_t waves = std::min<size_t>((count + 15) / 16, (size_t)256 * 32);
    const dim3 grid((unsigned)((waves + 3) / 4));
    hipEvent_t t0 = timing_begin(s);
    hipLaunchKernelGGL(k_bbp, grid, dim3(256), 0, s, d_pos, start, count, d_out, (unsigned long long*)d_next);
    g_launches.fetch_add(1, std::memory_order_relaxed);
    hipError_t e = hipGetLastError();
    timing_end("k_bbp", t0, s);
    return e;
}

} // namespace pdh
