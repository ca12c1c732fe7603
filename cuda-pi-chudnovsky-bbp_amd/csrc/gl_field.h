// ============================================================================
//  gl_field.h -- Goldilocks GF(p), p = 2^64 - 2^32 + 1, for gfx950 and host.
//
//  Same field and canonical [0, p) outputs as the reference's
//  include/goldilocks.h:83-189; the code is this library's own. Every
//  function is __host__ __device__ so the host builds the twiddle tables with
//  the exact arithmetic the kernels use.
//
//  Reduction uses 2^64 == 2^32 - 1 and 2^96 == -1 (mod p); on gfx950 the
//  64x64 product lowers to v_mad_u64_u32 / v_mul_hi_u32 partial products.
//
//  Root system. The library does not use the reference's roots 7^((p-1)/N):
//  any primitive N-th root gives the same (unique) integer product. It uses
//  one compatible system rho_{2^k} = rho^(2^(32-k)) built from a primitive
//  2^32-th root rho chosen so that rho_64 = 8 = 2^3. Then every root of order
//  <= 64 is a power of two, so radix-8 butterflies multiply by shifts only:
//  w8 = 2^24, w4 = 2^48, w2 = 2^96 = -1.
// ============================================================================
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace pdh {

typedef uint64_t u64;
typedef uint32_t u32;

constexpr u64 GL_P   = 0xFFFFFFFF00000001ULL;
constexpr u64 GL_EPS = 0xFFFFFFFFULL;          // 2^64 mod p

__host__ __device__ __forceinline__ u64 gl_mulhi(u64 a, u64 b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (u64)(((unsigned __int128)a * b) >> 64);
#endif
}

// a + b for canonical a, b. s + EPS wraps exactly when s >= p; a carry out of
// a + b means the true sum is s + 2^64 == s + EPS, which cannot wrap again.
__host__ __device__ __forceinline__ u64 gl_add(u64 a, u64 b) {
    u64 s = a + b;
    u64 t = s + GL_EPS;
    return ((s < a) | (t < s)) ? t : s;
}

// a - b for canonical a, b: on borrow the wrapped difference exceeds EPS, and
// subtracting EPS adds p back.
__host__ __device__ __forceinline__ u64 gl_sub(u64 a, u64 b) {
    u64 d = a - b;
    return (a < b) ? d - GL_EPS : d;
}

__host__ __device__ __forceinline__ u64 gl_neg(u64 a) { return a ? GL_P - a : 0; }

// (hi:lo) mod p: lo - hi_hi + hi_lo * EPS, one carry fold, one conditional
// subtract (the second carry fold of a general 128-bit input cannot happen
// after the first: t2 < hi_lo * EPS < 2^64 - 2^33).
__host__ __device__ __forceinline__ u64 gl_reduce(u64 lo, u64 hi) {
    u32 hh = (u32)(hi >> 32);
    u32 hl = (u32)hi;
    u64 t0 = lo - (u64)hh;
    if (lo < (u64)hh) t0 -= GL_EPS;
    u64 t1 = ((u64)hl << 32) - (u64)hl;
    u64 t2 = t0 + t1;
    if (t2 < t0) t2 += GL_EPS;
    u64 t3 = t2 + GL_EPS;
    return (t3 < t2) ? t3 : t2;
}

__host__ __device__ __forceinline__ u64 gl_mul(u64 a, u64 b) {
    return gl_reduce(a * b, gl_mulhi(a, b));
}

__host__ __device__ inline u64 gl_pow(u64 b, u64 e) {
    u64 r = 1;
    while (e) { if (e & 1) r = gl_mul(r, b); b = gl_mul(b, b); e >>= 1; }
    return r;
}
__host__ __device__ inline u64 gl_inv(u64 a) { return gl_pow(a, GL_P - 2); }

// x * 2^K mod p for a compile-time K in [0, 192); x canonical.
//   K >= 96      : 2^96 == -1, negate the K-96 product.
//   K < 64       : x*2^K = (x << K) + 2^64 (x >> (64-K)); general fold.
//   64 <= K < 96 : with j = K-64, L = x << j (mod 2^64), H = x >> (64-j):
//                  x*2^K = L*2^64 + H*2^128 == L0*2^32 - L0 - L1 - H*2^32,
//                  and L0*2^32 < p, L0 + L1 + H*2^32 < p.
template <int K>
__host__ __device__ __forceinline__ u64 gl_mulp2(u64 x) {
    static_assert(K >= 0 && K < 192, "shift exponent out of range");
    if constexpr (K == 0) {
        return x;
    } else if constexpr (K >= 96) {
        return gl_neg(gl_mulp2<K - 96>(x));
    } else if constexpr (K < 64) {
        return gl_reduce(x << K, x >> (64 - K));
    } else {
        constexpr int j = K - 64;
        u64 L = x << j;
        u64 H = (j == 0) ? 0 : (x >> ((64 - j) & 63));
        u64 L0 = (u32)L, L1 = L >> 32;
        return gl_sub(L0 << 32, (H << 32) + L0 + L1);
    }
}

// ---- the library's root system ---------------------------------------------
// rho: primitive 2^32-th root with rho^(2^26) == 8. Found once on the host:
// z = 7^((p-1)/2^32) is primitive of order 2^32; y = z^(2^26) is a primitive
// 64th root and 8 = y^u for exactly one odd u < 64; rho = z^u.
inline u64 gl_rho_host() {
    static u64 cached = 0;
    if (cached) return cached;
    u64 z = gl_pow(7, (GL_P - 1) >> 32);
    u64 y = gl_pow(z, 1ull << 26);
    u64 t = y;                       // y^u for odd u, stepping by y^2
    u64 y2 = gl_mul(y, y);
    for (u64 u = 1; u < 64; u += 2) {
        if (t == 8) { cached = gl_pow(z, u); return cached; }
        t = gl_mul(t, y2);
    }
    return 0;                        // unreachable: 8 is a primitive 64th root
}
// Primitive 2^k-th root of the system (k <= 32).
inline u64 gl_root_host(int k) { return gl_pow(gl_rho_host(), 1ull << (32 - k)); }

} // namespace pdh
