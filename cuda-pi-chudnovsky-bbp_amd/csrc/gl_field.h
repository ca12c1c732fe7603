// ============================================================================
//  gl_field.h -- Goldilocks GF(p), p = 2^64 - 2^32 + 1, for gfx950 and host.
//
//  Same field and canonical [0, p) outputs as the reference's
//  include/goldilocks.h:83-189; the code is this library's own. Every
//  function is __host__ __device__ so the host builds the twiddle tables with
//  the exact arithmetic the kernels use.
//
//  Reduction uses 2^64 == 2^32 - 1 and 2^96 == -1 (mod p) and no multiply;
//  on gfx950 the 64x64 product is four v_mad_u64_u32 partial products.
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

// (hh:hl:lo) mod p for a 128-bit value lo + hl*2^64 + hh*2^96, using
// 2^64 == 2^32 - 1 and 2^96 == -1: the value is congruent to
//   s1 - s2 + c1*EPS,  s1 = lo + hl*2^32 (mod 2^64, carry c1),  s2 = hl + hh,
// and with t = s1 - s2 (mod 2^64, borrow b) it is t + (c1 - b)*EPS, which
// stays inside [0, 2^64): if c1 = 1 and b = 0 then s1 < p so t + EPS < 2^64;
// if b = 1 and c1 = 0 then t >= 2^64 - 2^33 > EPS. One conditional subtract
// of p finishes. No multiply: hl*EPS is never formed.
__host__ __device__ __forceinline__ u64 gl_reduce3(u64 lo, u32 hl, u32 hh) {
    const u32 lo0 = (u32)lo, lo1 = (u32)(lo >> 32);
    const u32 s1hi = lo1 + hl;
    const bool c1 = s1hi < hl;
    const u64 s1 = ((u64)s1hi << 32) | lo0;
    const u64 s2 = (u64)hl + hh;
    const u64 t = s1 - s2;
    const bool b = s1 < s2;
    u64 r = t;
    if (c1 && !b) r = t + GL_EPS;
    if (b && !c1) r = t - GL_EPS;
    const u64 r2 = r + GL_EPS;             // wraps exactly when r >= p
    return (r2 < r) ? r2 : r;
}

__host__ __device__ __forceinline__ u64 gl_reduce(u64 lo, u64 hi) {
    return gl_reduce3(lo, (u32)hi, (u32)(hi >> 32));
}

// 64 x 64 -> 128 from four 32 x 32 + 64 multiply-adds (v_mad_u64_u32 on
// gfx950), then the multiply-free fold above.
__host__ __device__ __forceinline__ u64 gl_mul(u64 a, u64 b) {
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    const u64 p00 = (u64)a0 * b0;
    const u64 p01 = (u64)a0 * b1 + (p00 >> 32);
    const u64 p10 = (u64)a1 * b0 + (u32)p01;
    const u64 p11 = (u64)a1 * b1 + (p01 >> 32) + (p10 >> 32);   // < 2^64: no overflow
    const u64 lo = ((u64)(u32)p10 << 32) | (u32)p00;
    return gl_reduce3(lo, (u32)p11, (u32)(p11 >> 32));
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
        const u64 hi = x >> (64 - K);
        return gl_reduce3(x << K, (u32)hi, (u32)(hi >> 32));
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
inline u64 gl_rho_search() {
    u64 z = gl_pow(7, (GL_P - 1) >> 32);
    u64 y = gl_pow(z, 1ull << 26);
    u64 t = y;                       // y^u for odd u, stepping by y^2
    u64 y2 = gl_mul(y, y);
    for (u64 u = 1; u < 64; u += 2) {
        if (t == 8) return gl_pow(z, u);
        t = gl_mul(t, y2);
    }
    return 0;                        // unreachable: 8 is a primitive 64th root
}
// The search runs once per process; the initialisation of a function-local
// static is thread-safe in C++11, so concurrent first calls are too.
inline u64 gl_rho_host() {
    static const u64 rho = gl_rho_search();
    return rho;
}
// Primitive 2^k-th root of the system (k <= 32).
inline u64 gl_root_host(int k) { return gl_pow(gl_rho_host(), 1ull << (32 - k)); }

} // namespace pdh
