// ============================================================================
//  ntt3_kernels.hip -- three-prime NTT multiply for gfx950 (products with
//  N = 2^n > 2^13 in the reference's 16-bit digit count).
//
//  The product is a unique integer, so any exact convolution method gives the
//  reference's result (src/ntt_cuda.cu:179-243, include/ntt.h:47-91). This
//  engine convolves 32-bit digits (the limbs themselves) modulo three primes
//  below 2^30 and recombines each coefficient by CRT:
//
//      p0 = 105 * 2^23 + 1,  p1 = 107 * 2^23 + 1,  p2 = 119 * 2^23 + 1
//
//  A coefficient of a length-2^m cyclic convolution of 32-bit digits is at
//  most min(na, nb) (2^32 - 1)^2 < 2^86 for m <= 23 (na + nb <= 2^m), below
//  p0 p1 p2 (2^89.35), so it is recovered exactly. The transforms are half as
//  long as with 16-bit digits.
//
//  Lazy residues (Harvey's butterflies): inside the transforms values live in
//  [0, 2p), sums and differences in [0, 4p) (4p < 2^32), and a Montgomery
//  product (R = 2^32) of a b < 4p^2 is (a b + m p) / 2^32 < p (4p / 2^32 + 1)
//  < 2p without a final correction: 3 VALU ops (two v_mad_u64_u32, one
//  v_mul_lo_u32). A twiddled butterfly output costs sub + add + 3, a reduction
//  [0, 4p) -> [0, 2p) 2 ops (sub 2p, unsigned min). A radix-8 butterfly with
//  its seven twiddles is 96 ops (Goldilocks: about 450). Canonical residues
//  are only formed for the CRT.
//
//  Same four-step structure as the Goldilocks engine: column DIF per operand
//  and prime (k3_col_fwd), a fused row kernel per prime (outer twiddle,
//  forward A and B, pointwise, inverse, inverse twiddle with the scale), and
//  k3_col_inv, which runs the inverse columns of all three primes of a tile,
//  the CRT and each row segment's carry with carry-in 0, feeding the shared
//  carry scan. Values are canonical residues in [0, p); twiddles are stored
//  in Montgomery form (w R mod p) so that mm(x, wR) = x w.
// ============================================================================
#include "ntt_common.h"

namespace pdh {

// ---- host constants ---------------------------------------------------------------
static const uint32_t kP3[3] = {880803841u, 897581057u, 998244353u};      // 105, 107, 119 * 2^23 + 1 (ascending)
static const uint32_t kG3[3] = {26u, 3u, 3u};                              // generators of the unit groups
static const int kK3[3] = {23, 23, 23};                                    // 2-adicity of p - 1

uint32_t r3_mul(uint32_t a, uint32_t b, uint32_t p) { return (uint32_t)((uint64_t)a * b % p); }
uint32_t r3_pow(uint32_t a, uint64_t e, uint32_t p) {
    uint32_t r = 1;
    for (; e; e >>= 1, a = r3_mul(a, a, p)) if (e & 1) r = r3_mul(r, a, p);
    return r;
}
uint32_t r3_to_mont(uint32_t a, uint32_t p) { return (uint32_t)(((uint64_t)a << 32) % p); }
static uint32_t r3_inv(uint32_t a, uint32_t p) { return r3_pow(a, p - 2, p); }

// Compatible root system: w_{2^L} = rho^(2^(K - L)), rho = g^((p - 1) / 2^K).
uint32_t r3_root(int pi, int L) {
    const uint32_t p = kP3[pi];
    const uint32_t rho = r3_pow(kG3[pi], (p - 1) >> kK3[pi], p);
    return r3_pow(rho, (uint64_t)1 << (kK3[pi] - L), p);
}

const Primes3& primes3() {
    static const Primes3 PS = [] {
        Primes3 s{};
        for (int i = 0; i < 3; ++i) {
            const uint32_t p = kP3[i];
            uint32_t inv = 1;                                  // p^-1 mod 2^32 by Newton
            for (int k = 0; k < 5; ++k) inv *= 2u - p * inv;
            s.p[i].p = p;
            s.p[i].ninv = 0u - inv;
            const uint32_t w8 = r3_root(i, 3), w8i = r3_inv(w8, p);
            for (int k = 0; k < 4; ++k) {
                s.p[i].w8[k] = r3_to_mont(r3_pow(w8, k, p), p);
                s.p[i].w8i[k] = r3_to_mont(r3_pow(w8i, k, p), p);
            }
            const uint32_t w16 = r3_root(i, 4), w16i = r3_inv(w16, p);
            const int e16[3] = {1, 3, 9};
            for (int k = 0; k < 3; ++k) {
                s.p[i].w16[k] = r3_to_mont(r3_pow(w16, e16[k], p), p);
                s.p[i].w16i[k] = r3_to_mont(r3_pow(w16i, e16[k], p), p);
            }
        }
        const uint32_t p0 = kP3[0], p1 = kP3[1], p2 = kP3[2];
        s.inv_p0_m1 = r3_to_mont(r3_inv(p0 % p1, p1), p1);
        s.p0_m2 = r3_to_mont(p0 % p2, p2);
        s.inv_p0p1_m2 = r3_to_mont(r3_inv(r3_mul(p0 % p2, p1 % p2, p2), p2), p2);
        s.p0p1 = (uint64_t)p0 * p1;
        return s;
    }();
    return PS;
}

// ---- residue arithmetic -----------------------------------------------------------
__device__ __forceinline__ u32 umin32(u32 a, u32 b) { return a < b ? a : b; }
__device__ __forceinline__ u32 r_sub(u32 a, u32 b, u32 p) { const u32 d = a - b; return umin32(d, d + p); }
// a * b * 2^-32 mod p for a, b < p < 2^31: t + m p < 2^64 and (t + m p) / 2^32 < 2p.
__device__ __forceinline__ u32 mm(u32 a, u32 b, u32 p, u32 ninv) {
    const u64 t = (u64)a * b;
    const u32 m = (u32)t * ninv;
    const u32 r = (u32)(((u64)m * p + t) >> 32);
    return umin32(r, r - p);
}

// Lazy forms: red2 takes [0, 4p) to [0, 2p); mml is the Montgomery product
// without its final correction, valid for a b < 4p^2, result in [0, 2p).
__device__ __forceinline__ u32 red2(u32 x, u32 p2) { return umin32(x, x - p2); }
__device__ __forceinline__ u32 mml(u32 a, u32 b, u32 p, u32 ninv) {
    const u64 t = (u64)a * b;
    const u32 m = (u32)t * ninv;
    return (u32)(((u64)m * p + t) >> 32);
}
// (a - b) w R^-1 mod p as (a w + b (p - w)) R^-1: two multiply-adds take the
// place of the subtraction and its 2p offset. For a, b < 4p and a canonical
// twiddle w the sum is at most max(a, b) p < 4p^2 < p 2^32, so the result is
// in [0, 2p) as mml's.
__device__ __forceinline__ u32 mmd(u32 a, u32 b, u32 w, u32 p, u32 ninv) {
    const u64 t = (u64)a * w + (u64)b * (p - w);
    const u32 m = (u32)t * ninv;
    return (u32)(((u64)m * p + t) >> 32);
}

// ---- radix-q DFTs on lazy residues, natural order in and out ----------------------
// Inputs in [0, 2p). The DFT4 is two radix-2 halves: c0 = x0 + x2,
// c1 = x1 + x3, d0 = x0 - x2 and d1 = (x1 - x3) w4, each in [0, 2p); its
// outputs are Y0 = c0 + c1, Y1 = d0 + d1, Y2 = c0 - c1, Y3 = d0 - d1.
template <bool INV>
__device__ __forceinline__ void l_half4(u32 x0, u32 x1, u32 x2, u32 x3, u32& c0, u32& c1, u32& d0, u32& d1,
                                        const Prime3& P) {
    const u32 p2 = P.p << 1;
    c0 = red2(x0 + x2, p2); d0 = red2(x0 + p2 - x2, p2);
    c1 = red2(x1 + x3, p2);
    d1 = mmd(x1, x3, INV ? P.w8i[2] : P.w8[2], P.p, P.ninv);   // w4^{-1}, w4
}

// Output forms of a DFT's last radix-4 round: RAW leaves the outputs in
// [0, 4p); RED reduces each to [0, 2p); TW reduces output 0 and multiplies
// output k >= 1 by the twiddle w[k] (result in [0, 2p)), the differences
// by mmd, so the subtraction costs nothing extra.
enum : int { OUT_RAW = 0, OUT_RED = 1, OUT_TW = 2 };

template <int OUT>
__device__ __forceinline__ void out4(u32 c0, u32 c1, u32 d0, u32 d1, u32& y0, u32& y1, u32& y2, u32& y3,
                                     const u32* w, int k0, int ks, const Prime3& P) {
    const u32 p2 = P.p << 1;
    if constexpr (OUT == OUT_RAW) {
        y0 = c0 + c1; y2 = c0 + p2 - c1;
        y1 = d0 + d1; y3 = d0 + p2 - d1;
    } else if constexpr (OUT == OUT_RED) {
        y0 = red2(c0 + c1, p2); y2 = red2(c0 + p2 - c1, p2);
        y1 = red2(d0 + d1, p2); y3 = red2(d0 + p2 - d1, p2);
    } else {
        y0 = k0 == 0 ? red2(c0 + c1, p2) : mml(c0 + c1, w[k0], P.p, P.ninv);
        y1 = mml(d0 + d1, w[k0 + ks], P.p, P.ninv);
        y2 = mmd(c0, c1, w[k0 + 2 * ks], P.p, P.ninv);
        y3 = mmd(d0, d1, w[k0 + 3 * ks], P.p, P.ninv);
    }
}

// Inputs in [0, 2p), outputs in [0, 4p).
template <bool INV>
__device__ __forceinline__ void l_dft4(u32& x0, u32& x1, u32& x2, u32& x3, const Prime3& P) {
    u32 c0, c1, d0, d1;
    l_half4<INV>(x0, x1, x2, x3, c0, c1, d0, d1, P);
    out4<OUT_RAW>(c0, c1, d0, d1, x0, x1, x2, x3, nullptr, 0, 0, P);
}

// Y_{2k} = DFT4(x_t + x_{t+4})_k, Y_{2k+1} = DFT4((x_t - x_{t+4}) w8^t)_k.
template <bool INV>
__device__ __forceinline__ void l_dft8(u32* x, const Prime3& P) {
    const u32* w = INV ? P.w8i : P.w8;
    const u32 p2 = P.p << 1;
    u32 a0 = red2(x[0] + x[4], p2), b0 = red2(x[0] + p2 - x[4], p2);
    u32 a1 = red2(x[1] + x[5], p2), b1 = mmd(x[1], x[5], w[1], P.p, P.ninv);
    u32 a2 = red2(x[2] + x[6], p2), b2 = mmd(x[2], x[6], w[2], P.p, P.ninv);
    u32 a3 = red2(x[3] + x[7], p2), b3 = mmd(x[3], x[7], w[3], P.p, P.ninv);
    l_dft4<INV>(a0, a1, a2, a3, P);
    l_dft4<INV>(b0, b1, b2, b3, P);
    x[0] = a0; x[2] = a1; x[4] = a2; x[6] = a3;
    x[1] = b0; x[3] = b1; x[5] = b2; x[7] = b3;
}

// 16 = 4 x 4: with t = 4 t1 + t2 and k = k1 + 4 k2,
// Y_k = sum_t2 w4^(t2 k2) [w16^(t2 k1) sum_t1 w4^(t1 k1) x_{4 t1 + t2}]:
// four DFT4 over t1, nine twiddles w16^(t2 k1) (t2, k1 >= 1), four DFT4
// over t2. The middle values are reduced to [0, 2p) by the twiddle product or
// red2, so each DFT4 sees inputs in [0, 2p); the middle differences with a
// twiddle go through mmd. OUT selects the outputs' form (out4), with the
// twiddles w[k] in natural order for OUT_TW.
template <bool INV, int OUT = OUT_RAW>
__device__ __forceinline__ void l_dft16(u32* x, const Prime3& P, const u32* w = nullptr) {
    const u32 p2 = P.p << 1;
    const u32* w8 = INV ? P.w8i : P.w8;
    const u32* w16 = INV ? P.w16i : P.w16;
    // inner twiddles w16^(t2 k1) for t2, k1 = 1..3 (row t2 - 1):
    // 1 2 3 / 2 4 6 / 3 6 9 -> w16, w8, w16^3 / w8, w4, w8^3 / w16^3, w8^3, w16^9
    const u32 wi[3][3] = {{w16[0], w8[1], w16[1]}, {w8[1], w8[2], w8[3]}, {w16[1], w8[3], w16[2]}};
    u32 m[16];   // m[t2 + 4 k1]: the inner DFT at (t2, k1), in [0, 2p)
#pragma unroll
    for (int t2 = 0; t2 < 4; ++t2) {
        u32 c0, c1, d0, d1;
        l_half4<INV>(x[t2], x[t2 + 4], x[t2 + 8], x[t2 + 12], c0, c1, d0, d1, P);
        m[t2] = red2(c0 + c1, p2);
        if (t2 == 0) {
            m[4] = red2(d0 + d1, p2);
            m[8] = red2(c0 + p2 - c1, p2);
            m[12] = red2(d0 + p2 - d1, p2);
        } else {
            m[t2 + 4] = mml(d0 + d1, wi[t2 - 1][0], P.p, P.ninv);
            m[t2 + 8] = mmd(c0, c1, wi[t2 - 1][1], P.p, P.ninv);
            m[t2 + 12] = mmd(d0, d1, wi[t2 - 1][2], P.p, P.ninv);
        }
    }
    // the outer DFT4 over t2 of row k1 gives Y_{k1 + 4 k2}, k2 = 0..3
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        u32 c0, c1, d0, d1;
        l_half4<INV>(m[4 * k1], m[4 * k1 + 1], m[4 * k1 + 2], m[4 * k1 + 3], c0, c1, d0, d1, P);
        out4<OUT>(c0, c1, d0, d1, x[k1], x[k1 + 4], x[k1 + 8], x[k1 + 12], w, k1, 4, P);
    }
}

template <int Q, bool INV>
__device__ __forceinline__ void l_dftq(u32* x, const Prime3& P) {
    if constexpr (Q == 2) {
        const u32 a = x[0] + x[1], b = x[0] + (P.p << 1) - x[1];
        x[0] = a; x[1] = b;
    } else if constexpr (Q == 4) {
        l_dft4<INV>(x[0], x[1], x[2], x[3], P);
    } else if constexpr (Q == 8) {
        l_dft8<INV>(x, P);
    } else {
        l_dft16<INV>(x, P);
    }
}


// One lazy butterfly of a level, values in [0, 2p) in and out. Forward: DFT_Q
// (outputs in [0, 4p)), then each output c >= 1 times its twiddle (mml gives
// [0, 2p)) and output 0 reduced; without twiddles every output is reduced.
// Inverse: inputs c >= 1 times their twiddles first, then DFT_Q and reduce.
// Two inverse options let consecutive levels skip reductions: ROUT = false
// leaves the outputs in [0, 4p), and RIN0 reduces input 0 of a twiddled
// butterfly, so a level with RIN0 takes inputs in [0, 4p) (a twiddle product
// of a < 4p and a canonical w < p is (a w + m p) / 2^32 < p (4p / 2^32 + 1)
// < 2p since 4p < 2^32). A level without twiddles needs inputs in [0, 2p).
template <int Q, bool INV, bool RIN0 = false, bool ROUT = true>
__device__ __forceinline__ void lazy_butterfly(u32* x, const u32* w, bool has_tw, const Prime3& P) {
    const u32 p2 = P.p << 1;
    if constexpr (Q == 16) {
        // the outputs' reduction or twiddle product fused into the last round
        if (!INV) {
            if (has_tw) l_dft16<false, OUT_TW>(x, P, w);
            else l_dft16<false, OUT_RED>(x, P);
        } else {
            if (has_tw) {
                if constexpr (RIN0) x[0] = red2(x[0], p2);
#pragma unroll
                for (int c = 1; c < Q; ++c) x[c] = mml(x[c], w[c], P.p, P.ninv);
            }
            l_dft16<true, ROUT ? OUT_RED : OUT_RAW>(x, P);
        }
    } else if (!INV) {
        l_dftq<Q, false>(x, P);
        if (has_tw) {
            x[0] = red2(x[0], p2);
#pragma unroll
            for (int c = 1; c < Q; ++c) x[c] = mml(x[c], w[c], P.p, P.ninv);
        } else {
#pragma unroll
            for (int c = 0; c < Q; ++c) x[c] = red2(x[c], p2);
        }
    } else {
        if (has_tw) {
            if constexpr (RIN0) x[0] = red2(x[0], p2);
#pragma unroll
            for (int c = 1; c < Q; ++c) x[c] = mml(x[c], w[c], P.p, P.ninv);
        }
        l_dftq<Q, true>(x, P);
        if constexpr (ROUT) {
#pragma unroll
            for (int c = 0; c < Q; ++c) x[c] = red2(x[c], p2);
        }
    }
}

// Level twiddles, per prime: for each radix Q in {2, 4, 8, 16}, direction and
// sub-block size M = 2^logM (logM <= 14), the block at
// ((qi * 2 + INV) << 15) + 2^logM holds w_M^(+-j c) R mod p at j Q + c
// (j < M / Q, c < Q). A butterfly's Q - 1 twiddles are then contiguous and
// load from one address; w_M^(j c) equals the w_{2^L}^((j 2^(L - logM)) c)
// of the full sub-transform, so the table does not depend on L.
constexpr int TW3_BLOCK = 1 << 15;
template <int Q, bool INV>
__device__ __forceinline__ const u32* tw_level(const u32* tw_prime, int logM) {
    constexpr int qi = QLog<Q>::v - 1;
    return tw_prime + (size_t)((qi * 2 + (INV ? 1 : 0)) * TW3_BLOCK) + (1 << logM);
}

// One radix-Q level of an in-place 2^L-point sub-transform of NARR residue
// arrays in LDS (same index map as the Goldilocks lds_level): butterfly
// (beta, j) of sub-block size M = 2^logM = Q m reads i = beta M + j + t m.
// Forward: DFT_Q then output c times w_M^(j c); inverse: the mirror.
// tw is the prime's level-twiddle table.
template <int Q, bool INV, int NARR>
__device__ __forceinline__ void r_lds_level(u32* const* arr, int L, int logM, int tlog,
                                            const u32* __restrict__ tw, const Prime3& P) {
    constexpr int LQ = QLog<Q>::v;
    const int logm = logM - LQ;
    const int nbf = 1 << (L - LQ + tlog);
    const int tmask = (1 << tlog) - 1;
    const bool has_tw = logm > 0;
    for (int u = threadIdx.x; u < nbf; u += blockDim.x) {
        const int col = u & tmask;
        const int rest = u >> tlog;
        const int j = rest & ((1 << logm) - 1);
        const int base = ((rest >> logm) << logM) + j;
        u32 w[Q];
        if (has_tw) {
            const u32* __restrict__ tl = tw_level<Q, INV>(tw, logM) + j * Q;
#pragma unroll
            for (int c = 1; c < Q; ++c) w[c] = tl[c];
        }
        const int a0 = lpad((base << tlog) + col);
#pragma unroll
        for (int a = 0; a < NARR; ++a) {
            u32 x[Q];
#pragma unroll
            for (int t = 0; t < Q; ++t) x[t] = arr[a][a0 + pad_step(t << (logm + tlog))];
            lazy_butterfly<Q, INV>(x, w, has_tw, P);
#pragma unroll
            for (int t = 0; t < Q; ++t) arr[a][a0 + pad_step(t << (logm + tlog))] = x[t];
        }
    }
    __syncthreads();
}

// Radix-8 sub-transform of 2^L points: DIF = a radix-2^(L mod 3) level then
// radix-8 levels; the inverse is the exact mirror. Ends with a barrier.
template <bool INV, int NARR>
__device__ __forceinline__ void r_lds_dft(u32* const* arr, int L, int tlog, const u32* __restrict__ tw,
                                          const Prime3& P) {
    const int r0 = L % 3;
    if (!INV) {
        int logM = L;
        if (r0 == 1) { r_lds_level<2, false, NARR>(arr, L, logM, tlog, tw, P); logM -= 1; }
        if (r0 == 2) { r_lds_level<4, false, NARR>(arr, L, logM, tlog, tw, P); logM -= 2; }
        for (; logM >= 3; logM -= 3) r_lds_level<8, false, NARR>(arr, L, logM, tlog, tw, P);
    } else {
        for (int logM = 3; logM <= L - r0; logM += 3) r_lds_level<8, true, NARR>(arr, L, logM, tlog, tw, P);
        if (r0 == 1) r_lds_level<2, true, NARR>(arr, L, L, tlog, tw, P);
        if (r0 == 2) r_lds_level<4, true, NARR>(arr, L, L, tlog, tw, P);
    }
}

// The same level applied to the three residue arrays of one tile (one per
// prime) in one pass: the butterfly index math and the twiddle exponents are
// shared, each prime reads its own table (stride TW3_STRIDE) at the same
// index.
template <int Q, bool INV>
__device__ __forceinline__ void r_lds_level3(u32* const* arr, int L, int logM, int tlog,
                                             const u32* __restrict__ tw0, const Primes3& PS) {
    constexpr int LQ = QLog<Q>::v;
    const int logm = logM - LQ;
    const int nbf = 1 << (L - LQ + tlog);
    const int tmask = (1 << tlog) - 1;
    const bool has_tw = logm > 0;
    for (int u = threadIdx.x; u < nbf; u += blockDim.x) {
        const int col = u & tmask;
        const int rest = u >> tlog;
        const int j = rest & ((1 << logm) - 1);
        const int base = ((rest >> logm) << logM) + j;
        int idx[Q];
        const int a0 = lpad((base << tlog) + col);
#pragma unroll
        for (int t = 0; t < Q; ++t) idx[t] = a0 + pad_step(t << (logm + tlog));
#pragma unroll
        for (int pi = 0; pi < 3; ++pi) {
            const Prime3& P = PS.p[pi];
            const u32* __restrict__ tl = tw_level<Q, INV>(tw0 + (size_t)pi * TW3_STRIDE, logM) + j * Q;
            u32 x[Q];
#pragma unroll
            for (int t = 0; t < Q; ++t) x[t] = arr[pi][idx[t]];
            u32 w[Q];
            if (has_tw) {
#pragma unroll
                for (int c = 1; c < Q; ++c) w[c] = tl[c];
            }
            lazy_butterfly<Q, INV>(x, w, has_tw, P);
#pragma unroll
            for (int t = 0; t < Q; ++t) arr[pi][idx[t]] = x[t];
        }
    }
    __syncthreads();
}

template <bool INV>
__device__ __forceinline__ void r_lds_dft3(u32* const* arr, int L, int tlog, const u32* __restrict__ tw0,
                                           const Primes3& PS) {
    const int r0 = L % 3;
    if (!INV) {
        int logM = L;
        if (r0 == 1) { r_lds_level3<2, false>(arr, L, logM, tlog, tw0, PS); logM -= 1; }
        if (r0 == 2) { r_lds_level3<4, false>(arr, L, logM, tlog, tw0, PS); logM -= 2; }
        for (; logM >= 3; logM -= 3) r_lds_level3<8, false>(arr, L, logM, tlog, tw0, PS);
    } else {
        for (int logM = 3; logM <= L - r0; logM += 3) r_lds_level3<8, true>(arr, L, logM, tlog, tw0, PS);
        if (r0 == 1) r_lds_level3<2, true>(arr, L, L, tlog, tw0, PS);
        if (r0 == 2) r_lds_level3<4, true>(arr, L, L, tlog, tw0, PS);
    }
}

// The prime of a (wave-uniform) index, selected field by field so the
// kernel-argument struct is never indexed dynamically.
__device__ __forceinline__ Prime3 pick3(const Primes3& PS, int pi) {
    return pi == 0 ? PS.p[0] : (pi == 1 ? PS.p[1] : PS.p[2]);
}

// A limb below 2^32 < 5p brought into [0, 2p): [0, 5p) -> [0, 3p) -> [0, 2p).
__device__ __forceinline__ u32 r_reduce_limb(u32 v, u32 p) {
    const u32 p2 = p << 1;
    v = umin32(v, v - p2);
    return umin32(v, v - p2);
}

// ---- column pass, forward ---------------------------------------------------------
// Workgroup -> (job, operand, tile): the tile's limbs are read once and
// reduced modulo each prime into three LDS arrays, transformed together. The
// residue spectra of an operand sit in its workspace as three consecutive
// arrays of 2^n3 u32.
template <int T>
__global__ __launch_bounds__(1024) void k3_col_fwd(const MulJob* __restrict__ jobs, MulJob one, int ops, int n3,
                                                  int L1, const u32* __restrict__ tw3, Primes3 PS) {
    extern __shared__ u32 lds3[];
    const int L2 = n3 - L1;
    const size_t S = (size_t)1 << L2;
    const int R = 1 << L1;
    const unsigned tiles = (unsigned)(S / T);
    const unsigned bid = xcd_tile_order(blockIdx.x, gridDim.x);
    const unsigned job = bid / ((unsigned)ops * tiles);
    const unsigned rem = bid % ((unsigned)ops * tiles);
    const int op = (int)(rem / tiles);
    const size_t b0 = (size_t)(rem % tiles) * T;
    const MulJob J = job_of(jobs, job, one);
    const u32* __restrict__ src = op ? J.b : J.a;
    const size_t nl = op ? J.nb : J.na;
    u32* __restrict__ dst = (u32*)(op ? J.wsB : J.wsA);
    const int words = lds_words(R * T);
    for (int e = threadIdx.x; e < R * T; e += blockDim.x) {
        const int col = e & (T - 1), i = e / T;
        const size_t g = (size_t)i * S + b0 + col;
        const u32 v = g < nl ? src[g] : 0u;
        lds3[lpad(e)] = r_reduce_limb(v, PS.p[0].p);
        lds3[words + lpad(e)] = r_reduce_limb(v, PS.p[1].p);
        lds3[2 * words + lpad(e)] = r_reduce_limb(v, PS.p[2].p);
    }
    __syncthreads();
    u32* arr[3] = {lds3, lds3 + words, lds3 + 2 * words};
    r_lds_dft3<false>(arr, L1, QLogT<T>::v, tw3, PS);
    for (int e = threadIdx.x; e < R * T; e += blockDim.x) {
        const int col = e & (T - 1), i = e / T;
        const size_t g = (size_t)i * S + b0 + col;
        dst[g] = lds3[lpad(e)];
        dst[((size_t)1 << n3) + g] = lds3[words + lpad(e)];
        dst[((size_t)2 << n3) + g] = lds3[2 * words + lpad(e)];
    }
}

// ---- row pass ---------------------------------------------------------------------
// Workgroup -> (job, prime, row). Row r holds column frequency c = freq(r);
// element b is multiplied by w^(b c) on the way in and by w^(-b c) s on the
// way out, s = 2^-n3 R mod p (R undoes the pointwise Montgomery factor).
// Outer tables per prime: hi | lo | hi_inv | lo_inv, w^e R = mm(hi[e >> h], lo[e & (2^h - 1)]).
__device__ __forceinline__ u32 r_tw(const u32* __restrict__ hi, const u32* __restrict__ lo, int h, u64 e,
                                    const Prime3& P) {
    return mm(hi[e >> h], lo[e & (((u64)1 << h) - 1)], P.p, P.ninv);
}

template <bool SQUARE>
__global__ __launch_bounds__(1024) void k3_row_mul(const MulJob* __restrict__ jobs, MulJob one, int L1, int L2,
                                                  int lr_col, const u32* __restrict__ tw3,
                                                  const u32* __restrict__ o3, int h, Primes3 PS) {
    extern __shared__ u32 lds3[];
    const int R = 1 << L2;
    const int n3 = L1 + L2;
    u32* la = lds3;
    u32* lb = lds3 + lds_words(R);
    const unsigned rows = 1u << L1;
    const unsigned job = blockIdx.x / (3u * rows);
    const int pi = (int)((blockIdx.x / rows) % 3u);
    const int row = (int)(blockIdx.x % rows);
    const MulJob J = job_of(jobs, job, one);
    const Prime3 P = pick3(PS, pi);
    u32* __restrict__ A = (u32*)J.wsA + ((size_t)pi << n3);
    const u32* __restrict__ B = (const u32*)J.wsB + ((size_t)pi << n3);
    const size_t nhi = (size_t)1 << (n3 - h), nlo = (size_t)1 << h;
    const u32* hi = o3 + (size_t)pi * 2 * (nhi + nlo);
    const u32* lo = hi + nhi;
    const u32* hi_inv = lo + nlo;
    const u32* lo_inv = hi_inv + nhi;
    const u32* tw = tw3 + (size_t)pi * TW3_STRIDE;
    const size_t base = (size_t)row * R;
    const u64 c = freq_rt(row, L1, lr_col);
    const u64 N = (u64)1 << n3;
    {
        u32 w = r_tw(hi, lo, h, (c * threadIdx.x) & (N - 1), P);
        const u32 step = r_tw(hi, lo, h, (c * blockDim.x) & (N - 1), P);
        for (int e = threadIdx.x; e < R; e += blockDim.x) {
            la[lpad(e)] = mml(A[base + e], w, P.p, P.ninv);
            if constexpr (!SQUARE) lb[lpad(e)] = mml(B[base + e], w, P.p, P.ninv);
            w = mml(w, step, P.p, P.ninv);
        }
    }
    __syncthreads();
    if constexpr (SQUARE) {
        u32* arr[1] = {la};
        r_lds_dft<false, 1>(arr, L2, 0, tw, P);
        for (int e = threadIdx.x; e < R; e += blockDim.x) { const u32 v = la[lpad(e)]; la[lpad(e)] = mml(v, v, P.p, P.ninv); }
    } else {
        u32* arr[2] = {la, lb};
        r_lds_dft<false, 2>(arr, L2, 0, tw, P);
        for (int e = threadIdx.x; e < R; e += blockDim.x) la[lpad(e)] = mml(la[lpad(e)], lb[lpad(e)], P.p, P.ninv);
    }
    __syncthreads();
    u32* arr1[1] = {la};
    r_lds_dft<true, 1>(arr1, L2, 0, tw, P);
    {
        u32 w = r_tw(hi_inv, lo_inv, h, (c * threadIdx.x) & (N - 1), P);
        const u32 step = r_tw(hi, lo, h, (N - c * blockDim.x) & (N - 1), P);
        for (int e = threadIdx.x; e < R; e += blockDim.x) {
            A[base + e] = mml(la[lpad(e)], w, P.p, P.ninv);
            w = mml(w, step, P.p, P.ninv);
        }
    }
}

// ---- register-resident row pass (radix 16) -------------------------------------------
// A row of R = 2^L2 elements is spread over R / 16 threads, 16 elements each,
// held in registers. The transform runs in "sets": in set B a thread holds the
// 16 positions ((u >> B) << (B + 4)) | (t << B) | (u & (2^B - 1)), t < 16, and
// applies one radix-16 step at M = 2^(B + 4) (four radix-2 levels) to them.
// Sets run at B = L2 - 4, L2 - 8, ... and a last set at B = 0, which applies
// the remaining L2 mod 4 levels (radix 2, 4 or 8, no twiddles) when L2 is not
// a multiple of 4. Between sets the values pass through LDS (write in the old
// set's positions, barrier, read in the new set's): one barrier and one LDS
// round trip per four levels, where level-at-a-time passes take one per three
// levels plus the loads, stores and pointwise product through LDS. Set
// L2 - 4 is t R/16 + u, coalesced in global memory, so the outer twiddle is
// applied on load and store. A and B are transformed together (one barrier
// per exchange for both), the pointwise product is formed in registers and
// the inverse runs the sets backwards from where the forward ended.
template <int L2>
struct RowPlan {
    static constexpr int TPR = 1 << (L2 - 4);           // threads per row
    static constexpr int WG = TPR < 256 ? 256 : TPR;    // threads per workgroup
    static constexpr int G = WG / TPR;                  // rows per workgroup
    static constexpr int REM = L2 % 4;                  // levels of the last, partial set
    static constexpr int WORDS = (1 << L2) + (1 << (L2 - 4));   // one pad word per 16 (RowLay)
    static constexpr int NSET = L2 / 4 + (REM ? 1 : 0);
    // bottom bit of set s: L2 - 4 (s + 1) for the full sets, 0 for the partial one
    static constexpr int bottom(int s) { return (REM && s == NSET - 1) ? 0 : L2 - 4 * (s + 1); }
    static constexpr bool partial(int s) { return REM && s == NSET - 1; }
};

// Positions of set B of radix 2^LQ in a sub-transform whose element i sits at
// LDS word lpad((i << LOGT) | col): i(t) = ((u >> B) << (B + LQ)) | (t << B) |
// (u mod 2^B). The offsets of t > 0 from t = 0 are compile-time constants.
// LDS layouts: word e at e + K (e >> A), K pad words per 2^A words. The map
// is additive over bit-disjoint parts, so a set's offsets from its first
// word are compile-time constants. (A, K) = (3, 1) is the layout of the
// column kernels.
template <int A, int K> __device__ __forceinline__ int padak(int e) { return e + K * (e >> A); }

template <int B, int LQ, int LOGT, int A = 3, int K = 1>
__device__ __forceinline__ int set_base(int u, int col) {
    return padak<A, K>(((((u >> B) << (B + LQ)) | (u & ((1 << B) - 1))) << LOGT) | col);
}

template <int B, int LQ, int LOGT, int A = 3, int K = 1>
__device__ __forceinline__ void set_store(u32* __restrict__ l, int a0, const u32* x) {
#pragma unroll
    for (int t = 0; t < (1 << LQ); ++t) l[a0 + padak<A, K>((t << B) << LOGT)] = x[t];
}

template <int B, int LQ, int LOGT, int A = 3, int K = 1>
__device__ __forceinline__ void set_load(const u32* __restrict__ l, int a0, u32* x) {
#pragma unroll
    for (int t = 0; t < (1 << LQ); ++t) x[t] = l[a0 + padak<A, K>((t << B) << LOGT)];
}

// The row kernel's layout for the exchange between sets with bottoms B0 and
// B1 (either order). Every layout has one pad word per 16 on average, so a
// wavefront's block of 1,024 points always occupies the same 1,088 words
// (the wavefront-local exchanges rely on that); within the block the pads
// sit where they keep both sides of the exchange free of bank conflicts: a
// set's access for one t touches 64 words per wavefront, and with the
// exchange's layout they fall in 64 distinct banks (and each half-wavefront's
// 32 in 32 distinct banks). Lane bits map to point bits 0-5 for bottoms >= 6,
// to 0..B-1 and B+4..9 for bottoms B < 6; four pad words per 64 suit every
// pair whose lower bottom is at least 2, two per 32 a lower bottom of 1, one
// per 16 a lower bottom of 0 (the partial set). A single layout cannot serve
// all pairs: one pad per 8 words, the column kernels' layout, leaves every
// row access 2-way conflicted.
template <int B0, int B1>
struct RowLay {
    static constexpr int LO = B0 < B1 ? B0 : B1;
    static constexpr int A = LO == 0 ? 4 : LO == 1 ? 5 : 6;
    static constexpr int K = 1 << (A - 4);
};

// Twiddles of set S's radix-16 step: w_M^(j c), M = 2^(B + 4), j = u mod 2^B.
template <int L2, int S, bool INV>
__device__ __forceinline__ void set_tw(u32* w, int u, const u32* __restrict__ tw) {
    constexpr int B = RowPlan<L2>::bottom(S);
    if constexpr (!RowPlan<L2>::partial(S) && B > 0) {
        const uint4* tl = (const uint4*)(tw_level<16, INV>(tw, B + 4) + (u & ((1 << B) - 1)) * 16);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint4 v = tl[k];
            w[4 * k] = v.x; w[4 * k + 1] = v.y; w[4 * k + 2] = v.z; w[4 * k + 3] = v.w;
        }
    }
}

// Set S's step: the radix-16 butterfly, or for the partial set the
// 2^(4 - REM) butterflies of radix 2^REM over the low bits.
// The inverse runs its sets from the last one down to set 0: every set but
// set 0 leaves its outputs in [0, 4p) for the next set's twiddle products and
// input-0 reduction; set 0 reduces its outputs for the output twiddle, whose
// running power is only in [0, 2p). (The partial set is the inverse's first,
// so its inputs, the pointwise products, are in [0, 2p).)
template <int L2, int S, bool INV>
__device__ __forceinline__ void set_step(u32* x, const u32* w, const Prime3& P) {
    using RP = RowPlan<L2>;
    constexpr bool ROUT = !INV || S == 0;
    if constexpr (RP::partial(S)) {
        constexpr int Q = 1 << RP::REM;
#pragma unroll
        for (int hq = 0; hq < 16 / Q; ++hq) lazy_butterfly<Q, INV, INV, ROUT>(x + hq * Q, w, false, P);
    } else {
        lazy_butterfly<16, INV, INV, ROUT>(x, w, RP::bottom(S) > 0, P);
    }
}

// The exchange between sets with bottoms B0 and B1. A wavefront's 64
// consecutive threads u = 64 g + lane hold, in a set of bottom B <= 6, the
// positions ((u >> B) << (B + 4)) | (t << B) | (u mod 2^B) = 1024 g + ... :
// exactly the aligned block [1024 g, 1024 g + 1024) of its row, whatever
// B <= 6 is. An exchange between two such sets therefore stays inside the
// wavefront: its LDS writes and reads touch only its own block, which no
// other wavefront reads or writes until the next workgroup barrier, and a
// wavefront's LDS operations execute in order, so only the compiler must be
// kept from moving them (a wave barrier). The same holds when a whole row
// fits in one wavefront (TPR <= 64). Other exchanges need the workgroup
// barrier.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int L2, int B0, int B1>
__device__ __forceinline__ void row_exchange_sync() {
    if constexpr (RowPlan<L2>::TPR <= 64 || (B0 <= 6 && B1 <= 6)) wave_lds_sync();
    else __syncthreads();
}

// Forward of NOP operands (1 or 2, arrays l[0], l[1]): x holds set 0 on entry,
// the spectrum in the last set's positions on exit.
// STEP = false: the caller has already applied set S's step.
template <int L2, int NOP, int S = 0, bool STEP = true>
__device__ __forceinline__ void row_fwd(u32 (*x)[16], const u32* w, int u, u32* const* l, const u32* __restrict__ tw,
                                        const Prime3& P) {
    using RP = RowPlan<L2>;
    if constexpr (STEP) {
#pragma unroll
        for (int k = 0; k < NOP; ++k) set_step<L2, S, false>(x[k], w, P);
    }
    if constexpr (S + 1 < RP::NSET) {
        // the next set's twiddles are in flight during the exchange
        u32 wn[16];
        set_tw<L2, S + 1, false>(wn, u, tw);
        using XL = RowLay<RP::bottom(S), RP::bottom(S + 1)>;
        const int a0 = set_base<RP::bottom(S), 4, 0, XL::A, XL::K>(u, 0);
        const int a1 = set_base<RP::bottom(S + 1), 4, 0, XL::A, XL::K>(u, 0);
#pragma unroll
        for (int k = 0; k < NOP; ++k) set_store<RP::bottom(S), 4, 0, XL::A, XL::K>(l[k], a0, x[k]);
        row_exchange_sync<L2, RP::bottom(S), RP::bottom(S + 1)>();
#pragma unroll
        for (int k = 0; k < NOP; ++k) set_load<RP::bottom(S + 1), 4, 0, XL::A, XL::K>(l[k], a1, x[k]);
        row_fwd<L2, NOP, S + 1>(x, wn, u, l, tw, P);
    }
}

// Inverse: x holds the last set on entry, set 0 (natural order) on exit.
template <int L2, int S = RowPlan<L2>::NSET - 1>
__device__ __forceinline__ void row_inv(u32* x, const u32* w, int u, u32* __restrict__ l,
                                        const u32* __restrict__ tw, const Prime3& P) {
    using RP = RowPlan<L2>;
    set_step<L2, S, true>(x, w, P);
    if constexpr (S > 0) {
        u32 wn[16];
        set_tw<L2, S - 1, true>(wn, u, tw);
        using XL = RowLay<RP::bottom(S), RP::bottom(S - 1)>;
        set_store<RP::bottom(S), 4, 0, XL::A, XL::K>(l, set_base<RP::bottom(S), 4, 0, XL::A, XL::K>(u, 0), x);
        row_exchange_sync<L2, RP::bottom(S), RP::bottom(S - 1)>();
        set_load<RP::bottom(S - 1), 4, 0, XL::A, XL::K>(l, set_base<RP::bottom(S - 1), 4, 0, XL::A, XL::K>(u, 0), x);
        row_inv<L2, S - 1>(x, wn, u, l, tw, P);
    }
}

// The register budget is set for four waves per SIMD (128 VGPRs): at
// 2^13 points a row's LDS lets two 512-thread workgroups share a CU, which
// more registers would prevent.
template <int L2, bool SQUARE>
__global__ __launch_bounds__(RowPlan<L2>::WG, 4) void k3_row_reg(const MulJob* __restrict__ jobs, MulJob one, int L1,
                                                             int lr_col, const u32* __restrict__ tw3,
                                                             const u32* __restrict__ o3, int h, Primes3 PS) {
    using RP = RowPlan<L2>;
    constexpr int NOP = SQUARE ? 1 : 2;
    constexpr int R = 1 << L2;
    constexpr int TPR = RP::TPR;
    extern __shared__ u32 lds3[];
    const int n3 = L1 + L2;
    const int g = threadIdx.x / TPR;
    const int u = threadIdx.x % TPR;
    u32* l[2] = {lds3 + g * RP::WORDS, lds3 + (RP::G + g) * RP::WORDS};
    const unsigned wgs = (1u << L1) / RP::G;                       // workgroups per prime
    const unsigned job = blockIdx.x / (3u * wgs);
    const int pi = (int)((blockIdx.x / wgs) % 3u);
    const int row = (int)(blockIdx.x % wgs) * RP::G + g;
    const MulJob J = job_of(jobs, job, one);
    const Prime3 P = pick3(PS, pi);
    u32* __restrict__ A = (u32*)J.wsA + ((size_t)pi << n3) + (size_t)row * R;
    const u32* __restrict__ Bv = (const u32*)J.wsB + ((size_t)pi << n3) + (size_t)row * R;
    const size_t nhi = (size_t)1 << (n3 - h), nlo = (size_t)1 << h;
    const u32* hi = o3 + (size_t)pi * 2 * (nhi + nlo);
    const u32* lo = hi + nhi;
    const u32* hi_inv = lo + nlo;
    const u32* lo_inv = hi_inv + nhi;
    const u32* tw = tw3 + (size_t)pi * TW3_STRIDE;
    const u64 c = freq_rt(row, L1, lr_col);
    const u64 N = (u64)1 << n3;
    u32 x[NOP][16];
    // Every load of the prologue is issued before any arithmetic, in the order
    // the values are used (loads complete in issue order, so each use waits
    // only for the loads issued before it): the outer-twiddle table entries
    // and set 0's level twiddles, A's points (set 0, b = t R/16 + u), B's
    // points, and the inverse's outer-twiddle entries, which would otherwise
    // be fetched after the last exchange with nothing left to overlap them.
    // A's outer twiddle and set-0 step then run while B's points arrive.
    const u64 e0 = (c * (u64)u) & (N - 1), lmask = ((u64)1 << h) - 1;
    const u32 fhi = gld(hi + (e0 >> h)), flo = gld(lo + (e0 & lmask));
    u32 w0[16];
    set_tw<L2, 0, false>(w0, u, tw);
    sched_fence();
#pragma unroll
    for (int t = 0; t < 16; ++t) x[0][t] = gld(A + t * TPR + u);
    sched_fence();
    if constexpr (NOP == 2) {
#pragma unroll
        for (int t = 0; t < 16; ++t) x[1][t] = gld(Bv + t * TPR + u);
    }
    const u32 ihi = gld(hi_inv + (e0 >> h)), ilo = gld(lo_inv + (e0 & lmask));
    sched_fence();
    // outer twiddles w^(b c)
    u32 wt[16];
    {
        u32 w = mm(fhi, flo, P.p, P.ninv);
        const u32 step = r_tw(hi, lo, h, (c * (u64)TPR) & (N - 1), P);
#pragma unroll
        for (int t = 0; t < 16; ++t) { wt[t] = w; w = mml(w, step, P.p, P.ninv); }
    }
#pragma unroll
    for (int k = 0; k < NOP; ++k) {
#pragma unroll
        for (int t = 0; t < 16; ++t) x[k][t] = mml(x[k][t], wt[t], P.p, P.ninv);
        set_step<L2, 0, false>(x[k], w0, P);
        sched_fence();
    }
    const u32 winv = mm(ihi, ilo, P.p, P.ninv);
    row_fwd<L2, NOP, 0, false>(x, w0, u, l, tw, P);
#pragma unroll
    for (int t = 0; t < 16; ++t) x[0][t] = mml(x[0][t], x[NOP - 1][t], P.p, P.ninv);
    // the inverse's first exchange writes the positions this thread read last
    {
        u32 w[16];
        set_tw<L2, RP::NSET - 1, true>(w, u, tw);
        row_inv<L2>(x[0], w, u, l[0], tw, P);
    }
    u32 w = winv;
    const u32 step = r_tw(hi, lo, h, (N - c * (u64)TPR) & (N - 1), P);
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        gst(A + t * TPR + u, mml(x[0][t], w, P.p, P.ninv));
        w = mml(w, step, P.p, P.ninv);
    }
}

// ---- inverse column pass, CRT and the in-segment carry ------------------------------
// Garner: x0 = r0, x1 = (r1 - x0) p0^-1 mod p1, x2 = (r2 - x0 - x1 p0)(p0 p1)^-1
// mod p2, X = x0 + x1 p0 + x2 p0 p1 < p0 p1 p2, returned as lo32 + up * 2^32.
__device__ __forceinline__ void crt3(u32 r0, u32 r1, u32 r2, const Primes3& PS, u32& lo32, u64& up) {
    const Prime3& P1 = PS.p[1];
    const Prime3& P2 = PS.p[2];
    const u32 x0 = r0;                                              // < p0 < p1 < p2
    const u32 x1 = mm(r_sub(r1, x0, P1.p), PS.inv_p0_m1, P1.p, P1.ninv);
    const u32 t = r_sub(r_sub(r2, x0, P2.p), mm(x1, PS.p0_m2, P2.p, P2.ninv), P2.p);
    const u32 x2 = mm(t, PS.inv_p0p1_m2, P2.p, P2.ninv);
    const u64 a = (u64)x0 + (u64)x1 * PS.p[0].p;                   // < 2^62
    const u64 b = a + (u64)x2 * (u32)PS.p0p1;                       // < 2^62 + 2^63
    lo32 = (u32)b;
    up = (b >> 32) + (u64)x2 * (u32)(PS.p0p1 >> 32);
}

template <int T>
__global__ __launch_bounds__(1024) void k3_col_inv(const MulJob* __restrict__ jobs, MulJob one, int n3, int L1,
                                                  const u32* __restrict__ tw3, Primes3 PS, size_t nblk_max) {
    extern __shared__ u32 lds3[];
    const int L2 = n3 - L1;
    const size_t S = (size_t)1 << L2;
    const int R = 1 << L1;
    const unsigned tiles = (unsigned)(S / T);
    const unsigned bid = xcd_tile_order(blockIdx.x, gridDim.x);
    const MulJob J = job_of(jobs, bid / tiles, one);
    const size_t b0 = (size_t)(bid % tiles) * T;
    const int words = lds_words(R * T);
    for (int pi = 0; pi < 3; ++pi) {
        const u32* __restrict__ buf = (const u32*)J.wsA + ((size_t)pi << n3);
        u32* l = lds3 + pi * words;
        for (int e = threadIdx.x; e < R * T; e += blockDim.x) {
            const int col = e & (T - 1), i = e / T;
            l[lpad(e)] = buf[(size_t)i * S + b0 + col];
        }
    }
    __syncthreads();
    u32* arr[3] = {lds3, lds3 + words, lds3 + 2 * words};
    r_lds_dft3<true>(arr, L1, QLogT<T>::v, tw3, PS);
    // chunk = the T coefficients (= limbs) of one row segment; 2T is the
    // carry geometry's coefficient count for T limbs
    const CarryJob C = carry_job<2 * T>(J, nblk_max);
    for (int i = threadIdx.x; i < R; i += blockDim.x) {
        const size_t k0 = (size_t)i * S + b0;
        const size_t ch = k0 / T;
        if (ch >= C.nchunks) continue;
        u64 carry = 0;
#pragma unroll
        for (int q = 0; q < T; ++q) {
            const int e = lpad(i * T + q);
            u32 lo32; u64 up;
            crt3(umin32(lds3[e], lds3[e] - PS.p[0].p), umin32(lds3[words + e], lds3[words + e] - PS.p[1].p),
                 umin32(lds3[2 * words + e], lds3[2 * words + e] - PS.p[2].p), PS, lo32, up);
            const u64 s = (carry & 0xFFFFFFFFull) + lo32;
            if (k0 + q < C.nlimbs) C.out[k0 + q] = (u32)s;
            carry = (s >> 32) + (carry >> 32) + up;
        }
        C.C[ch] = carry;
    }
}

// ---- register-resident column passes (radix-8 sets) ---------------------------------
// A tile of T columns x R = 2^L1 rows, the three primes in turn through one
// LDS tile; a thread holds 8 rows of one column, WG = R/8 * T threads. Sets
// as in the row kernel but of radix 8 with the partial set first: set 0 holds
// rows t R/8 + u and applies the top L1 mod 3 levels (radix 2 or 4, twiddled)
// when L1 is not a multiple of 3, then radix-8 sets at bottoms
// L1 - L1 mod 3 - 3 s down to 0. That is the level order of the LDS-level
// column pass, so the frequency order (and the row kernel's c) is the same.
// Set 0 is read from (forward) or ends in (inverse) rows t R/8 + u, the last
// set rows 8u + t; both go to and from global memory directly, and the LDS
// only carries the exchanges between sets (and, inverse, the CRT values to
// the per-row carry).
template <int L1, int T>
struct ColPlan {
    static constexpr int LOGT = QLogT<T>::v;
    static constexpr int WG = (1 << (L1 - 3)) * T;
    static constexpr int REM = L1 % 3;
    static constexpr int NSET = L1 / 3 + (REM ? 1 : 0);
    static constexpr bool partial(int s) { return REM && s == 0; }
    static constexpr int bottom(int s) { return REM ? (s == 0 ? L1 - 3 : L1 - REM - 3 * s) : L1 - 3 * (s + 1); }
    // LDS layout of the tile (element (i, col) is word (i << LOGT) | col):
    // one pad word per 8. (Eight per 64 make every set access free of bank
    // conflicts, but the inverse pass's carry, one row segment of T words
    // per thread, then conflicts 8-way: 79 against 71 us at 2^9 x 16, and
    // the forward pass does not get faster.)
    static constexpr int PA = 3, PK = 1;
};

// Tiles the register column kernels are built for: at least one wave and at
// most 1,024 threads (R T / 8 of them).
constexpr bool col_reg_pair(int L1, int T) {
    const int lt = T == 4 ? 2 : T == 8 ? 3 : T == 16 ? 4 : T == 32 ? 5 : 99;
    return L1 >= 3 && L1 + lt >= 9 && L1 + lt <= 13;
}

// Level twiddles of a full radix-8 set S with bottom B > 0: w_M^(j c),
// M = 2^(B + 3), j = u mod 2^B, c = 1..7 (the partial set loads its own).
template <int L1, int T, int S, bool INV>
__device__ __forceinline__ void col_tw(u32* w, int u, const u32* __restrict__ tw) {
    using CP = ColPlan<L1, T>;
    if constexpr (!CP::partial(S) && CP::bottom(S) > 0) {
        constexpr int B = CP::bottom(S);
        const uint4* tl = (const uint4*)(tw_level<8, INV>(tw, B + 3) + (u & ((1 << B) - 1)) * 8);
        const uint4 v0 = tl[0], v1 = tl[1];
        w[1] = v0.y; w[2] = v0.z; w[3] = v0.w; w[4] = v1.x; w[5] = v1.y; w[6] = v1.z; w[7] = v1.w;
    }
}

template <int L1, int T, int S, bool INV>
__device__ __forceinline__ void col_step(u32* x, const u32* w, int u, const u32* __restrict__ tw, const Prime3& P) {
    using CP = ColPlan<L1, T>;
    if constexpr (CP::partial(S)) {
        // radix Q = 2^REM over the top bits: t = q 2^(3 - REM) + tlo, so the
        // butterfly of j = tlo R/8 + u takes x[q TL + tlo], q < Q
        constexpr int Q = 1 << CP::REM, TL = 8 / Q;
#pragma unroll
        for (int tlo = 0; tlo < TL; ++tlo) {
            const int j = tlo * (1 << (L1 - 3)) + u;
            const u32* __restrict__ tl = tw_level<Q, INV>(tw, L1) + j * Q;
            u32 y[Q], w[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) y[q] = x[q * TL + tlo];
#pragma unroll
            for (int c = 1; c < Q; ++c) w[c] = tl[c];
            lazy_butterfly<Q, INV, INV, true>(y, w, true, P);   // partial set: the inverse's last
#pragma unroll
            for (int q = 0; q < Q; ++q) x[q * TL + tlo] = y[q];
        }
    } else {
        // inverse: sets before set 0 leave outputs in [0, 4p) for the next
        // set's twiddle products and input-0 reduction; set 0 (the CRT's
        // input) reduces them
        lazy_butterfly<8, INV, INV, !INV || S == 0>(x, w, CP::bottom(S) > 0, P);
    }
}

// One prime's forward column sets: x holds set 0 (rows t R/8 + u) on entry,
// the last set (rows 8u + t) on exit; l is the tile's single LDS array.
template <int L1, int T, int S = 0>
__device__ __forceinline__ void col_fwd1(u32* x, const u32* w, int u, int col, u32* __restrict__ l,
                                         const u32* __restrict__ tw, const Prime3& P) {
    using CP = ColPlan<L1, T>;
    col_step<L1, T, S, false>(x, w, u, tw, P);
    if constexpr (S + 1 < CP::NSET) {
        u32 wn[8];
        col_tw<L1, T, S + 1, false>(wn, u, tw);
        constexpr int B0 = CP::bottom(S), B1 = CP::bottom(S + 1);
        set_store<B0, 3, CP::LOGT, CP::PA, CP::PK>(l, set_base<B0, 3, CP::LOGT, CP::PA, CP::PK>(u, col), x);
        __syncthreads();
        set_load<B1, 3, CP::LOGT, CP::PA, CP::PK>(l, set_base<B1, 3, CP::LOGT, CP::PA, CP::PK>(u, col), x);
        col_fwd1<L1, T, S + 1>(x, wn, u, col, l, tw, P);
    }
}

// A limb below 2^32 as a residue in [0, 2p) by one Montgomery product with
// R mod p (w8[0] is w8^0 R): v (R mod p) < 2^32 p, so the product is below 2p.
__device__ __forceinline__ u32 r_limb_mont(u32 v, const Prime3& P) { return mml(v, P.w8[0], P.p, P.ninv); }

// Forward column pass with the primes in sequence through one LDS tile: a
// third of the LDS of the three tiles at once, so two workgroups share a CU (at most
// 64 VGPRs) and one's loads, stores and barriers overlap the other's
// arithmetic. The limbs are loaded once and reduced per prime.
template <int L1, int T>
__global__ __launch_bounds__((ColPlan<L1, T>::WG), (ColPlan<L1, T>::WG >= 1024 ? 8 : 4))
void k3_col_fwd_seq(const MulJob* __restrict__ jobs, MulJob one, int ops, int n3, const u32* __restrict__ tw3,
                    Primes3 PS) {
    using CP = ColPlan<L1, T>;
    extern __shared__ u32 lds3[];
    constexpr int R = 1 << L1;
    const size_t S = (size_t)1 << (n3 - L1);
    const unsigned tiles = (unsigned)(S / T);
    const unsigned bid = xcd_tile_order(blockIdx.x, gridDim.x);
    const unsigned job = bid / ((unsigned)ops * tiles);
    const unsigned rem = bid % ((unsigned)ops * tiles);
    const int op = (int)(rem / tiles);
    const size_t b0 = (size_t)(rem % tiles) * T;
    const MulJob J = job_of(jobs, job, one);
    const u32* __restrict__ src = op ? J.b : J.a;
    const size_t nl = op ? J.nb : J.na;
    u32* __restrict__ dst = (u32*)(op ? J.wsB : J.wsA);
    const int col = threadIdx.x & (T - 1), u = threadIdx.x >> CP::LOGT;
    u32 v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const size_t g = (size_t)(t * (R / 8) + u) * S + b0 + col;
        v[t] = g < nl ? gld(src + g) : 0u;
    }
    sched_fence();
    u32 w[8];
    col_tw<L1, T, 0, false>(w, u, tw3);
    sched_fence();
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const Prime3& P = PS.p[p];
        const u32* tw = tw3 + (size_t)p * TW3_STRIDE;
        u32 x[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) x[t] = r_limb_mont(v[t], P);
        // the previous prime's last exchange has been read everywhere before
        // this prime's first exchange writes the tile
        if (p > 0) __syncthreads();
        col_fwd1<L1, T>(x, w, u, col, lds3, tw, P);
        if (p < 2) col_tw<L1, T, 0, false>(w, u, tw3 + (size_t)(p + 1) * TW3_STRIDE);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const size_t g = (size_t)((u << 3) | t) * S + b0 + col;
            gst(dst + ((size_t)p << n3) + g, x[t]);
        }
    }
}

// One prime's inverse column sets: x holds the last set (rows 8u + t) on
// entry, set 0 (rows t R/8 + u) on exit; l is the tile's single LDS array.
template <int L1, int T, int S = ColPlan<L1, T>::NSET - 1>
__device__ __forceinline__ void col_inv1(u32* x, const u32* w, int u, int col, u32* __restrict__ l,
                                         const u32* __restrict__ tw, const Prime3& P) {
    using CP = ColPlan<L1, T>;
    col_step<L1, T, S, true>(x, w, u, tw, P);
    if constexpr (S > 0) {
        u32 wn[8];
        col_tw<L1, T, S - 1, true>(wn, u, tw);
        constexpr int B0 = CP::bottom(S), B1 = CP::bottom(S - 1);
        set_store<B0, 3, CP::LOGT, CP::PA, CP::PK>(l, set_base<B0, 3, CP::LOGT, CP::PA, CP::PK>(u, col), x);
        __syncthreads();
        set_load<B1, 3, CP::LOGT, CP::PA, CP::PK>(l, set_base<B1, 3, CP::LOGT, CP::PA, CP::PK>(u, col), x);
        col_inv1<L1, T, S - 1>(x, wn, u, col, l, tw, P);
    }
}

// LDS words of k3_col_inv_seq: the exchange tile, or the carry's three
// arrays over half the rows, whichever is larger.
__host__ __device__ constexpr int col_inv_seq_words(int L1, int T) {
    return lds_words((1 << L1) * T) > 3 * lds_words((1 << (L1 - 1)) * T) ? lds_words((1 << L1) * T)
                                                                            : 3 * lds_words((1 << (L1 - 1)) * T);
}

// Inverse column pass with the primes in sequence through one LDS tile (as
// k3_col_fwd_seq), each prime's set 0 kept in registers; then the CRT in
// registers and each row segment's carry through LDS in two halves of the
// rows (rows t R/8 + u with t < 4, then t >= 4), so that the LDS is about
// half of what the three residue tiles at once would take, and two
// workgroups share a CU.
template <int L1, int T>
__global__ __launch_bounds__((ColPlan<L1, T>::WG), (ColPlan<L1, T>::WG >= 1024 ? 8 : 4))
void k3_col_inv_seq(const MulJob* __restrict__ jobs, MulJob one, int n3, const u32* __restrict__ tw3, Primes3 PS,
                    size_t nblk_max) {
    using CP = ColPlan<L1, T>;
    extern __shared__ u32 lds3[];
    constexpr int R = 1 << L1, RH = R / 2;
    constexpr int hwords = lds_words(RH * T);
    const size_t S = (size_t)1 << (n3 - L1);
    const unsigned tiles = (unsigned)(S / T);
    const unsigned bid = xcd_tile_order(blockIdx.x, gridDim.x);
    const MulJob J = job_of(jobs, bid / tiles, one);
    const size_t b0 = (size_t)(bid % tiles) * T;
    const int col = threadIdx.x & (T - 1), u = threadIdx.x >> CP::LOGT;
    // each prime's residues are loaded when its turn comes (all three up
    // front would not fit the register budget of two workgroups per CU)
    u32 x[3][8];
    u32 w[8];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const size_t g = (size_t)((u << 3) | t) * S + b0 + col;
            x[p][t] = gld((const u32*)J.wsA + ((size_t)p << n3) + g);
        }
        col_tw<L1, T, CP::NSET - 1, true>(w, u, tw3 + (size_t)p * TW3_STRIDE);
        sched_fence();
        // the previous prime's last exchange has been read everywhere before
        // this prime's first exchange writes the tile
        if (p > 0) __syncthreads();
        col_inv1<L1, T>(x[p], w, u, col, lds3, tw3 + (size_t)p * TW3_STRIDE, PS.p[p]);
    }
    // CRT of each coefficient (row t R/8 + u, column col) into (lo32, up)
    u32 lo[8];
    u64 up[8];
#pragma unroll
    for (int t = 0; t < 8; ++t)
        crt3(umin32(x[0][t], x[0][t] - PS.p[0].p), umin32(x[1][t], x[1][t] - PS.p[1].p),
             umin32(x[2][t], x[2][t] - PS.p[2].p), PS, lo[t], up[t]);
    // chunk = the T coefficients (= limbs) of one row segment, carry-in 0;
    // the chunk's limbs go back to array 0 in place. With the carry-out C the
    // chunk's threshold goes out too (2^(32T) - D when limbs 2..T-1 of D are
    // all ones, else never; limbs at or above na + nb count as zero), so the
    // carry scan reads (TH, C) and not the digits.
    const CarryJob C = carry_job<2 * T>(J, nblk_max);
    u32* l0 = lds3;
    u32* l1 = lds3 + hwords;
    u32* l2 = lds3 + 2 * hwords;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        __syncthreads();                        // the tile (or the previous half) is no longer read
#pragma unroll
        for (int t = 4 * h; t < 4 * h + 4; ++t) {
            const int e = padak<CP::PA, CP::PK>((((t - 4 * h) * (R / 8) + u) << CP::LOGT) | col);
            l0[e] = lo[t]; l1[e] = (u32)up[t]; l2[e] = (u32)(up[t] >> 32);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < RH; i += blockDim.x) {
            const size_t k0 = (size_t)(i + h * RH) * S + b0;
            const size_t ch = k0 / T;
            if (ch >= C.nchunks) continue;
            u64 carry = 0, lw = 0;
            bool ones = true;
#pragma unroll 2
            for (int q = 0; q < T; ++q) {
                const int e = padak<CP::PA, CP::PK>(i * T + q);
                const u64 hi = (u64)l1[e] | ((u64)l2[e] << 32);
                const u64 sum = (carry & 0xFFFFFFFFull) + l0[e];
                l0[e] = (u32)sum;
                const bool in = k0 + q < C.nlimbs;
                if (q < 2) lw |= (in ? (u64)(u32)sum : 0) << (32 * q);
                else ones &= in && (u32)sum == 0xFFFFFFFFu;
                carry = (sum >> 32) + (carry >> 32) + hi;
            }
            C.C[ch] = carry;
            C.TH[ch] = (ones && lw != 0) ? (u64)(0 - lw) : ~(u64)0;
        }
        __syncthreads();
        // limbs out with consecutive lanes on consecutive columns of a row
        // (T-limb runs), not one row per lane
#pragma unroll
        for (int k = 0; k < RH * T / CP::WG; ++k) {
            const int e = k * CP::WG + threadIdx.x;
            const size_t g = (size_t)(e / T + h * RH) * S + b0 + (e & (T - 1));
            if (g < C.nlimbs) gst(C.out + g, l0[padak<CP::PA, CP::PK>(e)]);
        }
    }
}

// ---- whole multiply in one workgroup, three primes ------------------------------------
// Products with N = 2^n <= 2^13 (the reference's 16-bit digit count): the
// 32-bit limbs over N' = 2^(n-1) points modulo each prime in six LDS arrays
// (A's three residues, then B's), level-at-a-time radix-8 transforms (DIF
// forward, DIT inverse, the j-major level twiddles), the pointwise product,
// one Montgomery product by 2^-(n-1) R^2 (undoing the pointwise R^-1 and the
// transform's factor N', with canonical residues out), the CRT into
// (lo32, up) over the three arrays and the carry as a block scan.

// Carry of one workgroup's coefficients (lo32 in a[0], up in a[1] | a[2] <<
// 32, up < 2^54): each thread takes k >= 2 consecutive coefficients, so that
// a chunk's carry-out with carry-in c < 2^57 is C or C + 1 (the function of
// (TH, C, C + 1) that the multi-workgroup carry composes); then a block scan
// gives every chunk its carry-in, and the chunks' limbs go out with the
// normalized length.
template <int NT>
__device__ __forceinline__ void block_carry3(u32* const* a, size_t nlimbs, u32* __restrict__ out,
                                             u64* __restrict__ out_len, CarryFn* wsum, u64* outs) {
    size_t k = (nlimbs + NT - 1) / NT;
    if (k < 2) k = 2;
    const size_t l0 = (size_t)threadIdx.x * k;
    const size_t l1 = l0 + k < nlimbs ? l0 + k : nlimbs;
    CarryFn f = cf_pad();
    if (l0 < nlimbs) {
        u64 carry = 0, lo = 0;
        bool ones = true;
        for (size_t l = l0; l < l1; ++l) {
            const int e = lpad((int)l);
            const u64 up = (u64)a[1][e] | ((u64)a[2][e] << 32);
            const u64 sum = (carry & 0xFFFFFFFFull) + a[0][e];
            const u32 limb = (u32)sum;
            carry = (sum >> 32) + (carry >> 32) + up;
            if (l - l0 < 2) lo |= (u64)limb << (32 * (l - l0));
            else ones &= limb == 0xFFFFFFFFu;
        }
        // a one-limb chunk can only be the last one, whose function only
        // reaches the padding threads after it
        f = CarryFn{(l1 - l0 >= 2 && ones && lo != 0) ? (u64)(0 - lo) : ~(u64)0, carry, carry + 1};
    }
    const CarryFn inc = block_scan(f, wsum);
    outs[threadIdx.x] = cf_apply(inc, 0);
    __syncthreads();
    u64 carry = threadIdx.x == 0 ? 0 : outs[threadIdx.x - 1];
    u64 top = 0;
    for (size_t l = l0; l < l1; ++l) {
        const int e = lpad((int)l);
        const u64 up = (u64)a[1][e] | ((u64)a[2][e] << 32);
        const u64 sum = (carry & 0xFFFFFFFFull) + a[0][e];
        const u32 limb = (u32)sum;
        carry = (sum >> 32) + (carry >> 32) + up;
        gst(out + l, limb);
        if (limb) top = l + 1;
    }
    top = wave_max(top);
    __syncthreads();                                  // outs[] is reused below
    if ((threadIdx.x & 63) == 0) outs[threadIdx.x >> 6] = top;
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 m = 0;
        for (int w = 0; w < NT / 64; ++w) m = outs[w] > m ? outs[w] : m;
        *out_len = m;
    }
}

template <bool SQUARE, int NT>
__global__ __launch_bounds__(NT) void k3_small_mul(const MulJob* __restrict__ jobs, MulJob one, int n3,
                                                   const u32* __restrict__ tw3, Primes3 PS, u32 sc0, u32 sc1,
                                                   u32 sc2) {
    extern __shared__ u32 lds3[];
    __shared__ CarryFn wsum[NT / 64];
    __shared__ u64 outs[NT];
    const MulJob J = job_of(jobs, blockIdx.x, one);
    const int R = 1 << n3;
    const int words = lds_words(R);
    u32* a[3] = {lds3, lds3 + words, lds3 + 2 * words};
    u32* b[3] = {lds3 + 3 * words, lds3 + 4 * words, lds3 + 5 * words};
    for (int e = threadIdx.x; e < R; e += NT) {
        const int i = lpad(e);
        const u32 va = (size_t)e < J.na ? gld(J.a + e) : 0u;
#pragma unroll
        for (int p = 0; p < 3; ++p) a[p][i] = r_limb_mont(va, PS.p[p]);
        if constexpr (!SQUARE) {
            const u32 vb = (size_t)e < J.nb ? gld(J.b + e) : 0u;
#pragma unroll
            for (int p = 0; p < 3; ++p) b[p][i] = r_limb_mont(vb, PS.p[p]);
        }
    }
    __syncthreads();
    r_lds_dft3<false>(a, n3, 0, tw3, PS);
    if constexpr (!SQUARE) r_lds_dft3<false>(b, n3, 0, tw3, PS);
    for (int e = threadIdx.x; e < R; e += NT) {
        const int i = lpad(e);
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const u32 x = a[p][i];
            a[p][i] = mml(x, SQUARE ? x : b[p][i], PS.p[p].p, PS.p[p].ninv);
        }
    }
    __syncthreads();
    r_lds_dft3<true>(a, n3, 0, tw3, PS);
    for (int e = threadIdx.x; e < R; e += NT) {
        const int i = lpad(e);
        u32 lo32;
        u64 up;
        crt3(mm(a[0][i], sc0, PS.p[0].p, PS.p[0].ninv), mm(a[1][i], sc1, PS.p[1].p, PS.p[1].ninv),
             mm(a[2][i], sc2, PS.p[2].p, PS.p[2].ninv), PS, lo32, up);
        a[0][i] = lo32; a[1][i] = (u32)up; a[2][i] = (u32)(up >> 32);
    }
    __syncthreads();
    block_carry3<NT>(a, J.na + J.nb, J.out, J.out_len, wsum, outs);
}

// Register-resident one-workgroup product for N' = 2^L2 = 2^8..2^12 points:
// the three primes as three "rows" of the row kernel's register sets (TPR =
// N'/16 threads per prime, 16 points of A and of B each, radix-16 sets and
// the per-exchange LDS layouts), with no outer twiddle (the whole transform
// is one row); the scale 2^-L2 R^2 per prime, the residues through LDS to the
// CRT and the block carry as in k3_small_mul. The workgroup is 3 TPR threads
// rounded up to whole waves; at L2 = 8 and 9 the primes share waves (their
// exchanges are wave-local, so the padding lanes take no barrier).
template <int L2>
struct SmallPlan {
    static constexpr int TPR = RowPlan<L2>::TPR;
    static constexpr int WG = (3 * TPR + 63) / 64 * 64;
    static constexpr int WORDS = RowPlan<L2>::WORDS;
    // LDS words of the CRT arrays; the carry's wave sums and thread outputs
    // follow them, in LDS the exchanges no longer use
    static constexpr int CRTW = 3 * lds_words(1 << L2);
    static constexpr size_t CARRY_BYTES = (size_t)(WG / 64) * sizeof(CarryFn) + (size_t)WG * 8;
    // waves per SIMD the register budget is set for: at 2^11 points, five
    // (at most 96 VGPRs), so that three workgroups of six waves share a CU
    // (their LDS, 52 KB each, fits three); elsewhere the compiler's choice
    static constexpr int MINW = L2 == 11 ? 5 : 1;
};
template <int L2, bool SQUARE>
constexpr size_t small_reg_lds() {
    constexpr size_t ex = (size_t)(SQUARE ? 3 : 6) * RowPlan<L2>::WORDS * 4;
    constexpr size_t crt = (size_t)SmallPlan<L2>::CRTW * 4 + SmallPlan<L2>::CARRY_BYTES;
    return ex > crt ? ex : crt;
}

template <int L2, bool SQUARE>
__global__ __launch_bounds__(SmallPlan<L2>::WG, SmallPlan<L2>::MINW) void k3_small_reg(const MulJob* __restrict__ jobs, MulJob one,
                                                                  const u32* __restrict__ tw3, Primes3 PS,
                                                                  u32 sc0, u32 sc1, u32 sc2) {
    using RP = RowPlan<L2>;
    using SP = SmallPlan<L2>;
    static_assert(L2 >= 8 && L2 <= 12, "register sets of 16 points per thread");
    static_assert(SP::TPR <= 64 || 3 * SP::TPR == SP::WG, "padding lanes only where exchanges are wave-local");
    constexpr int NOP = SQUARE ? 1 : 2;
    constexpr int TPR = SP::TPR;
    constexpr int R = 1 << L2;
    extern __shared__ u32 lds3[];
    static_assert(SP::CRTW % 2 == 0, "8-byte alignment of the carry's LDS");
    CarryFn* wsum = (CarryFn*)(lds3 + SP::CRTW);
    u64* outs = (u64*)(wsum + SP::WG / 64);
    const MulJob J = job_of(jobs, blockIdx.x, one);
    const int g = threadIdx.x / TPR;                 // the prime of this thread's row
    const int u = threadIdx.x % TPR;
    u32 x[NOP][16];
    if (g < 3) {
        const Prime3 P = pick3(PS, g);
        const u32* tw = tw3 + (size_t)g * TW3_STRIDE;
        u32* l[2] = {lds3 + g * RP::WORDS, lds3 + (3 + g) * RP::WORDS};
        u32 w0[16];
        set_tw<L2, 0, false>(w0, u, tw);
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const size_t b = (size_t)t * TPR + u;
            x[0][t] = r_limb_mont(b < J.na ? gld(J.a + b) : 0u, P);
            if constexpr (NOP == 2) x[1][t] = r_limb_mont(b < J.nb ? gld(J.b + b) : 0u, P);
        }
        row_fwd<L2, NOP>(x, w0, u, l, tw, P);
#pragma unroll
        for (int t = 0; t < 16; ++t) x[0][t] = mml(x[0][t], x[NOP - 1][t], P.p, P.ninv);
        {
            u32 w[16];
            set_tw<L2, RP::NSET - 1, true>(w, u, tw);
            row_inv<L2>(x[0], w, u, l[0], tw, P);
        }
        const u32 sc = g == 0 ? sc0 : (g == 1 ? sc1 : sc2);
#pragma unroll
        for (int t = 0; t < 16; ++t) x[0][t] = mm(x[0][t], sc, P.p, P.ninv);
    }
    __syncthreads();                                  // every exchange has been read
    if (g < 3) {
        u32* rp = lds3 + g * lds_words(R);
#pragma unroll
        for (int t = 0; t < 16; ++t) rp[lpad(t * TPR + u)] = x[0][t];
    }
    __syncthreads();
    u32* a[3] = {lds3, lds3 + lds_words(R), lds3 + 2 * lds_words(R)};
    for (int b = threadIdx.x; b < R; b += SP::WG) {
        const int i = lpad(b);
        u32 lo32;
        u64 up;
        crt3(a[0][i], a[1][i], a[2][i], PS, lo32, up);
        a[0][i] = lo32; a[1][i] = (u32)up; a[2][i] = (u32)(up >> 32);
    }
    __syncthreads();
    block_carry3<SP::WG>(a, J.na + J.nb, J.out, J.out_len, wsum, outs);
}

// ---- launcher -----------------------------------------------------------------------
// Row lengths of the register-resident row kernel and its LDS per workgroup.
#define PDH3_FOR_ROW(X) X(8) X(9) X(10) X(11) X(12) X(13) X(14)
#define PDH3_FOR_COL_L1(X) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11)
static size_t row_reg_lds(int L2, bool square) {
    const int tpr = 1 << (L2 - 4), wg = tpr < 256 ? 256 : tpr;
    return (size_t)(square ? 1 : 2) * (wg / tpr) * ((1 << L2) + (1 << (L2 - 4))) * 4;   // RowPlan::WORDS
}

template <class K>
static hipError_t set_lds3(K kernel, size_t bytes) {
    return hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <int T, int LL = 4>
static hipError_t set_col_reg_attrs() {
    constexpr size_t MAXLDS = 147456;
    if constexpr (LL <= 11) {
        if constexpr (col_reg_pair(LL, T)) {
            hipError_t e;
            if ((e = set_lds3(k3_col_fwd_seq<LL, T>, MAXLDS / 3)) != hipSuccess) return e;
            if ((e = set_lds3(k3_col_inv_seq<LL, T>, (size_t)col_inv_seq_words(LL, T) * 4)) != hipSuccess) return e;
        }
        return set_col_reg_attrs<T, LL + 1>();
    }
    return hipSuccess;
}

hipError_t init_kernel_attrs3() {
    constexpr size_t MAXLDS = 147456;
    hipError_t e;
#define PDH3_SET(TT) \
    if ((e = set_lds3(k3_col_fwd<TT>, MAXLDS)) != hipSuccess) return e; \
    if ((e = set_lds3(k3_col_inv<TT>, MAXLDS)) != hipSuccess) return e;
    PDH3_SET(4) PDH3_SET(8) PDH3_SET(16) PDH3_SET(32)
#undef PDH3_SET
    if ((e = set_lds3(k3_row_mul<true>, MAXLDS)) != hipSuccess) return e;
    if ((e = set_lds3(k3_row_mul<false>, MAXLDS)) != hipSuccess) return e;
#define PDH3_SET_ROW(LL) \
    if ((e = set_lds3(k3_row_reg<LL, true>, row_reg_lds(LL, true))) != hipSuccess) return e; \
    if ((e = set_lds3(k3_row_reg<LL, false>, row_reg_lds(LL, false))) != hipSuccess) return e;
    PDH3_FOR_ROW(PDH3_SET_ROW)
#undef PDH3_SET_ROW
#define PDH3_SET_SMALL(LL) \
    if ((e = set_lds3(k3_small_reg<LL, false>, small_reg_lds<LL, false>())) != hipSuccess) return e; \
    if ((e = set_lds3(k3_small_reg<LL, true>, small_reg_lds<LL, true>())) != hipSuccess) return e;
    PDH3_SET_SMALL(8) PDH3_SET_SMALL(9) PDH3_SET_SMALL(10) PDH3_SET_SMALL(11) PDH3_SET_SMALL(12)
#undef PDH3_SET_SMALL
    if ((e = set_col_reg_attrs<4>()) != hipSuccess) return e;
    if ((e = set_col_reg_attrs<8>()) != hipSuccess) return e;
    if ((e = set_col_reg_attrs<16>()) != hipSuccess) return e;
    return set_col_reg_attrs<32>();
}

#define PDH3_LAUNCH(name, s, ...) do { hipEvent_t _t0 = timing_begin(s); hipLaunchKernelGGL(__VA_ARGS__); \
    g_launches.fetch_add(1, std::memory_order_relaxed); hipError_t _e = hipGetLastError(); \
    timing_end(name, _t0, s); if (_e != hipSuccess) return _e; } while (0)

template <int T>
static hipError_t launch_cols3(const MulJob* d_jobs, const MulJob& one, unsigned njobs, int ops, int n3,
                               const NttShape& sh, const OuterTables& ot, hipStream_t s, bool inverse,
                               size_t nblk_max) {
    const Primes3& PS = primes3();
    const unsigned tiles = (unsigned)(((size_t)1 << (n3 - sh.L1)) / T);
    const int words = lds_words((1 << sh.L1) * T);
    if (sh.lr_row == 4) {
        bool done = false;
        switch (sh.L1) {
#define PDH3_COL_CASE(LL) \
        case LL: \
            if constexpr (col_reg_pair(LL, T)) { \
                constexpr int WG = ColPlan<LL, T>::WG; \
                if (!inverse) PDH3_LAUNCH("k3_col_fwd_seq", s, (k3_col_fwd_seq<LL, T>), dim3(njobs * ops * tiles), \
                                          dim3(WG), (size_t)words * 4, s, d_jobs, one, ops, n3, ot.tw3, PS); \
                else PDH3_LAUNCH("k3_col_inv_seq", s, (k3_col_inv_seq<LL, T>), dim3(njobs * tiles), dim3(WG), \
                                 (size_t)col_inv_seq_words(LL, T) * 4, s, d_jobs, one, n3, ot.tw3, PS, nblk_max); \
                done = true; \
            } \
            break;
        PDH3_FOR_COL_L1(PDH3_COL_CASE)
#undef PDH3_COL_CASE
        default: break;
        }
        if (done) return inverse ? launch_carry_scan(T, d_jobs, one, njobs, nblk_max, s, true) : hipSuccess;
    }
    if (!inverse) {
        PDH3_LAUNCH("k3_col_fwd", s, k3_col_fwd<T>, dim3(njobs * ops * tiles), dim3(1024), (size_t)3 * words * 4, s,
                    d_jobs, one, ops, n3, sh.L1, ot.tw3, PS);
        return hipSuccess;
    }
    PDH3_LAUNCH("k3_col_inv", s, k3_col_inv<T>, dim3(njobs * tiles), dim3(1024), (size_t)3 * words * 4, s,
                d_jobs, one, n3, sh.L1, ot.tw3, PS, nblk_max);
    return launch_carry_scan(T, d_jobs, one, njobs, nblk_max, s);
}

static hipError_t launch_cols3_T(const MulJob* d_jobs, const MulJob& one, unsigned njobs, int ops, int n3,
                                 const NttShape& sh, const OuterTables& ot, hipStream_t s, bool inverse,
                                 size_t nblk_max) {
    switch (sh.T) {
    case 4: return launch_cols3<4>(d_jobs, one, njobs, ops, n3, sh, ot, s, inverse, nblk_max);
    case 8: return launch_cols3<8>(d_jobs, one, njobs, ops, n3, sh, ot, s, inverse, nblk_max);
    case 16: return launch_cols3<16>(d_jobs, one, njobs, ops, n3, sh, ot, s, inverse, nblk_max);
    case 32: return launch_cols3<32>(d_jobs, one, njobs, ops, n3, sh, ot, s, inverse, nblk_max);
    default: return hipErrorInvalidValue;
    }
}

// One workgroup per product (N = 2^n <= 2^13): the scale 2^-(n-1) R^2 per
// prime (as Montgomery constants).
hipError_t launch_small3(const MulJob* d_jobs, const MulJob& one, unsigned njobs, int n, bool square,
                         const u32* tw3, hipStream_t s) {
    const int n3 = n - 1;
    const Primes3& PS = primes3();
    u32 sc[3];
    for (int pi = 0; pi < 3; ++pi) {
        const u32 p = PS.p[pi].p;
        const u32 inv = r3_pow(r3_pow(2, (uint64_t)n3, p), p - 2, p);   // 2^-n3
        sc[pi] = r3_to_mont(r3_to_mont(inv, p), p);
    }
    const dim3 g(njobs);
    if (n3 >= 8 && n3 <= 12) {
        // register-resident sets (k3_small_reg)
        switch (n3) {
#define PDH3_SMALL_REG(LL)                                                                                         \
        case LL:                                                                                                   \
            if (square) PDH3_LAUNCH("k3_small_reg", s, (k3_small_reg<LL, true>), g, dim3(SmallPlan<LL>::WG),       \
                                    (small_reg_lds<LL, true>()), s, d_jobs, one, tw3, PS, sc[0], sc[1], sc[2]);   \
            else PDH3_LAUNCH("k3_small_reg", s, (k3_small_reg<LL, false>), g, dim3(SmallPlan<LL>::WG),             \
                             (small_reg_lds<LL, false>()), s, d_jobs, one, tw3, PS, sc[0], sc[1], sc[2]);         \
            return hipSuccess;
        PDH3_SMALL_REG(8) PDH3_SMALL_REG(9) PDH3_SMALL_REG(10) PDH3_SMALL_REG(11) PDH3_SMALL_REG(12)
#undef PDH3_SMALL_REG
        default: break;
        }
    }
    // up to 2^7 points: radix-8 levels through LDS, one wave (at the sizes
    // the register sets take over, N'/4 threads had been the best of N'/8,
    // N'/4 and N'/2, and the sets ran 1.1-1.5x that)
    const size_t lds = (size_t)(square ? 3 : 6) * lds_words(1 << n3) * 4;
    if (square) PDH3_LAUNCH("k3_small_mul", s, (k3_small_mul<true, 64>), g, dim3(64), lds, s, d_jobs, one, n3, tw3, PS,
                            sc[0], sc[1], sc[2]);
    else PDH3_LAUNCH("k3_small_mul", s, (k3_small_mul<false, 64>), g, dim3(64), lds, s, d_jobs, one, n3, tw3, PS,
                     sc[0], sc[1], sc[2]);
    return hipSuccess;
}

hipError_t launch_class3(const MulJob* d_jobs, const MulJob& one, unsigned njobs, const NttShape& sh, bool square,
                         size_t nblk_max, const OuterTables& ot, hipStream_t s) {
    const int n3 = sh.n - 1;
    hipError_t e = launch_cols3_T(d_jobs, one, njobs, square ? 1 : 2, n3, sh, ot, s, false, nblk_max);
    if (e != hipSuccess) return e;
    const size_t lds = (size_t)(square ? 1 : 2) * lds_words(1 << sh.L2) * 4;
    const dim3 grid((unsigned)(njobs * 3u) << sh.L1);
    const Primes3& PS = primes3();
    if (sh.lr_row == 4) {
        switch (sh.L2) {
#define PDH3_ROW_CASE(LL) \
        case LL: { \
            using RP = RowPlan<LL>; \
            const dim3 g((unsigned)(njobs * 3u) * ((1u << sh.L1) / RP::G)); \
            if (square) PDH3_LAUNCH("k3_row_reg_sqr", s, (k3_row_reg<LL, true>), g, dim3(RP::WG), row_reg_lds(LL, true), s, \
                                    d_jobs, one, sh.L1, sh.lr_col, ot.tw3, ot.o3, ot.h3, PS); \
            else PDH3_LAUNCH("k3_row_reg", s, (k3_row_reg<LL, false>), g, dim3(RP::WG), row_reg_lds(LL, false), s, \
                             d_jobs, one, sh.L1, sh.lr_col, ot.tw3, ot.o3, ot.h3, PS); \
            break; \
        }
        PDH3_FOR_ROW(PDH3_ROW_CASE)
#undef PDH3_ROW_CASE
        default: return hipErrorInvalidValue;
        }
    } else if (square) {
        PDH3_LAUNCH("k3_row_sqr", s, k3_row_mul<true>, grid, dim3(1024), lds, s, d_jobs, one, sh.L1, sh.L2,
                    sh.lr_col, ot.tw3, ot.o3, ot.h3, PS);
    } else {
        PDH3_LAUNCH("k3_row_mul", s, k3_row_mul<false>, grid, dim3(1024), lds, s, d_jobs, one, sh.L1, sh.L2,
                    sh.lr_col, ot.tw3, ot.o3, ot.h3, PS);
    }
    return launch_cols3_T(d_jobs, one, njobs, 1, n3, sh, ot, s, true, nblk_max);
}

} // namespace pdh
