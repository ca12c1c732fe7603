// ============================================================================
//  ntt_kernels.hip -- Goldilocks NTT big-integer multiply for gfx950.
//
//  Replaces the reference's src/ntt_cuda.cu:68-243 (k_bit_reverse,
//  k_build_twiddles, log2(N) launches of k_butterfly, k_pointwise, k_scale,
//  host repack and host carry) with:
//
//    k_col_fwd   repack u32 limbs -> 16-bit digits (include/ntt.h:56-64,
//                fused), 2^L1-point DIF sub-transforms over the top L1 index
//                bits of T adjacent columns in LDS, outer twiddle w_N^(b*c).
//    k_row_mul   per row of 2^L2 points: DIF of A and B, pointwise product,
//                DIT inverse -- the forward second pass, the pointwise
//                multiply and the inverse first pass in one HBM round trip.
//    k_col_inv   inverse outer twiddle * N^-1, DIT column sub-transforms.
//    k_small_mul whole multiply in one workgroup for N <= 2^13.
//    k_carry_*   base-2^16 carry (include/ntt.h:70-91) as a parallel scan of
//                carry functions f(c) = (c >= thr) ? v1 : v0.
//
//  No bit-reversal pass: the forward transform is decimation in frequency
//  (natural order in, digit-reversed spectrum out) and the inverse is its
//  exact mirror (digit-reversed in, natural out). The spectrum order is not
//  observable in the product (it is a unique integer), so any order shared by
//  the forward and inverse transforms is correct.
//
//  Sub-transforms are radix-8 (radix-2/4 for the leftover bits) butterflies in
//  registers with shift-only internal twiddles (w8 = 2^24), exchanging through
//  LDS between radix levels; LDS addresses are padded one word in eight to
//  keep strided butterfly reads off the same banks.
// ============================================================================
#include "gl_field.h"
#include "pdh_internal.h"

namespace pdh {

std::atomic<uint64_t> g_launches{0};

__device__ __forceinline__ int lpad(int e) { return e + (e >> 3); }
__host__ __device__ constexpr int lds_words(int e) { return e + (e >> 3); }

// ---- radix-q DFTs, natural order in and out ---------------------------------
template <bool INV>
__device__ __forceinline__ void dft2(u64* x) {
    u64 a = gl_add(x[0], x[1]), b = gl_sub(x[0], x[1]);
    x[0] = a; x[1] = b;
}

template <bool INV>
__device__ __forceinline__ void dft4(u64& x0, u64& x1, u64& x2, u64& x3) {
    u64 c0 = gl_add(x0, x2), d0 = gl_sub(x0, x2);
    u64 c1 = gl_add(x1, x3), d1 = gl_sub(x1, x3);
    d1 = INV ? gl_mulp2<144>(d1) : gl_mulp2<48>(d1);      // w4^{-1}, w4
    x0 = gl_add(c0, c1); x2 = gl_sub(c0, c1);
    x1 = gl_add(d0, d1); x3 = gl_sub(d0, d1);
}

template <bool INV>
__device__ __forceinline__ void dft4a(u64* x) { dft4<INV>(x[0], x[1], x[2], x[3]); }

// Y_{2k} = DFT4(x_t + x_{t+4})_k, Y_{2k+1} = DFT4((x_t - x_{t+4}) w8^t)_k.
template <bool INV>
__device__ __forceinline__ void dft8(u64* x) {
    u64 a0 = gl_add(x[0], x[4]), b0 = gl_sub(x[0], x[4]);
    u64 a1 = gl_add(x[1], x[5]), b1 = gl_sub(x[1], x[5]);
    u64 a2 = gl_add(x[2], x[6]), b2 = gl_sub(x[2], x[6]);
    u64 a3 = gl_add(x[3], x[7]), b3 = gl_sub(x[3], x[7]);
    if (!INV) { b1 = gl_mulp2<24>(b1);  b2 = gl_mulp2<48>(b2);  b3 = gl_mulp2<72>(b3); }
    else      { b1 = gl_mulp2<168>(b1); b2 = gl_mulp2<144>(b2); b3 = gl_mulp2<120>(b3); }
    dft4<INV>(a0, a1, a2, a3);
    dft4<INV>(b0, b1, b2, b3);
    x[0] = a0; x[2] = a1; x[4] = a2; x[6] = a3;
    x[1] = b0; x[3] = b1; x[5] = b2; x[7] = b3;
}

template <int Q, bool INV>
__device__ __forceinline__ void dftq(u64* x) {
    if constexpr (Q == 2) dft2<INV>(x);
    else if constexpr (Q == 4) dft4a<INV>(x);
    else dft8<INV>(x);
}

// ---- one radix-Q level of an in-place 2^L-point sub-transform in LDS ---------
// NARR arrays, each holding T interleaved columns: element (i, col) of array
// a sits at arr[a][lpad(i*T + col)]. Sub-block size M (M = Q*m): butterfly
// (beta, j) reads i = beta*M + j + t*m, t < Q. Forward: DFT_Q then output c
// times w_M^(j*c). Inverse: times w_M^(-j*c), then inverse DFT_Q.
template <int Q, bool INV, int NARR>
__device__ __forceinline__ void lds_level(u64* const* arr, int L, int M, int T,
                                          const u64* __restrict__ twL) {
    const int m = M / Q;
    const int nbf = ((1 << L) / Q) * T;
    const int twscale = (1 << L) / M;
    const int mask = (1 << L) - 1;
    const int tlog = __ffs(T) - 1;
    for (int u = threadIdx.x; u < nbf; u += blockDim.x) {
        const int col = u & (T - 1);
        const int rest = u >> tlog;
        const int j = rest % m;
        const int base = (rest / m) * M + j;
        u64 w[Q];
        w[0] = 1;
#pragma unroll
        for (int c = 1; c < Q; ++c) {
            int e = (twscale * j * c) & mask;
            w[c] = twL[INV ? ((-e) & mask) : e];
        }
#pragma unroll
        for (int a = 0; a < NARR; ++a) {
            u64 x[Q];
#pragma unroll
            for (int t = 0; t < Q; ++t) x[t] = arr[a][lpad((base + t * m) * T + col)];
            if (!INV) {
                dftq<Q, false>(x);
#pragma unroll
                for (int c = 1; c < Q; ++c) x[c] = gl_mul(x[c], w[c]);
            } else {
#pragma unroll
                for (int c = 1; c < Q; ++c) x[c] = gl_mul(x[c], w[c]);
                dftq<Q, true>(x);
            }
#pragma unroll
            for (int t = 0; t < Q; ++t) arr[a][lpad((base + t * m) * T + col)] = x[t];
        }
    }
}

// Full 2^L-point sub-transform: forward = radix 2^(L mod 3) level then radix-8
// levels (M = 2^L ... 8); inverse = the same levels in reverse order.
// Callers synchronise before; this returns after a final barrier.
template <bool INV, int NARR>
__device__ void lds_dft(u64* const* arr, int L, int T, const u64* __restrict__ twL) {
    const int r0 = L % 3;
    if (!INV) {
        int M = 1 << L;
        if (r0 == 1) { lds_level<2, false, NARR>(arr, L, M, T, twL); M >>= 1; __syncthreads(); }
        if (r0 == 2) { lds_level<4, false, NARR>(arr, L, M, T, twL); M >>= 2; __syncthreads(); }
        for (; M >= 8; M >>= 3) { lds_level<8, false, NARR>(arr, L, M, T, twL); __syncthreads(); }
    } else {
        const int Mtop = (1 << L) >> r0;
        for (int M = 8; M <= Mtop; M <<= 3) { lds_level<8, true, NARR>(arr, L, M, T, twL); __syncthreads(); }
        if (r0 == 1) { lds_level<2, true, NARR>(arr, L, 1 << L, T, twL); __syncthreads(); }
        if (r0 == 2) { lds_level<4, true, NARR>(arr, L, 1 << L, T, twL); __syncthreads(); }
    }
}

// Frequency held at position i after the forward sub-transform: the position's
// digits (widths r0, 3, 3, ... from the top) are the frequency's digits from
// the bottom.
__device__ __forceinline__ u64 freq_of_pos(int i, int L) {
    const int r0 = L % 3;
    int left = L, sh = 0;
    u64 c = 0;
    if (r0) { left -= r0; c = (u64)((i >> left) & ((1 << r0) - 1)); sh = r0; }
    while (left > 0) { left -= 3; c |= (u64)((i >> left) & 7) << sh; sh += 3; }
    return c;
}

__device__ __forceinline__ u64 digit16(const u32* __restrict__ limbs, size_t nlimbs, size_t g) {
    size_t li = g >> 1;
    return li < nlimbs ? (u64)((limbs[li] >> (16 * (g & 1))) & 0xFFFFu) : 0;
}

// ---- column pass, forward ------------------------------------------------------
// Workgroup w handles columns b = w*T .. w*T+T-1 of the 2^L1 x 2^L2 view
// (element (i, b) at i*2^L2 + b).
template <int T, bool FROM_LIMBS>
__global__ __launch_bounds__(256) void k_col_fwd(const void* __restrict__ src, size_t nlimbs,
                                                 u64* __restrict__ dst, int n, int L1,
                                                 const u64* __restrict__ twL,
                                                 const u64* __restrict__ ohi,
                                                 const u64* __restrict__ olo, int h) {
    extern __shared__ u64 lds[];
    const int L2 = n - L1;
    const size_t S = (size_t)1 << L2;
    const int R = 1 << L1;
    const size_t b0 = (size_t)blockIdx.x * T;
    for (int e = threadIdx.x; e < R * T; e += blockDim.x) {
        const int col = e & (T - 1), i = e / T;
        const size_t g = (size_t)i * S + b0 + col;
        u64 v;
        if constexpr (FROM_LIMBS) v = digit16((const u32*)src, nlimbs, g);
        else v = ((const u64*)src)[g];
        lds[lpad(e)] = v;
    }
    __syncthreads();
    u64* arr[1] = {lds};
    lds_dft<false, 1>(arr, L1, T, twL + R);
    const u64 mask = ((u64)1 << h) - 1;
    for (int e = threadIdx.x; e < R * T; e += blockDim.x) {
        const int col = e & (T - 1), i = e / T;
        const u64 ex = (b0 + col) * freq_of_pos(i, L1);
        const u64 w = gl_mul(ohi[ex >> h], olo[ex & mask]);
        dst[(size_t)i * S + b0 + col] = gl_mul(lds[lpad(e)], w);
    }
}

// ---- column pass, inverse (times N^-1, folded into the inverse tables) ----------
template <int T>
__global__ __launch_bounds__(256) void k_col_inv(u64* __restrict__ buf, int n, int L1,
                                                 const u64* __restrict__ twL,
                                                 const u64* __restrict__ ohi,
                                                 const u64* __restrict__ olo, int h) {
    extern __shared__ u64 lds[];
    const int L2 = n - L1;
    const size_t S = (size_t)1 << L2;
    const int R = 1 << L1;
    const size_t b0 = (size_t)blockIdx.x * T;
    const u64 mask = ((u64)1 << h) - 1;
    for (int e = threadIdx.x; e < R * T; e += blockDim.x) {
        const int col = e & (T - 1), i = e / T;
        const u64 ex = (b0 + col) * freq_of_pos(i, L1);
        const u64 w = gl_mul(ohi[ex >> h], olo[ex & mask]);
        lds[lpad(e)] = gl_mul(buf[(size_t)i * S + b0 + col], w);
    }
    __syncthreads();
    u64* arr[1] = {lds};
    lds_dft<true, 1>(arr, L1, T, twL + R);
    for (int e = threadIdx.x; e < R * T; e += blockDim.x) {
        const int col = e & (T - 1), i = e / T;
        buf[(size_t)i * S + b0 + col] = lds[lpad(e)];
    }
}

// ---- row pass: forward A (and B), pointwise, inverse; result into A ----------------
template <bool SQUARE>
__global__ __launch_bounds__(256) void k_row_mul(u64* __restrict__ A, const u64* __restrict__ B,
                                                 int L2, const u64* __restrict__ twL) {
    extern __shared__ u64 lds[];
    const int R = 1 << L2;
    u64* la = lds;
    u64* lb = lds + lds_words(R);
    const size_t base = (size_t)blockIdx.x * R;
    for (int e = threadIdx.x; e < R; e += blockDim.x) {
        la[lpad(e)] = A[base + e];
        if constexpr (!SQUARE) lb[lpad(e)] = B[base + e];
    }
    __syncthreads();
    if constexpr (SQUARE) {
        u64* arr[1] = {la};
        lds_dft<false, 1>(arr, L2, 1, twL + R);
        for (int e = threadIdx.x; e < R; e += blockDim.x) { u64 v = la[lpad(e)]; la[lpad(e)] = gl_mul(v, v); }
    } else {
        u64* arr[2] = {la, lb};
        lds_dft<false, 2>(arr, L2, 1, twL + R);
        for (int e = threadIdx.x; e < R; e += blockDim.x) la[lpad(e)] = gl_mul(la[lpad(e)], lb[lpad(e)]);
    }
    __syncthreads();
    u64* arr1[1] = {la};
    lds_dft<true, 1>(arr1, L2, 1, twL + R);
    for (int e = threadIdx.x; e < R; e += blockDim.x) A[base + e] = la[lpad(e)];
}

// ---- whole multiply in one workgroup (N = 2^n <= 2^13) -----------------------------
template <bool SQUARE>
__global__ __launch_bounds__(256) void k_small_mul(const u32* __restrict__ a, size_t na,
                                                   const u32* __restrict__ b, size_t nb,
                                                   u64* __restrict__ coef, int n,
                                                   const u64* __restrict__ twL, u64 ninv) {
    extern __shared__ u64 lds[];
    const int R = 1 << n;
    u64* la = lds;
    u64* lb = lds + lds_words(R);
    for (int e = threadIdx.x; e < R; e += blockDim.x) {
        la[lpad(e)] = digit16(a, na, e);
        if constexpr (!SQUARE) lb[lpad(e)] = digit16(b, nb, e);
    }
    __syncthreads();
    if constexpr (SQUARE) {
        u64* arr[1] = {la};
        lds_dft<false, 1>(arr, n, 1, twL + R);
        for (int e = threadIdx.x; e < R; e += blockDim.x) { u64 v = la[lpad(e)]; la[lpad(e)] = gl_mul(v, v); }
    } else {
        u64* arr[2] = {la, lb};
        lds_dft<false, 2>(arr, n, 1, twL + R);
        for (int e = threadIdx.x; e < R; e += blockDim.x) la[lpad(e)] = gl_mul(la[lpad(e)], lb[lpad(e)]);
    }
    __syncthreads();
    u64* arr1[1] = {la};
    lds_dft<true, 1>(arr1, n, 1, twL + R);
    for (int e = threadIdx.x; e < R; e += blockDim.x) coef[e] = gl_mul(la[lpad(e)], ninv);
}

// ---- carry propagation --------------------------------------------------------------
// Chunk tau holds coefficients 16*tau .. 16*tau+15 -> limbs 8*tau .. 8*tau+7.
// With carry-in 0 it yields 256 bits D and carry-out C (< 2^41); with carry-in
// c the carry-out is C + [c >= 2^256 - D], and 2^256 - D fits in 64 bits only
// when limbs 2..7 of D are all ones. So each chunk is the function
// f(c) = c >= thr ? C+1 : C, and compositions keep that form.
struct CarryFn { u64 thr, v0, v1; };
__device__ __forceinline__ u64 cf_apply(const CarryFn& f, u64 c) { return c >= f.thr ? f.v1 : f.v0; }
// (later o earlier)
__device__ __forceinline__ CarryFn cf_compose(const CarryFn& later, const CarryFn& earlier) {
    return CarryFn{earlier.thr, cf_apply(later, earlier.v0), cf_apply(later, earlier.v1)};
}

constexpr int CARRY_NT = 256;
constexpr int CARRY_CH = 16;      // coefficients per thread

__device__ __forceinline__ void load_chunk(const u64* __restrict__ coef, size_t ncoef, size_t tau,
                                           u64* c) {
    const size_t g0 = tau * CARRY_CH;
#pragma unroll
    for (int q = 0; q < CARRY_CH; ++q) c[q] = (g0 + q < ncoef) ? coef[g0 + q] : 0;
}

__device__ __forceinline__ u64 run_carry(const u64* c, u64 carry, u32* limbs) {
#pragma unroll
    for (int q = 0; q < CARRY_CH / 2; ++q) {
        u64 cur = carry + c[2 * q];
        u32 lo16 = (u32)(cur & 0xFFFFu);
        carry = cur >> 16;
        u64 cur2 = carry + c[2 * q + 1];
        u32 hi16 = (u32)(cur2 & 0xFFFFu);
        carry = cur2 >> 16;
        limbs[q] = lo16 | (hi16 << 16);
    }
    return carry;
}

__device__ __forceinline__ CarryFn chunk_fn(const u64* c) {
    u32 limbs[CARRY_CH / 2];
    u64 C = run_carry(c, 0, limbs);
    bool ones = true;
#pragma unroll
    for (int q = 2; q < CARRY_CH / 2; ++q) ones &= (limbs[q] == 0xFFFFFFFFu);
    u64 lo = (u64)limbs[0] | ((u64)limbs[1] << 32);
    u64 thr = (ones && lo != 0) ? (u64)(0 - lo) : ~(u64)0;
    return CarryFn{thr, C, C + 1};
}

__device__ __forceinline__ CarryFn shfl_up_fn(const CarryFn& f, int d) {
    return CarryFn{__shfl_up(f.thr, d), __shfl_up(f.v0, d), __shfl_up(f.v1, d)};
}

// Inclusive scan of carry functions over the workgroup (thread 0 earliest).
__device__ CarryFn block_scan(CarryFn f, CarryFn* wsum) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        CarryFn g = shfl_up_fn(f, d);
        if (lane >= d) f = cf_compose(f, g);
    }
    if (lane == 63) wsum[wid] = f;
    __syncthreads();
    if (wid > 0) {
        CarryFn pre = wsum[0];
        for (int w = 1; w < wid; ++w) pre = cf_compose(wsum[w], pre);
        f = cf_compose(f, pre);
    }
    __syncthreads();
    return f;
}

__global__ __launch_bounds__(CARRY_NT) void k_carry_reduce(const u64* __restrict__ coef, size_t ncoef,
                                                           size_t nchunks, CarryFn* __restrict__ agg,
                                                           u64* __restrict__ out_len) {
    __shared__ CarryFn wsum[CARRY_NT / 64];
    const size_t tau = (size_t)blockIdx.x * CARRY_NT + threadIdx.x;
    CarryFn f{~(u64)0, 0, 0};
    if (tau < nchunks) {
        u64 c[CARRY_CH];
        load_chunk(coef, ncoef, tau, c);
        f = chunk_fn(c);
    }
    f = block_scan(f, wsum);
    if (threadIdx.x == CARRY_NT - 1) agg[blockIdx.x] = f;
    if (blockIdx.x == 0 && threadIdx.x == 0) *out_len = 0;
}

// One workgroup of 1024 threads: carry-in of every carry block, in tiles.
__global__ __launch_bounds__(1024) void k_carry_blocks(const CarryFn* __restrict__ agg, size_t nblocks,
                                                       u64* __restrict__ cin) {
    __shared__ CarryFn wsum[1024 / 64];
    __shared__ u64 tile_out[1024];
    u64 carry = 0;
    for (size_t t0 = 0; t0 < nblocks; t0 += 1024) {
        const size_t w = t0 + threadIdx.x;
        CarryFn f = (w < nblocks) ? agg[w] : CarryFn{~(u64)0, 0, 0};
        f = block_scan(f, wsum);
        const u64 outc = cf_apply(f, carry);
        tile_out[threadIdx.x] = outc;
        __syncthreads();
        if (w < nblocks) cin[w] = threadIdx.x == 0 ? carry : tile_out[threadIdx.x - 1];
        carry = tile_out[1023];
        __syncthreads();
    }
}

__global__ __launch_bounds__(CARRY_NT) void k_carry_apply(const u64* __restrict__ coef, size_t ncoef,
                                                          size_t nchunks, size_t nlimbs,
                                                          const u64* __restrict__ cin,
                                                          u32* __restrict__ out, u64* __restrict__ out_len) {
    __shared__ CarryFn wsum[CARRY_NT / 64];
    __shared__ u64 outs[CARRY_NT];
    const size_t tau = (size_t)blockIdx.x * CARRY_NT + threadIdx.x;
    u64 c[CARRY_CH];
    CarryFn f{~(u64)0, 0, 0};
    if (tau < nchunks) {
        load_chunk(coef, ncoef, tau, c);
        f = chunk_fn(c);
    }
    const u64 bc = cin[blockIdx.x];
    CarryFn inc = block_scan(f, wsum);
    outs[threadIdx.x] = cf_apply(inc, bc);
    __syncthreads();
    const u64 my_in = threadIdx.x == 0 ? bc : outs[threadIdx.x - 1];
    u64 top = 0;
    if (tau < nchunks) {
        u32 limbs[CARRY_CH / 2];
        run_carry(c, my_in, limbs);
        const size_t l0 = tau * (CARRY_CH / 2);
#pragma unroll
        for (int q = 0; q < CARRY_CH / 2; ++q) {
            if (l0 + q < nlimbs) {
                out[l0 + q] = limbs[q];
                if (limbs[q]) top = l0 + q + 1;
            }
        }
    }
    // wave max, one atomic per wave
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) { u64 o = __shfl_xor(top, d); top = o > top ? o : top; }
    if ((threadIdx.x & 63) == 0 && top) atomicMax((unsigned long long*)out_len, (unsigned long long)top);
}

__global__ void k_zero_len(u64* out_len) { *out_len = 0; }

// ---- host launchers -------------------------------------------------------------------
NttShape ntt_shape(int n) {
    NttShape s{n, 0, n, 1};
    if (n <= SMALL_MAX_LOG) return s;
    int L1 = n / 2 < COL_MAX_LOG ? n / 2 : COL_MAX_LOG;
    int L2 = n - L1;
    if (L2 > ROW_MAX_LOG) { L2 = ROW_MAX_LOG; L1 = n - L2; }
    s.L1 = L1; s.L2 = L2;
    // widest column tile that fits 144 KB of LDS and the row length
    int T = 64;
    while (T > 8 && ((size_t)lds_words((1 << L1) * T) * 8 > 147456 || T > (1 << L2))) T >>= 1;
    s.T = T;
    return s;
}

size_t carry_ws_bytes(size_t nlimbs) {
    size_t nchunks = (nlimbs + CARRY_CH / 2 - 1) / (CARRY_CH / 2);
    size_t nblk = (nchunks + CARRY_NT - 1) / CARRY_NT;
    return nblk * sizeof(CarryFn) + nblk * sizeof(u64) + 256;
}

template <class K>
static hipError_t set_lds(K kernel, size_t bytes) {
    return hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

#define PDH_TRY(x) do { hipError_t _e = (x); if (_e != hipSuccess) return _e; } while (0)
#define PDH_LAUNCHED() do { g_launches.fetch_add(1, std::memory_order_relaxed); PDH_TRY(hipGetLastError()); } while (0)
// Launch with optional event timing under the kernel's name.
#define PDH_LAUNCH(name, s, ...) do { hipEvent_t _t0 = timing_begin(s); hipLaunchKernelGGL(__VA_ARGS__); \
    PDH_LAUNCHED(); timing_end(name, _t0, s); } while (0)

template <int T>
static hipError_t launch_cols(const u32* a, size_t na, const u32* b, size_t nb, bool square,
                              u64* wsA, u64* wsB, const u64* twL, const OuterTables& ot,
                              const NttShape& sh, hipStream_t s, bool inverse) {
    const size_t lds = (size_t)lds_words((1 << sh.L1) * T) * 8;
    const dim3 grid((unsigned)(((size_t)1 << sh.L2) / T));
    if (!inverse) {
        PDH_TRY(set_lds(k_col_fwd<T, true>, lds));
        PDH_LAUNCH("k_col_fwd", s, (k_col_fwd<T, true>), grid, dim3(256), lds, s, (const void*)a, na, wsA,
                   sh.n, sh.L1, twL, ot.hi, ot.lo, ot.h);
        if (!square)
            PDH_LAUNCH("k_col_fwd", s, (k_col_fwd<T, true>), grid, dim3(256), lds, s, (const void*)b, nb, wsB,
                       sh.n, sh.L1, twL, ot.hi, ot.lo, ot.h);
    } else {
        PDH_TRY(set_lds(k_col_inv<T>, lds));
        PDH_LAUNCH("k_col_inv", s, (k_col_inv<T>), grid, dim3(256), lds, s, wsA, sh.n, sh.L1, twL,
                   ot.hi_inv, ot.lo_inv, ot.h);
    }
    return hipSuccess;
}

static hipError_t launch_cols_T(const u32* a, size_t na, const u32* b, size_t nb, bool square,
                                u64* wsA, u64* wsB, const u64* twL, const OuterTables& ot,
                                const NttShape& sh, hipStream_t s, bool inverse) {
    switch (sh.T) {
        case 8:  return launch_cols<8>(a, na, b, nb, square, wsA, wsB, twL, ot, sh, s, inverse);
        case 16: return launch_cols<16>(a, na, b, nb, square, wsA, wsB, twL, ot, sh, s, inverse);
        case 32: return launch_cols<32>(a, na, b, nb, square, wsA, wsB, twL, ot, sh, s, inverse);
        default: return launch_cols<64>(a, na, b, nb, square, wsA, wsB, twL, ot, sh, s, inverse);
    }
}

hipError_t launch_zero_len(uint64_t* out_len, hipStream_t s) {
    PDH_LAUNCH("k_zero_len", s, k_zero_len, dim3(1), dim3(1), 0, s, out_len);
    return hipSuccess;
}

static hipError_t launch_carry(const u64* coef, size_t ncoef, size_t nlimbs, u32* out, u64* out_len,
                               u64* carry_ws, hipStream_t s) {
    const size_t nchunks = (nlimbs + CARRY_CH / 2 - 1) / (CARRY_CH / 2);
    const size_t nblk = (nchunks + CARRY_NT - 1) / CARRY_NT;
    CarryFn* agg = (CarryFn*)carry_ws;
    u64* cin = (u64*)((char*)carry_ws + nblk * sizeof(CarryFn));
    PDH_LAUNCH("k_carry_reduce", s, k_carry_reduce, dim3((unsigned)nblk), dim3(CARRY_NT), 0, s, coef, ncoef,
               nchunks, agg, out_len);
    PDH_LAUNCH("k_carry_blocks", s, k_carry_blocks, dim3(1), dim3(1024), 0, s, (const CarryFn*)agg, nblk, cin);
    PDH_LAUNCH("k_carry_apply", s, k_carry_apply, dim3((unsigned)nblk), dim3(CARRY_NT), 0, s, coef, ncoef,
               nchunks, nlimbs, (const u64*)cin, out, out_len);
    return hipSuccess;
}

hipError_t launch_mul(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, bool square,
                      uint32_t* out, uint64_t* out_len, uint64_t* wsA, uint64_t* wsB, uint64_t* carry_ws,
                      const uint64_t* twL, const OuterTables& ot, const NttShape& sh, hipStream_t s) {
    const size_t N = (size_t)1 << sh.n;
    const size_t nlimbs = na + nb;
    if (sh.L1 == 0) {
        const size_t lds = (size_t)(square ? 1 : 2) * lds_words((int)N) * 8;
        const u64 ninv = gl_inv((u64)N);
        if (square) {
            PDH_TRY(set_lds(k_small_mul<true>, lds));
            PDH_LAUNCH("k_small_mul", s, k_small_mul<true>, dim3(1), dim3(256), lds, s, a, na, a, na, wsA, sh.n,
                       twL, ninv);
        } else {
            PDH_TRY(set_lds(k_small_mul<false>, lds));
            PDH_LAUNCH("k_small_mul", s, k_small_mul<false>, dim3(1), dim3(256), lds, s, a, na, b, nb, wsA, sh.n,
                       twL, ninv);
        }
    } else {
        PDH_TRY(launch_cols_T(a, na, b, nb, square, wsA, wsB, twL, ot, sh, s, false));
        const size_t lds = (size_t)(square ? 1 : 2) * lds_words(1 << sh.L2) * 8;
        const dim3 grid((unsigned)((size_t)1 << sh.L1));
        if (square) {
            PDH_TRY(set_lds(k_row_mul<true>, lds));
            PDH_LAUNCH("k_row_sqr", s, k_row_mul<true>, grid, dim3(256), lds, s, wsA, (const u64*)wsA, sh.L2, twL);
        } else {
            PDH_TRY(set_lds(k_row_mul<false>, lds));
            PDH_LAUNCH("k_row_mul", s, k_row_mul<false>, grid, dim3(256), lds, s, wsA, (const u64*)wsB, sh.L2, twL);
        }
        PDH_TRY(launch_cols_T(a, na, b, nb, square, wsA, wsB, twL, ot, sh, s, true));
    }
    return launch_carry(wsA, N, nlimbs, out, out_len, carry_ws, s);
}

} // namespace pdh
