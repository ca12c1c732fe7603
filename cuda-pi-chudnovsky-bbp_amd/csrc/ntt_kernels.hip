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
//                Its workgroups also run the base-2^16 carry (include/ntt.h:
//                70-91) inside each row segment of T coefficients.
//    k_carry_*   the carry across segments as a parallel scan of carry
//                functions f(c) = (c >= thr) ? v1 : v0.
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
#include "ntt_common.h"

namespace pdh {

std::atomic<uint64_t> g_launches{0};


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

// X_{c1 + 4 c2} = DFT4_b( w16^(b c1) DFT4_a(x_{b + 4a})_c1 )_c2 with
// w16 = 2^12: eight DFT4s and nine shift twiddles.
template <bool INV>
__device__ __forceinline__ void dft16(u64* x) {
#pragma unroll
    for (int b = 0; b < 4; ++b) dft4<INV>(x[b], x[b + 4], x[b + 8], x[b + 12]);
    // x[b + 4 c1] now holds y[b][c1]
    if (!INV) {
        x[5] = gl_mulp2<12>(x[5]);  x[9] = gl_mulp2<24>(x[9]);   x[13] = gl_mulp2<36>(x[13]);
        x[6] = gl_mulp2<24>(x[6]);  x[10] = gl_mulp2<48>(x[10]); x[14] = gl_mulp2<72>(x[14]);
        x[7] = gl_mulp2<36>(x[7]);  x[11] = gl_mulp2<72>(x[11]); x[15] = gl_mulp2<108>(x[15]);
    } else {   // w16^-k = 2^(192 - 12k)
        x[5] = gl_mulp2<180>(x[5]); x[9] = gl_mulp2<168>(x[9]);  x[13] = gl_mulp2<156>(x[13]);
        x[6] = gl_mulp2<168>(x[6]); x[10] = gl_mulp2<144>(x[10]); x[14] = gl_mulp2<120>(x[14]);
        x[7] = gl_mulp2<156>(x[7]); x[11] = gl_mulp2<120>(x[11]); x[15] = gl_mulp2<84>(x[15]);
    }
#pragma unroll
    for (int c1 = 0; c1 < 4; ++c1) dft4<INV>(x[4 * c1], x[4 * c1 + 1], x[4 * c1 + 2], x[4 * c1 + 3]);
    // x[4 c1 + c2] holds X_{c1 + 4 c2}: transpose the 4x4 in registers
    u64 t;
    t = x[1]; x[1] = x[4]; x[4] = t;   t = x[2]; x[2] = x[8]; x[8] = t;
    t = x[3]; x[3] = x[12]; x[12] = t; t = x[6]; x[6] = x[9]; x[9] = t;
    t = x[7]; x[7] = x[13]; x[13] = t; t = x[11]; x[11] = x[14]; x[14] = t;
}

template <int Q, bool INV>
__device__ __forceinline__ void dftq(u64* x) {
    if constexpr (Q == 2) dft2<INV>(x);
    else if constexpr (Q == 4) dft4a<INV>(x);
    else if constexpr (Q == 8) dft8<INV>(x);
    else dft16<INV>(x);
}


// ---- one radix-Q level of an in-place 2^L-point sub-transform in LDS ---------
// NARR arrays, each holding T interleaved columns: element (i, col) of array
// a sits at arr[a][lpad(i*T + col)]. Sub-block size M = 2^logM = Q*m:
// butterfly (beta, j) reads i = beta*M + j + t*m, t < Q. Forward: DFT_Q, then
// output c times w_M^(j*c). Inverse: times w_M^(-j*c), then inverse DFT_Q.
// The arrays share each butterfly's twiddles and are processed one after the
// other to keep register pressure at one array's Q values.
template <int Q, bool INV, int NARR>
__device__ __forceinline__ void lds_level(u64* const* arr, int L, int logM, int tlog,
                                          const u64* __restrict__ twL) {
    constexpr int LQ = QLog<Q>::v;
    const int logm = logM - LQ;
    const int nbf = 1 << (L - LQ + tlog);
    const int mask = (1 << L) - 1;
    const int tmask = (1 << tlog) - 1;
    const bool tw = logm > 0;                  // the last level (M == Q) has j == 0
    for (int u = threadIdx.x; u < nbf; u += blockDim.x) {
        const int col = u & tmask;
        const int rest = u >> tlog;
        const int j = rest & ((1 << logm) - 1);
        const int base = ((rest >> logm) << logM) + j;
        u64 w[Q];
        if (tw) {
            const int step = j << (L - logM);  // w_M^j = w_{2^L}^(j 2^(L-logM))
#pragma unroll
            for (int c = 1; c < Q; ++c) {
                const int e = (step * c) & mask;
                w[c] = twL[INV ? ((-e) & mask) : e];
            }
        }
        const int a0 = lpad((base << tlog) + col);
#pragma unroll
        for (int a = 0; a < NARR; ++a) {
            u64 x[Q];
#pragma unroll
            for (int t = 0; t < Q; ++t) x[t] = arr[a][a0 + pad_step(t << (logm + tlog))];
            if (!INV) {
                dftq<Q, false>(x);
                if (tw) {
#pragma unroll
                    for (int c = 1; c < Q; ++c) x[c] = gl_mul(x[c], w[c]);
                }
            } else {
                if (tw) {
#pragma unroll
                    for (int c = 1; c < Q; ++c) x[c] = gl_mul(x[c], w[c]);
                }
                dftq<Q, true>(x);
            }
#pragma unroll
            for (int t = 0; t < Q; ++t) arr[a][a0 + pad_step(t << (logm + tlog))] = x[t];
        }
    }
}

// Full 2^L-point sub-transform with radix 2^LR (LR = 3 or 4): forward = a
// radix-2^(L mod LR) level, then radix-2^LR levels (M = 2^L ... 2^LR);
// inverse = the same levels in reverse order. Callers synchronise before;
// this returns after a final barrier.
template <int Q, bool INV, int NARR>
__device__ __forceinline__ void lds_level_q(u64* const* arr, int L, int logM, int tlog, const u64* twL) {
    lds_level<Q, INV, NARR>(arr, L, logM, tlog, twL);
    __syncthreads();
}

template <int LR, bool INV, int NARR>
__device__ __forceinline__ void lds_dft(u64* const* arr, int L, int tlog, const u64* __restrict__ twL) {
    constexpr int RAD = 1 << LR;
    const int r0 = L % LR;
    if (!INV) {
        int logM = L;
        if (r0 == 1) { lds_level_q<2, false, NARR>(arr, L, logM, tlog, twL); logM -= 1; }
        if (r0 == 2) { lds_level_q<4, false, NARR>(arr, L, logM, tlog, twL); logM -= 2; }
        if (r0 == 3) { lds_level_q<8, false, NARR>(arr, L, logM, tlog, twL); logM -= 3; }
        for (; logM >= LR; logM -= LR) lds_level_q<RAD, false, NARR>(arr, L, logM, tlog, twL);
    } else {
        for (int logM = LR; logM <= L - r0; logM += LR) lds_level_q<RAD, true, NARR>(arr, L, logM, tlog, twL);
        if (r0 == 1) lds_level_q<2, true, NARR>(arr, L, L, tlog, twL);
        if (r0 == 2) lds_level_q<4, true, NARR>(arr, L, L, tlog, twL);
        if (r0 == 3) lds_level_q<8, true, NARR>(arr, L, L, tlog, twL);
    }
}

__device__ __forceinline__ u64 digit16(const u32* __restrict__ limbs, size_t nlimbs, size_t g) {
    size_t li = g >> 1;
    return li < nlimbs ? (u64)((limbs[li] >> (16 * (g & 1))) & 0xFFFFu) : 0;
}

// ---- column pass, forward ------------------------------------------------------
// Launch = a class of multiplies sharing n. Workgroup -> (job, operand, tile):
// tile t handles columns b = t*T .. t*T+T-1 of the 2^L1 x 2^L2 view (element
// (i, b) at i*2^L2 + b) of that operand. `jobs == nullptr` means the single
// job `one` passed by value.
template <int T, int LR>
__global__ __launch_bounds__(1024) void k_col_fwd(const MulJob* __restrict__ jobs, MulJob one, int ops,
                                                 int n, int L1, const u64* __restrict__ twL) {
    extern __shared__ u64 lds[];
    const int L2 = n - L1;
    const size_t S = (size_t)1 << L2;
    const int R = 1 << L1;
    const unsigned tiles = (unsigned)(S / T);
    const unsigned bid = xcd_tile_order(blockIdx.x, gridDim.x);
    const unsigned job = bid / (ops * tiles);
    const unsigned rem = bid % (ops * tiles);
    const int op = (int)(rem / tiles);
    const size_t b0 = (size_t)(rem % tiles) * T;
    const MulJob J = job_of(jobs, job, one);
    const u32* __restrict__ src = op ? J.b : J.a;
    const size_t nl = op ? J.nb : J.na;
    u64* __restrict__ dst = op ? J.wsB : J.wsA;
    for (int e = threadIdx.x; e < R * T; e += blockDim.x) {
        const int col = e & (T - 1), i = e / T;
        lds[lpad(e)] = digit16(src, nl, (size_t)i * S + b0 + col);
    }
    __syncthreads();
    u64* arr[1] = {lds};
    lds_dft<LR, false, 1>(arr, L1, QLogT<T>::v, twL + R);
    for (int e = threadIdx.x; e < R * T; e += blockDim.x) {
        const int col = e & (T - 1), i = e / T;
        dst[(size_t)i * S + b0 + col] = lds[lpad(e)];
    }
}

// ---- column pass, inverse ----------------------------------------------------------

// ---- row pass: outer twiddle, forward A (and B), pointwise, inverse, inverse
// outer twiddle times N^-1; result into A ----------------------------------------------
// Row r of the 2^L1 x 2^L2 view holds column-pass frequency c = freq(r); its
// element b is multiplied by w_N^(b c) on the way in and by w_N^(-b c) N^-1 on
// the way out. Thread t handles b = t + k*blockDim, so its twiddles form the
// chain w_N^(c t) * (w_N^(c blockDim))^k: two table lookups per thread.

__device__ __forceinline__ u64 tw_lookup(const u64* __restrict__ hi, const u64* __restrict__ lo, int h, u64 e) {
    return gl_mul(hi[e >> h], lo[e & (((u64)1 << h) - 1)]);
}

template <bool SQUARE, int LR>
__global__ __launch_bounds__(1024) void k_row_mul(const MulJob* __restrict__ jobs, MulJob one, int L1, int L2,
                                                 int lr_col, const u64* __restrict__ twL,
                                                 const u64* __restrict__ ohi, const u64* __restrict__ olo,
                                                 const u64* __restrict__ ohi_inv, const u64* __restrict__ olo_inv,
                                                 int h) {
    extern __shared__ u64 lds[];
    const int R = 1 << L2;
    u64* la = lds;
    u64* lb = lds + lds_words(R);
    const MulJob J = job_of(jobs, blockIdx.x >> L1, one);
    u64* __restrict__ A = J.wsA;
    const u64* __restrict__ B = J.wsB;
    const int row = (int)(blockIdx.x & ((1u << L1) - 1));
    const size_t base = (size_t)row * R;
    const u64 c = freq_rt(row, L1, lr_col);
    const u64 N = (u64)1 << (L1 + L2);
    {
        // exponents mod N (w_N has order N; blockDim may exceed the row length)
        u64 w = tw_lookup(ohi, olo, h, (c * threadIdx.x) & (N - 1));
        const u64 step = tw_lookup(ohi, olo, h, (c * blockDim.x) & (N - 1));
        for (int e = threadIdx.x; e < R; e += blockDim.x) {
            la[lpad(e)] = gl_mul(A[base + e], w);
            if constexpr (!SQUARE) lb[lpad(e)] = gl_mul(B[base + e], w);
            w = gl_mul(w, step);
        }
    }
    __syncthreads();
    if constexpr (SQUARE) {
        u64* arr[1] = {la};
        lds_dft<LR, false, 1>(arr, L2, 0, twL + R);
        for (int e = threadIdx.x; e < R; e += blockDim.x) { u64 v = la[lpad(e)]; la[lpad(e)] = gl_mul(v, v); }
    } else {
        u64* arr[2] = {la, lb};
        lds_dft<LR, false, 2>(arr, L2, 0, twL + R);
        for (int e = threadIdx.x; e < R; e += blockDim.x) la[lpad(e)] = gl_mul(la[lpad(e)], lb[lpad(e)]);
    }
    __syncthreads();
    u64* arr1[1] = {la};
    lds_dft<LR, true, 1>(arr1, L2, 0, twL + R);
    {
        u64 w = tw_lookup(ohi_inv, olo_inv, h, (c * threadIdx.x) & (N - 1));     // includes N^-1
        const u64 step = tw_lookup(ohi, olo, h, (N - c * blockDim.x) & (N - 1)); // w_N^(-c blockDim)
        for (int e = threadIdx.x; e < R; e += blockDim.x) {
            A[base + e] = gl_mul(la[lpad(e)], w);
            w = gl_mul(w, step);
        }
    }
}

template <int T>
__global__ __launch_bounds__(CARRY_NT) void k_carry_reduce(const MulJob* __restrict__ jobs, MulJob one,
                                                           size_t nblk_max) {
    using G = CarryGeom<T>;
    constexpr int KL = G::KL, CPT = G::CPT;
    __shared__ CarryFn wsum[CARRY_NT / 64];
    __shared__ u32 sD[G::LIMBS + G::LIMBS / 32];
    const MulJob J = job_of(jobs, blockIdx.x / nblk_max, one);
    const CarryJob C = carry_job<T>(J, nblk_max);
    const size_t blk = blockIdx.x % nblk_max;
    if (blk >= C.nblk) return;                       // whole workgroup leaves together
    stage_in<T>(C, blk * G::LIMBS, sD);
    __syncthreads();
    const size_t ch0 = (blk * CARRY_NT + threadIdx.x) * CPT;
    CarryFn f = cf_pad();
    bool first = true;
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        const size_t ch = ch0 + u;
        if (ch < C.nchunks) {
            u32 D[KL];
#pragma unroll
            for (int q = 0; q < KL; ++q) D[q] = sD[cpad((threadIdx.x * CPT + u) * KL + q)];
            const CarryFn g = chunk_fn<KL>(D, C.C[ch]);
            f = first ? g : cf_compose(g, f);
            first = false;
        } else if (!first) {
            f = cf_compose(cf_pad(), f);
        }
    }
    f = block_scan(f, wsum);
    if (threadIdx.x == CARRY_NT - 1) C.agg[blk] = f;
    if (blk == 0 && threadIdx.x == 0) *J.out_len = 0;
}

// The same reduce when the producer wrote the chunk thresholds: a chunk's
// function is (TH, C, C + 1), 16 bytes per chunk and no digits. (Folding the
// aggregate scan into the last workgroup to finish needs device-scope
// fences, which write back the L2 of each XCD; a separate launch is cheaper.)
template <int T>
__global__ __launch_bounds__(CARRY_NT) void k_carry_reduce_th(const MulJob* __restrict__ jobs, MulJob one,
                                                              size_t nblk_max) {
    using G = CarryGeom<T>;
    constexpr int CPT = G::CPT;
    __shared__ CarryFn wsum[CARRY_NT / 64];
    const MulJob J = job_of(jobs, blockIdx.x / nblk_max, one);
    const CarryJob C = carry_job<T>(J, nblk_max);
    const size_t blk = blockIdx.x % nblk_max;
    if (blk >= C.nblk) return;                       // whole workgroup leaves together
    const size_t ch0 = (blk * CARRY_NT + threadIdx.x) * CPT;
    CarryFn f = cf_pad();
    bool first = true;
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        const size_t ch = ch0 + u;
        if (ch < C.nchunks) {
            const u64 c = C.C[ch];
            const CarryFn g{C.TH[ch], c, c + 1};
            f = first ? g : cf_compose(g, f);
            first = false;
        } else if (!first) {
            f = cf_compose(cf_pad(), f);
        }
    }
    f = block_scan(f, wsum);
    if (threadIdx.x == CARRY_NT - 1) C.agg[blk] = f;
    if (blk == 0 && threadIdx.x == 0) *J.out_len = 0;
}

// The carry into carry workgroup blk: the composition of the aggregates of
// workgroups blk-1, blk-2, ... down to the nearest constant one (a constant
// function ignores its carry-in) or down to workgroup 0 (carry-in 0), taken
// in windows of CARRY_NT aggregates, one block scan per window. For a
// product that is not mostly all-ones limbs the aggregate of blk-1 is
// already constant, and the caller does not come here. Every thread of the
// workgroup calls it (it holds barriers); slot is LDS for three words.
__device__ __forceinline__ bool cf_constant(const CarryFn& f) { return f.thr == ~(u64)0 || f.v0 == f.v1; }

__device__ __noinline__ u64 carry_in_lookback(const CarryFn* __restrict__ agg, size_t blk, CarryFn* wsum,
                                              u64* slot) {
    CarryFn acc = cf_pad();
    bool have = false;
    size_t hi = blk;
    for (;;) {
        const size_t lo = hi > (size_t)CARRY_NT ? hi - CARRY_NT : 0;
        const size_t t = lo + threadIdx.x;
        const CarryFn F = block_scan(t < hi ? agg[t] : cf_pad(), wsum);   // thread 0 = workgroup lo
        if (threadIdx.x == hi - lo - 1) { slot[0] = F.thr; slot[1] = F.v0; slot[2] = F.v1; }
        __syncthreads();
        const CarryFn Fw{slot[0], slot[1], slot[2]};
        __syncthreads();
        acc = have ? cf_compose(acc, Fw) : Fw;
        have = true;
        if (lo == 0 || cf_constant(Fw)) return cf_apply(acc, 0);
        hi = lo;
    }
}

template <int T>
__global__ __launch_bounds__(CARRY_NT) void k_carry_apply(const MulJob* __restrict__ jobs, MulJob one,
                                                          size_t nblk_max) {
    using G = CarryGeom<T>;
    constexpr int KL = G::KL, CPT = G::CPT;
    __shared__ CarryFn wsum[CARRY_NT / 64];
    __shared__ u64 outs[CARRY_NT];
    __shared__ u64 wmax[CARRY_NT / 64];
    __shared__ u32 sD[G::LIMBS + G::LIMBS / 32];
    const MulJob J = job_of(jobs, blockIdx.x / nblk_max, one);
    const CarryJob C = carry_job<T>(J, nblk_max);
    // Highest limbs first: the block holding the top nonzero limb usually
    // runs early, and later blocks then skip the atomic on out_len.
    const size_t rb = blockIdx.x % nblk_max;
    if (rb >= C.nblk) return;
    const size_t blk = C.nblk - 1 - rb;
    const size_t L0 = blk * G::LIMBS;
    // the aggregate below and the carry-outs are loaded unconditionally (at
    // clamped indices) so that they are in flight with the staging loads
    const u64* pa = (const u64*)(C.agg + (blk > 0 ? blk - 1 : 0));
    const CarryFn pload{gld64(pa), gld64(pa + 1), gld64(pa + 2)};
    const size_t ch0 = (blk * CARRY_NT + threadIdx.x) * CPT;
    u64 Cc[CPT];
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        const size_t ch = ch0 + u;
        Cc[u] = gld64(C.C + (ch < C.nchunks ? ch : C.nchunks - 1));
    }
    stage_in<T>(C, L0, sD);
    const CarryFn prev = blk > 0 ? pload : cf_pad();
#pragma unroll
    for (int u = 0; u < CPT; ++u) if (ch0 + u >= C.nchunks) Cc[u] = 0;
    __syncthreads();
    CarryFn f = cf_pad();
    bool first = true;
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        const size_t ch = ch0 + u;
        if (ch < C.nchunks) {
            u32 D[KL];
#pragma unroll
            for (int q = 0; q < KL; ++q) D[q] = sD[cpad((threadIdx.x * CPT + u) * KL + q)];
            const CarryFn g = chunk_fn<KL>(D, Cc[u]);
            f = first ? g : cf_compose(g, f);
            first = false;
        } else if (!first) {
            f = cf_compose(cf_pad(), f);
        }
    }
    // this workgroup's carry-in, from the aggregates k_carry_reduce wrote
    u64 bc = 0;
    if (blk > 0) bc = cf_constant(prev) ? prev.v0 : carry_in_lookback(C.agg, blk, wsum, outs);
    CarryFn inc = block_scan(f, wsum);
    outs[threadIdx.x] = cf_apply(inc, bc);
    __syncthreads();
    u64 cin = threadIdx.x == 0 ? bc : outs[threadIdx.x - 1];
    u64 top = 0;
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        const size_t ch = ch0 + u;
        if (ch >= C.nchunks) break;
        const int w0 = (threadIdx.x * CPT + u) * KL;
        u32 D[KL];
#pragma unroll
        for (int q = 0; q < KL; ++q) D[q] = sD[cpad(w0 + q)];
        const u64 next = cf_apply(chunk_fn<KL>(D, Cc[u]), cin);
        // D + cin over KL limbs (the overflow is next's +1)
        const u64 lo = (u64)D[0] | ((u64)D[1] << 32);
        const u64 s = lo + cin;
        u32 carry = s < lo ? 1u : 0u;
        D[0] = (u32)s; D[1] = (u32)(s >> 32);
#pragma unroll
        for (int q = 2; q < KL; ++q) { const u32 t = D[q] + carry; carry = (t < carry) ? 1u : 0u; D[q] = t; }
        const size_t l0 = ch * KL;
#pragma unroll
        for (int q = 0; q < KL; ++q) {
            sD[cpad(w0 + q)] = D[q];
            if (D[q] && l0 + q < C.nlimbs) top = l0 + q + 1;
        }
        cin = next;
    }
    __syncthreads();
    stage_out<T>(C, L0, sD);
    top = wave_max(top);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = top;
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 m = 0;
        for (int w = 0; w < CARRY_NT / 64; ++w) m = wmax[w] > m ? wmax[w] : m;
        if (m && m > __atomic_load_n((unsigned long long*)J.out_len, __ATOMIC_RELAXED))
            atomicMax((unsigned long long*)J.out_len, (unsigned long long)m);
    }
}

// Inverse column pass with the first half of the carry: after the DIT column
// sub-transforms every row segment of T coefficients is a carry chunk; its
// digits D go to the product buffer and its carry-out to the carry scratch.
template <int T, int LR>
__global__ __launch_bounds__(1024) void k_col_inv(const MulJob* __restrict__ jobs, MulJob one, int n, int L1,
                                                 const u64* __restrict__ twL, size_t nblk_max) {
    constexpr int KL = T / 2;
    extern __shared__ u64 lds[];
    const int L2 = n - L1;
    const size_t S = (size_t)1 << L2;
    const int R = 1 << L1;
    const unsigned tiles = (unsigned)(S / T);
    const unsigned bid = xcd_tile_order(blockIdx.x, gridDim.x);
    const MulJob J = job_of(jobs, bid / tiles, one);
    const u64* __restrict__ buf = J.wsA;
    const size_t b0 = (size_t)(bid % tiles) * T;
    for (int e = threadIdx.x; e < R * T; e += blockDim.x) {
        const int col = e & (T - 1), i = e / T;
        lds[lpad(e)] = buf[(size_t)i * S + b0 + col];
    }
    __syncthreads();
    u64* arr[1] = {lds};
    lds_dft<LR, true, 1>(arr, L1, QLogT<T>::v, twL + R);
    const CarryJob C = carry_job<T>(J, nblk_max);
    for (int i = threadIdx.x; i < R; i += blockDim.x) {
        const size_t ch = ((size_t)i * S + b0) / T;
        if (ch >= C.nchunks) continue;
        const size_t l0 = ch * KL;
        u64 carry = 0;
        if constexpr (KL % 4 == 0 && KL <= 8) {
            u32 D[KL];
#pragma unroll
            for (int q = 0; q < KL; ++q) {
                const u64 c0 = carry + lds[lpad(i * T + 2 * q)];
                carry = c0 >> 16;
                const u64 c1 = carry + lds[lpad(i * T + 2 * q + 1)];
                carry = c1 >> 16;
                D[q] = (u32)(c0 & 0xFFFFu) | ((u32)(c1 & 0xFFFFu) << 16);
            }
            u32* dst = C.out + l0;
            if (l0 + KL <= C.nlimbs && ((uintptr_t)dst & 15) == 0) {
#pragma unroll
                for (int v = 0; v < KL / 4; ++v) ((uint4*)dst)[v] = uint4{D[4 * v], D[4 * v + 1], D[4 * v + 2], D[4 * v + 3]};
            } else {
#pragma unroll
                for (int q = 0; q < KL; ++q)
                    if (l0 + q < C.nlimbs) dst[q] = D[q];
            }
        } else {
#pragma unroll
            for (int q = 0; q < KL; ++q) {
                const u64 c0 = carry + lds[lpad(i * T + 2 * q)];
                carry = c0 >> 16;
                const u64 c1 = carry + lds[lpad(i * T + 2 * q + 1)];
                carry = c1 >> 16;
                if (l0 + q < C.nlimbs) C.out[l0 + q] = (u32)(c0 & 0xFFFFu) | ((u32)(c1 & 0xFFFFu) << 16);
            }
        }
        C.C[ch] = carry;
    }
}

// Carry of one product whose 2^n coefficients (scaled) sit in LDS: each of the
// blockDim.x threads takes k = ceil(nlimbs / blockDim.x) consecutive limbs.
__device__ __forceinline__ void block_carry(const u64* la, size_t nlimbs, u32* __restrict__ out, u64* __restrict__ out_len,
                            CarryFn* wsum, u64* outs) {
    const size_t k = (nlimbs + blockDim.x - 1) / blockDim.x;
    const size_t l0 = (size_t)threadIdx.x * k;
    const size_t l1 = l0 + k < nlimbs ? l0 + k : nlimbs;
    CarryFn f = cf_pad();
    if (l0 < nlimbs) {
        u64 carry = 0, lo = 0;
        bool ones = true;
        for (size_t l = l0; l < l1; ++l) {
            u64 cur = carry + la[lpad((int)(2 * l))];
            u32 lo16 = (u32)(cur & 0xFFFFu);
            carry = cur >> 16;
            u64 cur2 = carry + la[lpad((int)(2 * l + 1))];
            carry = cur2 >> 16;
            u32 limb = lo16 | ((u32)(cur2 & 0xFFFFu) << 16);
            if (l - l0 < 2) lo |= (u64)limb << (32 * (l - l0));
            else ones &= limb == 0xFFFFFFFFu;
        }
        u64 thr;
        if (l1 - l0 == 1) thr = ((u64)1 << 32) - lo;
        else thr = (ones && lo != 0) ? (u64)(0 - lo) : ~(u64)0;
        f = CarryFn{thr, carry, carry + 1};
    }
    CarryFn inc = block_scan(f, wsum);
    outs[threadIdx.x] = cf_apply(inc, 0);
    __syncthreads();
    u64 carry = threadIdx.x == 0 ? 0 : outs[threadIdx.x - 1];
    u64 top = 0;
    for (size_t l = l0; l < l1; ++l) {
        u64 cur = carry + la[lpad((int)(2 * l))];
        u32 lo16 = (u32)(cur & 0xFFFFu);
        carry = cur >> 16;
        u64 cur2 = carry + la[lpad((int)(2 * l + 1))];
        carry = cur2 >> 16;
        u32 limb = lo16 | ((u32)(cur2 & 0xFFFFu) << 16);
        out[l] = limb;
        if (limb) top = l + 1;
    }
    top = wave_max(top);
    __syncthreads();                                  // outs[] is reused below
    if ((threadIdx.x & 63) == 0) outs[threadIdx.x >> 6] = top;
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 m = 0;
        for (int w = 0; w < (int)(blockDim.x / 64); ++w) m = outs[w] > m ? outs[w] : m;
        *out_len = m;
    }
}

// ---- whole multiply in one workgroup (N = 2^n <= 2^13), carry included -------------
template <bool SQUARE, int NT, int LR>
__global__ __launch_bounds__(NT) void k_small_mul(const MulJob* __restrict__ jobs, MulJob one, int n,
                                                  const u64* __restrict__ twL, u64 ninv) {
    extern __shared__ u64 lds[];
    __shared__ CarryFn wsum[NT / 64];
    __shared__ u64 outs[NT];
    const MulJob J = job_of(jobs, blockIdx.x, one);
    const int R = 1 << n;
    u64* la = lds;
    u64* lb = lds + lds_words(R);
    for (int e = threadIdx.x; e < R; e += blockDim.x) {
        la[lpad(e)] = digit16(J.a, J.na, e);
        if constexpr (!SQUARE) lb[lpad(e)] = digit16(J.b, J.nb, e);
    }
    __syncthreads();
    if constexpr (SQUARE) {
        u64* arr[1] = {la};
        lds_dft<LR, false, 1>(arr, n, 0, twL + R);
        for (int e = threadIdx.x; e < R; e += blockDim.x) { u64 v = la[lpad(e)]; la[lpad(e)] = gl_mul(v, v); }
    } else {
        u64* arr[2] = {la, lb};
        lds_dft<LR, false, 2>(arr, n, 0, twL + R);
        for (int e = threadIdx.x; e < R; e += blockDim.x) la[lpad(e)] = gl_mul(la[lpad(e)], lb[lpad(e)]);
    }
    __syncthreads();
    u64* arr1[1] = {la};
    lds_dft<LR, true, 1>(arr1, n, 0, twL + R);
    for (int e = threadIdx.x; e < R; e += blockDim.x) la[lpad(e)] = gl_mul(la[lpad(e)], ninv);
    __syncthreads();
    block_carry(la, J.na + J.nb, J.out, J.out_len, wsum, outs);
}

__global__ void k_zero_len(u64* out_len) { *out_len = 0; }

// ---- host launchers -------------------------------------------------------------------
// ---- shapes: the default split per n, overridable by pdh_tune_set / autotune ----
// Defaults measured on MI355X with the shape sweep (the autotuner's search
// space): {L1, T} for n = 14..24; Goldilocks radix 8 for both passes.
static NttShape default_shape_eng(int n, int eng) {
    NttShape s{n, 0, n, 1, 3, 3, 0};
    if (n <= SMALL_MAX_LOG) {
        // one workgroup: Goldilocks over 2^n points, or the three primes over
        // 2^(n-1) points (radix-8 levels through LDS)
        if (eng == 1) { s.eng = 1; s.L2 = n - 1; }
        return s;
    }
    if (eng == 0) {
        static const int tab[MAX_LOG + 1][2] = {
            {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0},
            {5, 16}, {6, 8}, {7, 4}, {8, 8}, {8, 8}, {8, 16}, {8, 16}, {9, 16}, {9, 32}, {10, 16}, {11, 8}};
        s.L1 = tab[n][0];
        s.L2 = n - s.L1;
        s.T = tab[n][1];
        return s;
    }
    // three primes over 2^(n-1) points: {L1, T} for n = 14..24, measured on
    // MI355X with the shape sweep and the autotuner on distinct operands
    // (n = 23: 154.5 us at {9, 16} against 165.2 us at {8, 16});
    // lr_row 4: rows as register-resident radix-16 sets and, where the tile
    // has a register kernel, columns as register-resident radix-8 sets
    static const int tab3[MAX_LOG + 1][2] = {
        {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0},
        {4, 8}, {5, 8}, {6, 8}, {6, 8}, {5, 16}, {6, 16}, {6, 8}, {8, 8}, {8, 16}, {9, 16}, {9, 16}};
    s.eng = 1;
    s.lr_row = 4;
    s.L1 = tab3[n][0];
    s.L2 = n - 1 - s.L1;
    s.T = tab3[n][1];
    return s;
}
// Engine 1 everywhere from N = 4: on batches of equal products the
// one-workgroup three-prime kernel ran 1.6-1.9x the Goldilocks one at every
// N = 2^4..2^13 (2^9: 137 against 81 M/s; 2^13: 6.7 against 4.1 M/s).
static NttShape default_shape(int n) { return default_shape_eng(n, n >= 2 ? 1 : 0); }

static NttShape g_shapes[MAX_LOG + 1];
static bool g_shapes_init = false;

NttShape ntt_shape(int n) {
    if (!g_shapes_init) {
        for (int k = 0; k <= MAX_LOG; ++k) g_shapes[k] = default_shape(k);
        g_shapes_init = true;
    }
    return g_shapes[n];
}

bool shape_valid(const NttShape& s) {
    if (s.n < 1 || s.n > MAX_LOG) return false;
    if (s.lr_col < 3 || s.lr_col > 4 || s.lr_row < 3 || s.lr_row > 4) return false;
    if (s.L1 == 0)
        return s.n <= SMALL_MAX_LOG && s.T == 1 && ((s.eng == 0 && s.L2 == s.n) ||
                                                   (s.eng == 1 && s.n >= 2 && s.L2 == s.n - 1 && s.lr_col == 3 && s.lr_row == 3));
    if (s.n <= SMALL_MAX_LOG) return false;
    if (s.T != 4 && s.T != 8 && s.T != 16 && s.T != 32 && s.T != 64) return false;
    if (s.T > (1 << s.L2) || s.L1 < 1 || s.L2 < 3) return false;
    if (s.eng == 0)
        return s.L2 <= ROW_MAX_LOG && s.L1 + s.L2 == s.n && (size_t)lds_words((1 << s.L1) * s.T) * 8 <= 147456;
    // three primes: u32 rows of up to 2^14 points (two arrays fill the 144 KB);
    // radix-8 columns. lr_row 3: both passes as radix-8 levels through LDS;
    // lr_row 4 (rows of 2^8 points or more): register-resident passes, rows
    // of radix-16 sets and columns of radix-8 sets (tiles of 2^9 to 2^13
    // elements; other tiles keep the LDS-level column pass)
    if (s.eng != 1 || s.lr_col != 3 || s.T > 32 || s.L2 > ROW_MAX_LOG + 1) return false;
    if (s.lr_row == 4 && s.L2 < 8) return false;
    return s.L1 + s.L2 == s.n - 1 && (size_t)3 * lds_words((1 << s.L1) * s.T) * 4 <= 147456;
}

// eng = -1: the default engine for n
NttShape default_shape_for(int n, int eng) { return eng < 0 ? default_shape(n) : default_shape_eng(n, eng); }

static bool g_custom[MAX_LOG + 1];   // shape set by pdh_tune_set / pdh_tune_engine / autotune

bool set_shape(const NttShape& s, bool custom) {
    if (!shape_valid(s)) return false;
    ntt_shape(s.n);
    g_shapes[s.n] = s;
    g_custom[s.n] = custom;
    return true;
}

bool shape_custom(int n) { return g_custom[n]; }

// A class of many independent products already fills the GPU, so for the
// three-prime engine at n = 14..18 short columns and wide tiles (fewer
// levels, fewer workgroups) beat the single-product table: the 4,096-pair
// batch (limbs 64..65,536) takes 10.9 ms with L1 = 4, T = 32 against
// 15.5 ms with the single-product shapes. A shape set explicitly for n is
// used as it is.
NttShape batch_shape(int n, unsigned njobs) {
    const NttShape s = ntt_shape(n);
    if (g_custom[n] || njobs < 16 || s.eng != 1 || s.L1 == 0 || n > 18) return s;
    NttShape t = s;
    t.L1 = 4; t.L2 = n - 1 - 4; t.T = 32;
    return shape_valid(t) ? t : s;
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

// Every (tile, radix) instantiation the launcher can pick.
#define PDH_FOR_COL(X) X(4, 3) X(4, 4) X(8, 3) X(8, 4) X(16, 3) X(16, 4) X(32, 3) X(32, 4) X(64, 3) X(64, 4)

// Dynamic-LDS limits are raised once, before any launch (and never while a
// caller may be capturing a graph).
hipError_t init_kernel_attrs() {
    // 144 KB: the largest dynamic tile any launch uses (2 x 2^13 words + 1/8
    // padding); static __shared__ plus dynamic must stay within the 160 KB per
    // workgroup, which the runtime checks here.
    constexpr size_t MAXLDS = 147456;
#define PDH_SET_COL(TT, RR) PDH_TRY(set_lds(k_col_fwd<TT, RR>, MAXLDS)); PDH_TRY(set_lds(k_col_inv<TT, RR>, MAXLDS));
    PDH_FOR_COL(PDH_SET_COL)
#undef PDH_SET_COL
    PDH_TRY(set_lds(k_row_mul<true, 3>, MAXLDS)); PDH_TRY(set_lds(k_row_mul<false, 3>, MAXLDS));
    PDH_TRY(set_lds(k_row_mul<true, 4>, MAXLDS)); PDH_TRY(set_lds(k_row_mul<false, 4>, MAXLDS));
    PDH_TRY(set_lds((k_small_mul<true, 256, 3>), MAXLDS)); PDH_TRY(set_lds((k_small_mul<false, 256, 3>), MAXLDS));
    PDH_TRY(set_lds((k_small_mul<true, 1024, 3>), MAXLDS)); PDH_TRY(set_lds((k_small_mul<false, 1024, 3>), MAXLDS));
    PDH_TRY(set_lds((k_small_mul<true, 256, 4>), MAXLDS)); PDH_TRY(set_lds((k_small_mul<false, 256, 4>), MAXLDS));
    PDH_TRY(set_lds((k_small_mul<true, 1024, 4>), MAXLDS)); PDH_TRY(set_lds((k_small_mul<false, 1024, 4>), MAXLDS));
    return hipSuccess;
}

template <int T, int LR>
static hipError_t launch_cols(const MulJob* d_jobs, const MulJob& one, unsigned njobs, int ops,
                              const u64* twL, const OuterTables& ot, const NttShape& sh, hipStream_t s,
                              bool inverse, size_t nblk_max) {
    const size_t lds = (size_t)lds_words((1 << sh.L1) * T) * 8;
    const unsigned tiles = (unsigned)(((size_t)1 << sh.L2) / T);
    if (!inverse) {
        PDH_LAUNCH("k_col_fwd", s, (k_col_fwd<T, LR>), dim3(njobs * ops * tiles), dim3(1024), lds, s, d_jobs, one,
                   ops, sh.n, sh.L1, twL);
    } else {
        PDH_LAUNCH("k_col_inv", s, (k_col_inv<T, LR>), dim3(njobs * tiles), dim3(1024), lds, s, d_jobs, one,
                   sh.n, sh.L1, twL, nblk_max);
        PDH_TRY(launch_carry_scan(T / 2, d_jobs, one, njobs, nblk_max, s));
    }
    return hipSuccess;
}

static hipError_t launch_cols_T(const MulJob* d_jobs, const MulJob& one, unsigned njobs, int ops,
                                const u64* twL, const OuterTables& ot, const NttShape& sh, hipStream_t s,
                                bool inverse, size_t nblk_max) {
#define PDH_CASE_COL(TT, RR) if (sh.T == TT && sh.lr_col == RR) return launch_cols<TT, RR>(d_jobs, one, njobs, ops, twL, ot, sh, s, inverse, nblk_max);
    PDH_FOR_COL(PDH_CASE_COL)
#undef PDH_CASE_COL
    return hipErrorInvalidValue;
}

template <int T>
static hipError_t launch_carry_scan_T(const MulJob* d_jobs, const MulJob& one, unsigned njobs, size_t nblk_max,
                                      hipStream_t s, bool th) {
    const dim3 cgrid((unsigned)(njobs * nblk_max));
    if (th) PDH_LAUNCH("k_carry_reduce_th", s, k_carry_reduce_th<T>, cgrid, dim3(CARRY_NT), 0, s, d_jobs, one, nblk_max);
    else PDH_LAUNCH("k_carry_reduce", s, k_carry_reduce<T>, cgrid, dim3(CARRY_NT), 0, s, d_jobs, one, nblk_max);
    PDH_LAUNCH("k_carry_apply", s, k_carry_apply<T>, cgrid, dim3(CARRY_NT), 0, s, d_jobs, one, nblk_max);
    return hipSuccess;
}

// chunks of kl limbs = 2 kl 16-bit coefficients
hipError_t launch_carry_scan(int kl, const MulJob* d_jobs, const MulJob& one, unsigned njobs, size_t nblk_max,
                             hipStream_t s, bool th) {
    switch (kl) {
    case 2: return launch_carry_scan_T<4>(d_jobs, one, njobs, nblk_max, s, th);
    case 4: return launch_carry_scan_T<8>(d_jobs, one, njobs, nblk_max, s, th);
    case 8: return launch_carry_scan_T<16>(d_jobs, one, njobs, nblk_max, s, th);
    case 16: return launch_carry_scan_T<32>(d_jobs, one, njobs, nblk_max, s, th);
    case 32: return launch_carry_scan_T<64>(d_jobs, one, njobs, nblk_max, s, th);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_zero_len(uint64_t* out_len, hipStream_t s) {
    PDH_LAUNCH("k_zero_len", s, k_zero_len, dim3(1), dim3(1), 0, s, out_len);
    return hipSuccess;
}

// Carry blocks: every shape's workgroup covers CARRY_NT * 32 limbs.
size_t carry_nblk(size_t nlimbs) {
    return (nlimbs + CARRY_NT * 32 - 1) / (CARRY_NT * 32);   // CarryGeom::LIMBS
}

size_t carry_ws_bytes(size_t nlimbs) {
    const size_t nblk = carry_nblk(nlimbs);
    return 2 * (((nlimbs + 1) / 2 + 3) / 4 * 4) * sizeof(u64) + nblk * sizeof(CarryFn) + 256;
}

template <bool SQ, int NT>
static hipError_t launch_small(const MulJob* d_jobs, const MulJob& one, unsigned njobs, int n, int lr, size_t lds,
                               const u64* twL, u64 ninv, hipStream_t s) {
    if (lr == 3) PDH_LAUNCH("k_small_mul", s, (k_small_mul<SQ, NT, 3>), dim3(njobs), dim3(NT), lds, s, d_jobs, one, n, twL, ninv);
    else PDH_LAUNCH("k_small_mul", s, (k_small_mul<SQ, NT, 4>), dim3(njobs), dim3(NT), lds, s, d_jobs, one, n, twL, ninv);
    return hipSuccess;
}

// All jobs of a launch share n. For n > SMALL_MAX_LOG each job needs wsA, wsB of
// 2^n words and carry scratch of carry_ws_bytes(2^(n-1)) (or of its own
// nlimbs when it is launched alone: nblk_max is the per-job carry block count).
hipError_t launch_class(const MulJob* d_jobs, const MulJob& one, unsigned njobs, int n, bool square,
                        size_t nblk_max, const u64* twL, const OuterTables& ot, hipStream_t s,
                        const NttShape* shape) {
    const NttShape sh = shape ? *shape : ntt_shape(n);
    const size_t N = (size_t)1 << n;
    if (sh.L1 == 0 && sh.eng == 1) return launch_small3(d_jobs, one, njobs, n, square, ot.tw3, s);
    if (sh.L1 == 0) {
        const size_t lds = (size_t)(square ? 1 : 2) * lds_words((int)N) * 8;
        const u64 ninv = gl_inv((u64)N);
        // one workgroup per job: 1024 threads once the tile needs most of the LDS
        if (n >= 12) return square ? launch_small<true, 1024>(d_jobs, one, njobs, n, sh.lr_row, lds, twL, ninv, s)
                                   : launch_small<false, 1024>(d_jobs, one, njobs, n, sh.lr_row, lds, twL, ninv, s);
        return square ? launch_small<true, 256>(d_jobs, one, njobs, n, sh.lr_row, lds, twL, ninv, s)
                      : launch_small<false, 256>(d_jobs, one, njobs, n, sh.lr_row, lds, twL, ninv, s);
    }
    if (sh.eng == 1) return launch_class3(d_jobs, one, njobs, sh, square, nblk_max, ot, s);
    PDH_TRY(launch_cols_T(d_jobs, one, njobs, square ? 1 : 2, twL, ot, sh, s, false, nblk_max));
    const size_t lds = (size_t)(square ? 1 : 2) * lds_words(1 << sh.L2) * 8;
    const dim3 grid((unsigned)(njobs << sh.L1));
    if (sh.lr_row == 3) {
        if (square) PDH_LAUNCH("k_row_sqr", s, (k_row_mul<true, 3>), grid, dim3(1024), lds, s, d_jobs, one, sh.L1, sh.L2, sh.lr_col, twL, ot.hi, ot.lo, ot.hi_inv, ot.lo_inv, ot.h);
        else PDH_LAUNCH("k_row_mul", s, (k_row_mul<false, 3>), grid, dim3(1024), lds, s, d_jobs, one, sh.L1, sh.L2, sh.lr_col, twL, ot.hi, ot.lo, ot.hi_inv, ot.lo_inv, ot.h);
    } else {
        if (square) PDH_LAUNCH("k_row_sqr", s, (k_row_mul<true, 4>), grid, dim3(1024), lds, s, d_jobs, one, sh.L1, sh.L2, sh.lr_col, twL, ot.hi, ot.lo, ot.hi_inv, ot.lo_inv, ot.h);
        else PDH_LAUNCH("k_row_mul", s, (k_row_mul<false, 4>), grid, dim3(1024), lds, s, d_jobs, one, sh.L1, sh.L2, sh.lr_col, twL, ot.hi, ot.lo, ot.hi_inv, ot.lo_inv, ot.h);
    }
    return launch_cols_T(d_jobs, one, njobs, 1, twL, ot, sh, s, true, nblk_max);
}

hipError_t launch_mul(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, bool square,
                      uint32_t* out, uint64_t* out_len, uint64_t* wsA, uint64_t* wsB, uint64_t* carry_ws,
                      const uint64_t* twL, const OuterTables& ot, const NttShape& sh, hipStream_t s) {
    MulJob one{a, square ? a : b, out, out_len, na, square ? na : nb, wsA, wsB, carry_ws};
    return launch_class(nullptr, one, 1, sh.n, square, carry_nblk(na + nb), twL, ot, s, &sh);
}

} // namespace pdh
