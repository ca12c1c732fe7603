// ============================================================================
//  ntt_common.h -- device helpers shared by the Goldilocks and three-prime
//  transforms: LDS padding, digit-reversed frequency maps, XCD-aware tile
//  order, and the chunked carry (carry functions, per-job carry scratch,
//  LDS staging) that both engines' inverse column passes feed.
// ============================================================================
#pragma once
#include "gl_field.h"
#include "pdh_internal.h"

namespace pdh {

// Global-memory accesses through pointers read from job descriptors. Such
// pointers are generic, and generic (flat) loads count on the LDS counter as
// well as the vector-memory one, so every wait for an LDS exchange would also
// wait for each global load still in flight; the casts let the compiler emit
// global loads and stores instead.
typedef __attribute__((address_space(1))) u32 g_u32;
__device__ __forceinline__ u32 gld(const u32* p) { return *(const g_u32*)p; }
__device__ __forceinline__ void gst(u32* p, u32 v) { *(g_u32*)p = v; }
typedef u32 v4u __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) v4u g_v4u;
__device__ __forceinline__ uint4 gld4(const uint4* p) {
    const v4u v = *(const g_v4u*)p;
    return uint4{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ void gst4(uint4* p, uint4 v) { *(g_v4u*)p = v4u{v.x, v.y, v.z, v.w}; }
typedef __attribute__((address_space(1))) u64 g_u64;
__device__ __forceinline__ u64 gld64(const u64* p) { return *(const g_u64*)p; }
// Keeps the machine scheduler from moving instructions across this point
// (used to issue a group of loads before the arithmetic that consumes them).
__device__ __forceinline__ void sched_fence() { __builtin_amdgcn_sched_barrier(0); }

// The job descriptor of a workgroup, loaded through the constant address
// space (scalar loads into SGPRs): from the job array when one is given, else
// the single job passed as a kernel argument. A plain `jobs ? jobs[i] : one`
// selects between two generic pointers and loads the descriptor into VGPRs.
__device__ __forceinline__ MulJob job_of(const MulJob* jobs, size_t i, const MulJob& one) {
    typedef __attribute__((address_space(4))) const MulJob c_job;
    if (!jobs) return one;
    const c_job* p = (const c_job*)jobs + i;
    MulJob J;
    J.a = p->a; J.b = p->b; J.out = p->out; J.out_len = p->out_len; J.na = p->na; J.nb = p->nb;
    J.wsA = p->wsA; J.wsB = p->wsB; J.carry = p->carry;
    return J;
}

__device__ __forceinline__ int lpad(int e) { return e + (e >> 3); }
__host__ __device__ constexpr int lds_words(int e) { return e + (e >> 3); }

// LDS offset between element e0 of a butterfly and element e0 + d, d = t D
// with D = 2^(logm + tlog) the butterfly stride: lpad(e0 + d) - lpad(e0) =
// d + (d >> 3). For D >= 8 this is exact since D is a multiple of 8; for
// D < 8 the butterfly starts at e0 = (beta M + j) T + col with beta M T a
// multiple of Q D >= 8 and j T + col < D, so e0 mod 8 < D and adding the
// multiple d of D (D | 8) crosses a multiple of 8 exactly when d does. The
// offset is wave-uniform: one add per element instead of recomputing lpad.
__device__ __forceinline__ int pad_step(int d) { return d + (d >> 3); }

template <int T> struct QLogT { static constexpr int v = T == 1 ? 0 : T == 2 ? 1 : T == 4 ? 2 : T == 8 ? 3 : T == 16 ? 4 : T == 32 ? 5 : 6; };
template <int Q> struct QLog { static constexpr int v = Q == 2 ? 1 : Q == 4 ? 2 : Q == 8 ? 3 : 4; };

// Frequency held at position i after the forward sub-transform: the position's
// digits (widths L mod LR, LR, LR, ... from the top) are the frequency's
// digits from the bottom.
template <int LR>
__device__ __forceinline__ u64 freq_of_pos(int i, int L) {
    const int r0 = L % LR;
    int left = L, sh = 0;
    u64 c = 0;
    if (r0) { left -= r0; c = (u64)((i >> left) & ((1 << r0) - 1)); sh = r0; }
    while (left > 0) { left -= LR; c |= (u64)((i >> left) & ((1 << LR) - 1)) << sh; sh += LR; }
    return c;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2). The column passes read and write T-word row segments, so neighbouring
// tiles share HBM lines; renumbering gives XCD x the contiguous range
// [x G/8, (x+1) G/8) of the grid, so the tiles sharing a line run at the same
// time on the same L2 instead of each XCD fetching whole lines for a part.
__device__ __forceinline__ unsigned xcd_tile_order(unsigned b, unsigned G) {
    return (G & 7) ? b : (b & 7) * (G >> 3) + (b >> 3);
}


__device__ __forceinline__ u64 freq_rt(int i, int L, int lr) {
    const int r0 = L % lr;
    int left = L, sh = 0;
    u64 c = 0;
    if (r0) { left -= r0; c = (u64)((i >> left) & ((1 << r0) - 1)); sh = r0; }
    while (left > 0) { left -= lr; c |= (u64)((i >> left) & ((1 << lr) - 1)) << sh; sh += lr; }
    return c;
}

// ---- carry propagation --------------------------------------------------------------
// A chunk of k limbs (2k coefficients) with carry-in 0 yields 32k bits D and a
// carry-out C (< 2^41); with carry-in c the carry-out is C + [c >= 2^(32k) - D].
// For k >= 2, 2^(32k) - D fits in 64 bits only when limbs 2..k-1 of D are all
// ones (then it is 2^64 - D mod 2^64, or never if that is 2^64); for k = 1 it
// is 2^32 - D. So each chunk is the function f(c) = c >= thr ? C+1 : C, and
// compositions keep that form: the carry is a parallel scan of these triples.
struct CarryFn { u64 thr, v0, v1; };
__device__ __forceinline__ u64 cf_apply(const CarryFn& f, u64 c) { return c >= f.thr ? f.v1 : f.v0; }
// (later o earlier)
__device__ __forceinline__ CarryFn cf_compose(const CarryFn& later, const CarryFn& earlier) {
    return CarryFn{earlier.thr, cf_apply(later, earlier.v0), cf_apply(later, earlier.v1)};
}
// Padding entries (after the last real chunk) are the constant-0 function.
__device__ __forceinline__ CarryFn cf_pad() { return CarryFn{~(u64)0, 0, 0}; }

constexpr int CARRY_NT = 256;
// Multi-workgroup carry: each thread takes CPT = max(1, 32 / KL) consecutive
// chunks (KL limbs each, 32 limbs for every KL <= 32), a workgroup
// CARRY_NT * CPT chunks. (Against 16 limbs per thread, twice as many
// workgroups: 2^23 limbs took 9.0 + 21.3 us for the reduce and the apply
// against 7.9 + 19.5 us; 64 limbs per thread did not gain further.)
template <int T> struct CarryGeom {
    static constexpr int KL = T / 2;
    static constexpr int CPT = KL >= 32 ? 1 : 32 / KL;
    static constexpr int LIMBS = CARRY_NT * CPT * KL;          // limbs per workgroup
};
// LDS index of a workgroup's limb w: one pad word per 32 keeps the per-thread
// (32-limb stride) accesses on distinct banks.
__device__ __forceinline__ int cpad(int w) { return w + (w >> 5); }

__device__ __forceinline__ CarryFn shfl_up_fn(const CarryFn& f, int d) {
    return CarryFn{__shfl_up(f.thr, d), __shfl_up(f.v0, d), __shfl_up(f.v1, d)};
}

// Inclusive scan of carry functions over the workgroup (thread 0 earliest);
// wsum holds one entry per wave (at most 16).
__device__ __forceinline__ CarryFn block_scan(CarryFn f, CarryFn* wsum) {
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

__device__ __forceinline__ u64 wave_max(u64 v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) { u64 o = __shfl_xor(v, d); v = o > v ? o : v; }
    return v;
}

// ---- multi-workgroup carry ---------------------------------------------------------
// The inverse column pass holds, per workgroup, T adjacent coefficients of
// every row: a chunk of T coefficients = KL = T/2 limbs. It runs each chunk's
// carry with carry-in 0 and writes the chunk's digits D (KL limbs, straight
// into the product buffer) and its carry-out C into the carry scratch; the
// coefficients themselves never return to HBM. D and C determine the chunk's
// carry function (thr from D's limbs, v0 = C, v1 = C + 1), so two scans over
// (D, C) finish the carry: k_carry_reduce composes each workgroup's chunks
// into one aggregate per workgroup (k_carry_reduce_th reads the thresholds TH
// instead of D when the producer wrote them), and k_carry_apply takes its
// workgroup's carry-in from the aggregates below it (the nearest constant
// one ends the look-back) and adds each chunk's carry-in to D in place.
// Limbs at or above na + nb are zero in D (the product has na + nb limbs),
// so they are neither stored nor read.
//
// Scratch per job: C[ceil(nl/2)] (enough for T >= 4), then TH[ceil(nl/2)]
// (the chunk thresholds, when the producer writes them), then agg[nblk_max]
// CarryFn.
struct CarryJob {
    u32* out; size_t nlimbs, nchunks, nblk;
    u64* C; u64* TH; CarryFn* agg;
};
template <int T>
__device__ __forceinline__ CarryJob carry_job(const MulJob& J, size_t nblk_max) {
    using G = CarryGeom<T>;
    CarryJob c;
    c.out = J.out;
    c.nlimbs = J.na + J.nb;
    c.nchunks = (2 * c.nlimbs + T - 1) / T;
    c.nblk = (c.nchunks + CARRY_NT * G::CPT - 1) / (CARRY_NT * G::CPT);
    const size_t cap = ((c.nlimbs + 1) / 2 + 3) / 4 * 4;
    c.C = (u64*)J.carry;
    c.TH = c.C + cap;
    c.agg = (CarryFn*)(c.TH + cap);
    return c;
}

// Coalesced copy of the workgroup's limb range [L0, L0 + LIMBS) into LDS
// (zeros at and above nlimbs), and back. A whole, aligned range moves as
// 16-byte vectors, all of a thread's loads issued before the first LDS write.
template <int T>
__device__ __forceinline__ void stage_in(const CarryJob& c, size_t L0, u32* sD) {
    constexpr int LB = CarryGeom<T>::LIMBS;
    constexpr int NV = LB / 4 / CARRY_NT;                       // vectors per thread
    static_assert(LB % (4 * CARRY_NT) == 0, "whole vectors per thread");
    const u32* src = c.out + L0;
    if (L0 + LB <= c.nlimbs && ((uintptr_t)src & 15) == 0) {
        uint4 q[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) q[k] = gld4((const uint4*)src + k * CARRY_NT + threadIdx.x);
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const int v = k * CARRY_NT + threadIdx.x;
            sD[cpad(4 * v)] = q[k].x; sD[cpad(4 * v + 1)] = q[k].y; sD[cpad(4 * v + 2)] = q[k].z; sD[cpad(4 * v + 3)] = q[k].w;
        }
    } else {
        for (int w = threadIdx.x; w < LB; w += CARRY_NT) sD[cpad(w)] = (L0 + w < c.nlimbs) ? src[w] : 0u;
    }
}
template <int T>
__device__ __forceinline__ void stage_out(const CarryJob& c, size_t L0, const u32* sD) {
    constexpr int LB = CarryGeom<T>::LIMBS;
    constexpr int NV = LB / 4 / CARRY_NT;
    u32* dst = c.out + L0;
    if (L0 + LB <= c.nlimbs && ((uintptr_t)dst & 15) == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const int v = k * CARRY_NT + threadIdx.x;
            gst4((uint4*)dst + v, uint4{sD[cpad(4 * v)], sD[cpad(4 * v + 1)], sD[cpad(4 * v + 2)], sD[cpad(4 * v + 3)]});
        }
    } else {
        for (int w = threadIdx.x; w < LB; w += CARRY_NT)
            if (L0 + w < c.nlimbs) dst[w] = sD[cpad(w)];
    }
}

// f(c) = C + [D + c >= 2^(32 KL)]: the threshold 2^(32 KL) - D fits in 64
// bits only when limbs 2..KL-1 of D are all ones (KL >= 2).
template <int KL>
__device__ __forceinline__ CarryFn chunk_fn(const u32* D, u64 C) {
    bool ones = true;
#pragma unroll
    for (int q = 2; q < KL; ++q) ones &= D[q] == 0xFFFFFFFFu;
    const u64 lo = (u64)D[0] | ((u64)D[1] << 32);
    return CarryFn{(ones && lo != 0) ? (u64)(0 - lo) : ~(u64)0, C, C + 1};
}


} // namespace pdh
