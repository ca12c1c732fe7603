// ============================================================================
//  bs_kernels.hip -- the bottom of the Chudnovsky binary-splitting tree on the
//  device: the leaves and the first G levels, one workgroup per group of 2^G
//  consecutive terms.
//
//  The reference builds P, Q, T recursively on the host (src/chudnovsky.cpp:
//  64-102, leaves :70-83). P(a,b), Q(a,b), T(a,b) do not depend on where a
//  range is split, so a group's values equal the reference's for the same
//  range. Leaf k >= 1: p = (6k-5)(2k-1)(6k-1), q = k^3 C^3/24,
//  t = (-1)^k (A + Bk) p; leaf 0: p = q = 1, t = A. A merge is
//      P = P_L P_R,  Q = Q_L Q_R,  T = T_L Q_R + P_L T_R   (T signed).
//  With k < 2^28, p < 2^91 (3 limbs), q < 2^138 and |t| < 2^148 (5 limbs);
//  after l merges P, Q, T fit 3 2^l, 5 2^l and 6 2^l limbs.
//
//  Products are small (at most 320 x 320 limbs): each is computed by a team
//  of threads as 32 x 32 -> 64-bit partial products added into 64-bit LDS
//  column sums (low and high halves in adjacent columns, so no column can
//  overflow), then one thread normalizes the columns into limbs.
// ============================================================================
#include "pdh_internal.h"

namespace pdh {

namespace {

constexpr uint64_t A_CONST = 13591409ull;
constexpr uint64_t B_CONST = 545140134ull;
constexpr uint64_t C3_OVER_24 = 10939058860032000ull;

typedef unsigned __int128 u128;

constexpr int BS_NT = 256;
constexpr int BS_GMAX = 6;                      // 64 terms per group
constexpr int LVL_WORDS = 14 << BS_GMAX;        // P, Q, T slots of every node of one level

__device__ __forceinline__ int capP(int l) { return 3 << l; }
__device__ __forceinline__ int capQ(int l) { return 5 << l; }
__device__ __forceinline__ int capT(int l) { return 6 << l; }
__device__ __forceinline__ int node_words(int l) { return 14 << l; }
// column slots of one product at level l (children of level l-1): the widest
// product is T_L Q_R, (6 + 5) 2^(l-1) limbs, plus one column for the carry
__device__ __forceinline__ int col_slot(int l) { return (11 << (l - 1)) + 2; }
// all products of one level: 4 per pair, 2^(G-l) pairs: 4 2^(G-l) (11 2^(l-1) + 2)
// = 22 2^G + 8 2^(G-l) <= 26 2^G words
constexpr int COL_WORDS = 26 << BS_GMAX;

__device__ int u128_limbs(u128 v, uint32_t* out) {
    int n = 0;
    while (v) { out[n++] = (uint32_t)v; v >>= 32; }
    return n;
}

// |a| + |b| or |a| - |b| of little-endian limb strings; returns the length.
__device__ int mag_add(const uint32_t* a, int na, const uint32_t* b, int nb, uint32_t* r) {
    const int n = na > nb ? na : nb;
    uint64_t c = 0;
    for (int i = 0; i < n; ++i) {
        c += (uint64_t)(i < na ? a[i] : 0u) + (i < nb ? b[i] : 0u);
        r[i] = (uint32_t)c;
        c >>= 32;
    }
    int len = n;
    if (c) r[len++] = (uint32_t)c;
    return len;
}
__device__ int mag_cmp(const uint32_t* a, int na, const uint32_t* b, int nb) {
    if (na != nb) return na < nb ? -1 : 1;
    for (int i = na - 1; i >= 0; --i)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}
__device__ int mag_sub(const uint32_t* a, int na, const uint32_t* b, int nb, uint32_t* r) {   // |a| >= |b|
    int64_t brw = 0;
    for (int i = 0; i < na; ++i) {
        int64_t d = (int64_t)a[i] - (i < nb ? (int64_t)b[i] : 0) - brw;
        brw = d < 0;
        r[i] = (uint32_t)(d + (brw << 32));
    }
    int len = na;
    while (len && r[len - 1] == 0) --len;
    return len;
}

} // namespace

// Workgroup g: terms [k0 + g 2^G, min(n_terms, k0 + (g+1) 2^G)). Output slots
// of the group root: P at g 3 2^G, Q at g 5 2^G, T at g 6 2^G (limbs),
// lengths len[3g..3g+2], sign of T sgn[g].
__global__ __launch_bounds__(BS_NT) void k_bs_groups(uint64_t n_terms, int G, uint32_t* __restrict__ gP,
                                                     uint32_t* __restrict__ gQ, uint32_t* __restrict__ gT,
                                                     uint64_t* __restrict__ glen, int32_t* __restrict__ gsgn) {
    __shared__ uint32_t lv[2][LVL_WORDS];
    __shared__ int nlen[2][3 << BS_GMAX];          // per node: lengths of P, Q, T
    __shared__ int nsgn[2][1 << BS_GMAX];
    __shared__ unsigned long long col[COL_WORDS];
    __shared__ uint32_t prod[COL_WORDS];
    __shared__ int plen[4 << BS_GMAX];

    const uint64_t g = blockIdx.x;
    const uint64_t first = g << G;
    const int cnt0 = (int)(n_terms - first < ((uint64_t)1 << G) ? n_terms - first : ((uint64_t)1 << G));

    // leaves
    for (int t = threadIdx.x; t < cnt0; t += blockDim.x) {
        const uint64_t k = first + t;
        uint32_t* P = &lv[0][t * node_words(0)];
        uint32_t* Q = P + capP(0);
        uint32_t* T = Q + capQ(0);
        if (k == 0) {
            P[0] = 1; Q[0] = 1; T[0] = (uint32_t)A_CONST;
            nlen[0][3 * t] = 1; nlen[0][3 * t + 1] = 1; nlen[0][3 * t + 2] = 1; nsgn[0][t] = 1;
        } else {
            const u128 p = (u128)(6 * k - 5) * (2 * k - 1) * (6 * k - 1);
            const u128 q = (u128)k * k * k * C3_OVER_24;
            const uint64_t lin = A_CONST + B_CONST * k;
            // t = p lin: p < 2^91, lin < 2^57, as p_lo lin + (p_hi lin) 2^64
            const u128 lo = (u128)(uint64_t)p * lin;
            const u128 hi = (u128)(uint64_t)(p >> 64) * lin + (uint64_t)(lo >> 64);
            uint32_t tl[6];
            tl[0] = (uint32_t)lo; tl[1] = (uint32_t)((uint64_t)lo >> 32);
            tl[2] = (uint32_t)hi; tl[3] = (uint32_t)(hi >> 32); tl[4] = (uint32_t)(hi >> 64); tl[5] = (uint32_t)(hi >> 96);
            int nt = 6;
            while (nt && tl[nt - 1] == 0) --nt;
            for (int i = 0; i < nt; ++i) T[i] = tl[i];
            nlen[0][3 * t] = u128_limbs(p, P);
            nlen[0][3 * t + 1] = u128_limbs(q, Q);
            nlen[0][3 * t + 2] = nt;
            nsgn[0][t] = (k & 1) ? -1 : 1;
        }
    }
    __syncthreads();

    int cnt = cnt0;
    int cur = 0;
    for (int l = 1; l <= G; ++l) {
        const int nxt = cur ^ 1;
        const int nodes = (cnt + 1) / 2;
        const int pairs = cnt / 2;                  // nodes with a right child
        const int cs = col_slot(l);
        const int cw = node_words(l - 1);
        // clear the column sums of 4 products per pair
        for (int w = threadIdx.x; w < 4 * pairs * cs; w += blockDim.x) col[w] = 0;
        __syncthreads();
        // partial products: work item (pair, product, i) adds a_i * b_j for all j
        const int items_per_pair = 4 * capT(l - 1);
        for (int it = threadIdx.x; it < pairs * items_per_pair; it += blockDim.x) {
            const int pr = it / items_per_pair;
            const int rem = it - pr * items_per_pair;
            const int which = rem / capT(l - 1);
            const int i = rem - which * capT(l - 1);
            const uint32_t* L = &lv[cur][(2 * pr) * cw];
            const uint32_t* R = &lv[cur][(2 * pr + 1) * cw];
            const int* lL = &nlen[cur][3 * (2 * pr)];
            const int* lR = &nlen[cur][3 * (2 * pr + 1)];
            // which: 0 Q_L Q_R, 1 T_L Q_R, 2 P_L T_R, 3 P_L P_R
            const uint32_t* a; int na; const uint32_t* b; int nb;
            if (which == 0)      { a = L + capP(l - 1); na = lL[1]; b = R + capP(l - 1); nb = lR[1]; }
            else if (which == 1) { a = L + capP(l - 1) + capQ(l - 1); na = lL[2]; b = R + capP(l - 1); nb = lR[1]; }
            else if (which == 2) { a = L; na = lL[0]; b = R + capP(l - 1) + capQ(l - 1); nb = lR[2]; }
            else                 { a = L; na = lL[0]; b = R; nb = lR[0]; }
            if (i >= na) continue;
            unsigned long long* c = &col[(pr * 4 + which) * cs];
            const uint64_t ai = a[i];
            for (int j = 0; j < nb; ++j) {
                const uint64_t pp = ai * b[j];
                atomicAdd(&c[i + j], (unsigned long long)(uint32_t)pp);
                atomicAdd(&c[i + j + 1], (unsigned long long)(pp >> 32));
            }
        }
        __syncthreads();
        // normalize each product's columns into limbs
        for (int pp = threadIdx.x; pp < 4 * pairs; pp += blockDim.x) {
            const int pr = pp / 4, which = pp % 4;
            const int* lL = &nlen[cur][3 * (2 * pr)];
            const int* lR = &nlen[cur][3 * (2 * pr + 1)];
            const int na = which == 0 ? lL[1] : which == 1 ? lL[2] : lL[0];
            const int nb = which == 0 ? lR[1] : which == 1 ? lR[1] : which == 2 ? lR[2] : lR[0];
            const unsigned long long* c = &col[pp * cs];
            uint32_t* out = &prod[pp * cs];
            int len = 0;
            if (na && nb) {
                uint64_t carry = 0;
                for (int w = 0; w < na + nb + 1; ++w) {
                    carry += c[w];
                    out[w] = (uint32_t)carry;
                    carry >>= 32;
                }
                len = na + nb + 1;
                while (len && out[len - 1] == 0) --len;
            }
            plen[pp] = len;
        }
        __syncthreads();
        // combine: P = P_L P_R, Q = Q_L Q_R, T = T_L Q_R + P_L T_R (signed); an
        // odd last node moves up unchanged
        for (int nd = threadIdx.x; nd < nodes; nd += blockDim.x) {
            uint32_t* P = &lv[nxt][nd * node_words(l)];
            uint32_t* Q = P + capP(l);
            uint32_t* T = Q + capQ(l);
            int* ln = &nlen[nxt][3 * nd];
            if (nd >= pairs) {                          // no right child
                const uint32_t* S = &lv[cur][(2 * nd) * cw];
                const int* ls = &nlen[cur][3 * (2 * nd)];
                for (int i = 0; i < ls[0]; ++i) P[i] = S[i];
                for (int i = 0; i < ls[1]; ++i) Q[i] = S[capP(l - 1) + i];
                for (int i = 0; i < ls[2]; ++i) T[i] = S[capP(l - 1) + capQ(l - 1) + i];
                ln[0] = ls[0]; ln[1] = ls[1]; ln[2] = ls[2];
                nsgn[nxt][nd] = nsgn[cur][2 * nd];
                continue;
            }
            const uint32_t* qq = &prod[(nd * 4 + 0) * cs];
            const uint32_t* tq = &prod[(nd * 4 + 1) * cs];
            const uint32_t* pt = &prod[(nd * 4 + 2) * cs];
            const uint32_t* ppp = &prod[(nd * 4 + 3) * cs];
            const int lqq = plen[nd * 4], ltq = plen[nd * 4 + 1], lpt = plen[nd * 4 + 2], lpp = plen[nd * 4 + 3];
            for (int i = 0; i < lpp; ++i) P[i] = ppp[i];
            for (int i = 0; i < lqq; ++i) Q[i] = qq[i];
            ln[0] = lpp; ln[1] = lqq;
            const int s1 = ltq ? nsgn[cur][2 * nd] : 0;
            const int s2 = lpt ? nsgn[cur][2 * nd + 1] : 0;
            int lt, st;
            if (s1 == 0) { for (int i = 0; i < lpt; ++i) T[i] = pt[i]; lt = lpt; st = s2; }
            else if (s2 == 0) { for (int i = 0; i < ltq; ++i) T[i] = tq[i]; lt = ltq; st = s1; }
            else if (s1 == s2) { lt = mag_add(tq, ltq, pt, lpt, T); st = s1; }
            else {
                const int c = mag_cmp(tq, ltq, pt, lpt);
                if (c == 0) { lt = 0; st = 0; }
                else if (c > 0) { lt = mag_sub(tq, ltq, pt, lpt, T); st = s1; }
                else { lt = mag_sub(pt, lpt, tq, ltq, T); st = s2; }
            }
            ln[2] = lt;
            nsgn[nxt][nd] = st;
        }
        __syncthreads();
        cnt = nodes;
        cur = nxt;
    }
    // the group root (cnt == 1)
    const int* ln = &nlen[cur][0];
    const uint32_t* P = &lv[cur][0];
    const uint32_t* Q = P + capP(G);
    const uint32_t* T = Q + capQ(G);
    for (int i = threadIdx.x; i < ln[0]; i += blockDim.x) gP[g * (3u << G) + i] = P[i];
    for (int i = threadIdx.x; i < ln[1]; i += blockDim.x) gQ[g * (5u << G) + i] = Q[i];
    for (int i = threadIdx.x; i < ln[2]; i += blockDim.x) gT[g * (6u << G) + i] = T[i];
    if (threadIdx.x == 0) {
        glen[3 * g] = (uint64_t)ln[0]; glen[3 * g + 1] = (uint64_t)ln[1]; glen[3 * g + 2] = (uint64_t)ln[2];
        gsgn[g] = nsgn[cur][0];
    }
}

hipError_t launch_bs_groups(uint64_t n_terms, int G, uint32_t* dP, uint32_t* dQ, uint32_t* dT, uint64_t* dlen,
                            int32_t* dsgn, hipStream_t s) {
    const uint64_t groups = (n_terms + ((uint64_t)1 << G) - 1) >> G;
    if (groups == 0) return hipSuccess;
    hipEvent_t t0 = timing_begin(s);
    hipLaunchKernelGGL(k_bs_groups, dim3((unsigned)groups), dim3(BS_NT), 0, s, n_terms, G, dP, dQ, dT, dlen, dsgn);
    g_launches.fetch_add(1, std::memory_order_relaxed);
    hipError_t e = hipGetLastError();
    timing_end("k_bs_groups", t0, s);
    return e;
}

} // namespace pdh
