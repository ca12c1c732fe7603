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

// ---- signed sums T = sa A + sb B of many magnitude pairs ---------------------------
// Used for the merge of T above the groups (A = T_L Q_R and B = P_L T_R from
// the batched multiplies, with the children's T signs; T overwrites A) and for
// remainders elsewhere. T may alias A: every limb is read and written by the
// same thread. Every node is cut into tiles of SA_NT x SA_K limbs
// (one workgroup each; thread k owns SA_K consecutive limbs). A segment of
// limbs has a carry status for the sum, or a borrow status for the difference
// X - Y: GEN (carry/borrow out whatever comes in), PROP (out iff in), KILL
// (never out). For X - Y the borrow status of a segment is its comparison:
// X_seg < Y_seg GEN, equal PROP, greater KILL, so the tile statuses of A - B
// also give the order of |A| and |B| (the highest non-PROP tile decides) and
// flipping GEN and KILL gives the statuses of B - A. The carry into any
// segment is the status of the nearest lower non-PROP segment (GEN: 1,
// KILL: 0; none: 0). Pass 1 writes one status per tile; pass 2 composes the
// tile's carry-in from the lower tiles' statuses, recomputes its own segments,
// writes T and atomically raises the node's length.
namespace {

constexpr int SA_NT = 256;
constexpr int SA_K = BS_ADD_TILE / SA_NT;
static_assert(SA_NT * SA_K == BS_ADD_TILE, "tile size");
constexpr uint32_t ST_KILL = 0, ST_PROP = 1, ST_GEN = 2;

typedef uint32_t sa_v4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) sa_v4 sa_gv4;
typedef __attribute__((address_space(1))) uint32_t sa_gu32;

// A thread's SA_K limbs of one operand from base (a multiple of SA_K): as
// 16-byte vectors when they all lie below the length and the operand is
// 16-byte aligned, else limb by limb with zeros from the length on.
__device__ __forceinline__ void sa_load1(const uint32_t* a, uint64_t la, uint64_t base, uint32_t* x) {
    if (base + SA_K <= la && ((uintptr_t)a & 15) == 0) {
        const sa_gv4* p = (const sa_gv4*)(a + base);
#pragma unroll
        for (int q = 0; q < SA_K / 4; ++q) {
            const sa_v4 v = p[q];
            x[4 * q] = v.x; x[4 * q + 1] = v.y; x[4 * q + 2] = v.z; x[4 * q + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < SA_K; ++k) {
            const uint64_t i = base + k;
            x[k] = i < la ? ((const sa_gu32*)a)[i] : 0u;
        }
    }
}

__device__ __forceinline__ void sa_load(const uint32_t* a, uint64_t la, const uint32_t* b, uint64_t lb,
                                        uint64_t base, uint32_t* x, uint32_t* y) {
    sa_load1(a, la, base, x);
    sa_load1(b, lb, base, y);
}

__device__ __forceinline__ uint32_t seg_status(const uint32_t* x, const uint32_t* y, bool sub) {
    if (!sub) {
        uint64_t c = 0;
        uint32_t all = 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < SA_K; ++k) {
            c += (uint64_t)x[k] + y[k];
            all &= (uint32_t)c;
            c >>= 32;
        }
        return c ? ST_GEN : (all == 0xFFFFFFFFu ? ST_PROP : ST_KILL);
    }
    uint32_t st = ST_PROP;
#pragma unroll
    for (int k = SA_K - 1; k >= 0; --k)
        if (st == ST_PROP && x[k] != y[k]) st = x[k] < y[k] ? ST_GEN : ST_KILL;
    return st;
}

// The status of the nearest lower thread of the block whose status is not
// PROP (PROP if none), and in *agg the block's own status (that of its
// highest non-PROP thread). wv holds SA_NT / 64 words of LDS.
__device__ __forceinline__ uint32_t block_status_in(uint32_t st, uint32_t* wv, uint32_t* agg) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t m = __ballot(st != ST_PROP);
    const uint64_t lower = m & ((1ull << lane) - 1ull);
    const uint32_t from_lower = __shfl(st, lower ? 63 - __builtin_clzll(lower) : 0);
    const uint32_t wave_agg = __shfl(st, m ? 63 - __builtin_clzll(m) : 0);
    if (lane == 0) wv[w] = m ? wave_agg : ST_PROP;
    __syncthreads();
    uint32_t in = lower ? from_lower : ST_PROP;
    for (int v = w - 1; v >= 0 && in == ST_PROP; --v) in = wv[v];
    uint32_t a = ST_PROP;
    for (int v = SA_NT / 64 - 1; v >= 0 && a == ST_PROP; --v) a = wv[v];
    *agg = a;
    return in;
}

__device__ __forceinline__ uint32_t block_max_u32(uint32_t v, uint32_t* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t u = __shfl_xor(v, o);
        v = u > v ? u : v;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t r = 0;
    for (int i = 0; i < SA_NT / 64; ++i) r = red[i] > r ? red[i] : r;
    return r;
}

// The key (tile + 1) * 4 + status of the highest non-PROP tile among the
// node's tiles [0, hi) (0 if every one is PROP), looking down from hi - 1 a
// workgroup's width at a time and stopping at the first chunk that holds a
// non-PROP tile. A PROP tile has every limb of its sum 0xFFFFFFFF (or, for a
// difference, every limb equal), so the first chunk ends nearly every search
// and the cost no longer grows with the node's tile count.
__device__ __forceinline__ uint32_t top_non_prop(const uint8_t* __restrict__ status, uint32_t tile0, uint32_t hi,
                                                 uint32_t* red) {
    for (int64_t top = (int64_t)hi - 1; top >= 0; top -= SA_NT) {
        const int64_t i = top - (int64_t)threadIdx.x;
        uint32_t key = 0;
        if (i >= 0) {
            const uint32_t st = status[tile0 + i];
            if (st != ST_PROP) key = (uint32_t)(i + 1) * 4 + st;
        }
        key = block_max_u32(key, red);
        __syncthreads();                        // red is written again by the next chunk
        if (key) return key;
    }
    return 0;
}

struct SaView {
    uint32_t nd, j;
    uint64_t la, lb;
    bool sub;
};

__device__ __forceinline__ SaView sa_view(const BsAddNode& N, uint32_t nd, uint32_t t) {
    SaView v;
    v.nd = nd;
    v.j = t - N.tile0;
    v.la = N.sa ? *N.la : 0;
    v.lb = N.sb ? *N.lb : 0;
    v.sub = N.sa != 0 && N.sb != 0 && N.sa != N.sb;
    return v;
}

// nodes == nullptr: a single node, passed by value in `one`.
__global__ __launch_bounds__(SA_NT) void k_bs_add_status(const BsAddNode* __restrict__ nodes,
                                                         const uint32_t* __restrict__ tile_node, BsAddNode one,
                                                         uint8_t* __restrict__ status) {
    __shared__ uint32_t wv[SA_NT / 64];
    const uint32_t t = blockIdx.x;
    const uint32_t nd = nodes ? tile_node[t] : 0u;
    const BsAddNode N = nodes ? nodes[nd] : one;
    const SaView v = sa_view(N, nd, t);
    const uint64_t base = (uint64_t)v.j * BS_ADD_TILE + (uint64_t)threadIdx.x * SA_K;
    uint32_t x[SA_K], y[SA_K];
    sa_load(N.A, v.la, N.B, v.lb, base, x, y);
    uint32_t agg;
    (void)block_status_in(seg_status(x, y, v.sub), wv, &agg);
    if (threadIdx.x == 0) {
        status[t] = (uint8_t)agg;
        if (v.j == 0) *N.len = 0;
    }
}

__global__ __launch_bounds__(SA_NT) void k_bs_add_apply(const BsAddNode* __restrict__ nodes,
                                                        const uint32_t* __restrict__ tile_node, BsAddNode one,
                                                        const uint8_t* __restrict__ status) {
    __shared__ uint32_t wv[SA_NT / 64];
    __shared__ uint32_t red[2][SA_NT / 64];
    // highest tiles first (the last node's top tile first): the tile that
    // holds a node's top nonzero limb usually runs before the lower ones,
    // which then find the length already above theirs and skip the atomic
    const uint32_t t = gridDim.x - 1 - blockIdx.x;
    const uint32_t nd = nodes ? tile_node[t] : 0u;
    const BsAddNode N = nodes ? nodes[nd] : one;
    const SaView v = sa_view(N, nd, t);
    // nearest lower non-PROP tile (carry into this tile) and, for a
    // difference, the highest non-PROP tile (the order of |A| and |B|);
    // keys are (tile + 1) * 4 + status, 0 for none
    const uint32_t kin = top_non_prop(status, N.tile0, v.j, red[0]);
    const uint32_t kall = v.sub ? top_non_prop(status, N.tile0, N.ntiles, red[1]) : 0;
    const bool swap = v.sub && (kall & 3) == ST_GEN;          // |A| < |B|: T = sb (B - A)
    uint32_t tin = kin ? (kin & 3) : ST_PROP;
    if (swap && tin != ST_PROP) tin = ST_GEN + ST_KILL - tin;
    const uint64_t base = (uint64_t)v.j * BS_ADD_TILE + (uint64_t)threadIdx.x * SA_K;
    uint32_t x[SA_K], y[SA_K];
    if (swap) sa_load(N.B, v.lb, N.A, v.la, base, x, y);
    else sa_load(N.A, v.la, N.B, v.lb, base, x, y);
    uint32_t agg;
    const uint32_t in = block_status_in(seg_status(x, y, v.sub), wv, &agg);
    const uint32_t st_in = in != ST_PROP ? in : tin;
    uint64_t c = st_in == ST_GEN ? 1u : 0u;
    uint32_t r[SA_K];
#pragma unroll
    for (int k = 0; k < SA_K; ++k) {
        if (!v.sub) {
            c += (uint64_t)x[k] + y[k];
            r[k] = (uint32_t)c;
            c >>= 32;
        } else {
            const uint64_t d = (uint64_t)x[k] - y[k] - c;
            r[k] = (uint32_t)d;
            c = (d >> 32) & 1u;
        }
    }
    // T as 16-byte vectors where the segment lies below the capacity and T
    // is 16-byte aligned, else limb by limb up to the capacity
    uint32_t top = 0;
    if (base + SA_K <= N.cap && ((uintptr_t)N.T & 15) == 0) {
        sa_gv4* p = (sa_gv4*)(N.T + base);
#pragma unroll
        for (int q = 0; q < SA_K / 4; ++q) p[q] = sa_v4{r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]};
#pragma unroll
        for (int k = 0; k < SA_K; ++k)
            if (r[k]) top = (uint32_t)(k + 1);
    } else {
#pragma unroll
        for (int k = 0; k < SA_K; ++k) {
            if (base + k < N.cap) {
                ((sa_gu32*)N.T)[base + k] = r[k];
                if (r[k]) top = (uint32_t)(k + 1);
            }
        }
    }
    uint64_t hi = top ? base + top : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t u = __shfl_xor(hi, o);
        hi = u > hi ? u : hi;
    }
    if ((threadIdx.x & 63) == 0 && hi > __hip_atomic_load(N.len, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax((unsigned long long*)N.len, (unsigned long long)hi);
    if (v.j == 0 && threadIdx.x == 0) *N.sign = !v.sub ? (N.sa ? N.sa : N.sb) : (swap ? N.sb : N.sa);
}

} // namespace

hipError_t launch_bs_signed_add(const BsAddNode* d_nodes, const uint32_t* d_tile_node, uint32_t ntiles,
                                uint8_t* d_status, hipStream_t s, const BsAddNode* one) {
    if (ntiles == 0) return hipSuccess;
    const BsAddNode single = one ? *one : BsAddNode{};
    if (one) { d_nodes = nullptr; d_tile_node = nullptr; }
    hipEvent_t t0 = timing_begin(s);
    hipLaunchKernelGGL(k_bs_add_status, dim3(ntiles), dim3(SA_NT), 0, s, d_nodes, d_tile_node, single, d_status);
    g_launches.fetch_add(1, std::memory_order_relaxed);
    hipError_t e = hipGetLastError();
    timing_end("k_bs_add_status", t0, s);
    if (e != hipSuccess) return e;
    t0 = timing_begin(s);
    hipLaunchKernelGGL(k_bs_add_apply, dim3(ntiles), dim3(SA_NT), 0, s, d_nodes, d_tile_node, single, d_status);
    g_launches.fetch_add(1, std::memory_order_relaxed);
    e = hipGetLastError();
    timing_end("k_bs_add_apply", t0, s);
    return e;
}

} // namespace pdh
