// ============================================================================
//  chudnovsky_hip.cpp -- Chudnovsky pi with every multiply on the ntt-hip seam.
//
//  Same digits as the reference's compute_pi_chudnovsky (src/chudnovsky.cpp:
//  107-164): guard 16, prec = digits + 16, N = floor(prec / 14.1816...) + 2
//  terms (:54-59, :118-120), P, Q, T over [0, N) (:64-102), then
//      pi_scaled = floor( 426880 * Q * isqrt(10005 * 10^(2 prec)) / T )
//  (:130-140) and the first `digits` fractional digits of its decimal string
//  (:142-152). Every value on that path is an exact integer with a unique
//  definition, so a different algorithm yields the same digits. What
//  differs from the reference (the next rows of its own roadmap):
//
//  * Binary splitting runs bottom-up, level by level, in device memory: the
//    leaves and the first six levels (groups of 64 terms) in one kernel
//    (pdh_bs_groups), then one pdh_bs_merge_dev call per level (the level's
//    P, Q, T products as one batch and the signed T sum on the device); only
//    lengths and signs reach the host until the root. P(0,N), Q(0,N), T(0,N)
//    do not depend on where the range is split.
//  * Division is a Newton reciprocal with an exact floor correction (the
//    reference forwards div_fast to Knuth, src/bignum_div.cpp:170-178); isqrt
//    is a division-free Newton inverse square root and an exact correction
//    (reference: Newton over Knuth divisions, :189-204).
//  * to_decimal is divide-and-conquer on powers of ten (reference: O(d^2)
//    repeated /10^9, :239-267): on the device, every level's quotients and
//    remainders stay in HBM (batched products, shifts and signed sums) and
//    one kernel writes the leaves' digits.
//
//  Multiplies go through a Multiplier (device by default). A caller may pass
//  its own through pdh_pi_chudnovsky_cb -- the CPU tests do, with the oracle's
//  ntt-cpu -- which lets the algorithm be checked on a machine with no GPU.
// ============================================================================
#include "pidigits_hip.h"
#include "pdh_capi.h"
#include "pdh_pi.h"
#include "pdh_diag.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <future>
#include <memory>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace pidigits {
namespace chud {

typedef std::vector<uint32_t> Mag;
typedef unsigned __int128 u128;

struct SInt { Mag m; int s = 0; };            // sign s in {-1, 0, 1}, magnitude m
// dev: when set, Q and T (sign T.s) are in device memory (a DevRoot of the
// device backend) and the host magnitudes are empty
struct PQT { Mag P, Q; SInt T; std::shared_ptr<void> dev; };

// Test switch (pdh_newton_set_perturb): k != 0 moves every quotient and
// square-root estimate by k units, and every decimal-level quotient estimate
// by k (k > 0) or by -1 where it is at least 1 (k < 0), before the exact
// correction, so the correction branches run on every call.
static std::atomic<int> g_perturb{0};
// Correction steps taken by the exact floor corrections (quotients, square
// roots, decimal levels) in this process (pdh_newton_corrections).
static std::atomic<uint64_t> g_corrections{0};
// The estimates are within a few units (the test switch adds at most 8), so
// a correction that needs more steps than this means a wrong estimate: it
// is reported instead of being walked off one unit at a time.
static const int kMaxCorrections = 64;
static void count_correction(int& steps) {
    ++g_corrections;
    if (++steps > kMaxCorrections)
        throw HipError(PDH_E_HIP, "ntt-hip: a Newton estimate was off by more than 64 units");
}

struct Multiplier {
    std::function<Mag(const Mag&, const Mag&)> mul;
    std::function<Mag(const Mag&)> sqr;
    std::function<std::vector<Mag>(const std::vector<std::pair<const Mag*, const Mag*>>&)> mul_many;
    // optional: P, Q, T of consecutive ranges covering [0, N) in order (the
    // lower levels of the tree computed elsewhere; one node = the root, whose
    // P may be empty); empty = from the leaves
    std::function<std::vector<PQT>(uint64_t N)> groups;
    // optional: the decimal string of x > 0 (the levels run elsewhere); empty =
    // the host loop below. `ahead` is plan_decimal's handle or null.
    std::function<std::string(const std::vector<uint32_t>& x, const Multiplier& M,
                              const std::shared_ptr<void>& ahead)> to_decimal;
    // optional: start the divisors and reciprocals of to_decimal for a value
    // of xbits bits beside the caller's work (they depend only on the size);
    // the handle goes to to_decimal
    std::function<std::shared_ptr<void>(size_t xbits)> plan_decimal;
    // optional: floor(426880 Q sqrt(10005 10^(2 prec)) / T) computed elsewhere
    // (the square root and the division kept in device memory); empty = the
    // host backend below
    std::function<Mag(const PQT& r, uint64_t prec, std::chrono::steady_clock::time_point* ts)> pi_scaled;
    // optional: that value as "3." and `digits` decimals ("3" for none),
    // computed without it leaving the backend (ts[0..2] stamp the ends of
    // 10^prec, the square root and the division); replaces pi_scaled and
    // to_decimal
    std::function<std::string(const PQT& r, uint64_t prec, uint64_t digits, std::chrono::steady_clock::time_point* ts,
                              const std::shared_ptr<void>& ahead)> pi_decimal;
};

// ---- magnitude helpers (little-endian base 2^32, normalized) -----------------
static void trim(Mag& a) { while (!a.empty() && a.back() == 0) a.pop_back(); }

static size_t bitlen(const Mag& a) {
    if (a.empty()) return 0;
    return 32 * (a.size() - 1) + (32 - __builtin_clz(a.back()));
}

static int cmp(const Mag& a, const Mag& b) {
    if (a.size() != b.size()) return a.size() < b.size() ? -1 : 1;
    for (size_t i = a.size(); i-- > 0;)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

static Mag add(const Mag& a, const Mag& b) {
    const Mag& x = a.size() >= b.size() ? a : b;
    const Mag& y = a.size() >= b.size() ? b : a;
    Mag r(x.size() + 1);
    uint64_t c = 0;
    for (size_t i = 0; i < x.size(); ++i) {
        c += (uint64_t)x[i] + (i < y.size() ? y[i] : 0);
        r[i] = (uint32_t)c;
        c >>= 32;
    }
    r[x.size()] = (uint32_t)c;
    trim(r);
    return r;
}

// a - b, requires a >= b
static Mag sub(const Mag& a, const Mag& b) {
    Mag r(a.size());
    int64_t br = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        int64_t d = (int64_t)a[i] - (i < b.size() ? b[i] : 0) - br;
        br = d < 0;
        r[i] = (uint32_t)(d + (br << 32));
    }
    trim(r);
    return r;
}

static Mag shl(const Mag& a, size_t bits) {
    if (a.empty()) return a;
    const size_t w = bits / 32, s = bits % 32;
    Mag r(a.size() + w + 1, 0);
    for (size_t i = 0; i < a.size(); ++i) {
        r[i + w] |= a[i] << s;
        if (s) r[i + w + 1] |= a[i] >> (32 - s);
    }
    trim(r);
    return r;
}

static Mag shr(const Mag& a, size_t bits) {
    const size_t w = bits / 32, s = bits % 32;
    if (w >= a.size()) return {};
    Mag r(a.size() - w);
    for (size_t i = 0; i < r.size(); ++i) {
        uint64_t v = a[i + w];
        if (s && i + w + 1 < a.size()) v |= (uint64_t)a[i + w + 1] << 32;
        r[i] = (uint32_t)(v >> s);
    }
    trim(r);
    return r;
}

static Mag from_u128(u128 v) {
    Mag r;
    while (v) { r.push_back((uint32_t)v); v >>= 32; }
    return r;
}
static u128 to_u128(const Mag& a) {
    u128 v = 0;
    for (size_t i = a.size(); i-- > 0;) v = (v << 32) | a[i];
    return v;
}

static Mag mul_u32(const Mag& a, uint32_t m) {
    if (a.empty() || m == 0) return {};
    Mag r(a.size() + 1);
    uint64_t c = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        c += (uint64_t)a[i] * m;
        r[i] = (uint32_t)c;
        c >>= 32;
    }
    r[a.size()] = (uint32_t)c;
    trim(r);
    return r;
}

static Mag pow2(size_t e) { Mag r(e / 32 + 1, 0); r[e / 32] = 1u << (e % 32); return r; }

static void inc(Mag& a) {
    for (size_t i = 0; i < a.size(); ++i) if (++a[i] != 0) return;
    a.push_back(1);
}
static void dec(Mag& a) {            // a > 0
    for (size_t i = 0; i < a.size(); ++i) if (a[i]-- != 0) break;
    trim(a);
}

// ---- signed values (binary splitting's T) --------------------------------------
static SInt sadd(const SInt& a, const SInt& b) {
    if (a.s == 0) return b;
    if (b.s == 0) return a;
    if (a.s == b.s) return SInt{add(a.m, b.m), a.s};
    int c = cmp(a.m, b.m);
    if (c == 0) return SInt{};
    return c > 0 ? SInt{sub(a.m, b.m), a.s} : SInt{sub(b.m, a.m), b.s};
}

// ---- arithmetic backends ------------------------------------------------------------
// The Newton algorithms below are written once over a backend A with value
// type A::Num and the operations bits, shl, shr, add, sub (a >= b), cmp, pow2,
// mul, sqr, zero, inc, dec, mul_u32, from_u128 / to_u128 (small values), up /
// down (to and from host magnitudes) and tighten (make a value's size bound
// exact where a loose one would size later products too large). HostArith keeps values on the host
// and multiplies through the Multiplier (the CPU tests run the algorithms
// this way with the oracle's multiply); DevArith (further down) keeps them in
// device memory.
struct HostArith {
    typedef Mag Num;
    const Multiplier& M;
    size_t bits(const Mag& a) { return bitlen(a); }
    Mag shl(const Mag& a, size_t k) { return chud::shl(a, k); }
    Mag shr(const Mag& a, size_t k) { return chud::shr(a, k); }
    Mag add(const Mag& a, const Mag& b) { return chud::add(a, b); }
    Mag sub(const Mag& a, const Mag& b) { return chud::sub(a, b); }
    int cmp(const Mag& a, const Mag& b) { return chud::cmp(a, b); }
    Mag pow2(size_t e) { return chud::pow2(e); }
    Mag mul(const Mag& a, const Mag& b) { return M.mul(a, b); }
    Mag sqr(const Mag& a) { return M.sqr(a); }
    bool zero(const Mag& a) { return a.empty(); }
    Mag inc(Mag a) { chud::inc(a); return a; }
    Mag dec(Mag a) { chud::dec(a); return a; }
    Mag mul_u32(const Mag& a, uint32_t m) { return chud::mul_u32(a, m); }
    Mag from_u128(u128 v) { return chud::from_u128(v); }
    u128 to_u128(const Mag& a) { return chud::to_u128(a); }
    Mag up(const Mag& a) { return a; }
    Mag down(const Mag& a) { return a; }
    Mag pqt_q(const PQT& r) { return r.Q; }
    Mag pqt_t(const PQT& r) { return r.T.m; }
    void sync() {}
    void tighten(Mag&) {}
};

template <class A> struct SNum { typename A::Num m; int s = 0; };   // sign s in {-1, 0, 1}

template <class A> static SNum<A> sadd_t(A& o, const SNum<A>& a, const SNum<A>& b) {
    if (a.s == 0) return b;
    if (b.s == 0) return a;
    if (a.s == b.s) return SNum<A>{o.add(a.m, b.m), a.s};
    const int c = o.cmp(a.m, b.m);
    if (c == 0) return SNum<A>{o.from_u128(0), 0};
    return c > 0 ? SNum<A>{o.sub(a.m, b.m), a.s} : SNum<A>{o.sub(b.m, a.m), b.s};
}

// ---- Newton reciprocal: about floor(2^(2L) / d), L = bitlen(d) -----------------
// One Newton step V = V0 + V0 (2^(2L) - d V0) / 2^(2L) from V0 = V0h 2^sh
// (sh <= L - 2) squares V0's relative error. e = 2^(2L) - d V0 is formed at
// scale 2^sh, e_s = |2^(2L - sh) - d V0h|, and only e_s's bits from 2^k up,
// k = L - 2 - sh, enter the correction: the dropped part moves V by
// V0h e_lo 2^(2 sh) / 2^(2L) < 2^(L + 1 - sh + k + 2 sh - 2L) = 1/2.
template <class A>
static typename A::Num recip_step(A& o, const typename A::Num& d, size_t L, const typename A::Num& V0h, size_t sh) {
    typedef typename A::Num N;
    N P = o.mul(d, V0h);
    N two = o.pow2(2 * L - sh);
    const bool neg = o.cmp(P, two) > 0;
    N e = neg ? o.sub(P, two) : o.sub(two, P);
    const size_t k = L - 2 - sh;
    N corr = o.shr(o.mul(V0h, o.shr(e, k)), L + 2 - sh);
    N V0 = o.shl(V0h, sh);
    return neg ? o.sub(V0, corr) : o.add(V0, corr);
}

// The top h = L/2 + 8 bits of d give V_h ~ 2^(2h)/d_h; V0 = V_h 2^(L-h) has
// relative error ~2^(3-h), and one step takes it to within about two units
// (the truncations); callers correct exactly. L may be passed when known.
template <class A> static typename A::Num recip(A& o, const typename A::Num& d, size_t L = 0) {
    if (!L) L = o.bits(d);
    if (L <= 62) {
        const u128 v = ((u128)1 << (2 * L)) / o.to_u128(d);
        return o.from_u128(v);
    }
    const size_t h = L / 2 + 8;
    typename A::Num V = recip_step(o, d, L, recip(o, o.shr(d, L - h), h), L - h);
    o.tighten(V);
    return V;
}

// ---- exact floor division --------------------------------------------------------
// q ~ a / d from a reciprocal of d scaled to P = k + 32 bits (k = quotient
// bits): dp = d * 2^(P - nd) (truncated when nd > P), V ~ 2^(2P) / dp, so
// a / d ~ a V / 2^(2P - (P - nd)). The low bits of a beyond what k + P bits of
// precision need are dropped first. The estimate is within a few units; the
// remainder r = a - q d then moves q to the exact floor.
template <class A>
static void fix_quotient(A& o, const typename A::Num& a, const typename A::Num& d, const typename A::Num& qd,
                         typename A::Num& q, typename A::Num* rem) {
    typedef typename A::Num N;
    N r;
    int steps = 0;
    if (o.cmp(qd, a) > 0) {
        N deficit = o.sub(qd, a);                 // a - q d < 0
        for (;;) {
            q = o.dec(q);
            count_correction(steps);
            if (o.cmp(deficit, d) <= 0) { r = o.sub(d, deficit); break; }
            deficit = o.sub(deficit, d);
        }
    } else {
        r = o.sub(a, qd);
        while (o.cmp(r, d) >= 0) { r = o.sub(r, d); q = o.inc(q); count_correction(steps); }
    }
    if (rem) *rem = r;
}

// v moved by the test switch's k units (not below zero).
template <class A> static typename A::Num perturbed(A& o, typename A::Num v) {
    const int k = g_perturb.load();
    for (int i = 0; i < k; ++i) v = o.inc(v);
    for (int i = 0; i < -k && !o.zero(v); ++i) v = o.dec(v);
    return v;
}

template <class A>
static typename A::Num div_floor(A& o, const typename A::Num& a, const typename A::Num& d, typename A::Num* rem) {
    const size_t na = o.bits(a), nd = o.bits(d);
    if (na < nd || (na == nd && o.cmp(a, d) < 0)) { if (rem) *rem = a; return o.from_u128(0); }
    const size_t k = na - nd + 1;
    const size_t P = k + 32;
    const long sd = (long)P - (long)nd;
    typename A::Num dp = sd >= 0 ? o.shl(d, (size_t)sd) : o.shr(d, (size_t)(-sd));
    typename A::Num V = recip(o, dp, P);
    const size_t ta = na > k + P + 2 ? na - (k + P + 2) : 0;
    typename A::Num q = o.shr(o.mul(o.shr(a, ta), V), (size_t)((long)(2 * P) - sd - (long)ta));
    q = perturbed(o, q);
    fix_quotient(o, a, d, o.mul(q, d), q, rem);
    return q;
}

// ---- exact floor square root -------------------------------------------------------
// Y ~ 2^E / sqrt(s), E = bitlen(s) rounded up to even: an (E/2 + 1)-bit
// integer within a few units, by Newton on the inverse square root with no
// division. From Y_h for the top E_h ~ E/2 + 32 bits of s (s_h = s >> 2t,
// E_h = E - 2t), Y0 = Y_h 2^t has relative error ~2^(-E/4 - 13), and
//     Y = Y0 + Y0 (2^(2E) - s Y0^2) / 2^(2E + 1)
// squares it. Only what that step can see is formed: s without its low
// b = E/2 - 8 bits (s_lo Y0^2 < 2^(3E/2 - 6)), and e = 2^(2E) - s Y0^2 from
// 2^c up, c = 3E/2 - 2 (the dropped part moves Y by less than 1/2). L =
// bitlen(s) is passed down: s_h's is L - 2t.
template <class A> static typename A::Num rsqrt_approx(A& o, const typename A::Num& s, size_t L) {
    typedef typename A::Num N;
    const size_t E = L + (L & 1);
    if (E <= 100) {
        const long double v = (long double)o.to_u128(s);
        return o.from_u128((u128)(std::ldexp(1.0L, (int)E) / std::sqrt(v)));
    }
    size_t Eh = E / 2 + 32;
    Eh += Eh & 1;
    const size_t t = (E - Eh) / 2;
    const N Y0h = rsqrt_approx(o, o.shr(s, 2 * t), L - 2 * t);
    const N Y0 = o.shl(Y0h, t);
    const size_t b = E / 2 - 8, c = 3 * E / 2 - 2;
    N sy2 = o.mul(o.shr(s, b), o.sqr(Y0));        // s Y0^2 / 2^b, s truncated
    N two = o.pow2(2 * E - b);
    const bool neg = o.cmp(sy2, two) > 0;
    N e = neg ? o.sub(sy2, two) : o.sub(two, sy2);
    N corr = o.shr(o.mul(Y0, o.shr(e, c - b)), 2 * E + 1 - c);
    N Y = neg ? o.sub(Y0, corr) : o.add(Y0, corr);
    o.tighten(Y);
    return Y;
}

// x0 = s Y / 2^(E + G) from Y ~ 2^(E + G) / sqrt(s) (the inverse square root
// of s 2^(2G), G = 8 guard bits) is within about one unit; then
// r = s - x^2 moves x to the floor root with r updated by 2x -+ 1 per step
// (0 <= r <= 2x at the floor root), so only one square is formed.
template <class A> static typename A::Num isqrt(A& o, const typename A::Num& s) {
    typedef typename A::Num N;
    const size_t L = o.bits(s);
    if (L <= 100) {
        const u128 v = o.to_u128(s);
        u128 x = (u128)std::sqrt((long double)v);
        while (x * x > v) --x;
        while ((x + 1) * (x + 1) <= v) ++x;
        return o.from_u128(x);
    }
    const size_t G = 8;
    const size_t E = L + (L & 1);
    N Y = rsqrt_approx(o, o.shl(s, 2 * G), L + 2 * G);
    const size_t b = E / 2 - 2;                   // s_lo Y / 2^(E + G) < 1/2
    N x = perturbed(o, o.shr(o.mul(o.shr(s, b), Y), E + G - b));
    N x2 = o.sqr(x);
    const int c = o.cmp(s, x2);
    SNum<A> r{c >= 0 ? o.sub(s, x2) : o.sub(x2, s), c};
    const N one = o.from_u128(1);
    int steps = 0;
    while (r.s < 0) {                             // s - (x-1)^2 = r + 2x - 1
        r = sadd_t(o, r, SNum<A>{o.sub(o.shl(x, 1), one), 1});
        x = o.dec(x);
        count_correction(steps);
    }
    for (;;) {                                    // s - (x+1)^2 = r - 2x - 1
        N two_x = o.shl(x, 1);
        if (r.s == 0 || o.cmp(r.m, two_x) <= 0) break;
        r = sadd_t(o, r, SNum<A>{o.add(two_x, one), -1});
        x = o.inc(x);
        count_correction(steps);
    }
    return x;
}

// 10^e by squaring. Bounds are made exact after every step: 10^(2^k) has
// about 3.3 bits per decimal digit, so bounds doubled from one limb would
// overstate its size about tenfold.
template <class A> static typename A::Num pow10(A& o, uint64_t e) {
    typename A::Num result = o.from_u128(1), base = o.from_u128(10);
    while (e) {
        if (e & 1) { result = o.mul(result, base); o.tighten(result); }
        e >>= 1;
        if (e) { base = o.sqr(base); o.tighten(base); }
    }
    return result;
}

// ---- divide-and-conquer decimal conversion ---------------------------------------
// Widths are W = B * 2^K digits; level l splits each node of width B*2^(K-l)
// by D_l = 10^(B*2^(K-l-1)) with one reciprocal per level and the level's
// quotient/remainder products batched. The divisors come bottom-up, D_l =
// D_{l+1}^2, and so do their reciprocals: 1/D_l = (1/D_{l+1})^2, so V_{l+1}^2
// carries about half of V_l's bits and one Newton step completes it. Leaves
// of B digits convert on the host by repeated division by 10^9 (no big
// multiplies).
static void leaf_digits(Mag x, size_t width, char* out) {
    std::vector<uint32_t> chunks;                // base 10^9, least significant first
    while (!x.empty()) {
        uint64_t r = 0;
        for (size_t i = x.size(); i-- > 0;) {
            uint64_t cur = (r << 32) | x[i];
            x[i] = (uint32_t)(cur / 1000000000u);
            r = cur % 1000000000u;
        }
        trim(x);
        chunks.push_back((uint32_t)r);
    }
    size_t pos = width;
    for (uint32_t c : chunks) {
        for (int k = 0; k < 9 && pos > 0; ++k) { out[--pos] = (char)('0' + c % 10); c /= 10; }
    }
    while (pos > 0) out[--pos] = '0';
}

template <class A> struct DecimalPlan {
    int K = 0;
    size_t B = 0, W = 0;
    std::vector<typename A::Num> Ds, Vs;
    std::vector<size_t> Lds;
};

// Levels, leaf width and the divisors with their reciprocals for a value of
// xbits bits, on backend o.
template <class A> static DecimalPlan<A> decimal_plan(A& o, size_t xbits) {
    DecimalPlan<A> p;
    const size_t need = (size_t)((double)xbits * 0.30102999566398120) + 2;
    while ((need >> p.K) > 2304) ++p.K;
    const int K = p.K;
    p.B = ((need + ((size_t)1 << K) - 1) / ((size_t)1 << K) + 8) / 9 * 9;   // leaf width, multiple of 9
    p.W = p.B << K;
    // Ds[l] = 10^(B 2^(K-1-l)); Vs[l] ~ 2^(2L) / (Ds[l] 2^32), L = Ld + 32, i.e.
    // 2^(2 Ld + 32) / Ds[l]. Vs[l+1]^2 / 2^z with z = 4 Ld[l+1] + 32 - 2 Ld[l]
    // approximates Vs[l] to about L/2 + 14 bits; its top L/2 + 16 bits seed the step.
    p.Ds.resize(K);
    p.Vs.resize(K);
    p.Lds.resize(K);
    for (int l = K - 1; l >= 0; --l) {
        p.Ds[l] = l == K - 1 ? pow10(o, p.B) : o.sqr(p.Ds[l + 1]);
        o.tighten(p.Ds[l]);
        p.Lds[l] = o.bits(p.Ds[l]);
        const typename A::Num d32 = o.shl(p.Ds[l], 32);
        const size_t L = p.Lds[l] + 32;
        if (l == K - 1) { p.Vs[l] = recip(o, d32, L); continue; }
        const size_t z = 4 * p.Lds[l + 1] + 32 - 2 * p.Lds[l];
        const size_t sh = L / 2 - 15;
        p.Vs[l] = recip_step(o, d32, L, o.shr(o.sqr(p.Vs[l + 1]), z + sh), sh);
        o.tighten(p.Vs[l]);
    }
    return p;
}

// In place (a 10^8-digit string is not copied into a second buffer).
static std::string strip_zeros(std::string&& out) {
    const size_t nz = out.find_first_not_of('0');
    if (nz == std::string::npos) return "0";
    out.erase(0, nz);
    return std::move(out);
}

static std::string to_decimal(const Mag& x, const Multiplier& M, const std::shared_ptr<void>& ahead = nullptr) {
    if (x.empty()) return "0";
    if (M.to_decimal) return M.to_decimal(x, M, ahead);
    HostArith o{M};
    const DecimalPlan<HostArith> p = decimal_plan(o, bitlen(x));
    const int K = p.K;
    const size_t B = p.B;
    std::string out(p.W, '0');
    std::vector<Mag> nodes{x};
    for (int lvl = 0; lvl < K; ++lvl) {
        const Mag& D = p.Ds[lvl];
        const size_t Ld = p.Lds[lvl];
        // q ~ x V / 2^(2 Ld + 32), x < 10^(2 half)
        const Mag& V = p.Vs[lvl];
        std::vector<Mag> xs(nodes.size());
        std::vector<std::pair<const Mag*, const Mag*>> pairs;
        for (size_t i = 0; i < nodes.size(); ++i) xs[i] = nodes[i];
        for (size_t i = 0; i < nodes.size(); ++i) pairs.push_back({&xs[i], &V});
        std::vector<Mag> prods = M.mul_many(pairs);
        std::vector<Mag> qs(nodes.size());
        const int k = g_perturb.load();
        for (size_t i = 0; i < nodes.size(); ++i) {
            qs[i] = shr(prods[i], 2 * Ld + 32);
            for (int j = 0; j < k; ++j) inc(qs[i]);
            if (k < 0 && !qs[i].empty()) dec(qs[i]);
        }
        pairs.clear();
        for (size_t i = 0; i < nodes.size(); ++i) pairs.push_back({&qs[i], &D});
        std::vector<Mag> qds = M.mul_many(pairs);
        std::vector<Mag> next;
        next.reserve(2 * nodes.size());
        for (size_t i = 0; i < nodes.size(); ++i) {
            Mag q = qs[i], r;
            fix_quotient(o, nodes[i], D, qds[i], q, &r);
            next.push_back(std::move(q));
            next.push_back(std::move(r));
        }
        nodes.swap(next);
    }
    // the leaves are independent: convert them on several host threads
    const size_t nth = std::max<size_t>(1, std::min<size_t>({(size_t)std::thread::hardware_concurrency(), (size_t)16,
                                                            nodes.size() / 64 + 1}));
    std::vector<std::thread> pool;
    for (size_t t = 0; t < nth; ++t)
        pool.emplace_back([&, t]() {
            for (size_t i = t; i < nodes.size(); i += nth) leaf_digits(nodes[i], B, &out[i * B]);
        });
    for (auto& th : pool) th.join();
    return strip_zeros(std::move(out));
}

// ---- binary splitting, bottom-up with batched levels ---------------------------------

static const uint64_t A_CONST = 13591409ULL;
static const uint64_t B_CONST = 545140134ULL;
static const uint64_t C3_OVER_24 = 10939058860032000ULL;

static PQT leaf(uint64_t k) {
    PQT r;
    if (k == 0) { r.P = {1}; r.Q = {1}; r.T = SInt{from_u128(A_CONST), 1}; return r; }
    u128 p = (u128)(6 * k - 5) * (2 * k - 1) * (6 * k - 1);
    u128 q = (u128)k * k * k * C3_OVER_24;
    u128 lin = (u128)A_CONST + (u128)B_CONST * k;
    Mag pm = from_u128(p);
    r.P = pm;
    r.Q = from_u128(q);
    // T = P * (A + B k) may exceed 128 bits: multiply by the 64-bit factor in two halves
    Mag t = mul_u32(pm, (uint32_t)lin);
    Mag t2 = shl(mul_u32(pm, (uint32_t)(lin >> 32)), 32);
    r.T = SInt{add(t, t2), (k & 1) ? -1 : 1};
    return r;
}

static PQT binary_split(uint64_t N, const Multiplier& M) {
    std::vector<PQT> lvl;
    if (M.groups) {
        lvl = M.groups(N);
    } else {
        lvl.resize(N);
        for (uint64_t k = 0; k < N; ++k) lvl[k] = leaf(k);
    }
    while (lvl.size() > 1) {
        const size_t pairs = lvl.size() / 2;
        const bool root = lvl.size() == 2;
        std::vector<std::pair<const Mag*, const Mag*>> jobs;
        jobs.reserve(4 * pairs);
        for (size_t i = 0; i < pairs; ++i) {
            const PQT& L = lvl[2 * i];
            const PQT& R = lvl[2 * i + 1];
            jobs.push_back({&L.Q, &R.Q});
            jobs.push_back({&L.T.m, &R.Q});
            jobs.push_back({&L.P, &R.T.m});
            if (!root) jobs.push_back({&L.P, &R.P});      // the root's P is never used
        }
        std::vector<Mag> prod = M.mul_many(jobs);
        std::vector<PQT> next;
        next.reserve(pairs + 1);
        const size_t per = root ? 3 : 4;
        for (size_t i = 0; i < pairs; ++i) {
            const PQT& L = lvl[2 * i];
            const PQT& R = lvl[2 * i + 1];
            PQT r;
            r.Q = std::move(prod[per * i]);
            SInt a{std::move(prod[per * i + 1]), L.T.s};
            SInt b{std::move(prod[per * i + 2]), R.T.s};
            if (a.m.empty()) a.s = 0;
            if (b.m.empty()) b.s = 0;
            r.T = sadd(a, b);
            if (!root) r.P = std::move(prod[per * i + 3]);
            next.push_back(std::move(r));
        }
        if (lvl.size() & 1) next.push_back(std::move(lvl.back()));
        lvl.swap(next);
    }
    return std::move(lvl[0]);
}

uint64_t terms_for_digits(uint64_t digits) {              // src/chudnovsky.cpp:54-59
    const double DIGITS_PER_TERM = 14.181647462725477;
    uint64_t n = (uint64_t)((double)digits / DIGITS_PER_TERM) + 2;
    return n < 2 ? 2 : n;
}

// pi_scaled = floor(426880 Q isqrt(10005 10^(2 prec)) / T) (src/chudnovsky.cpp:
// 130-140) on backend o, as the backend's number; ts[0] and ts[1] stamp the
// ends of the 10^prec and square-root phases.
template <class A>
static typename A::Num pi_scaled_num(A& o, const PQT& r, uint64_t prec, std::chrono::steady_clock::time_point* ts) {
    typedef typename A::Num N;
    N one = pow10(o, prec);
    N radicand = o.mul_u32(o.sqr(one), 10005u);
    o.tighten(radicand);
    ts[0] = std::chrono::steady_clock::now();
    N sqrtC = isqrt(o, radicand);
    o.tighten(sqrtC);
    ts[1] = std::chrono::steady_clock::now();
    N numerator = o.mul(o.mul_u32(o.pqt_q(r), 426880u), sqrtC);
    return div_floor(o, numerator, o.pqt_t(r), (N*)nullptr);   // T > 0 for the full range
}

template <class A>
static Mag pi_scaled_t(A& o, const PQT& r, uint64_t prec, std::chrono::steady_clock::time_point* ts) {
    return o.down(pi_scaled_num(o, r, prec, ts));
}

std::string compute_pi(uint64_t digits, const Multiplier& M, double* phase_ms) {
    using clk = std::chrono::steady_clock;
    auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    const uint64_t guard = 16, prec = digits + guard;
    const uint64_t N = terms_for_digits(prec);
    auto t0 = clk::now();
    // pi_scaled is about pi 10^prec: its bit length is floor(log2 pi + prec
    // log2 10) + 1 unless that is within rounding of an integer, and a plan
    // made for another length is not used, so the decimal plan can start now
    std::shared_ptr<void> ahead;
    if (M.plan_decimal)
        ahead = M.plan_decimal((size_t)std::floor(1.6514961294723187 + (double)prec * 3.3219280948873623) + 1);
    PQT r = binary_split(N, M);
    auto t1 = clk::now();
    clk::time_point ts[3];
    std::string s;
    const bool formed = (bool)M.pi_decimal;
    if (formed) {
        s = M.pi_decimal(r, prec, digits, ts, ahead);
    } else {
        Mag pi_scaled;
        if (M.pi_scaled) {
            pi_scaled = M.pi_scaled(r, prec, ts);
        } else {
            HostArith o{M};
            pi_scaled = pi_scaled_t(o, r, prec, ts);
        }
        ts[2] = clk::now();
        s = to_decimal(pi_scaled, M, ahead);
    }
    const clk::time_point t2 = ts[0], t3 = ts[1], t4 = ts[2];
    auto t5 = clk::now();
    if (phase_ms) {
        phase_ms[0] = ms(t0, t1); phase_ms[1] = ms(t1, t2); phase_ms[2] = ms(t2, t3);
        phase_ms[3] = ms(t3, t4); phase_ms[4] = ms(t4, t5); phase_ms[5] = (double)N;
    }
    if (formed) return s;
    if (digits == 0) return "3";
    // "3." and the first `digits` decimals, in place in s (prec + 1 digits)
    if (s.size() > (size_t)digits + 1) s.resize((size_t)digits + 1);
    s.insert(s.begin() + 1, '.');
    return s;
}

// ---- the device multiplier -------------------------------------------------------------
static void check(int rc) {
    if (rc != PDH_OK) throw HipError(rc, std::string("ntt-hip: ") + pdh_last_error());
}

// One stream per host thread for the device-resident steps, created on first
// use and kept (creating and destroying a stream per call costs more than
// most of the calls).
static void* work_stream() {
    static thread_local void* st = nullptr;
    if (!st) check(pdh_stream_create(&st));
    return st;
}

// Any size: products beyond one transform are blocked inside the library.
static Mag dev_mul(const Mag& a, const Mag& b) {
    if (a.empty() || b.empty()) return {};
    return mul_ntt_hip(a, b);
}

static Mag dev_sqr(const Mag& a) {
    if (a.empty()) return {};
    return sqr_ntt_hip(a);
}

// One batched launch sequence for every pair that fits a transform.
static std::vector<Mag> dev_mul_many(const std::vector<std::pair<const Mag*, const Mag*>>& pairs) {
    std::vector<Mag> out(pairs.size());
    std::vector<size_t> idx;
    size_t la = 0, lb = 0, lo = 0;
    for (size_t i = 0; i < pairs.size(); ++i) {
        const Mag& a = *pairs[i].first;
        const Mag& b = *pairs[i].second;
        if (a.empty() || b.empty()) continue;
        if (a.size() + b.size() > ((size_t)1 << 23)) { out[i] = dev_mul(a, b); continue; }
        idx.push_back(i);
        la += a.size(); lb += b.size(); lo += a.size() + b.size();
    }
    if (idx.empty()) return out;
    std::vector<uint32_t> A(la), B(lb);
    std::vector<size_t> ao(idx.size()), bo(idx.size()), oo(idx.size());
    size_t pa = 0, pb = 0, po = 0;
    for (size_t k = 0; k < idx.size(); ++k) {
        const Mag& a = *pairs[idx[k]].first;
        const Mag& b = *pairs[idx[k]].second;
        ao[k] = pa; bo[k] = pb; oo[k] = po;
        std::memcpy(&A[pa], a.data(), a.size() * 4);
        std::memcpy(&B[pb], b.data(), b.size() * 4);
        pa += a.size(); pb += b.size(); po += a.size() + b.size();
    }
    void *dA = nullptr, *dB = nullptr, *dO = nullptr, *dL = nullptr, *st = nullptr;
    auto release = [&]() { pdh_dev_free_async(dA, st); pdh_dev_free_async(dB, st); pdh_dev_free_async(dO, st); pdh_dev_free_async(dL, st); };
    try {
        st = work_stream();
        check(pdh_dev_alloc(la * 4, &dA));
        check(pdh_dev_alloc(lb * 4, &dB));
        check(pdh_dev_alloc(lo * 4, &dO));
        check(pdh_dev_alloc(idx.size() * 8, &dL));
        check(pdh_copy_htod(dA, A.data(), la * 4, st));
        check(pdh_copy_htod(dB, B.data(), lb * 4, st));
        std::vector<pdh_mul_desc> d(idx.size());
        for (size_t k = 0; k < idx.size(); ++k) {
            d[k].a = (const uint32_t*)dA + ao[k]; d[k].na = pairs[idx[k]].first->size();
            d[k].b = (const uint32_t*)dB + bo[k]; d[k].nb = pairs[idx[k]].second->size();
            d[k].out = (uint32_t*)dO + oo[k]; d[k].out_len = (uint64_t*)dL + k;
        }
        check(pdh_mul_batch_dev(d.data(), d.size(), st));
        std::vector<uint32_t> O(lo);
        std::vector<uint64_t> lens(idx.size());
        check(pdh_copy_dtoh(O.data(), dO, lo * 4, st));
        check(pdh_copy_dtoh(lens.data(), dL, idx.size() * 8, st));
        check(pdh_stream_sync(st));
        for (size_t k = 0; k < idx.size(); ++k) out[idx[k]].assign(O.begin() + oo[k], O.begin() + oo[k] + lens[k]);
    } catch (...) {
        release();
        throw;
    }
    release();
    return out;
}

// The root of a device tree left in HBM: Q and T in blocks this handle owns,
// freed in stream order on `st` (the stream of the thread that built the
// tree, which is also the one that reads them) when the handle goes.
struct DevRoot {
    std::vector<void*> blocks;
    void* st = nullptr;
    const uint32_t* Q = nullptr;
    const uint32_t* T = nullptr;
    uint64_t nQ = 0, nT = 0;
    ~DevRoot() { for (void* b : blocks) (void)pdh_dev_free_async(b, st); }
};

// The tree on the device. The leaves and the first six levels come from one
// kernel (pdh_bs_groups, one node per 64 consecutive terms); every level above
// is one pdh_bs_merge_dev call (batched products and the signed T sum in HBM),
// each level in a fresh arena, the previous one freed. Only lengths and signs
// cross PCIe per level (products beyond one transform run blocked inside the
// merge). The returned level is the root, its P (never used) left empty.
static std::vector<PQT> dev_levels(uint64_t N) {
    const int G = 6;
    const uint64_t groups = (N + (1u << G) - 1) >> G;
    const size_t cp = (size_t)3 << G, cq = (size_t)5 << G, ct = (size_t)6 << G;
    void *dP = nullptr, *dQ = nullptr, *dT = nullptr, *dL = nullptr, *dS = nullptr, *st = nullptr, *arena = nullptr;
    auto release = [&]() { pdh_dev_free_async(dP, st); pdh_dev_free_async(dQ, st); pdh_dev_free_async(dT, st); pdh_dev_free_async(dL, st);
                           pdh_dev_free_async(dS, st); pdh_dev_free_async(arena, st); };
    auto r4 = [](uint64_t x) { return (x + 3) & ~(uint64_t)3; };
    std::vector<PQT> out;
    try {
        st = work_stream();
        check(pdh_dev_alloc(groups * cp * 4, &dP));
        check(pdh_dev_alloc(groups * cq * 4, &dQ));
        check(pdh_dev_alloc(groups * ct * 4, &dT));
        check(pdh_dev_alloc(groups * 3 * 8, &dL));
        check(pdh_dev_alloc(groups * 4, &dS));
        check(pdh_bs_groups(N, G, (uint32_t*)dP, (uint32_t*)dQ, (uint32_t*)dT, (uint64_t*)dL, (int32_t*)dS, st));
        std::vector<uint64_t> L(groups * 3);
        std::vector<int32_t> S(groups);
        check(pdh_copy_dtoh(L.data(), dL, L.size() * 8, st));
        check(pdh_copy_dtoh(S.data(), dS, S.size() * 4, st));
        check(pdh_stream_sync(st));
        std::vector<pdh_bs_node> lvl(groups);
        for (uint64_t g = 0; g < groups; ++g)
            lvl[g] = pdh_bs_node{(uint32_t*)dP + g * cp, L[3 * g], (uint32_t*)dQ + g * cq, L[3 * g + 1],
                                 (uint32_t*)dT + g * ct, L[3 * g + 2], S[g]};
        while (lvl.size() > 1) {
            const size_t pairs = lvl.size() / 2, n_out = (lvl.size() + 1) / 2;
            const bool root = lvl.size() == 2;
            uint64_t words = 0;
            std::vector<uint64_t> cP(n_out), cQ(n_out), cT(n_out);
            for (size_t i = 0; i < pairs; ++i) {
                const pdh_bs_node& A = lvl[2 * i];
                const pdh_bs_node& B = lvl[2 * i + 1];
                cP[i] = root ? 0 : A.nP + B.nP;
                cQ[i] = A.nQ + B.nQ;
                cT[i] = std::max(A.nT + B.nQ, A.nP + B.nT) + 1;
            }
            if (lvl.size() & 1) {
                const pdh_bs_node& v = lvl.back();
                cP[pairs] = root ? 0 : v.nP; cQ[pairs] = v.nQ; cT[pairs] = std::max<uint64_t>(v.nT, 1);
            }
            for (size_t i = 0; i < n_out; ++i) words += r4(cP[i]) + r4(cQ[i]) + r4(cT[i]);
            void* next = nullptr;
            check(pdh_dev_alloc(words * 4, &next));
            std::vector<pdh_bs_node> nl(n_out);
            uint32_t* w = (uint32_t*)next;
            for (size_t i = 0; i < n_out; ++i) {
                nl[i].P = cP[i] ? w : nullptr; w += r4(cP[i]);
                nl[i].Q = w; w += r4(cQ[i]);
                nl[i].T = w; w += r4(cT[i]);
            }
            const int rc = pdh_bs_merge_dev(lvl.data(), lvl.size(), nl.data(), root ? 1 : 0, st);
            if (rc != PDH_OK) { pdh_dev_free_async(next, st); check(rc); }
            // the inputs are no longer needed: the groups' buffers or the previous arena
            if (arena) { pdh_dev_free_async(arena, st); arena = nullptr; }
            else { pdh_dev_free_async(dP, st); pdh_dev_free_async(dQ, st); pdh_dev_free_async(dT, st); dP = dQ = dT = nullptr; }
            arena = next;
            lvl.swap(nl);
        }
        // the root stays in HBM: its blocks pass to a handle, and the tail
        // of the computation reads Q and T on this thread's stream
        const pdh_bs_node& v = lvl[0];
        auto root = std::make_shared<DevRoot>();
        root->st = st;
        if (arena) { root->blocks.push_back(arena); arena = nullptr; }
        else { root->blocks = {dP, dQ, dT}; dP = dQ = dT = nullptr; }
        root->Q = v.Q; root->nQ = v.nQ;
        root->T = v.T; root->nT = v.nT;
        out.resize(1);
        out[0].T.s = v.nT ? v.sT : 0;
        out[0].dev = root;
    } catch (...) {
        release();
        throw;
    }
    release();
    return out;
}

// ---- device backend -----------------------------------------------------------------
// Values live in device memory: a buffer of cap limbs (zero above the value's
// length) and the normalized length as a device word. ub is a host-side upper
// bound of that length, tight enough to size the next operation's output and
// the multiply's transform. Buffers are carved from per-backend chunks and
// released together when the backend ends, so the operations allocate nothing
// and run asynchronously on one stream; the host reads lengths, signs and
// small values only where the algorithm branches on them (bit lengths,
// comparisons, base cases). Uploads are synchronous copies into fresh
// buffers that no queued work uses yet.
struct DNum {
    uint32_t* p = nullptr;
    uint64_t* len = nullptr;
    uint64_t cap = 0, ub = 0;
};

class DevArith {
public:
    typedef DNum Num;
    explicit DevArith(void* st) : st_(st) {
        sign_ = (int32_t*)alloc(1).p;
        cmp_sign_ = (int32_t*)alloc(1).p;
        one_ = up(Mag{1});
    }
    ~DevArith() {
        for (void* c : chunks_) pdh_dev_free_async(c, st_);
    }
    DevArith(const DevArith&) = delete;
    DevArith& operator=(const DevArith&) = delete;

    void sync() { check(pdh_stream_sync(st_)); }
    void* stream() const { return st_; }
    void tighten(DNum& a) { a.ub = length(a); }
    DNum up(const Mag& a) {
        DNum x = alloc(a.size());
        const uint64_t n = a.size();
        if (n) check(pdh_copy_htod(x.p, a.data(), n * 4, nullptr));
        check(pdh_copy_htod(x.len, &n, 8, nullptr));
        x.ub = n;
        return x;
    }
    Mag down(const DNum& a) {
        Mag m(length(a));
        if (!m.empty()) check(pdh_copy_dtoh(m.data(), a.p, m.size() * 4, nullptr));
        return m;
    }
    size_t bits(const DNum& a) {
        const uint64_t n = length(a);
        if (!n) return 0;
        uint32_t top = 0;
        check(pdh_copy_dtoh(&top, a.p + n - 1, 4, nullptr));
        return 32 * (n - 1) + (32 - __builtin_clz(top));
    }
    bool zero(const DNum& a) { return length(a) == 0; }
    u128 to_u128(const DNum& a) {
        const uint64_t n = length(a);
        if (n > 4) throw HipError(PDH_E_ARG, "ntt-hip: to_u128 of a value above 2^128");
        uint32_t w[4] = {0, 0, 0, 0};
        if (n) check(pdh_copy_dtoh(w, a.p, n * 4, nullptr));
        return ((u128)w[3] << 96) | ((u128)w[2] << 64) | ((u128)w[1] << 32) | w[0];
    }
    DNum from_u128(u128 v) { return up(chud::from_u128(v)); }
    // A caller's device buffer of n limbs (the top ones may be zero) as a
    // normalized value in this backend's memory.
    DNum view(const uint32_t* p, uint64_t n) {
        DNum x = alloc(1);
        x.p = const_cast<uint32_t*>(p);
        x.cap = x.ub = std::max<uint64_t>(n, 1);
        check(pdh_copy_htod(x.len, &n, 8, nullptr));
        DNum y = sadd(x, one_, 0, x.ub);
        tighten(y);
        return y;
    }
    // Writes a into a caller's device buffer of cap limbs (zero above its
    // length) and its length into *len; PDH_E_CAPACITY if it does not fit.
    void store(const DNum& a, uint32_t* out, uint64_t cap, uint64_t* len) {
        if (length(a) > cap) throw HipError(PDH_E_CAPACITY, "ntt-hip: result longer than the output capacity");
        check(pdh_shift_dev(a.p, a.len, 0, out, std::max<uint64_t>(cap, 1), len, st_));
        sync();
    }
    DNum shl(const DNum& a, size_t k) {
        DNum x = alloc(a.ub + k / 32 + 1);
        check(pdh_shift_dev(a.p, a.len, (int64_t)k, x.p, x.cap, x.len, st_));
        return x;
    }
    DNum shr(const DNum& a, size_t k) {
        const uint64_t w = k / 32;
        DNum x = alloc(a.ub > w ? a.ub - w : 1);
        check(pdh_shift_dev(a.p, a.len, -(int64_t)k, x.p, x.cap, x.len, st_));
        return x;
    }
    DNum add(const DNum& a, const DNum& b) { return sadd(a, b, 1, std::max(a.ub, b.ub) + 1); }
    DNum sub(const DNum& a, const DNum& b) { return sadd(a, b, -1, a.ub); }   // a >= b
    int cmp(const DNum& a, const DNum& b) {
        DNum t = alloc(std::max(a.ub, b.ub) + 1);
        check(pdh_sadd_dev(a.p, a.len, 1, b.p, b.len, -1, t.p, t.cap, t.len, cmp_sign_, st_));
        const uint64_t n = length(t);
        if (!n) return 0;
        int32_t sg = 0;
        check(pdh_copy_dtoh(&sg, cmp_sign_, 4, nullptr));
        return sg;
    }
    DNum pow2(size_t e) { return shl(one_, e); }
    // Q and T of a binary-splitting root: copies of the device tree's
    // buffers when it was left in HBM, else uploads
    DNum pqt_q(const PQT& r) {
        const DevRoot* d = (const DevRoot*)r.dev.get();
        return d ? view(d->Q, d->nQ) : up(r.Q);
    }
    DNum pqt_t(const PQT& r) {
        const DevRoot* d = (const DevRoot*)r.dev.get();
        return d ? view(d->T, d->nT) : up(r.T.m);
    }
    DNum mul(const DNum& a, const DNum& b) {
        DNum x = alloc(a.ub + b.ub);
        check(pdh_mul_dev(a.p, a.ub, b.p, b.ub, x.p, x.len, st_));
        return x;
    }
    DNum sqr(const DNum& a) {
        DNum x = alloc(2 * a.ub);
        check(pdh_sqr_dev(a.p, a.ub, x.p, x.len, st_));
        return x;
    }
    DNum inc(const DNum& a) { return add(a, one_); }
    DNum dec(const DNum& a) { return sub(a, one_); }
    DNum mul_u32(const DNum& a, uint32_t m) { return mul(a, up(Mag{m})); }

private:
    DNum alloc(uint64_t cap) {
        cap = std::max<uint64_t>(cap, 1);
        const size_t bytes = ((cap * 4 + 255) & ~(size_t)255) + 256;
        if (bytes > left_) {
            const size_t cb = std::max(bytes, (size_t)64 << 20);
            void* c = nullptr;
            check(pdh_dev_alloc(cb, &c));
            chunks_.push_back(c);
            cur_ = (char*)c;
            left_ = cb;
        }
        DNum x;
        x.p = (uint32_t*)cur_;
        x.len = (uint64_t*)(cur_ + bytes - 256);
        x.cap = x.ub = cap;
        cur_ += bytes;
        left_ -= bytes;
        return x;
    }
    uint64_t length(const DNum& a) {
        sync();
        uint64_t n = 0;
        check(pdh_copy_dtoh(&n, a.len, 8, nullptr));
        return n;
    }
    DNum sadd(const DNum& a, const DNum& b, int sb, uint64_t cap) {
        DNum x = alloc(cap);
        check(pdh_sadd_dev(a.p, a.len, 1, b.p, b.len, sb, x.p, x.cap, x.len, sign_, st_));
        return x;
    }
    void* st_;
    std::vector<void*> chunks_;
    char* cur_ = nullptr;
    size_t left_ = 0;
    int32_t* sign_ = nullptr;
    int32_t* cmp_sign_ = nullptr;
    DNum one_;
};

// The test switch on a decimal level's quotient estimates (even slots of Xn,
// stride 2 Cn, device lengths XnL): + k each for k > 0, - 1 on those that are
// at least 1 for k < 0. The fix-up on the host then corrects them.
static void perturb_quotients(DevArith& o, uint32_t* Xn, uint64_t* XnL, uint64_t n, uint64_t Cn, void* st) {
    const int k = g_perturb.load();
    std::vector<uint64_t> ln(2 * n);
    check(pdh_copy_dtoh(ln.data(), XnL, 2 * n * 8, st));
    check(pdh_stream_sync(st));
    const DNum kk = o.up(Mag{(uint32_t)(k > 0 ? k : 1)});
    // the sums go to scratch with their own length words (a signed sum's
    // output length must not be its operand's length word), then back
    void* tmp = nullptr;
    const uint64_t tl = (n * Cn + 63) & ~(uint64_t)63;
    check(pdh_dev_alloc(tl * 4 + n * 8 + n * 4, &tmp));
    uint32_t* T = (uint32_t*)tmp;
    uint64_t* TL = (uint64_t*)(T + tl);
    int32_t* TS = (int32_t*)(TL + n);
    std::vector<pdh_sadd_desc> ad;
    std::vector<pdh_shr_desc> back;
    for (uint64_t i = 0; i < n; ++i) {
        if (k < 0 && ln[2 * i] == 0) continue;
        uint32_t* q = Xn + 2 * i * Cn;
        ad.push_back(pdh_sadd_desc{q, XnL + 2 * i, 1, kk.p, kk.len, k > 0 ? 1 : -1, T + i * Cn, Cn, TL + i, TS + i});
        back.push_back(pdh_shr_desc{T + i * Cn, TL + i, q, Cn, XnL + 2 * i, 0});
    }
    int rc = ad.empty() ? PDH_OK : pdh_sadd_batch_dev(ad.data(), ad.size(), st);
    if (rc == PDH_OK && !back.empty()) rc = pdh_shr_batch_dev(back.data(), back.size(), st);
    pdh_dev_free_async(tmp, st);
    check(rc);
}

// Divide-and-conquer decimal conversion in device memory. Level l holds 2^l
// nodes of width W_l = B 2^(K-l) digits at a fixed stride C_l limbs (values
// below 10^W_l, zero-padded, device lengths). Per level, with D = 10^(W_l/2)
// and its reciprocal V from the plan: P = X V (one batch), Q = P >> (2 Ld + 32)
// into the next level's even slots, QD = Q D (one batch), R = X - QD into the
// odd slots and T = |R| - D (signed sums). Q is floor(X / D) when R >= 0 and
// T < 0; otherwise (rare: the estimate is within a unit) the node is fixed
// by steps of D on the host. The leaves' digits come from one kernel; only
// the lengths, signs and the final W characters cross PCIe.
// A decimal plan made ahead on its own stream and host thread. Members are
// destroyed in reverse order: the task is waited for before its backend and
// plan go; the backend's buffers are released on the plan stream, so the
// user of the plan synchronizes its own stream first.
struct DevPlanAhead {
    size_t xbits = 0;
    std::unique_ptr<DevArith> o;
    DecimalPlan<DevArith> p;
    std::future<void> done;
};

// One plan stream per calling host thread, kept (as work_stream).
static void* plan_stream() {
    static thread_local void* st = nullptr;
    if (!st) check(pdh_stream_create(&st));
    return st;
}

static std::shared_ptr<void> dev_plan_decimal(size_t xbits) {
    auto h = std::make_shared<DevPlanAhead>();
    h->xbits = xbits;
    void* st = plan_stream();
    DevPlanAhead* raw = h.get();
    h->done = std::async(std::launch::async, [raw, st]() {
        raw->o.reset(new DevArith(st));
        raw->p = decimal_plan(*raw->o, raw->xbits);
        raw->o->sync();
    });
    return h;
}

// A value's known digit count and the form wanted: its first digit, then (for
// keep > 0) a point and the next keep digits.
struct DecForm { uint64_t ndig, keep; };

// The decimal digits of x (leading zeros stripped), or with `form` (and the
// first digit where form says) the formed string, *formed then set.
static std::string dev_decimal(DevArith& o, const DNum& x, const std::shared_ptr<void>& ahead,
                               const DecForm* form = nullptr, bool* formed = nullptr) {
    if (formed) *formed = false;
    void* st = o.stream();
    DevPlanAhead* pa = static_cast<DevPlanAhead*>(ahead.get());
    if (pa) pa->done.get();                      // the plan's work has run (or its error is rethrown)
    const size_t xb = o.bits(x);
    DecimalPlan<DevArith> own;
    if (!pa || pa->xbits != xb) own = decimal_plan(o, xb);
    const DecimalPlan<DevArith>& p = (pa && pa->xbits == xb) ? pa->p : own;
    const int K = p.K;
    const size_t B = p.B;
    auto cap_of = [](uint64_t w) { return (uint64_t)((double)w * 3.3219280948873623 / 32.0) + 3; };
    auto r64 = [](uint64_t v) { return (v + 63) & ~(uint64_t)63; };
    void *cur = nullptr, *nxt = nullptr, *aux = nullptr, *dout = nullptr;
    auto release = [&]() { pdh_dev_free_async(cur, st); pdh_dev_free_async(nxt, st); pdh_dev_free_async(aux, st); pdh_dev_free_async(dout, st); };
    std::string out;
    try {
        // a level: n nodes x C limbs, then n lengths
        uint64_t n = 1, C = cap_of(p.W);
        check(pdh_dev_alloc(r64(n * C) * 4 + n * 8, &cur));
        o.store(x, (uint32_t*)cur, C, (uint64_t*)((uint32_t*)cur + r64(C)));   // zero above x's length
        for (int l = 0; l < K; ++l) {
            uint32_t* X = (uint32_t*)cur;
            uint64_t* XL = (uint64_t*)(X + r64(n * C));
            const DNum& D = p.Ds[l];
            const DNum& V = p.Vs[l];
            const uint64_t nD = D.ub, nV = V.ub, Cn = cap_of(p.W >> (l + 1));
            // next level
            check(pdh_dev_alloc(r64(2 * n * Cn) * 4 + 2 * n * 8, &nxt));
            uint32_t* Xn = (uint32_t*)nxt;
            uint64_t* XnL = (uint64_t*)(Xn + r64(2 * n * Cn));
            // aux: P (n x (C + nV)) | QD (n x (Cn + nD)) | T (n x Cn) | lengths | signs
            const uint64_t oQD = r64(n * (C + nV)), oT = oQD + r64(n * (Cn + nD)), oEnd = oT + r64(n * Cn);
            check(pdh_dev_alloc(oEnd * 4 + 3 * n * 8 + 2 * n * 4, &aux));
            uint32_t* A = (uint32_t*)aux;
            uint32_t *dD = D.p, *dV = V.p, *dP = A, *dQD = A + oQD, *dT = A + oT;
            uint64_t* PL = (uint64_t*)(A + oEnd);
            uint64_t *QDL = PL + n, *TL = QDL + n;
            const uint64_t* DL = D.len;
            int32_t* RS = (int32_t*)(TL + n);
            int32_t* TS = RS + n;
            const uint64_t sh = 2 * p.Lds[l] + 32;
            // products beyond one transform (the top levels of very long
            // conversions) go one by one through the library's blocked path
            auto products = [&](const std::vector<pdh_mul_desc>& md) {
                bool fits = true;
                for (const pdh_mul_desc& d : md) fits = fits && d.na + d.nb <= ((uint64_t)1 << 23);
                if (fits) { check(pdh_mul_batch_dev(md.data(), md.size(), st)); return; }
                for (const pdh_mul_desc& d : md) check(pdh_mul_dev(d.a, d.na, d.b, d.nb, d.out, d.out_len, st));
                check(pdh_stream_sync(st));
            };
            std::vector<pdh_mul_desc> md(n);
            for (uint64_t i = 0; i < n; ++i)
                md[i] = pdh_mul_desc{X + i * C, C, dV, nV, dP + i * (C + nV), PL + i};
            products(md);
            std::vector<pdh_shr_desc> sd(n);
            for (uint64_t i = 0; i < n; ++i)
                sd[i] = pdh_shr_desc{dP + i * (C + nV), PL + i, Xn + 2 * i * Cn, Cn, XnL + 2 * i, sh};
            check(pdh_shr_batch_dev(sd.data(), n, st));
            if (g_perturb.load() != 0) perturb_quotients(o, Xn, XnL, n, Cn, st);
            for (uint64_t i = 0; i < n; ++i)
                md[i] = pdh_mul_desc{Xn + 2 * i * Cn, Cn, dD, nD, dQD + i * (Cn + nD), QDL + i};
            products(md);
            std::vector<pdh_sadd_desc> ad(n);
            for (uint64_t i = 0; i < n; ++i)
                ad[i] = pdh_sadd_desc{X + i * C, XL + i, 1, dQD + i * (Cn + nD), QDL + i, -1,
                                      Xn + (2 * i + 1) * Cn, Cn, XnL + 2 * i + 1, RS + i};
            check(pdh_sadd_batch_dev(ad.data(), n, st));
            for (uint64_t i = 0; i < n; ++i)
                ad[i] = pdh_sadd_desc{Xn + (2 * i + 1) * Cn, XnL + 2 * i + 1, 1, dD, DL, -1, dT + i * Cn, Cn, TL + i,
                                      TS + i};
            check(pdh_sadd_batch_dev(ad.data(), n, st));
            std::vector<int32_t> sg(2 * n);
            std::vector<uint64_t> ln(2 * n), tl(n);
            check(pdh_copy_dtoh(sg.data(), RS, 2 * n * 4, st));
            check(pdh_copy_dtoh(ln.data(), XnL, 2 * n * 8, st));
            check(pdh_copy_dtoh(tl.data(), TL, n * 8, st));
            check(pdh_stream_sync(st));
            for (uint64_t i = 0; i < n; ++i) {
                // signs are meaningful only for nonzero values: R < 0, or R > 0 and |R| - D >= 0
                const bool rnz = ln[2 * i + 1] != 0;
                const bool neg = rnz && sg[i] < 0;
                const bool ge = rnz && sg[i] > 0 && (tl[i] == 0 || sg[n + i] > 0);
                if (!neg && !ge) continue;
                Mag q(ln[2 * i]), rm(ln[2 * i + 1]);
                check(pdh_copy_dtoh(q.data(), Xn + 2 * i * Cn, q.size() * 4, st));
                check(pdh_copy_dtoh(rm.data(), Xn + (2 * i + 1) * Cn, rm.size() * 4, st));
                check(pdh_stream_sync(st));
                const Mag Dh = o.down(D);
                SInt r{rm, sg[i]};
                int steps = 0;
                while (r.s < 0) { r = sadd(r, SInt{Dh, 1}); dec(q); count_correction(steps); }
                while (r.s > 0 && cmp(r.m, Dh) >= 0) { r = sadd(r, SInt{Dh, -1}); inc(q); count_correction(steps); }
                Mag qz(Cn, 0), rz(Cn, 0);
                std::copy(q.begin(), q.end(), qz.begin());
                std::copy(r.m.begin(), r.m.end(), rz.begin());
                const uint64_t lens[2] = {q.size(), r.m.size()};
                check(pdh_copy_htod(Xn + 2 * i * Cn, qz.data(), Cn * 4, nullptr));
                check(pdh_copy_htod(Xn + (2 * i + 1) * Cn, rz.data(), Cn * 4, nullptr));
                check(pdh_copy_htod(XnL + 2 * i, lens, 16, nullptr));
            }
            pdh_dev_free_async(aux, st); aux = nullptr;
            pdh_dev_free_async(cur, st);
            cur = nxt; nxt = nullptr;
            n *= 2;
            C = Cn;
        }
        uint32_t* X = (uint32_t*)cur;
        uint64_t* XL = (uint64_t*)(X + r64(n * C));
        check(pdh_dev_alloc(p.W, &dout));
        check(pdh_decimal_leaves_dev(X, C, XL, n, (uint32_t)B, (char*)dout, st));
        const char* dd = (const char*)dout;
        if (form && formed && form->ndig >= 1 && form->ndig <= p.W && form->keep < form->ndig) {
            // only the digits wanted cross to the host, once the digit
            // before the first (if any) is seen to be 0 and the first not
            const size_t nz = p.W - form->ndig;
            char edge[2] = {'0', '0'};
            check(pdh_copy_dtoh(nz ? edge : edge + 1, dd + (nz ? nz - 1 : 0), nz ? 2 : 1, st));
            check(pdh_stream_sync(st));
            if (edge[0] == '0' && edge[1] != '0') {
                out.assign(form->keep ? form->keep + 2 : 1, '.');
                out[0] = edge[1];
                if (form->keep) check(pdh_copy_dtoh(&out[2], dd + nz + 1, form->keep, st));
                check(pdh_stream_sync(st));
                if (formed) *formed = true;
            }
        }
        if (!formed || !*formed) {
            out.assign(p.W, '0');
            check(pdh_copy_dtoh(&out[0], dout, p.W, st));
            check(pdh_stream_sync(st));
        }
    } catch (...) {
        release();
        (void)pdh_stream_sync(st);               // nothing queued still reads the plan when it is freed
        throw;
    }
    release();
    if (formed && *formed) return out;
    return strip_zeros(std::move(out));
}

static std::string dev_to_decimal(const Mag& x, const Multiplier&, const std::shared_ptr<void>& ahead) {
    DevArith o(work_stream());
    return dev_decimal(o, o.up(x), ahead);
}

static Mag dev_pi_scaled(const PQT& r, uint64_t prec, std::chrono::steady_clock::time_point* ts) {
    DevArith o(work_stream());
    return pi_scaled_t(o, r, prec, ts);
}

// The square root, the division and the decimal conversion with every value
// in HBM (the root's Q and T from the device tree, pi_scaled handed over to
// the conversion without a host copy). The division's memory goes (in stream
// order) once its result is copied into the conversion's backend.
static std::string dev_pi_decimal(const PQT& r, uint64_t prec, uint64_t digits,
                                  std::chrono::steady_clock::time_point* ts, const std::shared_ptr<void>& ahead) {
    void* st = work_stream();
    std::unique_ptr<DevArith> od(new DevArith(st));
    DNum xd = pi_scaled_num(*od, r, prec, ts);
    od->tighten(xd);
    ts[2] = std::chrono::steady_clock::now();
    DevArith o(st);
    const DNum x = o.view(xd.p, xd.ub);
    od.reset();
    // floor(pi 10^prec) lies in [3 10^prec, 4 10^prec): prec + 1 digits
    const DecForm f{prec + 1, digits};
    bool formed = false;
    std::string s = dev_decimal(o, x, ahead, &f, &formed);
    if (formed) return s;
    if (digits == 0) return s.substr(0, 1);
    if (s.size() > (size_t)digits + 1) s.resize((size_t)digits + 1);
    s.insert(s.begin() + 1, '.');
    return s;
}

Multiplier device_multiplier() {
    Multiplier M;
    M.mul = dev_mul;
    M.sqr = dev_sqr;
    M.mul_many = dev_mul_many;
    M.groups = dev_levels;
    M.to_decimal = dev_to_decimal;
    M.plan_decimal = dev_plan_decimal;
    M.pi_scaled = dev_pi_scaled;
    M.pi_decimal = dev_pi_decimal;
    return M;
}

} // namespace chud

std::string compute_pi_chudnovsky_hip(uint64_t digits, double* phase_ms) {
    return chud::compute_pi(digits, chud::device_multiplier(), phase_ms);
}

} // namespace pidigits

// ---- exact floor division and square root on the seam ------------------------
namespace {
using pidigits::chud::Mag;

Mag host_mag(const uint32_t* p, size_t n) {
    Mag m(p, p + n);
    while (!m.empty() && m.back() == 0) m.pop_back();
    return m;
}

int put(const Mag& v, uint32_t* out, size_t cap, size_t* len) {
    if (len) *len = v.size();
    if (!out) return PDH_OK;                          // only the length (or nothing) wanted
    if (v.size() > cap) return PDH_E_CAPACITY;
    if (!v.empty()) std::memcpy(out, v.data(), v.size() * 4);
    return PDH_OK;
}

pidigits::chud::Multiplier callback_multiplier(pdh_host_mul_fn mul) {
    using namespace pidigits::chud;
    Multiplier M;
    M.mul = [mul](const Mag& a, const Mag& b) -> Mag {
        if (a.empty() || b.empty()) return {};
        Mag r(a.size() + b.size());
        size_t n = 0;
        if (mul(a.data(), a.size(), b.data(), b.size(), r.data(), r.size(), &n) != 0)
            throw pidigits::HipError(PDH_E_HIP, "callback multiply failed");
        r.resize(n);
        return r;
    };
    M.sqr = [M](const Mag& a) { return M.mul(a, a); };
    M.mul_many = [M](const std::vector<std::pair<const Mag*, const Mag*>>& p) {
        std::vector<Mag> r(p.size());
        for (size_t i = 0; i < p.size(); ++i) r[i] = M.mul(*p[i].first, *p[i].second);
        return r;
    };
    return M;
}

// q, r of a / d (d != 0) and floor(sqrt(s)) on backend o.
template <class A> void divmod_on(A& o, const Mag& a, const Mag& d, Mag* q, Mag* r) {
    typename A::Num rem;
    typename A::Num qq = pidigits::chud::div_floor(o, o.up(a), o.up(d), &rem);
    *q = o.down(qq);
    *r = o.down(rem);
}

int divmod_any(pdh_host_mul_fn mul, const uint32_t* a, size_t na, const uint32_t* d, size_t nd, uint32_t* q,
               size_t q_cap, size_t* q_len, uint32_t* r, size_t r_cap, size_t* r_len) {
    if ((na && !a) || (nd && !d)) return PDH_E_ARG;
    const Mag A = host_mag(a, na), D = host_mag(d, nd);
    if (D.empty()) return PDH_E_ARG;                  // division by zero
    Mag Q, R;
    try {
        if (mul) {
            const pidigits::chud::Multiplier M = callback_multiplier(mul);
            pidigits::chud::HostArith o{M};
            divmod_on(o, A, D, &Q, &R);
        } else {
            pidigits::chud::DevArith o(pidigits::chud::work_stream());
            divmod_on(o, A, D, &Q, &R);
        }
    } catch (const pidigits::HipError& e) {
        return e.code;
    }
    const int rc = put(Q, q, q_cap, q_len);
    return rc ? rc : put(R, r, r_cap, r_len);
}

int isqrt_any(pdh_host_mul_fn mul, const uint32_t* s, size_t ns, uint32_t* x, size_t x_cap, size_t* x_len) {
    if (ns && !s) return PDH_E_ARG;
    const Mag S = host_mag(s, ns);
    Mag X;
    try {
        if (S.empty()) {
            X.clear();
        } else if (mul) {
            const pidigits::chud::Multiplier M = callback_multiplier(mul);
            pidigits::chud::HostArith o{M};
            X = o.down(pidigits::chud::isqrt(o, o.up(S)));
        } else {
            pidigits::chud::DevArith o(pidigits::chud::work_stream());
            X = o.down(pidigits::chud::isqrt(o, o.up(S)));
        }
    } catch (const pidigits::HipError& e) {
        return e.code;
    }
    return put(X, x, x_cap, x_len);
}
} // namespace

// ---- C ABI ------------------------------------------------------------------------------
extern "C" {

static int emit(const std::string& s, char* out, size_t cap) {
    if (s.size() + 1 > cap) return PDH_E_CAPACITY;
    std::memcpy(out, s.c_str(), s.size() + 1);
    return PDH_OK;
}

int pdh_pi_chudnovsky(uint64_t digits, char* out, size_t cap, double* phase_ms) {
    try {
        return emit(pidigits::compute_pi_chudnovsky_hip(digits, phase_ms), out, cap);
    } catch (const pidigits::HipError& e) {
        return e.code;
    }
}

int pdh_pi_chudnovsky_cb(uint64_t digits, pdh_host_mul_fn mul, char* out, size_t cap, double* phase_ms) {
    using namespace pidigits::chud;
    Multiplier M;
    M.mul = [mul](const Mag& a, const Mag& b) -> Mag {
        if (a.empty() || b.empty()) return {};
        Mag r(a.size() + b.size());
        size_t n = 0;
        if (mul(a.data(), a.size(), b.data(), b.size(), r.data(), r.size(), &n) != 0) throw pidigits::HipError(PDH_E_HIP, "callback multiply failed");
        r.resize(n);
        return r;
    };
    M.sqr = [M](const Mag& a) { return M.mul(a, a); };
    M.mul_many = [M](const std::vector<std::pair<const Mag*, const Mag*>>& p) {
        std::vector<Mag> r(p.size());
        for (size_t i = 0; i < p.size(); ++i) r[i] = M.mul(*p[i].first, *p[i].second);
        return r;
    };
    try {
        return emit(compute_pi(digits, M, phase_ms), out, cap);
    } catch (const pidigits::HipError& e) {
        return e.code;
    }
}

uint64_t pdh_chudnovsky_terms(uint64_t digits) { return pidigits::chud::terms_for_digits(digits + 16); }


int pdh_divmod(const uint32_t* a, size_t na, const uint32_t* d, size_t nd, uint32_t* q, size_t q_cap,
               size_t* q_len, uint32_t* r, size_t r_cap, size_t* r_len) {
    return divmod_any(nullptr, a, na, d, nd, q, q_cap, q_len, r, r_cap, r_len);
}

int pdh_divmod_cb(pdh_host_mul_fn mul, const uint32_t* a, size_t na, const uint32_t* d, size_t nd, uint32_t* q,
                  size_t q_cap, size_t* q_len, uint32_t* r, size_t r_cap, size_t* r_len) {
    if (!mul) return PDH_E_ARG;
    return divmod_any(mul, a, na, d, nd, q, q_cap, q_len, r, r_cap, r_len);
}

int pdh_isqrt(const uint32_t* s, size_t ns, uint32_t* x, size_t x_cap, size_t* x_len) {
    return isqrt_any(nullptr, s, ns, x, x_cap, x_len);
}

int pdh_isqrt_cb(pdh_host_mul_fn mul, const uint32_t* s, size_t ns, uint32_t* x, size_t x_cap, size_t* x_len) {
    if (!mul) return PDH_E_ARG;
    return isqrt_any(mul, s, ns, x, x_cap, x_len);
}

int pdh_divmod_dev(const uint32_t* d_a, size_t na, const uint32_t* d_d, size_t nd, uint32_t* d_q, size_t q_cap,
                   uint64_t* d_q_len, uint32_t* d_r, size_t r_cap, uint64_t* d_r_len, void* stream) {
    if ((na && !d_a) || !d_d || nd == 0 || (d_q && !d_q_len) || (d_r && !d_r_len)) return PDH_E_ARG;
    using namespace pidigits::chud;
    try {
        DevArith o(stream);
        const DNum A = o.view(d_a, na), D = o.view(d_d, nd);
        if (o.zero(D)) return PDH_E_ARG;
        DNum R;
        const DNum Q = div_floor(o, A, D, &R);
        if (d_q) o.store(Q, d_q, q_cap, d_q_len);
        if (d_r) o.store(R, d_r, r_cap, d_r_len);
    } catch (const pidigits::HipError& e) {
        return e.code;
    }
    return PDH_OK;
}

int pdh_isqrt_dev(const uint32_t* d_s, size_t ns, uint32_t* d_x, size_t x_cap, uint64_t* d_x_len, void* stream) {
    if ((ns && !d_s) || !d_x || !d_x_len) return PDH_E_ARG;
    using namespace pidigits::chud;
    try {
        DevArith o(stream);
        const DNum S = o.view(d_s, ns);
        o.store(o.zero(S) ? S : isqrt(o, S), d_x, x_cap, d_x_len);
    } catch (const pidigits::HipError& e) {
        return e.code;
    }
    return PDH_OK;
}

int pdh_newton_set_perturb(int k) {
    if (k < -8 || k > 8) return PDH_E_ARG;
    pidigits::chud::g_perturb.store(k);
    return PDH_OK;
}

uint64_t pdh_newton_corrections(void) { return pidigits::chud::g_corrections.load(); }

} // extern "C"
