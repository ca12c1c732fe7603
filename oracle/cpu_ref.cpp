// ============================================================================
//  oracle/cpu_ref.cpp -- TEST INFRASTRUCTURE ONLY.
//
//  A plain C++ restatement of the reference's CPU path for the multiply and
//  BBP hot path. It is the checker: only tests/, __graft_entry__.smoke() and
//  bench.py's cpu_baseline leg load it. The product library never links,
//  loads or calls it, and it is never the thing measured or shipped.
//
//  What each function follows (paths relative to the reference repository):
//    field         include/goldilocks.h:83-189
//    sizing/glue   include/ntt.h:37-109
//    NTT multiply  src/ntt_cpu.cpp:64-129   (radix-2 DIT, bit reversal first,
//                                            twiddle recurrence w *= w_len)
//    schoolbook    src/bignum.cpp:125-143
//    BBP           include/bbp_kernel.h:70-137, src/bbp_cpu.cpp:20-34
//
//  Pinning (tests/test_oracle.py): field closed forms and brute-force
//  unsigned __int128 % p; NTT multiply == schoolbook on random and
//  closed-form operands; carry/repack against Python big integers; BBP
//  against the reference's 64 hex digits and Bailey's 10^6 checkpoint; the
//  whole oracle against oracle/_ref (the reference's own ntt_cpu.cpp and
//  bbp_cpu.cpp compiled from source) on seeded inputs.
//
//  Built with -O2 -ffp-contract=off (no FMA contraction, no fast-math) so
//  the FP64 operation sequence of the BBP series is the reference's.
// ============================================================================
#include <cstdint>
#include <cstddef>
#include <cstring>
#include <cmath>
#include <vector>
#include <thread>
#include <atomic>
#include <utility>

namespace orc {

// ---- field: GF(p), p = 2^64 - 2^32 + 1 (goldilocks.h:83-85) ----------------
constexpr uint64_t P   = 0xFFFFFFFF00000001ULL;
constexpr uint64_t EPS = 0xFFFFFFFFULL;
constexpr uint64_t G   = 7ULL;

static inline uint64_t mulhi(uint64_t a, uint64_t b) {           // goldilocks.h:90-98
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
}
static inline uint64_t add(uint64_t a, uint64_t b) {             // goldilocks.h:104-109
    uint64_t s = a + b;
    if (s < a) s += EPS;
    if (s >= P) s -= P;
    return s;
}
static inline uint64_t sub(uint64_t a, uint64_t b) {             // goldilocks.h:114-118
    uint64_t d = a - b;
    if (a < b) d -= EPS;
    return d;
}
static inline uint64_t reduce128(uint64_t lo, uint64_t hi) {     // goldilocks.h:130-154
    uint32_t hh = (uint32_t)(hi >> 32);
    uint32_t hl = (uint32_t)(hi & 0xFFFFFFFFu);
    uint64_t t0 = lo - (uint64_t)hh;
    if (lo < (uint64_t)hh) t0 -= EPS;
    uint64_t t1 = (uint64_t)hl * EPS;
    uint64_t t2 = t0 + t1;
    if (t2 < t0) {
        uint64_t before = t2;
        t2 += EPS;
        if (t2 < before) t2 += EPS;
    }
    while (t2 >= P) t2 -= P;
    return t2;
}
static inline uint64_t mul(uint64_t a, uint64_t b) {             // goldilocks.h:158-162
    return reduce128(a * b, mulhi(a, b));
}
static inline uint64_t pow(uint64_t base, uint64_t e) {          // goldilocks.h:165-173
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = mul(r, base);
        base = mul(base, base);
        e >>= 1;
    }
    return r;
}
static inline uint64_t inv(uint64_t a) { return pow(a, P - 2); }            // :178-180
static inline uint64_t root_of_unity(uint64_t n) { return pow(G, (P - 1) / n); } // :187-189

// ---- NTT glue (ntt.h) --------------------------------------------------------
static inline size_t next_pow2(size_t x) { size_t n = 1; while (n < x) n <<= 1; return n; } // :37-41
static inline size_t ntt_length(size_t na, size_t nb) { return next_pow2(2 * na + 2 * nb); } // :47-50

static std::vector<uint64_t> repack_to16(const uint32_t* a, size_t na, size_t length) { // :56-64
    std::vector<uint64_t> out(length, 0);
    size_t k = 0;
    for (size_t i = 0; i < na; ++i) {
        out[k++] = a[i] & 0xFFFFu;
        out[k++] = (a[i] >> 16) & 0xFFFFu;
    }
    return out;
}

static std::vector<uint32_t> carry_propagate(const uint64_t* conv, size_t n) {  // :70-91
    std::vector<uint32_t> out;
    out.reserve(n / 2 + 2);
    uint64_t carry = 0;
    size_t i = 0;
    while (i < n || carry) {
        uint64_t cur = carry + (i < n ? conv[i] : 0);
        uint32_t lo16 = (uint32_t)(cur & 0xFFFFu);
        carry = cur >> 16;
        ++i;
        uint64_t cur2 = carry + (i < n ? conv[i] : 0);
        uint32_t hi16 = (uint32_t)(cur2 & 0xFFFFu);
        carry = cur2 >> 16;
        ++i;
        out.push_back(lo16 | (hi16 << 16));
    }
    while (!out.empty() && out.back() == 0) out.pop_back();
    return out;
}

static inline uint32_t bit_reverse(uint32_t i, int logn) {       // :95-102
    uint32_t r = 0;
    for (int b = 0; b < logn; ++b) { r = (r << 1) | (i & 1); i >>= 1; }
    return r;
}
static inline int ilog2(size_t n) { int l = 0; while ((size_t(1) << l) < n) ++l; return l; } // :105-109

// ---- radix-2 NTT (ntt_cpu.cpp:64-104) ---------------------------------------
static void ntt_transform(std::vector<uint64_t>& a, bool inverse) {
    const size_t n = a.size();
    if (n <= 1) return;
    const int logn = ilog2(n);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t j = bit_reverse(i, logn);
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        uint64_t w_len = root_of_unity((uint64_t)len);
        if (inverse) w_len = inv(w_len);
        const size_t half = len >> 1;
        for (size_t i = 0; i < n; i += len) {
            uint64_t w = 1;
            for (size_t j = 0; j < half; ++j) {
                uint64_t u = a[i + j];
                uint64_t v = mul(a[i + j + half], w);
                a[i + j] = add(u, v);
                a[i + j + half] = sub(u, v);
                w = mul(w, w_len);
            }
        }
    }
    if (inverse) {
        uint64_t ninv = inv((uint64_t)n);
        for (size_t i = 0; i < n; ++i) a[i] = mul(a[i], ninv);
    }
}

// ntt_cpu.cpp:111-129
static std::vector<uint32_t> mul_ntt(const uint32_t* a, size_t na, const uint32_t* b, size_t nb) {
    if (na == 0 || nb == 0) return {};
    const size_t n = ntt_length(na, nb);
    std::vector<uint64_t> fa = repack_to16(a, na, n);
    std::vector<uint64_t> fb = repack_to16(b, nb, n);
    ntt_transform(fa, false);
    ntt_transform(fb, false);
    for (size_t i = 0; i < n; ++i) fa[i] = mul(fa[i], fb[i]);
    ntt_transform(fa, true);
    return carry_propagate(fa.data(), fa.size());
}

// bignum.cpp:125-143
static std::vector<uint32_t> mul_schoolbook(const uint32_t* a, size_t na, const uint32_t* b, size_t nb) {
    if (na == 0 || nb == 0) return {};
    std::vector<uint32_t> out(na + nb, 0);
    for (size_t i = 0; i < na; ++i) {
        uint64_t carry = 0;
        uint64_t ai = a[i];
        for (size_t j = 0; j < nb; ++j) {
            uint64_t cur = (uint64_t)out[i + j] + ai * (uint64_t)b[j] + carry;
            out[i + j] = (uint32_t)cur;
            carry = cur >> 32;
        }
        out[i + nb] += (uint32_t)carry;
    }
    while (!out.empty() && out.back() == 0) out.pop_back();
    return out;
}

// ---- BBP (bbp_kernel.h:70-137) ------------------------------------------------
static inline uint64_t bbp_mulmod(uint64_t a, uint64_t b, uint64_t m) {  // :70-83
    if ((m >> 32) == 0) return (a * b) % m;
    uint64_t hi = mulhi(a, b);
    uint64_t lo = a * b;
    uint64_t rem = 0;
    for (int sh = 48; sh >= 0; sh -= 16) rem = ((rem << 16) | ((hi >> sh) & 0xFFFFu)) % m;
    for (int sh = 48; sh >= 0; sh -= 16) rem = ((rem << 16) | ((lo >> sh) & 0xFFFFu)) % m;
    return rem;
}
static inline uint64_t bbp_modpow16(uint64_t e, uint64_t m) {            // :86-96
    if (m == 1) return 0;
    uint64_t result = 1;
    uint64_t base = 16 % m;
    while (e) {
        if (e & 1) result = bbp_mulmod(result, base, m);
        base = bbp_mulmod(base, base, m);
        e >>= 1;
    }
    return result;
}
static double bbp_series(uint32_t j, uint64_t d) {                       // :99-122
    double sum = 0.0;
    for (uint64_t k = 0; k <= d; ++k) {
        uint64_t m = 8ull * k + j;
        uint64_t r = bbp_modpow16(d - k, m);
        sum += (double)r / (double)m;
        sum -= std::floor(sum);
    }
    double p16 = 1.0 / 16.0;
    uint64_t k = d + 1;
    while (p16 > 1e-18) {
        uint64_t m = 8ull * k + j;
        sum += p16 / (double)m;
        p16 /= 16.0;
        ++k;
    }
    sum -= std::floor(sum);
    return sum;
}
static int bbp_digit_value(uint64_t pos) {                               // :128-137
    uint64_t d = pos - 1;
    double s = 4.0 * bbp_series(1, d)
             - 2.0 * bbp_series(4, d)
             - 1.0 * bbp_series(5, d)
             - 1.0 * bbp_series(6, d);
    s -= std::floor(s);
    if (s < 0.0) s += 1.0;
    return (int)(16.0 * s);
}
static inline char to_hex_char(int v) { return (char)(v < 10 ? ('0' + v) : ('A' + (v - 10))); } // bbp_cpu.cpp:20-22

// Number of bbp_mulmod calls bbp_modpow16(e, m) makes for m > 1:
// one squaring per bit of e plus one multiply per set bit.
static inline uint64_t modpow_calls(uint64_t e) {
    return e ? (uint64_t)(64 - __builtin_clzll(e)) + (uint64_t)__builtin_popcountll(e) : 0;
}

// Sum over e in [0, x) of popcount(e): closed form per bit.
static uint64_t popcount_prefix(uint64_t x) {
    uint64_t total = 0;
    for (int b = 0; b < 63; ++b) {
        uint64_t period = 1ull << (b + 1), half = 1ull << b;
        uint64_t full = x / period, rem = x % period;
        total += full * half + (rem > half ? rem - half : 0);
    }
    return total;
}
// Sum over e in [0, x) of bitlen(e).
static uint64_t bitlen_prefix(uint64_t x) {
    uint64_t total = 0;
    for (int L = 1; L < 64; ++L) {           // e in [2^(L-1), 2^L) has bitlen L
        uint64_t lo = 1ull << (L - 1), hi = (L < 63) ? (1ull << L) : ~0ull;
        if (x <= lo) break;
        uint64_t cnt = (x < hi ? x : hi) - lo;
        total += cnt * (uint64_t)L;
    }
    return total;
}

template <class F>
static void parallel_for(size_t n, int threads, F&& f) {
    if (threads <= 1 || n < 2) { for (size_t i = 0; i < n; ++i) f(i); return; }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&]() {
            for (size_t i; (i = next.fetch_add(1)) < n;) f(i);
        });
    for (auto& th : pool) th.join();
}

static int emit(const std::vector<uint32_t>& v, uint32_t* out, size_t cap, size_t* len) {
    if (len) *len = v.size();
    if (v.size() > cap) return -1;
    if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(uint32_t));
    return 0;
}

} // namespace orc

extern "C" {

uint64_t orc_gl_add(uint64_t a, uint64_t b) { return orc::add(a, b); }
uint64_t orc_gl_sub(uint64_t a, uint64_t b) { return orc::sub(a, b); }
uint64_t orc_gl_mul(uint64_t a, uint64_t b) { return orc::mul(a, b); }
uint64_t orc_gl_pow(uint64_t a, uint64_t e) { return orc::pow(a, e); }
uint64_t orc_gl_inv(uint64_t a) { return orc::inv(a); }
uint64_t orc_gl_root_of_unity(uint64_t n) { return orc::root_of_unity(n); }
uint64_t orc_gl_reduce128(uint64_t lo, uint64_t hi) { return orc::reduce128(lo, hi); }
size_t   orc_ntt_length(size_t na, size_t nb) { return orc::ntt_length(na, nb); }

void orc_repack_to16(const uint32_t* a, size_t na, uint64_t* out, size_t length) {
    std::vector<uint64_t> v = orc::repack_to16(a, na, length);
    std::memcpy(out, v.data(), length * sizeof(uint64_t));
}
int orc_carry_propagate(const uint64_t* conv, size_t n, uint32_t* out, size_t cap, size_t* len) {
    return orc::emit(orc::carry_propagate(conv, n), out, cap, len);
}
// Forward or inverse length-n transform in place (n a power of two).
void orc_ntt_transform(uint64_t* a, size_t n, int inverse) {
    std::vector<uint64_t> v(a, a + n);
    orc::ntt_transform(v, inverse != 0);
    std::memcpy(a, v.data(), n * sizeof(uint64_t));
}
int orc_mul_ntt(const uint32_t* a, size_t na, const uint32_t* b, size_t nb,
                uint32_t* out, size_t cap, size_t* len) {
    return orc::emit(orc::mul_ntt(a, na, b, nb), out, cap, len);
}
int orc_mul_schoolbook(const uint32_t* a, size_t na, const uint32_t* b, size_t nb,
                       uint32_t* out, size_t cap, size_t* len) {
    return orc::emit(orc::mul_schoolbook(a, na, b, nb), out, cap, len);
}

// Batch of independent multiplies, pairs spread over `threads` host threads,
// each multiply single-threaded (the reference's ntt-cpu is single-threaded).
// a_off/b_off/out_off index into the packed arrays; out_len[i] receives the
// normalized length. Returns 0, or -1 if an output slot is too small.
int orc_mul_ntt_batch(const uint32_t* a, const uint64_t* a_off, const uint64_t* na,
                      const uint32_t* b, const uint64_t* b_off, const uint64_t* nb,
                      uint32_t* out, const uint64_t* out_off, uint64_t* out_len,
                      size_t count, int threads) {
    std::atomic<int> bad{0};
    orc::parallel_for(count, threads, [&](size_t i) {
        std::vector<uint32_t> r = orc::mul_ntt(a + a_off[i], na[i], b + b_off[i], nb[i]);
        out_len[i] = r.size();
        if (r.size() > na[i] + nb[i]) { bad = 1; return; }
        if (!r.empty()) std::memcpy(out + out_off[i], r.data(), r.size() * sizeof(uint32_t));
    });
    return bad ? -1 : 0;
}

int  orc_bbp_digit_value(uint64_t pos) { return orc::bbp_digit_value(pos); }
double orc_bbp_series(uint32_t j, uint64_t d) { return orc::bbp_series(j, d); }
uint64_t orc_bbp_modpow16(uint64_t e, uint64_t m) { return orc::bbp_modpow16(e, m); }

// Uppercase hex digits for positions [start, start+count) (bbp_cpu.cpp:28-34).
void orc_bbp_hex_range(uint64_t start, uint32_t count, char* out, int threads) {
    orc::parallel_for(count, threads, [&](size_t i) {
        out[i] = orc::to_hex_char(orc::bbp_digit_value(start + i));
    });
}
// Digit values (0..15) at arbitrary 1-based positions.
void orc_bbp_positions(const uint64_t* pos, size_t n, uint8_t* vals, int threads) {
    orc::parallel_for(n, threads, [&](size_t i) {
        vals[i] = (uint8_t)orc::bbp_digit_value(pos[i]);
    });
}

// Exact count of bbp_mulmod calls the reference makes for one position:
// sum over j in {1,4,5,6} and k = 0..d of calls(16^(d-k) mod 8k+j); the
// j = 1, k = 0 term has m = 1 and makes none.
uint64_t orc_bbp_mulmod_count(uint64_t pos) {
    uint64_t d = pos - 1;
    uint64_t per_series = orc::bitlen_prefix(d + 1) + orc::popcount_prefix(d + 1);
    return 4 * per_series - orc::modpow_calls(d);
}
uint64_t orc_bbp_mulmod_count_brute(uint64_t pos) {
    uint64_t d = pos - 1, total = 0;
    const uint32_t js[4] = {1, 4, 5, 6};
    for (uint32_t j : js)
        for (uint64_t k = 0; k <= d; ++k)
            if (8 * k + j != 1) total += orc::modpow_calls(d - k);
    return total;
}

} // extern "C"
