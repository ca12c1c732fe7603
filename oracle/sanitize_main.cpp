// sanitize_main.cpp -- test infrastructure: drives the host side of the
// Chudnovsky engine (host/chudnovsky_hip.cpp through its callback entry
// points) and the oracle's multiply (cpu_ref.cpp), both compiled with
// AddressSanitizer and UndefinedBehaviorSanitizer by `make -C oracle
// sanitize`; no device is used. tests/test_sanitize_cpu.py runs it and
// compares what it prints with the golden fixtures.
//
//   pdh_sanitize pi DIGITS [PERTURB]   pi through pdh_pi_chudnovsky_cb, every
//                                      multiply by orc_mul_ntt; prints the string
//   pdh_sanitize arith                 divmod / isqrt through the callbacks on
//                                      seeded cases, checked by schoolbook
//                                      products; prints "arith ok N"
//   pdh_sanitize fail K                the K-th callback multiply fails: the
//                                      entry point must return an error (and
//                                      leak nothing); prints "fail rc=R calls=C"
#include "pdh_pi.h"
#include "pdh_diag.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

extern "C" int orc_mul_ntt(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out,
                           size_t out_cap, size_t* out_len);

namespace {

using Mag = std::vector<uint32_t>;

Mag school(const Mag& a, const Mag& b) {
    Mag r(a.size() + b.size() + 1, 0);
    for (size_t i = 0; i < a.size(); ++i) {
        uint64_t c = 0;
        for (size_t j = 0; j < b.size(); ++j) {
            const uint64_t t = (uint64_t)a[i] * b[j] + r[i + j] + c;
            r[i + j] = (uint32_t)t;
            c = t >> 32;
        }
        for (size_t k = i + b.size(); c; ++k) {
            const uint64_t t = (uint64_t)r[k] + c;
            r[k] = (uint32_t)t;
            c = t >> 32;
        }
    }
    while (!r.empty() && r.back() == 0) r.pop_back();
    return r;
}

Mag add(const Mag& a, const Mag& b) {
    Mag r(std::max(a.size(), b.size()) + 1, 0);
    uint64_t c = 0;
    for (size_t i = 0; i + 1 < r.size(); ++i) {
        const uint64_t t = (uint64_t)(i < a.size() ? a[i] : 0) + (i < b.size() ? b[i] : 0) + c;
        r[i] = (uint32_t)t;
        c = t >> 32;
    }
    r.back() = (uint32_t)c;
    while (!r.empty() && r.back() == 0) r.pop_back();
    return r;
}

int cmp(const Mag& a, const Mag& b) {
    if (a.size() != b.size()) return a.size() < b.size() ? -1 : 1;
    for (size_t i = a.size(); i-- > 0;)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

Mag norm(Mag a) {
    while (!a.empty() && a.back() == 0) a.pop_back();
    return a;
}

uint64_t g_state = 0x2545F4914F6CDD1DULL;
uint32_t rnd() {
    g_state ^= g_state << 13;
    g_state ^= g_state >> 7;
    g_state ^= g_state << 17;
    return (uint32_t)(g_state >> 16);
}

Mag random_mag(size_t n, bool all_ones = false) {
    Mag a(n);
    for (auto& v : a) v = all_ones ? 0xFFFFFFFFu : rnd();
    if (n) a.back() |= 0x80000000u;
    return a;
}

long g_calls = 0, g_fail_at = -1;
int failing_mul(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out, size_t out_cap,
                size_t* out_len) {
    if (++g_calls == g_fail_at) return -1;
    return orc_mul_ntt(a, na, b, nb, out, out_cap, out_len);
}

int check_divmod(const Mag& a, const Mag& d) {
    Mag q(a.size() + 1), r(d.size() + 1);
    size_t nq = 0, nr = 0;
    int rc = pdh_divmod_cb(orc_mul_ntt, a.data(), a.size(), d.data(), d.size(), q.data(), q.size(), &nq, r.data(),
                           r.size(), &nr);
    if (rc) return rc;
    q.resize(nq);
    r.resize(nr);
    if (cmp(norm(r), norm(d)) >= 0) return 100;
    return cmp(add(school(norm(q), norm(d)), norm(r)), norm(a)) == 0 ? 0 : 101;
}

int check_isqrt(const Mag& s) {
    Mag x(s.size() / 2 + 2);
    size_t nx = 0;
    int rc = pdh_isqrt_cb(orc_mul_ntt, s.data(), s.size(), x.data(), x.size(), &nx);
    if (rc) return rc;
    x.resize(nx);
    x = norm(x);
    // x^2 <= s < (x + 1)^2
    if (cmp(school(x, x), norm(s)) > 0) return 102;
    const Mag x1 = add(x, Mag{1});
    return cmp(school(x1, x1), norm(s)) > 0 ? 0 : 103;
}

int run_arith() {
    int n = 0, rc = 0;
    const size_t sizes[][2] = {{2, 1}, {5, 3}, {40, 17}, {300, 120}, {1000, 999}, {64, 64}, {2000, 33}};
    for (int perturb : {0, 3, -3}) {
        pdh_newton_set_perturb(perturb);
        for (const auto& s : sizes) {
            const Mag a = random_mag(s[0]), d = random_mag(s[1]);
            if ((rc = check_divmod(a, d))) return std::fprintf(stderr, "divmod %zu/%zu rc=%d\n", s[0], s[1], rc), 1;
            ++n;
        }
        // all-ones dividends over divisors 2^k - 1, and an exact multiple
        for (size_t k : {1, 2, 31, 77}) {
            Mag d(k, 0xFFFFFFFFu);
            if ((rc = check_divmod(random_mag(200, true), d))) return std::fprintf(stderr, "divmod ones rc=%d\n", rc), 1;
            ++n;
        }
        const Mag d = random_mag(90), q = random_mag(150);
        if ((rc = check_divmod(school(q, d), d))) return std::fprintf(stderr, "divmod exact rc=%d\n", rc), 1;
        ++n;
        for (size_t ns : {1, 2, 7, 64, 501}) {
            if ((rc = check_isqrt(random_mag(ns)))) return std::fprintf(stderr, "isqrt %zu rc=%d\n", ns, rc), 1;
            const Mag r = random_mag(ns);
            const Mag sq = school(r, r);
            if ((rc = check_isqrt(sq))) return std::fprintf(stderr, "isqrt square rc=%d\n", rc), 1;
            // one below a square: (r^2 - 1) has floor sqrt r - 1
            Mag below = sq;
            for (size_t i = 0; i < below.size(); ++i)
                if (below[i]--) break;
            if ((rc = check_isqrt(below))) return std::fprintf(stderr, "isqrt below rc=%d\n", rc), 1;
            n += 3;
        }
    }
    pdh_newton_set_perturb(0);
    std::printf("arith ok %d\n", n);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "pi" && argc >= 3) {
        const uint64_t digits = std::strtoull(argv[2], nullptr, 10);
        if (argc >= 4) pdh_newton_set_perturb(std::atoi(argv[3]));
        std::vector<char> out(digits + 64);
        double ph[6] = {0};
        const int rc = pdh_pi_chudnovsky_cb(digits, orc_mul_ntt, out.data(), out.size(), ph);
        if (rc) return std::fprintf(stderr, "pi rc=%d\n", rc), 1;
        std::printf("%s\n", out.data());
        return 0;
    }
    if (mode == "arith") return run_arith();
    if (mode == "fail" && argc >= 3) {
        g_fail_at = std::atol(argv[2]);
        std::vector<char> out(20000);
        double ph[6] = {0};
        const int rc = pdh_pi_chudnovsky_cb(10000, failing_mul, out.data(), out.size(), ph);
        std::printf("fail rc=%d calls=%ld\n", rc, g_calls);
        // the engine still works after a failed call
        const long calls = g_calls;
        g_fail_at = -1;
        const int rc2 = pdh_pi_chudnovsky_cb(1000, failing_mul, out.data(), out.size(), ph);
        std::printf("after rc=%d calls=%ld\n%s\n", rc2, g_calls - calls, out.data());
        return 0;
    }
    return 2;
}
