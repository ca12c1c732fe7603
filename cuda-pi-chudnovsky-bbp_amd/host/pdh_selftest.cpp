// ============================================================================
//  pdh_selftest -- the reference self-test's multiply and BBP checks
//  (tests/self_test.cu:64-78, 124-136) restated against the C++ mirror.
//  Needs a gfx950 GPU; exits non-zero on any failure. Its schoolbook is the
//  test's own ground truth, not part of the library.
// ============================================================================
#include "pidigits_hip.h"
#include "pdh_capi.h"
#include "pdh_diag.h"

#include <cstdio>
#include <string>
#include <vector>

using namespace pidigits;

static int g_fail = 0, g_pass = 0;
#define CHECK(cond, msg) do { if (cond) ++g_pass; else { ++g_fail; std::printf("  FAIL: %s\n", msg); } } while (0)

static uint64_t s_state = 0x9e3779b97f4a7c15ULL;
static uint64_t rnd64() { s_state ^= s_state << 13; s_state ^= s_state >> 7; s_state ^= s_state << 17; return s_state; }
static uint32_t rnd32() { return (uint32_t)rnd64(); }
static std::vector<uint32_t> randmag(size_t n) {
    std::vector<uint32_t> v(n);
    for (auto& x : v) x = rnd32();
    while (!v.empty() && v.back() == 0) v.pop_back();
    return v;
}
static std::vector<uint32_t> schoolbook(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
    if (a.empty() || b.empty()) return {};
    std::vector<uint32_t> out(a.size() + b.size(), 0);
    for (size_t i = 0; i < a.size(); ++i) {
        uint64_t carry = 0;
        for (size_t j = 0; j < b.size(); ++j) {
            uint64_t cur = (uint64_t)out[i + j] + (uint64_t)a[i] * b[j] + carry;
            out[i + j] = (uint32_t)cur;
            carry = cur >> 32;
        }
        out[i + b.size()] += (uint32_t)carry;
    }
    while (!out.empty() && out.back() == 0) out.pop_back();
    return out;
}

int main() {
    std::printf("================ pdh_selftest ================\n");
    if (!hip_available()) { std::printf("no gfx950 device: %s\n", pdh_last_error()); return 2; }
    uint64_t launches = pdh_launch_count();
    std::printf("[1] ntt-hip multiply == schoolbook\n");
    for (int t = 0; t < 120; ++t) {
        size_t na = 1 + rnd32() % 1500, nb = 1 + rnd32() % 1500;
        auto a = randmag(na), b = randmag(nb);
        auto sb = schoolbook(a, b);
        CHECK(mul_ntt_hip(a, b) == sb, "ntt-hip == schoolbook");
        CHECK(sqr_ntt_hip(a) == schoolbook(a, a), "ntt-hip square == schoolbook");
    }
    CHECK(pdh_launch_count() > launches, "device kernels ran");
    std::printf("[2] Chudnovsky pi, every multiply on the device\n");
    {
        static const char* ref100 = "1415926535897932384626433832795028841971693993751058209749445923078164062862089986280348253421170679";
        std::string p = compute_pi_chudnovsky_hip(8000);
        CHECK(p.size() == 8002 && p.compare(0, 2, "3.") == 0, "8000-digit result has the right shape");
        CHECK(p.compare(2, 100, ref100) == 0, "first 100 digits match");
    }
    std::printf("[3] BBP hex digits\n");
    CHECK(bbp_hex_digits_hip(1, 64) == "243F6A8885A308D313198A2E03707344A4093822299F31D0082EFA98EC4E6C89",
          "BBP 1..64 == reference");
    CHECK(bbp_hex_digits_hip(1000000, 10) == "26C65E52CB", "BBP checkpoint @1e6");
    std::printf("%d checks passed, %d failed\n", g_pass, g_fail);
    return g_fail ? 1 : 0;
}
