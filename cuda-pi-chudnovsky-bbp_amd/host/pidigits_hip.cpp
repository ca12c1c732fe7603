// C++ mirror of the reference seams over the C ABI (see pidigits_hip.h).
#include "pidigits_hip.h"
#include "pdh_capi.h"

namespace pidigits {

static void check(int rc) {
    if (rc != PDH_OK) throw HipError(rc, std::string("ntt-hip: ") + pdh_last_error());
}

bool hip_available() { return pdh_device_available() != 0; }

std::vector<uint32_t> mul_ntt_hip(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
    if (a.empty() || b.empty()) return {};
    std::vector<uint32_t> out(a.size() + b.size());
    size_t n = 0;
    check(pdh_mul(a.data(), a.size(), b.data(), b.size(), out.data(), out.size(), &n));
    out.resize(n);
    return out;
}

std::vector<uint32_t> sqr_ntt_hip(const std::vector<uint32_t>& a) {
    if (a.empty()) return {};
    std::vector<uint32_t> out(2 * a.size());
    size_t n = 0;
    check(pdh_sqr(a.data(), a.size(), out.data(), out.size(), &n));
    out.resize(n);
    return out;
}

static char to_hex_char(int v) { return (char)(v < 10 ? ('0' + v) : ('A' + (v - 10))); }

std::string bbp_hex_digits_hip(uint64_t start, uint32_t count) {
    if (count == 0) return std::string();
    std::vector<uint8_t> v(count);
    check(pdh_bbp_range(start, count, v.data()));
    std::string s(count, '0');
    for (uint32_t i = 0; i < count; ++i) s[i] = to_hex_char(v[i]);
    return s;
}

char bbp_hex_digit_hip(uint64_t position) { return bbp_hex_digits_hip(position, 1)[0]; }

std::vector<uint8_t> bbp_values_hip(const std::vector<uint64_t>& positions) {
    std::vector<uint8_t> v(positions.size());
    if (!positions.empty()) check(pdh_bbp_positions(positions.data(), positions.size(), v.data()));
    return v;
}

} // namespace pidigits
