// ============================================================================
//  pidigits_hip.h -- C++ mirror of the reference's multiply and BBP seams,
//  backed by the ntt-hip C ABI (include/pdh_capi.h).
//
//  Same names, types and argument meaning as the reference's backends:
//    mul_ntt_hip          <-> mul_ntt_cuda   (include/bignum.h:200-201)
//    hip_available        <-> cuda_available (include/bignum.h:205)
//    bbp_hex_digits_hip   <-> bbp_hex_digits_cuda (include/bbp.h:34)
//    bbp_hex_digit_hip    <-> bbp_hex_digit  (include/bbp.h:27)
//  None of these names is declared by the reference's headers, so this
//  header can be included next to its bignum.h, bbp.h and chudnovsky.h: a
//  reference build adds NttHip to its own MulBackend (include/bignum.h:77-83)
//  and routes `case MulBackend::NttHip` in mul_dispatch here. Alternatively
//  seam_cuda_names.cpp defines the reference's own mul_ntt_cuda,
//  cuda_available and bbp_hex_digits_cuda on these functions, so the
//  reference's sources link against this library unchanged.
//
//  Error behaviour differs on purpose: the reference falls back to its CPU
//  path on any device error (src/ntt_cuda.cu:182,241; src/bbp_cuda.cu:42,64).
//  These functions throw pidigits::HipError instead, so a product computed
//  on the host is never mistaken for a device result.
// ============================================================================
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace pidigits {

struct HipError : std::runtime_error {
    int code;
    HipError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

bool hip_available();

// Normalized product of little-endian base-2^32 magnitudes (empty = zero).
std::vector<uint32_t> mul_ntt_hip(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b);

// a * a with one forward transform.
std::vector<uint32_t> sqr_ntt_hip(const std::vector<uint32_t>& a);

// Uppercase hex digits at 1-based positions [start, start + count).
std::string bbp_hex_digits_hip(uint64_t start, uint32_t count);
char bbp_hex_digit_hip(uint64_t position);

// Digit values (0..15) at arbitrary positions.
std::vector<uint8_t> bbp_values_hip(const std::vector<uint64_t>& positions);

// "3." + the first `digits` decimal digits of pi, every multiply on the
// device (reference: compute_pi_chudnovsky, src/chudnovsky.cpp:107-164).
// phase_ms, if given, receives 6 values (see pdh_pi.h).
std::string compute_pi_chudnovsky_hip(uint64_t digits, double* phase_ms = nullptr);

} // namespace pidigits
