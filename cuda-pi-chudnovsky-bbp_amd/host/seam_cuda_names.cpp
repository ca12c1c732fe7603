// ============================================================================
//  seam_cuda_names.cpp -- the reference's device-backend symbols on ntt-hip.
//
//  The reference's include/bignum.h:197-205 and include/bbp.h:34 declare
//      std::vector<uint32_t> pidigits::mul_ntt_cuda(const std::vector<uint32_t>&,
//                                                   const std::vector<uint32_t>&);
//      bool                  pidigits::cuda_available();
//      std::string           pidigits::bbp_hex_digits_cuda(uint64_t, uint32_t);
//  and expect them from src/ntt_cuda.cu / src/bbp_cuda.cu, or from a stub
//  translation unit in builds without CUDA (include/bignum.h:197-199). This
//  file is such a unit for an MI355X build: linked in their place
//  (libpdh_seam.so), it lets the reference's unchanged bignum.cpp
//  (mul_dispatch, src/bignum.cpp:204-232), chudnovsky.cpp, main.cpp and
//  tests/self_test.cu run every NTT multiply and every BBP block on the
//  device. Same types and results; errors throw pidigits::HipError instead
//  of falling back to the host (there is no host compute path here).
// ============================================================================
#include "pidigits_hip.h"

#include <cstdint>
#include <string>
#include <vector>

namespace pidigits {

std::vector<uint32_t> mul_ntt_cuda(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
    return mul_ntt_hip(a, b);
}

bool cuda_available() { return hip_available(); }

std::string bbp_hex_digits_cuda(uint64_t start, uint32_t count) { return bbp_hex_digits_hip(start, count); }

} // namespace pidigits
