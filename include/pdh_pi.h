/* ===========================================================================
 *  pdh_pi.h -- C ABI of libpidigits_hip.so: the callers of the ntt-hip seam
 *  that the build adds on the host side (Chudnovsky with Newton division and
 *  square root and divide-and-conquer decimal output, every multiply on the
 *  device). Same status codes as pdh_capi.h.
 * ========================================================================= */
#ifndef PDH_PI_H
#define PDH_PI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* "3." followed by the first `digits` decimal digits of pi, NUL-terminated,
 * into out[cap]. phase_ms (optional, 6 doubles) receives the wall time in ms
 * of binary splitting, 10^prec and the radicand, the square root, the final
 * multiply and division, the decimal conversion, and the number of terms.
 * Replaces: ChudnovskyResult compute_pi_chudnovsky(uint64_t digits, bool)
 * (reference include/chudnovsky.h, src/chudnovsky.cpp:107-164) run with every
 * multiply on the device; the digits are the same. */
int pdh_pi_chudnovsky(uint64_t digits, char* out, size_t cap, double* phase_ms);

/* The same computation with every multiply done by `mul` (host buffers,
 * pdh_mul's contract); lets the algorithm run without a device. */
typedef int (*pdh_host_mul_fn)(const uint32_t* a, size_t na, const uint32_t* b, size_t nb,
                               uint32_t* out, size_t out_cap, size_t* out_len);
int pdh_pi_chudnovsky_cb(uint64_t digits, pdh_host_mul_fn mul, char* out, size_t cap, double* phase_ms);

/* Number of series terms for `digits` digits (16 guard digits included),
 * as the reference's chudnovsky_terms_for_digits(digits + 16)
 * (src/chudnovsky.cpp:54-59, 118-120). */
uint64_t pdh_chudnovsky_terms(uint64_t digits);

#ifdef __cplusplus
}
#endif
#endif /* PDH_PI_H */
