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

/* ---- serving: a continuous-batching request loop on one GPU ----------------
 * No counterpart in the reference (one computation per process run,
 * src/main.cpp:93-179). Requests wait in arrival order until the HBM budget
 * and the in-flight limit admit them; admitted requests run while earlier
 * ones are still on the device (multiplies admitted together share one
 * batched plan; BBP requests run in order on their own stream; Chudnovsky
 * requests run on a host worker whose multiplies go to the device). Results
 * stay in device memory (multiplies) until fetched. */
typedef struct pdh_server pdh_server;
int pdh_server_create(uint64_t hbm_budget_bytes, int max_inflight, pdh_server** srv);
int pdh_server_submit_mul(pdh_server* s, const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint64_t* id);
int pdh_server_submit_pi(pdh_server* s, uint64_t digits, uint64_t* id);
int pdh_server_submit_bbp(pdh_server* s, uint64_t start, uint32_t count, uint64_t* id);
/* Blocks until request `id` has finished; latency = finish - submission. */
int pdh_server_wait(pdh_server* s, uint64_t id, double* latency_ms);
/* Copies the result (multiply: normalized u32 limbs; pi: "3.14..." chars;
 * BBP: digit values 0..15), *len in elements, and releases the request. */
int pdh_server_result(pdh_server* s, uint64_t id, void* out, size_t cap, size_t* len);
/* Finishes the requests already running, then frees the server. */
int pdh_server_destroy(pdh_server* s);

#ifdef __cplusplus
}
#endif
#endif /* PDH_PI_H */
