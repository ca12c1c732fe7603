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
int pdh_divmod_cb(pdh_host_mul_fn mul, const uint32_t* a, size_t na, const uint32_t* d, size_t nd, uint32_t* q,
                  size_t q_cap, size_t* q_len, uint32_t* r, size_t r_cap, size_t* r_len);
int pdh_isqrt_cb(pdh_host_mul_fn mul, const uint32_t* s, size_t ns, uint32_t* x, size_t x_cap, size_t* x_len);

/* ---- exact floor division and square root -------------------------------
 * The reference's div_fast and isqrt slots (src/bignum_div.cpp:170-178 and
 * :189-204; include/bignum.h:220,227): q = floor(a / d), r = a - q d for
 * d > 0, and floor(sqrt(s)), on little-endian base-2^32 magnitudes (top
 * limbs may be zero; results normalized). Computed by Newton iterations
 * whose multiplies run on the device (a reciprocal with exact floor
 * correction; a division-free inverse square root, then x -> floor root by
 * r = s - x^2), so the cost is a few multiplies, not the reference's
 * quadratic Knuth division. PDH_E_ARG for d = 0; PDH_E_CAPACITY if a result
 * is longer than its capacity (its length is still written). A null q or r
 * skips that output.
 *
 * pdh_divmod, pdh_isqrt: host buffers. pdh_divmod_dev, pdh_isqrt_dev: device
 * buffers on `stream` (inputs of na, nd, ns limbs; outputs zero-padded to
 * their capacity, normalized lengths in device uint64_t words); the Newton
 * steps read lengths and signs back, so these synchronise the stream.
 * pdh_divmod_cb, pdh_isqrt_cb: the same algorithm with every multiply done
 * by `mul` on host buffers (the CPU tests run it with the oracle's
 * multiply). */
int pdh_divmod(const uint32_t* a, size_t na, const uint32_t* d, size_t nd, uint32_t* q, size_t q_cap,
               size_t* q_len, uint32_t* r, size_t r_cap, size_t* r_len);
int pdh_isqrt(const uint32_t* s, size_t ns, uint32_t* x, size_t x_cap, size_t* x_len);
int pdh_divmod_dev(const uint32_t* d_a, size_t na, const uint32_t* d_d, size_t nd, uint32_t* d_q, size_t q_cap,
                   uint64_t* d_q_len, uint32_t* d_r, size_t r_cap, uint64_t* d_r_len, void* stream);
int pdh_isqrt_dev(const uint32_t* d_s, size_t ns, uint32_t* d_x, size_t x_cap, uint64_t* d_x_len, void* stream);

/* Number of series terms for `digits` digits (16 guard digits included),
 * as the reference's chudnovsky_terms_for_digits(digits + 16)
 * (src/chudnovsky.cpp:54-59, 118-120). */
uint64_t pdh_chudnovsky_terms(uint64_t digits);

/* ---- serving: a continuous-batching request loop on one GPU ----------------
 * No counterpart in the reference (one computation per process run,
 * src/main.cpp:93-179). Requests wait until the HBM budget (bytes each
 * request may hold on the device) and the in-flight limit admit them, in
 * arrival order except that a request short of budget lets later ones pass
 * for up to 20 ms; admitted requests run while earlier ones are still on
 * the device (multiplies admitted together share one batched plan on a
 * stream of their own; BBP and Chudnovsky requests run on two host workers
 * each, whose device work overlaps with everything else) and complete as
 * soon as their own work has finished. Results stay in device memory
 * (multiplies) until fetched. */
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
/* Counters: stats[0] requests submitted, [1] completed, [2] requests that
 * found the budget short at least once and waited, [3] admissions that
 * passed an earlier waiting request (backfill), [4] peak admitted bytes,
 * [5] bytes held now, [6] the budget. Writes min(n, 7) values. */
int pdh_server_stats(pdh_server* s, uint64_t* stats, int n);
/* Drops requests not yet admitted, finishes the running ones, then frees
 * the server. */
int pdh_server_destroy(pdh_server* s);

#ifdef __cplusplus
}
#endif
#endif /* PDH_PI_H */
