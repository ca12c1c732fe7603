/* ===========================================================================
 *  pdh_capi.h -- C ABI of the MI355X (gfx950) ntt-hip multiply + BBP engine.
 *
 *  The only interface host code sees of HIP: plain pointers, sizes and int
 *  status codes; no C++ types, no exceptions, no HIP types in signatures
 *  (streams are passed as void*, a hipStream_t or NULL for the null stream).
 *
 *  Big integers are little-endian base-2^32 limb arrays, as the reference's
 *  BigInt::mag (include/bignum.h:108). Products are written normalized
 *  (no most-significant zero limbs; length 0 for zero), as mul_ntt_cpu
 *  returns them (src/ntt_cpu.cpp:111-129, include/ntt.h:89).
 *
 *  Every function returns PDH_OK (0) or a negative PDH_E* code; the text of
 *  the last error on the calling thread is in pdh_last_error(). There is no
 *  CPU fallback: a missing device or a failed HIP call is reported, never
 *  silently replaced by a host computation.
 * ========================================================================= */
#ifndef PDH_CAPI_H
#define PDH_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PDH_OK            0
#define PDH_E_NODEVICE   -1   /* no usable gfx950 device                    */
#define PDH_E_HIP        -2   /* a HIP runtime call or kernel launch failed */
#define PDH_E_ARG        -3   /* bad argument (null pointer, zero position) */
#define PDH_E_CAPACITY   -4   /* output buffer smaller than required        */
#define PDH_E_TOO_LARGE  -5   /* a batch product beyond one 2^24 transform  */

/* Text of the last error on this thread ("" if none). */
const char* pdh_last_error(void);

/* 1 if a gfx950 device is usable, else 0.
 * Replaces: bool cuda_available()  (reference include/bignum.h:205,
 * src/ntt_cuda.cu:165-173). */
int pdh_device_available(void);

/* Transform length N = next_pow2(2*na + 2*nb) the multiply uses
 * (reference include/ntt.h:47-50). */
size_t pdh_ntt_length(size_t na, size_t nb);

/* out = a * b, host buffers. out_cap >= na + nb limbs; *out_len receives the
 * normalized length. na == 0 or nb == 0 gives *out_len = 0. Any size: a
 * product of more than 2^23 limbs (beyond one 2^24 transform) is assembled
 * on the device from 2^22-limb block products.
 * Replaces: std::vector<uint32_t> mul_ntt_cuda(const std::vector<uint32_t>&,
 * const std::vector<uint32_t>&)  (reference include/bignum.h:200-201,
 * src/ntt_cuda.cu:179-243). */
int pdh_mul(const uint32_t* a, size_t na, const uint32_t* b, size_t nb,
            uint32_t* out, size_t out_cap, size_t* out_len);

/* out = a * a with one forward transform, host buffers, out_cap >= 2*na.
 * Replaces: mul_ntt_cuda(a, a) (the reference has no squaring path; same
 * product, include/bignum.h:200-201). */
int pdh_sqr(const uint32_t* a, size_t na, uint32_t* out, size_t out_cap, size_t* out_len);

/* Device-resident multiply: d_a, d_b, d_out are device pointers; d_out holds
 * na + nb limbs (the top limb may be zero); *d_out_len (device uint64_t)
 * receives the normalized length. Enqueued on `stream`, asynchronous. Calls
 * on different streams (from any threads) share the library's workspace in
 * turn: each waits on the device, not on the host, for the previous user.
 * Same product as pdh_mul; the operands stay in HBM (no PCIe traffic).
 * Replaces: the device part of mul_ntt_cuda (reference src/ntt_cuda.cu:
 * 198-233) without its per-call cudaMalloc/H2D/D2H/host carry. */
int pdh_mul_dev(const uint32_t* d_a, size_t na, const uint32_t* d_b, size_t nb,
                uint32_t* d_out, uint64_t* d_out_len, void* stream);

/* Device-resident square, d_out holds 2*na limbs. */
int pdh_sqr_dev(const uint32_t* d_a, size_t na, uint32_t* d_out, uint64_t* d_out_len,
                void* stream);

/* Allocation-free variant of pdh_mul_dev. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments and product as pdh_mul_dev, for na + nb <= PDH_NOALLOC_MAX_LIMBS
 * (2^20 limbs; larger products fail with PDH_E_TOO_LARGE). Its transforms
 * run in a fixed ~40 MiB scratch array in device memory that every GPU the
 * library runs on holds until the process ends; the call itself makes no
 * hipMalloc, hipFree or other allocating call. The library must have been
 * set up first: call pdh_device_available() once (set-up resolves the
 * scratch address and builds the twiddle tables; it allocates), with the
 * same device current as for the calls. Calls on one stream run in order;
 * calls on different streams take the scratch in turn (each stream waits on
 * the device for the previous user), so they do not overlap. */
#define PDH_NOALLOC_MAX_LIMBS ((size_t)1 << 20)
int pdh_mul_dev_noalloc(const uint32_t* d_a, size_t na, const uint32_t* d_b, size_t nb,
                        uint32_t* d_out, uint64_t* d_out_len, void* stream);

/* ---- device memory and streams for C/C++ callers ----------------------------
 * Thin wrappers so host code that builds batches or keeps operands resident
 * never includes HIP headers. Copies with a stream are asynchronous.
 * pdh_dev_free waits for the device, as hipFree does; pdh_dev_free_async
 * returns at once and the block is reused only after the work enqueued on
 * `stream` up to the call has run (the block's last use must be on that
 * stream, or ordered before it). pdh_stream_query: 1 if every operation on
 * the stream has finished, 0 if some are still running, < 0 on error. */
int pdh_dev_alloc(size_t bytes, void** ptr);
int pdh_dev_free(void* ptr);
int pdh_dev_free_async(void* ptr, void* stream);
int pdh_copy_htod(void* dst, const void* src, size_t bytes, void* stream);
int pdh_copy_dtoh(void* dst, const void* src, size_t bytes, void* stream);
int pdh_stream_create(void** stream);
int pdh_stream_destroy(void* stream);
int pdh_stream_sync(void* stream);
int pdh_stream_query(void* stream);

/* ---- batched multiplies ----------------------------------------------------
 * Many independent multiplies of uneven sizes (e.g. one level of the
 * reference's binary-splitting tree, src/chudnovsky.cpp:64-102, whose P, Q, T
 * products the reference issues one mul_dispatch call at a time,
 * src/bignum.cpp:204-232). All pointers in a descriptor are device pointers;
 * out holds na + nb limbs and *out_len (device) receives the normalized
 * length, exactly as pdh_mul_dev would write them. */
typedef struct pdh_mul_desc {
    const uint32_t* a; uint64_t na;
    const uint32_t* b; uint64_t nb;
    uint32_t* out;
    uint64_t* out_len;
} pdh_mul_desc;

typedef struct pdh_batch pdh_batch;   /* opaque plan */

/* Plan a batch once: group by transform length, allocate the batch's own
 * workspace, upload the job tables. The descriptors are copied; the device
 * buffers they point to must stay valid while the plan is run. */
int pdh_batch_create(const pdh_mul_desc* descs, size_t n, pdh_batch** plan);
/* Enqueue every multiply of the plan on `stream` (one launch sequence per
 * transform-length class, classes on parallel internal streams joined back
 * to `stream`). After pdh_batch_capture the run replays one hipGraph. */
int pdh_batch_run(pdh_batch* plan, void* stream);
/* Capture the plan's launch sequence into a hipGraph (once). */
int pdh_batch_capture(pdh_batch* plan, void* stream);
/* Waits for the plan's work on the device, then frees it. */
int pdh_batch_destroy(pdh_batch* plan);
/* One-shot: create, run on `stream` and release the plan. Returns without
 * waiting for the device: the products and lengths are complete once
 * `stream` has passed this call (stream order); the descriptors are copied
 * before the call returns. */
int pdh_mul_batch_dev(const pdh_mul_desc* descs, size_t n, void* stream);

/* ---- transform shape tuning ----------------------------------------------
 * A 2^n-point transform (n > 13) runs as a column pass over the top L1 index
 * bits (T adjacent columns per workgroup, radix 2^lr_col butterflies) and a
 * fused row pass over the low n - L1 bits (radix 2^lr_row); n <= 13 runs in
 * one workgroup (L1 = 0; radix 2^lr_row on engine 0, radix 8 on engine 1).
 * On the three-prime engine lr_col is 3 and lr_row selects the kernels: 4
 * (default, rows of 2^8 points or more) register-resident radix-16 rows and
 * radix-8 columns, 3 radix-8 levels through LDS. Every valid shape gives the
 * same product; only speed differs. No counterpart in the reference (its
 * radix-2 transform has no parameters, src/ntt_cuda.cu:136-159). */
int pdh_tune_get(int n, int* L1, int* T, int* lr_col, int* lr_row);
/* PDH_E_ARG if the shape cannot be launched. */
int pdh_tune_set(int n, int L1, int T, int lr_col, int lr_row);
/* Engine for products with N = 2^n (n >= 2): 0 = Goldilocks over 2^n
 * points (16-bit digits, the reference's field), 1 = three primes below 2^31
 * over 2^(n-1) points (32-bit digits, CRT); the default is 1, for the
 * one-workgroup sizes (N <= 2^13) as for the four-step ones. engine = -1 only
 * reads. Switching resets that n to the engine's default shape, after which
 * pdh_tune_get/set use that engine's split (L1 + L2 = n - engine). Every
 * engine and shape gives the same product bits. */
int pdh_tune_engine(int n, int engine, int* current);
/* Time every candidate shape for each n in [n_lo, n_hi] on `stream` (reps
 * timed runs each, on seeded operands in scratch buffers) and keep the
 * fastest. Allocates and frees its scratch; the work is synchronous. */
int pdh_autotune(int n_lo, int n_hi, int reps, void* stream);
/* Write the shape of every transform size (defaults, set and tuned ones) to
 * a text file, so that a later process can skip the autotuning; read such a
 * file back. A file is checked whole before anything changes: an unreadable
 * file, another format or an unlaunchable shape gives PDH_E_ARG and leaves
 * every shape as it was. Sizes the file does not name keep their shapes.
 * Neither needs a device. No counterpart in the reference. */
int pdh_tune_save(const char* path);
int pdh_tune_load(const char* path);

/* Hex digit values (0..15) at 1-based positions start .. start+count-1, host
 * buffer. Position 1 is the first hex digit after the point (value 2).
 * Replaces: std::string bbp_hex_digits_cuda(uint64_t start, uint32_t count)
 * (reference include/bbp.h:34, src/bbp_cuda.cu:40-70); the caller maps
 * values to '0'-'9','A'-'F' as src/bbp_cpu.cpp:20-22 does. A range of 16
 * or more positions goes by digit blocks: one exact 64-bit fixed-point
 * evaluation at every G-th position (G = 7 up to position 65,536, 5 at
 * 10^8) decides the digits of the G positions from it on wherever the
 * accumulated error bound keeps them clear of a digit boundary; the rest
 * take the per-position path of pdh_bbp_positions_dev. The digits equal the
 * reference's position for position. */
int pdh_bbp_range(uint64_t start, uint32_t count, uint8_t* values);

/* Device-output variant of pdh_bbp_range (same digit blocks): d_values is a
 * device buffer of `count` bytes, the work is enqueued on `stream`, which
 * the call synchronises. Replaces the same bbp_hex_digits_cuda
 * (include/bbp.h:34) for callers that keep the digits in device memory. */
int pdh_bbp_range_dev(uint64_t start, uint32_t count, uint8_t* d_values, void* stream);

/* Digit values at arbitrary 1-based positions (host buffers). New: the
 * reference only offers contiguous blocks (bbp.h:27-34) or one position
 * (char bbp_hex_digit(uint64_t), bbp.h:27, src/bbp_cpu.cpp:24-26). */
int pdh_bbp_positions(const uint64_t* positions, size_t n, uint8_t* values);

/* Device-resident variant: d_positions and d_values are device pointers;
 * enqueued on `stream`. Every term is computed in parallel: first as exact
 * 64-bit fixed-point fractions, which decide the digit whenever it lies
 * provably far from a digit boundary; the rare rest (about 4e-6 per position
 * at 10^8) is recomputed with the reference's FP64 fold order. The call
 * synchronises `stream` twice: after a scan of the positions (the largest,
 * and how close each group of 64 lies, which picks the term kernel) and to
 * find the undecided positions. Up to position 2^27 - 1 more than 64
 * positions are sorted on the device first (rocPRIM radix sort) and the
 * digits scattered back to the caller's order, since the term kernel's cost
 * falls with the gaps between the sorted positions of each group of 64; any
 * order gives the same digits at the same cost. Positions must lie in
 * [1, 2^45] (every modulus 8k + j below 2^48, the range where the
 * reference's bbp_mulmod is exact, include/bbp_kernel.h:63-83); the host
 * entry points reject others with PDH_E_ARG before any device work, and this
 * one writes 0xFF for them and returns PDH_E_ARG. */
int pdh_bbp_positions_dev(const uint64_t* d_positions, size_t n, uint8_t* d_values,
                          void* stream);

/* The bottom of the Chudnovsky binary-splitting tree on the device (the
 * reference computes leaves and merges on the host, src/chudnovsky.cpp:
 * 64-102, leaves :70-83): for every group g of 2^group_log consecutive
 * terms of [0, n_terms), P, Q and T of the group, written to device buffers
 * at g * 3 * 2^group_log, g * 5 * 2^group_log and g * 6 * 2^group_log limbs
 * (normalized little-endian magnitudes), their lengths to d_len[3g..3g+2]
 * and the sign of T (-1, 0, 1) to d_sign[g]. The values do not depend on
 * how a range is split, so they are the reference's P(a,b), Q(a,b), T(a,b).
 * group_log in [1, 6], n_terms <= 2^28. Enqueued on `stream`. */
int pdh_bs_groups(uint64_t n_terms, int group_log, uint32_t* d_P, uint32_t* d_Q, uint32_t* d_T,
                  uint64_t* d_len, int32_t* d_sign, void* stream);

/* One binary-splitting node: device magnitudes P, Q, |T| (little-endian,
 * normalized, lengths in limbs on the host) and the sign of T. */
typedef struct pdh_bs_node {
    uint32_t* P; uint64_t nP;
    uint32_t* Q; uint64_t nQ;
    uint32_t* T; uint64_t nT;
    int32_t sT;
} pdh_bs_node;

/* One level of the reference's merge (src/chudnovsky.cpp:93-100,
 *     P = P_L P_R,  Q = Q_L Q_R,  T = T_L Q_R + P_L T_R)
 * for the pairs (in[2i], in[2i+1]), entirely in device memory: the products
 * run as one batch and the signed sum of T on the device (two passes over
 * 4,096-limb tiles). out[i].P, .Q, .T are device buffers set by the caller
 * with room for nP_L + nP_R, nQ_L + nQ_R and
 * max(nT_L + nQ_R, nP_L + nT_R) + 1 limbs; they must not overlap the inputs.
 * With n_in odd, out[n_in / 2] receives a copy of the last input node.
 * skip_p != 0 leaves every output P unwritten (nP = 0; the root's P is never
 * used). On return (the stream is synchronized) out[i].nP, .nQ, .nT, .sT
 * hold the results. Every input needs nP > 0 and nQ > 0; nT = 0 means T = 0.
 * Products beyond one transform (more than 2^23 limbs) run blocked. */
int pdh_bs_merge_dev(const pdh_bs_node* in, size_t n_in, pdh_bs_node* out, int skip_p, void* stream);

/* ---- device-resident magnitudes ---------------------------------------------
 * Building blocks of the reference's to_decimal (src/bignum_div.cpp:239-267,
 * repeated division of the whole number by 10^9) done divide-and-conquer on
 * device buffers: quotient estimates by shifts, remainders by signed sums,
 * and the leaves' digits. Lengths are device uint64_t normalized limb counts,
 * so a level's steps need no host reads of the values. */

/* out = sa*a + sb*b for every descriptor (sa, sb in {-1, 0, 1}; 0 means that
 * operand is zero and is not read). out may be a itself, but out_len must
 * not be a_len or b_len (the kernel clears it while other workgroups still
 * read the operands' lengths; PDH_E_ARG if it is). Limbs below out_cap
 * are written and the result must fit in them; *out_len receives its
 * normalized length and *out_sign its sign (0 for zero). Two passes over
 * 4,096-limb tiles of every output, enqueued on `stream`; the call returns
 * without waiting for the device (results in stream order). */
typedef struct pdh_sadd_desc {
    const uint32_t* a; const uint64_t* a_len; int32_t sa;
    const uint32_t* b; const uint64_t* b_len; int32_t sb;
    uint32_t* out; uint64_t out_cap;
    uint64_t* out_len; int32_t* out_sign;
} pdh_sadd_desc;
int pdh_sadd_batch_dev(const pdh_sadd_desc* descs, size_t n, void* stream);

/* out = floor(a / 2^bits) for every descriptor; *a_len is a's normalized
 * length (device). Limbs below out_cap (at least 1) are written; the result
 * must fit. *out_len receives the normalized length. Enqueued on `stream`;
 * returns without waiting for the device (results in stream order). */
typedef struct pdh_shr_desc {
    const uint32_t* a; const uint64_t* a_len;
    uint32_t* out; uint64_t out_cap;
    uint64_t* out_len;
    uint64_t bits;
} pdh_shr_desc;
int pdh_shr_batch_dev(const pdh_shr_desc* descs, size_t n, void* stream);

/* Single operations on one device magnitude each, enqueued on `stream` with
 * no host synchronization (lengths stay on the device), for Newton
 * iterations kept in HBM. pdh_sadd_dev: out = sa*a + sb*b as in
 * pdh_sadd_batch_dev (out may be a; its tile statuses live in the library's
 * scratch, which calls on different streams take in turn, as pdh_mul_dev's
 * workspace). pdh_shift_dev: out = a * 2^bits for bits >= 0,
 * floor(a / 2^-bits) for bits < 0; limbs below out_cap (at least 1) are
 * written and the result must fit; *out_len receives its normalized length.
 * No counterpart in the reference (its BigInt shifts and adds are host
 * code, src/bignum.cpp). */
int pdh_sadd_dev(const uint32_t* a, const uint64_t* a_len, int32_t sa, const uint32_t* b, const uint64_t* b_len,
                 int32_t sb, uint32_t* out, uint64_t out_cap, uint64_t* out_len, int32_t* out_sign, void* stream);
int pdh_shift_dev(const uint32_t* a, const uint64_t* a_len, int64_t bits, uint32_t* out, uint64_t out_cap,
                  uint64_t* out_len, void* stream);

/* Decimal digits of n values below 10^width (value i at d_x + i*stride limbs,
 * normalized length d_len[i]): width ASCII digits each, most significant
 * first and zero-padded, at d_out + i*width. width is a positive multiple of
 * 9. The values are overwritten. Enqueued on `stream`. */
int pdh_decimal_leaves_dev(uint32_t* d_x, uint64_t stride, const uint64_t* d_len, uint64_t n, uint32_t width,
                           char* d_out, void* stream);

/* Test switches, the launch counter, per-kernel timing and a microbenchmark
 * are declared in pdh_diag.h: they are not part of the seam. */

#ifdef __cplusplus
}
#endif
#endif /* PDH_CAPI_H */
