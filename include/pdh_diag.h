/* ===========================================================================
 *  pdh_diag.h -- test switches and diagnostics of the ntt-hip libraries.
 *
 *  None of these is part of the multiply or BBP seam (pdh_capi.h, pdh_pi.h)
 *  and none has a counterpart in the reference: they let tests force rare
 *  code paths, prove that device code ran, and let benchmarks time kernels.
 *  Same status codes as pdh_capi.h.
 * ========================================================================= */
#ifndef PDH_DIAG_H
#define PDH_DIAG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- exported by libpdh.so ------------------------------------------------ */

/* pdh_bbp_set_exact(mode): 0 (default) the fixed-point path with the
 * FP64-order fallback; 1 the FP64-order (chunked) path for every position;
 * 2 the fixed-point path with every odd index (in sorted order) sent to the
 * fallback as well, to exercise it; 3 one lane per series in the reference's
 * FP64 order for every position (wavefronts take groups of positions from a
 * work-queue counter in the library's context, so such calls on different
 * streams must not overlap on the device). pdh_bbp_exact_count is the
 * number of positions computed by the chunked FP64-order path in this
 * process. */
int pdh_bbp_set_exact(int mode);
uint64_t pdh_bbp_exact_count(void);
/* pdh_bbp_set_layout(mode): the fixed-point path's term kernel. 0 (default)
 * chooses per call: moduli in lanes (64 terms k of one series per
 * wavefront, stepping through each group of 64 positions in sorted order)
 * for more than 64 positions when the largest is at most 2^27 - 1; else
 * positions in lanes
 * (one exponent per wavefront, its bits scalar) when at least three lanes
 * in four stay busy -- the positions of each group of 64, in the order
 * given, lie close together -- else terms in lanes. 1 forces terms in
 * lanes, 2 positions in lanes, 3 moduli in lanes (where it applies, else as
 * 0). The digits are the same either way. */
int pdh_bbp_set_layout(int mode);
/* The device's measured v_mad_u64_u32 issue rate (lane operations per
 * second, whole GPU), the integer-multiply ceiling BBP throughput is quoted
 * against. Synchronous; allocates and frees its own buffer. */
int pdh_probe_mad_rate(double* lane_ops_per_s);

/* Number of kernels this library has launched in this process. Lets a test
 * assert that the device path really ran (the reference's self-test could
 * not tell its CPU fallback from the GPU, tests/self_test.cu:66). */
uint64_t pdh_launch_count(void);

/* Per-kernel device timing for benchmarks and profiling (no counterpart in
 * the reference, which times whole calls with std::chrono,
 * src/chudnovsky.cpp:108-157). pdh_timing_enable(1) clears the record and
 * brackets every later launch with hipEvents on its stream; 0 stops.
 * pdh_timing_read waits for the recorded events, writes up to max_kernels
 * kernel names ('\n'-separated, NUL-terminated) with their summed durations
 * in ms and launch counts, and clears the record. */
int pdh_timing_enable(int on);
int pdh_timing_read(char* names, size_t names_cap, double* total_ms, uint64_t* counts,
                    int max_kernels, int* n_kernels);


/* ---- exported by libpidigits_hip.so ------------------------------------------ */

/* Test switch (no counterpart in the reference): k != 0 (|k| <= 8) moves
 * every Newton quotient and square-root estimate by k units, and every
 * decimal-conversion quotient estimate by k (k > 0) or by -1 where it is at
 * least 1 (k < 0), before the exact correction, so that the correction
 * branches run; results are unchanged. 0 (default) turns it off. */
int pdh_newton_set_perturb(int k);
/* Correction steps the exact floor corrections have taken in this process
 * (quotient, square-root and decimal-level steps together). */
uint64_t pdh_newton_corrections(void);

#ifdef __cplusplus
}
#endif
#endif /* PDH_DIAG_H */
