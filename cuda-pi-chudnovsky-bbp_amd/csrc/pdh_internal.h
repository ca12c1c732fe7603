// ============================================================================
//  pdh_internal.h -- declarations shared by the kernel translation units and
//  the C-ABI layer. Not part of the public interface.
// ============================================================================
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <atomic>

namespace pdh {

// How a product with N = 2^n (16-bit digit count, include/ntt.h:47-50) is
// transformed. n <= SMALL_MAX_LOG runs as one workgroup per multiply
// (L1 = 0): engine 0 over 2^n Goldilocks points (L2 = n), engine 1 over
// 2^(n-1) points of the three primes (L2 = n - 1, the default). Above that, engine 0 is the Goldilocks four-step over 2^n
// points (L1 + L2 = n) and engine 1 the three-prime four-step over 32-bit
// digits, 2^(n-1) points (L1 + L2 = n - 1): a column pass over the top L1
// index bits (T columns per workgroup) and a row pass over the low L2 bits.
struct NttShape {
    int n;       // log2 N
    int L1;      // column sub-transform bits (0 for the one-workgroup path)
    int L2;      // row sub-transform bits
    int T;       // columns per column-pass workgroup
    int lr_col;  // log2 of the column pass radix (3 or 4; 3 for engine 1)
    int lr_row;  // log2 of the row pass (and one-workgroup path) radix
    int eng;     // 0 Goldilocks, 1 three primes
};

constexpr int SMALL_MAX_LOG = 13;   // 2 x 2^13 u64 (+1/8 padding) = 144 KB LDS
constexpr int MAX_LOG       = 24;
constexpr int ROW_MAX_LOG   = 13;
constexpr int COL_MAX_LOG   = 11;

NttShape ntt_shape(int n);
bool shape_valid(const NttShape& s);
// false if the shape is not launchable; custom = false marks an engine's
// default shape (batches then keep their own shape rule)
bool set_shape(const NttShape& s, bool custom = true);
bool shape_custom(int n);   // set by pdh_tune_set / pdh_tune_engine / autotune / a loaded file
NttShape default_shape_for(int n, int eng);
// Shape for a launch of njobs independent products of size class n.
NttShape batch_shape(int n, unsigned njobs);

// Device tables for one device: twL[2^L + x] = w_{2^L}^x for L <= 13, and
// the per-n outer tables (w_N^e = hi[e >> h] * lo[e & (2^h-1)], forward; the
// inverse pair carries N^-1 in hi).
struct OuterTables {
    const uint64_t* hi;
    const uint64_t* lo;
    const uint64_t* hi_inv;
    const uint64_t* lo_inv;
    int h;
    // three-prime engine: per prime (stride TW3_STRIDE) the level twiddles
    // w_M^(+-j c) R mod p, j-major per radix, direction and M = 2^logM
    // (ntt3_kernels.hip), and per n the outer tables of all three primes (o3, h3)
    const uint32_t* tw3;
    const uint32_t* o3;
    int h3;
};

// The three primes, their Montgomery constants (R = 2^32) and radix-8
// twiddles, and the CRT constants; see ntt3_kernels.hip.
struct Prime3 {
    uint32_t p, ninv;        // ninv = -p^-1 mod 2^32
    uint32_t w8[4], w8i[4];  // w8^k R mod p and w8^-k R mod p, k = 0..3
    uint32_t w16[3], w16i[3];  // w16^k R mod p and w16^-k R mod p, k = 1, 3, 9
};
struct Primes3 {
    Prime3 p[3];
    uint32_t inv_p0_m1;      // p0^-1 R mod p1
    uint32_t p0_m2;          // p0 R mod p2
    uint32_t inv_p0p1_m2;    // (p0 p1)^-1 R mod p2
    uint64_t p0p1;
};
constexpr size_t TW3_STRIDE = (size_t)8 << 15;    // 4 radices x 2 directions x 2^15
const Primes3& primes3();
// w_{2^L} for prime pi (L <= 24), normal form, and the host's modular helpers.
uint32_t r3_root(int pi, int L);
uint32_t r3_mul(uint32_t a, uint32_t b, uint32_t p);
uint32_t r3_pow(uint32_t a, uint64_t e, uint32_t p);
uint32_t r3_to_mont(uint32_t a, uint32_t p);

// One multiply as the kernels see it (device-visible when batched).
struct MulJob {
    const uint32_t* a;
    const uint32_t* b;
    uint32_t* out;          // na + nb limbs
    uint64_t* out_len;      // normalized length (device)
    uint64_t na, nb;
    uint64_t* wsA;          // 2^n words (product coefficients end here)
    uint64_t* wsB;          // 2^n words (unused by the one-workgroup path)
    void* carry;            // carry scratch (unused by the one-workgroup path)
};

// Launchers (ntt_kernels.hip, ntt3_kernels.hip). All enqueue on `s`, return hipError_t.
hipError_t init_kernel_attrs();
hipError_t init_kernel_attrs3();
hipError_t launch_class3(const MulJob* d_jobs, const MulJob& one, unsigned njobs, const NttShape& sh, bool square,
                         size_t nblk_max, const OuterTables& ot, hipStream_t s);
// One workgroup per product on the three primes (N = 2^n <= 2^13; tw3 = the
// level twiddle table).
hipError_t launch_small3(const MulJob* d_jobs, const MulJob& one, unsigned njobs, int n, bool square,
                         const uint32_t* tw3, hipStream_t s);
// The carry scan over chunks of `kl` limbs (digits and carry-outs already in
// the product buffer and the carry scratch; with `th`, also the chunk
// thresholds, so the reduce reads no digits).
hipError_t launch_carry_scan(int kl, const MulJob* d_jobs, const MulJob& one, unsigned njobs, size_t nblk_max,
                             hipStream_t s, bool th = false);
// Every job in a launch shares n; jobs == nullptr launches the single job `one`.
hipError_t launch_class(const MulJob* d_jobs, const MulJob& one, unsigned njobs, int n, bool square,
                        size_t nblk_max, const uint64_t* twL, const OuterTables& ot, hipStream_t s,
                        const NttShape* shape = nullptr);
size_t carry_nblk(size_t nlimbs);
hipError_t launch_mul(const uint32_t* a, size_t na, const uint32_t* b, size_t nb,
                      bool square, uint32_t* out, uint64_t* out_len,
                      uint64_t* wsA, uint64_t* wsB, uint64_t* carry_ws,
                      const uint64_t* twL, const OuterTables& ot, const NttShape& sh,
                      hipStream_t s);
hipError_t launch_zero_len(uint64_t* out_len, hipStream_t s);

// Bytes of carry scratch launch_mul needs for a product of `nlimbs` limbs.
size_t carry_ws_bytes(size_t nlimbs);

// BBP (bbp_kernels.hip). launch_bbp: one lane per series, for many
// positions. For few positions (workspace bbp_few_ws_bytes, dmax = largest
// position - 1): launch_bbp_fx computes every term in parallel as an exact
// 64-bit fixed-point fraction and writes the digit, or 0xFE where the
// reference's FP64 rounding could decide it differently; launch_bbp_chunked
// computes terms in parallel and folds them per series in the reference's
// FP64 order (odd_undecided, for tests, also marks every odd index
// undecided). With block > 1 and no d_pos, launch_bbp_fx evaluates the
// `count` anchors start + block t and writes the digits of the nout
// positions start .. start + nout - 1 (block of them per anchor, 0xFE where
// undecided); `tk` picks its term kernel: terms in lanes (k_bbp_fx_terms),
// positions in lanes (k_bbp_fx_lanes) or moduli in lanes (k_bbp_fx_moduli,
// for dmax <= BBP_MODULI_DMAX). launch_bbp_group_scan: over groups of 64
// entries of d_pos, d_out3[0] = max(position - 1), d_out3[1] = sum of
// positions, d_out3[2] = sum of 64 x the group's largest position (the
// caller zeroes the three words).
hipError_t launch_bbp(const uint64_t* d_pos, uint64_t start, size_t count, uint8_t* d_out, uint64_t* d_next,
                      hipStream_t s);
hipError_t launch_bbp_fx(const uint64_t* d_pos, uint64_t start, size_t count, uint64_t dmax,
                         uint8_t* d_out, void* ws, int odd_undecided, hipStream_t s, uint32_t block = 1,
                         size_t nout = 0, int tk = 0);
enum : int { BBP_TK_TERMS = 0, BBP_TK_LANES = 1, BBP_TK_MODULI = 2 };
constexpr uint64_t BBP_MODULI_DMAX_HOST = ((uint64_t)1 << 27) - 2;   // largest d the moduli kernel takes
hipError_t launch_bbp_chunked(const uint64_t* d_pos, uint64_t start, size_t count, uint64_t dmax,
                              uint8_t* d_out, void* ws, hipStream_t s);
hipError_t launch_bbp_group_scan(const uint64_t* d_pos, size_t count, uint64_t* d_out3, hipStream_t s);
// Device sort of positions (by their low 27 bits) with their indices, and
// the scatter of values computed in that order back to the caller's order
// (out[perm[i]] = in[i]). bbp_sort_positions with temp null only sets
// *temp_bytes.
hipError_t bbp_sort_positions(const uint64_t* d_pos, size_t n, uint64_t* d_sorted, uint32_t* d_perm, void* temp,
                              size_t* temp_bytes, hipStream_t s);
hipError_t launch_bbp_scatter(const uint8_t* d_in, const uint32_t* d_perm, size_t n, uint8_t* d_out, hipStream_t s);
// Measured issue rate of v_mad_u64_u32 (lane operations per second, whole GPU).
hipError_t probe_mad_rate(double* lane_ops_per_s);
size_t bbp_few_ws_bytes(size_t count);

// Chudnovsky binary splitting (bs_kernels.hip): P, Q, T of every group of
// 2^G consecutive terms of [0, n_terms) into slots of 3, 5, 6 x 2^G limbs.
hipError_t launch_bs_groups(uint64_t n_terms, int G, uint32_t* dP, uint32_t* dQ, uint32_t* dT, uint64_t* dlen,
                            int32_t* dsgn, hipStream_t s);

// Signed sums (bs_kernels.hip): for every node, T = sa A + sb B with A
// (device length *la) and B (length *lb); T may be A's own buffer; sa, sb in
// {-1, 0, 1} (0: the operand is zero). Nodes are cut into BS_ADD_TILE-limb
// tiles numbered tile0 .. tile0 + ntiles - 1 in one flat list (d_tile_node
// maps a tile to its node); T is written below cap limbs (the sum must fit),
// *len receives T's normalized length and *sign its sign (meaningless when
// the length is 0). d_status: one byte per tile.
constexpr uint32_t BS_ADD_TILE = 4096;
struct BsAddNode {
    uint32_t* T;
    const uint32_t* A;
    const uint32_t* B;
    const uint64_t* la;
    const uint64_t* lb;
    uint64_t cap;
    uint64_t* len;
    int32_t* sign;
    uint32_t tile0, ntiles;
    int32_t sa, sb;
};
// one != nullptr: a single node passed by value (d_nodes, d_tile_node unused).
hipError_t launch_bs_signed_add(const BsAddNode* d_nodes, const uint32_t* d_tile_node, uint32_t ntiles,
                                uint8_t* d_status, hipStream_t s, const BsAddNode* one = nullptr);

// Decimal conversion helpers (dec_kernels.hip). Shifts: out = a >> bits (or
// a << bits when left) for every node, over a flat list of SHR_TILE-limb
// tiles of the outputs (limbs below out_cap written; *out_len normalized from
// *a_len, which must be normalized). Leaves: for value i at x + i stride
// (normalized length x_len[i], below 10^width, width a multiple of 9), width
// ASCII digits at out + i width, most significant first; the values are
// overwritten.
constexpr uint32_t SHR_TILE = 2048;
struct ShiftNode {
    const uint32_t* a;
    const uint64_t* a_len;
    uint32_t* out;
    uint64_t* out_len;
    uint64_t out_cap;
    uint64_t bits;
    uint32_t tile0, ntiles;
    uint32_t left;
};
hipError_t launch_shift_batch(const ShiftNode* d_nodes, const uint32_t* d_tile_node, uint32_t ntiles, hipStream_t s,
                              const ShiftNode* one = nullptr);
hipError_t launch_fill_zero(uint32_t* p, uint64_t n, hipStream_t s);
hipError_t launch_decimal_leaves(uint32_t* x, uint64_t stride, const uint64_t* x_len, uint64_t n, uint32_t width,
                                 char* out, hipStream_t s);

extern std::atomic<uint64_t> g_launches;

// Optional per-kernel timing (pdh_timing_enable): when on, each launch is
// bracketed by hipEvents on its own stream and the durations are summed by
// kernel name when read. timing_begin returns nullptr when timing is off.
hipEvent_t timing_begin(hipStream_t s);
void timing_end(const char* name, hipEvent_t ev0, hipStream_t s);

} // namespace pdh
