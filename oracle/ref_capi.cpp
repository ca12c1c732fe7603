// ============================================================================
//  oracle/ref_capi.cpp -- TEST INFRASTRUCTURE ONLY.
//
//  A C-ABI wrapper, written for this repository, around the reference's own
//  CPU functions. oracle/Makefile compiles it together with the reference's
//  src/ntt_cpu.cpp and src/bbp_cpu.cpp, read in place from the reference
//  checkout, into oracle/_ref/libref.so. Nothing of the reference is copied
//  here; only its public declarations (bignum.h:194-195, bbp.h:27-30) and
//  the inline bbp_modpow16 of bbp_kernel.h are included and called.
//
//  Used to pin the restatement in cpu_ref.cpp and as the "reference" CPU
//  baseline in bench.py when the library has been built.
// ============================================================================
#include "bignum.h"
#include "bbp.h"
#include "bbp_kernel.h"

#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <atomic>
#include <vector>

namespace {
template <class F>
void parallel_for(size_t n, int threads, F&& f) {
    if (threads <= 1 || n < 2) { for (size_t i = 0; i < n; ++i) f(i); return; }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&]() { for (size_t i; (i = next.fetch_add(1)) < n;) f(i); });
    for (auto& th : pool) th.join();
}
}

extern "C" {

int ref_mul_ntt_cpu(const uint32_t* a, size_t na, const uint32_t* b, size_t nb,
                    uint32_t* out, size_t cap, size_t* len) {
    std::vector<uint32_t> va(a, a + na), vb(b, b + nb);
    std::vector<uint32_t> r = pidigits::mul_ntt_cpu(va, vb);
    *len = r.size();
    if (r.size() > cap) return -1;
    if (!r.empty()) std::memcpy(out, r.data(), r.size() * sizeof(uint32_t));
    return 0;
}

int ref_mul_ntt_cpu_batch(const uint32_t* a, const uint64_t* a_off, const uint64_t* na,
                          const uint32_t* b, const uint64_t* b_off, const uint64_t* nb,
                          uint32_t* out, const uint64_t* out_off, uint64_t* out_len,
                          size_t count, int threads) {
    std::atomic<int> bad{0};
    parallel_for(count, threads, [&](size_t i) {
        std::vector<uint32_t> va(a + a_off[i], a + a_off[i] + na[i]);
        std::vector<uint32_t> vb(b + b_off[i], b + b_off[i] + nb[i]);
        std::vector<uint32_t> r = pidigits::mul_ntt_cpu(va, vb);
        out_len[i] = r.size();
        if (r.size() > na[i] + nb[i]) { bad = 1; return; }
        if (!r.empty()) std::memcpy(out + out_off[i], r.data(), r.size() * sizeof(uint32_t));
    });
    return bad ? -1 : 0;
}

// Uppercase hex for [start, start+count), split into per-thread blocks of
// bbp_hex_digits_cpu calls.
void ref_bbp_hex_range(uint64_t start, uint32_t count, char* out, int threads) {
    const uint32_t block = 64;
    size_t nblocks = (count + block - 1) / block;
    parallel_for(nblocks, threads, [&](size_t bi) {
        uint32_t off = (uint32_t)(bi * block);
        uint32_t c = count - off < block ? count - off : block;
        std::string s = pidigits::bbp_hex_digits_cpu(start + off, c);
        std::memcpy(out + off, s.data(), c);
    });
}

void ref_bbp_positions(const uint64_t* pos, size_t n, char* out, int threads) {
    parallel_for(n, threads, [&](size_t i) { out[i] = pidigits::bbp_hex_digit(pos[i]); });
}

// The reference's 16^e mod m (include/bbp_kernel.h:86-96, with its 128-bit
// bbp_mulmod branch for m >= 2^32, :70-83).
uint64_t ref_bbp_modpow16(uint64_t e, uint64_t m) { return pidigits::bbp_modpow16(e, m); }

} // extern "C"
