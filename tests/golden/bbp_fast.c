/* bbp_fast.c -- BBP hex digits at arbitrary positions on the host, fast
 * enough to write a 4,096-position fixture over [10^6, 10^8] in minutes.
 * Fixture tooling only (make_golden.py builds and loads it).
 *
 * Follows the reference's digit computation (include/bbp_kernel.h:86-137)
 * term for term: for each series j in {1, 4, 5, 6} and k = 0..d,
 *     sum += (double)r / (double)m;  sum -= floor(sum);   r = 16^(d-k) mod m,
 * then the same 1e-18 tail, the same 4 S1 - 2 S4 - S5 - S6 combination and
 * the same final floor and scaling, all in IEEE double in the same order.
 * Only the exact integers r are computed differently: r is unique, so any
 * exact method gives the reference's bits.
 *
 *   m = 1 gives 0 (checked first, as bbp_modpow16 does), e = 0 gives 1;
 *   otherwise m = 2^s mo with mo odd, and for 4e >= s (always, s <= 2)
 *   16^e mod m = 2^s (2^(4e - s) mod mo), with 2^E mod mo by a left-to-right
 *   square-and-double ladder in Montgomery form (R = 2^32, mo < 2^30),
 *   eight terms per AVX-512 vector and four vectors interleaved; moduli from
 *   2^30 up to 2^48 (positions up to 2^45) take the same ladder with R = 2^64
 *   one term at a time.
 *
 * Checked against the reference's own bbp_cpu (compiled from source) on
 * positions 1..65,536 and on the committed sampled positions before use. */
#include <immintrin.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define BLK 4096   /* terms per block and series, a multiple of 32 */

static uint32_t neg_inv32(uint32_t m) { /* -m^-1 mod 2^32, m odd */
    uint32_t x = m;                        /* m * m = 1 mod 8: 3 bits */
    for (int i = 0; i < 5; ++i) x *= 2u - m * x;
    return 0u - x;
}

/* r[i] = 2^E[i] mod mo[i] for n lanes (n a multiple of 32), mo odd in [3, 2^30).
 * The ladder keeps x in [0, 2 mo): for x < 2 mo, REDC(x^2) < (4 mo^2 + 2^32 mo)
 * / 2^32 < 2 mo since 4 mo < 2^32, and a doubling is reduced by 2 mo. */
static void pow2_mod(const uint64_t* E, const uint64_t* mo, const uint64_t* mi, const uint64_t* one,
                     uint64_t* r, int n) {
    for (int base = 0; base < n; base += 32) {
        __m512i x[4], M[4], M2[4], MI[4], e[4];
        uint64_t emax = 0;
        for (int i = 0; i < 32; ++i) emax |= E[base + i];
        for (int q = 0; q < 4; ++q) {
            M[q] = _mm512_loadu_si512((const void*)(mo + base + 8 * q));
            M2[q] = _mm512_add_epi64(M[q], M[q]);
            MI[q] = _mm512_loadu_si512((const void*)(mi + base + 8 * q));
            e[q] = _mm512_loadu_si512((const void*)(E + base + 8 * q));
            x[q] = _mm512_loadu_si512((const void*)(one + base + 8 * q));
        }
        for (int b = emax ? 63 - __builtin_clzll(emax) : -1; b >= 0; --b) {
            const __m512i bit = _mm512_set1_epi64((long long)(1ull << b));
            for (int q = 0; q < 4; ++q) {
                const __m512i t = _mm512_mul_epu32(x[q], x[q]);
                const __m512i u = _mm512_mul_epu32(t, MI[q]);
                __m512i s = _mm512_srli_epi64(_mm512_add_epi64(t, _mm512_mul_epu32(u, M[q])), 32);
                s = _mm512_mask_add_epi64(s, _mm512_test_epi64_mask(e[q], bit), s, s);
                x[q] = _mm512_min_epu64(s, _mm512_sub_epi64(s, M2[q]));
            }
        }
        for (int q = 0; q < 4; ++q) {       /* out of Montgomery form, into [0, mo) */
            const __m512i u = _mm512_mul_epu32(x[q], MI[q]);
            __m512i s = _mm512_srli_epi64(_mm512_add_epi64(x[q], _mm512_mul_epu32(u, M[q])), 32);
            s = _mm512_min_epu64(s, _mm512_sub_epi64(s, M[q]));
            _mm512_storeu_si512((void*)(r + base + 8 * q), s);
        }
    }
}

/* 2^E mod mo for odd mo in [2^30, 2^48) (moduli of positions above 2^27),
 * one term at a time: the same square-and-double ladder in Montgomery form
 * with R = 2^64 on 128-bit products. For x < 2 mo, REDC(x^2) < (4 mo^2 +
 * 2^64 mo) / 2^64 < 2 mo; a doubling and two conditional subtractions keep
 * x in [0, mo) between steps. */
static uint64_t pow2_mod64(uint64_t E, uint64_t mo) {
    uint64_t x = mo;                            /* mo^-1 mod 2^64: 3 -> 96 bits */
    for (int i = 0; i < 5; ++i) x *= 2u - mo * x;
    const uint64_t ninv = 0u - x;
    const uint64_t one = (uint64_t)(((unsigned __int128)1 << 64) % mo);
    uint64_t r = one;
    for (int b = E ? 63 - __builtin_clzll(E) : -1; b >= 0; --b) {
        const unsigned __int128 t = (unsigned __int128)r * r;
        const uint64_t m = (uint64_t)t * ninv;
        r = (uint64_t)((t + (unsigned __int128)m * mo) >> 64);
        if ((E >> b) & 1) r <<= 1;
        if (r >= mo) r -= mo;
        if (r >= mo) r -= mo;
    }
    const uint64_t m = r * ninv;
    r = (uint64_t)(((unsigned __int128)r + (unsigned __int128)m * mo) >> 64);
    return r >= mo ? r - mo : r;
}

typedef struct {
    uint64_t E[BLK], mo[BLK], mi[BLK], one[BLK], r[BLK];
    uint8_t big[BLK];
    double q[4][BLK];
} Work;

/* 16^(d-k) mod m for k = 0 and k = d, the only terms the ladder does not
 * cover (m = 1, e = 0, or m a power of two), exactly as bbp_modpow16. */
static uint64_t edge_term(uint64_t d, uint64_t k, uint64_t m) {
    if (m == 1) return 0;
    if (d == k) return 1;
    return 0;                                   /* m = 4, e >= 1: 16^e = 0 mod 4 */
}

/* q[jj][i] = (double)(16^(d-k) mod m) / (double)m, k = k0 + i, m = 8k + j.
 * With 8k + j = 2^s mo (s = 0, 2, 0, 1 for j = 1, 4, 5, 6), the ladder gives
 * 2^(4e - s) mod mo, which is shifted back by s. */
static void block_terms(Work* w, uint64_t d, uint64_t k0, int len) {
    static const uint32_t J[4] = {1, 4, 5, 6};
    static const int S[4] = {0, 2, 0, 1};
    const int n = (len + 31) & ~31;
    for (int jj = 0; jj < 4; ++jj) {
        const uint64_t j = J[jj];
        const int s = S[jj];
        for (int i = 0; i < n; ++i) {           /* branch-free; padding lanes get e = 0 */
            const uint64_t k = k0 + (uint64_t)i;
            const uint64_t e = k <= d ? d - k : 0;
            uint64_t mo = (8 * k + j) >> s;
            w->big[i] = mo >= (1u << 30);
            if (w->big[i]) {                         /* the 64-bit ladder below; a dummy lane here */
                w->r[i] = e ? pow2_mod64(4 * e - (uint64_t)s, mo) : 1;
                mo = 3;
            }
            mo = mo < 3 ? 3 : mo;
            uint32_t x = (uint32_t)mo;              /* -mo^-1 mod 2^32 by Newton */
            for (int it = 0; it < 5; ++it) x *= 2u - (uint32_t)mo * x;
            const uint64_t q = (uint64_t)(4294967296.0 / (double)mo);
            int64_t one = (int64_t)((1ull << 32) - q * mo);
            one += one < 0 ? (int64_t)mo : 0;
            one -= one >= (int64_t)mo ? (int64_t)mo : 0;
            w->E[i] = e ? 4 * e - (uint64_t)s : 0;
            w->mo[i] = mo;
            w->mi[i] = (uint32_t)(0u - x);
            w->one[i] = (uint64_t)one;
        }
        uint64_t keep[BLK];
        for (int i = 0; i < n; ++i) keep[i] = w->r[i];
        pow2_mod(w->E, w->mo, w->mi, w->one, w->r, n);
        for (int i = 0; i < n; ++i) if (w->big[i]) w->r[i] = keep[i];
        for (int i = 0; i < len; ++i) {
            const uint64_t m = 8 * (k0 + (uint64_t)i) + j;
            w->q[jj][i] = (double)(w->r[i] << s) / (double)m;
        }
        if (k0 == 0 && (j >> s) == 1) w->q[jj][0] = (double)edge_term(d, 0, j) / (double)j;   /* m = 1, 4 */
        if (d - k0 < (uint64_t)len) {
            const uint64_t m = 8 * d + j;
            w->q[jj][d - k0] = (double)edge_term(d, d, m) / (double)m;
        }
    }
}

static double tail(uint32_t j, uint64_t d, double sum) {   /* bbp_kernel.h:112-121 */
    double p16 = 1.0 / 16.0;
    uint64_t k = d + 1;
    while (p16 > 1e-18) {
        const uint64_t m = 8ull * k + j;
        sum += p16 / (double)m;
        p16 /= 16.0;
        ++k;
    }
    sum -= floor(sum);
    return sum;
}

static int digit_value(uint64_t pos, Work* w) {
    const uint64_t d = pos - 1;
    double s1 = 0.0, s4 = 0.0, s5 = 0.0, s6 = 0.0;
    for (uint64_t k0 = 0; k0 <= d; k0 += BLK) {
        const int len = (int)((d + 1 - k0) < BLK ? (d + 1 - k0) : BLK);
        block_terms(w, d, k0, len);
        for (int i = 0; i < len; ++i) {     /* four independent chains, each in the reference's order */
            s1 += w->q[0][i]; s1 -= floor(s1);
            s4 += w->q[1][i]; s4 -= floor(s4);
            s5 += w->q[2][i]; s5 -= floor(s5);
            s6 += w->q[3][i]; s6 -= floor(s6);
        }
    }
    s1 = tail(1, d, s1); s4 = tail(4, d, s4); s5 = tail(5, d, s5); s6 = tail(6, d, s6);
    double s = 4.0 * s1 - 2.0 * s4 - 1.0 * s5 - 1.0 * s6;   /* bbp_kernel.h:128-137 */
    s -= floor(s);
    if (s < 0.0) s += 1.0;
    return (int)(16.0 * s);
}

typedef struct {
    const uint64_t* pos;
    uint8_t* vals;
    size_t n;
    size_t next;
    pthread_mutex_t mu;
} Job;

static void* worker(void* arg) {
    Job* job = (Job*)arg;
    Work* w = (Work*)aligned_alloc(64, sizeof(Work));
    if (!w) return NULL;
    for (;;) {
        pthread_mutex_lock(&job->mu);
        const size_t i = job->next++;
        pthread_mutex_unlock(&job->mu);
        if (i >= job->n) break;
        job->vals[i] = (uint8_t)digit_value(job->pos[i], w);
    }
    free(w);
    return NULL;
}

/* Digit values (0..15) at 1-based positions, positions <= 2^45 (m < 2^48). */
int bbp_fast_positions(const uint64_t* pos, size_t n, uint8_t* vals, int threads) {
    for (size_t i = 0; i < n; ++i)
        if (pos[i] == 0 || pos[i] > (1ull << 45)) return -1;
    if (threads < 1) threads = 1;
    Job job = {pos, vals, n, 0, PTHREAD_MUTEX_INITIALIZER};
    memset(vals, 0xFF, n);
    pthread_t th[64];
    if (threads > 64) threads = 64;
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, worker, &job);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    for (size_t i = 0; i < n; ++i)
        if (vals[i] > 15) return -2;
    return 0;
}

/* 16^e mod m by the same route, for the unit check against brute force. */
uint64_t bbp_fast_modpow16(uint64_t e, uint64_t m) {
    if (m == 1) return 0;
    if (e == 0) return 1;
    const int s = __builtin_ctzll(m);
    const uint64_t mo = m >> s;
    if (mo == 1) return 0;
    if (mo >= (1u << 30)) return pow2_mod64(4 * e - (uint64_t)s, mo) << s;
    uint64_t E[32], MO[32], MI[32], ONE[32], R[32];
    for (int i = 0; i < 32; ++i) {
        E[i] = 4 * e - (uint64_t)s; MO[i] = mo; MI[i] = neg_inv32((uint32_t)mo); ONE[i] = (1ull << 32) % mo;
    }
    pow2_mod(E, MO, MI, ONE, R, 32);
    return R[0] << s;
}
