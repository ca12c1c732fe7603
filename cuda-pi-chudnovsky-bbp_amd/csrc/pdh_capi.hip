// ============================================================================
//  pdh_capi.hip -- the C ABI (include/pdh_capi.h): device selection, twiddle
//  tables, workspace, host<->device staging and error reporting.
//
//  Memory: the library keeps one grow-only workspace per process on the
//  current device (two N-point u64 spectra, the carry scratch, and staging
//  for the host-buffer entry points) plus the twiddle tables, so repeated
//  calls allocate nothing once the largest size has been seen. The
//  reference allocates and frees 24N bytes on every call
//  (src/ntt_cuda.cu:198-201,237).
//
//  Calls are serialised by a mutex while they enqueue. Device-resident calls
//  on different streams share the one workspace: the stream of the latest
//  call that used a piece of it records an event, and a call on another
//  stream makes its stream wait for that event before its own launches
//  (ScratchOrder), so such calls may be issued from any thread on any
//  stream. The host entry points synchronise before returning. BBP calls
//  take their scratch per call from the block cache and hold the mutex only
//  while they read the configuration, so they overlap with everything else.
// ============================================================================
#include "pdh_capi.h"
#include "pdh_diag.h"
#include "gl_field.h"
#include "pdh_internal.h"

#include <mutex>
#include <string>
#include <vector>
#include <map>
#include <numeric>
#include <algorithm>
#include <cstring>
#include <cstdio>

namespace pdh {
namespace {

thread_local std::string t_err;

int fail(int code, const std::string& msg) { t_err = msg; return code; }
int fail_hip(hipError_t e, const char* what) {
    return fail(PDH_E_HIP, std::string("[hip] ") + what + " failed: " + hipGetErrorString(e));
}

#define HIPCALL(expr) do { hipError_t _e = (expr); if (_e != hipSuccess) return fail_hip(_e, #expr); } while (0)

struct Outer { u64* hi = nullptr; u64* lo = nullptr; u64* hi_inv = nullptr; u64* lo_inv = nullptr; int h = 0;
               u32* o3 = nullptr; int h3 = 0; };

// Reuse of one piece of the context's scratch by calls on different streams.
// The stream of the latest call that used it records `ev` after its
// launches; a call on another stream first makes its stream wait for that
// event (no host synchronisation). Calls on one stream are ordered by the
// stream itself. Streams under graph capture neither wait nor record: a
// capturing caller orders its own work.
struct ScratchOrder {
    hipEvent_t ev = nullptr;
    hipStream_t last = nullptr;
    bool used = false;
};

bool capturing(hipStream_t s) {
    if (!s) return false;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) { (void)hipGetLastError(); return false; }
    return st != hipStreamCaptureStatusNone;
}

int scratch_acquire(ScratchOrder& o, hipStream_t s) {
    if (o.used && o.last != s && !capturing(s)) HIPCALL(hipStreamWaitEvent(s, o.ev, 0));
    return PDH_OK;
}

int scratch_release(ScratchOrder& o, hipStream_t s) {
    if (capturing(s)) return PDH_OK;
    if (!o.ev) HIPCALL(hipEventCreateWithFlags(&o.ev, hipEventDisableTiming));
    HIPCALL(hipEventRecord(o.ev, s));
    o.last = s;
    o.used = true;
    return PDH_OK;
}

// The host waits until the latest user of the scratch is done (before a
// synchronous host-to-device copy into it).
int scratch_wait_host(ScratchOrder& o) {
    if (o.used) HIPCALL(hipEventSynchronize(o.ev));
    return PDH_OK;
}

// ---- scratch of the allocation-free multiply (pdh_mul_dev_noalloc) -----------
// Two spectra of up to 2^21 u64 words (either engine) and the carry scratch
// of a product of up to 2^20 limbs, in a file-scope device array: every GPU
// the library runs on holds it until the process ends. Host code never
// passes `scratch` itself to a kernel or a copy (on the host the name stands
// for a host-side shadow, not the device address): set-up resolves the
// device address once with hipGetSymbolAddress and keeps it in the context.
constexpr int NOALLOC_MAX_LOG = 21;
constexpr size_t NOALLOC_WS_BYTES = (size_t)8 << NOALLOC_MAX_LOG;
constexpr size_t NOALLOC_CARRY_BYTES = ((size_t)8 << 20) + ((size_t)64 << 10);
constexpr size_t NOALLOC_SCRATCH_BYTES = 2 * NOALLOC_WS_BYTES + NOALLOC_CARRY_BYTES;
static_assert(PDH_NOALLOC_MAX_LIMBS * 2 <= ((size_t)1 << NOALLOC_MAX_LOG), "limit and transform size disagree");
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

struct Ctx {
    std::mutex mu;
    int state = 0;          // 0 unknown, 1 ready, -1 no device
    int device = -1;        // the device current at set-up; every call must run with it current
    std::string why;
    u64* twL = nullptr;     // 2^14 entries: twL[2^L + x] = w_{2^L}^x
    u32* tw3 = nullptr;     // three primes x 2^15 entries: w_{2^L}^x R mod p
    u64* bbp_next = nullptr;  // work-queue counter of k_bbp
    std::map<int, Outer> outer;
    u64* wsA = nullptr; u64* wsB = nullptr; size_t ws_words = 0;
    u64* carry = nullptr; size_t carry_bytes = 0;
    ScratchOrder ws_order;           // wsA, wsB, carry
    int bbp_layout = 0;              // pdh_bbp_set_layout mode
    int bbp_exact = 0;               // pdh_bbp_set_exact mode
    std::atomic<uint64_t> bbp_exact_count{0};   // positions computed by the ordered path
    uint8_t* sadd_status = nullptr; size_t sadd_status_n = 0;   // tile statuses of pdh_sadd_dev
    ScratchOrder sadd_order;
    char* blk_scratch = nullptr; size_t blk_scratch_n = 0;      // block product and lengths of blocked multiplies
    ScratchOrder blk_order;
    std::vector<hipStream_t> sh_streams;   // internal streams of one-shot batches
    std::vector<hipEvent_t> sh_events;
    char* noalloc_ptr = nullptr;           // device address of `scratch` on `device`
    std::string noalloc_why;               // why it is unavailable, if it is
    ScratchOrder noalloc_order;
};

Ctx& ctx() { static Ctx c; return c; }

// Device blocks for scratch, batch workspaces and pdh_dev_alloc are kept for
// reuse instead of being returned with hipFree, which waits for the whole
// device and costs tens of microseconds per call. A block re-enters the cache
// only when no work can still use it: internal users free it after
// synchronizing the stream that used it; pdh_dev_free waits for the device,
// as hipFree would; pdh_dev_free_async records an event on the stream of the
// block's last use and the block turns idle once that event has completed
// (checked without waiting by every allocation, waited for only when the
// device runs out of memory). An allocation takes the smallest cached block
// of at least the requested size and at most twice it; at most kCacheMax
// bytes are kept, and blocks above kBlockMax are never cached.
//
// Small requests (up to kSmallMax) take power-of-two classes, and a class
// that runs dry is refilled with kSlab blocks carved from one allocation;
// small blocks stay with the process (never trimmed), so the many small
// tables of a run (job tables, node and tile maps, statuses) cost no
// allocation calls after the first of each class.
struct BlockCache {
    std::mutex mu;
    std::multimap<size_t, void*> idle;     // size -> block
    std::map<void*, size_t> size_of;       // every block the cache handed out
    size_t idle_bytes = 0;                 // large idle blocks only
    std::vector<std::pair<void*, hipEvent_t>> pending;   // freed once the event completes
    std::vector<hipEvent_t> ev_pool;
};
BlockCache& bcache() { static BlockCache b; return b; }
constexpr size_t kCacheMax = (size_t)8 << 30;
constexpr size_t kBlockMax = (size_t)2 << 30;
constexpr size_t kSmallMax = (size_t)1 << 20;
constexpr int kSlab = 16;

void cache_trim(BlockCache& b) {   // caller holds b.mu; small blocks stay
    for (auto it = b.idle.begin(); it != b.idle.end();) {
        if (it->first <= kSmallMax) { ++it; continue; }
        b.size_of.erase(it->second);
        (void)hipFree(it->second);
        it = b.idle.erase(it);
    }
    b.idle_bytes = 0;
}

void cache_put_locked(BlockCache& b, void* p) {
    auto it = b.size_of.find(p);
    if (it == b.size_of.end()) { (void)hipFree(p); return; }
    const size_t sz = it->second;
    if (sz <= kSmallMax) {                 // a slab block: back to its class
        b.idle.emplace(sz, p);
        return;
    }
    if (sz > kBlockMax || b.idle_bytes + sz > kCacheMax) {
        b.size_of.erase(it);
        (void)hipFree(p);
        return;
    }
    b.idle.emplace(sz, p);
    b.idle_bytes += sz;
}

// Pending blocks whose last use has completed become idle; with `wait`, every
// pending block is waited for. Caller holds b.mu.
void cache_reap(BlockCache& b, bool wait) {
    for (size_t i = 0; i < b.pending.size();) {
        const hipError_t q = wait ? hipEventSynchronize(b.pending[i].second) : hipEventQuery(b.pending[i].second);
        if (q == hipErrorNotReady) { (void)hipGetLastError(); ++i; continue; }
        void* p = b.pending[i].first;
        b.ev_pool.push_back(b.pending[i].second);
        b.pending[i] = b.pending.back();
        b.pending.pop_back();
        cache_put_locked(b, p);
    }
}

hipError_t cache_alloc(void** p, size_t bytes) {
    bytes = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
    BlockCache& b = bcache();
    std::lock_guard<std::mutex> g(b.mu);
    if (!b.pending.empty()) cache_reap(b, false);
    if (bytes <= kSmallMax) {
        size_t cls = 4096;
        while (cls < bytes) cls <<= 1;
        auto it = b.idle.find(cls);
        if (it == b.idle.end()) {
            char* slab = nullptr;
            hipError_t e = hipMalloc((void**)&slab, cls * kSlab);
            if (e != hipSuccess) return e;
            for (int i = 0; i < kSlab; ++i) {
                b.size_of[slab + i * cls] = cls;
                b.idle.emplace(cls, slab + i * cls);
            }
            it = b.idle.find(cls);
        }
        *p = it->second;
        b.idle.erase(it);
        return hipSuccess;
    }
    auto it = b.idle.lower_bound(bytes);
    while (it != b.idle.end() && it->first <= kSmallMax) ++it;
    if (it != b.idle.end() && it->first <= 2 * bytes) {
        *p = it->second;
        b.idle_bytes -= it->first;
        b.idle.erase(it);
        return hipSuccess;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory && (!b.idle.empty() || !b.pending.empty())) {
        (void)hipGetLastError();
        cache_reap(b, true);
        cache_trim(b);
        e = hipMalloc(p, bytes);
    }
    if (e == hipSuccess) b.size_of[*p] = bytes;
    return e;
}

void cache_free(void* p) {
    if (!p) return;
    BlockCache& b = bcache();
    std::lock_guard<std::mutex> g(b.mu);
    cache_put_locked(b, p);
}

// The block becomes reusable once the work enqueued on `s` so far has run.
hipError_t cache_free_after(void* p, hipStream_t s) {
    if (!p) return hipSuccess;
    BlockCache& b = bcache();
    std::lock_guard<std::mutex> g(b.mu);
    hipEvent_t ev = nullptr;
    if (!b.ev_pool.empty()) { ev = b.ev_pool.back(); b.ev_pool.pop_back(); }
    else {
        hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    hipError_t e = hipEventRecord(ev, s);
    if (e != hipSuccess) { b.ev_pool.push_back(ev); return e; }
    b.pending.emplace_back(p, ev);
    return hipSuccess;
}

// Page-locked host blocks for staging descriptor tables: an async copy from
// pageable memory runs synchronously, so the tables of calls that return
// before their work has run go through these. A block is reused once the
// copy that read it has run (an event per block); classes are powers of two
// from 4 KB, and blocks stay allocated until the process ends.
struct PinCache {
    std::mutex mu;
    std::multimap<size_t, void*> idle;
    std::map<void*, size_t> size_of;
    std::vector<std::pair<void*, hipEvent_t>> pending;
    std::vector<hipEvent_t> ev_pool;
};
PinCache& pcache() { static PinCache b; return b; }

void pin_reap(PinCache& b) {   // caller holds b.mu
    for (size_t i = 0; i < b.pending.size();) {
        const hipError_t q = hipEventQuery(b.pending[i].second);
        if (q == hipErrorNotReady) { (void)hipGetLastError(); ++i; continue; }
        b.idle.emplace(b.size_of[b.pending[i].first], b.pending[i].first);
        b.ev_pool.push_back(b.pending[i].second);
        b.pending[i] = b.pending.back();
        b.pending.pop_back();
    }
}

hipError_t pin_alloc(void** p, size_t bytes) {
    size_t cls = 4096;
    while (cls < bytes) cls <<= 1;
    PinCache& b = pcache();
    std::lock_guard<std::mutex> g(b.mu);
    if (!b.pending.empty()) pin_reap(b);
    auto it = b.idle.find(cls);
    if (it != b.idle.end()) {
        *p = it->second;
        b.idle.erase(it);
        return hipSuccess;
    }
    hipError_t e = hipHostMalloc(p, cls, hipHostMallocDefault);
    if (e == hipSuccess) b.size_of[*p] = cls;
    return e;
}

// The staging block becomes reusable once the work enqueued on `s` so far
// (its copy among it) has run.
hipError_t pin_free_after(void* p, hipStream_t s) {
    if (!p) return hipSuccess;
    PinCache& b = pcache();
    std::lock_guard<std::mutex> g(b.mu);
    hipEvent_t ev = nullptr;
    if (!b.ev_pool.empty()) { ev = b.ev_pool.back(); b.ev_pool.pop_back(); }
    else {
        hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    hipError_t e = hipEventRecord(ev, s);
    if (e != hipSuccess) { b.ev_pool.push_back(ev); return e; }
    b.pending.emplace_back(p, ev);
    return hipSuccess;
}

// Device scratch of one call, from the block cache. Blocks go back when
// the call ends, once the work enqueued on its stream so far has run.
struct CallScratch {
    hipStream_t st;
    std::vector<void*> blocks;
    explicit CallScratch(hipStream_t s) : st(s) {}
    ~CallScratch() { for (void* p : blocks) (void)cache_free_after(p, st); }
    CallScratch(const CallScratch&) = delete;
    CallScratch& operator=(const CallScratch&) = delete;
    template <class T> int get(T** p, size_t bytes) {
        void* q = nullptr;
        hipError_t e = cache_alloc(&q, bytes);
        if (e != hipSuccess) return fail_hip(e, "scratch allocation");
        blocks.push_back(q);
        *p = (T*)q;
        return PDH_OK;
    }
};

// Internal streams for host-buffer calls (multiply, BBP), one per concurrent
// call, kept for reuse.
struct StreamPool {
    std::mutex mu;
    std::vector<hipStream_t> idle;
};
StreamPool& spool() { static StreamPool p; return p; }

struct PooledStream {
    hipStream_t s = nullptr;
    ~PooledStream() {
        if (!s) return;
        StreamPool& p = spool();
        std::lock_guard<std::mutex> g(p.mu);
        p.idle.push_back(s);
    }
    int get() {
        StreamPool& p = spool();
        {
            std::lock_guard<std::mutex> g(p.mu);
            if (!p.idle.empty()) { s = p.idle.back(); p.idle.pop_back(); return PDH_OK; }
        }
        HIPCALL(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        return PDH_OK;
    }
};

void noalloc_setup(Ctx& c);

int init_locked(Ctx& c) {
    if (c.state == 1) return PDH_OK;
    if (c.state == -1) return fail(PDH_E_NODEVICE, c.why);
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) {
        c.state = -1;
        c.why = "[hip] no HIP device available";
        return fail(PDH_E_NODEVICE, c.why);
    }
    int dev = 0;
    HIPCALL(hipGetDevice(&dev));
    c.device = dev;
    hipDeviceProp_t prop;
    HIPCALL(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        c.state = -1;
        c.why = std::string("[hip] device is ") + prop.gcnArchName + ", this library is built for gfx950";
        return fail(PDH_E_NODEVICE, c.why);
    }
    HIPCALL(init_kernel_attrs());
    HIPCALL(init_kernel_attrs3());
    // twL[2^L + x] = w_{2^L}^x, L = 1..13
    std::vector<u64> tw((size_t)1 << 14, 0);
    for (int L = 1; L <= 13; ++L) {
        const u64 w = gl_root_host(L);
        u64 v = 1;
        for (size_t x = 0; x < ((size_t)1 << L); ++x) { tw[((size_t)1 << L) + x] = v; v = gl_mul(v, w); }
    }
    HIPCALL(hipMalloc(&c.twL, tw.size() * sizeof(u64)));
    HIPCALL(hipMemcpy(c.twL, tw.data(), tw.size() * sizeof(u64), hipMemcpyHostToDevice));
    // level twiddles of the three-prime engine, per prime: block (qi, dir) of
    // 2^15 words for radix Q = 2^(qi+1), then at 2^logM + j Q + c the value
    // w_{2^logM}^(+-j c) R mod p (Montgomery form), j < 2^logM / Q, c < Q
    std::vector<u32> t3(3 * TW3_STRIDE, 0);
    for (int pi = 0; pi < 3; ++pi) {
        const u32 p = primes3().p[pi].p;
        for (int qi = 0; qi < 4; ++qi) {
            const int Q = 2 << qi;
            for (int dir = 0; dir < 2; ++dir) {
                u32* blk = t3.data() + pi * TW3_STRIDE + (size_t)(qi * 2 + dir) * ((size_t)1 << 15);
                for (int logM = qi + 1; logM <= 14; ++logM) {
                    const u32 w0 = r3_root(pi, logM);
                    const u32 w = dir ? r3_pow(w0, p - 2, p) : w0;
                    const size_t nj = ((size_t)1 << logM) / Q;
                    for (size_t j = 0; j < nj; ++j) {
                        const u32 wj = r3_pow(w, j, p);
                        u32 v = 1;
                        for (int c = 0; c < Q; ++c) {
                            blk[((size_t)1 << logM) + j * Q + c] = r3_to_mont(v, p);
                            v = r3_mul(v, wj, p);
                        }
                    }
                }
            }
        }
    }
    HIPCALL(hipMalloc(&c.bbp_next, sizeof(u64)));
    HIPCALL(hipMalloc(&c.tw3, t3.size() * sizeof(u32)));
    HIPCALL(hipMemcpy(c.tw3, t3.data(), t3.size() * sizeof(u32), hipMemcpyHostToDevice));
    c.state = 1;
    noalloc_setup(c);
    return PDH_OK;
}

int outer_for(Ctx& c, int n, OuterTables* ot) {
    auto it = c.outer.find(n);
    if (it == c.outer.end()) {
        Outer o;
        o.h = (n + 1) / 2;
        const size_t nlo = (size_t)1 << o.h, nhi = (size_t)1 << (n - o.h);
        const u64 w = gl_root_host(n), wi = gl_inv(w);
        const u64 ninv = gl_inv((u64)1 << n);
        std::vector<u64> lo(nlo), lo_inv(nlo), hi(nhi), hi_inv(nhi);
        u64 v = 1, vi = 1;
        for (size_t x = 0; x < nlo; ++x) { lo[x] = v; lo_inv[x] = vi; v = gl_mul(v, w); vi = gl_mul(vi, wi); }
        const u64 step = v, step_i = vi;              // w^(2^h), w^(-2^h)
        v = 1; vi = ninv;
        for (size_t y = 0; y < nhi; ++y) { hi[y] = v; hi_inv[y] = vi; v = gl_mul(v, step); vi = gl_mul(vi, step_i); }
        HIPCALL(hipMalloc(&o.lo, nlo * 8));
        HIPCALL(hipMalloc(&o.lo_inv, nlo * 8));
        HIPCALL(hipMalloc(&o.hi, nhi * 8));
        HIPCALL(hipMalloc(&o.hi_inv, nhi * 8));
        HIPCALL(hipMemcpy(o.lo, lo.data(), nlo * 8, hipMemcpyHostToDevice));
        HIPCALL(hipMemcpy(o.lo_inv, lo_inv.data(), nlo * 8, hipMemcpyHostToDevice));
        HIPCALL(hipMemcpy(o.hi, hi.data(), nhi * 8, hipMemcpyHostToDevice));
        HIPCALL(hipMemcpy(o.hi_inv, hi_inv.data(), nhi * 8, hipMemcpyHostToDevice));
        // three primes over 2^(n-1) points, per prime: hi | lo | hi_inv | lo_inv
        // with w^e R = mm(hi[e >> h3], lo[e & (2^h3 - 1)]) and the inverse pair
        // carrying s = 2^-(n-1) R (undoing the pointwise Montgomery factor)
        const int n3 = n - 1;
        o.h3 = (n3 + 1) / 2;
        const size_t l3 = (size_t)1 << o.h3, h3n = (size_t)1 << (n3 - o.h3);
        std::vector<u32> t((size_t)3 * 2 * (l3 + h3n));
        for (int pi = 0; pi < 3; ++pi) {
            const u32 p = primes3().p[pi].p;
            const u32 w = r3_root(pi, n3), wi = r3_pow(w, p - 2, p);
            const u32 sc = r3_mul(r3_pow(r3_pow(2, n3, p), p - 2, p), r3_to_mont(1, p), p);   // 2^-n3 R
            u32* hi3 = t.data() + (size_t)pi * 2 * (l3 + h3n);
            u32* lo3 = hi3 + h3n;
            u32* hi3i = lo3 + l3;
            u32* lo3i = hi3i + h3n;
            u32 v = 1, vi = 1;
            for (size_t x = 0; x < l3; ++x) {
                lo3[x] = r3_to_mont(v, p); lo3i[x] = r3_to_mont(vi, p);
                v = r3_mul(v, w, p); vi = r3_mul(vi, wi, p);
            }
            const u32 st = v, sti = vi;                   // w^(2^h3), w^(-2^h3)
            v = 1; vi = sc;
            for (size_t y = 0; y < h3n; ++y) {
                hi3[y] = r3_to_mont(v, p); hi3i[y] = r3_to_mont(vi, p);
                v = r3_mul(v, st, p); vi = r3_mul(vi, sti, p);
            }
        }
        HIPCALL(hipMalloc(&o.o3, t.size() * 4));
        HIPCALL(hipMemcpy(o.o3, t.data(), t.size() * 4, hipMemcpyHostToDevice));
        it = c.outer.emplace(n, o).first;
    }
    ot->hi = it->second.hi; ot->lo = it->second.lo;
    ot->hi_inv = it->second.hi_inv; ot->lo_inv = it->second.lo_inv; ot->h = it->second.h;
    ot->tw3 = c.tw3; ot->o3 = it->second.o3; ot->h3 = it->second.h3;
    return PDH_OK;
}

// Set-up of pdh_mul_dev_noalloc, with the context's device current: the
// device address of `scratch`, the outer twiddle tables of every transform
// size it can use and its ordering event, so that the entry point itself
// allocates nothing. A failure here leaves the rest of the library usable
// and makes only that entry point report it.
void noalloc_setup(Ctx& c) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c.noalloc_why = std::string("[hip] hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
        return;
    }
    if (carry_ws_bytes(PDH_NOALLOC_MAX_LIMBS) > NOALLOC_CARRY_BYTES) {
        c.noalloc_why = "allocation-free multiply: carry scratch larger than reserved";
        return;
    }
    for (int n = SMALL_MAX_LOG + 1; n <= NOALLOC_MAX_LOG; ++n) {
        OuterTables ot{};
        if (outer_for(c, n, &ot) != PDH_OK) { c.noalloc_why = t_err; return; }
    }
    if (hipEventCreateWithFlags(&c.noalloc_order.ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        c.noalloc_why = "allocation-free multiply: event creation failed";
        return;
    }
    c.noalloc_ptr = (char*)p;
}

template <class P>
int grow(P** p, size_t* have, size_t need, size_t elem) {
    if (*have >= need) return PDH_OK;
    if (*p) { HIPCALL(hipDeviceSynchronize()); HIPCALL(hipFree(*p)); }   // queued work may still use it
    *p = nullptr;
    *have = 0;
    HIPCALL(hipMalloc((void**)p, need * elem));
    *have = need;
    return PDH_OK;
}

int ensure_ws(Ctx& c, size_t N, size_t nlimbs) {
    if (c.ws_words < N) {
        if (c.wsA) HIPCALL(hipDeviceSynchronize());    // queued work on any stream may still use it
        if (c.wsA) HIPCALL(hipFree(c.wsA));
        if (c.wsB) HIPCALL(hipFree(c.wsB));
        c.wsA = c.wsB = nullptr;
        c.ws_words = 0;
        HIPCALL(hipMalloc(&c.wsA, N * 8));
        HIPCALL(hipMalloc(&c.wsB, N * 8));
        c.ws_words = N;
    }
    return grow(&c.carry, &c.carry_bytes, carry_ws_bytes(nlimbs), 1);
}

int log2_exact(size_t N) { int l = 0; while (((size_t)1 << l) < N) ++l; return l; }

int mul_dev_locked(Ctx& c, const u32* a, size_t na, const u32* b, size_t nb, bool square,
                   u32* out, u64* out_len, hipStream_t s);

// Products beyond one transform (na + nb > 2^23 limbs): a and b in blocks of
// 2^22 limbs, each block product one 2^24-point transform into scratch,
// added into out at its offset by the single-node signed sum (every partial
// sum is at most the product, so it fits from the offset to the end), out
// zeroed first and its length normalized by a last pass. Everything is
// enqueued on s; the scratch is the context's, as for the other
// device-resident calls.
int mul_blocked_locked(Ctx& c, const u32* a, size_t na, const u32* b, size_t nb, u32* out, u64* out_len,
                       hipStream_t s) {
    const size_t K = (size_t)1 << 22, total = na + nb;
    const size_t ba = (na + K - 1) / K, bb = (nb + K - 1) / K;
    std::vector<u64> wl;                               // window lengths: total - offset, then total
    for (size_t i = 0; i < ba; ++i)
        for (size_t j = 0; j < bb; ++j) wl.push_back(total - (i + j) * K);
    wl.push_back(total);
    const size_t need = ((2 * K * 4 + 255) & ~(size_t)255) + 256 + wl.size() * 8 + 64;
    int rc = grow(&c.blk_scratch, &c.blk_scratch_n, need, 1);
    if (rc) return rc;
    const size_t nt = (total + BS_ADD_TILE - 1) / BS_ADD_TILE;
    if ((rc = grow(&c.sadd_status, &c.sadd_status_n, nt, 1))) return rc;
    char* w = (char*)c.blk_scratch;
    u32* P = (u32*)w;
    u64* PL = (u64*)(w + ((2 * K * 4 + 255) & ~(size_t)255));
    u64* dwl = PL + 1;
    u64* slen = dwl + wl.size();
    int32_t* ssgn = (int32_t*)(slen + 1);
    // the scratch may still serve earlier work on any stream: the host waits
    // for it before the synchronous upload of the window lengths
    if ((rc = scratch_wait_host(c.blk_order))) return rc;
    HIPCALL(hipMemcpy(dwl, wl.data(), wl.size() * 8, hipMemcpyHostToDevice));
    if ((rc = scratch_acquire(c.sadd_order, s))) return rc;
    hipError_t e = launch_fill_zero(out, total, s);
    if (e != hipSuccess) return fail_hip(e, "zero fill launch");
    size_t k = 0;
    for (size_t i = 0; i < ba; ++i) {
        for (size_t j = 0; j < bb; ++j, ++k) {
            const size_t ka = std::min(K, na - i * K), kb = std::min(K, nb - j * K);
            if ((rc = mul_dev_locked(c, a + i * K, ka, b + j * K, kb, false, P, PL, s))) return rc;
            u32* win = out + (i + j) * K;
            const BsAddNode node{win, win, P, dwl + k, PL, wl[k], slen, ssgn, 0u,
                                 (uint32_t)((wl[k] + BS_ADD_TILE - 1) / BS_ADD_TILE), 1, 1};
            e = launch_bs_signed_add(nullptr, nullptr, node.ntiles, c.sadd_status, s, &node);
            if (e != hipSuccess) return fail_hip(e, "block sum launch");
        }
    }
    const BsAddNode norm{out, out, out, dwl + k, dwl + k, total, out_len, ssgn, 0u, (uint32_t)nt, 1, 0};
    e = launch_bs_signed_add(nullptr, nullptr, (uint32_t)nt, c.sadd_status, s, &norm);
    if (e != hipSuccess) return fail_hip(e, "normalize launch");
    if ((rc = scratch_release(c.sadd_order, s))) return rc;
    return scratch_release(c.blk_order, s);
}

int mul_dev_locked(Ctx& c, const u32* a, size_t na, const u32* b, size_t nb, bool square,
                   u32* out, u64* out_len, hipStream_t s) {
    if (na == 0 || (!square && nb == 0)) {
        hipError_t e = launch_zero_len(out_len, s);
        if (e != hipSuccess) return fail_hip(e, "zero-length launch");
        return PDH_OK;
    }
    if (square) { b = a; nb = na; }
    const size_t N = pdh_ntt_length(na, nb);
    const int n = log2_exact(N);
    if (n > MAX_LOG) return mul_blocked_locked(c, a, na, b, nb, out, out_len, s);
    int rc = ensure_ws(c, N, na + nb);
    if (rc) return rc;
    OuterTables ot{nullptr, nullptr, nullptr, nullptr, 0, c.tw3, nullptr, 0};   // tw3: the small three-prime path
    const NttShape sh = ntt_shape(n);
    if (sh.L1 > 0 && (rc = outer_for(c, n, &ot))) return rc;
    if ((rc = scratch_acquire(c.ws_order, s))) return rc;
    hipError_t e = launch_mul(a, na, b, nb, square, out, out_len, c.wsA, c.wsB, c.carry, c.twL, ot, sh, s);
    if (e != hipSuccess) return fail_hip(e, "multiply launch");
    return scratch_release(c.ws_order, s);
}

// Host-buffer multiply: operands staged into per-call device scratch on a
// pooled stream; the context mutex is held only while the transforms are
// enqueued, not while the host waits for them.
int mul_host(const u32* a, size_t na, const u32* b, size_t nb, bool square,
             u32* out, size_t out_cap, size_t* out_len) {
    if (!out_len || (na && !a) || (!square && nb && !b)) return fail(PDH_E_ARG, "null pointer argument");
    if (square) nb = na;
    if (na == 0 || nb == 0) { *out_len = 0; return PDH_OK; }
    if (!out || out_cap < na + nb) return fail(PDH_E_CAPACITY, "output capacity below na + nb limbs");
    Ctx& c = ctx();
    {
        std::lock_guard<std::mutex> g(c.mu);
        int rc = init_locked(c);
        if (rc) return rc;
    }
    PooledStream ps;
    int rc = ps.get();
    if (rc) return rc;
    const hipStream_t st = ps.s;
    const size_t nl = na + nb;
    u64 len = 0;
    {
        CallScratch sc(st);
        char* w = nullptr;
        const size_t oa = 0, ob = (na * 4 + 255) & ~(size_t)255;
        const size_t oo = ob + (square ? 0 : (nb * 4 + 255) & ~(size_t)255);
        const size_t ol = oo + ((nl * 4 + 255) & ~(size_t)255);
        if ((rc = sc.get(&w, ol + 8))) return rc;
        u32* da = (u32*)(w + oa);
        u32* db = square ? da : (u32*)(w + ob);
        u32* dout = (u32*)(w + oo);
        u64* dlen = (u64*)(w + ol);
        HIPCALL(hipMemcpyAsync(da, a, na * 4, hipMemcpyHostToDevice, st));
        if (!square) HIPCALL(hipMemcpyAsync(db, b, nb * 4, hipMemcpyHostToDevice, st));
        {
            std::lock_guard<std::mutex> g(c.mu);
            rc = mul_dev_locked(c, da, na, db, nb, square, dout, dlen, st);
        }
        if (rc) return rc;
        HIPCALL(hipMemcpyAsync(&len, dlen, 8, hipMemcpyDeviceToHost, st));
        HIPCALL(hipStreamSynchronize(st));
        if (len > nl) return fail(PDH_E_HIP, "device returned an invalid product length");
        if (len) {
            HIPCALL(hipMemcpyAsync(out, dout, len * 4, hipMemcpyDeviceToHost, st));
            HIPCALL(hipStreamSynchronize(st));
        }
    }
    *out_len = (size_t)len;
    return PDH_OK;
}

constexpr uint64_t BBP_POS_MAX = (uint64_t)1 << 45;   // 8(pos-1)+6 < 2^48, where the reference's mulmod is exact

// Digits per anchor of a contiguous range whose largest d (position - 1) is
// dmax: the largest G <= 8 with 16^(G-1) e1 <= 2^46, e1 = 4 (dmax + 15) + 2
// (the fixed-point error bound of an anchor). The last digit of a block is
// then undecided with probability about 2^-13; G is 7 up to d = 65,535 and
// 5 at d = 10^8.
uint32_t bbp_block_digits(uint64_t dmax) {
    const uint64_t e1 = 4 * (dmax + 15) + 2;
    uint32_t G = 1;
    while (G < 8 && (e1 << (4 * G)) <= ((uint64_t)1 << 46)) ++G;
    return G;
}

// The BBP configuration a call runs with (read under the context mutex at
// its start; the call itself holds no lock).
struct BbpCfg {
    int exact = 0;     // pdh_bbp_set_exact mode
    int layout = 0;    // pdh_bbp_set_layout mode
};

int bbp_few(const BbpCfg& cfg, const uint64_t* d_pos, const uint64_t* h_pos, uint64_t start, size_t count,
            uint64_t dmax, uint8_t* d_out, uint8_t* h_vals, hipStream_t st, bool blocks = false,
            double share = -1.0);

// Share of lanes the positions-in-lanes term kernel keeps busy: positions
// h_pos[i] (or start + i stride) in groups of 64, sum of positions over 64 x
// each group's largest (a lane idles for exponents above its own d).
double lane_share(const uint64_t* h_pos, uint64_t start, uint64_t stride, size_t count) {
    double busy = 0, total = 0;
    for (size_t g = 0; g < count; g += 64) {
        uint64_t mx = 0;
        const size_t end = std::min(count, g + 64);
        for (size_t i = g; i < end; ++i) {
            const uint64_t p = h_pos ? h_pos[i] : start + i * stride;
            busy += (double)p;
            mx = std::max(mx, p);
        }
        total += 64.0 * (double)mx;
    }
    return total > 0 ? busy / total : 0.0;
}

// Positions in lanes costs about 0.5-0.75 of terms in lanes per busy lane
// (measured: 1..65,536, share 0.99, 27 against 36 ms; 512 sorted positions
// in [10^6, 10^8], share 0.89, 0.47 against 0.80 s), so it pays from a
// share of about 0.6 up (64 such positions have 0.52).
// Moduli in lanes, where it applies (largest position up to 2^27 - 1), is
// the default for more than 64 positions: measured against the better of
// the other two it takes 10.5 against 23.6 ms on 1..65,536, 0.28 against
// 0.44 s on 512 sorted positions in [10^6, 10^8] and 1.7 against 4.7 s on
// 4,096 unsorted ones (sorted on the device first). Its cost has a part set
// by the largest position alone (every block of 64 terms sets up its lanes
// and runs their long ladders however few positions its group holds), so
// for up to 64 positions the other two serve: 1 position near 5*10^7 takes
// 1.0 against 33 ms, 64 spread ones 62 against 70 ms. pdh_bbp_set_layout 1
// and 2 force the others, 3 forces it (the default rule where it does not
// apply).
int bbp_term_kernel(const BbpCfg& cfg, double share, uint64_t dmax, size_t count) {
    if (dmax <= BBP_MODULI_DMAX_HOST && (cfg.layout == 3 || (cfg.layout == 0 && count > 64))) return BBP_TK_MODULI;
    if (cfg.layout == 1) return BBP_TK_TERMS;
    return cfg.layout == 2 || share >= 0.6 ? BBP_TK_LANES : BBP_TK_TERMS;
}

// A contiguous range by digit blocks: one fixed-point evaluation per G
// positions (k_bbp_fx_block), then the per-position paths for the digits
// left undecided.
int bbp_blocks(const BbpCfg& cfg, uint64_t start, size_t count, uint32_t G, uint8_t* d_out, uint8_t* h_vals,
               hipStream_t st) {
    CallScratch sc(st);
    const size_t A = (count + G - 1) / G;
    void* ws = nullptr;
    int rc = sc.get(&ws, bbp_few_ws_bytes(A));
    if (rc) return rc;
    hipError_t e = launch_bbp_fx(nullptr, start, A, start - 1 + (A - 1) * (uint64_t)G, d_out, ws,
                                 cfg.exact == 2, st, G, count,
                                 bbp_term_kernel(cfg, lane_share(nullptr, start, G, A), start - 1 + (A - 1) * (uint64_t)G, A));
    if (e != hipSuccess) return fail_hip(e, "bbp launch");
    HIPCALL(hipMemcpyAsync(h_vals, d_out, count, hipMemcpyDeviceToHost, st));
    HIPCALL(hipStreamSynchronize(st));
    std::vector<size_t> und;
    std::vector<uint64_t> upos;
    for (size_t i = 0; i < count; ++i)
        if (h_vals[i] == 0xFE) { und.push_back(i); upos.push_back(start + i); }
    if (und.empty()) return PDH_OK;
    const size_t nu = und.size();
    uint64_t* dpos = nullptr;
    uint8_t* dval = nullptr;
    if ((rc = sc.get(&dpos, nu * 8)) || (rc = sc.get(&dval, nu))) return rc;
    HIPCALL(hipMemcpyAsync(dpos, upos.data(), nu * 8, hipMemcpyHostToDevice, st));
    std::vector<uint8_t> fixed(nu);
    // under the test mode 2 (every odd index undecided) those digits all take the ordered path
    BbpCfg sub = cfg;
    if (sub.exact == 2) sub.exact = 1;
    if ((rc = bbp_few(sub, dpos, upos.data(), 0, nu, upos.back() - 1, dval, fixed.data(), st))) return rc;
    for (size_t i = 0; i < nu; ++i) h_vals[und[i]] = fixed[i];
    HIPCALL(hipMemcpyAsync(d_out, h_vals, count, hipMemcpyHostToDevice, st));
    HIPCALL(hipStreamSynchronize(st));
    return PDH_OK;
}

// The fixed-point path, then the ordered FP64 path for every position it
// leaves undecided (or for all of them when pdh_bbp_set_exact is 1). d_pos (or start for a contiguous
// range) and d_out are device buffers; h_pos, when given, holds the same
// positions on the host. The values end in d_out and in h_vals. Synchronises
// `st`. A contiguous range with `blocks` set goes by digit blocks.
int bbp_few(const BbpCfg& cfg, const uint64_t* d_pos, const uint64_t* h_pos, uint64_t start, size_t count,
            uint64_t dmax, uint8_t* d_out, uint8_t* h_vals, hipStream_t st, bool blocks, double share) {
    if (blocks && !d_pos && cfg.exact != 1 && count >= 16) {
        const uint32_t G = bbp_block_digits(dmax);
        if (G > 1) return bbp_blocks(cfg, start, count, G, d_out, h_vals, st);
    }
    CallScratch sc(st);
    void* ws = nullptr;
    int rc = sc.get(&ws, bbp_few_ws_bytes(count));
    if (rc) return rc;
    if (cfg.exact == 1) {
        hipError_t e = launch_bbp_chunked(d_pos, start, count, dmax, d_out, ws, st);
        if (e != hipSuccess) return fail_hip(e, "bbp launch");
        HIPCALL(hipMemcpyAsync(h_vals, d_out, count, hipMemcpyDeviceToHost, st));
        HIPCALL(hipStreamSynchronize(st));
        ctx().bbp_exact_count += count;
        return PDH_OK;
    }
    if (share < 0) share = d_pos && !h_pos ? 0.0 : lane_share(h_pos, start, 1, count);
    hipError_t e = launch_bbp_fx(d_pos, start, count, dmax, d_out, ws, cfg.exact == 2, st, 1, 0,
                                 bbp_term_kernel(cfg, share, dmax, count));
    if (e != hipSuccess) return fail_hip(e, "bbp launch");
    HIPCALL(hipMemcpyAsync(h_vals, d_out, count, hipMemcpyDeviceToHost, st));
    HIPCALL(hipStreamSynchronize(st));
    std::vector<size_t> und;
    for (size_t i = 0; i < count; ++i)
        if (h_vals[i] == 0xFE) und.push_back(i);
    if (und.empty()) return PDH_OK;
    std::vector<uint64_t> all;
    if (d_pos && !h_pos) {
        all.resize(count);
        HIPCALL(hipMemcpyAsync(all.data(), d_pos, count * 8, hipMemcpyDeviceToHost, st));
        HIPCALL(hipStreamSynchronize(st));
        h_pos = all.data();
    }
    std::vector<uint64_t> upos(und.size());
    uint64_t umax = 0;
    for (size_t i = 0; i < und.size(); ++i) {
        upos[i] = h_pos ? h_pos[und[i]] : start + und[i];
        umax = std::max(umax, upos[i] - 1);
    }
    const size_t nu = und.size();
    uint64_t* fpos = nullptr;
    uint8_t* fval = nullptr;
    void* fws = nullptr;
    if ((rc = sc.get(&fpos, nu * 8)) || (rc = sc.get(&fval, nu)) || (rc = sc.get(&fws, bbp_few_ws_bytes(nu))))
        return rc;
    HIPCALL(hipMemcpyAsync(fpos, upos.data(), nu * 8, hipMemcpyHostToDevice, st));
    e = launch_bbp_chunked(fpos, 0, nu, umax, fval, fws, st);
    if (e != hipSuccess) return fail_hip(e, "bbp launch");
    std::vector<uint8_t> fixed(nu);
    HIPCALL(hipMemcpyAsync(fixed.data(), fval, nu, hipMemcpyDeviceToHost, st));
    HIPCALL(hipStreamSynchronize(st));
    for (size_t i = 0; i < nu; ++i) h_vals[und[i]] = fixed[i];
    HIPCALL(hipMemcpyAsync(d_out, h_vals, count, hipMemcpyHostToDevice, st));
    HIPCALL(hipStreamSynchronize(st));
    ctx().bbp_exact_count += nu;
    return PDH_OK;
}

// Initialises the context if needed and reads the BBP configuration.
int bbp_begin(BbpCfg* cfg) {
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = init_locked(c);
    if (rc) return rc;
    cfg->exact = c.bbp_exact;
    cfg->layout = c.bbp_layout;
    return PDH_OK;
}

int bbp_host(const uint64_t* positions, uint64_t start, size_t count, uint8_t* values) {
    if (count == 0) return PDH_OK;
    if (!values) return fail(PDH_E_ARG, "null output pointer");
    std::vector<uint64_t> sorted;
    std::vector<size_t> order;
    if (positions) {
        for (size_t i = 0; i < count; ++i)
            if (positions[i] < 1 || positions[i] > BBP_POS_MAX)
                return fail(PDH_E_ARG, "BBP position outside [1, 2^45]");
        // Similar positions share a wavefront: sort, compute, scatter back.
        order.resize(count);
        std::iota(order.begin(), order.end(), (size_t)0);
        std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return positions[x] < positions[y]; });
        sorted.resize(count);
        for (size_t i = 0; i < count; ++i) sorted[i] = positions[order[i]];
    } else if (start < 1 || start - 1 + count > BBP_POS_MAX) {
        return fail(PDH_E_ARG, "BBP position outside [1, 2^45]");
    }
    BbpCfg cfg;
    int rc = bbp_begin(&cfg);
    if (rc) return rc;
    PooledStream ps;
    if ((rc = ps.get())) return rc;
    const hipStream_t st = ps.s;
    std::vector<uint8_t> tmp(count);
    {
        CallScratch sc(st);
        uint64_t* dpos = nullptr;
        uint8_t* dval = nullptr;
        if ((positions && (rc = sc.get(&dpos, count * 8))) || (rc = sc.get(&dval, count))) return rc;
        if (positions) HIPCALL(hipMemcpyAsync(dpos, sorted.data(), count * 8, hipMemcpyHostToDevice, st));
        if (cfg.exact != 3) {
            const uint64_t dmax = positions ? sorted.back() - 1 : start + count - 2;
            rc = bbp_few(cfg, dpos, positions ? sorted.data() : nullptr, start, count, dmax, dval, tmp.data(), st,
                         !positions);
            if (rc) return rc;
        } else {
            Ctx& c = ctx();                        // the work-queue counter is the context's
            std::lock_guard<std::mutex> g(c.mu);
            hipError_t e = launch_bbp(dpos, start, count, dval, c.bbp_next, st);
            if (e != hipSuccess) return fail_hip(e, "bbp launch");
            HIPCALL(hipMemcpyAsync(tmp.data(), dval, count, hipMemcpyDeviceToHost, st));
            HIPCALL(hipStreamSynchronize(st));
        }
    }
    if (positions) for (size_t i = 0; i < count; ++i) values[order[i]] = tmp[i];
    else std::memcpy(values, tmp.data(), count);
    return PDH_OK;
}

// ---- optional per-kernel event timing -----------------------------------------
struct TimingRec { const char* name; hipEvent_t a, b; };
struct Timing {
    std::mutex mu;
    bool on = false;
    std::vector<TimingRec> recs;
    std::vector<hipEvent_t> pool;
    hipEvent_t get() {
        if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    void recycle() {
        for (auto& r : recs) { pool.push_back(r.a); pool.push_back(r.b); }
        recs.clear();
    }
};
Timing& timing() { static Timing t; return t; }

} // namespace

hipEvent_t timing_begin(hipStream_t s) {
    Timing& t = timing();
    std::lock_guard<std::mutex> g(t.mu);
    if (!t.on) return nullptr;
    hipEvent_t e = t.get();
    if (e) (void)hipEventRecord(e, s);
    return e;
}

void timing_end(const char* name, hipEvent_t ev0, hipStream_t s) {
    if (!ev0) return;
    Timing& t = timing();
    std::lock_guard<std::mutex> g(t.mu);
    hipEvent_t e = t.get();
    if (!e) { t.pool.push_back(ev0); return; }
    (void)hipEventRecord(e, s);
    t.recs.push_back(TimingRec{name, ev0, e});
}

} // namespace pdh

using namespace pdh;

extern "C" {

const char* pdh_last_error(void) { return t_err.c_str(); }

int pdh_device_available(void) {
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    return init_locked(c) == PDH_OK ? 1 : 0;
}

size_t pdh_ntt_length(size_t na, size_t nb) {
    size_t n = 1, want = 2 * na + 2 * nb;
    while (n < want) n <<= 1;
    return n;
}

int pdh_mul(const uint32_t* a, size_t na, const uint32_t* b, size_t nb,
            uint32_t* out, size_t out_cap, size_t* out_len) {
    return mul_host(a, na, b, nb, false, out, out_cap, out_len);
}

int pdh_sqr(const uint32_t* a, size_t na, uint32_t* out, size_t out_cap, size_t* out_len) {
    return mul_host(a, na, a, na, true, out, out_cap, out_len);
}

int pdh_mul_dev(const uint32_t* d_a, size_t na, const uint32_t* d_b, size_t nb,
                uint32_t* d_out, uint64_t* d_out_len, void* stream) {
    if (!d_out_len || (na && nb && (!d_a || !d_b || !d_out))) return fail(PDH_E_ARG, "null pointer argument");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = init_locked(c);
    if (rc) return rc;
    return mul_dev_locked(c, d_a, na, d_b, nb, false, d_out, d_out_len, (hipStream_t)stream);
}

int pdh_sqr_dev(const uint32_t* d_a, size_t na, uint32_t* d_out, uint64_t* d_out_len, void* stream) {
    if (!d_out_len || (na && (!d_a || !d_out))) return fail(PDH_E_ARG, "null pointer argument");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = init_locked(c);
    if (rc) return rc;
    return mul_dev_locked(c, d_a, na, d_a, na, true, d_out, d_out_len, (hipStream_t)stream);
}

int pdh_mul_dev_noalloc(const uint32_t* d_a, size_t na, const uint32_t* d_b, size_t nb,
                        uint32_t* d_out, uint64_t* d_out_len, void* stream) {
    if (!d_out_len || (na && nb && (!d_a || !d_b || !d_out))) return fail(PDH_E_ARG, "null pointer argument");
    if (na + nb > PDH_NOALLOC_MAX_LIMBS)
        return fail(PDH_E_TOO_LARGE, "pdh_mul_dev_noalloc: na + nb = " + std::to_string(na + nb) +
                                         " limbs, above its limit of 2^20 (use pdh_mul_dev)");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    if (c.state != 1)
        return fail(PDH_E_NODEVICE, "pdh_mul_dev_noalloc: the library is not set up; call pdh_device_available() "
                                    "first (set-up allocates, this entry point does not)");
    if (!c.noalloc_ptr) return fail(PDH_E_HIP, c.noalloc_why);
    int dev = -1;
    HIPCALL(hipGetDevice(&dev));
    if (dev != c.device) return fail(PDH_E_ARG, "pdh_mul_dev_noalloc: the current device is not the one set up");
    const hipStream_t s = (hipStream_t)stream;
    int rc = scratch_acquire(c.noalloc_order, s);
    if (rc) return rc;
    hipError_t e;
    if (na == 0 || nb == 0) {
        e = launch_zero_len(d_out_len, s);
    } else {
        const int n = log2_exact(pdh_ntt_length(na, nb));
        const NttShape sh = ntt_shape(n);
        OuterTables ot{};
        ot.tw3 = c.tw3;
        if (sh.L1 > 0) {
            auto it = c.outer.find(n);                 // built at set-up
            if (it == c.outer.end()) return fail(PDH_E_HIP, "pdh_mul_dev_noalloc: twiddle table missing");
            const Outer& o = it->second;
            ot = OuterTables{o.hi, o.lo, o.hi_inv, o.lo_inv, o.h, c.tw3, o.o3, o.h3};
        }
        u64* wsA = (u64*)c.noalloc_ptr;
        u64* wsB = (u64*)(c.noalloc_ptr + NOALLOC_WS_BYTES);
        u64* carry = (u64*)(c.noalloc_ptr + 2 * NOALLOC_WS_BYTES);
        e = launch_mul(d_a, na, d_b, nb, false, d_out, d_out_len, wsA, wsB, carry, c.twL, ot, sh, s);
    }
    if (e != hipSuccess) return fail_hip(e, "allocation-free multiply launch");
    return scratch_release(c.noalloc_order, s);
}

int pdh_bbp_range(uint64_t start, uint32_t count, uint8_t* values) {
    return bbp_host(nullptr, start, count, values);
}

int pdh_bbp_range_dev(uint64_t start, uint32_t count, uint8_t* d_values, void* stream) {
    if (count == 0) return PDH_OK;
    if (!d_values) return fail(PDH_E_ARG, "null output pointer");
    if (start < 1 || start - 1 + count > BBP_POS_MAX) return fail(PDH_E_ARG, "BBP position outside [1, 2^45]");
    BbpCfg cfg;
    int rc = bbp_begin(&cfg);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    std::vector<uint8_t> tmp(count);
    if (cfg.exact != 3)
        return bbp_few(cfg, nullptr, nullptr, start, count, start + count - 2, d_values, tmp.data(), st, true);
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    hipError_t e = launch_bbp(nullptr, start, count, d_values, c.bbp_next, st);
    if (e != hipSuccess) return fail_hip(e, "bbp launch");
    return PDH_OK;
}

int pdh_bbp_positions(const uint64_t* positions, size_t n, uint8_t* values) {
    if (n && !positions) return fail(PDH_E_ARG, "null positions pointer");
    return bbp_host(positions, 0, n, values);
}

int pdh_bbp_positions_dev(const uint64_t* d_positions, size_t n, uint8_t* d_values, void* stream) {
    if (n == 0) return PDH_OK;
    if (!d_positions || !d_values) return fail(PDH_E_ARG, "null pointer argument");
    BbpCfg cfg;
    int rc = bbp_begin(&cfg);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    if (cfg.exact != 3) {
        // both few-position paths size their grids by the largest position
        // and choose the term layout by how close the positions of each group of 64 lie
        uint64_t sc[3] = {0, 0, 0};
        {
            CallScratch scr(st);
            uint64_t* scan = nullptr;
            if ((rc = scr.get(&scan, 24))) return rc;
            e = hipMemsetAsync(scan, 0, 24, st);
            if (e == hipSuccess) e = launch_bbp_group_scan(d_positions, n, scan, st);
            if (e == hipSuccess) e = hipMemcpyAsync(sc, scan, 24, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) return fail_hip(e, "bbp position scan");
        }
        const double share = sc[2] ? (double)sc[1] / (double)sc[2] : 0.0;
        std::vector<uint8_t> tmp(n);
        if (n > 64 && cfg.exact != 1 && bbp_term_kernel(cfg, share, sc[0], n) == BBP_TK_MODULI) {
            // moduli in lanes steps a lane from one sorted position of its
            // group to the next at a cost that grows with the gap: sort the
            // positions on the device, compute in that order, scatter back
            CallScratch scr(st);
            size_t tb = 0;
            if ((e = bbp_sort_positions(d_positions, n, nullptr, nullptr, nullptr, &tb, st)) != hipSuccess)
                return fail_hip(e, "bbp position sort");
            uint64_t* dsorted = nullptr;
            uint32_t* dperm = nullptr;
            uint8_t* dvals = nullptr;
            void* dtemp = nullptr;
            if ((rc = scr.get(&dsorted, n * 8)) || (rc = scr.get(&dperm, n * 4)) || (rc = scr.get(&dvals, n)) ||
                (rc = scr.get(&dtemp, tb ? tb : 1)))
                return rc;
            if ((e = bbp_sort_positions(d_positions, n, dsorted, dperm, dtemp, &tb, st)) != hipSuccess)
                return fail_hip(e, "bbp position sort");
            if ((rc = bbp_few(cfg, dsorted, nullptr, 0, n, sc[0], dvals, tmp.data(), st, false, share))) return rc;
            if ((e = launch_bbp_scatter(dvals, dperm, n, d_values, st)) != hipSuccess)
                return fail_hip(e, "bbp scatter");
        } else {
            if ((rc = bbp_few(cfg, d_positions, nullptr, 0, n, sc[0], d_values, tmp.data(), st, false, share)))
                return rc;
        }
        for (size_t i = 0; i < n; ++i)
            if (tmp[i] == 0xFF) return fail(PDH_E_ARG, "BBP position outside [1, 2^45] (its value is 0xFF)");
        return PDH_OK;
    }
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    e = launch_bbp(d_positions, 0, n, d_values, c.bbp_next, st);
    if (e != hipSuccess) return fail_hip(e, "bbp launch");
    return PDH_OK;
}

int pdh_bs_groups(uint64_t n_terms, int group_log, uint32_t* d_P, uint32_t* d_Q, uint32_t* d_T, uint64_t* d_len,
                  int32_t* d_sign, void* stream) {
    if (group_log < 1 || group_log > 6) return fail(PDH_E_ARG, "group_log must lie in [1, 6]");
    if (n_terms > ((uint64_t)1 << 28)) return fail(PDH_E_ARG, "more than 2^28 terms");
    if (n_terms == 0) return PDH_OK;
    if (!d_P || !d_Q || !d_T || !d_len || !d_sign) return fail(PDH_E_ARG, "null pointer argument");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = init_locked(c);
    if (rc) return rc;
    hipError_t e = launch_bs_groups(n_terms, group_log, d_P, d_Q, d_T, d_len, d_sign, (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "binary-splitting groups");
    return PDH_OK;
}

int pdh_bbp_set_exact(int mode) {
    if (mode < 0 || mode > 3) return fail(PDH_E_ARG, "pdh_bbp_set_exact: mode must be 0, 1, 2 or 3");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    c.bbp_exact = mode;
    return PDH_OK;
}

int pdh_bbp_set_layout(int mode) {
    if (mode < 0 || mode > 3) return fail(PDH_E_ARG, "pdh_bbp_set_layout: mode must be 0, 1, 2 or 3");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    c.bbp_layout = mode;
    return PDH_OK;
}

uint64_t pdh_bbp_exact_count(void) { return ctx().bbp_exact_count.load(); }

uint64_t pdh_launch_count(void) { return g_launches.load(); }

int pdh_probe_mad_rate(double* lane_ops_per_s) {
    if (!lane_ops_per_s) return fail(PDH_E_ARG, "null pointer argument");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = init_locked(c);
    if (rc) return rc;
    hipError_t e = probe_mad_rate(lane_ops_per_s);
    if (e != hipSuccess) return fail_hip(e, "probe launch");
    return PDH_OK;
}

// ---- batched multiplies --------------------------------------------------------------
struct pdh_batch {
    struct Class { int n; unsigned njobs; MulJob* d_jobs; size_t nblk_max; OuterTables ot; NttShape shape; };
    std::vector<Class> classes;
    std::vector<uint64_t*> zero_lens;     // jobs with an empty operand
    char* arena = nullptr;
    MulJob* d_jobs = nullptr;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;       // fork event + one join event per stream
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t last = nullptr;           // stream of the latest run
    hipStream_t up = nullptr;             // stream of the job table's async upload (one-shot batches)
    bool ran = false;
    bool shared = false;                  // streams and events are the context's (one-shot batches)
};

namespace {

int batch_enqueue(pdh_batch* p, hipStream_t s) {
    Ctx& c = ctx();
    for (uint64_t* z : p->zero_lens) {
        hipError_t e = launch_zero_len(z, s);
        if (e != hipSuccess) return fail_hip(e, "zero-length launch");
    }
    if (p->classes.empty()) return PDH_OK;
    // fork: every internal stream waits for work already on s
    HIPCALL(hipEventRecord(p->events[0], s));
    for (size_t i = 0; i < p->streams.size(); ++i) HIPCALL(hipStreamWaitEvent(p->streams[i], p->events[0], 0));
    for (size_t k = 0; k < p->classes.size(); ++k) {
        const pdh_batch::Class& cl = p->classes[k];
        hipStream_t si = p->streams[k % p->streams.size()];
        MulJob none{};
        hipError_t e = launch_class(cl.d_jobs, none, cl.njobs, cl.n, false, cl.nblk_max, c.twL, cl.ot, si, &cl.shape);
        if (e != hipSuccess) return fail_hip(e, "batch launch");
    }
    // join
    for (size_t i = 0; i < p->streams.size(); ++i) {
        HIPCALL(hipEventRecord(p->events[1 + i], p->streams[i]));
        HIPCALL(hipStreamWaitEvent(s, p->events[1 + i], 0));
    }
    return PDH_OK;
}

void batch_free(pdh_batch* p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->ran) (void)hipStreamSynchronize(p->last);   // the caller's stream joins every branch
    if (!p->shared) {
        for (auto st : p->streams) (void)hipStreamDestroy(st);
        for (auto ev : p->events) (void)hipEventDestroy(ev);
    }
    if (p->up) {                          // an upload may still be in flight (error paths)
        (void)cache_free_after(p->arena, p->up);
        (void)cache_free_after(p->d_jobs, p->up);
    } else {
        cache_free(p->arena);
        cache_free(p->d_jobs);
    }
    delete p;
}

// A one-shot plan after its run: its blocks return to the cache once the
// caller's stream (which joins every branch of the run) has passed this
// point; the host does not wait.
void batch_release_after(pdh_batch* p, hipStream_t s) {
    (void)cache_free_after(p->arena, s);
    (void)cache_free_after(p->d_jobs, s);
    delete p;
}

// One-shot batches (created, run once and released) share the context's
// streams and events: their enqueues are serialised by the context mutex, a
// shared stream runs its work in order, and a stream's wait on an event takes
// the event's latest record at the time of the wait, so a later batch's
// records do not move an earlier batch's waits.
int shared_streams(Ctx& c, size_t n, pdh_batch* p) {
    while (c.sh_streams.size() < 4) {
        hipStream_t st;
        HIPCALL(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        c.sh_streams.push_back(st);
    }
    while (c.sh_events.size() < 5) {
        hipEvent_t ev;
        HIPCALL(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        c.sh_events.push_back(ev);
    }
    p->streams.assign(c.sh_streams.begin(), c.sh_streams.begin() + n);
    p->events.assign(c.sh_events.begin(), c.sh_events.begin() + n + 1);
    p->shared = true;
    return PDH_OK;
}

int batch_create_impl(const pdh_mul_desc* descs, size_t n, pdh_batch** plan, bool shared,
                      hipStream_t up = nullptr);

} // namespace

int pdh_batch_create(const pdh_mul_desc* descs, size_t n, pdh_batch** plan) {
    return batch_create_impl(descs, n, plan, false);
}

} // extern "C"

namespace {

// up: upload the job table asynchronously on this stream through a pinned
// staging block (the one-shot path); else a synchronous copy.
int batch_create_impl(const pdh_mul_desc* descs, size_t n, pdh_batch** plan, bool shared, hipStream_t up) {
    if (!plan || (n && !descs)) return fail(PDH_E_ARG, "null argument");
    *plan = nullptr;
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = init_locked(c);
    if (rc) return rc;
    pdh_batch* p = new pdh_batch();
    std::map<int, std::vector<size_t>> by_n;
    for (size_t i = 0; i < n; ++i) {
        const pdh_mul_desc& d = descs[i];
        if (!d.out_len) { batch_free(p); return fail(PDH_E_ARG, "null out_len in descriptor"); }
        if (d.na == 0 || d.nb == 0) { p->zero_lens.push_back(d.out_len); continue; }
        if (!d.a || !d.b || !d.out) { batch_free(p); return fail(PDH_E_ARG, "null pointer in descriptor"); }
        const int ln = log2_exact(pdh_ntt_length(d.na, d.nb));
        if (ln > MAX_LOG) {
            batch_free(p);
            return fail(PDH_E_TOO_LARGE, "batch product of " + std::to_string(d.na) + " x " + std::to_string(d.nb) +
                                             " limbs needs a transform beyond 2^24");
        }
        by_n[ln].push_back(i);
    }
    // workspace arena: per job of class n > SMALL_MAX_LOG two spectra + carry scratch
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t arena = 0, njobs_total = 0;
    for (auto& kv : by_n) {
        njobs_total += kv.second.size();
        if (kv.first > SMALL_MAX_LOG)
            arena += kv.second.size() * (2 * align(((size_t)8) << kv.first) + align(carry_ws_bytes((size_t)1 << (kv.first - 1))));
    }
    if (arena) {
        hipError_t e = cache_alloc((void**)&p->arena, arena);
        if (e != hipSuccess) { batch_free(p); return fail_hip(e, "batch workspace allocation"); }
    }
    std::vector<MulJob> jobs;
    jobs.reserve(njobs_total);
    size_t off = 0;
    for (auto& kv : by_n) {
        const int ln = kv.first;
        for (size_t i : kv.second) {
            const pdh_mul_desc& d = descs[i];
            MulJob J{d.a, d.b, d.out, d.out_len, d.na, d.nb, nullptr, nullptr, nullptr};
            if (ln > SMALL_MAX_LOG) {
                J.wsA = (uint64_t*)(p->arena + off); off += align((size_t)8 << ln);
                J.wsB = (uint64_t*)(p->arena + off); off += align((size_t)8 << ln);
                J.carry = p->arena + off; off += align(carry_ws_bytes((size_t)1 << (ln - 1)));
            }
            jobs.push_back(J);
        }
    }
    if (njobs_total) {
        const size_t jb = njobs_total * sizeof(MulJob);
        hipError_t e = cache_alloc((void**)&p->d_jobs, jb);
        if (e == hipSuccess && up) {
            void* pin = nullptr;
            p->up = up;
            e = pin_alloc(&pin, jb);
            if (e == hipSuccess) {
                std::memcpy(pin, jobs.data(), jb);
                e = hipMemcpyAsync(p->d_jobs, pin, jb, hipMemcpyHostToDevice, up);
                const hipError_t e2 = pin_free_after(pin, up);
                if (e == hipSuccess) e = e2;
            }
        } else if (e == hipSuccess) {
            e = hipMemcpy(p->d_jobs, jobs.data(), jb, hipMemcpyHostToDevice);
        }
        if (e != hipSuccess) { batch_free(p); return fail_hip(e, "batch job table upload"); }
    }
    size_t first = 0;
    for (auto& kv : by_n) {
        pdh_batch::Class cl;
        cl.n = kv.first;
        cl.njobs = (unsigned)kv.second.size();
        cl.d_jobs = p->d_jobs + first;
        cl.nblk_max = carry_nblk((size_t)1 << (kv.first - 1));
        cl.shape = batch_shape(kv.first, cl.njobs);
        cl.ot = OuterTables{nullptr, nullptr, nullptr, nullptr, 0, c.tw3, nullptr, 0};
        if (kv.first > SMALL_MAX_LOG && (rc = outer_for(c, kv.first, &cl.ot))) { batch_free(p); return rc; }
        first += kv.second.size();
        p->classes.push_back(cl);
    }
    // largest classes first: they have the longest critical path
    std::reverse(p->classes.begin(), p->classes.end());
    const size_t nstreams = std::min<size_t>(4, std::max<size_t>(1, p->classes.size()));
    if (shared) {
        if ((rc = shared_streams(c, nstreams, p))) { batch_free(p); return rc; }
        *plan = p;
        return PDH_OK;
    }
    for (size_t i = 0; i < nstreams; ++i) {
        hipStream_t st;
        hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) { batch_free(p); return fail_hip(e, "batch stream"); }
        p->streams.push_back(st);
    }
    for (size_t i = 0; i < nstreams + 1; ++i) {
        hipEvent_t ev;
        hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) { batch_free(p); return fail_hip(e, "batch event"); }
        p->events.push_back(ev);
    }
    *plan = p;
    return PDH_OK;
}

} // namespace

extern "C" {

int pdh_batch_run(pdh_batch* p, void* stream) {
    if (!p) return fail(PDH_E_ARG, "null plan");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    p->last = (hipStream_t)stream;
    p->ran = true;
    if (p->exec) {
        HIPCALL(hipGraphLaunch(p->exec, (hipStream_t)stream));
        g_launches.fetch_add(1, std::memory_order_relaxed);
        return PDH_OK;
    }
    return batch_enqueue(p, (hipStream_t)stream);
}

int pdh_batch_capture(pdh_batch* p, void* stream) {
    if (!p) return fail(PDH_E_ARG, "null plan");
    if (!stream) return fail(PDH_E_ARG, "graph capture needs a non-null stream");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    if (p->exec) return PDH_OK;
    hipStream_t s = (hipStream_t)stream;
    HIPCALL(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = batch_enqueue(p, s);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(s, &graph);
    if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return fail_hip(e, "hipStreamEndCapture");
    p->graph = graph;
    HIPCALL(hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0));
    return PDH_OK;
}

int pdh_batch_destroy(pdh_batch* p) {
    if (!p) return PDH_OK;
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    batch_free(p);
    return PDH_OK;
}

int pdh_mul_batch_dev(const pdh_mul_desc* descs, size_t n, void* stream) {
    pdh_batch* p = nullptr;
    int rc = batch_create_impl(descs, n, &p, true, (hipStream_t)stream);
    if (rc) return rc;
    rc = pdh_batch_run(p, stream);
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    if (rc == PDH_OK) batch_release_after(p, (hipStream_t)stream);
    else batch_free(p);
    return rc;
}

// One merge level of binary splitting in device memory (src/chudnovsky.cpp:
// 93-100): the pairs' Q, T_L Q_R, P_L T_R and P products as one batch, with
// T_L Q_R written straight into the output T buffer and P_L T_R into
// scratch, then the signed sum in place (launch_bs_signed_add). Only the
// lengths and signs come back to the host.
int pdh_bs_merge_dev(const pdh_bs_node* in, size_t n_in, pdh_bs_node* out, int skip_p, void* stream) {
    if (n_in == 0) return PDH_OK;
    if (!in || !out) return fail(PDH_E_ARG, "null pointer argument");
    for (size_t i = 0; i < n_in; ++i) {
        const pdh_bs_node& v = in[i];
        if (v.nP == 0 || v.nQ == 0 || !v.P || !v.Q || (v.nT && !v.T) || v.sT < -1 || v.sT > 1)
            return fail(PDH_E_ARG, "pdh_bs_merge_dev: every input needs P, Q > 0, T's buffer and a sign in {-1, 0, 1}");
    }
    const size_t pairs = n_in / 2, n_out = (n_in + 1) / 2;
    for (size_t i = 0; i < n_out; ++i)
        if (!out[i].Q || !out[i].T || (!skip_p && !out[i].P)) return fail(PDH_E_ARG, "pdh_bs_merge_dev: null output buffer");
    Ctx& c = ctx();
    {
        std::lock_guard<std::mutex> g(c.mu);
        int rc = init_locked(c);
        if (rc) return rc;
    }
    hipStream_t s = (hipStream_t)stream;
    auto eff_nT = [](const pdh_bs_node& v) { return v.sT ? v.nT : (uint64_t)0; };
    auto al256 = [](size_t x) { return (x + 255) & ~(size_t)255; };

    // scratch: B products | product lengths (4 per pair) | T lengths | T signs | add nodes | tile map | statuses
    std::vector<uint64_t> boff(pairs), capT(pairs);
    uint64_t bwords = 0, ntiles = 0;
    for (size_t i = 0; i < pairs; ++i) {
        const pdh_bs_node& L = in[2 * i];
        const pdh_bs_node& R = in[2 * i + 1];
        const uint64_t nA = eff_nT(L) ? eff_nT(L) + R.nQ : 0;
        const uint64_t nB = eff_nT(R) ? L.nP + eff_nT(R) : 0;
        boff[i] = bwords;
        bwords += (nB + 3) & ~(uint64_t)3;
        capT[i] = std::max(nA, nB) + 1;
        ntiles += (capT[i] + BS_ADD_TILE - 1) / BS_ADD_TILE;
    }
    if (ntiles > 0xFFFFFFFFull) return fail(PDH_E_TOO_LARGE, "pdh_bs_merge_dev: too many tiles");
    const size_t o_lens = al256(bwords * 4);
    const size_t o_olen = o_lens + al256(pairs * 4 * 8);
    const size_t o_osgn = o_olen + pairs * 8;
    const size_t o_nodes = al256(o_osgn + pairs * 4);
    const size_t o_tmap = o_nodes + al256(pairs * sizeof(BsAddNode));
    const size_t o_stat = o_tmap + al256(ntiles * 4);
    const size_t bytes = o_stat + al256(ntiles);
    char* ws = nullptr;
    pdh_batch* plan = nullptr;
    auto cleanup = [&]() {
        if (plan) pdh_batch_destroy(plan);
        if (ws) { (void)hipStreamSynchronize(s); cache_free(ws); }
    };
    if (pairs) {
        hipError_t e = cache_alloc((void**)&ws, bytes);
        if (e != hipSuccess) return fail_hip(e, "pdh_bs_merge_dev: scratch allocation");
    }
    uint32_t* dB = (uint32_t*)ws;
    uint64_t* dlens = (uint64_t*)(ws + o_lens);
    uint64_t* dolen = (uint64_t*)(ws + o_olen);
    int32_t* dosgn = (int32_t*)(ws + o_osgn);

    // products
    std::vector<pdh_mul_desc> jobs;
    jobs.reserve(4 * pairs);
    std::vector<char> stage(o_stat - o_nodes);
    BsAddNode* an = (BsAddNode*)stage.data();
    uint32_t* tmap = (uint32_t*)(stage.data() + (o_tmap - o_nodes));
    uint32_t tile = 0;
    for (size_t i = 0; i < pairs; ++i) {
        const pdh_bs_node& L = in[2 * i];
        const pdh_bs_node& R = in[2 * i + 1];
        uint64_t* li = dlens + 4 * i;
        jobs.push_back(pdh_mul_desc{L.Q, L.nQ, R.Q, R.nQ, out[i].Q, li + 0});
        jobs.push_back(pdh_mul_desc{L.T, eff_nT(L), R.Q, R.nQ, out[i].T, li + 1});
        jobs.push_back(pdh_mul_desc{L.P, L.nP, R.T, eff_nT(R), dB + boff[i], li + 2});
        if (!skip_p) jobs.push_back(pdh_mul_desc{L.P, L.nP, R.P, R.nP, out[i].P, li + 3});
        const uint32_t nt = (uint32_t)((capT[i] + BS_ADD_TILE - 1) / BS_ADD_TILE);
        an[i] = BsAddNode{out[i].T, out[i].T, dB + boff[i], li + 1, li + 2, capT[i], dolen + i, dosgn + i, tile, nt,
                          eff_nT(L) ? L.sT : 0, eff_nT(R) ? R.sT : 0};
        for (uint32_t k = 0; k < nt; ++k) tmap[tile + k] = (uint32_t)i;
        tile += nt;
    }
    int rc = PDH_OK;
    hipError_t e = hipSuccess;
    if (pairs) {
        e = hipMemcpyAsync(ws + o_nodes, stage.data(), stage.size(), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) { cleanup(); return fail_hip(e, "pdh_bs_merge_dev: node upload"); }
        // products beyond one transform (the top levels of very long runs)
        // go one by one through the blocked multiply, after the batch
        std::vector<pdh_mul_desc> small, big;
        for (const pdh_mul_desc& d : jobs) (d.na + d.nb > ((uint64_t)1 << 23) ? big : small).push_back(d);
        rc = batch_create_impl(small.data(), small.size(), &plan, true);
        if (rc == PDH_OK) rc = pdh_batch_run(plan, stream);
        if (rc != PDH_OK) { std::string m = pdh_last_error(); cleanup(); return fail(rc, m); }
        {
            std::lock_guard<std::mutex> g(c.mu);
            for (const pdh_mul_desc& d : big)
                if ((rc = mul_dev_locked(c, d.a, d.na, d.b, d.nb, false, d.out, d.out_len, s))) break;
            if (rc == PDH_OK)
                e = launch_bs_signed_add((const BsAddNode*)(ws + o_nodes), (const uint32_t*)(ws + o_tmap), tile,
                                         (uint8_t*)(ws + o_stat), s);
        }
        if (rc != PDH_OK) { std::string m = pdh_last_error(); cleanup(); return fail(rc, m); }
        if (e != hipSuccess) { cleanup(); return fail_hip(e, "pdh_bs_merge_dev: signed add"); }
    }
    if (n_in & 1) {
        const pdh_bs_node& v = in[n_in - 1];
        pdh_bs_node& o = out[n_out - 1];
        if (!skip_p) e = hipMemcpyAsync(o.P, v.P, v.nP * 4, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(o.Q, v.Q, v.nQ * 4, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess && eff_nT(v)) e = hipMemcpyAsync(o.T, v.T, eff_nT(v) * 4, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) { cleanup(); return fail_hip(e, "pdh_bs_merge_dev: copy of the odd node"); }
        o.nP = skip_p ? 0 : v.nP;
        o.nQ = v.nQ;
        o.nT = eff_nT(v);
        o.sT = o.nT ? v.sT : 0;
    }
    std::vector<char> back(o_osgn + pairs * 4 - o_lens);
    if (pairs) e = hipMemcpyAsync(back.data(), ws + o_lens, back.size(), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { cleanup(); return fail_hip(e, "pdh_bs_merge_dev: result lengths"); }
    const uint64_t* hl = (const uint64_t*)back.data();
    const uint64_t* ho = (const uint64_t*)(back.data() + (o_olen - o_lens));
    const int32_t* hs = (const int32_t*)(back.data() + (o_osgn - o_lens));
    for (size_t i = 0; i < pairs; ++i) {
        out[i].nQ = hl[4 * i];
        out[i].nP = skip_p ? 0 : hl[4 * i + 3];
        out[i].nT = ho[i];
        out[i].sT = ho[i] ? hs[i] : 0;
    }
    cleanup();
    return PDH_OK;
}

} // extern "C"

namespace {

// Uploads a node table (nodes, then the tile -> node map) into one device
// allocation through a pinned staging block, runs `launch` on it, and hands
// both blocks back once the stream has passed the launches (no host wait).
template <class Node, class Launch>
int run_tiled(const std::vector<Node>& nodes, const std::vector<uint32_t>& tmap, size_t extra, hipStream_t s,
              const char* what, Launch launch) {
    auto al256 = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_tmap = al256(nodes.size() * sizeof(Node));
    const size_t o_extra = o_tmap + al256(tmap.size() * 4);
    char* ws = nullptr;
    void* pin = nullptr;
    hipError_t e = cache_alloc((void**)&ws, o_extra + al256(extra));
    if (e != hipSuccess) return fail_hip(e, what);
    e = pin_alloc(&pin, o_extra);
    if (e != hipSuccess) { cache_free(ws); return fail_hip(e, what); }
    std::memcpy(pin, nodes.data(), nodes.size() * sizeof(Node));
    std::memcpy((char*)pin + o_tmap, tmap.data(), tmap.size() * 4);
    e = hipMemcpyAsync(ws, pin, o_extra, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        Ctx& c = ctx();
        std::lock_guard<std::mutex> g(c.mu);
        e = launch((const Node*)ws, (const uint32_t*)(ws + o_tmap), (uint32_t)tmap.size(), ws + o_extra);
    }
    const hipError_t e2 = pin_free_after(pin, s);
    const hipError_t e3 = cache_free_after(ws, s);
    if (e == hipSuccess) e = e2 != hipSuccess ? e2 : e3;
    return e == hipSuccess ? PDH_OK : fail_hip(e, what);
}

int dev_ready() {
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    return init_locked(c);
}

} // namespace

extern "C" {

int pdh_sadd_batch_dev(const pdh_sadd_desc* d, size_t n, void* stream) {
    if (n == 0) return PDH_OK;
    if (!d) return fail(PDH_E_ARG, "null pointer argument");
    uint64_t ntiles = 0;
    for (size_t i = 0; i < n; ++i) {
        const pdh_sadd_desc& x = d[i];
        if (x.sa < -1 || x.sa > 1 || x.sb < -1 || x.sb > 1 || (x.sa && (!x.a || !x.a_len)) ||
            (x.sb && (!x.b || !x.b_len)) || !x.out || x.out_cap == 0 || !x.out_len || !x.out_sign)
            return fail(PDH_E_ARG, "pdh_sadd_batch_dev: bad descriptor");
        if ((x.sa && x.out_len == x.a_len) || (x.sb && x.out_len == x.b_len))
            return fail(PDH_E_ARG, "pdh_sadd_batch_dev: out_len must not be an operand's length word");
        ntiles += (x.out_cap + BS_ADD_TILE - 1) / BS_ADD_TILE;
    }
    if (ntiles > 0xFFFFFFFFull) return fail(PDH_E_TOO_LARGE, "pdh_sadd_batch_dev: too many tiles");
    int rc = dev_ready();
    if (rc) return rc;
    std::vector<BsAddNode> nodes(n);
    std::vector<uint32_t> tmap;
    tmap.reserve(ntiles);
    for (size_t i = 0; i < n; ++i) {
        const pdh_sadd_desc& x = d[i];
        const uint32_t nt = (uint32_t)((x.out_cap + BS_ADD_TILE - 1) / BS_ADD_TILE);
        nodes[i] = BsAddNode{x.out, x.a, x.b, x.a_len, x.b_len, x.out_cap, x.out_len, x.out_sign,
                             (uint32_t)tmap.size(), nt, x.sa, x.sb};
        tmap.insert(tmap.end(), nt, (uint32_t)i);
    }
    hipStream_t s = (hipStream_t)stream;
    return run_tiled(nodes, tmap, tmap.size(), s, "pdh_sadd_batch_dev",
                     [&](const BsAddNode* dn, const uint32_t* dt, uint32_t nt, char* extra) {
                         return launch_bs_signed_add(dn, dt, nt, (uint8_t*)extra, s);
                     });
}

int pdh_shr_batch_dev(const pdh_shr_desc* d, size_t n, void* stream) {
    if (n == 0) return PDH_OK;
    if (!d) return fail(PDH_E_ARG, "null pointer argument");
    uint64_t ntiles = 0;
    for (size_t i = 0; i < n; ++i) {
        const pdh_shr_desc& x = d[i];
        if (!x.a || !x.a_len || !x.out || x.out_cap == 0 || !x.out_len)
            return fail(PDH_E_ARG, "pdh_shr_batch_dev: bad descriptor");
        ntiles += (x.out_cap + SHR_TILE - 1) / SHR_TILE;
    }
    if (ntiles > 0xFFFFFFFFull) return fail(PDH_E_TOO_LARGE, "pdh_shr_batch_dev: too many tiles");
    int rc = dev_ready();
    if (rc) return rc;
    std::vector<ShiftNode> nodes(n);
    std::vector<uint32_t> tmap;
    tmap.reserve(ntiles);
    for (size_t i = 0; i < n; ++i) {
        const pdh_shr_desc& x = d[i];
        const uint32_t nt = (uint32_t)((x.out_cap + SHR_TILE - 1) / SHR_TILE);
        nodes[i] = ShiftNode{x.a, x.a_len, x.out, x.out_len, x.out_cap, x.bits, (uint32_t)tmap.size(), nt, 0u};
        tmap.insert(tmap.end(), nt, (uint32_t)i);
    }
    hipStream_t s = (hipStream_t)stream;
    return run_tiled(nodes, tmap, 0, s, "pdh_shr_batch_dev",
                     [&](const ShiftNode* dn, const uint32_t* dt, uint32_t nt, char*) {
                         return launch_shift_batch(dn, dt, nt, s);
                     });
}

int pdh_sadd_dev(const uint32_t* a, const uint64_t* a_len, int32_t sa, const uint32_t* b, const uint64_t* b_len,
                 int32_t sb, uint32_t* out, uint64_t out_cap, uint64_t* out_len, int32_t* out_sign, void* stream) {
    if (sa < -1 || sa > 1 || sb < -1 || sb > 1 || (sa && (!a || !a_len)) || (sb && (!b || !b_len)) || !out ||
        out_cap == 0 || !out_len || !out_sign)
        return fail(PDH_E_ARG, "pdh_sadd_dev: bad argument");
    if ((sa && out_len == a_len) || (sb && out_len == b_len))
        return fail(PDH_E_ARG, "pdh_sadd_dev: out_len must not be an operand's length word");
    const uint64_t nt = (out_cap + BS_ADD_TILE - 1) / BS_ADD_TILE;
    if (nt > 0xFFFFFFFFull) return fail(PDH_E_TOO_LARGE, "pdh_sadd_dev: too many tiles");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = init_locked(c);
    if (rc) return rc;
    if ((rc = grow(&c.sadd_status, &c.sadd_status_n, nt, 1))) return rc;
    const BsAddNode one{out, a, b, a_len, b_len, out_cap, out_len, out_sign, 0u, (uint32_t)nt, sa, sb};
    if ((rc = scratch_acquire(c.sadd_order, (hipStream_t)stream))) return rc;
    hipError_t e = launch_bs_signed_add(nullptr, nullptr, (uint32_t)nt, c.sadd_status, (hipStream_t)stream, &one);
    if (e != hipSuccess) return fail_hip(e, "signed sum launch");
    return scratch_release(c.sadd_order, (hipStream_t)stream);
}

int pdh_shift_dev(const uint32_t* a, const uint64_t* a_len, int64_t bits, uint32_t* out, uint64_t out_cap,
                  uint64_t* out_len, void* stream) {
    if (!a || !a_len || !out || out_cap == 0 || !out_len) return fail(PDH_E_ARG, "pdh_shift_dev: bad argument");
    const uint64_t nt = (out_cap + SHR_TILE - 1) / SHR_TILE;
    if (nt > 0xFFFFFFFFull) return fail(PDH_E_TOO_LARGE, "pdh_shift_dev: too many tiles");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = init_locked(c);
    if (rc) return rc;
    const uint64_t mag = bits < 0 ? (uint64_t)0 - (uint64_t)bits : (uint64_t)bits;
    const ShiftNode one{a, a_len, out, out_len, out_cap, mag, 0u, (uint32_t)nt, bits < 0 ? 0u : 1u};
    hipError_t e = launch_shift_batch(nullptr, nullptr, (uint32_t)nt, (hipStream_t)stream, &one);
    return e == hipSuccess ? PDH_OK : fail_hip(e, "shift launch");
}

int pdh_decimal_leaves_dev(uint32_t* d_x, uint64_t stride, const uint64_t* d_len, uint64_t n, uint32_t width,
                           char* d_out, void* stream) {
    if (n == 0) return PDH_OK;
    if (!d_x || !d_len || !d_out) return fail(PDH_E_ARG, "null pointer argument");
    if (width == 0 || width % 9 || stride == 0) return fail(PDH_E_ARG, "pdh_decimal_leaves_dev: width must be a positive multiple of 9, stride positive");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = init_locked(c);
    if (rc) return rc;
    hipError_t e = launch_decimal_leaves(d_x, stride, d_len, n, width, d_out, (hipStream_t)stream);
    return e == hipSuccess ? PDH_OK : fail_hip(e, "decimal leaves launch");
}

int pdh_dev_alloc(size_t bytes, void** ptr) {
    if (!ptr) return fail(PDH_E_ARG, "null pointer argument");
    *ptr = nullptr;
    Ctx& c = ctx();
    { std::lock_guard<std::mutex> g(c.mu); int rc = init_locked(c); if (rc) return rc; }
    if (bytes == 0) return PDH_OK;
    HIPCALL(cache_alloc(ptr, bytes));
    return PDH_OK;
}
int pdh_dev_free(void* ptr) {
    if (!ptr) return PDH_OK;
    HIPCALL(hipDeviceSynchronize());       // as hipFree: no work may still use the block
    cache_free(ptr);
    return PDH_OK;
}
int pdh_dev_free_async(void* ptr, void* stream) {
    if (!ptr) return PDH_OK;
    HIPCALL(cache_free_after(ptr, (hipStream_t)stream));
    return PDH_OK;
}
int pdh_copy_htod(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return PDH_OK;
    if (!dst || !src) return fail(PDH_E_ARG, "null pointer argument");
    if (stream) HIPCALL(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    else HIPCALL(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return PDH_OK;
}
int pdh_copy_dtoh(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return PDH_OK;
    if (!dst || !src) return fail(PDH_E_ARG, "null pointer argument");
    if (stream) HIPCALL(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    else HIPCALL(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return PDH_OK;
}
int pdh_stream_create(void** stream) {
    if (!stream) return fail(PDH_E_ARG, "null pointer argument");
    Ctx& c = ctx();
    { std::lock_guard<std::mutex> g(c.mu); int rc = init_locked(c); if (rc) return rc; }
    hipStream_t s;
    HIPCALL(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return PDH_OK;
}
int pdh_stream_destroy(void* stream) {
    if (stream) HIPCALL(hipStreamDestroy((hipStream_t)stream));
    return PDH_OK;
}
int pdh_stream_sync(void* stream) {
    HIPCALL(hipStreamSynchronize((hipStream_t)stream));
    return PDH_OK;
}
int pdh_stream_query(void* stream) {
    const hipError_t e = hipStreamQuery((hipStream_t)stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) { (void)hipGetLastError(); return 0; }
    return fail_hip(e, "hipStreamQuery");
}

int pdh_tune_get(int n, int* L1, int* T, int* lr_col, int* lr_row) {
    if (n < 1 || n > MAX_LOG || !L1 || !T || !lr_col || !lr_row) return fail(PDH_E_ARG, "bad argument");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    NttShape sh = ntt_shape(n);
    *L1 = sh.L1; *T = sh.T; *lr_col = sh.lr_col; *lr_row = sh.lr_row;
    return PDH_OK;
}

int pdh_tune_set(int n, int L1, int T, int lr_col, int lr_row) {
    if (n < 1 || n > MAX_LOG) return fail(PDH_E_ARG, "bad argument");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    const int eng = ntt_shape(n).eng;
    NttShape sh{n, L1, n - eng - L1, T, lr_col, lr_row, eng};
    if (L1 == 0) { sh.L2 = n - eng; sh.T = 1; sh.lr_col = 3; }
    if (!set_shape(sh)) return fail(PDH_E_ARG, "shape cannot be launched");
    return PDH_OK;
}

int pdh_tune_engine(int n, int engine, int* current) {
    if (n < 1 || n > MAX_LOG || engine < -1 || engine > 1) return fail(PDH_E_ARG, "bad argument");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    if (engine >= 0 && engine != ntt_shape(n).eng) {
        if (n < 2) return fail(PDH_E_ARG, "no product has N = 2");
        set_shape(default_shape_for(n, engine), engine != default_shape_for(n, -1).eng);
    }
    if (current) *current = ntt_shape(n).eng;
    return PDH_OK;
}

int pdh_autotune(int n_lo, int n_hi, int reps, void* stream) {
    if (n_lo < 2 || n_hi > MAX_LOG || n_lo > n_hi || reps < 1) return fail(PDH_E_ARG, "bad autotune range");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    int rc = init_locked(c);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const size_t maxlimbs = (size_t)1 << (n_hi - 2);
    u32 *a = nullptr, *b = nullptr, *out = nullptr;
    u64* len = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
        if (out) (void)hipFree(out);
        if (len) (void)hipFree(len);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    hipError_t e = hipMalloc(&a, maxlimbs * 4);
    if (e == hipSuccess) e = hipMalloc(&b, maxlimbs * 4);
    if (e == hipSuccess) e = hipMalloc(&out, maxlimbs * 8);
    if (e == hipSuccess) e = hipMalloc(&len, 8);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        // distinct operands: with a == b the column pass reads every limb
        // twice from cache and favours other shapes than a real product
        std::vector<u32> h(maxlimbs);
        uint64_t x = 0x9e3779b97f4a7c15ULL;
        for (auto& v : h) { x ^= x << 13; x ^= x >> 7; x ^= x << 17; v = (u32)x | 1u; }
        e = hipMemcpy(a, h.data(), maxlimbs * 4, hipMemcpyHostToDevice);
        for (auto& v : h) { x ^= x << 13; x ^= x >> 7; x ^= x << 17; v = (u32)x | 1u; }
        if (e == hipSuccess) e = hipMemcpy(b, h.data(), maxlimbs * 4, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) { cleanup(); return fail_hip(e, "autotune setup"); }
    for (int n = n_lo; n <= n_hi; ++n) {
        const size_t na = (size_t)1 << (n - 2);
        std::vector<NttShape> cands;
        const NttShape def = ntt_shape(n);
        if (n <= SMALL_MAX_LOG) {
            // one workgroup: Goldilocks at either radix, or the three primes
            for (int lr = 3; lr <= 4; ++lr) cands.push_back(NttShape{n, 0, n, 1, 3, lr, 0});
            cands.push_back(default_shape_for(n, 1));
        } else {
            for (int L1 = def.L1 - 2; L1 <= def.L1 + 2; ++L1)
                for (int T : {4, 8, 16, 32, 64})
                    for (int lc = 3; lc <= 4; ++lc)
                        for (int lr = 3; lr <= 4; ++lr) {
                            NttShape sh{n, L1, n - def.eng - L1, T, lc, lr, def.eng};
                            if (shape_valid(sh)) cands.push_back(sh);
                        }
        }
        NttShape best = def;
        float best_ms = 1e30f;
        for (const NttShape& sh : cands) {
            set_shape(sh);
            rc = mul_dev_locked(c, a, na, b, na, false, out, len, s);          // warm-up
            if (rc) { set_shape(def); cleanup(); return rc; }
            e = hipEventRecord(e0, s);
            for (int r = 0; r < reps && e == hipSuccess; ++r)
                if ((rc = mul_dev_locked(c, a, na, b, na, false, out, len, s))) { set_shape(def); cleanup(); return rc; }
            if (e == hipSuccess) e = hipEventRecord(e1, s);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) { set_shape(def); cleanup(); return fail_hip(e, "autotune timing"); }
            if (ms < best_ms) { best_ms = ms; best = sh; }
        }
        set_shape(best);
    }
    cleanup();
    return PDH_OK;
}

// Shape files: a first line "pdh-shapes 1 gfx950", then one line per
// transform size, "n eng L1 L2 T lr_col lr_row custom", for n = 1..MAX_LOG.
// A file is read whole and checked (every shape must be launchable and name
// its own n once) before any shape changes, so a bad file changes nothing.
static const char kShapeMagic[] = "pdh-shapes 1 gfx950";

int pdh_tune_save(const char* path) {
    if (!path) return fail(PDH_E_ARG, "bad argument");
    Ctx& c = ctx();
    std::string text = std::string(kShapeMagic) + "\n";
    {
        std::lock_guard<std::mutex> g(c.mu);
        for (int n = 1; n <= MAX_LOG; ++n) {
            const NttShape s = ntt_shape(n);
            char line[96];
            std::snprintf(line, sizeof line, "%d %d %d %d %d %d %d %d\n", s.n, s.eng, s.L1, s.L2, s.T,
                          s.lr_col, s.lr_row, shape_custom(n) ? 1 : 0);
            text += line;
        }
    }
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(PDH_E_ARG, "cannot open the shape file for writing");
    const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
    if (std::fclose(f) != 0 || !ok) return fail(PDH_E_ARG, "cannot write the shape file");
    return PDH_OK;
}

int pdh_tune_load(const char* path) {
    if (!path) return fail(PDH_E_ARG, "bad argument");
    FILE* f = std::fopen(path, "r");
    if (!f) return fail(PDH_E_ARG, "cannot open the shape file");
    char line[256];
    bool ok = std::fgets(line, sizeof line, f) != nullptr && std::strncmp(line, kShapeMagic, sizeof kShapeMagic - 1) == 0 &&
              (line[sizeof kShapeMagic - 1] == '\n' || line[sizeof kShapeMagic - 1] == 0);
    std::vector<std::pair<NttShape, bool>> shapes;
    bool seen[MAX_LOG + 1] = {};
    while (ok && std::fgets(line, sizeof line, f)) {
        NttShape s{};
        int custom = 0, used = 0;
        if (std::sscanf(line, "%d %d %d %d %d %d %d %d %n", &s.n, &s.eng, &s.L1, &s.L2, &s.T, &s.lr_col,
                        &s.lr_row, &custom, &used) != 8 || line[used] != 0 || (custom != 0 && custom != 1) ||
            !shape_valid(s) || seen[s.n]) { ok = false; break; }
        seen[s.n] = true;
        shapes.push_back({s, custom == 1});
    }
    std::fclose(f);
    if (!ok) return fail(PDH_E_ARG, "shape file is not a pdh-shapes 1 gfx950 file of launchable shapes");
    Ctx& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    for (const auto& [s, custom] : shapes) set_shape(s, custom);
    return PDH_OK;
}

int pdh_timing_enable(int on) {
    Timing& t = timing();
    std::lock_guard<std::mutex> g(t.mu);
    t.recycle();
    t.on = on != 0;
    return PDH_OK;
}

int pdh_timing_read(char* names, size_t names_cap, double* total_ms, uint64_t* counts,
                    int max_kernels, int* n_kernels) {
    if (!names || !total_ms || !counts || !n_kernels || max_kernels <= 0) return fail(PDH_E_ARG, "null argument");
    Timing& t = timing();
    std::lock_guard<std::mutex> g(t.mu);
    std::vector<std::string> order;
    std::map<std::string, std::pair<double, uint64_t>> acc;
    for (auto& r : t.recs) {
        HIPCALL(hipEventSynchronize(r.b));
        float ms = 0.f;
        HIPCALL(hipEventElapsedTime(&ms, r.a, r.b));
        auto it = acc.find(r.name);
        if (it == acc.end()) { order.push_back(r.name); it = acc.emplace(r.name, std::make_pair(0.0, (uint64_t)0)).first; }
        it->second.first += ms;
        it->second.second += 1;
    }
    t.recycle();
    std::string joined;
    int k = 0;
    for (auto& nm : order) {
        if (k >= max_kernels) break;
        total_ms[k] = acc[nm].first;
        counts[k] = acc[nm].second;
        joined += nm;
        joined += '\n';
        ++k;
    }
    if (joined.size() + 1 > names_cap) return fail(PDH_E_CAPACITY, "names buffer too small");
    std::memcpy(names, joined.c_str(), joined.size() + 1);
    *n_kernels = k;
    return PDH_OK;
}

} // extern "C"
