// ============================================================================
//  serve_hip.cpp -- a continuous-batching request loop on one GPU.
//
//  Requests (multiply, Chudnovsky pi, BBP hex digits) arrive at any time and
//  wait until the HBM budget and the in-flight limit admit them. Admission
//  goes in arrival order, but a request that does not fit the budget lets
//  later, smaller requests pass it (backfill) until it has waited
//  kStarveMs; from then on nothing passes it, so it is admitted as soon as
//  the requests ahead of it release enough memory. With nothing running, a
//  request is admitted over the budget (so that one larger than the budget,
//  or a client that waits on it while holding unfetched results, cannot
//  stall the server) at once if nothing is held, else after kIdleOverMs, the
//  time the client has to fetch the results whose products it still holds.
//  Admitted requests start while earlier ones are still on the device:
//
//  * multiplies admitted in one scheduler tick whose transform is of
//    moderate size (N <= 2^18) are grouped into one batched plan
//    (pdh_batch_*), larger ones get a plan each; every plan has its own
//    workspace and runs on its own stream from the server's pool, so their
//    kernels overlap on the device. Operands and products stay in HBM until
//    the result is fetched.
//  * BBP requests go to kBbpWorkers host workers, Chudnovsky requests to
//    kPiWorkers; each worker's device work runs on its own streams (the
//    library orders the shared scratch between streams on the device), so
//    they overlap with each other and with the multiplies.
//
//  A completion thread polls the running plans' streams and completes each
//  request as soon as its own stream drains, in any order. Memory is given
//  back with stream-ordered frees (no device-wide synchronisation). Latency
//  is completion minus submission. The reference has no serving layer: it
//  computes one request per process run (src/main.cpp:93-179).
// ============================================================================
#include "pidigits_hip.h"
#include "pdh_capi.h"
#include "pdh_pi.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

using clk = std::chrono::steady_clock;

enum Kind { MUL = 1, PI = 2, BBP = 3 };
enum State { QUEUED = 0, RUNNING = 1, DONE = 2, FAILED = 3 };

constexpr int kBbpWorkers = 2;
constexpr int kPiWorkers = 2;
constexpr double kStarveMs = 20.0;          // a request that waited this long stops being passed
constexpr double kIdleOverMs = 50.0;        // nothing running, results held: wait this long before going over budget
constexpr size_t kSmallN = (size_t)1 << 18; // multiplies up to this transform length are grouped

struct Request {
    uint64_t id = 0;
    Kind kind = MUL;
    clk::time_point t_arrival, t_done;
    std::atomic<int> state{QUEUED};
    int err = 0;
    size_t bytes = 0;              // admission estimate of device memory held while running
    size_t resident = 0;           // of which held until the result is fetched (multiply arena)
    bool waited_mem = false;       // was passed over at least once for lack of budget
    // multiply
    std::vector<uint32_t> a, b;
    void* d_arena = nullptr;       // a | b | out | len, resident until fetched
    size_t na = 0, nb = 0;
    bool empty_operand = false;
    // pi
    uint64_t digits = 0;
    std::string text;
    // bbp
    uint64_t start = 0;
    uint32_t count = 0;
    std::vector<uint8_t> values;
};

size_t mul_arena_bytes(size_t na, size_t nb) { return 8 * (na + nb) + 16; }

size_t mul_bytes(size_t na, size_t nb) {
    const size_t N = pdh_ntt_length(na, nb);
    return mul_arena_bytes(na, nb) + 16 * N + (1u << 20);     // operands+product, two spectra, scratch
}

struct Job {
    void* stream = nullptr;
    pdh_batch* plan = nullptr;
    std::vector<Request*> reqs;
};

uint32_t* arena_out(Request* r) { return (uint32_t*)r->d_arena + r->na + r->nb; }
uint64_t* arena_len(Request* r) {
    return (uint64_t*)(((uintptr_t)(arena_out(r) + r->na + r->nb) + 7) & ~(uintptr_t)7);
}

} // namespace

struct pdh_server {
    size_t budget = 0, inuse = 0, peak = 0;
    int max_inflight = 0, inflight = 0;
    clk::time_point idle_since = clk::now();   // when inflight last fell to 0
    uint64_t next_id = 1;
    uint64_t n_submitted = 0, n_done = 0, n_waited_mem = 0, n_backfilled = 0;
    bool stop = false;
    std::mutex mu;
    std::condition_variable cv_sched, cv_done, cv_jobs, cv_work;
    std::deque<Request*> pending;
    std::map<uint64_t, std::unique_ptr<Request>> all;
    std::list<Job> running;                 // multiply plans on the device
    std::deque<Request*> pi_queue, bbp_queue;
    std::vector<void*> stream_pool;
    std::thread sched, completer;
    std::vector<std::thread> workers;

    // caller holds mu
    void release_locked(size_t bytes) { inuse -= std::min(inuse, bytes); cv_sched.notify_all(); }

    void finish(Request* r, int state, size_t release) {
        std::lock_guard<std::mutex> g(mu);
        r->t_done = clk::now();
        r->state = state;
        release_locked(release);
        if (--inflight == 0) idle_since = r->t_done;
        ++n_done;
        cv_done.notify_all();
    }

    void* get_stream() {
        {
            std::lock_guard<std::mutex> g(mu);
            if (!stream_pool.empty()) { void* s = stream_pool.back(); stream_pool.pop_back(); return s; }
        }
        void* s = nullptr;
        return pdh_stream_create(&s) == PDH_OK ? s : nullptr;
    }
    void put_stream(void* s) {
        std::lock_guard<std::mutex> g(mu);
        stream_pool.push_back(s);
    }

    // A failed multiply gives back everything it holds, its arena included.
    void fail_mul(Request* r, int rc, void* stream) {
        if (r->d_arena) pdh_dev_free_async(r->d_arena, stream);
        r->d_arena = nullptr;
        r->err = rc;
        finish(r, FAILED, r->bytes);
    }

    void launch_muls(std::vector<Request*>& rs) {
        if (rs.empty()) return;
        Job job;
        job.stream = get_stream();
        int rc = job.stream ? PDH_OK : PDH_E_HIP;
        std::vector<pdh_mul_desc> d;
        for (size_t i = 0; rc == PDH_OK && i < rs.size(); ++i) {
            Request* r = rs[i];
            rc = pdh_dev_alloc(mul_arena_bytes(r->na, r->nb), &r->d_arena);
            if (rc) break;
            uint32_t* A = (uint32_t*)r->d_arena;
            uint32_t* B = A + r->na;
            rc = pdh_copy_htod(A, r->a.data(), 4 * r->na, job.stream);
            if (!rc) rc = pdh_copy_htod(B, r->b.data(), 4 * r->nb, job.stream);
            d.push_back(pdh_mul_desc{A, r->na, B, r->nb, arena_out(r), arena_len(r)});
        }
        if (!rc) rc = pdh_batch_create(d.data(), d.size(), &job.plan);
        if (!rc) rc = pdh_batch_run(job.plan, job.stream);
        if (rc) {
            if (job.plan) pdh_batch_destroy(job.plan);
            for (Request* r : rs) fail_mul(r, rc, job.stream);
            if (job.stream) put_stream(job.stream);
            return;
        }
        job.reqs = rs;
        std::lock_guard<std::mutex> g(mu);
        running.push_back(std::move(job));
        cv_jobs.notify_all();
    }

    // One scheduler tick: admit while slots and budget allow (see the file
    // comment for the order), then start everything admitted.
    void sched_loop() {
        for (;;) {
            std::vector<Request*> adm;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_sched.wait_for(lk, std::chrono::microseconds(500), [&] { return stop || !pending.empty(); });
                if (stop) return;
                const clk::time_point now = clk::now();
                bool blocked = false;
                for (auto it = pending.begin(); it != pending.end() && inflight < max_inflight;) {
                    Request* r = *it;
                    const double waited = std::chrono::duration<double, std::milli>(now - r->t_arrival).count();
                    const double idle = std::chrono::duration<double, std::milli>(now - idle_since).count();
                    const bool fits = inuse + r->bytes <= budget ||
                                      (inflight == 0 && (inuse == 0 || idle >= kIdleOverMs));
                    if (!fits) {
                        if (!r->waited_mem) { r->waited_mem = true; ++n_waited_mem; }
                        if (waited >= kStarveMs) break;        // no more passing this one
                        blocked = true;
                        ++it;
                        continue;
                    }
                    if (blocked) ++n_backfilled;
                    it = pending.erase(it);
                    ++inflight;
                    inuse += r->bytes;
                    peak = std::max(peak, inuse);
                    r->state = RUNNING;
                    adm.push_back(r);
                }
            }
            std::vector<Request*> small;
            for (Request* r : adm) {
                if (r->kind == MUL) {
                    if (pdh_ntt_length(r->na, r->nb) <= kSmallN) small.push_back(r);
                    else { std::vector<Request*> one{r}; launch_muls(one); }
                } else {
                    std::lock_guard<std::mutex> g(mu);
                    (r->kind == BBP ? bbp_queue : pi_queue).push_back(r);
                    cv_work.notify_all();
                }
            }
            launch_muls(small);
        }
    }

    // Completes each plan as soon as its stream has drained, in any order.
    void complete_loop() {
        for (;;) {
            std::vector<Job> done;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_jobs.wait_for(lk, std::chrono::microseconds(200), [&] { return stop || !running.empty(); });
                if (running.empty()) { if (stop) return; continue; }
                for (auto it = running.begin(); it != running.end();) {
                    const int q = pdh_stream_query(it->stream);
                    if (q == 0) { ++it; continue; }
                    done.push_back(std::move(*it));
                    it = running.erase(it);
                }
            }
            if (done.empty()) { std::this_thread::sleep_for(std::chrono::microseconds(20)); continue; }
            for (Job& job : done) {
                const int rc = pdh_stream_query(job.stream) == 1 ? PDH_OK : pdh_stream_sync(job.stream);
                if (job.plan) pdh_batch_destroy(job.plan);
                for (Request* r : job.reqs) {
                    if (rc) { fail_mul(r, rc, job.stream); continue; }
                    finish(r, DONE, r->bytes - r->resident);   // the arena stays until the result is fetched
                }
                put_stream(job.stream);
            }
        }
    }

    void work_loop(Kind kind) {
        for (;;) {
            Request* r;
            {
                std::unique_lock<std::mutex> lk(mu);
                std::deque<Request*>& q = kind == BBP ? bbp_queue : pi_queue;
                cv_work.wait(lk, [&] { return stop || !q.empty(); });
                if (q.empty()) return;
                r = q.front();
                q.pop_front();
            }
            int st = DONE;
            try {
                if (kind == BBP) {
                    r->values.resize(r->count);
                    const int rc = pdh_bbp_range(r->start, r->count, r->values.data());
                    if (rc) { r->err = rc; st = FAILED; }
                } else {
                    r->text = pidigits::compute_pi_chudnovsky_hip(r->digits);
                }
            } catch (const pidigits::HipError& e) {
                r->err = e.code;
                st = FAILED;
            }
            finish(r, st, r->bytes);
        }
    }

    uint64_t submit(std::unique_ptr<Request> r) {
        std::lock_guard<std::mutex> g(mu);
        r->id = next_id++;
        r->t_arrival = clk::now();
        Request* raw = r.get();
        all[raw->id] = std::move(r);
        pending.push_back(raw);
        ++n_submitted;
        cv_sched.notify_all();
        return raw->id;
    }
};

extern "C" {

int pdh_server_create(uint64_t hbm_budget_bytes, int max_inflight, pdh_server** srv) {
    if (!srv || max_inflight < 1) return PDH_E_ARG;
    if (!pdh_device_available()) return PDH_E_NODEVICE;
    pdh_server* s = new pdh_server();
    s->budget = hbm_budget_bytes;
    s->max_inflight = max_inflight;
    s->sched = std::thread([s] { s->sched_loop(); });
    s->completer = std::thread([s] { s->complete_loop(); });
    for (int i = 0; i < kBbpWorkers; ++i) s->workers.emplace_back([s] { s->work_loop(BBP); });
    for (int i = 0; i < kPiWorkers; ++i) s->workers.emplace_back([s] { s->work_loop(PI); });
    *srv = s;
    return PDH_OK;
}

int pdh_server_submit_mul(pdh_server* s, const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint64_t* id) {
    if (!s || !id || (na && !a) || (nb && !b)) return PDH_E_ARG;
    if (na + nb > ((size_t)1 << 23)) return PDH_E_TOO_LARGE;
    std::unique_ptr<Request> r(new Request());
    r->kind = MUL;
    r->a.assign(a, a + na);
    r->b.assign(b, b + nb);
    r->na = na; r->nb = nb;
    if (na == 0 || nb == 0) {                  // zero product: a one-limb stand-in, result reported empty
        r->a.assign(1, 0); r->b.assign(1, 0); r->na = r->nb = 1;
        r->empty_operand = true;
    }
    r->bytes = mul_bytes(r->na, r->nb);
    r->resident = mul_arena_bytes(r->na, r->nb);
    *id = s->submit(std::move(r));
    return PDH_OK;
}

int pdh_server_submit_pi(pdh_server* s, uint64_t digits, uint64_t* id) {
    if (!s || !id) return PDH_E_ARG;
    std::unique_ptr<Request> r(new Request());
    r->kind = PI;
    r->digits = digits;
    r->bytes = 64 * (size_t)digits + ((size_t)1 << 28);
    *id = s->submit(std::move(r));
    return PDH_OK;
}

int pdh_server_submit_bbp(pdh_server* s, uint64_t start, uint32_t count, uint64_t* id) {
    if (!s || !id || start < 1 || count == 0 || start - 1 + count > ((uint64_t)1 << 45)) return PDH_E_ARG;
    std::unique_ptr<Request> r(new Request());
    r->kind = BBP;
    r->start = start;
    r->count = count;
    r->bytes = ((size_t)1 << 28);              // the BBP scratch of a range, bounded
    *id = s->submit(std::move(r));
    return PDH_OK;
}

int pdh_server_wait(pdh_server* s, uint64_t id, double* latency_ms) {
    if (!s) return PDH_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return PDH_E_ARG;
    Request* r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->state == DONE || r->state == FAILED; });
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_arrival).count();
    return r->state == DONE ? PDH_OK : (r->err ? r->err : PDH_E_HIP);
}

// Copies the result (multiply: normalized u32 limbs; pi: "3.1415..."; BBP:
// digit values) and releases the request and its resident buffers, also
// when the request failed (its error code is returned).
int pdh_server_result(pdh_server* s, uint64_t id, void* out, size_t cap, size_t* len) {
    if (!s || !len) return PDH_E_ARG;
    const int wrc = pdh_server_wait(s, id, nullptr);
    std::unique_ptr<Request> r;
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->all.find(id);
        if (it == s->all.end()) return PDH_E_ARG;
        r = std::move(it->second);
        s->all.erase(it);
    }
    if (wrc) { *len = 0; return wrc; }
    if (r->kind == MUL) {
        uint64_t nl = 0;
        int rc = pdh_copy_dtoh(&nl, arena_len(r.get()), 8, nullptr);
        if (r->empty_operand) nl = 0;
        *len = (size_t)nl;
        if (!rc && nl > cap) rc = PDH_E_CAPACITY;
        if (!rc && nl) rc = pdh_copy_dtoh(out, arena_out(r.get()), 4 * nl, nullptr);
        pdh_dev_free_async(r->d_arena, nullptr);       // its stream drained before the request completed
        std::lock_guard<std::mutex> g(s->mu);
        s->release_locked(r->resident);
        return rc;
    }
    const void* src = r->kind == PI ? (const void*)r->text.data() : (const void*)r->values.data();
    const size_t n = r->kind == PI ? r->text.size() : r->values.size();
    *len = n;
    if (n > cap) return PDH_E_CAPACITY;
    std::memcpy(out, src, n);
    return PDH_OK;
}

int pdh_server_stats(pdh_server* s, uint64_t* stats, int n) {
    if (!s || !stats || n < 1) return PDH_E_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    const uint64_t v[7] = {s->n_submitted, s->n_done, s->n_waited_mem, s->n_backfilled, (uint64_t)s->peak,
                           (uint64_t)s->inuse, (uint64_t)s->budget};
    for (int i = 0; i < n && i < 7; ++i) stats[i] = v[i];
    return PDH_OK;
}

int pdh_server_destroy(pdh_server* s) {
    if (!s) return PDH_OK;
    {
        // requests still waiting for admission are dropped; admitted ones finish
        std::unique_lock<std::mutex> lk(s->mu);
        s->pending.clear();
        s->cv_done.wait(lk, [&] { return s->inflight == 0; });
        s->stop = true;
        s->cv_sched.notify_all();
        s->cv_jobs.notify_all();
        s->cv_work.notify_all();
    }
    s->sched.join();
    s->completer.join();
    for (auto& t : s->workers) t.join();
    for (auto& kv : s->all) if (kv.second->d_arena) pdh_dev_free(kv.second->d_arena);
    for (void* st : s->stream_pool) pdh_stream_destroy(st);
    delete s;
    return PDH_OK;
}

} // extern "C"
