// ============================================================================
//  serve_hip.cpp -- a continuous-batching request loop on one GPU.
//
//  Requests (multiply, Chudnovsky pi, BBP hex digits) arrive at any time and
//  wait in FIFO order until the HBM budget and the in-flight limit admit
//  them; admitted requests start while earlier ones are still on the device:
//
//  * multiplies admitted in one scheduler tick whose transform fits one
//    workgroup class of moderate size (N <= 2^18) are grouped into one
//    batched plan (pdh_batch_*), larger ones get a plan each; every plan has
//    its own workspace and stream, so their kernels overlap on the device.
//    Operands and products stay resident in HBM until the result is fetched.
//  * BBP requests run in arrival order on one BBP stream (they share the
//    library's BBP workspace), overlapping with the multiplies.
//  * Chudnovsky requests run on a host worker thread whose multiplies and
//    batched tree levels go to the device like everyone else's.
//
//  A completion thread waits on each plan's stream and stamps the request's
//  finish time; latency is finish - arrival. The reference has no serving
//  layer: it computes one request per process run (src/main.cpp:93-179).
// ============================================================================
#include "pidigits_hip.h"
#include "pdh_capi.h"
#include "pdh_pi.h"

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
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

struct Request {
    uint64_t id = 0;
    Kind kind = MUL;
    clk::time_point t_arrival, t_done;
    std::atomic<int> state{QUEUED};
    int err = 0;
    size_t bytes = 0;              // admission estimate of device memory held while running
    // multiply
    std::vector<uint32_t> a, b;
    void* d_arena = nullptr;       // a | b | out | len, resident until fetched
    size_t na = 0, nb = 0;
    // pi
    uint64_t digits = 0;
    std::string text;
    // bbp
    uint64_t start = 0;
    uint32_t count = 0;
    std::vector<uint8_t> values;
};

size_t mul_bytes(size_t na, size_t nb) {
    const size_t N = pdh_ntt_length(na, nb);
    return 8 * (na + nb) + 16 + 16 * N + (1u << 20);          // operands+product, two spectra, scratch
}

struct Job {
    void* stream = nullptr;
    pdh_batch* plan = nullptr;
    void* d_pos = nullptr;          // BBP positions (freed at completion)
    void* d_val = nullptr;
    std::vector<Request*> reqs;
    size_t work_bytes = 0;          // released at completion
};

} // namespace

struct pdh_server {
    size_t budget = 0, inuse = 0;
    int max_inflight = 0, inflight = 0;
    uint64_t next_id = 1;
    bool stop = false;
    std::mutex mu;
    std::condition_variable cv_sched, cv_done, cv_jobs, cv_pi;
    std::deque<Request*> pending;
    std::map<uint64_t, std::unique_ptr<Request>> all;
    std::deque<Job> jobs;
    std::deque<Request*> pi_queue;
    void* bbp_stream = nullptr;
    std::thread sched, completer, piworker;

    void finish(Request* r, int state, size_t release) {
        std::lock_guard<std::mutex> g(mu);
        r->t_done = clk::now();
        r->state = state;
        inuse -= release;
        --inflight;
        cv_done.notify_all();
        cv_sched.notify_all();
    }

    int launch_muls(std::vector<Request*>& rs) {
        if (rs.empty()) return PDH_OK;
        Job job;
        job.reqs = rs;
        int rc = pdh_stream_create(&job.stream);
        std::vector<pdh_mul_desc> d(rs.size());
        for (size_t i = 0; rc == PDH_OK && i < rs.size(); ++i) {
            Request* r = rs[i];
            const size_t n = r->na + r->nb;
            rc = pdh_dev_alloc(8 * n + 16, &r->d_arena);
            if (rc) break;
            uint32_t* A = (uint32_t*)r->d_arena;
            uint32_t* B = A + r->na;
            uint32_t* O = B + r->nb;
            uint64_t* L = (uint64_t*)(((uintptr_t)(O + n) + 7) & ~(uintptr_t)7);
            rc = pdh_copy_htod(A, r->a.data(), 4 * r->na, job.stream);
            if (!rc) rc = pdh_copy_htod(B, r->b.data(), 4 * r->nb, job.stream);
            d[i] = pdh_mul_desc{A, r->na, B, r->nb, O, L};
            job.work_bytes += r->bytes - (8 * n + 16);
        }
        if (!rc) rc = pdh_batch_create(d.data(), d.size(), &job.plan);
        if (!rc) rc = pdh_batch_run(job.plan, job.stream);
        if (rc) {
            if (job.plan) pdh_batch_destroy(job.plan);
            pdh_stream_destroy(job.stream);
            for (Request* r : rs) { pdh_dev_free(r->d_arena); r->d_arena = nullptr; r->err = rc; finish(r, FAILED, r->bytes); }
            return rc;
        }
        std::lock_guard<std::mutex> g(mu);
        jobs.push_back(std::move(job));
        cv_jobs.notify_all();
        return PDH_OK;
    }

    void launch_bbp(Request* r) {
        Job job;
        job.reqs = {r};
        job.stream = bbp_stream;
        int rc = pdh_dev_alloc(r->count, &job.d_val);
        if (!rc) rc = pdh_bbp_range_dev(r->start, r->count, (uint8_t*)job.d_val, bbp_stream);
        if (rc) {
            pdh_dev_free(job.d_pos);
            pdh_dev_free(job.d_val);
            r->err = rc;
            finish(r, FAILED, r->bytes);
            return;
        }
        job.work_bytes = r->bytes;
        std::lock_guard<std::mutex> g(mu);
        jobs.push_back(std::move(job));
        cv_jobs.notify_all();
    }

    // One scheduler tick: admit in arrival order while budget and slots allow,
    // then start everything admitted.
    void sched_loop() {
        for (;;) {
            std::vector<Request*> adm;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_sched.wait_for(lk, std::chrono::microseconds(500), [&] { return stop || !pending.empty(); });
                if (stop) return;
                while (!pending.empty()) {
                    Request* r = pending.front();
                    if (inflight >= max_inflight || (inflight > 0 && inuse + r->bytes > budget)) break;
                    pending.pop_front();
                    ++inflight;
                    inuse += r->bytes;
                    r->state = RUNNING;
                    adm.push_back(r);
                }
            }
            std::vector<Request*> small;
            for (Request* r : adm) {
                if (r->kind == MUL) {
                    if (pdh_ntt_length(r->na, r->nb) <= ((size_t)1 << 18)) small.push_back(r);
                    else { std::vector<Request*> one{r}; launch_muls(one); }
                } else if (r->kind == BBP) {
                    launch_bbp(r);
                } else {
                    std::lock_guard<std::mutex> g(mu);
                    pi_queue.push_back(r);
                    cv_pi.notify_all();
                }
            }
            launch_muls(small);
        }
    }

    void complete_loop() {
        for (;;) {
            Job job;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_jobs.wait(lk, [&] { return stop || !jobs.empty(); });
                if (jobs.empty()) return;
                job = std::move(jobs.front());
                jobs.pop_front();
            }
            int rc = pdh_stream_sync(job.stream);
            for (Request* r : job.reqs) {
                if (r->kind == BBP && rc == PDH_OK) {
                    r->values.resize(r->count);
                    rc = pdh_copy_dtoh(r->values.data(), job.d_val, r->count, nullptr);
                }
                r->err = rc;
                finish(r, rc ? FAILED : DONE, r->kind == BBP ? r->bytes : r->bytes - (8 * (r->na + r->nb) + 16));
            }
            if (job.plan) pdh_batch_destroy(job.plan);
            if (job.stream != bbp_stream) pdh_stream_destroy(job.stream);
            pdh_dev_free(job.d_pos);
            pdh_dev_free(job.d_val);
        }
    }

    void pi_loop() {
        for (;;) {
            Request* r;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_pi.wait(lk, [&] { return stop || !pi_queue.empty(); });
                if (pi_queue.empty()) return;
                r = pi_queue.front();
                pi_queue.pop_front();
            }
            int st = DONE;
            try {
                r->text = pidigits::compute_pi_chudnovsky_hip(r->digits);
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
    int rc = pdh_stream_create(&s->bbp_stream);
    if (rc) { delete s; return rc; }
    s->sched = std::thread([s] { s->sched_loop(); });
    s->completer = std::thread([s] { s->complete_loop(); });
    s->piworker = std::thread([s] { s->pi_loop(); });
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
    r->bytes = mul_bytes(na, nb);
    if (na == 0 || nb == 0) {                  // zero product: nothing to run
        r->a.assign(1, 0); r->b.assign(1, 0); r->na = r->nb = 1;
    }
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
    if (!s || !id || start < 1 || count == 0 || start - 1 + count > ((uint64_t)1 << 28)) return PDH_E_ARG;
    std::unique_ptr<Request> r(new Request());
    r->kind = BBP;
    r->start = start;
    r->count = count;
    r->bytes = ((size_t)1 << 28);              // the chunked BBP workspace
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
// digit values) and releases the request and its resident buffers.
int pdh_server_result(pdh_server* s, uint64_t id, void* out, size_t cap, size_t* len) {
    if (!s || !len) return PDH_E_ARG;
    int rc = pdh_server_wait(s, id, nullptr);
    if (rc) return rc;
    std::unique_ptr<Request> r;
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->all.find(id);
        r = std::move(it->second);
        s->all.erase(it);
    }
    if (r->kind == MUL) {
        const size_t n = r->na + r->nb;
        uint32_t* O = (uint32_t*)r->d_arena + n;
        uint64_t* L = (uint64_t*)(((uintptr_t)(O + n) + 7) & ~(uintptr_t)7);
        uint64_t nl = 0;
        rc = pdh_copy_dtoh(&nl, L, 8, nullptr);
        if (r->a.size() == 1 && r->a[0] == 0) nl = 0;   // submitted with an empty operand
        *len = (size_t)nl;
        if (!rc && nl > cap) rc = PDH_E_CAPACITY;
        if (!rc && nl) rc = pdh_copy_dtoh(out, O, 4 * nl, nullptr);
        pdh_dev_free(r->d_arena);
        std::lock_guard<std::mutex> g(s->mu);
        s->inuse -= 8 * n + 16;
        return rc;
    }
    const void* src = r->kind == PI ? (const void*)r->text.data() : (const void*)r->values.data();
    const size_t n = r->kind == PI ? r->text.size() : r->values.size();
    *len = n;
    if (n > cap) return PDH_E_CAPACITY;
    std::memcpy(out, src, n);
    return PDH_OK;
}

int pdh_server_destroy(pdh_server* s) {
    if (!s) return PDH_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv_sched.notify_all();
        s->cv_jobs.notify_all();
        s->cv_pi.notify_all();
    }
    s->sched.join();
    s->piworker.join();
    s->completer.join();
    for (auto& kv : s->all) if (kv.second->d_arena) pdh_dev_free(kv.second->d_arena);
    pdh_stream_destroy(s->bbp_stream);
    delete s;
    return PDH_OK;
}

} // extern "C"
