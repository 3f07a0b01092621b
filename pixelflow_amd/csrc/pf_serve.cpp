// pf_serve.cpp -- continuous-batching server over one resident batch (pf_batch_*).
//
// A worker thread owns the batch and its stream.  Hosts submit requests (a case: its
// descriptor, porosity and number of steps) from any thread; the worker, between step
// chunks,
//   1. admits queued requests in arrival order while a slot is free and the arena's
//      largest free range holds the case (pf_batch_case_bytes), a later request never
//      overtaking an earlier one that does not fit yet;
//   2. advances every resident case by min(chunk, fewest remaining steps) steps in shared
//      launches (pf_batch_step);
//   3. retires cases that reached their step count -- their requested fields are
//      downloaded into host memory first -- and cases whose step became non-finite.
// A case's arithmetic does not depend on its batch companions, so every result equals a
// run of that case alone.  Request state, timestamps and results are read with
// pf_serve_poll / pf_serve_times / pf_serve_result.  No reference counterpart: the
// reference runs one case per process.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pixelflow.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    pf_desc desc{};
    std::vector<double> raw;
    double threshold = 0;
    int64_t nsteps = 0;
    int32_t fields = 0;          // bit f keeps field f (PF_FIELD_*) at retirement
    int state = PF_SERVE_QUEUED;
    int slot = -1;
    int64_t steps_done = 0;
    int64_t sor_iters = 0;       // SOR iterations over all its steps
    double t_submit = 0, t_admit = -1, t_finish = -1;
    pf_stats last{};
    pf_stats worst{};             // per field, the maximum over its steps (NaN kept)
    std::map<int, std::vector<double>> out;   // field -> (m+2)(n+2)(l+2) host array
    std::string error;
};

}  // namespace

struct pf_server {
    pf_batch *batch = nullptr;
    int max_cases = 0;
    int chunk = 1;
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work;    // worker waits for requests
    std::condition_variable cv_done;    // hosts wait for results
    std::deque<std::shared_ptr<Request>> queue;
    std::map<int64_t, std::shared_ptr<Request>> all;
    std::map<int, std::shared_ptr<Request>> running;   // slot -> request (worker only)
    int64_t next_id = 1;
    bool stop = false;
    std::string fatal;
    std::thread worker;
    int64_t steps_run = 0;
    double busy_s = 0;
    int waiters = 0;                    // threads inside pf_serve_wait

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void retire(pf_server *s, const std::shared_ptr<Request> &r, bool failed, const std::string &why) {
    // fields first, outside the lock (the worker owns the batch)
    std::map<int, std::vector<double>> out;
    if (!failed) {
        const pf_desc &d = r->desc;
        const int64_t ldi = d.m + 2, ldj = d.n + 2, lz = (d.dim == 3) ? d.l + 2 : 1;
        for (int f = 0; f <= PF_FIELD_POROSITY; ++f) {
            if (!(r->fields & (1 << f)) || (f == PF_FIELD_W && d.dim == 2)) continue;
            std::vector<double> buf((size_t)(ldi * ldj * lz));
            if (pf_batch_download(s->batch, r->slot, f, buf.data(), ldi, ldj) != PF_OK) {
                failed = true;
                break;
            }
            out[f] = std::move(buf);
        }
    }
    pf_batch_remove(s->batch, r->slot);
    std::lock_guard<std::mutex> lk(s->mu);
    r->out = std::move(out);
    r->state = failed ? PF_SERVE_FAILED : PF_SERVE_DONE;
    if (failed) r->error = why.empty() ? std::string("field download failed") : why;
    r->t_finish = s->now();
    s->cv_done.notify_all();
}

void worker_loop(pf_server *s) {
    std::vector<pf_stats> st;
    for (;;) {
        // admission
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty() || !s->running.empty(); });
            if (s->stop) return;
            while (!s->queue.empty() && (int)s->running.size() < s->max_cases) {
                std::shared_ptr<Request> r = s->queue.front();
                const int64_t need = pf_batch_case_bytes(&r->desc);
                if (need < 0 || need > pf_batch_free_bytes(s->batch)) break;   // wait for memory (FIFO)
                s->queue.pop_front();
                lk.unlock();
                int slot = -1;
                const int rc = pf_batch_add(s->batch, &r->desc, r->raw.data(), r->threshold, &slot);
                lk.lock();
                if (rc != PF_OK) {
                    r->state = PF_SERVE_FAILED;
                    r->error = pf_batch_last_error(s->batch);
                    r->t_finish = s->now();
                    s->cv_done.notify_all();
                    continue;
                }
                r->slot = slot;
                r->state = PF_SERVE_RUNNING;
                r->t_admit = s->now();
                r->raw.clear();
                r->raw.shrink_to_fit();
                s->running[slot] = r;
            }
            if (s->running.empty()) {
                // the head of the queue does not fit an empty arena: it never will
                if (!s->queue.empty()) {
                    std::shared_ptr<Request> r = s->queue.front();
                    s->queue.pop_front();
                    r->state = PF_SERVE_FAILED;
                    r->error = "case larger than the server's arena";
                    r->t_finish = s->now();
                    s->cv_done.notify_all();
                }
                continue;
            }
        }
        // one chunk for every resident case
        int64_t todo = INT64_MAX;
        for (auto &kv : s->running) todo = std::min(todo, kv.second->nsteps - kv.second->steps_done);
        const int n = (int)std::max<int64_t>(1, std::min<int64_t>(s->chunk, todo));
        st.assign((size_t)n * s->max_cases, pf_stats{});
        const auto t = Clock::now();
        const int rc = pf_batch_step(s->batch, n, st.data());
        const double dt = std::chrono::duration<double>(Clock::now() - t).count();
        if (rc != PF_OK && rc != PF_EDIVERGED) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->fatal = pf_batch_last_error(s->batch);
            for (auto &kv : s->running) {
                pf_batch_remove(s->batch, kv.first);
                kv.second->state = PF_SERVE_FAILED;
                kv.second->error = s->fatal;
                kv.second->t_finish = s->now();
            }
            s->running.clear();
            s->cv_done.notify_all();
            continue;
        }
        std::vector<std::shared_ptr<Request>> finished, diverged;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->steps_run += n;
            s->busy_s += dt;
            for (auto &kv : s->running) {
                Request &r = *kv.second;
                r.steps_done += n;
                bool finite = true;
                for (int q = 0; q < n; ++q) {
                    const pf_stats &x = st[(size_t)q * s->max_cases + kv.first];
                    // residual is -1 where it is not evaluated; NaN or infinite where it is
                    finite = finite && std::isfinite(x.sor_error) && std::isfinite(x.residual);
                }
                for (int q = 0; q < n; ++q) r.sor_iters += st[(size_t)q * s->max_cases + kv.first].sor_iters;
                r.last = st[(size_t)(n - 1) * s->max_cases + kv.first];
                for (int q = 0; q < n; ++q) {
                    const pf_stats &x = st[(size_t)q * s->max_cases + kv.first];
                    auto worse = [](double a, double b) { return (b > a || b != b) ? b : a; };
                    r.worst.sor_error = worse(r.worst.sor_error, x.sor_error);
                    r.worst.residual = worse(r.worst.residual, x.residual);
                    r.worst.sor_iters = std::max(r.worst.sor_iters, x.sor_iters);
                    r.worst.cycles = std::max(r.worst.cycles, x.cycles);
                }
                if (!finite) diverged.push_back(kv.second);
                else if (r.steps_done >= r.nsteps) finished.push_back(kv.second);
            }
            for (auto &r : finished) s->running.erase(r->slot);
            for (auto &r : diverged) s->running.erase(r->slot);
        }
        for (auto &r : finished) retire(s, r, false, "");
        for (auto &r : diverged) retire(s, r, true, "diverged: non-finite SOR error or residual");
    }
}

}  // namespace

extern "C" int pf_serve_create(const pf_desc *kind, int max_cases, int64_t arena_bytes, int chunk, pf_server **out) {
    if (!out || !kind || max_cases < 1 || chunk < 1) return PF_EINVAL;
    *out = nullptr;
    if (arena_bytes <= 0) {
        int64_t free_b = 0, total_b = 0;
        if (pf_device_memory(&free_b, &total_b) != PF_OK) return PF_EHIP;
        arena_bytes = free_b / 2;
    }
    pf_server *s = new (std::nothrow) pf_server();
    if (!s) return PF_ENOMEM;
    const int rc = pf_batch_create(kind, max_cases, arena_bytes, &s->batch);
    if (rc != PF_OK) {
        delete s;
        return rc;
    }
    s->max_cases = max_cases;
    s->chunk = chunk;
    s->t0 = Clock::now();
    int dev = 0;
    (void)hipGetDevice(&dev);
    s->worker = std::thread([s, dev] {
        (void)hipSetDevice(dev);
        worker_loop(s);
    });
    *out = s;
    return PF_OK;
}

extern "C" int pf_serve_destroy(pf_server *s) {
    if (!s) return PF_EINVAL;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    {
        // the worker has stopped: whatever is still queued or running will never finish
        std::lock_guard<std::mutex> lk(s->mu);
        const double t = s->now();
        for (auto &kv : s->all) {
            Request &r = *kv.second;
            if (r.state == PF_SERVE_QUEUED || r.state == PF_SERVE_RUNNING) {
                r.state = PF_SERVE_FAILED;
                r.error = "server stopped";
                r.t_finish = t;
            }
        }
        s->queue.clear();
        s->running.clear();
        s->cv_done.notify_all();
    }
    {
        // threads blocked in pf_serve_wait wake on the failed state; let them leave first
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&] { return s->waiters == 0; });
    }
    pf_batch_destroy(s->batch);
    delete s;
    return PF_OK;
}

extern "C" pf_batch *pf_serve_batch(pf_server *s) { return s ? s->batch : nullptr; }

extern "C" int pf_serve_submit(pf_server *s, const pf_desc *desc, const double *raw_interior, double threshold,
                               int64_t nsteps, int32_t fields, int64_t *request_id) {
    if (!s || !desc || !raw_interior || nsteps < 1 || !request_id) return PF_EINVAL;
    if (pf_batch_case_bytes(desc) < 0) return PF_EINVAL;
    auto r = std::make_shared<Request>();
    r->desc = *desc;
    const size_t cells = (size_t)desc->m * desc->n * (desc->dim == 3 ? desc->l : 1);
    r->raw.assign(raw_interior, raw_interior + cells);
    r->threshold = threshold;
    r->nsteps = nsteps;
    r->fields = fields;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        r->t_submit = s->now();
        s->all[r->id] = r;
        s->queue.push_back(r);
        *request_id = r->id;
    }
    s->cv_work.notify_all();
    return PF_OK;
}

extern "C" int pf_serve_poll(pf_server *s, int64_t id, int *state, int64_t *steps_done, int64_t *sor_iters,
                             pf_stats *last) {
    if (!s || !state) return PF_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return PF_EINVAL;
    *state = it->second->state;
    if (steps_done) *steps_done = it->second->steps_done;
    if (sor_iters) *sor_iters = it->second->sor_iters;
    if (last) *last = it->second->last;
    return PF_OK;
}

extern "C" int pf_serve_worst(pf_server *s, int64_t id, pf_stats *worst) {
    if (!s || !worst) return PF_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return PF_EINVAL;
    *worst = it->second->worst;
    return PF_OK;
}

extern "C" int pf_serve_wait(pf_server *s, int64_t id, double timeout_s, int *state) {
    if (!s || !state) return PF_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return PF_EINVAL;
    std::shared_ptr<Request> r = it->second;
    auto finished = [&] { return r->state == PF_SERVE_DONE || r->state == PF_SERVE_FAILED; };
    ++s->waiters;
    if (timeout_s < 0) s->cv_done.wait(lk, finished);
    else s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished);
    *state = r->state;
    // pf_serve_destroy waits until no thread is left in here before it frees the server
    if (--s->waiters == 0) s->cv_done.notify_all();
    return PF_OK;
}

extern "C" int pf_serve_times(pf_server *s, int64_t id, double *t_submit, double *t_admit, double *t_finish) {
    if (!s) return PF_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return PF_EINVAL;
    if (t_submit) *t_submit = it->second->t_submit;
    if (t_admit) *t_admit = it->second->t_admit;
    if (t_finish) *t_finish = it->second->t_finish;
    return PF_OK;
}

extern "C" int pf_serve_result(pf_server *s, int64_t id, int field, double *host, int64_t ldi, int64_t ldj) {
    if (!s || !host) return PF_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end() || it->second->state != PF_SERVE_DONE) return PF_ESTATE;
    const Request &r = *it->second;
    auto f = r.out.find(field);
    if (f == r.out.end()) return PF_EINVAL;
    const pf_desc &d = r.desc;
    const int64_t X = d.m + 2, Y = d.n + 2, Z = (d.dim == 3) ? d.l + 2 : 1;
    if (ldi < X || ldj < Y) return PF_EINVAL;
    for (int64_t k = 0; k < Z; ++k)
        for (int64_t j = 0; j < Y; ++j)
            std::memcpy(host + ldi * (j + ldj * k), f->second.data() + X * (j + Y * k), sizeof(double) * (size_t)X);
    return PF_OK;
}

extern "C" int pf_serve_release(pf_server *s, int64_t id) {
    if (!s) return PF_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return PF_EINVAL;
    if (it->second->state != PF_SERVE_DONE && it->second->state != PF_SERVE_FAILED) return PF_ESTATE;
    s->all.erase(it);
    return PF_OK;
}

extern "C" int pf_serve_counters(pf_server *s, int64_t *steps_run, double *busy_s, int *queued, int *running) {
    if (!s) return PF_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    if (steps_run) *steps_run = s->steps_run;
    if (busy_s) *busy_s = s->busy_s;
    if (queued) *queued = (int)s->queue.size();
    if (running) *running = (int)s->running.size();
    return PF_OK;
}

extern "C" const char *pf_serve_error(pf_server *s, int64_t id) {
    // a copy per calling thread: the request may be released, and the worker may set the
    // server's fatal message, while the caller still reads the text
    thread_local std::string text;
    if (!s) return "";
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    text = (it == s->all.end()) ? s->fatal : it->second->error;
    return text.c_str();
}
