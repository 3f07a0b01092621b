// A stand-in for the batch C-ABI (pf_batch_*), linked with the library's own serving loop
// (pixelflow_amd/csrc/pf_serve.cpp) into a host-only test library, so the loop's policy
// is tested on the CPU: admission by free arena memory in arrival order, chunked steps
// that stop at each request's step count, retirement with the fields downloaded first,
// failure of diverged and oversized cases, and teardown.  Arena use is cells x a set
// number of bytes; a step advances every resident case at once.  With the gate on, every
// pf_batch_step waits until the test grants it, so the test decides how the worker's
// steps interleave with its own calls; with adds held, pf_batch_add waits until they are
// released, so requests submitted meanwhile are all queued when admission goes on.  Every
// call is logged as a line of text.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "pixelflow.h"

struct pf_batch {
    int max_cases = 0;
    int64_t arena = 0;
    std::map<int, pf_desc> cases;          // slot -> descriptor
    std::map<int, int64_t> steps;          // slot -> steps taken
};

namespace {

std::mutex g_mu;
std::condition_variable g_cv;
int64_t g_bytes_per_cell = 100;
bool g_gate = false;
bool g_hold_adds = false;
int g_granted = 0;                         // step calls the test has allowed
int g_step_calls = 0;
int g_blocked = 0;                         // step calls waiting at the gate
int g_diverge_m = -1;                      // a case with this m goes non-finite ...
int64_t g_diverge_after = 2;               // ... once it has taken this many steps
std::string g_log;

int64_t cells_of(const pf_desc &d) { return (int64_t)d.m * d.n * (d.dim == 3 ? d.l : 1); }

void log_line(const std::string &s) { g_log += s + "\n"; }

}  // namespace

extern "C" {

void fake_reset(int64_t bytes_per_cell) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_bytes_per_cell = bytes_per_cell;
    g_gate = false;
    g_hold_adds = false;
    g_granted = 0;
    g_step_calls = 0;
    g_blocked = 0;
    g_diverge_m = -1;
    g_diverge_after = 2;
    g_log.clear();
}

void fake_set_gate(int on) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_gate = on != 0;
    g_cv.notify_all();
}

void fake_hold_adds(int on) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_hold_adds = on != 0;
    g_cv.notify_all();
}

void fake_grant_steps(int n) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_granted += n;
    g_cv.notify_all();
}

// Wait (at most timeout_s) until the worker has finished `n` step calls in all.
int fake_wait_step_calls(int n, double timeout_s) {
    std::unique_lock<std::mutex> lk(g_mu);
    g_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return g_step_calls >= n; });
    return g_step_calls;
}

// Wait (at most timeout_s) until a step call is waiting at the gate; returns 1 if one is.
int fake_wait_blocked(double timeout_s) {
    std::unique_lock<std::mutex> lk(g_mu);
    g_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return g_blocked > 0 && g_granted == 0; });
    return (g_blocked > 0 && g_granted == 0) ? 1 : 0;
}

void fake_diverge(int m, int64_t after_steps) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_diverge_m = m;
    g_diverge_after = after_steps;
}

// Copy of the call log into buf (cap bytes with the terminating zero); returns its length.
int fake_log(char *buf, int cap) {
    std::lock_guard<std::mutex> lk(g_mu);
    const int n = (int)std::min<size_t>(g_log.size(), (size_t)std::max(0, cap - 1));
    std::memcpy(buf, g_log.data(), (size_t)n);
    if (cap > 0) buf[n] = 0;
    return (int)g_log.size();
}

int pf_device_memory(int64_t *free_bytes, int64_t *total_bytes) {
    if (free_bytes) *free_bytes = (int64_t)1 << 34;
    if (total_bytes) *total_bytes = (int64_t)1 << 35;
    return PF_OK;
}

int pf_batch_create(const pf_desc *kind, int max_cases, int64_t arena_bytes, pf_batch **out) {
    if (!kind || !out || max_cases < 1) return PF_EINVAL;
    pf_batch *b = new pf_batch();
    b->max_cases = max_cases;
    b->arena = arena_bytes;
    *out = b;
    std::lock_guard<std::mutex> lk(g_mu);
    log_line("create " + std::to_string(max_cases) + " " + std::to_string(arena_bytes));
    return PF_OK;
}

int pf_batch_destroy(pf_batch *b) {
    delete b;
    std::lock_guard<std::mutex> lk(g_mu);
    log_line("destroy");
    return PF_OK;
}

const char *pf_batch_last_error(const pf_batch *) { return "stand-in batch"; }

int64_t pf_batch_case_bytes(const pf_desc *d) {
    if (!d || d->m < 2 || d->n < 2) return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    return cells_of(*d) * g_bytes_per_cell;
}

int64_t pf_batch_free_bytes(const pf_batch *b) {
    std::lock_guard<std::mutex> lk(g_mu);
    int64_t used = 0;
    for (const auto &kv : b->cases) used += cells_of(kv.second) * g_bytes_per_cell;
    return b->arena - used;
}

int pf_batch_add(pf_batch *b, const pf_desc *d, const double *raw, double, int *slot) {
    if (!b || !d || !raw || !slot) return PF_EINVAL;
    {
        std::unique_lock<std::mutex> lk(g_mu);
        g_cv.wait(lk, [] { return !g_hold_adds; });
    }
    if (pf_batch_case_bytes(d) > pf_batch_free_bytes(b)) return PF_ENOMEM;
    std::lock_guard<std::mutex> lk(g_mu);
    int s = 0;
    while (b->cases.count(s)) ++s;
    if (s >= b->max_cases) return PF_EINVAL;
    b->cases[s] = *d;
    b->steps[s] = 0;
    *slot = s;
    log_line("add " + std::to_string(s) + " m=" + std::to_string(d->m));
    return PF_OK;
}

int pf_batch_remove(pf_batch *b, int slot) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!b || !b->cases.count(slot)) return PF_EINVAL;
    b->cases.erase(slot);
    b->steps.erase(slot);
    log_line("remove " + std::to_string(slot));
    return PF_OK;
}

int pf_batch_step(pf_batch *b, int nsteps, pf_stats *stats) {
    std::unique_lock<std::mutex> lk(g_mu);
    if (g_gate) {
        ++g_blocked;
        g_cv.notify_all();
        g_cv.wait(lk, [&] { return g_granted > 0 || !g_gate; });
        --g_blocked;
        if (g_granted > 0) --g_granted;
    }
    std::string line = "step " + std::to_string(nsteps);
    bool diverged = false;
    for (auto &kv : b->steps) {
        const int s = kv.first;
        kv.second += nsteps;
        line += " " + std::to_string(s);
        const bool bad = b->cases[s].m == g_diverge_m && kv.second >= g_diverge_after;
        for (int q = 0; q < nsteps; ++q) {
            pf_stats &o = stats[(size_t)q * b->max_cases + s];
            o.sor_error = (bad && q == nsteps - 1) ? std::nan("") : 1e-3;
            o.residual = -1.0;
            o.sor_iters = 5;
            o.cycles = 0;
        }
        diverged = diverged || bad;
    }
    log_line(line);
    ++g_step_calls;
    g_cv.notify_all();
    return diverged ? PF_EDIVERGED : PF_OK;
}

// Every value of field f of the case in `slot` is 10 * slot + f.
int pf_batch_download(pf_batch *b, int slot, int field, double *host, int64_t ldi, int64_t ldj) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!b || !b->cases.count(slot) || !host) return PF_EINVAL;
    const pf_desc &d = b->cases[slot];
    const int64_t X = d.m + 2, Y = d.n + 2, Z = (d.dim == 3) ? d.l + 2 : 1;
    for (int64_t k = 0; k < Z; ++k)
        for (int64_t j = 0; j < Y; ++j)
            for (int64_t i = 0; i < X; ++i) host[i + ldi * (j + ldj * k)] = 10.0 * slot + field;
    log_line("download " + std::to_string(slot) + " " + std::to_string(field));
    return PF_OK;
}

}  // extern "C"
