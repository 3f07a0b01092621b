// Lifecycle driver of the batch and serving C-ABI, built with host AddressSanitizer
// (pixelflow_amd/build.py --asan: every host translation unit of the library and this file
// instrumented, device code unchanged).  It runs the sequences whose host bookkeeping is the
// most involved -- the autotuner's sweep over configurations (each one a drain, a graph drop,
// a capture and an instantiation), continuous batching (cases removed and added between
// steps, tables and step histories rebuilt), multigrid tolerance-mode batches with their side
// streams, and a server destroyed with requests still queued -- and exits non-zero on any
// error code.  A heap misuse in the library's host code stops the run with an ASan report.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pixelflow.h"

#define CHECK(call)                                                                          \
    do {                                                                                     \
        int rc_ = (call);                                                                    \
        if (rc_ != PF_OK) {                                                                  \
            std::fprintf(stderr, "[lifecycle] %s:%d %s -> %d\n", __FILE__, __LINE__, #call, rc_); \
            std::exit(1);                                                                    \
        }                                                                                    \
    } while (0)

static void say(const char *what) {
    std::printf("[lifecycle] %s\n", what);
    std::fflush(stdout);
}

// A 2D uniform-inflow case of m x n cells with a cylinder of porosity 0 (tanh interface,
// diameter 0.3 n), cell Reynolds number 1 and a time step of 0.1 dx / U.
static pf_desc desc_2d(int m, int n, int precision) {
    pf_desc d{};
    d.dim = 2;
    d.m = m;
    d.n = n;
    d.l = 1;
    d.bc = PF_BC_UNIFORM;
    d.nonslip = 1;
    d.iter_max = 40;
    d.precision = precision;
    d.inner_iters = 20;
    d.dx = 1.0 / (m - 1);
    d.dy = d.dx;
    d.dz = 1.0;
    d.dt = 0.1 * d.dx;
    d.xnue = d.dx;
    d.density = 1.0;
    d.thickness = 1.5;
    d.inlet_velocity = 1.0;
    d.omega = 1.7;
    return d;
}

static pf_desc desc_room(int s) {
    pf_desc d{};
    d.dim = 3;
    d.m = d.n = d.l = s;
    d.bc = PF_BC_AIR;
    d.nonslip = 1;
    d.iter_max = 20;
    d.precision = PF_PREC_FP64;
    const int faces[6] = {0, 0, 1, 2, 0, 0};   // inlet on the east face, outlet on the west
    for (int f = 0; f < 6; ++f) d.face[f] = faces[f];
    d.dx = d.dy = d.dz = 1.0 / (s - 1);
    d.dt = 0.2 * d.dx;
    d.xnue = d.dx;
    d.density = 1.0;
    d.thickness = 1.5;
    d.inlet_velocity = 1.0;
    d.omega = 1.7;
    return d;
}

static std::vector<double> cylinder(int m, int n) {
    std::vector<double> phi((size_t)m * n);
    const double cx = 0.3 * m, cy = 0.5 * n, r = 0.15 * n;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            const double d = std::sqrt((i - cx) * (i - cx) + (j - cy) * (j - cy)) - r;
            phi[(size_t)i + (size_t)m * j] = 0.5 * std::tanh(d / 1.5) + 0.5;
        }
    return phi;
}

static int add_2d(pf_batch *b, int m, int n, int precision, double tol = 0.0, int solver = PF_SOLVER_SOR) {
    pf_desc d = desc_2d(m, n, precision);
    d.residual_tol = tol;
    d.solver = solver;
    if (tol > 0) d.max_cycles = 30;
    std::vector<double> phi = cylinder(m, n);
    int slot = -1;
    CHECK(pf_batch_add(b, &d, phi.data(), 1e-6, &slot));
    return slot;
}

// A step that reports PF_EDIVERGED still advanced every case (the autotuner times such
// steps too); the cases whose statistics are not finite are listed once.
static int g_diverged_calls = 0;

static void step(pf_batch *b, int nsteps, int max_cases = 256) {
    std::vector<pf_stats> st((size_t)nsteps * max_cases);
    const int rc = pf_batch_step(b, nsteps, st.data());
    if (rc == PF_EDIVERGED) {
        if (g_diverged_calls++ == 0) {
            for (int s = 0; s < nsteps; ++s)
                for (int slot = 0; slot < max_cases; ++slot) {
                    const pf_stats &o = st[(size_t)s * max_cases + slot];
                    if (pf_batch_case_steps(b, slot) >= 0 && !std::isfinite(o.sor_error))
                        std::printf("[lifecycle] step %d slot %d: sor_error %g residual %g\n", s, slot, o.sor_error,
                                    o.residual);
                }
            std::fflush(stdout);
        }
    } else {
        CHECK(rc);
    }
    CHECK(pf_batch_sync(b));
}

// The autotuner's candidate sweep (pixelflow_amd/autotune.py) over 2D cases of uneven size.
static void autotune_sequence(int precision) {
    pf_desc kind = desc_2d(64, 32, precision);
    pf_batch *b = nullptr;
    CHECK(pf_batch_create(&kind, 256, (int64_t)1 << 30, &b));
    std::vector<int> slots;
    for (int c = 0; c < 48; ++c) slots.push_back(add_2d(b, 48 + 8 * (c % 13), 24 + 4 * (c % 7), precision));
    const int tiles[3] = {32, 64, 128}, blocks[4] = {0, 512, 1024, 2048}, inner[3] = {10, 20, 40};
    for (int t : tiles)
        for (int pb : blocks)
            for (int k : inner) {
                pf_batch_config cfg{pb, 1, k, t};
                CHECK(pf_batch_set_config(b, &cfg));
                step(b, 1);   // captures the graphs
                step(b, 2);
            }
    say(precision == PF_PREC_MIXED ? "autotune sweep (mixed) done" : "autotune sweep (fp64) done");
    // continuous batching: every third case leaves, new ones arrive, between graph replays
    for (size_t i = 0; i < slots.size(); i += 3) CHECK(pf_batch_remove(b, slots[i]));
    for (int c = 0; c < 20; ++c) add_2d(b, 40 + 12 * (c % 9), 20 + 6 * (c % 5), precision);
    step(b, 3);
    pf_batch_config eager{0, 0, 20, 64};
    CHECK(pf_batch_set_config(b, &eager));
    step(b, 2);
    pf_batch_config graph{1024, 1, 20, 32};
    CHECK(pf_batch_set_config(b, &graph));
    step(b, 12);   // longer than any earlier call: the step histories grow, the graphs are dropped
    step(b, 1);
    CHECK(pf_batch_destroy(b));
    say("continuous batching done");
}

static void rooms_sequence() {
    pf_desc kind = desc_room(24);
    pf_batch *b = nullptr;
    CHECK(pf_batch_create(&kind, 16, (int64_t)1 << 30, &b));   // max_cases 16
    std::vector<int> slots;
    for (int s : {24, 28, 32, 20, 26}) {
        pf_desc d = desc_room(s);
        std::vector<double> phi((size_t)s * s * s, 1.0);
        for (int k = s / 3; k < s / 2; ++k)   // a box of furniture
            for (int j = 0; j < s / 3; ++j)
                for (int i = s / 4; i < s / 2; ++i) phi[(size_t)i + (size_t)s * (j + (size_t)s * k)] = 1e-6;
        int slot = -1;
        CHECK(pf_batch_add(b, &d, phi.data(), 1e-6, &slot));
        slots.push_back(slot);
    }
    for (int pb : {0, 512}) {
        pf_batch_config cfg{pb, 1, 0, 0};
        CHECK(pf_batch_set_config(b, &cfg));
        step(b, 2, 16);
    }
    CHECK(pf_batch_remove(b, slots[1]));
    step(b, 2, 16);
    CHECK(pf_batch_destroy(b));
    say("3D rooms done");
}

static void multigrid_sequence(int precision) {
    pf_desc kind = desc_2d(64, 32, precision);
    kind.residual_tol = 1e-9;
    kind.solver = PF_SOLVER_MG;
    kind.max_cycles = 30;
    pf_batch *b = nullptr;
    CHECK(pf_batch_create(&kind, 32, (int64_t)1 << 30, &b));
    std::vector<int> slots;
    for (int c = 0; c < 6; ++c) slots.push_back(add_2d(b, 48 + 16 * c, 24 + 8 * (c % 3), precision, 1e-9, PF_SOLVER_MG));
    step(b, 2, 32);
    CHECK(pf_batch_remove(b, slots[2]));
    CHECK(pf_batch_remove(b, slots[4]));
    add_2d(b, 96, 40, precision, 1e-9, PF_SOLVER_MG);
    step(b, 3, 32);
    CHECK(pf_batch_destroy(b));
    say(precision == PF_PREC_MIXED ? "multigrid batch (mixed) done" : "multigrid batch (fp64) done");
}

static void serve_sequence() {
    pf_desc kind = desc_2d(64, 32, PF_PREC_MIXED);
    pf_server *s = nullptr;
    CHECK(pf_serve_create(&kind, 8, (int64_t)256 << 20, 5, &s));
    std::vector<int64_t> ids;
    for (int r = 0; r < 14; ++r) {
        const int m = 48 + 16 * (r % 6), n = 24 + 8 * (r % 4);
        pf_desc d = desc_2d(m, n, PF_PREC_MIXED);
        std::vector<double> phi = cylinder(m, n);
        int64_t id = -1;
        CHECK(pf_serve_submit(s, &d, phi.data(), 1e-6, 20 + 5 * (r % 3), 1u << PF_FIELD_U, &id));
        ids.push_back(id);
    }
    int state = -1;
    CHECK(pf_serve_wait(s, ids[0], 60.0, &state));
    std::vector<double> u((size_t)(48 + 2) * (24 + 2));
    if (state == PF_SERVE_DONE) {
        CHECK(pf_serve_result(s, ids[0], PF_FIELD_U, u.data(), 48 + 2, 24 + 2));
        CHECK(pf_serve_release(s, ids[0]));
    }
    CHECK(pf_serve_destroy(s));   // later requests are still queued or running
    say("server destroyed with requests in flight");
}

int main() {
    autotune_sequence(PF_PREC_MIXED);
    autotune_sequence(PF_PREC_FP64);
    rooms_sequence();
    multigrid_sequence(PF_PREC_MIXED);
    multigrid_sequence(PF_PREC_FP64);
    serve_sequence();
    std::printf("[lifecycle] step calls reporting a diverged case: %d\n", g_diverged_calls);
    say("all sequences done");
    return 0;
}
