// pf_api.hip -- the C-ABI of include/pixelflow.h: contexts, host<->device staging and the
// per-step launch sequence on one HIP stream.  Everything stays resident in HBM between
// calls; the host only moves data at set-up and when a caller asks for a field.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pixelflow.h"
#include "pf_host.h"
#include "pf_kernels.h"
#include "pf_mg.h"

struct pf_ctx {
    pf_desc d{};
    PfGeom g{};
    pf::StepArgs a{};
    hipStream_t st = nullptr;
    int device = 0;
    // natural layout
    double *u = nullptr, *v = nullptr, *w = nullptr, *uo = nullptr, *vo = nullptr, *wo = nullptr;
    double *phi = nullptr, *div = nullptr, *dout = nullptr;
    double *force_buf = nullptr;   // body force: per-block partial sums, then the 6 totals
    double *P2 = nullptr;          // fused and 2D blocked sweeps: second pressure buffer (ping-pong)
    int tb2_k = 0, tb2_r = 0;      // 2D blocked sweep: iterations per launch, rows per block (0: off)
    float *E2 = nullptr;           // fused sweep, mixed precision: second correction buffer
    bool fused = false;            // one launch per SOR iteration (PF_FUSED=1; see pf_fused.h)
    // colour-split layout (2 * HS doubles each)
    double *P = nullptr, *BB = nullptr, *AW = nullptr, *AS = nullptr, *AB = nullptr, *PHIs = nullptr;
    // SOR sweep kernel: 0 stored face coefficients, 1 faces rebuilt from porosity with IEEE
    // division, 2 the same with the checked constant-divisor division (chosen when
    // k_faces_check finds every face identical to IEEE division).
    int sor_variant = 2;
    int sor_forced = -1;
    int xcd_remap = 1;              // XCD-contiguous block order for FP32 and residual kernels
    int xcd_remap64 = 1;            // the same for FP64 SOR sweeps
    int sweep_kb = 1;               // 3D FP64 sweeps: cells per thread (1, 2 or 4)
    int sweep_kb32 = 1;             // 3D FP32 sweeps: cells per thread
    int sweep_xw = 64;              // colour slots per wave row in the sweeps (64 or 32)
    unsigned long long *err = nullptr;       // one entry per step of the last pf_step
    unsigned long long *err_part = nullptr;  // per-wave partial maxima of the current step
    unsigned long long *res_part = nullptr;  // mixed: per-wave max |r/ap|
    unsigned long long *res_val = nullptr;   // mixed: reduced residual of the last evaluation
    float *E32 = nullptr, *R32 = nullptr, *PHI32 = nullptr;   // mixed: colour-split FP32 arrays
    // FP64 sweeps: the colour-split porosity coded in 4 bytes per value (PHC), with the
    // table of the values FP32 cannot hold (PHTAB, room for PHI_TAB_MAX); used when
    // phi_coded (on past 512 MB of sweep arrays; PF_PHI_CODE=0/1 in the environment decides)
    float *PHC = nullptr;
    double *PHTAB = nullptr;
    bool phi_code_on = false, phi_coded = false;
    bool phi_code_forced = false;   // PF_PHI_CODE=1: code even where unit waves dominate
    // FP64 sweeps: one byte per 32 colour slots of a row, 1 where every porosity operand of
    // those cells is 1 (unit_index in pf_layout.h); waves of flagged chunks skip their
    // porosity loads.  unit_cells: cells of both colours in such waves, of the FP64 sweeps
    // or, in mixed precision, of the FP32 sweeps (PF_UNIT_PHI=0 in the environment turns
    // the flags off)
    uint8_t *UNIT = nullptr;
    int64_t unit_cells = 0;
    uint8_t *UNIT32 = nullptr;      // mixed precision: the same flags of the FP32 porosity
    bool unit_half = false;         // unit waves hold half the cells or more: the residual and
                                    // the multigrid's level-0 kernels use the flags too
    int err_cap = 0, nparts = 0;
    bool have_phi = false, have_init = false;
    bool timing = false;
    std::vector<hipEvent_t> evs;   // timing mode: one (start, end) pair per step around the SOR
    double sweep_us_avg = 0.0;
    int64_t sweep_launches = 0;
    std::vector<double> mixed_res;
    std::vector<int> mixed_cycles, mixed_iters;
    double solve_us_avg = 0.0;      // timing mode: average pressure-solve time per step
    // tolerance mode with PF_SOLVER_MG: the multigrid-preconditioned solver of the correction
    // equation, FP32 for mixed precision, FP64 otherwise
    pf::MgSolver<float> *mg32 = nullptr;
    pf::MgSolver<double> *mg64 = nullptr;
    std::string msg;
};

static thread_local std::string g_create_error;

// Distinct porosity values that FP32 cannot hold, at most, for the coded sweep (8 MB table;
// a benchmark sphere has about 10^4, the 512x256x256 boxes a few hundred).
static const size_t PHI_TAB_MAX = (size_t)1 << 20;

// On a failure the context's stream is drained before the early return, so no asynchronous
// copy from or into a host buffer of the calling function is still in flight afterwards.
#define PF_TRY_HIP(ctx, call)                                                                   \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            if ((ctx)->st) (void)hipStreamSynchronize((ctx)->st);                               \
            (ctx)->msg = std::string(#call) + ": " + hipGetErrorString(e_);                     \
            return PF_EHIP;                                                                     \
        }                                                                                       \
    } while (0)

static int fail(pf_ctx *c, int code, const std::string &m) {
    c->msg = m;
    return code;
}

static bool desc_ok(const pf_desc *d, std::string &why) { return pf::desc_valid(d, why); }

extern "C" int pf_abi_version(void) { return PF_ABI_VERSION; }

// The temporally blocked 2D sweep's iterations per launch and rows per block, or 0 and 0
// where it is not used: FP64, 2D, and the shape within what the kernel holds.  By default 2
// iterations and 1 row (on the 256x128 cylinder, ms per step: two-launch sweeps 0.86, K,R =
// 1,1 0.60, 2,1 0.47, 2,2 0.54, 3,1 0.48, 4,1 0.52 with porosity in LDS, 4,8 0.80); rows
// wider than about 310 cells do not fit LDS and keep the two-launch sweeps.
// PF_SOR_TB2="K,R" sets them ("0,0": off).
static void tb2_params(const PfGeom &g, const pf_desc *d, int &K, int &R) {
    K = R = 0;
    if (d->precision != PF_PREC_FP64 || g.dim != 2) return;
    int k = 2, r = 1;
    const char *e = std::getenv("PF_SOR_TB2");
    if (e && *e && std::sscanf(e, "%d,%d", &k, &r) != 2) k = r = 0;
    if (k > 0 && r > 0 && pf::sor_tb2_ok(g, d->bc, k, r)) {
        K = k;
        R = r;
    }
}

// The colour sweeps' unit-porosity flags, of the FP64 porosity for FP64 and of the FP32
// porosity for mixed precision (off with PF_UNIT_PHI=0 in the environment).
static bool unit_wanted(const pf_desc *d) {
    const char *u = std::getenv("PF_UNIT_PHI");
    return !(u && *u && std::atoi(u) == 0);
}

extern "C" int64_t pf_footprint_bytes(const pf_desc *d) {
    std::string why;
    if (!desc_ok(d, why)) return -1;
    PfGeom g = pf_make_geom(d->dim, d->m, d->n, d->l);
    const int nat = (d->dim == 3) ? 8 : 6;
    const int spl = (d->dim == 3) ? 6 : 5;
    const int64_t interior = (int64_t)d->m * d->n * (d->dim == 3 ? d->l : 1);
    const int64_t mixed = (d->precision == PF_PREC_MIXED) ? (int64_t)4 * 3 * 2 * g.HS : 0;
    const int64_t mg = (d->residual_tol > 0 && d->solver == PF_SOLVER_MG)
                           ? pf::mg_bytes(g, d->bc, d->precision == PF_PREC_MIXED)
                           : 0;
    const char *pc = std::getenv("PF_PHI_CODE");
    const bool code_on = (pc && *pc) ? std::atoi(pc) != 0 : interior * 24 > ((int64_t)512 << 20);
    const int64_t coded =
        (d->precision == PF_PREC_FP64 && code_on) ? (int64_t)4 * 2 * g.HS + (int64_t)8 * PHI_TAB_MAX : 0;
    int tk = 0, tr = 0;
    tb2_params(g, d, tk, tr);
    const int64_t p2 = (tk > 0) ? (int64_t)8 * 2 * g.HS : 0;
    const int64_t unit = unit_wanted(d) ? unit_flags_size(g) * (d->precision == PF_PREC_MIXED ? 2 : 1) : 0;
    return (int64_t)8 * (nat * g.NS + spl * 2 * g.HS + interior) + mixed + mg + coded + p2 + unit + 4096;
}

static int alloc(pf_ctx *c, double **p, int64_t n) {
    hipError_t e = hipMalloc((void **)p, (size_t)n * sizeof(double));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, PF_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    e = hipMemsetAsync(*p, 0, (size_t)n * sizeof(double), c->st);
    if (e != hipSuccess) return fail(c, PF_EHIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    return PF_OK;
}

static void release(pf_ctx *c) {
    double **all[] = {&c->u, &c->v, &c->w, &c->uo, &c->vo, &c->wo, &c->phi, &c->div, &c->dout,
                      &c->P, &c->BB, &c->AW, &c->AS, &c->AB, &c->PHIs, &c->force_buf, &c->P2};
    for (double **p : all)
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    if (c->err) { (void)hipFree(c->err); c->err = nullptr; }
    if (c->err_part) { (void)hipFree(c->err_part); c->err_part = nullptr; }
    if (c->res_part) { (void)hipFree(c->res_part); c->res_part = nullptr; }
    if (c->res_val) { (void)hipFree(c->res_val); c->res_val = nullptr; }
    float **f32[] = {&c->E32, &c->R32, &c->PHI32, &c->E2, &c->PHC};
    for (float **p : f32)
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    if (c->PHTAB) { (void)hipFree(c->PHTAB); c->PHTAB = nullptr; }
    if (c->UNIT) { (void)hipFree(c->UNIT); c->UNIT = nullptr; }
    if (c->UNIT32) { (void)hipFree(c->UNIT32); c->UNIT32 = nullptr; }
    if (c->mg32) { pf::mg_destroy(*c->mg32); delete c->mg32; c->mg32 = nullptr; }
    if (c->mg64) { pf::mg_destroy(*c->mg64); delete c->mg64; c->mg64 = nullptr; }
    for (hipEvent_t e : c->evs) (void)hipEventDestroy(e);
    c->evs.clear();
    if (c->st) (void)hipStreamDestroy(c->st);
    c->st = nullptr;
}

extern "C" int pf_create(const pf_desc *desc, pf_ctx **out) {
    if (!out) return PF_EINVAL;
    *out = nullptr;
    std::string why;
    if (!desc_ok(desc, why)) {
        g_create_error = why;
        return PF_EINVAL;
    }
    pf_ctx *c = new (std::nothrow) pf_ctx();
    if (!c) return PF_ENOMEM;
    c->d = *desc;
    c->g = pf_make_geom(desc->dim, desc->m, desc->n, desc->l);
    c->a = pf::make_step_args(*desc);

    int rc;
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    if (e != hipSuccess) {
        g_create_error = std::string("HIP init: ") + hipGetErrorString(e);
        release(c);
        delete c;
        return PF_EHIP;
    }
    const PfGeom &g = c->g;
    const int64_t interior = (int64_t)g.m * g.n * g.l;
    double **nat[] = {&c->u, &c->v, &c->w, &c->uo, &c->vo, &c->wo, &c->phi, &c->div};
    for (double **p : nat) {
        if ((p == &c->w || p == &c->wo) && g.dim == 2) continue;
        if ((rc = alloc(c, p, g.NS)) != PF_OK) goto bad;
    }
    if ((rc = alloc(c, &c->dout, interior)) != PF_OK) goto bad;
    {
        double **spl[] = {&c->P, &c->BB, &c->AW, &c->AS, &c->AB, &c->PHIs};
        for (double **p : spl) {
            if (p == &c->AB && g.dim == 2) continue;
            if ((rc = alloc(c, p, 2 * g.HS)) != PF_OK) goto bad;
        }
    }
    c->nparts = pf::sor_blocks(g);
    if (hipMalloc((void **)&c->err_part, sizeof(unsigned long long) * (size_t)c->nparts) != hipSuccess ||
        hipMalloc((void **)&c->res_part, sizeof(unsigned long long) * (size_t)c->nparts) != hipSuccess ||
        hipMalloc((void **)&c->res_val, sizeof(unsigned long long)) != hipSuccess) {
        rc = fail(c, PF_ENOMEM, "hipMalloc of the reduction slots failed");
        goto bad;
    }
    pf::launch_clear_u64(c->st, c->err_part, c->nparts);
    pf::launch_clear_u64(c->st, c->res_part, c->nparts);
    {
        const char *fv = std::getenv("PF_FUSED");
        if (fv && *fv) c->fused = std::atoi(fv) != 0;
        c->fused = c->fused && pf::sor_fused_ok(g, desc->bc);
    }
    tb2_params(g, desc, c->tb2_k, c->tb2_r);
    if ((c->fused || c->tb2_k > 0) && desc->precision == PF_PREC_FP64 && (rc = alloc(c, &c->P2, 2 * g.HS)) != PF_OK)
        goto bad;
    if (desc->precision == PF_PREC_MIXED) {
        float **f32[] = {&c->E32, &c->R32, &c->PHI32, &c->E2};
        for (float **p : f32) {
            if (p == &c->E2 && !c->fused) continue;
            if (hipMalloc((void **)p, sizeof(float) * (size_t)(2 * g.HS)) != hipSuccess) {
                rc = fail(c, PF_ENOMEM, "hipMalloc of the FP32 arrays failed");
                goto bad;
            }
            if (hipMemsetAsync(*p, 0, sizeof(float) * (size_t)(2 * g.HS), c->st) != hipSuccess) {
                rc = fail(c, PF_EHIP, "hipMemsetAsync failed");
                goto bad;
            }
        }
    }
    {
        // coded porosity where the FP64 sweep's arrays (24 B per cell) are well past the
        // last-level cache: 512x256x256 179 against 184 us per sweep; at 256x128x128, inside
        // the cache, it was slower (18.5 against 17.9 us).  pf_set_porosity leaves it unused
        // where unit-porosity waves hold half the cells or more (see derive_from_porosity)
        const int64_t cells = (int64_t)g.m * g.n * (g.dim == 3 ? g.l : 1);
        c->phi_code_on = cells * 24 > ((int64_t)512 << 20);
        const char *pc = std::getenv("PF_PHI_CODE");
        if (pc && *pc) {
            c->phi_code_on = std::atoi(pc) != 0;
            c->phi_code_forced = c->phi_code_on;
        }
    }
    if (desc->precision == PF_PREC_FP64 && c->phi_code_on) {
        if (hipMalloc((void **)&c->PHC, sizeof(float) * (size_t)(2 * g.HS)) != hipSuccess ||
            hipMalloc((void **)&c->PHTAB, sizeof(double) * PHI_TAB_MAX) != hipSuccess) {
            rc = fail(c, PF_ENOMEM, "hipMalloc of the coded porosity failed");
            goto bad;
        }
    }
    // FP64 flags for every precision (the FP64 residual reads the FP64 porosity), FP32 flags
    // for the mixed-precision sweeps
    if (unit_wanted(desc) &&
        (hipMalloc((void **)&c->UNIT, (size_t)unit_flags_size(g)) != hipSuccess ||
         (desc->precision == PF_PREC_MIXED &&
          hipMalloc((void **)&c->UNIT32, (size_t)unit_flags_size(g)) != hipSuccess))) {
        rc = fail(c, PF_ENOMEM, "hipMalloc of the unit-porosity flags failed");
        goto bad;
    }
    if (hipStreamSynchronize(c->st) != hipSuccess) {
        rc = fail(c, PF_EHIP, "stream synchronize failed after allocation");
        goto bad;
    }
    {
        const char *v = std::getenv("PF_SOR_VARIANT");
        if (v && *v) c->sor_forced = std::atoi(v);
        // FP32 sweeps whose arrays (12 B per cell) fit well inside the 256 MB last-level
        // cache run 6-8 % faster with two cells per thread (all loads issued together);
        // past that size one cell per thread is faster (measured on 256x128x128, 192^3
        // and 512x256x256).  FP64 sweeps are fastest with one.
        const int64_t cells = (int64_t)g.m * g.n * (g.dim == 3 ? g.l : 1);
        c->sweep_kb32 = (cells * 12 <= ((int64_t)128 << 20)) ? 2 : 1;
        const char *kb = std::getenv("PF_SWEEP_KB");
        if (kb && *kb) {
            const int v = std::atoi(kb);
            c->sweep_kb = c->sweep_kb32 = (v == 1 || v == 2 || v == 4) ? v : 1;
        }
        // FP64 sweeps whose arrays (24 B per cell) are well past the last-level cache run
        // 7-9 % faster in hardware block order (512x256x256: 183 against 199 us per
        // sweep); at 384x192x192 the two orders tie and below that the remap wins.
        c->xcd_remap64 = (cells * 24 > ((int64_t)512 << 20)) ? 0 : 1;
        c->sweep_xw = pf::sweep_wave_width(g);
        const char *xw = std::getenv("PF_SWEEP_XW");
        if (xw && *xw) c->sweep_xw = (std::atoi(xw) == 32) ? 32 : 64;
        const char *x = std::getenv("PF_XCD_REMAP");
        if (x && *x) c->xcd_remap = c->xcd_remap64 = std::atoi(x) != 0;
    }
    if (desc->residual_tol > 0 && desc->solver == PF_SOLVER_MG) {
        pf::MgCase mc;
        mc.g = g;
        mc.a = c->a;
        mc.P = c->P;
        mc.BB = c->BB;
        mc.phi = c->phi;
        mc.PHIs = c->PHIs;
        mc.PHI32 = c->PHI32;
        mc.PHC = c->PHC;
        mc.PHTAB = c->PHTAB;
        mc.UNIT = c->UNIT;
        mc.UNIT32 = c->UNIT32;
        mc.kb32 = c->sweep_kb32;
        mc.kb64 = c->sweep_kb;
        mc.xw = c->sweep_xw;
        mc.remap32 = c->xcd_remap;
        mc.remap64 = c->xcd_remap64;
        std::string why;
        bool ok;
        if (desc->precision == PF_PREC_MIXED) {
            c->mg32 = new (std::nothrow) pf::MgSolver<float>();
            ok = c->mg32 && pf::mg_create(*c->mg32, mc, &why);
        } else {
            c->mg64 = new (std::nothrow) pf::MgSolver<double>();
            ok = c->mg64 && pf::mg_create(*c->mg64, mc, &why);
        }
        if (!ok) {
            rc = fail(c, PF_ENOMEM, why.empty() ? std::string("multigrid solver allocation failed") : why);
            goto bad;
        }
    }
    *out = c;
    return PF_OK;
bad:
    g_create_error = c->msg;
    release(c);
    delete c;
    return rc;
}

extern "C" int pf_destroy(pf_ctx *c) {
    if (!c) return PF_EINVAL;
    if (c->st) (void)hipStreamSynchronize(c->st);
    release(c);
    delete c;
    return PF_OK;
}

extern "C" const char *pf_last_error(const pf_ctx *c) {
    return c ? c->msg.c_str() : g_create_error.c_str();
}

extern "C" int pf_sync(pf_ctx *c) {
    if (!c) return PF_EINVAL;
    PF_TRY_HIP(c, hipStreamSynchronize(c->st));
    return PF_OK;
}

// ------------------------------------------------------------------------------------
// Porosity: grid_conditions* (src/omp_parallel/lib/grid.f90) on the host, once.
// ------------------------------------------------------------------------------------
// Time-invariant data derived from porosity: stored faces, the colour-split porosity and
// the choice of sweep kernel.
static int derive_from_porosity(pf_ctx *c) {
    const PfGeom &g = c->g;
    pf::launch_faces(c->st, g, c->phi, c->AW, c->AS, c->AB, c->d.dt, c->d.dx, c->d.dy, g.dim == 3 ? c->d.dz : 1.0);
    pf::launch_split(c->st, g, c->phi, c->PHIs);
    if (c->PHI32) pf::launch_split32(c->st, g, c->phi, c->PHI32);
    int *bad = nullptr;
    PF_TRY_HIP(c, hipMalloc((void **)&bad, sizeof(int)));
    int zero = 0, hbad = 1;
    PF_TRY_HIP(c, hipMemcpyAsync(bad, &zero, sizeof(int), hipMemcpyHostToDevice, c->st));
    pf::launch_faces_check(c->st, g, c->phi, c->AW, c->AS, c->AB, c->a, bad);
    hipError_t e = hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    (void)hipFree(bad);
    PF_TRY_HIP(c, e);
    PF_TRY_HIP(c, hipGetLastError());
    c->sor_variant = (hbad == 0) ? 2 : 1;
    if (c->sor_forced >= 0 && c->sor_forced <= 2 && !(c->sor_forced == 2 && hbad != 0)) c->sor_variant = c->sor_forced;
    c->unit_cells = 0;
    if (c->UNIT) {
        pf::launch_unit_flags(c->st, g, c->PHIs, c->UNIT);
        std::vector<uint8_t> h((size_t)unit_flags_size(g));
        PF_TRY_HIP(c, hipMemcpyAsync(h.data(), c->UNIT, h.size(), hipMemcpyDeviceToHost, c->st));
        PF_TRY_HIP(c, hipStreamSynchronize(c->st));
        // counted where the two-launch colour sweeps run (not the stored-face, fused or 2D
        // blocked sweeps, which read porosity or coefficients everywhere)
        if (c->sor_variant != 0 && !c->fused && c->tb2_k == 0)
            c->unit_cells = pf::unit_wave_cells(g, h, c->sweep_xw, g.dim == 3 ? c->sweep_kb : 1);
    }
    if (c->UNIT32 && c->PHI32) {
        pf::launch_unit_flags(c->st, g, c->PHI32, c->UNIT32);
        std::vector<uint8_t> h((size_t)unit_flags_size(g));
        PF_TRY_HIP(c, hipMemcpyAsync(h.data(), c->UNIT32, h.size(), hipMemcpyDeviceToHost, c->st));
        PF_TRY_HIP(c, hipStreamSynchronize(c->st));
        if (!c->fused) c->unit_cells = pf::unit_wave_cells(g, h, c->sweep_xw, g.dim == 3 ? c->sweep_kb32 : 1);
    }
    // coded porosity only where unit waves hold less than half the cells: a unit wave reads
    // no porosity, and at 512x256x256 with 65 % of the cells in unit waves the sweep on
    // stored FP64 porosity took 165 us against 195 us coded
    c->phi_coded = false;
    const int64_t ncell = (int64_t)g.m * g.n * (g.dim == 3 ? g.l : 1);
    if (c->PHC && c->PHTAB && (c->phi_code_forced || 2 * c->unit_cells < ncell)) {
        std::vector<double> h((size_t)(2 * g.HS));
        PF_TRY_HIP(c, hipMemcpyAsync(h.data(), c->PHIs, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
        PF_TRY_HIP(c, hipStreamSynchronize(c->st));
        std::vector<float> code;
        std::vector<double> tab;
        if (pf::code_porosity(h, code, tab, PHI_TAB_MAX)) {
            PF_TRY_HIP(c, hipMemcpy(c->PHC, code.data(), code.size() * sizeof(float), hipMemcpyHostToDevice));
            if (!tab.empty())
                PF_TRY_HIP(c, hipMemcpy(c->PHTAB, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
            c->phi_coded = true;
        }
    }
    // the multigrid operators follow the porosity (rebuilt at the next solve)
    const bool unit_ops = 2 * c->unit_cells >= ncell && c->unit_cells > 0;
    c->unit_half = unit_ops;
    if (c->mg32) {
        c->mg32->c.fast = (c->sor_variant == 2);
        c->mg32->c.unit_ops = unit_ops;
        c->mg32->built = false;
    }
    if (c->mg64) {
        c->mg64->c.fast = (c->sor_variant == 2);
        c->mg64->c.coded = c->phi_coded;
        c->mg64->c.unit_ops = unit_ops;
        c->mg64->built = false;
    }
    c->have_phi = true;
    return PF_OK;
}

extern "C" int pf_set_porosity(pf_ctx *c, const double *raw, double threshold) {
    if (!c || !raw) return PF_EINVAL;
    std::vector<double> phi;
    pf::host_porosity(c->g, c->d.bc, raw, threshold, phi);
    PF_TRY_HIP(c, hipMemcpyAsync(c->phi, phi.data(), phi.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
    return derive_from_porosity(c);
}

// ------------------------------------------------------------------------------------
// Field staging between the reference's host storage and the device layouts.
// ------------------------------------------------------------------------------------
static double *field_ptr(pf_ctx *c, int f, bool &split) {
    split = false;
    switch (f) {
        case PF_FIELD_U: return c->u;
        case PF_FIELD_V: return c->v;
        case PF_FIELD_W: return c->g.dim == 3 ? c->w : nullptr;
        case PF_FIELD_P: split = true; return c->P;
        case PF_FIELD_POROSITY: return c->phi;
        default: return nullptr;
    }
}

extern "C" int pf_upload(pf_ctx *c, int field, const double *host, int64_t ldi, int64_t ldj) {
    if (!c || !host) return PF_EINVAL;
    const PfGeom &g = c->g;
    if (ldi < g.m + 2 || ldj < g.n + 2) return fail(c, PF_EINVAL, "leading dimensions smaller than m+2, n+2");
    bool split;
    double *dst = field_ptr(c, field, split);
    if (!dst) return fail(c, PF_EINVAL, "unknown field");
    const int64_t size = split ? 2 * g.HS : g.NS;
    std::vector<double> stage((size_t)size, 0.0);
    PF_TRY_HIP(c, hipStreamSynchronize(c->st));
    if (split) PF_TRY_HIP(c, hipMemcpy(stage.data(), dst, (size_t)size * sizeof(double), hipMemcpyDeviceToHost));
    pf::from_host(g, split, host, ldi, ldj, stage);
    PF_TRY_HIP(c, hipMemcpy(dst, stage.data(), (size_t)size * sizeof(double), hipMemcpyHostToDevice));
    if (field == PF_FIELD_POROSITY) return derive_from_porosity(c);
    return PF_OK;
}

extern "C" int pf_download(pf_ctx *c, int field, double *host, int64_t ldi, int64_t ldj) {
    if (!c || !host) return PF_EINVAL;
    const PfGeom &g = c->g;
    if (ldi < g.m + 2 || ldj < g.n + 2) return fail(c, PF_EINVAL, "leading dimensions smaller than m+2, n+2");
    bool split;
    double *src = field_ptr(c, field, split);
    if (!src) return fail(c, PF_EINVAL, "unknown field");
    const int64_t size = split ? 2 * g.HS : g.NS;
    std::vector<double> stage((size_t)size);
    PF_TRY_HIP(c, hipStreamSynchronize(c->st));
    PF_TRY_HIP(c, hipMemcpy(stage.data(), src, (size_t)size * sizeof(double), hipMemcpyDeviceToHost));
    pf::to_host(g, split, stage, host, ldi, ldj);
    return PF_OK;
}

// ------------------------------------------------------------------------------------
// initial_conditions + boundary
// ------------------------------------------------------------------------------------
extern "C" int pf_initial(pf_ctx *c) {
    if (!c) return PF_EINVAL;
    if (!c->have_phi) return fail(c, PF_ESTATE, "pf_initial before pf_set_porosity");
    const PfGeom &g = c->g;
    const pf_desc &d = c->d;
    double u0, v0;
    pf::initial_values(d, u0, v0);
    std::vector<double> nat, spl, phi;
    if (d.inlet_porosity) {
        phi.resize((size_t)g.NS);
        PF_TRY_HIP(c, hipStreamSynchronize(c->st));
        PF_TRY_HIP(c, hipMemcpy(phi.data(), c->phi, phi.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    auto fill = [&](double val, bool split) { pf::fill_interior(g, val, split, split ? spl : nat); };
    fill(u0, false);
    if (d.inlet_porosity) pf::scale_interior(g, phi, nat);
    PF_TRY_HIP(c, hipMemcpy(c->u, nat.data(), nat.size() * sizeof(double), hipMemcpyHostToDevice));
    fill(v0, false);
    if (d.inlet_porosity) pf::scale_interior(g, phi, nat);
    PF_TRY_HIP(c, hipMemcpy(c->v, nat.data(), nat.size() * sizeof(double), hipMemcpyHostToDevice));
    if (g.dim == 3) {
        fill(0.0, false);
        PF_TRY_HIP(c, hipMemcpy(c->w, nat.data(), nat.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    fill(d.outlet_pressure, true);
    PF_TRY_HIP(c, hipMemcpy(c->P, spl.data(), spl.size() * sizeof(double), hipMemcpyHostToDevice));
    pf::launch_boundary(c->st, g, c->a, c->phi, c->u, c->v, c->w, c->P);
    PF_TRY_HIP(c, hipGetLastError());
    PF_TRY_HIP(c, hipStreamSynchronize(c->st));
    c->have_init = true;
    return PF_OK;
}

// ------------------------------------------------------------------------------------
// The time step
// ------------------------------------------------------------------------------------
static void enqueue_sor(pf_ctx *c, unsigned long long *slot, int iters) {
    const PfGeom &g = c->g;
    const bool periodic = (c->d.bc == PF_BC_UNIFORM);
    const int fy = periodic && (g.n % 2 == 0);
    const int fz = periodic && g.dim == 3 && (g.l % 2 == 0);
    const int ry = periodic && !fy;
    const int rz = periodic && g.dim == 3 && !fz;
    // 3D sweeps (i+j+k) even first (:493), 2D sweeps (i+j) odd first (:339)
    const int first = (g.dim == 3) ? 0 : 1;
    int done = 0;
    if (c->tb2_k > 0 && c->sor_variant != 0 && iters >= 2) {
        // K iterations per launch, P -> P2 -> P ...: an even number of launches leaves the
        // result in P, so a launch of two or more iterations is split when the count is odd
        std::vector<int> ks;
        for (int left = iters; left > 0;) {
            const int k = std::min(left, c->tb2_k);
            ks.push_back(k);
            left -= k;
        }
        if (ks.size() % 2 == 1)
            for (size_t q = ks.size(); q-- > 0;)
                if (ks[q] >= 2) {
                    const int half = ks[q] / 2;
                    ks[q] -= half;
                    ks.insert(ks.begin() + (std::ptrdiff_t)q + 1, half);
                    break;
                }
        if (ks.size() % 2 == 0) {
            for (size_t q = 0; q < ks.size(); ++q) {
                const double *in = (q % 2 == 0) ? c->P : c->P2;
                double *out = (q % 2 == 0) ? c->P2 : c->P;
                pf::launch_sor_tb2(c->st, g, c->a, first, ks[q], c->tb2_r, in, out, c->PHIs, c->BB, c->phi, slot,
                                   c->sor_variant == 2);
            }
            done = iters;
        }
    }
    if (c->fused && c->sor_variant != 0 && iters - done >= 2) {
        // pairs of fused iterations P -> P2 -> P; P2 takes P's ghost values first (the
        // x ghosts and the air room's faces are not written by the sweep)
        (void)hipMemcpyAsync(c->P2, c->P, sizeof(double) * (size_t)(2 * g.HS), hipMemcpyDeviceToDevice, c->st);
        const bool fast = (c->sor_variant == 2);
        const int track = 1;   // 3D tracks the second colour inside the kernel, 2D both
        for (; done + 2 <= iters; done += 2) {
            pf::launch_sor_fused(c->st, g, c->a, c->P, c->P2, c->PHIs, c->BB, c->phi, slot, track, fast);
            pf::launch_sor_fused(c->st, g, c->a, c->P2, c->P, c->PHIs, c->BB, c->phi, slot, track, fast);
        }
    }
    for (int it = done; it < iters; ++it) {
        for (int half = 0; half < 2; ++half) {
            pf::SweepArgs s{};
            s.colour = half == 0 ? first : 1 - first;
            s.fuse_y = fy;
            s.fuse_z = fz;
            // 3D reduces |dp| after the second sweep only (:575-583); 2D after both (:351, :385)
            s.track_err = (g.dim == 2) || (half == 1);
            s.xcd_remap = c->xcd_remap64;
            s.gx = s.gy = 1;
            s.kb = c->sweep_kb;
            s.xw = c->sweep_xw;
            if (c->sor_variant == 0)
                pf::launch_sor(c->st, g, c->a, s, c->P, c->AW, c->AS, c->AB, c->BB, c->phi, slot);
            else
                pf::launch_sor_phi(c->st, g, c->a, s, c->P, c->PHIs, c->BB, c->phi, slot, c->sor_variant == 2,
                                   c->phi_coded ? c->PHC : nullptr, c->PHTAB, c->UNIT);
            if (ry || rz) pf::launch_ghost_refresh(c->st, g, ry, rz, c->P);
        }
    }
}

// Mixed precision pressure solve of one step (see pf_mixed.hip).  Returns cycles run;
// *res receives the last evaluated residual (or -1): in tolerance mode the residual of the
// final pressure, in fixed-count mode (report_residual) the residual after the step's solve.
static int enqueue_mixed(pf_ctx *c, unsigned long long *slot, double *res, int *iters) {
    const PfGeom &g = c->g;
    const bool periodic = (c->d.bc == PF_BC_UNIFORM);
    const int fy = periodic && (g.n % 2 == 0);
    const int fz = periodic && g.dim == 3 && (g.l % 2 == 0);
    const int ry = periodic && !fy;
    const int rz = periodic && g.dim == 3 && !fz;
    const int first = (g.dim == 3) ? 0 : 1;
    const bool fast = (c->sor_variant == 2);
    const int K = c->d.inner_iters > 0 ? c->d.inner_iters : (c->d.iter_max > 0 ? c->d.iter_max : 1);
    const bool tol_mode = c->d.residual_tol > 0;
    const int maxc = tol_mode ? (c->d.max_cycles > 0 ? c->d.max_cycles : 1000) : (c->d.iter_max + K - 1) / K;
    int remaining = c->d.iter_max, cycles = 0;
    *res = -1.0;
    *iters = 0;
    auto residual = [&](bool norm) -> int {
        for (int half = 0; half < 2; ++half) {
            pf::SweepArgs s{};
            s.colour = half;
            s.fuse_y = fy;
            s.fuse_z = fz;
            s.track_err = norm ? 1 : 0;
            s.xcd_remap = c->xcd_remap;
            pf::launch_residual(c->st, g, c->a, s, c->P, c->PHIs, c->BB, c->phi, c->R32, c->E32, c->res_part, fast,
                                c->unit_half ? c->UNIT : nullptr);
        }
        if (!norm) return PF_OK;
        pf::launch_reduce_err(c->st, c->res_part, c->nparts, c->res_val);
        unsigned long long bits = 0;
        PF_TRY_HIP(c, hipMemcpyAsync(&bits, c->res_val, sizeof(bits), hipMemcpyDeviceToHost, c->st));
        PF_TRY_HIP(c, hipStreamSynchronize(c->st));
        std::memcpy(res, &bits, sizeof(double));
        return PF_OK;
    };
    // tolerance mode: a residual before every cycle and one after the last allowed cycle
    // (so *res is the residual of the returned pressure)
    for (int cyc = 0; tol_mode ? cyc <= maxc : cyc < maxc; ++cyc) {
        int rc;
        if (tol_mode) {
            if ((rc = residual(true)) != PF_OK) return rc;
            // a non-finite residual never converges: stop at once (pf_step reports it)
            if (*res <= c->d.residual_tol || cyc == maxc || !std::isfinite(*res)) break;
        } else {
            if (remaining <= 0) break;
            if ((rc = residual(false)) != PF_OK) return rc;
        }
        const int kk = tol_mode ? K : (remaining < K ? remaining : K);
        int done = 0;
        if (c->fused && kk >= 2) {
            // E2's non-periodic ghosts are 0 like E's (zeroed at allocation, never written)
            for (; done + 2 <= kk; done += 2) {
                pf::launch_sor32_fused(c->st, g, c->a, c->E32, c->E2, c->PHI32, c->R32, c->phi, slot, 1);
                pf::launch_sor32_fused(c->st, g, c->a, c->E2, c->E32, c->PHI32, c->R32, c->phi, slot, 1);
            }
        }
        for (int it = done; it < kk; ++it) {
            for (int half = 0; half < 2; ++half) {
                pf::SweepArgs s{};
                s.colour = half == 0 ? first : 1 - first;
                s.fuse_y = fy;
                s.fuse_z = fz;
                s.track_err = (g.dim == 2) || (half == 1);
                s.xcd_remap = c->xcd_remap;
                s.kb = c->sweep_kb32;
                s.xw = c->sweep_xw;
                pf::launch_sor32(c->st, g, c->a, s, c->E32, c->PHI32, c->R32, c->phi, slot, c->UNIT32);
                if (ry || rz) pf::launch_ghost_refresh32(c->st, g, ry, rz, c->E32);
            }
        }
        pf::launch_update(c->st, g, c->E32, c->P);
        remaining -= kk;
        *iters += kk;
        ++cycles;
    }
    if (!tol_mode && c->d.report_residual) {
        int rc = residual(true);
        if (rc != PF_OK) return rc;
    }
    return cycles;
}

// Tolerance-mode solve with the multigrid-preconditioned inner solver (pf_mg.h): cycles of
// the FP64 residual of the reference operator, an inner BiCGStab solve with V-cycles to
// inner_tol, and the FP64 update p += e, until max|r/ap| <= residual_tol.  The step's
// diagnostic (max |e| of its updates) is left in the solver's err_slots.
template <typename T>
static int enqueue_mg(pf_ctx *c, pf::MgSolver<T> &S, double *res, int *iters) {
    pf::MgControl ctl;
    ctl.tol = c->d.residual_tol;
    ctl.max_cycles = c->d.max_cycles > 0 ? c->d.max_cycles : 50;
    ctl.rtol = c->d.inner_tol > 0 ? c->d.inner_tol : 1e-3;
    ctl.maxit = c->d.inner_iters > 0 ? c->d.inner_iters : 30;
    std::vector<pf::MgStepStats> out;
    PF_TRY_HIP(c, pf::mg_step_solve<T>({&S}, {c->st}, ctl, out));
    *res = out[0].residual;
    *iters = out[0].vcycles;
    return out[0].cycles;
}

// FP64 solve to a residual tolerance: cycles of inner_iters reference SOR iterations with an
// FP64 residual check before each (the FP64 counterpart of the mixed tolerance mode).
static int enqueue_fp64_tol(pf_ctx *c, double *res, int *iters) {
    const PfGeom &g = c->g;
    const int K = c->d.inner_iters > 0 ? c->d.inner_iters : (c->d.iter_max > 0 ? c->d.iter_max : 1);
    const int maxc = c->d.max_cycles > 0 ? c->d.max_cycles : 1000;
    const bool fast = (c->sor_variant == 2);
    int cycles = 0;
    *iters = 0;
    for (int cyc = 0; cyc <= maxc; ++cyc) {
        for (int half = 0; half < 2; ++half) {
            pf::SweepArgs s{};
            s.colour = half;
            s.track_err = 1;
            s.xcd_remap = c->xcd_remap;
            pf::launch_residual(c->st, g, c->a, s, c->P, c->PHIs, c->BB, c->phi, (float *)nullptr, (float *)nullptr, c->res_part,
                                fast, c->unit_half ? c->UNIT : nullptr);
        }
        pf::launch_reduce_err(c->st, c->res_part, c->nparts, c->res_val);
        unsigned long long bits = 0;
        PF_TRY_HIP(c, hipMemcpyAsync(&bits, c->res_val, sizeof(bits), hipMemcpyDeviceToHost, c->st));
        PF_TRY_HIP(c, hipStreamSynchronize(c->st));
        std::memcpy(res, &bits, sizeof(double));
        if (*res <= c->d.residual_tol || cyc == maxc || !std::isfinite(*res)) break;
        enqueue_sor(c, c->err_part, K);
        *iters += K;
        ++cycles;
    }
    return cycles;
}

extern "C" int pf_step(pf_ctx *c, int nsteps, pf_stats *stats) {
    if (!c || nsteps < 0) return PF_EINVAL;
    if (!c->have_init) return fail(c, PF_ESTATE, "pf_step before pf_initial");
    if (nsteps == 0) return PF_OK;
    const PfGeom &g = c->g;
    if (c->err_cap < nsteps) {
        if (c->err) PF_TRY_HIP(c, hipFree(c->err));
        c->err = nullptr;
        PF_TRY_HIP(c, hipMalloc((void **)&c->err, sizeof(unsigned long long) * (size_t)nsteps));
        c->err_cap = nsteps;
    }
    pf::launch_clear_u64(c->st, c->err, nsteps);
    c->mixed_res.clear();
    c->mixed_cycles.clear();
    c->mixed_iters.clear();
    if (c->timing) {
        while (c->evs.size() < (size_t)(2 * nsteps)) {
            hipEvent_t e;
            PF_TRY_HIP(c, hipEventCreate(&e));
            c->evs.push_back(e);
        }
    }
    for (int s = 0; s < nsteps; ++s) {
        std::swap(c->u, c->uo);
        std::swap(c->v, c->vo);
        std::swap(c->w, c->wo);
        pf::launch_div(c->st, g, c->a, c->uo, c->vo, c->wo, c->div);
        pf::launch_predict(c->st, g, c->a, c->uo, c->vo, c->wo, c->div, c->phi, c->u, c->v, c->w);
        pf::launch_rhs(c->st, g, c->a, c->u, c->v, c->w, c->phi, c->P, c->BB);
        if (c->timing) PF_TRY_HIP(c, hipEventRecord(c->evs[(size_t)(2 * s)], c->st));
        if (c->d.precision == PF_PREC_MIXED) {
            double res = -1.0;
            int iters = 0;
            const int cyc = c->mg32 ? enqueue_mg(c, *c->mg32, &res, &iters)
                                    : enqueue_mixed(c, c->err_part, &res, &iters);
            if (cyc < 0) return cyc;
            c->mixed_res.push_back(res);
            c->mixed_cycles.push_back(cyc);
            c->mixed_iters.push_back(iters);
        } else if (c->d.residual_tol > 0) {
            double res = -1.0;
            int iters = 0;
            const int cyc = c->mg64 ? enqueue_mg(c, *c->mg64, &res, &iters) : enqueue_fp64_tol(c, &res, &iters);
            if (cyc < 0) return cyc;
            c->mixed_res.push_back(res);
            c->mixed_cycles.push_back(cyc);
            c->mixed_iters.push_back(iters);
        } else {
            enqueue_sor(c, c->err_part, c->d.iter_max);
        }
        if (c->mg32) pf::launch_reduce_err(c->st, c->mg32->err_slots, c->mg32->nslots, c->err + s);
        else if (c->mg64) pf::launch_reduce_err(c->st, c->mg64->err_slots, c->mg64->nslots, c->err + s);
        else pf::launch_reduce_err(c->st, c->err_part, c->nparts, c->err + s);
        if (c->timing) PF_TRY_HIP(c, hipEventRecord(c->evs[(size_t)(2 * s + 1)], c->st));
        pf::launch_project(c->st, g, c->a, c->P, c->u, c->v, c->w);
        pf::launch_boundary(c->st, g, c->a, c->phi, c->u, c->v, c->w, c->P);
        PF_TRY_HIP(c, hipGetLastError());
    }
    std::vector<unsigned long long> e((size_t)nsteps);
    PF_TRY_HIP(c, hipMemcpyAsync(e.data(), c->err, sizeof(unsigned long long) * (size_t)nsteps, hipMemcpyDeviceToHost, c->st));
    PF_TRY_HIP(c, hipStreamSynchronize(c->st));
    if (c->timing && c->d.iter_max > 0) {
        double sweep_ms = 0.0;
        for (int s = 0; s < nsteps; ++s) {
            float ms = 0.f;
            PF_TRY_HIP(c, hipEventElapsedTime(&ms, c->evs[(size_t)(2 * s)], c->evs[(size_t)(2 * s + 1)]));
            sweep_ms += ms;
        }
        c->solve_us_avg = sweep_ms * 1000.0 / (double)nsteps;
        c->sweep_launches = (int64_t)nsteps * 2 * c->d.iter_max;
        c->sweep_us_avg = sweep_ms * 1000.0 / (double)c->sweep_launches;
    }
    bool finite = true;
    for (int s = 0; s < nsteps; ++s) {
        double v;
        std::memcpy(&v, &e[(size_t)s], sizeof(double));
        const bool cyc = (c->d.precision == PF_PREC_MIXED) || c->d.residual_tol > 0;
        if (!std::isfinite(v) || (cyc && !std::isfinite(c->mixed_res[(size_t)s]))) finite = false;
        if (stats) {
            stats[s].sor_error = v;
            stats[s].residual = cyc ? c->mixed_res[(size_t)s] : -1.0;
            stats[s].sor_iters = cyc ? c->mixed_iters[(size_t)s] : c->d.iter_max;
            stats[s].cycles = cyc ? c->mixed_cycles[(size_t)s] : 0;
        }
    }
    if (!finite) return fail(c, PF_EDIVERGED, "non-finite SOR error or residual: the run diverged");
    return PF_OK;
}

extern "C" int pf_divergence(pf_ctx *c, double *out) {
    if (!c || !out) return PF_EINVAL;
    const PfGeom &g = c->g;
    pf::launch_divergence_out(c->st, g, c->a, c->phi, c->u, c->v, c->w, c->dout);
    PF_TRY_HIP(c, hipGetLastError());
    const size_t n = (size_t)g.m * g.n * g.l;
    PF_TRY_HIP(c, hipMemcpyAsync(out, c->dout, n * sizeof(double), hipMemcpyDeviceToHost, c->st));
    PF_TRY_HIP(c, hipStreamSynchronize(c->st));
    return PF_OK;
}

namespace pf {
// The body force of the context's current fields, reduced in the device buffer `work`
// (force_blocks(g) * 6 partial sums, then the six totals), the totals copied to host `out`.
int force_into(pf_ctx *c, double *work, double *out) {
    const PfGeom &g = c->g;
    double *tot = work + (size_t)pf::force_blocks(g) * 6;
    // every partial the final reduction reads is written first, so nothing needs clearing
    pf::launch_force(c->st, g, c->a, c->phi, c->u, c->v, c->w, c->P, work, tot);
    PF_TRY_HIP(c, hipGetLastError());
    PF_TRY_HIP(c, hipMemcpyAsync(out, tot, 6 * sizeof(double), hipMemcpyDeviceToHost, c->st));
    PF_TRY_HIP(c, hipStreamSynchronize(c->st));
    return PF_OK;
}
bool ctx_ready(const pf_ctx *c) { return c->have_init; }
int ctx_device(const pf_ctx *c) { return c->device; }
const PfGeom &ctx_geom(const pf_ctx *c) { return c->g; }
int ctx_fail(pf_ctx *c, int code, const std::string &m) { return fail(c, code, m); }
}  // namespace pf

extern "C" int pf_force(pf_ctx *c, double *out) {
    if (!c || !out) return PF_EINVAL;
    if (!c->have_init) return fail(c, PF_ESTATE, "pf_force before pf_initial");
    const int nb = pf::force_blocks(c->g);
    if (!c->force_buf)
        PF_TRY_HIP(c, hipMalloc((void **)&c->force_buf, sizeof(double) * ((size_t)nb * 6 + 6)));
    return pf::force_into(c, c->force_buf, out);
}

extern "C" int pf_sor_fused(const pf_ctx *c) {
    if (!c) return PF_EINVAL;
    if (!c->fused) return 0;
    return (c->d.precision == PF_PREC_MIXED || c->sor_variant != 0) ? 1 : 0;
}

extern "C" int pf_set_timing(pf_ctx *c, int on) {
    if (!c) return PF_EINVAL;
    c->timing = on != 0;
    return PF_OK;
}

extern "C" int pf_last_sweep_time_us(const pf_ctx *c, double *avg_us, int64_t *launches) {
    if (!c || !avg_us || !launches) return PF_EINVAL;
    *avg_us = c->sweep_us_avg;
    *launches = c->sweep_launches;
    return PF_OK;
}

extern "C" int pf_sor_kernel(const pf_ctx *c) { return c ? c->sor_variant : PF_EINVAL; }

extern "C" int pf_sor_coded(const pf_ctx *c) { return c ? (c->phi_coded ? 1 : 0) : PF_EINVAL; }

extern "C" int pf_sor_unit_cells(const pf_ctx *c, int64_t *cells) {
    if (!c || !cells) return PF_EINVAL;
    *cells = c->unit_cells;
    return PF_OK;
}

// Bandwidth calibration: copies `bytes` (rounded down to 16) device to device `iters` times
// on a private stream and returns the average GB/s (read + write bytes per second).
extern "C" int pf_probe_copy(int64_t bytes, int iters, double *gbps) {
    if (bytes < 16 || iters < 1 || !gbps) return PF_EINVAL;
    const int64_t n = (bytes / 16) * 2;
    double *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PF_OK;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return PF_EHIP;
    if (hipMalloc((void **)&a, (size_t)n * 8) != hipSuccess || hipMalloc((void **)&b, (size_t)n * 8) != hipSuccess) {
        rc = PF_ENOMEM;
    } else if (hipMemsetAsync(a, 0, (size_t)n * 8, st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
               hipEventCreate(&e1) != hipSuccess) {
        rc = PF_EHIP;
    } else {
        pf::launch_copy_probe(st, a, b, n);  // warm-up
        (void)hipEventRecord(e0, st);
        for (int t = 0; t < iters; ++t) pf::launch_copy_probe(st, a, b, n);
        (void)hipEventRecord(e1, st);
        float ms = 0.f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = PF_EHIP;
        else *gbps = (double)(2 * n * 8) * iters / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    (void)hipStreamDestroy(st);
    return rc;
}
