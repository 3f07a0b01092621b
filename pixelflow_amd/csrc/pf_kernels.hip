// pf_kernels.hip -- gfx950 kernels of the FP64 time step (2D and 3D, uniform and air).
//
// Every arithmetic expression keeps the reference's Fortran evaluation order (a*b/c*d is
// ((a*b)/c)*d, -a-b is (-a)-b) and is compiled with -ffp-contract=off and IEEE division,
// so a cell's result is the same IEEE FP64 value the reference computes.  Kernels are
// HBM-bandwidth bound stencils: one thread per cell, x-contiguous wave rows, no MFMA.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "pf_cells.h"
#include "pf_device.h"
#include "pf_fused.h"
#include "pf_kernels.h"

namespace pf {


// ------------------------------------------------------------------------------------
// Poisson face coefficients (time invariant; solve_p 3D :390-400, 2D :265-268).
// AW(i) = dt*max(small,(phi(i)+phi(i-1))*0.5)/dx/dx == aw(i) == ae(i-1), likewise AS, AB.
// ------------------------------------------------------------------------------------
__global__ void k_faces(PfGeom g, const double *__restrict__ phi, double *__restrict__ AW,
                        double *__restrict__ AS, double *__restrict__ AB, double dt, double dx,
                        double dy, double dz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = (g.dim == 3) ? (int)blockIdx.z + 1 : 0;
    if (i > g.m + 1 || j > g.n + 1) return;
    const int64_t o = pf_off(g, i, j, k);
    const int64_t so = pf_soff(g, i, j, k);
    const double pc = phi[o];
    AW[so] = ((dt * fmax_ref(PF_SMALL, (pc + phi[o - 1]) * 0.5)) / dx) / dx;
    AS[so] = ((dt * fmax_ref(PF_SMALL, (pc + phi[o - g.PX]) * 0.5)) / dy) / dy;
    if (g.dim == 3)
        AB[so] = ((dt * fmax_ref(PF_SMALL, (pc + phi[o - (int64_t)g.PX * g.Y]) * 0.5)) / dz) / dz;
}

// ------------------------------------------------------------------------------------
// div of the old velocity with its ghost values (3D :186-222, air :189-237, 2D :172-194).
// Ghost values are recomputed from the wrapped interior cell rather than copied, so one
// launch produces interior and ghosts with no intra-kernel dependency.
// ------------------------------------------------------------------------------------


// Threads of the x dimension cover i = 1..m; those at i = 1 and i = m also take the ghost
// cells i = 0 and i = m + 1 of their row, so that a row of 256 cells fills four waves instead
// of leaving a fifth one with two live lanes.
__global__ void k_div(PfGeom g, StepArgs a, const double *__restrict__ uo, const double *__restrict__ vo,
                      const double *__restrict__ wo, double *__restrict__ div) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = (g.dim == 3) ? (int)blockIdx.z : 0;
    if (i > g.m || j > g.n + 1) return;
    if (i == 1) div_cell(g, a, uo, vo, wo, div, 0, j, k);
    if (i == g.m) div_cell(g, a, uo, vo, wo, div, g.m + 1, j, k);
    div_cell(g, a, uo, vo, wo, div, i, j, k);
}

// ------------------------------------------------------------------------------------
// Momentum predictor (3D :228-381, air :276-428, 2D :200-258).  Writes every cell of the
// new velocity: interior cells predicted, ghost cells copied from the old velocity (the
// reference predicts the interior of u after u_old = u, so the ghosts keep u_old's).
// ------------------------------------------------------------------------------------


__global__ void k_predict3(PfGeom g, StepArgs a, const double *__restrict__ uo, const double *__restrict__ vo,
                           const double *__restrict__ wo, const double *__restrict__ div,
                           const double *__restrict__ po, double *__restrict__ u, double *__restrict__ v,
                           double *__restrict__ w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;  // x ghosts as in k_div
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > g.m || j > g.n + 1) return;
    if (i == 1) predict3_cell(g, a, uo, vo, wo, div, po, u, v, w, 0, j, k);
    if (i == g.m) predict3_cell(g, a, uo, vo, wo, div, po, u, v, w, g.m + 1, j, k);
    predict3_cell(g, a, uo, vo, wo, div, po, u, v, w, i, j, k);
}



__global__ void k_predict2(PfGeom g, StepArgs a, const double *__restrict__ uo, const double *__restrict__ vo,
                           const double *__restrict__ div, const double *__restrict__ po,
                           double *__restrict__ u, double *__restrict__ v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;  // x ghosts as in k_div
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > g.m || j > g.n + 1) return;
    if (i == 1) predict2_cell(g, a, uo, vo, div, po, u, v, 0, j);
    if (i == g.m) predict2_cell(g, a, uo, vo, div, po, u, v, g.m + 1, j);
    predict2_cell(g, a, uo, vo, div, po, u, v, i, j);
}

// ------------------------------------------------------------------------------------
// RHS bb of the predicted velocity (3D :404-409, 2D :271-274) with the outlet terms of
// boundary_matrix, written in the colour-split layout.
// ------------------------------------------------------------------------------------


__global__ void k_rhs(PfGeom g, StepArgs a, const double *__restrict__ u, const double *__restrict__ v,
                      const double *__restrict__ w, const double *__restrict__ po, const double *__restrict__ P,
                      double *__restrict__ BB) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = (g.dim == 3) ? (int)blockIdx.z + 1 : 0;
    if (i > g.m || j > g.n) return;
    rhs_cell(g, a, u, v, w, po, P, BB, i, j, k);
}

// ------------------------------------------------------------------------------------
// Red/black SOR half-sweep (3D :492-516 / :549-572, 2D :339-352 / :373-386).
// One launch updates every interior cell of colour c in place.  This equals the
// reference's p_old form: all neighbours of a colour-c cell have colour 1-c and are not
// written during the sweep; ghost cells are read as stored.  With an even periodic extent
// the ghost copy of a freshly updated boundary cell has the same colour and no cell of
// this sweep reads it, so the refresh of the reference (:463-480) is fused here as a
// second store; odd extents use k_ghost_refresh after the sweep instead.
// The SOR error diagnostic is the max |p - p_old| over the updated cells, reduced per
// wave with DPP cross-lane steps (wave_umax), per block through LDS, then one atomic max
// per block (this stored-face kernel; the porosity sweeps keep one slot per wave).
// ------------------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(256) void k_sor(PfGeom g, StepArgs a, SweepArgs s, const double *__restrict__ Poth,
                                             double *__restrict__ Pown, const double *__restrict__ AWo,
                                             const double *__restrict__ AWx, const double *__restrict__ ASo,
                                             const double *__restrict__ ASx, const double *__restrict__ ABo,
                                             const double *__restrict__ ABx, const double *__restrict__ BBo,
                                             const double *__restrict__ phi, unsigned long long *err_slot) {
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = (DIM == 3) ? (int)blockIdx.z + 1 : 0;
    const int sh = (s.colour + j + k) & 1;
    // thread t of a row takes the t-th interior cell of the colour: i = 2t+1 (sh = 1) or 2t+2
    const int h = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1 - sh;
    const int i = 2 * h + sh;
    const bool active = (j <= g.n) && (i <= g.m);
    double err = 0.;
    if (active) {
        const int64_t HP = g.HP, HPY = (int64_t)g.HP * g.Y;
        const int64_t o = (int64_t)h + HP * ((int64_t)j + (int64_t)g.Y * k);
        const double pc = Pown[o];
        const double pW = Poth[o - 1 + sh], pE = Poth[o + sh];
        const double pS = Poth[o - HP], pN = Poth[o + HP];
        Coef c;
        c.aw = AWo[o];
        c.ae = AWx[o + sh];
        c.as = ASo[o];
        c.an = ASx[o + HP];
        double pB = 0., pT = 0.;
        if (DIM == 3) {
            pB = Poth[o - HPY];
            pT = Poth[o + HPY];
            c.ab = ABo[o];
            c.at = ABx[o + HPY];
        } else {
            c.ab = 0.;
            c.at = 0.;
        }
        double ap;
        if (DIM == 3) ap = (((((-c.ae) - c.aw) - c.an) - c.as) - c.at) - c.ab;
        else ap = (((-c.ae) - c.aw) - c.an) - c.as;
        if (is_edge_row(g, a, i, j, k)) row_rules<false>(g, a, i, j, k, c, nullptr, nullptr, phi, 0.);
        const double bb = BBo[o];
        double num;
        if (DIM == 3)
            num = (((((bb - (c.ae * pE)) - (c.aw * pW)) - (c.an * pN)) - (c.as * pS)) - (c.at * pT)) - (c.ab * pB);
        else
            num = (((bb - (c.ae * pE)) - (c.aw * pW)) - (c.an * pN)) - (c.as * pS);
        const double pn = ((num / ap) * a.omega) + (pc * a.omega_c);
        Pown[o] = pn;
        if (s.fuse_y) {
            if (j == g.n) Pown[o - HP * (int64_t)g.n] = pn;          // ghost (i, 0, k)
            if (j == 1) Pown[o + HP * (int64_t)g.n] = pn;            // ghost (i, n+1, k)
        }
        if (DIM == 3 && s.fuse_z) {
            if (k == g.l) Pown[o - HPY * (int64_t)g.l] = pn;         // ghost (i, j, 0)
            if (k == 1) Pown[o + HPY * (int64_t)g.l] = pn;           // ghost (i, j, l+1)
        }
        err = fabs(pn - pc);
    }
    if (!s.track_err) return;
    // |dp| >= 0, so its IEEE bit pattern orders like an unsigned integer
    const unsigned long long e = wave_umax(__double_as_longlong(err));
    __shared__ unsigned long long red[4];
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    if ((lane & 63) == 0) red[lane >> 6] = e;
    __syncthreads();
    if (lane == 0) {
        unsigned long long b = red[0];
        const int nw = (blockDim.x * blockDim.y + 63) >> 6;
        for (int q = 1; q < nw; ++q) b = (red[q] > b) ? red[q] : b;
        // one slot per block: no two blocks of a launch touch the same address
        const int64_t bid = (int64_t)blockIdx.x + (int64_t)gridDim.x * ((int64_t)blockIdx.y + (int64_t)gridDim.y * blockIdx.z);
        unsigned long long *slot = err_slot + bid;
        atomicMax(slot, b);   // result unread: the block need not wait for the slot
    }
}

// Max of the partial slots of one step into that step's entry; the slots are cleared
// for the next step.  One block.
__global__ void k_reduce_err(unsigned long long *slots, int nslots, unsigned long long *step_err) {
    // eight independent loads in flight per thread before any slot is cleared (a sweep of
    // the 256x128x128 grid leaves 32768 wave slots: four rounds instead of 32 dependent ones)
    constexpr int U = 8;
    unsigned long long e = 0ull;
    for (int base = threadIdx.x; base < nslots; base += U * blockDim.x) {
        unsigned long long v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = base + u * (int)blockDim.x;
            v[u] = (t < nslots) ? slots[t] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = base + u * (int)blockDim.x;
            e = (v[u] > e) ? v[u] : e;
            if (t < nslots) slots[t] = 0ull;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o2 = __shfl_xor(e, off, 64);
        e = (o2 > e) ? o2 : e;
    }
    __shared__ unsigned long long red[16];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long b = red[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) b = (red[q] > b) ? red[q] : b;
        *step_err = b;
    }
}

// Variant of the sweep that rebuilds the six face coefficients of each updated cell from
// porosity (helpers in pf_device.h).
// One cell of the sweep (FP64, faces rebuilt from porosity); returns |p_new - p_old|.

// PT = double: porosity as stored FP64; PT = float: porosity coded in 4 bytes
// (phi_decode, table `tab`), the same doubles for 4 fewer bytes per cell and launch.
// UNIT: `unit` holds the unit-porosity chunk flags (unit_index in pf_layout.h); a wave whose
// lanes all lie in flagged chunks skips its porosity loads and the face arithmetic.
template <int DIM, bool FAST, int KB, typename PT, bool UNIT>
__global__ __launch_bounds__(256) void k_sor_phi(PfGeom g, StepArgs a, SweepArgs s, const double *__restrict__ Poth,
                                                 double *__restrict__ Pown, const PT *__restrict__ PHIo,
                                                 const PT *__restrict__ PHIx, const double *__restrict__ BBo,
                                                 const double *__restrict__ phi, unsigned long long *err_slot,
                                                 const double *__restrict__ tab, const uint8_t *__restrict__ unit) {
    const SweepThread t = sweep_thread<DIM, KB>(g, s);
    const bool active = t.j <= g.n && t.i <= g.m;
    double err = 0.;
    SweepIn<double> x[KB];
    int64_t o[KB];
    int ok = 1;
    if (active) {
        // the cells' unit flags first (loads return in order: the vote waits for these
        // only), then the pressure and RHS operands, all in flight together; the flags are
        // combined after the operand loads are issued
        uint32_t fl[KB];
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            const int k = (DIM == 3 && t.kbase + 2 * q > g.l) ? g.l : t.kbase + 2 * q;
            fl[q] = UNIT ? (uint32_t)gmem(unit)[unit_index(g, s.colour, t.j, k, t.h + t.sh - 1)] : 0u;
        }
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            // a plane past the last one loads plane l instead (in bounds) and is not updated
            const int k = (DIM == 3 && t.kbase + 2 * q > g.l) ? g.l : t.kbase + 2 * q;
            o[q] = (int64_t)t.h + (int64_t)g.HP * ((int64_t)t.j + (int64_t)g.Y * k);
            sweep_gather_unit<DIM>(g, t.sh, o[q], Poth, Pown, BBo, x[q]);
        }
#pragma unroll
        for (int q = 0; q < KB; ++q)
            if (!(DIM == 3 && t.kbase + 2 * q > g.l)) ok &= (int)fl[q];
    }
    // the same in every lane, so the branches below are the wave's
    const bool all_unit = UNIT && __all(ok) != 0;
    if (active && all_unit) {
        // every porosity operand is 1: the unit faces, and num / ap by the wave's divisor (ap
        // is the same for every cell of such a wave)
        const DivK kap = divk_make(a.unit_ap);
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            const int k = t.kbase + 2 * q;
            if (DIM == 3 && k > g.l) break;
            const SweepIn<double> &c = x[q];
            DivFast dv;
            double pn = sor_unit_value<DIM>(g, a, t.i, t.j, k, c.c, c.w, c.e, c.s, c.n, c.b, c.t, c.rhs, phi, dv, kap);
            if (dv.slow)
                pn = sor_value<DIM, FAST>(g, a, t.i, t.j, k, c.c, c.w, c.e, c.s, c.n, c.b, c.t, 1., 1., 1., 1., 1.,
                                          (DIM == 3) ? 1. : 0., (DIM == 3) ? 1. : 0., c.rhs, phi);
            sweep_store<DIM>(g, s, t.j, k, o[q], pn, Pown);
            err = max_keep_nan(err, fabs(pn - c.c));
        }
    } else if (active) {
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            if constexpr (std::is_same<PT, float>::value)
                sweep_gather_phi_coded<DIM>(g, t.sh, o[q], PHIo, PHIx, tab, x[q]);
            else
                sweep_gather_phi<DIM>(g, t.sh, o[q], PHIo, PHIx, x[q]);
        }
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            const int k = t.kbase + 2 * q;
            if (DIM == 3 && k > g.l) break;
            const SweepIn<double> &c = x[q];
            const double pn = sor_value<DIM, FAST>(g, a, t.i, t.j, k, c.c, c.w, c.e, c.s, c.n, c.b, c.t, c.fc, c.fw,
                                                   c.fe, c.fs, c.fn, c.fb, c.ft, c.rhs, phi);
            sweep_store<DIM>(g, s, t.j, k, o[q], pn, Pown);
            const double e = fabs(pn - c.c);
            err = max_keep_nan(err, e);
        }
    }
    if (s.track_err) wave_max_to_slot(err, err_slot);
}

// ------------------------------------------------------------------------------------
// Temporally blocked 2D colour sweeps (FP64, faces from porosity) for grids whose sweeps
// are launch-bound.  A block owns R rows and holds them with 2K halo rows on each side
// (rows wrap periodically in y; n even) in LDS: p and bb of both colours and every cell's
// coefficients, built once per launch from porosity by sor_coefs.  It
// runs K red/black iterations there: every half-sweep updates the rows whose neighbours
// are still exact, one row fewer on each side per half-sweep, so after 2K half-sweeps the
// owned rows hold what K iterations of the two-launch sweep give, bit for bit (the same
// sor_value on the same operands; a half-sweep's updates read only the other colour).  The
// owned rows and their periodic mirrors go to the other pressure buffer: blocks do not wait
// for each other, so a block must not overwrite rows another block has yet to read.  |dp|
// is folded over the owned rows only, for each of the 2K half-sweeps, as 2D tracks both
// colours.  Rows are (r, h) = (x / H, x % H) with x / H = umulhi(x, hmul) (tb2_magic).
// ------------------------------------------------------------------------------------
PF_HD uint32_t tb2_magic(uint32_t d) { return d <= 1 ? 0u : (uint32_t)((((uint64_t)1) << 32) / d + 1); }

template <bool FAST>
__global__ __launch_bounds__(TB2_THREADS) void k_sor_tb2(PfGeom g, StepArgs a, int first, int K, int R,
                                                         uint32_t hmul, const double *__restrict__ Pin,
                                                         double *__restrict__ Pout, const double *__restrict__ PHIs,
                                                         const double *__restrict__ BB,
                                                         const double *__restrict__ phi,
                                                         unsigned long long *err_slot) {
    extern __shared__ double tb[];
    __shared__ int jrow[TB2_MAX_ROWS];          // the grid row held in each local row
    const int H = g.H, NR = R + 4 * K;
    const uint32_t NH = (uint32_t)NR * (uint32_t)H;
    // per colour: p, bb and the cell's coefficients (ae, aw, an, as after the boundary
    // rows' rules, and ap), computed once per launch from porosity
    double *lp[2], *lb[2], *le[2], *lw[2], *ln[2], *ls[2], *la[2];
    for (int c = 0; c < 2; ++c) {
        double *base = tb + (size_t)c * TB2_ARRAYS / 2 * NH;
        lp[c] = base;
        lb[c] = base + NH;
        le[c] = base + 2 * NH;
        lw[c] = base + 3 * NH;
        ln[c] = base + 4 * NH;
        ls[c] = base + 5 * NH;
        la[c] = base + 6 * NH;
    }
    const int jb = (int)blockIdx.x * R + 1;   // the first owned row
    const int v0 = jb - 2 * K;                // the row held in local row 0
    const int64_t HS = g.HS, HP = g.HP;
    const uint32_t tid = threadIdx.x;
    for (int r = (int)tid; r < NR; r += TB2_THREADS) jrow[r] = ((v0 + r - 1) % g.n + g.n) % g.n + 1;
    __syncthreads();
    for (uint32_t x = tid; x < NH; x += TB2_THREADS) {
        const uint32_t r = hmul ? __umulhi(x, hmul) : x;
        const int h = (int)(x - r * (uint32_t)H);
        const int j = jrow[r];
        const int64_t o = (int64_t)h + HP * j;
        lp[0][x] = Pin[o];
        lp[1][x] = Pin[HS + o];
        lb[0][x] = BB[o];
        lb[1][x] = BB[HS + o];
        if (r < 1 || (int)r > NR - 2) continue;   // the first and last local rows are never updated
        const int64_t os = (int64_t)h + HP * jrow[r - 1], on = (int64_t)h + HP * jrow[r + 1];
        for (int c = 0; c < 2; ++c) {
            const int sh = (c + j) & 1;
            const int i = 2 * h + sh;
            if (i < 1 || i > g.m) continue;
            const int64_t own = (int64_t)c * HS, oth = (int64_t)(1 - c) * HS;
            Coef cf;
            const double ap = sor_coefs<2, FAST>(g, a, i, j, 0, PHIs[own + o], PHIs[oth + o - 1 + sh],
                                                 PHIs[oth + o + sh], PHIs[oth + os], PHIs[oth + on], 0., 0., phi, cf);
            le[c][x] = cf.ae;
            lw[c][x] = cf.aw;
            ln[c][x] = cf.an;
            ls[c][x] = cf.as;
            la[c][x] = ap;
        }
    }
    __syncthreads();
    double err = 0.;
    for (int s = 1; s <= 2 * K; ++s) {
        const int c = ((s & 1) == 1) ? first : 1 - first;
        const double *Xo = lp[1 - c];
        double *Xc = lp[c];
        const uint32_t x0 = (uint32_t)s * (uint32_t)H, x1 = (uint32_t)(NR - s) * (uint32_t)H;
        for (uint32_t x = x0 + tid; x < x1; x += TB2_THREADS) {
            const uint32_t r = hmul ? __umulhi(x, hmul) : x;
            const int h = (int)(x - r * (uint32_t)H);
            const int sh = (c + jrow[r]) & 1;
            const int i = 2 * h + sh;
            if (i < 1 || i > g.m) continue;
            const double pc = Xc[x];
            Coef cf;
            cf.ae = le[c][x];
            cf.aw = lw[c][x];
            cf.an = ln[c][x];
            cf.as = ls[c][x];
            cf.at = cf.ab = 0.;
            const double pn = sor_apply<2>(a, cf, la[c][x], lb[c][x], pc, Xo[x - 1 + sh], Xo[x + sh], Xo[x - H],
                                           Xo[x + H], 0., 0.);
            Xc[x] = pn;
            if ((int)r >= 2 * K && (int)r < 2 * K + R) err = max_keep_nan(err, fabs(pn - pc));
        }
        __syncthreads();
    }
    // the owned rows (both colours, every slot) and their periodic mirrors
    const uint32_t xo0 = (uint32_t)(2 * K) * (uint32_t)H, xo1 = (uint32_t)(2 * K + R) * (uint32_t)H;
    for (uint32_t x = xo0 + tid; x < xo1; x += TB2_THREADS) {
        const uint32_t r = hmul ? __umulhi(x, hmul) : x;
        const uint32_t h = x - r * (uint32_t)H;
        const int j = v0 + (int)r;
        if (j > g.n) continue;
        const int64_t o = (int64_t)h + HP * j;
        Pout[o] = lp[0][x];
        Pout[HS + o] = lp[1][x];
        if (j == 1) {
            Pout[o + HP * g.n] = lp[0][x];
            Pout[HS + o + HP * g.n] = lp[1][x];
        }
        if (j == g.n) {
            Pout[o - HP * g.n] = lp[0][x];
            Pout[HS + o - HP * g.n] = lp[1][x];
        }
    }
    wave_max_to_slot(err, err_slot);
}

int64_t sor_tb2_lds_bytes(const PfGeom &g, int K, int R) { return (int64_t)TB2_ARRAYS * (R + 4 * K) * g.H * 8; }

bool sor_tb2_ok(const PfGeom &g, int bc, int K, int R) {
    if (g.dim != 2 || bc != PF_BC_UNIFORM || (g.n % 2) != 0 || K < 1 || R < 1) return false;
    if (R + 4 * K > g.n || R + 4 * K > TB2_MAX_ROWS) return false;   // distinct rows of the grid
    if (sor_tb2_lds_bytes(g, K, R) > TB2_LDS_MAX) return false;
    // the row split x / H by multiplier is exact for every x a block uses (x * H < 2^32)
    const uint64_t nh = (uint64_t)(R + 4 * K) * (uint64_t)g.H;
    return nh * (uint64_t)g.H < ((uint64_t)1 << 32);
}

void launch_sor_tb2(hipStream_t st, const PfGeom &g, const StepArgs &a, int first, int K, int R, const double *Pin,
                    double *Pout, const double *PHIs, const double *BB, const double *phi, unsigned long long *slot,
                    bool fast) {
    const size_t lds = (size_t)sor_tb2_lds_bytes(g, K, R);
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)k_sor_tb2<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)TB2_LDS_MAX);
        (void)hipFuncSetAttribute((const void *)k_sor_tb2<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)TB2_LDS_MAX);
        attr_set = true;
    }
    const unsigned blocks = (unsigned)((g.n + R - 1) / R);
    const uint32_t hmul = tb2_magic((uint32_t)g.H);
    if (fast)
        hipLaunchKernelGGL(k_sor_tb2<true>, dim3(blocks), dim3(TB2_THREADS), lds, st, g, a, first, K, R, hmul, Pin, Pout, PHIs,
                           BB, phi, slot);
    else
        hipLaunchKernelGGL(k_sor_tb2<false>, dim3(blocks), dim3(TB2_THREADS), lds, st, g, a, first, K, R, hmul, Pin, Pout,
                           PHIs, BB, phi, slot);
}

// Porosity (natural layout, ghosts included) -> colour-split copy.
__global__ void k_split(PfGeom g, const double *__restrict__ src, double *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > g.m + 1 || j > g.n + 1) return;
    dst[pf_soff(g, i, j, k)] = src[pf_off(g, i, j, k)];
}

// Unit-porosity chunk flags (unit_index in pf_layout.h) of a colour-split porosity of
// type T (the FP64 one, or the FP32 one of the mixed-precision sweeps, whose unit values
// are those that round to 1).  One thread per slot position of every (colour, plane, row);
// a row holds an even number of chunks, so each wave covers two whole chunks of one row and
// a ballot of the failing lanes gives both flags.  Rows and planes outside the interior get
// 0 (no sweep reads them).
template <typename T>
__global__ __launch_bounds__(256) void k_unit_flags(PfGeom g, const T *__restrict__ PHIs, uint8_t *__restrict__ unit,
                                                    int64_t nslots) {
    // nslots is a multiple of 64: every wave lies wholly inside the slots or wholly past them
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int RW = unit_chunks(g) * 32;
    const int x = (int)(t % RW);
    const int64_t r = t / RW;
    const int j = (int)(r % g.Y);
    const int64_t r2 = r / g.Y;
    const int k = (int)(r2 % g.Z);
    const int c = (int)(r2 / g.Z);
    const bool interior = t < nslots && j >= 1 && j <= g.n && (g.dim == 2 || (k >= 1 && k <= g.l));
    bool fail = !interior;
    if (interior) {
        const int sh = (c + j + k) & 1;
        const int h = x + 1 - sh, i = 2 * h + sh;   // i >= 1 for every x >= 0
        if (i <= g.m) {
            const int64_t HP = g.HP, HPY = (int64_t)g.HP * g.Y;
            const T *own = PHIs + (int64_t)c * g.HS, *oth = PHIs + (int64_t)(1 - c) * g.HS;
            const T one = (T)1;
            const int64_t o = pf_hoff(g, h, j, k);
            bool u = own[o] == one && oth[o - 1 + sh] == one && oth[o + sh] == one && oth[o - HP] == one &&
                     oth[o + HP] == one;
            if (g.dim == 3) u = u && oth[o - HPY] == one && oth[o + HPY] == one;
            fail = !u;
        }
    }
    const unsigned long long bad = __ballot(fail);
    const int lane = (int)(threadIdx.x & 63);
    if (t < nslots && (lane == 0 || lane == 32)) {
        const unsigned half = (lane == 0) ? (unsigned)bad : (unsigned)(bad >> 32);
        unit[t >> 5] = interior ? (half == 0u ? 1 : 0) : 0;
    }
}

void launch_unit_flags(hipStream_t st, const PfGeom &g, const double *PHIs, uint8_t *unit) {
    const int64_t n = unit_flags_size(g) * 32;   // slots: a multiple of 64 (two chunks per row at least)
    hipLaunchKernelGGL(k_unit_flags<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, PHIs, unit, n);
}
void launch_unit_flags(hipStream_t st, const PfGeom &g, const float *PHI32, uint8_t *unit) {
    const int64_t n = unit_flags_size(g) * 32;
    hipLaunchKernelGGL(k_unit_flags<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, PHI32, unit, n);
}

template <int DIM, int KB>
static int64_t unit_wave_cells_t(const PfGeom &g, const std::vector<uint8_t> &unit, int xw) {
    int64_t cells = 0;
    for (int colour = 0; colour < 2; ++colour) {
        SweepArgs s{};
        s.colour = colour;
        s.xw = xw;
        s.kb = KB;
        const SweepShape sp = sweep_shape(g, KB, xw);
        sweep_set_grid(s, sp.gx, sp.gy, sp.blocks());
        for (int L = 0; L < sp.blocks(); ++L)
            for (int ty = 0; ty < 4; ++ty) {   // one wave: 64 threads of one threadIdx.y
                bool all = true;
                int64_t nw = 0;
                for (int tx = 0; tx < 64; ++tx) {
                    const SweepThread t = sweep_thread_at<DIM, KB>(g, s, L, tx, ty);
                    if (t.j > g.n || t.i > g.m) continue;
                    for (int q = 0; q < KB; ++q) {
                        const int k = t.kbase + 2 * q;
                        if (DIM == 3 && k > g.l) break;
                        all = all && unit[(size_t)unit_index(g, colour, t.j, k, t.h + t.sh - 1)] != 0;
                        ++nw;
                    }
                }
                if (all) cells += nw;
            }
    }
    return cells;
}

int64_t unit_wave_cells(const PfGeom &g, const std::vector<uint8_t> &unit, int xw, int kb) {
    if ((int64_t)unit.size() != unit_flags_size(g)) return 0;
    xw = (xw == 32) ? 32 : 64;
    if (g.dim == 2) return unit_wave_cells_t<2, 1>(g, unit, xw);
    if (kb == 2) return unit_wave_cells_t<3, 2>(g, unit, xw);
    if (kb == 4) return unit_wave_cells_t<3, 4>(g, unit, xw);
    return unit_wave_cells_t<3, 1>(g, unit, xw);
}

// Compares the fast face formula with IEEE division for every face k_faces stored; any
// difference sets *bad (never expected; the solver then keeps true division).
__global__ void k_faces_check(PfGeom g, const double *__restrict__ phi, const double *__restrict__ AW,
                              const double *__restrict__ AS, const double *__restrict__ AB, StepArgs a,
                              int *__restrict__ bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = (g.dim == 3) ? (int)blockIdx.z + 1 : 0;
    if (i > g.m + 1 || j > g.n + 1) return;
    const int64_t o = pf_off(g, i, j, k);
    const int64_t so = pf_soff(g, i, j, k);
    const double pc = phi[o];
    bool ok = __double_as_longlong(face_coef_fast(pc, phi[o - 1], a.dt, a.dx, a.rdx)) == __double_as_longlong(AW[so]);
    ok = ok && __double_as_longlong(face_coef_fast(pc, phi[o - g.PX], a.dt, a.dy, a.rdy)) == __double_as_longlong(AS[so]);
    if (g.dim == 3)
        ok = ok && __double_as_longlong(face_coef_fast(pc, phi[o - (int64_t)g.PX * g.Y], a.dt, a.dz, a.rdz)) ==
                       __double_as_longlong(AB[so]);
    if (!ok) atomicOr(bad, 1);
}

// The same check without stored faces: every face of the case evaluated both ways.
__global__ void k_faces_check_direct(PfGeom g, const double *__restrict__ phi, StepArgs a, int *__restrict__ bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = (g.dim == 3) ? (int)blockIdx.z + 1 : 0;
    if (i > g.m + 1 || j > g.n + 1) return;
    const int64_t o = pf_off(g, i, j, k);
    const double pc = phi[o];
    bool ok = __double_as_longlong(face_coef_fast(pc, phi[o - 1], a.dt, a.dx, a.rdx)) ==
              __double_as_longlong(face_coef(pc, phi[o - 1], a.dt, a.dx));
    ok = ok && __double_as_longlong(face_coef_fast(pc, phi[o - g.PX], a.dt, a.dy, a.rdy)) ==
                   __double_as_longlong(face_coef(pc, phi[o - g.PX], a.dt, a.dy));
    if (g.dim == 3) {
        const double pb = phi[o - (int64_t)g.PX * g.Y];
        ok = ok && __double_as_longlong(face_coef_fast(pc, pb, a.dt, a.dz, a.rdz)) ==
                       __double_as_longlong(face_coef(pc, pb, a.dt, a.dz));
    }
    if (!ok) atomicOr(bad, 1);
}

// Periodic ghost refresh of p for odd extents (3D :463-480 / :520-537 / :588-605,
// 2D :322-327 / :356-361 / :392-397): y ghosts for i in 1..m, k in 1..l; z ghosts for
// i in 1..m, j in 1..n.
__global__ void k_ghost_refresh(PfGeom g, int do_y, int do_z, double *__restrict__ P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int t = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > g.m) return;
    if (do_y && t <= g.l && (g.dim == 3 || t == 1)) {
        const int k = (g.dim == 3) ? t : 0;
        P[pf_soff(g, i, 0, k)] = P[pf_soff(g, i, g.n, k)];
        P[pf_soff(g, i, g.n + 1, k)] = P[pf_soff(g, i, 1, k)];
    }
    if (do_z && g.dim == 3 && t <= g.n) {
        P[pf_soff(g, i, t, 0)] = P[pf_soff(g, i, t, g.l)];
        P[pf_soff(g, i, t, g.l + 1)] = P[pf_soff(g, i, t, 1)];
    }
}

// ------------------------------------------------------------------------------------
// Projection (3D main :115-123, 2D main :108-113).
// ------------------------------------------------------------------------------------

__global__ void k_project(PfGeom g, StepArgs a, const double *__restrict__ P, double *__restrict__ u,
                          double *__restrict__ v, double *__restrict__ w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = (g.dim == 3) ? (int)blockIdx.z + 1 : 0;
    if (i > g.m || j > g.n) return;
    project_cell(g, a, P, u, v, w, i, j, k);
}

// ------------------------------------------------------------------------------------
// boundary for the uniform-inflow solvers (3D :691-748, 2D :477-501).  The reference
// applies inlet and outlet rows for interior (j,k), then copies y-periodic ghosts over the
// full i and k ranges, then z-periodic over full i and j ranges.  The final value of any
// boundary cell is therefore V(i, wrap(j), wrap(k)), V being the state after the x rows;
// each thread evaluates that directly, so one launch has no intra-kernel dependency.
// The launch covers the x planes i in {0,1,m+1}, the y planes and the z planes.
// ------------------------------------------------------------------------------------
// Final value of one boundary cell of the uniform-inflow solvers (i in {0,1,m+1}, or j, k
// a ghost index).


__global__ void k_boundary_uniform(PfGeom g, StepArgs a, const double *__restrict__ po, double *__restrict__ u,
                                   double *__restrict__ v, double *__restrict__ w, double *__restrict__ P) {
    const int m = g.m, n = g.n, l = g.l;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t Y = g.Y, Z = g.Z, X = m + 2;
    const int64_t nx = 3 * Y * Z, ny = 2 * X * Z, nz = (g.dim == 3) ? 2 * X * Y : 0;
    int i, j, k;
    if (t < nx) {
        const int sel = (int)(t / (Y * Z));
        const int64_t r = t % (Y * Z);
        i = (sel == 0) ? 0 : (sel == 1 ? 1 : m + 1);
        j = (int)(r % Y);
        k = (int)(r / Y);
    } else if (t < nx + ny) {
        const int64_t r = t - nx;
        const int sel = (int)(r / (X * Z));
        const int64_t q = r % (X * Z);
        j = (sel == 0) ? 0 : n + 1;
        i = (int)(q % X);
        k = (int)(q / X);
    } else if (t < nx + ny + nz) {
        const int64_t r = t - nx - ny;
        const int sel = (int)(r / (X * Y));
        const int64_t q = r % (X * Y);
        k = (sel == 0) ? 0 : l + 1;
        i = (int)(q % X);
        j = (int)(q / X);
    } else {
        return;
    }
    if (g.dim == 2) k = 0;
    boundary_uniform_cell(g, a, po, u, v, w, P, i, j, k);
}

// boundary of the air-conditioning solver, one face per launch in the reference's order
// top, bottom, west, east, north, south (ibm_3d_air_condition_omp_cpu.f90:890-1167);
// each launch covers the face's full index range, ghost rows included.

__global__ void k_boundary_air(PfGeom g, StepArgs a, int face, const double *__restrict__ po, double *__restrict__ u,
                               double *__restrict__ v, double *__restrict__ w, double *__restrict__ P) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y * blockDim.y + threadIdx.y;
    boundary_air_cell(g, a, face, po, u, v, w, P, r, q);
}

// ------------------------------------------------------------------------------------
// output_divergent_2d/3d (output.f90:217-224, 929-940) into a dense interior array.
// ------------------------------------------------------------------------------------

__global__ void k_divergence_out(PfGeom g, StepArgs a, const double *__restrict__ po, const double *__restrict__ u,
                                 const double *__restrict__ v, const double *__restrict__ w, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = (g.dim == 3) ? (int)blockIdx.z + 1 : 0;
    if (i > g.m || j > g.n) return;
    divout_cell(g, a, po, u, v, w, out, i, j, k);
}

// ------------------------------------------------------------------------------------
// Force on the immersed body (output_force_log_2d, output.f90:264-284; output_force_log_3d,
// :1118-1145): per interior cell the pressure term -dV*p*2*phi*(1-phi)/(thickness*d)*n and
// the wall-force term dV*32*rho*nu*(phi*(1-phi)/(thickness*d))**2*u, each in the Fortran
// evaluation order.  The sum runs in a fixed tree (each thread strides over cells, then a
// block tree, then k_force_final over the blocks), so it is deterministic but not the
// reference's serial order: results agree to rounding, not bit for bit.
// partial[b*6 + q]: q = 0..2 pressure x, y, z; 3..5 viscous x, y, z.
// ------------------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(256) void k_force(PfGeom g, StepArgs a, const double *__restrict__ po,
                                               const double *__restrict__ u, const double *__restrict__ v,
                                               const double *__restrict__ w, const double *__restrict__ P,
                                               double *__restrict__ partial) {
    const int64_t cells = (int64_t)g.m * g.n * (DIM == 3 ? g.l : 1);
    const double dx = a.dx, dy = a.dy, dz = a.dz, th = a.thickness;
    const double alpha = 32.0, rho = a.density, nu = a.xnue;
    double f[6] = {0., 0., 0., 0., 0., 0.};
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(c % g.m) + 1;
        const int64_t r = c / g.m;
        const int j = (int)(r % g.n) + 1;
        const int k = (DIM == 3) ? (int)(r / g.n) + 1 : 0;
        const int64_t o = pf_off(g, i, j, k);
        const int64_t sy = g.PX, sz = (int64_t)g.PX * g.Y;
        const double gx = (po[o + 1] - po[o - 1]) * 0.5;
        const double gy = (po[o + sy] - po[o - sy]) * 0.5;
        const double gz = (DIM == 3) ? (po[o + sz] - po[o - sz]) * 0.5 : 0.;
        const double na = (DIM == 3) ? sqrt(((gx * gx) + (gy * gy)) + (gz * gz)) : sqrt((gx * gx) + (gy * gy));
        const double den = fmax_ref(na, PF_SMALL);
        const double ph = po[o], q = 1.0 - ph;
        const double pc = P[pf_soff(g, i, j, k)];
        const double dv = (DIM == 3) ? ((dx * dy) * dz) : (dx * dy);
        const double base = (((dv * pc) * 2) * ph) * q;
        const double cv = ((dv * alpha) * rho) * nu;
        const double tx = (ph * q) / (th * dx), ty = (ph * q) / (th * dy);
        f[0] += -((base / (th * dx)) * (gx / den));
        f[1] += -((base / (th * dy)) * (gy / den));
        f[3] += (cv * (tx * tx)) * u[o];
        f[4] += (cv * (ty * ty)) * v[o];
        if (DIM == 3) {
            const double tz = (ph * q) / (th * dz);
            f[2] += -((base / (th * dz)) * (gz / den));
            f[5] += (cv * (tz * tz)) * w[o];
        }
    }
    __shared__ double red[6][256];
    for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = f[q];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int q = 0; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[(int64_t)blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_force_final(const double *__restrict__ partial, int nblocks,
                                                     double *__restrict__ out) {
    double f[6] = {0., 0., 0., 0., 0., 0.};
    for (int b = (int)threadIdx.x; b < nblocks; b += 256)
        for (int q = 0; q < 6; ++q) f[q] += partial[(int64_t)b * 6 + q];
    __shared__ double red[6][256];
    for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = f[q];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int q = 0; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = red[threadIdx.x][0];
}

// One red/black SOR iteration in one launch (pf_fused.h): reads A, writes B.  One block per
// tile, XCD-contiguous tile order; the per-block running max of |dp| goes to err_slot.
template <int DIM, int KIND, bool FAST>
__global__ __launch_bounds__(256) void k_sor_fused(PfGeom g, StepArgs a, FusedGeom fg,
                                                   FusedArgs<typename FusedT<KIND>::T> p,
                                                   unsigned long long *err_slot, int track) {
    const int nb = (int)gridDim.x;
    const int L = xcd_logical_block((int)blockIdx.x, nb);
    int x0, j0, k0, k1;
    fused_tile_coords(g, fg, L, x0, j0, k0, k1);
    const double e = fused_tile<DIM, KIND, FAST>(g, a, fg, p, x0, j0, k0, k1);
    if (track) fused_block_max(e, err_slot + L);
}

// Clear the per-step diagnostics with a kernel (graph-capture safe, unlike a memset node).
__global__ void k_clear_u64(unsigned long long *p, int count) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) p[t] = 0ull;
}

// ------------------------------------------------------------------------------------
// Launch wrappers
// ------------------------------------------------------------------------------------
static inline dim3 grid2(int nx, int ny, int nz, dim3 b) {
    return dim3((unsigned)((nx + (int)b.x - 1) / (int)b.x), (unsigned)((ny + (int)b.y - 1) / (int)b.y), (unsigned)nz);
}

void launch_faces(hipStream_t st, const PfGeom &g, const double *phi, double *AW, double *AS, double *AB,
                  double dt, double dx, double dy, double dz) {
    dim3 b(64, 4, 1);
    hipLaunchKernelGGL(k_faces, grid2(g.m + 1, g.n + 1, g.dim == 3 ? g.l + 1 : 1, b), b, 0, st, g, phi, AW, AS, AB,
                       dt, dx, dy, dz);
}

void launch_div(hipStream_t st, const PfGeom &g, const StepArgs &a, const double *uo, const double *vo,
                const double *wo, double *div) {
    dim3 b(64, 4, 1);
    hipLaunchKernelGGL(k_div, grid2(g.m, g.n + 2, g.dim == 3 ? g.l + 2 : 1, b), b, 0, st, g, a, uo, vo, wo, div);
}

void launch_predict(hipStream_t st, const PfGeom &g, const StepArgs &a, const double *uo, const double *vo,
                    const double *wo, const double *div, const double *phi, double *u, double *v, double *w) {
    dim3 b(64, 4, 1);
    if (g.dim == 3)
        hipLaunchKernelGGL(k_predict3, grid2(g.m, g.n + 2, g.l + 2, b), b, 0, st, g, a, uo, vo, wo, div, phi, u, v, w);
    else
        hipLaunchKernelGGL(k_predict2, grid2(g.m, g.n + 2, 1, b), b, 0, st, g, a, uo, vo, div, phi, u, v);
}

void launch_rhs(hipStream_t st, const PfGeom &g, const StepArgs &a, const double *u, const double *v,
                const double *w, const double *phi, const double *P, double *BB) {
    dim3 b(64, 4, 1);
    hipLaunchKernelGGL(k_rhs, grid2(g.m, g.n, g.dim == 3 ? g.l : 1, b), b, 0, st, g, a, u, v, w, phi, P, BB);
}

void launch_sor(hipStream_t st, const PfGeom &g, const StepArgs &a, const SweepArgs &s, double *P,
                const double *AW, const double *AS, const double *AB, const double *BB, const double *phi,
                unsigned long long *err_slot) {
    const int c = s.colour;
    const int64_t own = (int64_t)c * g.HS, oth = (int64_t)(1 - c) * g.HS;
    dim3 b(64, 4, 1);
    dim3 gr = grid2((g.m + 1) / 2, g.n, g.dim == 3 ? g.l : 1, b);
    if (g.dim == 3)
        hipLaunchKernelGGL(k_sor<3>, gr, b, 0, st, g, a, s, P + oth, P + own, AW + own, AW + oth, AS + own, AS + oth,
                           AB + own, AB + oth, BB + own, phi, err_slot);
    else
        hipLaunchKernelGGL(k_sor<2>, gr, b, 0, st, g, a, s, P + oth, P + own, AW + own, AW + oth, AS + own, AS + oth,
                           AW + own, AW + oth, BB + own, phi, err_slot);
}

void launch_sor_phi(hipStream_t st, const PfGeom &g, const StepArgs &a, const SweepArgs &s, double *P,
                    const double *PHIs, const double *BB, const double *phi, unsigned long long *err_slot,
                    bool fast, const float *PHC, const double *tab, const uint8_t *unit) {
    const int c = s.colour;
    const int64_t own = (int64_t)c * g.HS, oth = (int64_t)(1 - c) * g.HS;
    const int KB = (g.dim == 3) ? s.kb : 1;
    SweepArgs s1 = s;
    s1.xw = (s.xw == 32) ? 32 : 64;   // an unset width means one row per wave
    const SweepShape sp = sweep_shape(g, KB, s1.xw);
    sweep_set_grid(s1, sp.gx, sp.gy, sp.blocks());
    dim3 b(64, 4, 1), gr((unsigned)sp.blocks());
#define PF_SOR_PHI_U(D, F, K, U) hipLaunchKernelGGL((k_sor_phi<D, F, K, double, U>), gr, b, 0, st, g, a, s1, P + oth, \
                                                   P + own, PHIs + own, PHIs + oth, BB + own, phi, err_slot, \
                                                   (const double *)nullptr, unit)
#define PF_SOR_PHC_U(D, F, K, U) hipLaunchKernelGGL((k_sor_phi<D, F, K, float, U>), gr, b, 0, st, g, a, s1, P + oth, \
                                                   P + own, PHC + own, PHC + oth, BB + own, phi, err_slot, tab, unit)
#define PF_SOR_PHI(D, F, K) do { if (unit) PF_SOR_PHI_U(D, F, K, true); else PF_SOR_PHI_U(D, F, K, false); } while (0)
#define PF_SOR_PHC(D, F, K) do { if (unit) PF_SOR_PHC_U(D, F, K, true); else PF_SOR_PHC_U(D, F, K, false); } while (0)
    if (PHC && tab && KB == 1) {
        if (g.dim == 3) { if (fast) PF_SOR_PHC(3, true, 1); else PF_SOR_PHC(3, false, 1); }
        else { if (fast) PF_SOR_PHC(2, true, 1); else PF_SOR_PHC(2, false, 1); }
    } else if (g.dim == 3) {
        if (fast) {
            if (KB == 1) PF_SOR_PHI(3, true, 1); else if (KB == 2) PF_SOR_PHI(3, true, 2); else PF_SOR_PHI(3, true, 4);
        } else {
            if (KB == 1) PF_SOR_PHI(3, false, 1); else if (KB == 2) PF_SOR_PHI(3, false, 2); else PF_SOR_PHI(3, false, 4);
        }
    } else {
        if (fast) PF_SOR_PHI(2, true, 1); else PF_SOR_PHI(2, false, 1);
    }
#undef PF_SOR_PHI
#undef PF_SOR_PHC
#undef PF_SOR_PHI_U
#undef PF_SOR_PHC_U
}

bool code_porosity(const std::vector<double> &phis, std::vector<float> &code, std::vector<double> &tab,
                   size_t max_tab) {
    const size_t cap = std::min(max_tab, (size_t)1 << PHI_CODE_INDEX_BITS);
    code.resize(phis.size());
    tab.clear();
    std::unordered_map<uint64_t, uint32_t> index;
    for (size_t q = 0; q < phis.size(); ++q) {
        const double v = phis[q];
        const float f = (float)v;
        if ((double)f == v) {   // exact in FP32 (a NaN never compares equal: it goes to the table)
            code[q] = f;
            continue;
        }
        uint64_t bits;
        std::memcpy(&bits, &v, sizeof bits);
        auto it = index.find(bits);
        uint32_t id;
        if (it == index.end()) {
            if (tab.size() >= cap) return false;
            id = (uint32_t)tab.size();
            index.emplace(bits, id);
            tab.push_back(v);
        } else {
            id = it->second;
        }
        const uint32_t cbits = 0x7FC00000u | id;   // quiet NaN, index in the low bits
        std::memcpy(&code[q], &cbits, sizeof cbits);
    }
    return true;
}

void launch_faces_check_direct(hipStream_t st, const PfGeom &g, const double *phi, const StepArgs &a, int *bad) {
    dim3 b(64, 4, 1);
    hipLaunchKernelGGL(k_faces_check_direct, grid2(g.m + 1, g.n + 1, g.dim == 3 ? g.l + 1 : 1, b), b, 0, st, g, phi, a,
                       bad);
}

void launch_faces_check(hipStream_t st, const PfGeom &g, const double *phi, const double *AW, const double *AS,
                        const double *AB, const StepArgs &a, int *bad) {
    dim3 b(64, 4, 1);
    hipLaunchKernelGGL(k_faces_check, grid2(g.m + 1, g.n + 1, g.dim == 3 ? g.l + 1 : 1, b), b, 0, st, g, phi, AW, AS,
                       AB, a, bad);
}

void launch_split(hipStream_t st, const PfGeom &g, const double *src, double *dst) {
    dim3 b(64, 4, 1);
    hipLaunchKernelGGL(k_split, grid2(g.m + 2, g.n + 2, g.Z, b), b, 0, st, g, src, dst);
}

void launch_ghost_refresh(hipStream_t st, const PfGeom &g, int do_y, int do_z, double *P) {
    dim3 b(64, 4, 1);
    const int t = (g.dim == 3) ? (g.l > g.n ? g.l : g.n) : 1;
    hipLaunchKernelGGL(k_ghost_refresh, grid2(g.m, t, 1, b), b, 0, st, g, do_y, do_z, P);
}

void launch_project(hipStream_t st, const PfGeom &g, const StepArgs &a, const double *P, double *u, double *v,
                    double *w) {
    dim3 b(64, 4, 1);
    hipLaunchKernelGGL(k_project, grid2(g.m, g.n, g.dim == 3 ? g.l : 1, b), b, 0, st, g, a, P, u, v, w);
}

void launch_boundary(hipStream_t st, const PfGeom &g, const StepArgs &a, const double *phi, double *u, double *v,
                     double *w, double *P) {
    if (a.bc == PF_BC_UNIFORM) {
        const int64_t X = g.m + 2;
        const int64_t total = 3 * (int64_t)g.Y * g.Z + 2 * X * g.Z + (g.dim == 3 ? 2 * X * g.Y : 0);
        const int b = 256;
        hipLaunchKernelGGL(k_boundary_uniform, dim3((unsigned)((total + b - 1) / b)), dim3(b), 0, st, g, a, phi, u, v,
                           w, P);
        return;
    }
    dim3 b(64, 4, 1);
    const int order[6] = {0, 1, 2, 3, 4, 5};  // top, bottom, west, east, north, south
    for (int f : order) {
        dim3 gr;
        if (f <= 1) gr = grid2(g.m + 2, g.n + 2, 1, b);
        else if (f <= 3) gr = grid2(g.n + 2, g.l + 2, 1, b);
        else gr = grid2(g.m + 2, g.l + 2, 1, b);
        hipLaunchKernelGGL(k_boundary_air, gr, b, 0, st, g, a, f, phi, u, v, w, P);
    }
}

void launch_divergence_out(hipStream_t st, const PfGeom &g, const StepArgs &a, const double *phi, const double *u,
                           const double *v, const double *w, double *out) {
    dim3 b(64, 4, 1);
    hipLaunchKernelGGL(k_divergence_out, grid2(g.m, g.n, g.dim == 3 ? g.l : 1, b), b, 0, st, g, a, phi, u, v, w, out);
}

int force_blocks(const PfGeom &g) {
    const int64_t cells = (int64_t)g.m * g.n * (g.dim == 3 ? g.l : 1);
    const int64_t b = (cells + 255) / 256;
    return (int)(b < 1024 ? b : 1024);
}

void launch_force(hipStream_t st, const PfGeom &g, const StepArgs &a, const double *phi, const double *u,
                  const double *v, const double *w, const double *P, double *partial, double *out) {
    const int nb = force_blocks(g);
    if (g.dim == 3)
        hipLaunchKernelGGL(k_force<3>, dim3((unsigned)nb), dim3(256), 0, st, g, a, phi, u, v, w, P, partial);
    else
        hipLaunchKernelGGL(k_force<2>, dim3((unsigned)nb), dim3(256), 0, st, g, a, phi, u, v, w, P, partial);
    hipLaunchKernelGGL(k_force_final, dim3(1), dim3(256), 0, st, partial, nb, out);
}

int sweep_wave_width(const PfGeom &g) { return sweep_xw(g); }

int sor_blocks(const PfGeom &g) {
    // slots for every kernel that reduces into them: one per wave (four per 256-thread
    // block) of the stored-face sweep, the KB = 1, 2, 4 sweeps and the one-cell-per-thread
    // residual, one per block of the fused sweep
    dim3 b(64, 4, 1);
    dim3 gr = grid2((g.m + 1) / 2, g.n, g.dim == 3 ? g.l : 1, b);
    int n = (int)(gr.x * gr.y * gr.z);
    for (int kb : {1, 2, 4})
        for (int xw : {32, 64}) {
            const int v = sweep_shape(g, kb, xw).blocks();
            n = v > n ? v : n;
        }
    // the temporally blocked 2D sweep: one row per block, TB2_THREADS / 64 waves per block
    if (g.dim == 2) n = g.n * (TB2_THREADS / 256) > n ? g.n * (TB2_THREADS / 256) : n;
    n *= 4;
    const FusedGeom f = fused_geom(g, 1);
    const int nf = f.gx * f.gy * f.gz;
    return nf > n ? nf : n;
}

bool sor_fused_ok(const PfGeom &g, int bc) { return fused_ok(g, bc == PF_BC_UNIFORM); }

void launch_sor_fused(hipStream_t st, const PfGeom &g, const StepArgs &a, const double *A, double *B,
                      const double *F, const double *R, const double *phi, unsigned long long *slot, int track,
                      bool fast) {
    const FusedGeom f = fused_geom(g, a.bc == PF_BC_UNIFORM);
    const dim3 gr((unsigned)(f.gx * f.gy * f.gz));
    FusedArgs<double> p{A, B, F, R, phi, 0.f, 0.f, 0.f};
    if (g.dim == 3) {
        if (fast) hipLaunchKernelGGL((k_sor_fused<3, 0, true>), gr, dim3(256), 0, st, g, a, f, p, slot, track);
        else hipLaunchKernelGGL((k_sor_fused<3, 0, false>), gr, dim3(256), 0, st, g, a, f, p, slot, track);
    } else {
        if (fast) hipLaunchKernelGGL((k_sor_fused<2, 0, true>), gr, dim3(256), 0, st, g, a, f, p, slot, track);
        else hipLaunchKernelGGL((k_sor_fused<2, 0, false>), gr, dim3(256), 0, st, g, a, f, p, slot, track);
    }
}

void launch_sor32_fused(hipStream_t st, const PfGeom &g, const StepArgs &a, const float *A, float *B, const float *F,
                        const float *R, const double *phi, unsigned long long *slot, int track) {
    const FusedGeom f = fused_geom(g, a.bc == PF_BC_UNIFORM);
    const dim3 gr((unsigned)(f.gx * f.gy * f.gz));
    FusedArgs<float> p{A, B, F, R, phi, (float)(a.dt / (a.dx * a.dx)), (float)(a.dt / (a.dy * a.dy)),
                       (float)(a.dt / (a.dz * a.dz))};
    if (g.dim == 3) hipLaunchKernelGGL((k_sor_fused<3, 1, true>), gr, dim3(256), 0, st, g, a, f, p, slot, track);
    else hipLaunchKernelGGL((k_sor_fused<2, 1, true>), gr, dim3(256), 0, st, g, a, f, p, slot, track);
}

void launch_reduce_err(hipStream_t st, unsigned long long *slots, int nslots, unsigned long long *step_err) {
    hipLaunchKernelGGL(k_reduce_err, dim3(1), dim3(1024), 0, st, slots, nslots, step_err);
}

void launch_clear_u64(hipStream_t st, unsigned long long *p, int count) {
    hipLaunchKernelGGL(k_clear_u64, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, p, count);
}

// Plain device copy (bandwidth calibration: a kernel whose HBM bytes are known exactly).
// One 16-byte element per thread over a grid that covers the array: on MI355X this form
// copies 1-4 GB at 6.2-6.5 TB/s, where a grid-stride loop over 8192 blocks reached
// 4.5-4.9 TB/s.
__global__ void k_copy_probe(const double2 *__restrict__ src, double2 *__restrict__ dst, int64_t n2) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n2) dst[t] = src[t];
}

void launch_copy_probe(hipStream_t st, const double *src, double *dst, int64_t n) {
    const int64_t n2 = n / 2;
    const int64_t blocks = (n2 + 255) / 256;
    hipLaunchKernelGGL(k_copy_probe, dim3((unsigned)blocks), dim3(256), 0, st, (const double2 *)src, (double2 *)dst,
                       n2);
}

}  // namespace pf
