// pf_mixed.hip -- mixed-precision pressure solve: FP64 defect correction around FP32
// red/black SOR sweeps.
//
// One cycle:   r = bb - A p            (FP64, the reference's operator, faces rebuilt exactly)
//              A e = r, e0 = 0         (K red/black SOR iterations in FP32)
//              p = p + e               (FP64)
// SOR is a linear stationary iteration, so in exact arithmetic K iterations on A e = r from
// e = 0 move p exactly as K iterations on A p = bb from p do: with cycles of K iterations
// summing to iter_max the mixed path follows the reference's fixed-iteration semantics and
// differs from it only by the FP32 rounding of the corrections.  With a residual tolerance
// it instead stops on the FP64 Jacobi-scaled residual max|r / ap|.
// Storage per cycle: e, r and porosity in FP32 (half the bytes of an FP64 sweep); p, bb
// and the residual/update passes stay FP64.
#include <hip/hip_runtime.h>

#include "pf_cells.h"
#include "pf_device.h"
#include "pf_kernels.h"

namespace pf {

// ------------------------------------------------------------------------------------
// FP64 residual of the cells of one colour (launched for both colours): r = bb - A p
// stored as FP32 R, e cleared at the cell and at the periodic ghost slot it feeds, and the
// wave max of |r / ap| into the wave's slot.
// ------------------------------------------------------------------------------------

// UNIT: `unit` holds the unit-porosity chunk flags of the FP64 porosity; a wave (one row of
// 64 slots) whose cells all lie in flagged chunks takes the unit faces.
template <int DIM, bool FAST, typename TR, bool UNIT>
__global__ __launch_bounds__(256) void k_residual(PfGeom g, StepArgs a, SweepArgs s, const double *__restrict__ P,
                                                  const double *__restrict__ PHIs, const double *__restrict__ BB,
                                                  const double *__restrict__ phi, TR *__restrict__ R,
                                                  TR *__restrict__ E, unsigned long long *slots,
                                                  const uint8_t *__restrict__ unit) {
    const CellIdx q = sweep_cell<DIM>(g, s);
    bool all_unit = false;
    if constexpr (UNIT)
        all_unit = __all(!q.active || gmem(unit)[unit_index(g, s.colour, q.j, q.k, q.h + q.sh - 1)] != 0) != 0;
    double res = 0.;
    if (q.active)
        res = residual_cell<DIM, FAST, TR>(g, a, s.colour, q.i, q.j, q.k, q.sh, q.o, P, PHIs, BB, phi, R, E, all_unit);
    if (s.track_err) wave_max_to_slot(res, slots);
}

// ------------------------------------------------------------------------------------
// FP32 red/black SOR half-sweep on A e = r (same structure as the FP64 sweep).
// ------------------------------------------------------------------------------------

// UNIT: `unit` holds the unit-porosity chunk flags of the FP32 porosity; a wave whose lanes
// all lie in flagged chunks skips its porosity loads and uses 1 for them (the same floats).
template <int DIM, int KB, bool UNIT>
__global__ __launch_bounds__(256) void k_sor32(PfGeom g, StepArgs a, SweepArgs s, const float *__restrict__ Eoth,
                                               float *__restrict__ Eown, const float *__restrict__ F32o,
                                               const float *__restrict__ F32x, const float *__restrict__ Ro,
                                               const double *__restrict__ phi, float cdx, float cdy, float cdz,
                                               unsigned long long *err_slot, const uint8_t *__restrict__ unit) {
    const SweepThread t = sweep_thread<DIM, KB>(g, s);
    const bool active = t.j <= g.n && t.i <= g.m;
    double err = 0.;
    SweepIn<float> x[KB];
    int64_t o[KB];
    int ok = 1;
    if (active) {
        // the unit flags first (the vote waits for these only), then the correction and
        // residual operands
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
            sweep_gather_unit<DIM>(g, t.sh, o[q], Eoth, Eown, Ro, x[q]);
        }
#pragma unroll
        for (int q = 0; q < KB; ++q)
            if (!(DIM == 3 && t.kbase + 2 * q > g.l)) ok &= (int)fl[q];
    }
    const bool all_unit = UNIT && __all(ok) != 0;   // the wave's branch
    if (active && !all_unit) {
#pragma unroll
        for (int q = 0; q < KB; ++q) sweep_gather_phi32<DIM>(g, t.sh, o[q], F32o, F32x, x[q]);
    }
    if (active) {
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            const int k = t.kbase + 2 * q;
            if (DIM == 3 && k > g.l) break;
            const SweepIn<float> &c = x[q];
            const float en = all_unit
                                 ? sor32_value<DIM>(g, a, t.i, t.j, k, c.c, c.w, c.e, c.s, c.n, c.b, c.t, 1.f, 1.f, 1.f,
                                                    1.f, 1.f, 1.f, 1.f, c.rhs, phi, cdx, cdy, cdz)
                                 : sor32_value<DIM>(g, a, t.i, t.j, k, c.c, c.w, c.e, c.s, c.n, c.b, c.t, c.fc, c.fw,
                                                    c.fe, c.fs, c.fn, c.fb, c.ft, c.rhs, phi, cdx, cdy, cdz);
            sweep_store<DIM>(g, s, t.j, k, o[q], en, Eown);
            const double e = fabs((double)en - (double)c.c);
            err = max_keep_nan(err, e);
        }
    }
    if (s.track_err) wave_max_to_slot(err, err_slot);
}

// p += e over every slot of both colours: interior cells and the periodic ghost slots
// (which hold e of their source cell, so the ghost stays equal to its source).
// One element per thread (the launch covers n).
template <typename TE>
__global__ void k_update(const TE *__restrict__ E, double *__restrict__ P, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) P[t] = P[t] + (double)E[t];
}

// p += e as k_update, with the max |e| of the update (NaN kept) folded into the slots of
// the block's waves: the step's diagnostic when a solve reports no sweep error.
template <typename TE>
__global__ __launch_bounds__(256) void k_update_track(const TE *__restrict__ E, double *__restrict__ P, int64_t n,
                                                      unsigned long long *slots) {
    double mx = 0.;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const double e = (double)E[t];
        P[t] = P[t] + e;
        mx = max_keep_nan(mx, fabs(e));
    }
    wave_max_to_slot(mx, slots);
}

__global__ void k_ghost_refresh32(PfGeom g, int do_y, int do_z, float *__restrict__ E) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int t = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > g.m) return;
    if (do_y && t <= g.l && (g.dim == 3 || t == 1)) {
        const int k = (g.dim == 3) ? t : 0;
        E[pf_soff(g, i, 0, k)] = E[pf_soff(g, i, g.n, k)];
        E[pf_soff(g, i, g.n + 1, k)] = E[pf_soff(g, i, 1, k)];
    }
    if (do_z && g.dim == 3 && t <= g.n) {
        E[pf_soff(g, i, t, 0)] = E[pf_soff(g, i, t, g.l)];
        E[pf_soff(g, i, t, g.l + 1)] = E[pf_soff(g, i, t, 1)];
    }
}

__global__ void k_split32(PfGeom g, const double *__restrict__ src, float *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > g.m + 1 || j > g.n + 1) return;
    dst[pf_soff(g, i, j, k)] = (float)src[pf_off(g, i, j, k)];
}

// ------------------------------------------------------------------------------------
static inline dim3 sweep_grid(const PfGeom &g, SweepArgs &s, dim3 b) {
    const unsigned gx = (unsigned)(((g.m + 1) / 2 + (int)b.x - 1) / (int)b.x);
    const unsigned gy = (unsigned)((g.n + (int)b.y - 1) / (int)b.y);
    const unsigned gz = (unsigned)(g.dim == 3 ? g.l : 1);
    sweep_set_grid(s, (int)gx, (int)gy, (int)(gx * gy * gz));
    return dim3(gx * gy * gz);
}

template <typename TR>
static void residual_any(hipStream_t st, const PfGeom &g, const StepArgs &a, SweepArgs s, const double *P,
                         const double *PHIs, const double *BB, const double *phi, TR *R, TR *E,
                         unsigned long long *slots, bool fast, const uint8_t *unit) {
    dim3 b(64, 4, 1);
    dim3 gr = sweep_grid(g, s, b);
#define PF_RES_U(D, F, U) hipLaunchKernelGGL((k_residual<D, F, TR, U>), gr, b, 0, st, g, a, s, P, PHIs, BB, phi, R, E, \
                                             slots, unit)
#define PF_RES(D, F) do { if (unit) PF_RES_U(D, F, true); else PF_RES_U(D, F, false); } while (0)
    if (g.dim == 3) { if (fast) PF_RES(3, true); else PF_RES(3, false); }
    else { if (fast) PF_RES(2, true); else PF_RES(2, false); }
#undef PF_RES
#undef PF_RES_U
}
void launch_residual(hipStream_t st, const PfGeom &g, const StepArgs &a, SweepArgs s, const double *P,
                     const double *PHIs, const double *BB, const double *phi, float *R, float *E,
                     unsigned long long *slots, bool fast, const uint8_t *unit) {
    residual_any<float>(st, g, a, s, P, PHIs, BB, phi, R, E, slots, fast, unit);
}
void launch_residual(hipStream_t st, const PfGeom &g, const StepArgs &a, SweepArgs s, const double *P,
                     const double *PHIs, const double *BB, const double *phi, double *R, double *E,
                     unsigned long long *slots, bool fast, const uint8_t *unit) {
    residual_any<double>(st, g, a, s, P, PHIs, BB, phi, R, E, slots, fast, unit);
}

void launch_sor32(hipStream_t st, const PfGeom &g, const StepArgs &a, SweepArgs s, float *E, const float *F32,
                  const float *R, const double *phi, unsigned long long *err_slot, const uint8_t *unit) {
    const int KB = (g.dim == 3) ? s.kb : 1;
    s.xw = (s.xw == 32) ? 32 : 64;   // an unset width means one row per wave
    const SweepShape sp = sweep_shape(g, KB, s.xw);
    sweep_set_grid(s, sp.gx, sp.gy, sp.blocks());
    dim3 b(64, 4, 1), gr((unsigned)sp.blocks());
    const int64_t own = (int64_t)s.colour * g.HS, oth = (int64_t)(1 - s.colour) * g.HS;
    const float cdx = (float)(a.dt / (a.dx * a.dx)), cdy = (float)(a.dt / (a.dy * a.dy)),
                cdz = (float)(a.dt / (a.dz * a.dz));
#define PF_S32_U(D, K, U) hipLaunchKernelGGL((k_sor32<D, K, U>), gr, b, 0, st, g, a, s, E + oth, E + own, F32 + own, \
                                             F32 + oth, R + own, phi, cdx, cdy, cdz, err_slot, unit)
#define PF_S32(D, K) do { if (unit) PF_S32_U(D, K, true); else PF_S32_U(D, K, false); } while (0)
    if (g.dim == 3) {
        if (KB == 1) PF_S32(3, 1); else if (KB == 2) PF_S32(3, 2); else PF_S32(3, 4);
    } else {
        PF_S32(2, 1);
    }
#undef PF_S32
#undef PF_S32_U
}

void launch_update(hipStream_t st, const PfGeom &g, const float *E, double *P) {
    hipLaunchKernelGGL(k_update<float>, dim3((unsigned)((2 * g.HS + 255) / 256)), dim3(256), 0, st, E, P, 2 * g.HS);
}
void launch_update(hipStream_t st, const PfGeom &g, const double *E, double *P) {
    hipLaunchKernelGGL(k_update<double>, dim3((unsigned)((2 * g.HS + 255) / 256)), dim3(256), 0, st, E, P, 2 * g.HS);
}
// nslots: the slots available (four per block are used)
template <typename TE>
static void update_track_any(hipStream_t st, const PfGeom &g, const TE *E, double *P, unsigned long long *slots,
                             int nslots) {
    int nb = nslots / 4;
    nb = nb < 1 ? 1 : (nb > 2048 ? 2048 : nb);
    hipLaunchKernelGGL(k_update_track<TE>, dim3((unsigned)nb), dim3(256), 0, st, E, P, 2 * g.HS, slots);
}
void launch_update_track(hipStream_t st, const PfGeom &g, const float *E, double *P, unsigned long long *slots,
                         int nslots) {
    update_track_any<float>(st, g, E, P, slots, nslots);
}
void launch_update_track(hipStream_t st, const PfGeom &g, const double *E, double *P, unsigned long long *slots,
                         int nslots) {
    update_track_any<double>(st, g, E, P, slots, nslots);
}

void launch_ghost_refresh32(hipStream_t st, const PfGeom &g, int do_y, int do_z, float *E) {
    dim3 b(64, 4, 1);
    const int t = (g.dim == 3) ? (g.l > g.n ? g.l : g.n) : 1;
    hipLaunchKernelGGL(k_ghost_refresh32, dim3((unsigned)((g.m + 63) / 64), (unsigned)((t + 3) / 4)), b, 0, st, g,
                       do_y, do_z, E);
}

void launch_split32(hipStream_t st, const PfGeom &g, const double *src, float *dst) {
    dim3 b(64, 4, 1);
    hipLaunchKernelGGL(k_split32, dim3((unsigned)((g.m + 2 + 63) / 64), (unsigned)((g.n + 2 + 3) / 4), (unsigned)g.Z), b,
                       0, st, g, src, dst);
}

}  // namespace pf
