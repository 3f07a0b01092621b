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

#include "pf_device.h"
#include "pf_kernels.h"

namespace pf {

// FP32 face coefficient: dt/d^2 * max(small, (pa+pb)/2) (the FP64 path's expression,
// regrouped; exact order does not matter for a correction that the FP64 residual checks).
__device__ __forceinline__ float face32(float pa, float pb, float cdt) {
    return fmaxf(1.e-6f, (pa + pb) * 0.5f) * cdt;
}

__device__ __forceinline__ void block_max_to_slot(double v, unsigned long long *slots) {
    unsigned long long e = __double_as_longlong(v);
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o2 = __shfl_xor(e, off, 64);
        e = (o2 > e) ? o2 : e;
    }
    __shared__ unsigned long long red[4];
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    if ((lane & 63) == 0) red[lane >> 6] = e;
    __syncthreads();
    if (lane == 0) {
        unsigned long long b = red[0];
        const int nw = (blockDim.x * blockDim.y + 63) >> 6;
        for (int q = 1; q < nw; ++q) b = (red[q] > b) ? red[q] : b;
        unsigned long long *slot = slots + blockIdx.x;
        if (b > *slot) *slot = b;
    }
}

struct CellIdx {
    int i, j, k, sh, h;
    int64_t o;
    bool active;
};

template <int DIM>
__device__ __forceinline__ CellIdx sweep_cell(const PfGeom &g, const SweepArgs &s) {
    CellIdx c;
    const int nb = (int)gridDim.x;
    const int L = s.xcd_remap ? xcd_logical_block((int)blockIdx.x, nb) : (int)blockIdx.x;
    const int bx = L % s.gx, byz = L / s.gx;
    const int by = byz % s.gy, bz = byz / s.gy;
    c.j = by * blockDim.y + threadIdx.y + 1;
    c.k = (DIM == 3) ? bz + 1 : 0;
    c.sh = (s.colour + c.j + c.k) & 1;
    c.h = (int)(bx * blockDim.x + threadIdx.x) + 1 - c.sh;
    c.i = 2 * c.h + c.sh;
    c.active = (c.j <= g.n) && (c.i <= g.m);
    c.o = (int64_t)c.h + (int64_t)g.HP * ((int64_t)c.j + (int64_t)g.Y * c.k);
    return c;
}

// ------------------------------------------------------------------------------------
// FP64 residual of the cells of one colour (launched for both colours) -> FP32 r; clears
// e at the cell and at the periodic ghost slot the cell feeds; block max of |r / ap|.
// ------------------------------------------------------------------------------------
template <int DIM, bool FAST>
__global__ __launch_bounds__(256) void k_residual(PfGeom g, StepArgs a, SweepArgs s, const double *__restrict__ P,
                                                  const double *__restrict__ PHIs, const double *__restrict__ BB,
                                                  const double *__restrict__ phi, float *__restrict__ R,
                                                  float *__restrict__ E, unsigned long long *slots) {
    const CellIdx q = sweep_cell<DIM>(g, s);
    double res = 0.;
    if (q.active) {
        const int64_t own = (int64_t)s.colour * g.HS, oth = (int64_t)(1 - s.colour) * g.HS;
        const int64_t HP = g.HP, HPY = (int64_t)g.HP * g.Y, o = q.o;
        const double *Po = P + own, *Px = P + oth, *Fo = PHIs + own, *Fx = PHIs + oth;
        const double pc = Po[o], pW = Px[o - 1 + q.sh], pE = Px[o + q.sh], pS = Px[o - HP], pN = Px[o + HP];
        const double fc = Fo[o], fW = Fx[o - 1 + q.sh], fE = Fx[o + q.sh], fS = Fx[o - HP], fN = Fx[o + HP];
        Coef c;
        c.ae = face_any<FAST>(fE, fc, a.dt, a.dx, a.rdx);
        c.aw = face_any<FAST>(fc, fW, a.dt, a.dx, a.rdx);
        c.an = face_any<FAST>(fN, fc, a.dt, a.dy, a.rdy);
        c.as = face_any<FAST>(fc, fS, a.dt, a.dy, a.rdy);
        double pB = 0., pT = 0.;
        c.at = 0.;
        c.ab = 0.;
        if (DIM == 3) {
            pB = Px[o - HPY];
            pT = Px[o + HPY];
            c.at = face_any<FAST>(Fx[o + HPY], fc, a.dt, a.dz, a.rdz);
            c.ab = face_any<FAST>(fc, Fx[o - HPY], a.dt, a.dz, a.rdz);
        }
        const double ap = (DIM == 3) ? (((((-c.ae) - c.aw) - c.an) - c.as) - c.at) - c.ab
                                     : (((-c.ae) - c.aw) - c.an) - c.as;
        if (is_edge_row(g, a, q.i, q.j, q.k)) row_rules<false>(g, a, q.i, q.j, q.k, c, nullptr, nullptr, phi, 0.);
        double r = BB[own + o];
        r = r - c.ae * pE;
        r = r - c.aw * pW;
        r = r - c.an * pN;
        r = r - c.as * pS;
        if (DIM == 3) {
            r = r - c.at * pT;
            r = r - c.ab * pB;
        }
        r = r - ap * pc;
        if (R) R[own + o] = (float)r;
        if (E) E[own + o] = 0.f;
        if (E && a.bc == PF_BC_UNIFORM) {  // periodic ghost slots of e that a sweep or refresh writes
            if (q.j == g.n) E[pf_soff(g, q.i, 0, q.k)] = 0.f;
            if (q.j == 1) E[pf_soff(g, q.i, g.n + 1, q.k)] = 0.f;
            if (DIM == 3) {
                if (q.k == g.l) E[pf_soff(g, q.i, q.j, 0)] = 0.f;
                if (q.k == 1) E[pf_soff(g, q.i, q.j, g.l + 1)] = 0.f;
            }
        }
        res = fabs(r / ap);
    }
    if (s.track_err) block_max_to_slot(res, slots);
}

// ------------------------------------------------------------------------------------
// FP32 red/black SOR half-sweep on A e = r (same structure as the FP64 sweep).
// ------------------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(256) void k_sor32(PfGeom g, StepArgs a, SweepArgs s, const float *__restrict__ Eoth,
                                               float *__restrict__ Eown, const float *__restrict__ F32o,
                                               const float *__restrict__ F32x, const float *__restrict__ Ro,
                                               const double *__restrict__ phi, float cdx, float cdy, float cdz,
                                               unsigned long long *err_slot) {
    const CellIdx q = sweep_cell<DIM>(g, s);
    double err = 0.;
    if (q.active) {
        const int64_t HP = g.HP, HPY = (int64_t)g.HP * g.Y, o = q.o;
        const float ec = Eown[o];
        const float eW = Eoth[o - 1 + q.sh], eE = Eoth[o + q.sh], eS = Eoth[o - HP], eN = Eoth[o + HP];
        const float fc = F32o[o];
        CoefT<float> c;
        c.ae = face32(F32x[o + q.sh], fc, cdx);
        c.aw = face32(fc, F32x[o - 1 + q.sh], cdx);
        c.an = face32(F32x[o + HP], fc, cdy);
        c.as = face32(fc, F32x[o - HP], cdy);
        float eB = 0.f, eT = 0.f;
        c.at = 0.f;
        c.ab = 0.f;
        if (DIM == 3) {
            eB = Eoth[o - HPY];
            eT = Eoth[o + HPY];
            c.at = face32(F32x[o + HPY], fc, cdz);
            c.ab = face32(fc, F32x[o - HPY], cdz);
        }
        const float ap = -(c.ae + c.aw + c.an + c.as + c.at + c.ab);
        if (is_edge_row(g, a, q.i, q.j, q.k)) row_rules<false, float>(g, a, q.i, q.j, q.k, c, nullptr, nullptr, phi, 0.);
        float num = Ro[o] - c.ae * eE - c.aw * eW - c.an * eN - c.as * eS;
        if (DIM == 3) num = num - c.at * eT - c.ab * eB;
        const float om = (float)a.omega;
        // the approximate reciprocal measured more accurate end-to-end than IEEE division
        // here: the error of a fixed-count FP32 iterate is set by rounding in the iteration
        const float en = num * __builtin_amdgcn_rcpf(ap) * om + ec * (1.f - om);
        Eown[o] = en;
        if (s.fuse_y) {
            if (q.j == g.n) Eown[o - HP * (int64_t)g.n] = en;
            if (q.j == 1) Eown[o + HP * (int64_t)g.n] = en;
        }
        if (DIM == 3 && s.fuse_z) {
            if (q.k == g.l) Eown[o - HPY * (int64_t)g.l] = en;
            if (q.k == 1) Eown[o + HPY * (int64_t)g.l] = en;
        }
        err = fabs((double)en - (double)ec);
    }
    if (s.track_err) block_max_to_slot(err, err_slot);
}

// p += e over every slot of both colours: interior cells and the periodic ghost slots
// (which hold e of their source cell, so the ghost stays equal to its source).
__global__ void k_update(const float *__restrict__ E, double *__restrict__ P, int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        P[t] = P[t] + (double)E[t];
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
    s.gx = (int)gx;
    s.gy = (int)gy;
    return dim3(gx * gy * gz);
}

void launch_residual(hipStream_t st, const PfGeom &g, const StepArgs &a, SweepArgs s, const double *P,
                     const double *PHIs, const double *BB, const double *phi, float *R, float *E,
                     unsigned long long *slots, bool fast) {
    dim3 b(64, 4, 1);
    dim3 gr = sweep_grid(g, s, b);
#define PF_RES(D, F) hipLaunchKernelGGL((k_residual<D, F>), gr, b, 0, st, g, a, s, P, PHIs, BB, phi, R, E, slots)
    if (g.dim == 3) { if (fast) PF_RES(3, true); else PF_RES(3, false); }
    else { if (fast) PF_RES(2, true); else PF_RES(2, false); }
#undef PF_RES
}

void launch_sor32(hipStream_t st, const PfGeom &g, const StepArgs &a, SweepArgs s, float *E, const float *F32,
                  const float *R, const double *phi, unsigned long long *err_slot) {
    dim3 b(64, 4, 1);
    dim3 gr = sweep_grid(g, s, b);
    const int64_t own = (int64_t)s.colour * g.HS, oth = (int64_t)(1 - s.colour) * g.HS;
    const float cdx = (float)(a.dt / (a.dx * a.dx)), cdy = (float)(a.dt / (a.dy * a.dy)),
                cdz = (float)(a.dt / (a.dz * a.dz));
    if (g.dim == 3)
        hipLaunchKernelGGL(k_sor32<3>, gr, b, 0, st, g, a, s, E + oth, E + own, F32 + own, F32 + oth, R + own, phi,
                           cdx, cdy, cdz, err_slot);
    else
        hipLaunchKernelGGL(k_sor32<2>, gr, b, 0, st, g, a, s, E + oth, E + own, F32 + own, F32 + oth, R + own, phi,
                           cdx, cdy, cdz, err_slot);
}

void launch_update(hipStream_t st, const PfGeom &g, const float *E, double *P) {
    hipLaunchKernelGGL(k_update, dim3(256 * 8), dim3(256), 0, st, E, P, 2 * g.HS);
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
