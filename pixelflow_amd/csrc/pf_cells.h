// pf_cells.h -- the arithmetic of one cell for every stage of the step.  Single-case
// kernels (pf_kernels.hip, pf_mixed.hip) and batched kernels (pf_batch.hip) call these same
// functions, so every path evaluates identical expressions.
#pragma once
#include <hip/hip_runtime.h>

#include "pf_device.h"
#include "pf_kernels.h"

namespace pf {

__device__ __forceinline__ double div3_at(const PfGeom &g, const StepArgs &a, const double *uo_,
                                          const double *vo_, const double *wo_, int i, int j, int k) {
    const auto uo = gmem(uo_);
    const auto vo = gmem(vo_);
    const auto wo = gmem(wo_);
    const int64_t o = pf_off(g, i, j, k);
    const int64_t sy = g.PX, sz = (int64_t)g.PX * g.Y;
    return ((((uo[o + 1] - uo[o - 1]) / a.dx) * 0.5) + (((vo[o + sy] - vo[o - sy]) / a.dy) * 0.5)) +
           (((wo[o + sz] - wo[o - sz]) / a.dz) * 0.5);
}

__device__ __forceinline__ double div2_at(const PfGeom &g, const StepArgs &a, const double *uo_,
                                          const double *vo_, int i, int j) {
    const auto uo = gmem(uo_);
    const auto vo = gmem(vo_);
    const int64_t o = pf_off(g, i, j, 0);
    // the 2D code divides the v-difference by dx (ibm_2d_uniform_omp_cpu.f90:175-176)
    return (((uo[o + 1] - uo[o - 1]) / a.dx) * .5) + (((vo[o + g.PX] - vo[o - g.PX]) / a.dx) * .5);
}

__device__ __forceinline__ void div_cell(const PfGeom &g, const StepArgs &a, const double *__restrict__ uo,
                                         const double *__restrict__ vo, const double *__restrict__ wo,
                                         double *__restrict__ div_, int i, int j, int k) {
    const auto div = gmem(div_);
    const int m = g.m, n = g.n, l = g.l;
    const bool gi = (i == 0 || i == m + 1), gj = (j == 0 || j == n + 1);
    const bool gk = (g.dim == 3) && (k == 0 || k == l + 1);
    const int ng = (int)gi + (int)gj + (int)gk;
    if (ng == 0) {
        div[pf_off(g, i, j, k)] = (g.dim == 3) ? div3_at(g, a, uo, vo, wo, i, j, k) : div2_at(g, a, uo, vo, i, j);
        return;
    }
    if (ng > 1) return;  // edges and corners are never written nor read
    double val = 0.;
    if (a.bc == PF_BC_UNIFORM && !gi) {  // y / z periodic ghosts; x ghosts are 0
        const int jj = (j == 0) ? n : (j == n + 1 ? 1 : j);
        const int kk = (g.dim == 3) ? ((k == 0) ? l : (k == l + 1 ? 1 : k)) : 0;
        val = (g.dim == 3) ? div3_at(g, a, uo, vo, wo, i, jj, kk) : div2_at(g, a, uo, vo, i, jj);
    }
    div[pf_off(g, i, j, k)] = val;
}

// Divisors of the momentum predictor: the grid spacings and the squared wall thicknesses
// (thickness * d)^2 of the wall force, each prepared once per wave (DivK).
struct PredK {
    DivK x, y, z, tx, ty, tz;
};
__device__ __forceinline__ PredK pred_consts(const StepArgs &a, int dim) {
    PredK K;
    K.x = divk_make(a.dx);
    K.y = divk_make(a.dy);
    const double thx = a.thickness * a.dx;
    K.tx = divk_make(thx * thx);
    if (dim == 3) {
        K.z = divk_make(a.dz);
        const double thy = a.thickness * a.dy, thz = a.thickness * a.dz;
        K.ty = divk_make(thy * thy);
        K.tz = divk_make(thz * thz);
    } else {
        K.z = K.ty = K.tz = K.x;
    }
    return K;
}

// Operands of the predictor at one cell: porosity p, divergence d and the old velocity
// components at the cell (c) and its neighbours (E/W in x, N/S in y, T/B in z).
struct PredIn {
    double pc, pE, pW, pN, pS, pT, pB;
    double dc, dE, dW, dN, dS, dT, dB;
    double uc, uE, uW, uN, uS, uT, uB;
    double vc, vE, vW, vN, vS, vT, vB;
    double wc, wE, wW, wN, wS, wT, wB;
};

// The 3D predictor (:233-376) with every division by a spacing or a squared thickness
// taken by dv, in the reference's operand order.  A gradient half such as
// ((uE - uW) / dx) * 0.5 appears twice in one expression and again in a neighbouring
// component; it is evaluated once (gux), the same value each time.
template <class DIV>
__device__ __forceinline__ void predict3_values(const StepArgs &a, const PredIn &x, const PredK &K, DIV &dv,
                                                double r3[3]) {
    const double dt = a.dt, xnue = a.xnue, xlambda = a.xlambda;
    const double wall = PF_ALPHA;
    const double pc = x.pc;
    const double gux = dv(x.uE - x.uW, K.x) * 0.5, guy = dv(x.uN - x.uS, K.y) * 0.5, guz = dv(x.uT - x.uB, K.z) * 0.5;
    const double gvx = dv(x.vE - x.vW, K.x) * 0.5, gvy = dv(x.vN - x.vS, K.y) * 0.5, gvz = dv(x.vT - x.vB, K.z) * 0.5;
    const double gwx = dv(x.wE - x.wW, K.x) * 0.5, gwy = dv(x.wN - x.wS, K.y) * 0.5, gwz = dv(x.wT - x.wB, K.z) * 0.5;
    const double dpx = x.pE - x.pW, dpy = x.pN - x.pS, dpz = x.pT - x.pB;
    {   // u (:233-271)
        double r = x.uc - (dv((dt * x.uc) * (x.uE - x.uW), K.x) * 0.5);
        r = r - (dv((dt * x.vc) * (x.uN - x.uS), K.y) * 0.5);
        r = r - (dv((dt * x.wc) * (x.uT - x.uB), K.z) * 0.5);
        r = r + dv(dv((dt * xnue) * ((x.uE - (2. * x.uc)) + x.uW), K.x), K.x);
        r = r + dv(dv((dt * xnue) * ((x.uN - (2. * x.uc)) + x.uS), K.y), K.y);
        r = r + dv(dv((dt * xnue) * ((x.uT - (2. * x.uc)) + x.uB), K.z), K.z);
        r = r + (dv((dt * (xnue + xlambda)) * (x.dE - x.dW), K.x) * 0.5);
        const double t1 = dv(((gux + gux) * xnue) * dpx, K.x) * 0.5;
        const double t2 = dv(((guy + gvx) * xnue) * dpy, K.y) * 0.5;
        const double t3 = dv(((guz + gwx) * xnue) * dpz, K.z) * 0.5;
        const double t4 = (dv(x.dc * dpx, K.x) * 0.5) * xlambda;
        r = r + ((dt * (((t1 + t2) + t3) + t4)) / pc);
        if (a.nonslip) r = r - ((((dv((dt * xnue) * x.uc, K.tx) * wall) * pc) * (1. - pc)) * (1. - pc));
        r3[0] = r;
    }
    {   // v (:286-323)
        double r = x.vc - (dv((dt * x.uc) * (x.vE - x.vW), K.x) * 0.5);
        r = r - (dv((dt * x.vc) * (x.vN - x.vS), K.y) * 0.5);
        r = r - (dv((dt * x.wc) * (x.vT - x.vB), K.z) * 0.5);
        r = r + dv(dv((dt * xnue) * ((x.vE - (2. * x.vc)) + x.vW), K.x), K.x);
        r = r + dv(dv((dt * xnue) * ((x.vN - (2. * x.vc)) + x.vS), K.y), K.y);
        r = r + dv(dv((dt * xnue) * ((x.vT - (2. * x.vc)) + x.vB), K.z), K.z);
        r = r + (dv((dt * (xnue + xlambda)) * (x.dN - x.dS), K.y) * 0.5);
        const double t1 = dv(((gvx + guy) * xnue) * dpx, K.x) * 0.5;
        const double t2 = dv(((gvy + gvy) * xnue) * dpy, K.y) * 0.5;
        const double t3 = dv(((gvz + gwy) * xnue) * dpz, K.z) * 0.5;
        const double t4 = (dv(x.dc * dpy, K.y) * 0.5) * xlambda;
        r = r + ((dt * (((t1 + t2) + t3) + t4)) / pc);
        if (a.nonslip) r = r - ((((dv((dt * xnue) * x.vc, K.ty) * wall) * pc) * (1. - pc)) * (1. - pc));
        r3[1] = r;
    }
    {   // w (:339-376)
        double r = x.wc - (dv((dt * x.uc) * (x.wE - x.wW), K.x) * 0.5);
        r = r - (dv((dt * x.vc) * (x.wN - x.wS), K.y) * 0.5);
        r = r - (dv((dt * x.wc) * (x.wT - x.wB), K.z) * 0.5);
        r = r + dv(dv((dt * xnue) * ((x.wE - (2. * x.wc)) + x.wW), K.x), K.x);
        r = r + dv(dv((dt * xnue) * ((x.wN - (2. * x.wc)) + x.wS), K.y), K.y);
        r = r + dv(dv((dt * xnue) * ((x.wT - (2. * x.wc)) + x.wB), K.z), K.z);
        r = r + (dv((dt * (xnue + xlambda)) * (x.dT - x.dB), K.z) * 0.5);
        const double t1 = dv(((gwx + guz) * xnue) * dpx, K.x) * 0.5;
        const double t2 = dv(((gwy + gvz) * xnue) * dpy, K.y) * 0.5;
        const double t3 = dv(((gwz + gwz) * xnue) * dpz, K.z) * 0.5;
        const double t4 = (dv(x.dc * dpz, K.z) * 0.5) * xlambda;
        r = r + ((dt * (((t1 + t2) + t3) + t4)) / pc);
        if (a.nonslip) r = r - ((((dv((dt * xnue) * x.wc, K.tz) * wall) * pc) * (1. - pc)) * (1. - pc));
        r3[2] = r;
    }
}

// The 2D predictor (ibm_2d_uniform_omp_cpu.f90:204-255); the v wall force uses dx as well
// (:254).
template <class DIV>
__device__ __forceinline__ void predict2_values(const StepArgs &a, const PredIn &x, const PredK &K, DIV &dv,
                                                double r2[2]) {
    const double dt = a.dt, xnue = a.xnue, xlambda = a.xlambda;
    const double pc = x.pc;
    const double gux = dv(x.uE - x.uW, K.x) * .5, guy = dv(x.uN - x.uS, K.y) * .5;
    const double gvx = dv(x.vE - x.vW, K.x) * .5, gvy = dv(x.vN - x.vS, K.y) * .5;
    const double dpx = x.pE - x.pW, dpy = x.pN - x.pS;
    {   // u (:204-224)
        double r = x.uc - (dt * (dv(x.uc * (x.uE - x.uW), K.x) / 2.));
        r = r - (dt * (dv(x.vc * (x.uN - x.uS), K.y) / 2.));
        r = r + dv(dv((dt * xnue) * ((x.uE - (2. * x.uc)) + x.uW), K.x), K.x);
        r = r + dv(dv((dt * xnue) * ((x.uN - (2. * x.uc)) + x.uS), K.y), K.y);
        r = r + (dv((dt * (xnue + xlambda)) * (x.dE - x.dW), K.x) * .5);
        const double t1 = dv(((gux + gux) * xnue) * dpx, K.x) * .5;
        const double t2 = dv(((guy + gvx) * xnue) * dpy, K.y) * .5;
        const double t3 = (dv(x.dc * dpx, K.x) * 0.5) * xlambda;
        r = r + ((dt * ((t1 + t2) + t3)) / pc);
        if (a.nonslip) r = r - ((((dv((dt * xnue) * x.uc, K.tx) * PF_ALPHA) * pc) * (1. - pc)) * (1. - pc));
        r2[0] = r;
    }
    {   // v (:235-255)
        double r = x.vc - (dt * (dv(x.uc * (x.vE - x.vW), K.x) / 2.));
        r = r - (dt * (dv(x.vc * (x.vN - x.vS), K.y) / 2.));
        r = r + dv(dv((dt * xnue) * ((x.vE - (2. * x.vc)) + x.vW), K.x), K.x);
        r = r + dv(dv((dt * xnue) * ((x.vN - (2. * x.vc)) + x.vS), K.y), K.y);
        r = r + (dv((dt * (xnue + xlambda)) * (x.dN - x.dS), K.y) * .5);
        const double t1 = dv(((gvx + guy) * xnue) * dpx, K.x) * .5;
        const double t2 = dv(((gvy + gvy) * xnue) * dpy, K.y) * .5;
        const double t3 = (dv(x.dc * dpy, K.y) * 0.5) * xlambda;
        r = r + ((dt * ((t1 + t2) + t3)) / pc);
        if (a.nonslip) r = r - ((((dv((dt * xnue) * x.vc, K.tx) * PF_ALPHA) * pc) * (1. - pc)) * (1. - pc));
        r2[1] = r;
    }
}

__device__ __forceinline__ void predict3_load(const PfGeom &g, const double *uo_, const double *vo_, const double *wo_,
                                              const double *div_, const double *po_, int64_t o, PredIn &x) {
    const auto uo = gmem(uo_);
    const auto vo = gmem(vo_);
    const auto wo = gmem(wo_);
    const auto div = gmem(div_);
    const auto po = gmem(po_);
    const int64_t sy = g.PX, sz = (int64_t)g.PX * g.Y;
    x.pc = po[o]; x.pE = po[o + 1]; x.pW = po[o - 1]; x.pN = po[o + sy]; x.pS = po[o - sy]; x.pT = po[o + sz]; x.pB = po[o - sz];
    x.dc = div[o]; x.dE = div[o + 1]; x.dW = div[o - 1]; x.dN = div[o + sy]; x.dS = div[o - sy]; x.dT = div[o + sz]; x.dB = div[o - sz];
    x.uc = uo[o]; x.uE = uo[o + 1]; x.uW = uo[o - 1]; x.uN = uo[o + sy]; x.uS = uo[o - sy]; x.uT = uo[o + sz]; x.uB = uo[o - sz];
    x.vc = vo[o]; x.vE = vo[o + 1]; x.vW = vo[o - 1]; x.vN = vo[o + sy]; x.vS = vo[o - sy]; x.vT = vo[o + sz]; x.vB = vo[o - sz];
    x.wc = wo[o]; x.wE = wo[o + 1]; x.wW = wo[o - 1]; x.wN = wo[o + sy]; x.wS = wo[o - sy]; x.wT = wo[o + sz]; x.wB = wo[o - sz];
}

__device__ __forceinline__ void predict3_cell(const PfGeom &g, const StepArgs &a, const double *__restrict__ uo_,
                                              const double *__restrict__ vo_, const double *__restrict__ wo_,
                                              const double *__restrict__ div_, const double *__restrict__ po_,
                                              double *__restrict__ u_, double *__restrict__ v_, double *__restrict__ w_,
                                              int i, int j, int k) {
    const auto u = gmem(u_);
    const auto v = gmem(v_);
    const auto w = gmem(w_);
    const int64_t o = pf_off(g, i, j, k);
    if (i == 0 || i == g.m + 1 || j == 0 || j == g.n + 1 || k == 0 || k == g.l + 1) {
        u[o] = gmem(uo_)[o]; v[o] = gmem(vo_)[o]; w[o] = gmem(wo_)[o];
        return;
    }
    PredIn x;
    predict3_load(g, uo_, vo_, wo_, div_, po_, o, x);
    const PredK K = pred_consts(a, 3);
    double r[3];
    DivFast fast;
    predict3_values(a, x, K, fast, r);
    if (fast.slow) {
        // the rare thread whose dividends left the range in which DivFast reuses the
        // divisor's reciprocal: the cell again with true division, from operands loaded
        // again (at an offset the compiler cannot fold, so the common path need not keep
        // its operands live for this one)
        predict3_load(g, uo_, vo_, wo_, div_, po_, o + opaque_zero(), x);
        DivIeee ieee;
        predict3_values(a, x, K, ieee, r);
    }
    u[o] = r[0]; v[o] = r[1]; w[o] = r[2];
}

__device__ __forceinline__ void predict2_load(const PfGeom &g, const double *uo_, const double *vo_,
                                              const double *div_, const double *po_, int64_t o, PredIn &x) {
    const auto uo = gmem(uo_);
    const auto vo = gmem(vo_);
    const auto div = gmem(div_);
    const auto po = gmem(po_);
    const int64_t sy = g.PX;
    x.pc = po[o]; x.pE = po[o + 1]; x.pW = po[o - 1]; x.pN = po[o + sy]; x.pS = po[o - sy];
    x.dc = div[o]; x.dE = div[o + 1]; x.dW = div[o - 1]; x.dN = div[o + sy]; x.dS = div[o - sy];
    x.uc = uo[o]; x.uE = uo[o + 1]; x.uW = uo[o - 1]; x.uN = uo[o + sy]; x.uS = uo[o - sy];
    x.vc = vo[o]; x.vE = vo[o + 1]; x.vW = vo[o - 1]; x.vN = vo[o + sy]; x.vS = vo[o - sy];
    x.pT = x.pB = x.dT = x.dB = x.uT = x.uB = x.vT = x.vB = 0.;
    x.wc = x.wE = x.wW = x.wN = x.wS = x.wT = x.wB = 0.;
}

__device__ __forceinline__ void predict2_cell(const PfGeom &g, const StepArgs &a, const double *__restrict__ uo_,
                                              const double *__restrict__ vo_, const double *__restrict__ div_,
                                              const double *__restrict__ po_, double *__restrict__ u_,
                                              double *__restrict__ v_, int i, int j) {
    const auto u = gmem(u_);
    const auto v = gmem(v_);
    const int64_t o = pf_off(g, i, j, 0);
    if (i == 0 || i == g.m + 1 || j == 0 || j == g.n + 1) {
        u[o] = gmem(uo_)[o]; v[o] = gmem(vo_)[o];
        return;
    }
    PredIn x;
    predict2_load(g, uo_, vo_, div_, po_, o, x);
    const PredK K = pred_consts(a, 2);
    double r[2];
    DivFast fast;
    predict2_values(a, x, K, fast, r);
    if (fast.slow) {  // as in predict3_cell
        predict2_load(g, uo_, vo_, div_, po_, o + opaque_zero(), x);
        DivIeee ieee;
        predict2_values(a, x, K, ieee, r);
    }
    u[o] = r[0]; v[o] = r[1];
}

__device__ __forceinline__ double bb3_at(const PfGeom &g, const StepArgs &a, const double *u_, const double *v_,
                                         const double *w_, const double *po_, int i, int j, int k) {
    const auto u = gmem(u_);
    const auto v = gmem(v_);
    const auto w = gmem(w_);
    const auto po = gmem(po_);
    const int64_t o = pf_off(g, i, j, k);
    const int64_t sy = g.PX, sz = (int64_t)g.PX * g.Y;
    const double pc = po[o], density = a.density;
    return (((((((po[o + 1] * u[o]) + (pc * u[o + 1])) * 0.5) - (((po[o - 1] * u[o]) + (pc * u[o - 1])) * 0.5)) * density) / a.dx) +
            ((((((po[o + sy] * v[o]) + (pc * v[o + sy])) * 0.5) - (((po[o - sy] * v[o]) + (pc * v[o - sy])) * 0.5)) * density) / a.dy)) +
           ((((((po[o + sz] * w[o]) + (pc * w[o + sz])) * 0.5) - (((po[o - sz] * w[o]) + (pc * w[o - sz])) * 0.5)) * density) / a.dz);
}

__device__ __forceinline__ double bb2_at(const PfGeom &g, const StepArgs &a, const double *u_, const double *v_,
                                         const double *po_, int i, int j) {
    const auto u = gmem(u_);
    const auto v = gmem(v_);
    const auto po = gmem(po_);
    const int64_t o = pf_off(g, i, j, 0);
    const int64_t sy = g.PX;
    const double pc = po[o], density = a.density;
    return ((((((po[o + 1] * u[o]) + (pc * u[o + 1])) * 0.5) - (((po[o - 1] * u[o]) + (pc * u[o - 1])) * 0.5)) * density) / a.dx) +
           ((((((po[o + sy] * v[o]) + (pc * v[o + sy])) * 0.5) - (((po[o - sy] * v[o]) + (pc * v[o - sy])) * 0.5)) * density) / a.dy);
}

__device__ __forceinline__ void rhs_cell(const PfGeom &g, const StepArgs &a, const double *__restrict__ u,
                                         const double *__restrict__ v, const double *__restrict__ w,
                                         const double *__restrict__ po_, const double *__restrict__ P,
                                         double *__restrict__ BB_, int i, int j, int k) {
    const auto po = gmem(po_);
    const auto BB = gmem(BB_);
    double bb = (g.dim == 3) ? bb3_at(g, a, u, v, w, po_, i, j, k) : bb2_at(g, a, u, v, po_, i, j);
    if (is_edge_row(g, a, i, j, k)) {
        // the row's six coefficients, evaluated exactly as solve_p does (3D :390-400)
        const int64_t o = pf_off(g, i, j, k);
        const int64_t sy = g.PX, sz = (int64_t)g.PX * g.Y;
        const double pc = po[o];
        Coef c;
        c.ae = face_coef(po[o + 1], pc, a.dt, a.dx);
        c.aw = face_coef(pc, po[o - 1], a.dt, a.dx);
        c.an = face_coef(po[o + sy], pc, a.dt, a.dy);
        c.as = face_coef(pc, po[o - sy], a.dt, a.dy);
        c.at = (g.dim == 3) ? face_coef(po[o + sz], pc, a.dt, a.dz) : 0.;
        c.ab = (g.dim == 3) ? face_coef(pc, po[o - sz], a.dt, a.dz) : 0.;
        double bbk1 = 0.;
        if (a.bc == PF_BC_AIR && k == g.l && a.face[0] == 2) bbk1 = bb3_at(g, a, u, v, w, po_, i, j, 1);
        row_rules<true>(g, a, i, j, k, c, &bb, P, po_, bbk1);
    }
    BB[pf_soff(g, i, j, k)] = bb;
}

// The SOR update of one cell from the values around it: pressure of the cell and its six
// neighbours, porosity of the same seven cells, the RHS (3D :492-516 / :549-572, 2D
// :339-352 / :373-386).  Shared by the two-launch sweep below and the fused sweep.
// The cell's six coefficients after the boundary rows' rules, and ap (the sum before them).
template <int DIM, bool FAST>
__device__ __forceinline__ double sor_coefs(const PfGeom &g, const StepArgs &a, int i, int j, int k, double fc,
                                            double fW, double fE, double fS, double fN, double fB, double fT,
                                            const double *__restrict__ phi, Coef &c) {
    const double dt = a.dt;
    c.ae = face_any<FAST>(fE, fc, dt, a.dx, a.rdx);
    c.aw = face_any<FAST>(fc, fW, dt, a.dx, a.rdx);
    c.an = face_any<FAST>(fN, fc, dt, a.dy, a.rdy);
    c.as = face_any<FAST>(fc, fS, dt, a.dy, a.rdy);
    if (DIM == 3) {
        c.at = face_any<FAST>(fT, fc, dt, a.dz, a.rdz);
        c.ab = face_any<FAST>(fc, fB, dt, a.dz, a.rdz);
    } else {
        c.at = 0.;
        c.ab = 0.;
    }
    double ap;
    if (DIM == 3) ap = (((((-c.ae) - c.aw) - c.an) - c.as) - c.at) - c.ab;
    else ap = (((-c.ae) - c.aw) - c.an) - c.as;
    if (is_edge_row(g, a, i, j, k)) row_rules<false>(g, a, i, j, k, c, nullptr, nullptr, phi, 0.);
    return ap;
}

// The SOR update from the coefficients, ap and the neighbours.
template <int DIM>
__device__ __forceinline__ double sor_apply(const StepArgs &a, const Coef &c, double ap, double bb, double pc,
                                            double pW, double pE, double pS, double pN, double pB, double pT) {
    double num;
    if (DIM == 3)
        num = (((((bb - (c.ae * pE)) - (c.aw * pW)) - (c.an * pN)) - (c.as * pS)) - (c.at * pT)) - (c.ab * pB);
    else
        num = (((bb - (c.ae * pE)) - (c.aw * pW)) - (c.an * pN)) - (c.as * pS);
    return ((num / ap) * a.omega) + (pc * a.omega_c);
}

template <int DIM, bool FAST>
__device__ __forceinline__ double sor_value(const PfGeom &g, const StepArgs &a, int i, int j, int k, double pc,
                                            double pW, double pE, double pS, double pN, double pB, double pT,
                                            double fc, double fW, double fE, double fS, double fN, double fB,
                                            double fT, double bb, const double *__restrict__ phi) {
    Coef c;
    const double ap = sor_coefs<DIM, FAST>(g, a, i, j, k, fc, fW, fE, fS, fN, fB, fT, phi, c);
    return sor_apply<DIM>(a, c, ap, bb, pc, pW, pE, pS, pN, pB, pT);
}

// The SOR update of a cell whose seven porosity operands are 1: the six faces are the
// per-axis values StepArgs unit_ax/ay/az (face_any of two unit porosities, which the
// host evaluates with the same IEEE operations; the fast path's faces were checked
// against IEEE division), ap is unit_ap, the boundary rows' rules apply as in sor_coefs,
// and num / ap is taken by dv with the wave's divisor kap.  The caller evaluates the cell
// with sor_value where dv reports a dividend outside its range.
template <int DIM, class DIV>
__device__ __forceinline__ double sor_unit_value(const PfGeom &g, const StepArgs &a, int i, int j, int k,
                                                 double pc, double pW, double pE, double pS, double pN, double pB,
                                                 double pT, double bb, const double *__restrict__ phi, DIV &dv,
                                                 const DivK &kap) {
    Coef c;
    c.ae = a.unit_ax;
    c.aw = a.unit_ax;
    c.an = a.unit_ay;
    c.as = a.unit_ay;
    c.at = (DIM == 3) ? a.unit_az : 0.;
    c.ab = (DIM == 3) ? a.unit_az : 0.;
    if (is_edge_row(g, a, i, j, k)) row_rules<false>(g, a, i, j, k, c, nullptr, nullptr, phi, 0.);
    double num;
    if (DIM == 3)
        num = (((((bb - (c.ae * pE)) - (c.aw * pW)) - (c.an * pN)) - (c.as * pS)) - (c.at * pT)) - (c.ab * pB);
    else
        num = (((bb - (c.ae * pE)) - (c.aw * pW)) - (c.an * pN)) - (c.as * pS);
    return (dv(num, kap) * a.omega) + (pc * a.omega_c);
}

// The operands of one colour-sweep cell: the unknown at the cell and at its six neighbours
// (other colour), porosity at the same seven places, and the right-hand side.
template <typename T>
struct SweepIn {
    T c, w, e, s, n, b, t;
    T fc, fw, fe, fs, fn, fb, ft;
    T rhs;
};

// Load the operands of the cell at colour-split offset o (row shift sh).  Kernels that
// update several cells per thread gather all of them before computing any, so the loads
// of every cell are in flight together: in-place stores of the sweep only touch cells of
// its own colour, which no gathered neighbour is, and the cell's own slot.
template <int DIM, typename T>
__device__ __forceinline__ void sweep_gather(const PfGeom &g, int sh, int64_t o, const T *__restrict__ Xoth,
                                             const T *__restrict__ Xown, const T *__restrict__ Fo,
                                             const T *__restrict__ Fx, const T *__restrict__ Ro, SweepIn<T> &x) {
    const int64_t HP = g.HP, HPY = (int64_t)g.HP * g.Y;
    const auto X = gmem(Xoth), F = gmem(Fx);
    x.c = gmem(Xown)[o];
    x.w = X[o - 1 + sh];
    x.e = X[o + sh];
    x.s = X[o - HP];
    x.n = X[o + HP];
    x.fc = gmem(Fo)[o];
    x.fw = F[o - 1 + sh];
    x.fe = F[o + sh];
    x.fs = F[o - HP];
    x.fn = F[o + HP];
    x.rhs = gmem(Ro)[o];
    if (DIM == 3) {
        x.b = X[o - HPY];
        x.t = X[o + HPY];
        x.fb = F[o - HPY];
        x.ft = F[o + HPY];
    } else {
        x.b = x.t = x.fb = x.ft = (T)0;
    }
}

// Porosity coded in 4 bytes, exactly: a value that FP32 holds exactly is stored as that
// float; any other value as a quiet NaN whose low 22 bits index a table of the exact FP64
// values.  The decoded double is the stored double, bit for bit.
constexpr uint32_t PHI_CODE_INDEX_BITS = 22;
__device__ __forceinline__ double phi_decode(float c, const double *__restrict__ tab) {
    return (c == c) ? (double)c : gmem(tab)[__float_as_uint(c) & ((1u << PHI_CODE_INDEX_BITS) - 1u)];
}

// sweep_gather for the FP64 sweep with coded porosity (same loads, 4-byte porosity).
template <int DIM>
__device__ __forceinline__ void sweep_gather_coded(const PfGeom &g, int sh, int64_t o, const double *__restrict__ Xoth,
                                                   const double *__restrict__ Xown, const float *__restrict__ Co,
                                                   const float *__restrict__ Cx, const double *__restrict__ tab,
                                                   const double *__restrict__ Ro, SweepIn<double> &x) {
    const int64_t HP = g.HP, HPY = (int64_t)g.HP * g.Y;
    const auto X = gmem(Xoth);
    const auto C = gmem(Cx);
    x.c = gmem(Xown)[o];
    x.w = X[o - 1 + sh];
    x.e = X[o + sh];
    x.s = X[o - HP];
    x.n = X[o + HP];
    const float cc = gmem(Co)[o], cw = C[o - 1 + sh], ce = C[o + sh], cs = C[o - HP], cn = C[o + HP];
    x.rhs = gmem(Ro)[o];
    float cb = 0.f, ct = 0.f;
    if (DIM == 3) {
        x.b = X[o - HPY];
        x.t = X[o + HPY];
        cb = C[o - HPY];
        ct = C[o + HPY];
    } else {
        x.b = x.t = 0.;
    }
    x.fc = phi_decode(cc, tab);
    x.fw = phi_decode(cw, tab);
    x.fe = phi_decode(ce, tab);
    x.fs = phi_decode(cs, tab);
    x.fn = phi_decode(cn, tab);
    if (DIM == 3) {
        x.fb = phi_decode(cb, tab);
        x.ft = phi_decode(ct, tab);
    } else {
        x.fb = x.ft = 0.;
    }
}

// sweep_gather for a wave whose every porosity operand is 1: the same pressure and
// right-hand-side loads, no porosity loads, and 1 in the seven porosity operands.  The
// update then evaluates the same IEEE operations on the same values as with the loaded
// ones, so its result is the same double.
template <int DIM, typename T>
__device__ __forceinline__ void sweep_gather_unit(const PfGeom &g, int sh, int64_t o, const T *__restrict__ Xoth,
                                                  const T *__restrict__ Xown, const T *__restrict__ Ro, SweepIn<T> &x) {
    const int64_t HP = g.HP, HPY = (int64_t)g.HP * g.Y;
    const auto X = gmem(Xoth);
    x.c = gmem(Xown)[o];
    x.w = X[o - 1 + sh];
    x.e = X[o + sh];
    x.s = X[o - HP];
    x.n = X[o + HP];
    x.rhs = gmem(Ro)[o];
    if (DIM == 3) {
        x.b = X[o - HPY];
        x.t = X[o + HPY];
    } else {
        x.b = x.t = (T)0;
    }
    x.fc = x.fw = x.fe = x.fs = x.fn = (T)1;
    x.fb = x.ft = (DIM == 3) ? (T)1 : (T)0;
}

// The porosity operands alone (the rest of sweep_gather), as stored FP64 and coded.
template <int DIM>
__device__ __forceinline__ void sweep_gather_phi(const PfGeom &g, int sh, int64_t o, const double *__restrict__ Fo,
                                                 const double *__restrict__ Fx, SweepIn<double> &x) {
    const int64_t HP = g.HP, HPY = (int64_t)g.HP * g.Y;
    const auto F = gmem(Fx);
    x.fc = gmem(Fo)[o];
    x.fw = F[o - 1 + sh];
    x.fe = F[o + sh];
    x.fs = F[o - HP];
    x.fn = F[o + HP];
    if (DIM == 3) {
        x.fb = F[o - HPY];
        x.ft = F[o + HPY];
    } else {
        x.fb = x.ft = 0.;
    }
}
template <int DIM>
__device__ __forceinline__ void sweep_gather_phi32(const PfGeom &g, int sh, int64_t o, const float *__restrict__ Fo,
                                                   const float *__restrict__ Fx, SweepIn<float> &x) {
    const int64_t HP = g.HP, HPY = (int64_t)g.HP * g.Y;
    const auto F = gmem(Fx);
    x.fc = gmem(Fo)[o];
    x.fw = F[o - 1 + sh];
    x.fe = F[o + sh];
    x.fs = F[o - HP];
    x.fn = F[o + HP];
    if (DIM == 3) {
        x.fb = F[o - HPY];
        x.ft = F[o + HPY];
    } else {
        x.fb = x.ft = 0.f;
    }
}
template <int DIM>
__device__ __forceinline__ void sweep_gather_phi_coded(const PfGeom &g, int sh, int64_t o, const float *__restrict__ Co,
                                                       const float *__restrict__ Cx, const double *__restrict__ tab,
                                                       SweepIn<double> &x) {
    const int64_t HP = g.HP, HPY = (int64_t)g.HP * g.Y;
    const auto C = gmem(Cx);
    const float cc = gmem(Co)[o], cw = C[o - 1 + sh], ce = C[o + sh], cs = C[o - HP], cn = C[o + HP];
    float cb = 0.f, ct = 0.f;
    if (DIM == 3) {
        cb = C[o - HPY];
        ct = C[o + HPY];
    }
    x.fc = phi_decode(cc, tab);
    x.fw = phi_decode(cw, tab);
    x.fe = phi_decode(ce, tab);
    x.fs = phi_decode(cs, tab);
    x.fn = phi_decode(cn, tab);
    if (DIM == 3) {
        x.fb = phi_decode(cb, tab);
        x.ft = phi_decode(ct, tab);
    } else {
        x.fb = x.ft = 0.;
    }
}

// Store a new value at the cell and, with an even periodic extent, at the ghost slot that
// mirrors it (same colour; no cell of the sweep reads it).
template <int DIM, typename T>
__device__ __forceinline__ void sweep_store(const PfGeom &g, const SweepArgs &s, int j, int k, int64_t o, T v,
                                            T *__restrict__ Xown) {
    const int64_t HP = g.HP, HPY = (int64_t)g.HP * g.Y;
    const auto X = gmem(Xown);
    X[o] = v;
    if (s.fuse_y) {
        if (j == g.n) X[o - HP * (int64_t)g.n] = v;
        if (j == 1) X[o + HP * (int64_t)g.n] = v;
    }
    if (DIM == 3 && s.fuse_z) {
        if (k == g.l) X[o - HPY * (int64_t)g.l] = v;
        if (k == 1) X[o + HPY * (int64_t)g.l] = v;
    }
}

template <int DIM, bool FAST>
__device__ __forceinline__ double sor_phi_cell(const PfGeom &g, const StepArgs &a, const SweepArgs &s, int i, int j,
                                               int k, int sh, int64_t o, const double *__restrict__ Poth,
                                               double *__restrict__ Pown, const double *__restrict__ PHIo,
                                               const double *__restrict__ PHIx, const double *__restrict__ BBo,
                                               const double *__restrict__ phi) {
    SweepIn<double> c;
    sweep_gather<DIM>(g, sh, o, Poth, Pown, PHIo, PHIx, BBo, c);
    const double pn = sor_value<DIM, FAST>(g, a, i, j, k, c.c, c.w, c.e, c.s, c.n, c.b, c.t, c.fc, c.fw, c.fe, c.fs,
                                           c.fn, c.fb, c.ft, c.rhs, phi);
    sweep_store<DIM>(g, s, j, k, o, pn, Pown);
    return fabs(pn - c.c);
}

__device__ __forceinline__ void project_cell(const PfGeom &g, const StepArgs &a, const double *__restrict__ P_,
                                             double *__restrict__ u_, double *__restrict__ v_, double *__restrict__ w_,
                                             int i, int j, int k) {
    const auto P = gmem(P_);
    const auto u = gmem(u_);
    const auto v = gmem(v_);
    const auto w = gmem(w_);
    const int64_t o = pf_off(g, i, j, k);
    const double f = a.dt / a.density;
    u[o] = u[o] - (((f * (P[pf_soff(g, i + 1, j, k)] - P[pf_soff(g, i - 1, j, k)])) / a.dx) * 0.5);
    v[o] = v[o] - (((f * (P[pf_soff(g, i, j + 1, k)] - P[pf_soff(g, i, j - 1, k)])) / a.dy) * 0.5);
    if (g.dim == 3)
        w[o] = w[o] - (((f * (P[pf_soff(g, i, j, k + 1)] - P[pf_soff(g, i, j, k - 1)])) / a.dz) * 0.5);
}

__device__ __forceinline__ void boundary_uniform_cell(const PfGeom &g, const StepArgs &a,
                                                      const double *__restrict__ po_, double *__restrict__ u_,
                                                      double *__restrict__ v_, double *__restrict__ w_,
                                                      double *__restrict__ P_, int i, int j, int k) {
    const auto po = gmem(po_);
    const auto u = gmem(u_);
    const auto v = gmem(v_);
    const auto w = gmem(w_);
    const auto P = gmem(P_);
    const int m = g.m, n = g.n, l = g.l;
    // an i = 1 cell with interior (j, k) only takes the inlet velocity (its p is untouched)
    const int jj = (j == 0) ? n : (j == n + 1 ? 1 : j);
    const int kk = (g.dim == 3) ? ((k == 0) ? l : (k == l + 1 ? 1 : k)) : 0;
    double uu, vv, ww, pp;
    if (i == 0 || i == 1) {
        uu = a.u_in; vv = a.v_in; ww = 0.;
        if (a.inlet_porosity) {      // u(1,j) = U cos * porosity(1,j); u(0,j) = u(1,j)
            const double f = po[pf_off(g, 1, jj, kk)];
            uu = uu * f;
            vv = vv * f;
        }
        pp = (i == 0) ? P[pf_soff(g, 2, jj, kk)] : P[pf_soff(g, 1, jj, kk)];
    } else if (i == m + 1) {
        if (m - 1 == 1) { uu = a.u_in; vv = a.v_in; ww = 0.; }
        else {
            const int64_t s = pf_off(g, m - 1, jj, kk);
            uu = u[s]; vv = v[s]; ww = (g.dim == 3) ? w[s] : 0.;
        }
        pp = a.outlet_pressure;
    } else {
        const int64_t s = pf_off(g, i, jj, kk);
        uu = u[s]; vv = v[s]; ww = (g.dim == 3) ? w[s] : 0.;
        pp = P[pf_soff(g, i, jj, kk)];
    }
    const int64_t o = pf_off(g, i, j, k);
    u[o] = uu;
    v[o] = vv;
    if (g.dim == 3) w[o] = ww;
    P[pf_soff(g, i, j, k)] = pp;
}

__device__ __forceinline__ void boundary_air_cell(const PfGeom &g, const StepArgs &a, int face,
                                                  const double *__restrict__ po_, double *__restrict__ u_,
                                                  double *__restrict__ v_, double *__restrict__ w_,
                                                  double *__restrict__ P_, int r, int q) {
    const auto po = gmem(po_);
    const auto u = gmem(u_);
    const auto v = gmem(v_);
    const auto w = gmem(w_);
    const auto P = gmem(P_);
    const int m = g.m, n = g.n, l = g.l;
    const double U = a.inlet_velocity, pout = a.outlet_pressure;
#define OFF(ii, jj, kk) pf_off(g, ii, jj, kk)
#define SOF(ii, jj, kk) pf_soff(g, ii, jj, kk)
    if (face == 0 || face == 1) {          // top (k = l) / bottom (k = 1): r = i, q = j
        const int i = r, j = q;
        if (i > m + 1 || j > n + 1) return;
        if (face == 0) {
            const int code = a.face[0];
            const bool wall = (code == 0) || (po[OFF(i, j, l)] < 0.9);
            if (wall) {
                u[OFF(i, j, l)] = 0.; v[OFF(i, j, l)] = 0.; w[OFF(i, j, l)] = 0.;
                w[OFF(i, j, l + 1)] = -w[OFF(i, j, l - 1)];
                P[SOF(i, j, l + 1)] = P[SOF(i, j, l - 1)];
            } else if (code == 1) {
                u[OFF(i, j, l)] = 0.; v[OFF(i, j, l)] = 0.; w[OFF(i, j, l)] = -U;
                u[OFF(i, j, l + 1)] = 0.; v[OFF(i, j, l + 1)] = 0.; w[OFF(i, j, l + 1)] = -U;
                P[SOF(i, j, l + 1)] = P[SOF(i, j, l - 1)];
            } else {
                u[OFF(i, j, l + 1)] = u[OFF(i, j, l - 1)];
                v[OFF(i, j, l + 1)] = v[OFF(i, j, l - 1)];
                w[OFF(i, j, l + 1)] = w[OFF(i, j, l - 1)];
                P[SOF(i, j, l + 1)] = pout;
            }
        } else {
            const int code = a.face[1];
            // the inlet branch tests porosity(i,j,l) as written at :948
            const bool wall = (code == 0) || (code == 1 && po[OFF(i, j, l)] < 0.9) ||
                              (code == 2 && po[OFF(i, j, 1)] < 0.9);
            if (wall) {
                u[OFF(i, j, 1)] = 0.; v[OFF(i, j, 1)] = 0.; w[OFF(i, j, 1)] = 0.;
                w[OFF(i, j, 0)] = -w[OFF(i, j, 2)];
                P[SOF(i, j, 0)] = P[SOF(i, j, 2)];
            } else if (code == 1) {
                u[OFF(i, j, 1)] = 0.; v[OFF(i, j, 1)] = 0.; w[OFF(i, j, 1)] = U;
                u[OFF(i, j, 0)] = 0.; v[OFF(i, j, 0)] = 0.; w[OFF(i, j, 0)] = U;
                P[SOF(i, j, 0)] = P[SOF(i, j, 2)];
            } else {
                u[OFF(i, j, 0)] = u[OFF(i, j, 1)];
                v[OFF(i, j, 0)] = v[OFF(i, j, 1)];
                w[OFF(i, j, 0)] = w[OFF(i, j, 1)];
                P[SOF(i, j, 0)] = pout;
            }
        }
        return;
    }
    if (face == 2 || face == 3) {          // west (i = 1) / east (i = m): r = j, q = k
        const int j = r, k = q;
        if (j > n + 1 || k > l + 1) return;
        if (face == 2) {
            const int code = a.face[3];
            const bool wall = (code == 0) || (po[OFF(1, j, k)] < 0.9);
            if (wall) {
                u[OFF(1, j, k)] = 0.; v[OFF(1, j, k)] = 0.; w[OFF(1, j, k)] = 0.;
                u[OFF(0, j, k)] = -u[OFF(2, j, k)];
                P[SOF(0, j, k)] = P[SOF(2, j, k)];
            } else if (code == 1) {
                u[OFF(1, j, k)] = U; v[OFF(1, j, k)] = 0.; w[OFF(1, j, k)] = 0.;
                u[OFF(0, j, k)] = U; v[OFF(0, j, k)] = 0.; w[OFF(0, j, k)] = 0.;
                P[SOF(0, j, k)] = P[SOF(2, j, k)];
            } else {
                u[OFF(0, j, k)] = u[OFF(1, j, k)];
                v[OFF(0, j, k)] = v[OFF(1, j, k)];
                w[OFF(0, j, k)] = w[OFF(1, j, k)];
                P[SOF(0, j, k)] = pout;
            }
        } else {
            const int code = a.face[2];
            const bool wall = (code == 0) || (po[OFF(m, j, k)] < 0.9);
            if (wall) {
                u[OFF(m, j, k)] = 0.; v[OFF(m, j, k)] = 0.; w[OFF(m, j, k)] = 0.;
                u[OFF(m + 1, j, k)] = -u[OFF(m - 1, j, k)];
                P[SOF(m + 1, j, k)] = P[SOF(m - 1, j, k)];
            } else if (code == 1) {
                u[OFF(m, j, k)] = -U; v[OFF(m, j, k)] = 0.; w[OFF(m, j, k)] = 0.;
                u[OFF(m + 1, j, k)] = -U; v[OFF(m + 1, j, k)] = 0.; w[OFF(m + 1, j, k)] = 0.;
                P[SOF(m + 1, j, k)] = P[SOF(m - 1, j, k)];
            } else {
                u[OFF(m + 1, j, k)] = u[OFF(m, j, k)];
                v[OFF(m + 1, j, k)] = v[OFF(m, j, k)];
                w[OFF(m + 1, j, k)] = w[OFF(m, j, k)];
                P[SOF(m + 1, j, k)] = pout;
            }
        }
        return;
    }
    {                                      // north (j = n) / south (j = 1): r = i, q = k
        const int i = r, k = q;
        if (i > m + 1 || k > l + 1) return;
        if (face == 4) {
            const int code = a.face[4];
            const bool wall = (code == 0) || (po[OFF(i, n, k)] < 0.9);
            if (wall) {
                u[OFF(i, n, k)] = 0.; v[OFF(i, n, k)] = 0.; w[OFF(i, n, k)] = 0.;
                u[OFF(i, n + 1, k)] = -u[OFF(i, n - 1, k)];   // u, as written at :1087
                P[SOF(i, n + 1, k)] = P[SOF(i, n - 1, k)];
            } else if (code == 1) {
                u[OFF(i, n, k)] = -U; v[OFF(i, n, k)] = 0.; w[OFF(i, n, k)] = 0.;
                u[OFF(i, n + 1, k)] = -U; v[OFF(i, n + 1, k)] = 0.; w[OFF(i, n + 1, k)] = 0.;
                P[SOF(i, n + 1, k)] = P[SOF(i, n - 1, k)];
            } else {
                u[OFF(i, n + 1, k)] = u[OFF(i, n, k)];
                v[OFF(i, n + 1, k)] = v[OFF(i, n, k)];
                w[OFF(i, n + 1, k)] = w[OFF(i, n, k)];
                P[SOF(i, n + 1, k)] = pout;
            }
        } else {
            const int code = a.face[5];
            const bool wall = (code == 0) || (po[OFF(i, 1, k)] < 0.9);
            if (wall) {
                u[OFF(i, 1, k)] = 0.; v[OFF(i, 1, k)] = 0.; w[OFF(i, 1, k)] = 0.;
                v[OFF(i, 0, k)] = -v[OFF(i, 2, k)];
                P[SOF(i, 0, k)] = P[SOF(i, 2, k)];
            } else if (code == 1) {
                u[OFF(i, 1, k)] = U; v[OFF(i, 1, k)] = 0.; w[OFF(i, 1, k)] = 0.;
                u[OFF(i, 0, k)] = U; v[OFF(i, 0, k)] = 0.; w[OFF(i, 0, k)] = 0.;
                P[SOF(i, 0, k)] = P[SOF(i, 2, k)];
            } else {
                u[OFF(i, 0, k)] = u[OFF(i, 2, k)];    // (i,2,k), as written at :1153
                v[OFF(i, 0, k)] = v[OFF(i, 2, k)];
                w[OFF(i, 0, k)] = w[OFF(i, 2, k)];
                P[SOF(i, 0, k)] = pout;
            }
        }
    }
#undef OFF
#undef SOF
}

__device__ __forceinline__ void divout_cell(const PfGeom &g, const StepArgs &a, const double *__restrict__ po,
                                            const double *__restrict__ u, const double *__restrict__ v,
                                            const double *__restrict__ w, double *__restrict__ out, int i, int j,
                                            int k) {
    const int64_t o = pf_off(g, i, j, k);
    const int64_t sy = g.PX, sz = (int64_t)g.PX * g.Y;
    const double pc = po[o];
    double r = ((((po[o + 1] * u[o]) + (pc * u[o + 1])) / 2.) - (((po[o - 1] * u[o]) + (pc * u[o - 1])) / 2.)) / a.dx;
    r = r + (((((po[o + sy] * v[o]) + (pc * v[o + sy])) / 2.) - (((po[o - sy] * v[o]) + (pc * v[o - sy])) / 2.)) / a.dy);
    if (g.dim == 3)
        r = r + (((((po[o + sz] * w[o]) + (pc * w[o + sz])) / 2.) - (((po[o - sz] * w[o]) + (pc * w[o - sz])) / 2.)) / a.dz);
    const int64_t d = (int64_t)(i - 1) + (int64_t)g.m * ((int64_t)(j - 1) + (int64_t)g.n * (g.dim == 3 ? k - 1 : 0));
    out[d] = r;
}

struct CellIdx {
    int i, j, k, sh, h;
    int64_t o;
    bool active;
};

__device__ __forceinline__ float face32(float pa, float pb, float cdt) {
    return fmaxf(1.e-6f, (pa + pb) * 0.5f) * cdt;
}

template <int DIM>
__device__ __forceinline__ CellIdx sweep_cell(const PfGeom &g, const SweepArgs &s) {
    CellIdx c;
    const int nb = (int)gridDim.x;
    const int L = s.xcd_remap ? xcd_logical_block((int)blockIdx.x, nb) : (int)blockIdx.x;
    const int byz = (int)grid_div((uint32_t)L, (uint32_t)s.gx, s.mgx), bx = L - byz * s.gx;
    const int bz = (int)grid_div((uint32_t)byz, (uint32_t)s.gy, s.mgy), by = byz - bz * s.gy;
    c.j = by * blockDim.y + threadIdx.y + 1;
    c.k = (DIM == 3) ? bz + 1 : 0;
    c.sh = (s.colour + c.j + c.k) & 1;
    c.h = (int)(bx * blockDim.x + threadIdx.x) + 1 - c.sh;
    c.i = 2 * c.h + c.sh;
    c.active = (c.j <= g.n) && (c.i <= g.m);
    c.o = (int64_t)c.h + (int64_t)g.HP * ((int64_t)c.j + (int64_t)g.Y * c.k);
    return c;
}

// The FP64 residual r = bb - A p of one cell, stored as TR (float for the mixed solve,
// double for the FP64 one); E (optional) is cleared at the cell and at the periodic ghost
// slot it feeds.  Returns |r / ap|.
template <int DIM, bool FAST, typename TR = float>
__device__ __forceinline__ double residual_cell(const PfGeom &g, const StepArgs &a, int colour, int i, int j, int k,
                                                int sh, int64_t o, const double *__restrict__ P_,
                                                const double *__restrict__ PHIs_, const double *__restrict__ BB_,
                                                const double *__restrict__ phi, TR *__restrict__ R_,
                                                TR *__restrict__ E_, bool unit = false) {
    const auto P = gmem(P_);
    const auto PHIs = gmem(PHIs_);
    const auto BB = gmem(BB_);
    const auto R = gmem(R_);
    const auto E = gmem(E_);
    const int64_t own = (int64_t)colour * g.HS, oth = (int64_t)(1 - colour) * g.HS;
    const int64_t HP = g.HP, HPY = (int64_t)g.HP * g.Y;
    const auto Po = P + own, Px = P + oth, Fo = PHIs + own, Fx = PHIs + oth;
    const double pc = Po[o], pW = Px[o - 1 + sh], pE = Px[o + sh], pS = Px[o - HP], pN = Px[o + HP];
    double pB = 0., pT = 0.;
    if (DIM == 3) {
        pB = Px[o - HPY];
        pT = Px[o + HPY];
    }
    Coef c;
    double ap;
    if (unit) {
        // every porosity operand is 1 (a unit-porosity chunk): the unit faces and their sum
        c.ae = a.unit_ax;
        c.aw = a.unit_ax;
        c.an = a.unit_ay;
        c.as = a.unit_ay;
        c.at = (DIM == 3) ? a.unit_az : 0.;
        c.ab = (DIM == 3) ? a.unit_az : 0.;
        ap = a.unit_ap;
    } else {
        const double fc = Fo[o], fW = Fx[o - 1 + sh], fE = Fx[o + sh], fS = Fx[o - HP], fN = Fx[o + HP];
        c.ae = face_any<FAST>(fE, fc, a.dt, a.dx, a.rdx);
        c.aw = face_any<FAST>(fc, fW, a.dt, a.dx, a.rdx);
        c.an = face_any<FAST>(fN, fc, a.dt, a.dy, a.rdy);
        c.as = face_any<FAST>(fc, fS, a.dt, a.dy, a.rdy);
        c.at = 0.;
        c.ab = 0.;
        if (DIM == 3) {
            c.at = face_any<FAST>(Fx[o + HPY], fc, a.dt, a.dz, a.rdz);
            c.ab = face_any<FAST>(fc, Fx[o - HPY], a.dt, a.dz, a.rdz);
        }
        ap = (DIM == 3) ? (((((-c.ae) - c.aw) - c.an) - c.as) - c.at) - c.ab : (((-c.ae) - c.aw) - c.an) - c.as;
    }
    if (is_edge_row(g, a, i, j, k)) row_rules<false>(g, a, i, j, k, c, nullptr, nullptr, phi, 0.);
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
    if (R) R[own + o] = (TR)r;
    if (E) E[own + o] = TR(0);
    if (E && a.bc == PF_BC_UNIFORM) {  // periodic ghost slots of e that a sweep or refresh writes
        if (j == g.n) E[pf_soff(g, i, 0, k)] = TR(0);
        if (j == 1) E[pf_soff(g, i, g.n + 1, k)] = TR(0);
        if (DIM == 3) {
            if (k == g.l) E[pf_soff(g, i, j, 0)] = TR(0);
            if (k == 1) E[pf_soff(g, i, j, g.l + 1)] = TR(0);
        }
    }
    return fabs(r / ap);
}

// FP32 SOR update of the correction e on A e = r (mixed precision), from the values
// around the cell; shared by the two-launch and the fused sweep.
template <int DIM>
__device__ __forceinline__ float sor32_value(const PfGeom &g, const StepArgs &a, int i, int j, int k, float ec,
                                             float eW, float eE, float eS, float eN, float eB, float eT, float fc,
                                             float fW, float fE, float fS, float fN, float fB, float fT, float rr,
                                             const double *__restrict__ phi, float cdx, float cdy, float cdz) {
    CoefT<float> c;
    c.ae = face32(fE, fc, cdx);
    c.aw = face32(fc, fW, cdx);
    c.an = face32(fN, fc, cdy);
    c.as = face32(fc, fS, cdy);
    c.at = 0.f;
    c.ab = 0.f;
    if (DIM == 3) {
        c.at = face32(fT, fc, cdz);
        c.ab = face32(fc, fB, cdz);
    }
    const float ap = -(c.ae + c.aw + c.an + c.as + c.at + c.ab);
    if (is_edge_row(g, a, i, j, k)) row_rules<false, float>(g, a, i, j, k, c, nullptr, nullptr, phi, 0.);
    float num = rr - c.ae * eE - c.aw * eW - c.an * eN - c.as * eS;
    if (DIM == 3) num = num - c.at * eT - c.ab * eB;
    const float om = (float)a.omega;
    // the approximate reciprocal measured more accurate end-to-end than IEEE division here:
    // the error of a fixed-count FP32 iterate is set by rounding inside the iteration
    return num * __builtin_amdgcn_rcpf(ap) * om + ec * (1.f - om);
}

template <int DIM>
__device__ __forceinline__ double sor32_cell(const PfGeom &g, const StepArgs &a, const SweepArgs &s, int i, int j,
                                             int k, int sh, int64_t o, const float *__restrict__ Eoth,
                                             float *__restrict__ Eown, const float *__restrict__ F32o,
                                             const float *__restrict__ F32x, const float *__restrict__ Ro,
                                             const double *__restrict__ phi, float cdx, float cdy, float cdz) {
    SweepIn<float> c;
    sweep_gather<DIM>(g, sh, o, Eoth, Eown, F32o, F32x, Ro, c);
    const float en = sor32_value<DIM>(g, a, i, j, k, c.c, c.w, c.e, c.s, c.n, c.b, c.t, c.fc, c.fw, c.fe, c.fs, c.fn,
                                      c.fb, c.ft, c.rhs, phi, cdx, cdy, cdz);
    sweep_store<DIM>(g, s, j, k, o, en, Eown);
    return fabs((double)en - (double)c.c);
}

}  // namespace pf
