// pf_host.h -- host-side helpers shared by the single-case and the batch contexts:
// descriptor checks, step constants, porosity set-up (grid_conditions*), initial fields,
// and conversion between the reference's host storage and the device layouts.
#pragma once
#include <cmath>
#include <cstddef>
#include <string>
#include <vector>

#include "../../include/pixelflow.h"
#include "pf_kernels.h"
#include "pf_layout.h"

namespace pf {

inline bool desc_valid(const pf_desc *d, std::string &why) {
    if (!d) { why = "null descriptor"; return false; }
    if (d->dim != 2 && d->dim != 3) { why = "dim must be 2 or 3"; return false; }
    if (d->m < 3 || d->n < 2 || (d->dim == 3 && d->l < 2)) { why = "extents too small (m>=3, n>=2, l>=2)"; return false; }
    if (d->bc != PF_BC_UNIFORM && d->bc != PF_BC_AIR) { why = "bc must be PF_BC_UNIFORM or PF_BC_AIR"; return false; }
    if (d->bc == PF_BC_AIR && d->dim != 3) { why = "the air-conditioning solver is 3D"; return false; }
    for (int f = 0; f < 6; ++f)
        if (d->face[f] < 0 || d->face[f] > 2) { why = "face codes must be 0, 1 or 2"; return false; }
    if (d->iter_max < 0) { why = "iter_max < 0"; return false; }
    if (d->inlet_porosity && (d->dim != 2 || d->bc != PF_BC_UNIFORM)) {
        why = "inlet_porosity is the 2D uniform-inflow (backward step) variant";
        return false;
    }
    if (d->precision != PF_PREC_FP64 && d->precision != PF_PREC_MIXED) { why = "precision must be FP64 or MIXED"; return false; }
    if (d->inner_iters < 0 || d->max_cycles < 0 || !(d->residual_tol >= 0) || !(d->inner_tol >= 0)) {
        why = "inner_iters, max_cycles >= 0 and residual_tol, inner_tol >= 0";
        return false;
    }
    if (d->solver != PF_SOLVER_SOR && d->solver != PF_SOLVER_MG) { why = "solver must be SOR or MG"; return false; }
    if (!(d->dx > 0) || !(d->dy > 0) || (d->dim == 3 && !(d->dz > 0)) || !(d->dt > 0) || !(d->density != 0)) {
        why = "dx, dy, dz, dt must be > 0 and density != 0";
        return false;
    }
    const double cells = (double)d->m * d->n * (d->dim == 3 ? d->l : 1);
    if (cells > 2.0e9) { why = "grid too large"; return false; }
    return true;
}

// Step constants.  Inlet components as in boundary: 3D uses AoA/1300
// (ibm_3d_uniform_omp_cpu.f90:695-696), 2D AoA/180 (ibm_2d_uniform_omp_cpu.f90:478-479).
inline StepArgs make_step_args(const pf_desc &d) {
    StepArgs a{};
    a.bc = d.bc;
    a.nonslip = d.nonslip ? 1 : 0;
    a.inlet_porosity = d.inlet_porosity ? 1 : 0;
    for (int f = 0; f < 6; ++f) a.face[f] = d.face[f];
    a.dt = d.dt;
    a.dx = d.dx;
    a.dy = d.dy;
    a.dz = (d.dim == 3) ? d.dz : 1.0;
    a.rdx = 1.0 / a.dx;
    a.rdy = 1.0 / a.dy;
    a.rdz = 1.0 / a.dz;
    a.xnue = d.xnue;
    a.xlambda = d.xlambda;
    a.density = d.density;
    a.thickness = d.thickness;
    a.omega = d.omega;
    a.omega_c = 1. - d.omega;
    const double pi = std::atan(1.) * 4.;
    const double ang = (d.dim == 3) ? ((d.aoa / 1300.) * pi) : ((d.aoa / 180.) * pi);
    a.u_in = d.inlet_velocity * std::cos(ang);
    a.v_in = d.inlet_velocity * std::sin(ang);
    a.inlet_velocity = d.inlet_velocity;
    a.outlet_pressure = d.outlet_pressure;
    // face_coef(1, 1, dt, d): max(small, (1 + 1) * 0.5) is 1 and dt * 1 is dt, exactly
    a.unit_ax = (a.dt / a.dx) / a.dx;
    a.unit_ay = (a.dt / a.dy) / a.dy;
    a.unit_az = (d.dim == 3) ? (a.dt / a.dz) / a.dz : 0.;
    a.unit_ap = (d.dim == 3) ? (((((-a.unit_ax) - a.unit_ax) - a.unit_ay) - a.unit_ay) - a.unit_az) - a.unit_az
                             : (((-a.unit_ax) - a.unit_ax) - a.unit_ay) - a.unit_ay;
    return a;
}

// grid_conditions* (src/omp_parallel/lib/grid.f90): clamp to threshold, fill ghosts, in
// the natural device layout.  2D :94-105; air (zero gradient) :217-242; uniform (x outlet,
// y/z periodic) :351-377.
inline void host_porosity(const PfGeom &g, int bc, const double *raw, double threshold, std::vector<double> &phi) {
    const int m = g.m, n = g.n, l = g.l;
    phi.assign((size_t)g.NS, 0.0);
    auto at = [&](int i, int j, int k) -> double & { return phi[(size_t)pf_off(g, i, j, k)]; };
    const int kmin = (g.dim == 3) ? 1 : 0, kmax = (g.dim == 3) ? l : 0;
    for (int k = kmin; k <= kmax; ++k)
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= m; ++i) {
                const double vraw = raw[(size_t)(i - 1) + (size_t)m * ((size_t)(j - 1) + (size_t)n * (size_t)(k - kmin))];
                at(i, j, k) = (vraw > threshold) ? vraw : threshold;   // max(poro_val, threshold)
            }
    if (g.dim == 2) {
        for (int j = 1; j <= n + 1; ++j) { at(0, j, 0) = at(1, j, 0); at(m + 1, j, 0) = at(m, j, 0); }
        for (int i = 0; i <= m + 1; ++i) { at(i, 0, 0) = at(i, n, 0); at(i, n + 1, 0) = at(i, 1, 0); }
    } else if (bc == PF_BC_AIR) {
        for (int j = 0; j <= n + 1; ++j)
            for (int k = 0; k <= l + 1; ++k) { at(0, j, k) = at(1, j, k); at(m + 1, j, k) = at(m, j, k); }
        for (int i = 0; i <= m + 1; ++i)
            for (int k = 0; k <= l + 1; ++k) { at(i, 0, k) = at(i, 1, k); at(i, n + 1, k) = at(i, n, k); }
        for (int i = 0; i <= m + 1; ++i)
            for (int j = 0; j <= n + 1; ++j) { at(i, j, 0) = at(i, j, 1); at(i, j, l + 1) = at(i, j, l); }
    } else {
        for (int j = 1; j <= n + 1; ++j)
            for (int k = 1; k <= l + 1; ++k) { at(0, j, k) = at(1, j, k); at(m + 1, j, k) = at(m, j, k); }
        for (int i = 0; i <= m + 1; ++i)
            for (int k = 0; k <= l + 1; ++k) { at(i, 0, k) = at(i, n, k); at(i, n + 1, k) = at(i, 1, k); }
        for (int i = 0; i <= m + 1; ++i)
            for (int j = 0; j <= n + 1; ++j) { at(i, j, 0) = at(i, j, l); at(i, j, l + 1) = at(i, j, 1); }
    }
}

// initial_conditions: 3D uniform u = U cos(AoA/360 pi) (ibm_3d_uniform_omp_cpu.f90:782-783),
// 2D u = U cos(AoA/180 pi) (ibm_2d_uniform_omp_cpu.f90:560-561), air all zero
// (ibm_3d_air_condition_omp_cpu.f90:1200-1203); p = outlet_pressure.  Interior only.
inline void initial_values(const pf_desc &d, double &u0, double &v0) {
    u0 = 0.;
    v0 = 0.;
    if (d.bc != PF_BC_UNIFORM) return;
    const double pi = std::atan(1.) * 4.;
    const double ang = (d.dim == 3) ? ((d.aoa / 360.) * pi) : ((d.aoa / 180.) * pi);
    u0 = d.inlet_velocity * std::cos(ang);
    v0 = d.inlet_velocity * std::sin(ang);
}

inline void fill_interior(const PfGeom &g, double val, bool split, std::vector<double> &buf) {
    buf.assign((size_t)(split ? 2 * g.HS : g.NS), 0.0);
    const int kmin = (g.dim == 3) ? 1 : 0, kmax = (g.dim == 3) ? g.l : 0;
    for (int k = kmin; k <= kmax; ++k)
        for (int j = 1; j <= g.n; ++j)
            for (int i = 1; i <= g.m; ++i)
                buf[(size_t)(split ? pf_soff(g, i, j, k) : pf_off(g, i, j, k))] = val;
}

// Backward-step initial velocity: U*cos(AoA/180*pi)*porosity(i,j)
// (ibm_2d_backstep_omp_cpu.f90:615-618): the uniform value times each interior cell's
// porosity, in that order.  buf and phi are in the natural layout.
inline void scale_interior(const PfGeom &g, const std::vector<double> &phi, std::vector<double> &buf) {
    const int kmin = (g.dim == 3) ? 1 : 0, kmax = (g.dim == 3) ? g.l : 0;
    for (int k = kmin; k <= kmax; ++k)
        for (int j = 1; j <= g.n; ++j)
            for (int i = 1; i <= g.m; ++i) {
                const size_t o = (size_t)pf_off(g, i, j, k);
                buf[o] = buf[o] * phi[o];
            }
}

// device layout <-> reference host storage (Fortran order, leading dimensions ldi, ldj)
inline void to_host(const PfGeom &g, bool split, const std::vector<double> &stage, double *host, int64_t ldi,
                    int64_t ldj) {
    for (int k = 0; k < g.Z; ++k)
        for (int j = 0; j < g.Y; ++j)
            for (int i = 0; i < g.m + 2; ++i)
                host[(size_t)i + (size_t)ldi * ((size_t)j + (size_t)ldj * (size_t)k)] =
                    stage[(size_t)(split ? pf_soff(g, i, j, k) : pf_off(g, i, j, k))];
}

inline void from_host(const PfGeom &g, bool split, const double *host, int64_t ldi, int64_t ldj,
                      std::vector<double> &stage) {
    for (int k = 0; k < g.Z; ++k)
        for (int j = 0; j < g.Y; ++j)
            for (int i = 0; i < g.m + 2; ++i)
                stage[(size_t)(split ? pf_soff(g, i, j, k) : pf_off(g, i, j, k))] =
                    host[(size_t)i + (size_t)ldi * ((size_t)j + (size_t)ldj * (size_t)k)];
}

}  // namespace pf
