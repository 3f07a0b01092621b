/*
 * pf_oracle.c -- CPU restatement of PixelFlow's time step.  TEST INFRASTRUCTURE ONLY:
 * the results the HIP path is compared with, and the CPU baseline timed beside it.  The product never
 * links, loads or calls this file.
 *
 * Each function restates the reference's Fortran (paths relative to the reference tree)
 * in the same arithmetic order, with the same loop nests and OpenMP work sharing, built
 * with -ffp-contract=off so that IEEE FP64 results match the reference compiled with
 * -fdefault-real-8.  Fortran evaluates  a*b/c*d  as ((a*b)/c)*d  and  -a-b  as (-a)-b;
 * every expression below is parenthesised to make that order explicit.
 *
 * Parity of this file is pinned by tests/test_oracle_golden.py against fixtures written
 * by the reference build itself (tests/golden/gen_golden.py) and by closed-form checks.
 */
#include "pf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct pfo_ctx {
    pfo_case c;
    size_t ntot;
    double *u_old, *v_old, *w_old, *div;
    double *ap, *ae, *aw, *an, *as, *at, *ab, *bb, *p_old;
};

#define SMALL 1.e-6
#define ALPHA 32.0

/* full-array offset helpers (ghost layers included) */
#define I3(i, j, k) ((size_t)(i) + (size_t)(M2) * ((size_t)(j) + (size_t)(N2) * (size_t)(k)))
#define I2(i, j) ((size_t)(i) + (size_t)(M2) * (size_t)(j))

static double fmaxf64(double a, double b) { return (a > b) ? a : b; }

static size_t full_size(const pfo_case *c) {
    size_t s = (size_t)(c->m + 2) * (size_t)(c->n + 2);
    if (c->dim == 3) s *= (size_t)(c->l + 2);
    return s;
}

pfo_ctx *pfo_create(const pfo_case *c) {
    if (!c || (c->dim != 2 && c->dim != 3) || c->m < 2 || c->n < 2 || (c->dim == 3 && c->l < 2))
        return NULL;
    pfo_ctx *x = (pfo_ctx *)calloc(1, sizeof(pfo_ctx));
    if (!x) return NULL;
    x->c = *c;
    x->ntot = full_size(c);
    /* The reference keeps these as SAVEd (zero-initialised) locals (-fno-automatic). */
    double **arrs[] = {&x->u_old, &x->v_old, &x->w_old, &x->div, &x->ap, &x->ae, &x->aw,
                       &x->an,    &x->as,    &x->at,    &x->ab,  &x->bb, &x->p_old};
    for (size_t a = 0; a < sizeof(arrs) / sizeof(arrs[0]); ++a) {
        *arrs[a] = (double *)calloc(x->ntot, sizeof(double));
        if (!*arrs[a]) { pfo_destroy(x); return NULL; }
    }
    return x;
}

void pfo_destroy(pfo_ctx *x) {
    if (!x) return;
    free(x->u_old); free(x->v_old); free(x->w_old); free(x->div);
    free(x->ap); free(x->ae); free(x->aw); free(x->an); free(x->as);
    free(x->at); free(x->ab); free(x->bb); free(x->p_old);
    free(x);
}

/* ------------------------------------------------------------------------------------ */
/* grid_conditions (src/omp_parallel/lib/grid.f90:36-47, 92-106),                        */
/* grid_conditions_wall (grid.f90:151-158, 215-243),                                     */
/* grid_conditions_yz_periodic (grid.f90:285-292, 349-377)                               */
/* ------------------------------------------------------------------------------------ */
void pfo_setup_porosity(const pfo_case *c, const double *raw, double threshold, double *phi) {
    const int m = c->m, n = c->n, l = c->l, M2 = m + 2, N2 = n + 2;
    memset(phi, 0, full_size(c) * sizeof(double));
    if (c->dim == 2) {
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= m; ++i)
                phi[I2(i, j)] = fmaxf64(raw[(size_t)(i - 1) + (size_t)m * (j - 1)], threshold);
        for (int j = 1; j <= n + 1; ++j) {          /* grid.f90:94-97 */
            phi[I2(0, j)] = phi[I2(1, j)];
            phi[I2(m + 1, j)] = phi[I2(m, j)];
        }
        for (int i = 0; i <= m + 1; ++i) {          /* grid.f90:102-105 */
            phi[I2(i, 0)] = phi[I2(i, n)];
            phi[I2(i, n + 1)] = phi[I2(i, 1)];
        }
        return;
    }
    for (int k = 1; k <= l; ++k)
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= m; ++i)
                phi[I3(i, j, k)] = fmaxf64(
                    raw[(size_t)(i - 1) + (size_t)m * ((size_t)(j - 1) + (size_t)n * (k - 1))],
                    threshold);
    if (c->bc == PFO_BC_AIR) {
        for (int j = 0; j <= n + 1; ++j)            /* grid.f90:217-222 */
            for (int k = 0; k <= l + 1; ++k) {
                phi[I3(0, j, k)] = phi[I3(1, j, k)];
                phi[I3(m + 1, j, k)] = phi[I3(m, j, k)];
            }
        for (int i = 0; i <= m + 1; ++i)            /* grid.f90:227-232 */
            for (int k = 0; k <= l + 1; ++k) {
                phi[I3(i, 0, k)] = phi[I3(i, 1, k)];
                phi[I3(i, n + 1, k)] = phi[I3(i, n, k)];
            }
        for (int i = 0; i <= m + 1; ++i)            /* grid.f90:237-242 */
            for (int j = 0; j <= n + 1; ++j) {
                phi[I3(i, j, 0)] = phi[I3(i, j, 1)];
                phi[I3(i, j, l + 1)] = phi[I3(i, j, l)];
            }
        return;
    }
    for (int j = 1; j <= n + 1; ++j)                /* grid.f90:351-356 */
        for (int k = 1; k <= l + 1; ++k) {
            phi[I3(0, j, k)] = phi[I3(1, j, k)];
            phi[I3(m + 1, j, k)] = phi[I3(m, j, k)];
        }
    for (int i = 0; i <= m + 1; ++i)                /* grid.f90:361-366 */
        for (int k = 0; k <= l + 1; ++k) {
            phi[I3(i, 0, k)] = phi[I3(i, n, k)];
            phi[I3(i, n + 1, k)] = phi[I3(i, 1, k)];
        }
    for (int i = 0; i <= m + 1; ++i)                /* grid.f90:372-377 */
        for (int j = 0; j <= n + 1; ++j) {
            phi[I3(i, j, 0)] = phi[I3(i, j, l)];
            phi[I3(i, j, l + 1)] = phi[I3(i, j, 1)];
        }
}

/* ------------------------------------------------------------------------------------ */
/* boundary: 2D src/omp_parallel/ibm_2d_uniform_omp_cpu.f90:460-538;                     */
/* 3D ibm_3d_uniform_omp_cpu.f90:669-752; air ibm_3d_air_condition_omp_cpu.f90:873-1170  */
/* ------------------------------------------------------------------------------------ */
static void boundary_2d(const pfo_case *c, const double *phi, double *u, double *v, double *p) {
    const int m = c->m, n = c->n, M2 = m + 2;
    const double pai = atan(1.) * 4.;
    for (int j = 1; j <= n; ++j) {
        u[I2(1, j)] = c->inlet_velocity * cos(((c->aoa / 180.) * pai));
        v[I2(1, j)] = c->inlet_velocity * sin(((c->aoa / 180.) * pai));
        if (c->inlet_porosity) {   /* ibm_2d_backstep_omp_cpu.f90:533-534 */
            u[I2(1, j)] = u[I2(1, j)] * phi[I2(1, j)];
            v[I2(1, j)] = v[I2(1, j)] * phi[I2(1, j)];
        }
        u[I2(0, j)] = u[I2(1, j)];
        v[I2(0, j)] = v[I2(1, j)];
        p[I2(0, j)] = p[I2(2, j)];
    }
    for (int j = 1; j <= n; ++j) {
        u[I2(m + 1, j)] = u[I2(m - 1, j)];
        v[I2(m + 1, j)] = v[I2(m - 1, j)];
        p[I2(m + 1, j)] = c->outlet_pressure;
    }
    for (int i = 0; i <= m + 1; ++i) {
        u[I2(i, 0)] = u[I2(i, n)];
        v[I2(i, 0)] = v[I2(i, n)];
        p[I2(i, 0)] = p[I2(i, n)];
        u[I2(i, n + 1)] = u[I2(i, 1)];
        v[I2(i, n + 1)] = v[I2(i, 1)];
        p[I2(i, n + 1)] = p[I2(i, 1)];
    }
}

static void boundary_3d_uniform(const pfo_case *c, double *u, double *v, double *w, double *p) {
    const int m = c->m, n = c->n, l = c->l, M2 = m + 2, N2 = n + 2;
    const double pi = atan(1.) * 4.;
    const double uin = c->inlet_velocity * cos(((c->aoa / 1300.) * pi));
    const double vin = c->inlet_velocity * sin(((c->aoa / 1300.) * pi));
#pragma omp parallel
    {
#pragma omp for
        for (int k = 1; k <= l; ++k)
            for (int j = 1; j <= n; ++j) {
                u[I3(1, j, k)] = uin;
                v[I3(1, j, k)] = vin;
                w[I3(1, j, k)] = 0.;
                u[I3(0, j, k)] = u[I3(1, j, k)];
                v[I3(0, j, k)] = v[I3(1, j, k)];
                w[I3(0, j, k)] = w[I3(1, j, k)];
                p[I3(0, j, k)] = p[I3(2, j, k)];
            }
#pragma omp for
        for (int k = 1; k <= l; ++k)
            for (int j = 1; j <= n; ++j) {
                u[I3(m + 1, j, k)] = u[I3(m - 1, j, k)];
                v[I3(m + 1, j, k)] = v[I3(m - 1, j, k)];
                w[I3(m + 1, j, k)] = w[I3(m - 1, j, k)];
                p[I3(m + 1, j, k)] = c->outlet_pressure;
            }
#pragma omp for
        for (int k = 0; k <= l + 1; ++k)
            for (int i = 0; i <= m + 1; ++i) {
                u[I3(i, 0, k)] = u[I3(i, n, k)];
                v[I3(i, 0, k)] = v[I3(i, n, k)];
                w[I3(i, 0, k)] = w[I3(i, n, k)];
                p[I3(i, 0, k)] = p[I3(i, n, k)];
                u[I3(i, n + 1, k)] = u[I3(i, 1, k)];
                v[I3(i, n + 1, k)] = v[I3(i, 1, k)];
                w[I3(i, n + 1, k)] = w[I3(i, 1, k)];
                p[I3(i, n + 1, k)] = p[I3(i, 1, k)];
            }
#pragma omp for
        for (int j = 0; j <= n + 1; ++j)
            for (int i = 0; i <= m + 1; ++i) {
                u[I3(i, j, 0)] = u[I3(i, j, l)];
                v[I3(i, j, 0)] = v[I3(i, j, l)];
                w[I3(i, j, 0)] = w[I3(i, j, l)];
                p[I3(i, j, 0)] = p[I3(i, j, l)];
                u[I3(i, j, l + 1)] = u[I3(i, j, 1)];
                v[I3(i, j, l + 1)] = v[I3(i, j, 1)];
                w[I3(i, j, l + 1)] = w[I3(i, j, 1)];
                p[I3(i, j, l + 1)] = p[I3(i, j, 1)];
            }
    }
}

/* Air-conditioning faces: face codes 0 wall, 1 inlet (where porosity >= 0.9), 2 outlet
 * (where porosity >= 0.9).  Order top, bottom, west, east, north, south and the index
 * ranges (ghost rows included) follow ibm_3d_air_condition_omp_cpu.f90:890-1167.  The
 * bottom inlet tests porosity(i,j,l) as written at :948. */
static void boundary_3d_air(const pfo_case *c, const double *phi, double *u, double *v,
                            double *w, double *p) {
    const int m = c->m, n = c->n, l = c->l, M2 = m + 2, N2 = n + 2;
    const double U = c->inlet_velocity, po = c->outlet_pressure;
    const int top = c->face[0], bottom = c->face[1], east = c->face[2], west = c->face[3],
              north = c->face[4], south = c->face[5];
    for (int i = 0; i <= m + 1; ++i)                /* top :891-935 */
        for (int j = 0; j <= n + 1; ++j) {
            int wall = (top == 0) || (top != 0 && phi[I3(i, j, l)] < 0.9);
            if (wall) {
                u[I3(i, j, l)] = 0.; v[I3(i, j, l)] = 0.; w[I3(i, j, l)] = 0.;
                w[I3(i, j, l + 1)] = -w[I3(i, j, l - 1)];
                p[I3(i, j, l + 1)] = p[I3(i, j, l - 1)];
            } else if (top == 1) {
                u[I3(i, j, l)] = 0.; v[I3(i, j, l)] = 0.; w[I3(i, j, l)] = -U;
                u[I3(i, j, l + 1)] = u[I3(i, j, l)];
                v[I3(i, j, l + 1)] = v[I3(i, j, l)];
                w[I3(i, j, l + 1)] = w[I3(i, j, l)];
                p[I3(i, j, l + 1)] = p[I3(i, j, l - 1)];
            } else {
                u[I3(i, j, l + 1)] = u[I3(i, j, l - 1)];
                v[I3(i, j, l + 1)] = v[I3(i, j, l - 1)];
                w[I3(i, j, l + 1)] = w[I3(i, j, l - 1)];
                p[I3(i, j, l + 1)] = po;
            }
        }
    for (int i = 0; i <= m + 1; ++i)                /* bottom :938-982 */
        for (int j = 0; j <= n + 1; ++j) {
            int wall = (bottom == 0) ||
                       (bottom == 1 && phi[I3(i, j, l)] < 0.9) ||
                       (bottom == 2 && phi[I3(i, j, 1)] < 0.9);
            if (wall) {
                u[I3(i, j, 1)] = 0.; v[I3(i, j, 1)] = 0.; w[I3(i, j, 1)] = 0.;
                w[I3(i, j, 0)] = -w[I3(i, j, 2)];
                p[I3(i, j, 0)] = p[I3(i, j, 2)];
            } else if (bottom == 1) {
                u[I3(i, j, 1)] = 0.; v[I3(i, j, 1)] = 0.; w[I3(i, j, 1)] = U;
                u[I3(i, j, 0)] = u[I3(i, j, 1)];
                v[I3(i, j, 0)] = v[I3(i, j, 1)];
                w[I3(i, j, 0)] = w[I3(i, j, 1)];
                p[I3(i, j, 0)] = p[I3(i, j, 2)];
            } else {
                u[I3(i, j, 0)] = u[I3(i, j, 1)];
                v[I3(i, j, 0)] = v[I3(i, j, 1)];
                w[I3(i, j, 0)] = w[I3(i, j, 1)];
                p[I3(i, j, 0)] = po;
            }
        }
    for (int j = 0; j <= n + 1; ++j)                /* west :985-1029 */
        for (int k = 0; k <= l + 1; ++k) {
            int wall = (west == 0) || (west != 0 && phi[I3(1, j, k)] < 0.9);
            if (wall) {
                u[I3(1, j, k)] = 0.; v[I3(1, j, k)] = 0.; w[I3(1, j, k)] = 0.;
                u[I3(0, j, k)] = -u[I3(2, j, k)];
                p[I3(0, j, k)] = p[I3(2, j, k)];
            } else if (west == 1) {
                u[I3(1, j, k)] = U; v[I3(1, j, k)] = 0.; w[I3(1, j, k)] = 0.;
                u[I3(0, j, k)] = u[I3(1, j, k)];
                v[I3(0, j, k)] = v[I3(1, j, k)];
                w[I3(0, j, k)] = w[I3(1, j, k)];
                p[I3(0, j, k)] = p[I3(2, j, k)];
            } else {
                u[I3(0, j, k)] = u[I3(1, j, k)];
                v[I3(0, j, k)] = v[I3(1, j, k)];
                w[I3(0, j, k)] = w[I3(1, j, k)];
                p[I3(0, j, k)] = po;
            }
        }
    for (int j = 0; j <= n + 1; ++j)                /* east :1032-1078 */
        for (int k = 0; k <= l + 1; ++k) {
            int wall = (east == 0) || (east != 0 && phi[I3(m, j, k)] < 0.9);
            if (wall) {
                u[I3(m, j, k)] = 0.; v[I3(m, j, k)] = 0.; w[I3(m, j, k)] = 0.;
                u[I3(m + 1, j, k)] = -u[I3(m - 1, j, k)];
                p[I3(m + 1, j, k)] = p[I3(m - 1, j, k)];
            } else if (east == 1) {
                u[I3(m, j, k)] = -U; v[I3(m, j, k)] = 0.; w[I3(m, j, k)] = 0.;
                u[I3(m + 1, j, k)] = u[I3(m, j, k)];
                v[I3(m + 1, j, k)] = v[I3(m, j, k)];
                w[I3(m + 1, j, k)] = w[I3(m, j, k)];
                p[I3(m + 1, j, k)] = p[I3(m - 1, j, k)];
            } else {
                u[I3(m + 1, j, k)] = u[I3(m, j, k)];
                v[I3(m + 1, j, k)] = v[I3(m, j, k)];
                w[I3(m + 1, j, k)] = w[I3(m, j, k)];
                p[I3(m + 1, j, k)] = po;
            }
        }
    for (int i = 0; i <= m + 1; ++i)                /* north :1081-1120 */
        for (int k = 0; k <= l + 1; ++k) {
            int wall = (north == 0) || (north != 0 && phi[I3(i, n, k)] < 0.9);
            if (wall) {
                u[I3(i, n, k)] = 0.; v[I3(i, n, k)] = 0.; w[I3(i, n, k)] = 0.;
                u[I3(i, n + 1, k)] = -u[I3(i, n - 1, k)];
                p[I3(i, n + 1, k)] = p[I3(i, n - 1, k)];
            } else if (north == 1) {
                u[I3(i, n, k)] = -U; v[I3(i, n, k)] = 0.; w[I3(i, n, k)] = 0.;
                u[I3(i, n + 1, k)] = u[I3(i, n, k)];
                v[I3(i, n + 1, k)] = v[I3(i, n, k)];
                w[I3(i, n + 1, k)] = w[I3(i, n, k)];
                p[I3(i, n + 1, k)] = p[I3(i, n - 1, k)];
            } else {
                u[I3(i, n + 1, k)] = u[I3(i, n, k)];
                v[I3(i, n + 1, k)] = v[I3(i, n, k)];
                w[I3(i, n + 1, k)] = w[I3(i, n, k)];
                p[I3(i, n + 1, k)] = po;
            }
        }
    for (int i = 0; i <= m + 1; ++i)                /* south :1123-1167 */
        for (int k = 0; k <= l + 1; ++k) {
            int wall = (south == 0) || (south != 0 && phi[I3(i, 1, k)] < 0.9);
            if (wall) {
                u[I3(i, 1, k)] = 0.; v[I3(i, 1, k)] = 0.; w[I3(i, 1, k)] = 0.;
                v[I3(i, 0, k)] = -v[I3(i, 2, k)];
                p[I3(i, 0, k)] = p[I3(i, 2, k)];
            } else if (south == 1) {
                u[I3(i, 1, k)] = U; v[I3(i, 1, k)] = 0.; w[I3(i, 1, k)] = 0.;
                u[I3(i, 0, k)] = u[I3(i, 1, k)];
                v[I3(i, 0, k)] = v[I3(i, 1, k)];
                w[I3(i, 0, k)] = w[I3(i, 1, k)];
                p[I3(i, 0, k)] = p[I3(i, 2, k)];
            } else {
                u[I3(i, 0, k)] = u[I3(i, 2, k)];
                v[I3(i, 0, k)] = v[I3(i, 2, k)];
                w[I3(i, 0, k)] = w[I3(i, 2, k)];
                p[I3(i, 0, k)] = po;
            }
        }
}

static void boundary_any(const pfo_case *c, const double *phi, double *u, double *v, double *w,
                         double *p) {
    if (c->dim == 2) boundary_2d(c, phi, u, v, p);
    else if (c->bc == PFO_BC_AIR) boundary_3d_air(c, phi, u, v, w, p);
    else boundary_3d_uniform(c, u, v, w, p);
}

/* initial_conditions: 2D ibm_2d_uniform_omp_cpu.f90:558-564; 3D ibm_3d_uniform_omp_cpu.f90:
 * 778-789 (note AoA/360 there); air ibm_3d_air_condition_omp_cpu.f90:1196-1206. */
void pfo_initial(const pfo_case *c, const double *phi, double *u, double *v, double *w,
                 double *p) {
    const int m = c->m, n = c->n, l = c->l, M2 = m + 2, N2 = n + 2;
    const size_t nt = full_size(c);
    memset(u, 0, nt * sizeof(double));
    memset(v, 0, nt * sizeof(double));
    memset(p, 0, nt * sizeof(double));
    if (w) memset(w, 0, nt * sizeof(double));
    if (c->dim == 2) {
        const double pai = atan(1.) * 4.;
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= m; ++i) {
                u[I2(i, j)] = c->inlet_velocity * cos(((c->aoa / 180.) * pai));
                v[I2(i, j)] = c->inlet_velocity * sin(((c->aoa / 180.) * pai));
                if (c->inlet_porosity) {   /* ibm_2d_backstep_omp_cpu.f90:615-618 */
                    u[I2(i, j)] = u[I2(i, j)] * phi[I2(i, j)];
                    v[I2(i, j)] = v[I2(i, j)] * phi[I2(i, j)];
                }
                p[I2(i, j)] = c->outlet_pressure;
            }
    } else {
        const double pi = atan(1.) * 4.;
        const int air = (c->bc == PFO_BC_AIR);
        for (int k = 1; k <= l; ++k)
            for (int j = 1; j <= n; ++j)
                for (int i = 1; i <= m; ++i) {
                    u[I3(i, j, k)] = air ? 0. : c->inlet_velocity * cos(((c->aoa / 360.) * pi));
                    v[I3(i, j, k)] = air ? 0. : c->inlet_velocity * sin(((c->aoa / 360.) * pi));
                    w[I3(i, j, k)] = 0.;
                    p[I3(i, j, k)] = c->outlet_pressure;
                }
    }
    boundary_any(c, phi, u, v, w, p);
}

/* ------------------------------------------------------------------------------------ */
/* 2D step: ibm_2d_uniform_omp_cpu.f90:84-118 (main loop body), solve_p :142-286,         */
/* solve_matrix_vec_omp :293-406, boundrary_matrix :410-454                              */
/* ------------------------------------------------------------------------------------ */
static double sor_2d(pfo_ctx *x, double *p) {
    const pfo_case *c = &x->c;
    const int m = c->m, n = c->n, M2 = m + 2;
    const double om = c->omega;
    double *p_old = x->p_old;
    const double *ap = x->ap, *ae = x->ae, *aw = x->aw, *an = x->an, *as = x->as, *bb = x->bb;
    double error = 0.0;
    for (int iter = 1; iter <= c->iter_max; ++iter) {
        for (int half = 0; half < 2; ++half) {
            for (int i = 1; i <= m; ++i) {          /* :322-327 / :356-361 */
                p[I2(i, 0)] = p[I2(i, n)];
                p[I2(i, n + 1)] = p[I2(i, 1)];
            }
#pragma omp parallel for
            for (int i = 0; i <= m + 1; ++i)        /* :329-335 (i outer, j inner) */
                for (int j = 0; j <= n + 1; ++j) p_old[I2(i, j)] = p[I2(i, j)];
            const int kstart = half == 0 ? 2 : 1;   /* :339 even space, :373 odd space */
            double err_h = 0.0;
#pragma omp parallel for reduction(max : err_h)
            for (int kk = kstart; kk <= m * n; kk += 2) {
                int j = (kk - 1) / m + 1;
                int i = kk - (j - 1) * m;
                if (m % 2 == 0 && j % 2 == 0) i = (half == 0) ? i - 1 : i + 1;   /* :344 / :378 */
                const size_t o = I2(i, j);
                p[o] = ((((((bb[o] - (ae[o] * p_old[I2(i + 1, j)])) - (aw[o] * p_old[I2(i - 1, j)])) -
                           (an[o] * p_old[I2(i, j + 1)])) - (as[o] * p_old[I2(i, j - 1)])) /
                         ap[o]) * om) + (p_old[o] * (1. - om));
                double d = fabs(p[o] - p_old[o]);
                err_h = fmaxf64(err_h, d);
            }
            error = fmaxf64(error, err_h);
        }
    }
    for (int i = 1; i <= m; ++i) {                  /* :393-396 */
        p[I2(i, 0)] = p[I2(i, n)];
        p[I2(i, n + 1)] = p[I2(i, 1)];
    }
    return error;
}

static double step_2d(pfo_ctx *x, const double *phi, double *u, double *v, double *p) {
    const pfo_case *c = &x->c;
    const int m = c->m, n = c->n, M2 = m + 2;
    const double dx = c->dx, dy = c->dy, dt = c->dt, xnue = c->xnue, xlambda = c->xlambda;
    const double density = c->density, thickness = c->thickness;
    double *uo = x->u_old, *vo = x->v_old, *div = x->div;
    double *ap = x->ap, *ae = x->ae, *aw = x->aw, *an = x->an, *as = x->as, *bb = x->bb;
    const double *po = phi;

#pragma omp parallel for
    for (int i = 0; i <= m + 1; ++i)                /* :89-94 */
        for (int j = 0; j <= n + 1; ++j) {
            uo[I2(i, j)] = u[I2(i, j)];
            vo[I2(i, j)] = v[I2(i, j)];
        }

#pragma omp parallel
    {
#pragma omp for
        for (int i = 1; i <= m; ++i)                /* div :173-178 (v-term divided by dx) */
            for (int j = 1; j <= n; ++j)
                div[I2(i, j)] = (((uo[I2(i + 1, j)] - uo[I2(i - 1, j)]) / dx) * .5) +
                                (((vo[I2(i, j + 1)] - vo[I2(i, j - 1)]) / dx) * .5);
#pragma omp for
        for (int j = 1; j <= n; ++j) { div[I2(0, j)] = 0.; div[I2(m + 1, j)] = 0.; }
#pragma omp for
        for (int i = 1; i <= m; ++i) { div[I2(i, 0)] = div[I2(i, n)]; div[I2(i, n + 1)] = div[I2(i, 1)]; }

#pragma omp for
        for (int i = 1; i <= m; ++i)                /* u predictor :201-226 */
            for (int j = 1; j <= n; ++j) {
                const size_t o = I2(i, j);
                const double uc = uo[o];
                double r = uc - (dt * (((uc * (uo[I2(i + 1, j)] - uo[I2(i - 1, j)])) / dx) / 2.));
                r = r - (dt * (((vo[o] * (uo[I2(i, j + 1)] - uo[I2(i, j - 1)])) / dy) / 2.));
                r = r + ((((dt * xnue) * ((uo[I2(i + 1, j)] - (2. * uc)) + uo[I2(i - 1, j)])) / dx) / dx);
                r = r + ((((dt * xnue) * ((uo[I2(i, j + 1)] - (2. * uc)) + uo[I2(i, j - 1)])) / dy) / dy);
                r = r + ((((dt * (xnue + xlambda)) * (div[I2(i + 1, j)] - div[I2(i - 1, j)])) / dx) * .5);
                {
                    const double t1 = (((((((uo[I2(i + 1, j)] - uo[I2(i - 1, j)]) / dx) * .5) +
                                          (((uo[I2(i + 1, j)] - uo[I2(i - 1, j)]) / dx) * .5)) * xnue) *
                                        (po[I2(i + 1, j)] - po[I2(i - 1, j)])) / dx) * .5;
                    const double t2 = (((((((uo[I2(i, j + 1)] - uo[I2(i, j - 1)]) / dy) * .5) +
                                          (((vo[I2(i + 1, j)] - vo[I2(i - 1, j)]) / dx) * .5)) * xnue) *
                                        (po[I2(i, j + 1)] - po[I2(i, j - 1)])) / dy) * .5;
                    const double t3 = (((div[o] * (po[I2(i + 1, j)] - po[I2(i - 1, j)])) / dx) * 0.5) * xlambda;
                    r = r + ((dt * ((t1 + t2) + t3)) / po[o]);
                }
                if (c->nonslip) {
                    const double th = thickness * dx;
                    r = r - (((((((dt * xnue) * uc) / (th * th)) * ALPHA) * po[o]) * (1. - po[o])) * (1. - po[o]));
                }
                u[o] = r;
            }
#pragma omp for
        for (int i = 1; i <= m; ++i)                /* v predictor :232-257 */
            for (int j = 1; j <= n; ++j) {
                const size_t o = I2(i, j);
                const double vc = vo[o];
                double r = vc - (dt * (((uo[o] * (vo[I2(i + 1, j)] - vo[I2(i - 1, j)])) / dx) / 2.));
                r = r - (dt * (((vc * (vo[I2(i, j + 1)] - vo[I2(i, j - 1)])) / dy) / 2.));
                r = r + ((((dt * xnue) * ((vo[I2(i + 1, j)] - (2. * vc)) + vo[I2(i - 1, j)])) / dx) / dx);
                r = r + ((((dt * xnue) * ((vo[I2(i, j + 1)] - (2. * vc)) + vo[I2(i, j - 1)])) / dy) / dy);
                r = r + ((((dt * (xnue + xlambda)) * (div[I2(i, j + 1)] - div[I2(i, j - 1)])) / dy) * .5);
                {
                    const double t1 = (((((((vo[I2(i + 1, j)] - vo[I2(i - 1, j)]) / dx) * .5) +
                                          (((uo[I2(i, j + 1)] - uo[I2(i, j - 1)]) / dy) * .5)) * xnue) *
                                        (po[I2(i + 1, j)] - po[I2(i - 1, j)])) / dx) * .5;
                    const double t2 = (((((((vo[I2(i, j + 1)] - vo[I2(i, j - 1)]) / dy) * .5) +
                                          (((vo[I2(i, j + 1)] - vo[I2(i, j - 1)]) / dy) * .5)) * xnue) *
                                        (po[I2(i, j + 1)] - po[I2(i, j - 1)])) / dy) * .5;
                    const double t3 = (((div[o] * (po[I2(i, j + 1)] - po[I2(i, j - 1)])) / dy) * 0.5) * xlambda;
                    r = r + ((dt * ((t1 + t2) + t3)) / po[o]);
                }
                if (c->nonslip) {
                    const double th = thickness * dx;   /* the 2D v wall force uses dx (:254) */
                    r = r - (((((((dt * xnue) * vc) / (th * th)) * ALPHA) * po[o]) * (1. - po[o])) * (1. - po[o]));
                }
                v[o] = r;
            }
#pragma omp for
        for (int i = 1; i <= m; ++i)                /* coefficients + RHS :263-277 */
            for (int j = 1; j <= n; ++j) {
                const size_t o = I2(i, j);
                ae[o] = ((dt * fmaxf64(SMALL, (po[I2(i + 1, j)] + po[o]) * 0.5)) / dx) / dx;
                aw[o] = ((dt * fmaxf64(SMALL, (po[o] + po[I2(i - 1, j)]) * 0.5)) / dx) / dx;
                an[o] = ((dt * fmaxf64(SMALL, (po[I2(i, j + 1)] + po[o]) * 0.5)) / dy) / dy;
                as[o] = ((dt * fmaxf64(SMALL, (po[o] + po[I2(i, j - 1)]) * 0.5)) / dy) / dy;
                ap[o] = (((-ae[o]) - aw[o]) - an[o]) - as[o];
                bb[o] = ((((((po[I2(i + 1, j)] * u[o]) + (po[o] * u[I2(i + 1, j)])) * 0.5) -
                           (((po[I2(i - 1, j)] * u[o]) + (po[o] * u[I2(i - 1, j)])) * 0.5)) * density) / dx) +
                        ((((((po[I2(i, j + 1)] * v[o]) + (po[o] * v[I2(i, j + 1)])) * 0.5) -
                           (((po[I2(i, j - 1)] * v[o]) + (po[o] * v[I2(i, j - 1)])) * 0.5)) * density) / dy);
            }
    }
    for (int j = 1; j <= n; ++j) {                  /* boundrary_matrix :424-436 */
        ae[I2(1, j)] = ae[I2(1, j)] + aw[I2(1, j)];
        aw[I2(1, j)] = 0.;
    }
    for (int j = 1; j <= n; ++j) {
        bb[I2(m, j)] = bb[I2(m, j)] + (ae[I2(m, j)] * p[I2(m + 1, j)]);
        ae[I2(m, j)] = 0.; aw[I2(m, j)] = 0.; an[I2(m, j)] = 0.; as[I2(m, j)] = 0.;
    }
    const double err = sor_2d(x, p);
#pragma omp parallel for
    for (int j = 1; j <= n; ++j)                    /* projection :108-113 */
        for (int i = 1; i <= m; ++i) {
            const size_t o = I2(i, j);
            u[o] = u[o] - ((((dt / density) * (p[I2(i + 1, j)] - p[I2(i - 1, j)])) / dx) * 0.5);
            v[o] = v[o] - ((((dt / density) * (p[I2(i, j + 1)] - p[I2(i, j - 1)])) / dy) * 0.5);
        }
    boundary_2d(c, phi, u, v, p);
    return err;
}

/* ------------------------------------------------------------------------------------ */
/* 3D step: ibm_3d_uniform_omp_cpu.f90:85-128 (main loop body), solve_p :151-426,        */
/* solve_matrix_vec_omp :433-614, boundrary_matrix :618-663; the air-conditioning        */
/* variant ibm_3d_air_condition_omp_cpu.f90: div ghosts :210-237, SOR :480-661 (no       */
/* periodic refresh), boundary_matrix :665-867.                                          */
/* ------------------------------------------------------------------------------------ */
static void sor3_refresh(const pfo_case *c, double *p) {
    const int m = c->m, n = c->n, l = c->l, M2 = m + 2, N2 = n + 2;
#pragma omp parallel
    {
#pragma omp for
        for (int i = 1; i <= m; ++i)                /* :463-470 */
            for (int k = 1; k <= l; ++k) {
                p[I3(i, 0, k)] = p[I3(i, n, k)];
                p[I3(i, n + 1, k)] = p[I3(i, 1, k)];
            }
#pragma omp for
        for (int i = 1; i <= m; ++i)                /* :473-480 */
            for (int j = 1; j <= n; ++j) {
                p[I3(i, j, 0)] = p[I3(i, j, l)];
                p[I3(i, j, l + 1)] = p[I3(i, j, 1)];
            }
    }
}

static double sor_3d(pfo_ctx *x, double *p) {
    const pfo_case *c = &x->c;
    const int m = c->m, n = c->n, l = c->l, M2 = m + 2, N2 = n + 2;
    const int periodic = (c->bc != PFO_BC_AIR);
    const double om = c->omega;
    double *p_old = x->p_old;
    const double *ap = x->ap, *ae = x->ae, *aw = x->aw, *an = x->an, *as = x->as;
    const double *at = x->at, *ab = x->ab, *bb = x->bb;
    const long mnl = (long)m * n * l;
    double error = 0.0;
    for (int iter = 1; iter <= c->iter_max; ++iter) {
        for (int half = 0; half < 2; ++half) {
            if (periodic) sor3_refresh(c, p);
#pragma omp parallel for
            for (int i = 0; i <= m + 1; ++i)        /* :482-490 (i outer, k inner) */
                for (int j = 0; j <= n + 1; ++j)
                    for (int k = 0; k <= l + 1; ++k) p_old[I3(i, j, k)] = p[I3(i, j, k)];
            const long start = (half == 0) ? 2 : 1; /* :493 even space, :550 odd space */
#pragma omp parallel for
            for (long ii = start; ii <= mnl; ii += 2) {
                int k = (int)((ii - 1) / ((long)m * n)) + 1;
                int j = (int)(((ii - 1) / m + 1) - (long)(k - 1) * n);
                int i = (int)((ii - (long)(j - 1) * m) - (long)(k - 1) * m * n);
                if (m % 2 != 0 && n % 2 == 0 && k % 2 == 0) {          /* :499-508 / :556-565 */
                    if (i % 2 != 0) j = (half == 0) ? j - 1 : j + 1;
                    else j = (half == 0) ? j + 1 : j - 1;
                } else if (m % 2 == 0 && (j % 2 + k % 2 == 1)) {
                    i = (half == 0) ? i - 1 : i + 1;
                }
                const size_t o = I3(i, j, k);
                p[o] = ((((((((bb[o] - (ae[o] * p_old[I3(i + 1, j, k)])) - (aw[o] * p_old[I3(i - 1, j, k)])) -
                             (an[o] * p_old[I3(i, j + 1, k)])) - (as[o] * p_old[I3(i, j - 1, k)])) -
                           (at[o] * p_old[I3(i, j, k + 1)])) - (ab[o] * p_old[I3(i, j, k - 1)])) / ap[o]) * om) +
                       (p_old[o] * (1. - om));
            }
        }
        double err_i = 0.0;                          /* :575-583, after the odd sweep only */
#pragma omp parallel for reduction(max : err_i)
        for (int i = 1; i <= m; ++i)
            for (int j = 1; j <= n; ++j)
                for (int k = 1; k <= l; ++k) {
                    double d = fabs(p[I3(i, j, k)] - p_old[I3(i, j, k)]);
                    err_i = fmaxf64(err_i, d);
                }
        error = fmaxf64(error, err_i);
    }
    if (periodic) sor3_refresh(c, p);              /* :588-605 */
    return error;
}

static void boundary_matrix_air(pfo_ctx *x, const double *phi, const double *p) {
    const pfo_case *c = &x->c;
    const int m = c->m, n = c->n, l = c->l, M2 = m + 2, N2 = n + 2;
    double *ae = x->ae, *aw = x->aw, *an = x->an, *as = x->as, *at = x->at, *ab = x->ab, *bb = x->bb;
    const int top = c->face[0], bottom = c->face[1], east = c->face[2], west = c->face[3],
              north = c->face[4], south = c->face[5];
#define ZERO6(o) do { ae[o] = 0.; aw[o] = 0.; an[o] = 0.; as[o] = 0.; at[o] = 0.; ab[o] = 0.; } while (0)
    for (int i = 0; i <= m + 1; ++i)                /* top :686-714 */
        for (int j = 0; j <= n + 1; ++j) {
            const size_t o = I3(i, j, l);
            if (top == 0 || top == 1 || phi[o] < 0.9) { ab[o] = ab[o] + at[o]; at[o] = 0.; }
            else if (top == 2) {   /* :702 reads bb(i,j,1) as written */
                bb[o] = bb[I3(i, j, 1)] + (at[o] * p[I3(i, j, l + 1)]);
                ZERO6(o);
            }
        }
    for (int i = 0; i <= m + 1; ++i)                /* bottom :717-745 */
        for (int j = 0; j <= n + 1; ++j) {
            const size_t o = I3(i, j, 1);
            if (bottom == 0 || bottom == 1 || phi[o] < 0.9) { at[o] = at[o] + ab[o]; ab[o] = 0.; }
            else { bb[o] = bb[o] + (ab[o] * p[I3(i, j, 0)]); ZERO6(o); }
        }
    for (int j = 0; j <= n + 1; ++j)                /* east :748-776 */
        for (int k = 0; k <= l + 1; ++k) {
            const size_t o = I3(m, j, k);
            if (east == 0 || east == 1 || phi[o] < 0.9) { aw[o] = aw[o] + ae[o]; ae[o] = 0.; }
            else { bb[o] = bb[o] + (ae[o] * p[I3(m + 1, j, k)]); ZERO6(o); }
        }
    for (int j = 0; j <= n + 1; ++j)                /* west :779-807 */
        for (int k = 0; k <= l + 1; ++k) {
            const size_t o = I3(1, j, k);
            if (west == 0 || west == 1 || phi[o] < 0.9) { ae[o] = ae[o] + aw[o]; aw[o] = 0.; }
            else { bb[o] = bb[o] + (aw[o] * p[I3(0, j, k)]); ZERO6(o); }
        }
    for (int i = 0; i <= m + 1; ++i)                /* north :810-837 */
        for (int k = 0; k <= l + 1; ++k) {
            const size_t o = I3(i, n, k);
            if (north == 0 || north == 1 || phi[o] < 0.9) { as[o] = as[o] + an[o]; an[o] = 0.; }
            else { bb[o] = bb[o] + (an[o] * p[I3(i, n + 1, k)]); ZERO6(o); }
        }
    for (int i = 0; i <= m + 1; ++i)                /* south :840-864 */
        for (int k = 0; k <= l + 1; ++k) {
            const size_t o = I3(i, 1, k);
            if (south == 0 || south == 1 || phi[o] < 0.9) { an[o] = an[o] + as[o]; as[o] = 0.; }
            else { bb[o] = bb[o] + (as[o] * p[I3(i, 0, k)]); ZERO6(o); }
        }
#undef ZERO6
}

static double step_3d(pfo_ctx *x, const double *phi, double *u, double *v, double *w, double *p) {
    const pfo_case *c = &x->c;
    const int m = c->m, n = c->n, l = c->l, M2 = m + 2, N2 = n + 2;
    const int air = (c->bc == PFO_BC_AIR);
    const double dx = c->dx, dy = c->dy, dz = c->dz, dt = c->dt, xnue = c->xnue, xlambda = c->xlambda;
    const double density = c->density, thickness = c->thickness;
    double *uo = x->u_old, *vo = x->v_old, *wo = x->w_old, *div = x->div;
    double *ap = x->ap, *ae = x->ae, *aw = x->aw, *an = x->an, *as = x->as;
    double *at = x->at, *ab = x->ab, *bb = x->bb;
    const double *po = phi;

#pragma omp parallel for
    for (int k = 0; k <= l + 1; ++k)                /* main :90-98 */
        for (int j = 0; j <= n + 1; ++j)
            for (int i = 0; i <= m + 1; ++i) {
                const size_t o = I3(i, j, k);
                uo[o] = u[o]; vo[o] = v[o]; wo[o] = w[o];
            }

#pragma omp parallel
    {
#pragma omp for
        for (int k = 1; k <= l; ++k)                /* div :186-194 */
            for (int j = 1; j <= n; ++j)
                for (int i = 1; i <= m; ++i)
                    div[I3(i, j, k)] = ((((uo[I3(i + 1, j, k)] - uo[I3(i - 1, j, k)]) / dx) * 0.5) +
                                        (((vo[I3(i, j + 1, k)] - vo[I3(i, j - 1, k)]) / dy) * 0.5)) +
                                       (((wo[I3(i, j, k + 1)] - wo[I3(i, j, k - 1)]) / dz) * 0.5);
        if (!air) {
#pragma omp for
            for (int k = 1; k <= l; ++k)            /* :198-204 */
                for (int j = 1; j <= n; ++j) { div[I3(0, j, k)] = 0.; div[I3(m + 1, j, k)] = 0.; }
#pragma omp for
            for (int k = 1; k <= l; ++k)            /* :207-213 */
                for (int i = 1; i <= m; ++i) { div[I3(i, 0, k)] = div[I3(i, n, k)]; div[I3(i, n + 1, k)] = div[I3(i, 1, k)]; }
#pragma omp for
            for (int j = 1; j <= n; ++j)            /* :216-222 */
                for (int i = 1; i <= m; ++i) { div[I3(i, j, 0)] = div[I3(i, j, l)]; div[I3(i, j, l + 1)] = div[I3(i, j, 1)]; }
        } else {
#pragma omp for
            for (int i = 1; i <= m; ++i)            /* air :211-217 */
                for (int j = 1; j <= n; ++j) { div[I3(i, j, 0)] = 0.; div[I3(i, j, l + 1)] = 0.; }
#pragma omp for
            for (int j = 1; j <= n; ++j)            /* air :221-227 */
                for (int k = 1; k <= l; ++k) { div[I3(0, j, k)] = 0.; div[I3(m + 1, j, k)] = 0.; }
#pragma omp for
            for (int i = 1; i <= m; ++i)            /* air :230-236 */
                for (int k = 1; k <= l; ++k) { div[I3(i, 0, k)] = 0.; div[I3(i, n + 1, k)] = 0.; }
        }

#pragma omp for
        for (int k = 1; k <= l; ++k)                /* u predictor :229-275 */
            for (int j = 1; j <= n; ++j)
                for (int i = 1; i <= m; ++i) {
                    const size_t o = I3(i, j, k);
                    const double uc = uo[o];
                    const double uE = uo[I3(i + 1, j, k)], uW = uo[I3(i - 1, j, k)];
                    const double uN = uo[I3(i, j + 1, k)], uS = uo[I3(i, j - 1, k)];
                    const double uT = uo[I3(i, j, k + 1)], uB = uo[I3(i, j, k - 1)];
                    double r = uc - ((((dt * uc) * (uE - uW)) / dx) * 0.5);
                    r = r - ((((dt * vo[o]) * (uN - uS)) / dy) * 0.5);
                    r = r - ((((dt * wo[o]) * (uT - uB)) / dz) * 0.5);
                    r = r + ((((dt * xnue) * ((uE - (2. * uc)) + uW)) / dx) / dx);
                    r = r + ((((dt * xnue) * ((uN - (2. * uc)) + uS)) / dy) / dy);
                    r = r + ((((dt * xnue) * ((uT - (2. * uc)) + uB)) / dz) / dz);
                    r = r + ((((dt * (xnue + xlambda)) * (div[I3(i + 1, j, k)] - div[I3(i - 1, j, k)])) / dx) * 0.5);
                    {
                        const double t1 = ((((((((uE - uW) / dx) * 0.5) + (((uE - uW) / dx) * 0.5)) * xnue) *
                                             (po[I3(i + 1, j, k)] - po[I3(i - 1, j, k)])) / dx) * 0.5);
                        const double t2 = ((((((((uN - uS) / dy) * 0.5) +
                                                (((vo[I3(i + 1, j, k)] - vo[I3(i - 1, j, k)]) / dx) * 0.5)) * xnue) *
                                              (po[I3(i, j + 1, k)] - po[I3(i, j - 1, k)])) / dy) * 0.5);
                        const double t3 = ((((((((uT - uB) / dz) * 0.5) +
                                                (((wo[I3(i + 1, j, k)] - wo[I3(i - 1, j, k)]) / dx) * 0.5)) * xnue) *
                                              (po[I3(i, j, k + 1)] - po[I3(i, j, k - 1)])) / dz) * 0.5);
                        const double t4 = ((((div[o] * (po[I3(i + 1, j, k)] - po[I3(i - 1, j, k)])) / dx) * 0.5) * xlambda);
                        r = r + ((dt * (((t1 + t2) + t3) + t4)) / po[o]);
                    }
                    if (c->nonslip) {
                        const double th = thickness * dx;
                        r = r - (((((((dt * xnue) * uc) / (th * th)) * ALPHA) * po[o]) * (1. - po[o])) * (1. - po[o]));
                    }
                    u[o] = r;
                }
#pragma omp for
        for (int k = 1; k <= l; ++k)                /* v predictor :282-327 */
            for (int j = 1; j <= n; ++j)
                for (int i = 1; i <= m; ++i) {
                    const size_t o = I3(i, j, k);
                    const double vc = vo[o];
                    const double vE = vo[I3(i + 1, j, k)], vW = vo[I3(i - 1, j, k)];
                    const double vN = vo[I3(i, j + 1, k)], vS = vo[I3(i, j - 1, k)];
                    const double vT = vo[I3(i, j, k + 1)], vB = vo[I3(i, j, k - 1)];
                    double r = vc - ((((dt * uo[o]) * (vE - vW)) / dx) * 0.5);
                    r = r - ((((dt * vc) * (vN - vS)) / dy) * 0.5);
                    r = r - ((((dt * wo[o]) * (vT - vB)) / dz) * 0.5);
                    r = r + ((((dt * xnue) * ((vE - (2. * vc)) + vW)) / dx) / dx);
                    r = r + ((((dt * xnue) * ((vN - (2. * vc)) + vS)) / dy) / dy);
                    r = r + ((((dt * xnue) * ((vT - (2. * vc)) + vB)) / dz) / dz);
                    r = r + ((((dt * (xnue + xlambda)) * (div[I3(i, j + 1, k)] - div[I3(i, j - 1, k)])) / dy) * 0.5);
                    {
                        const double t1 = ((((((((vE - vW) / dx) * 0.5) +
                                                (((uo[I3(i, j + 1, k)] - uo[I3(i, j - 1, k)]) / dy) * 0.5)) * xnue) *
                                              (po[I3(i + 1, j, k)] - po[I3(i - 1, j, k)])) / dx) * 0.5);
                        const double t2 = ((((((((vN - vS) / dy) * .5) + (((vN - vS) / dy) * 0.5)) * xnue) *
                                             (po[I3(i, j + 1, k)] - po[I3(i, j - 1, k)])) / dy) * 0.5);
                        const double t3 = ((((((((vT - vB) / dz) * .5) +
                                                (((wo[I3(i, j + 1, k)] - wo[I3(i, j - 1, k)]) / dy) * 0.5)) * xnue) *
                                              (po[I3(i, j, k + 1)] - po[I3(i, j, k - 1)])) / dz) * 0.5);
                        const double t4 = ((((div[o] * (po[I3(i, j + 1, k)] - po[I3(i, j - 1, k)])) / dy) * 0.5) * xlambda);
                        r = r + ((dt * (((t1 + t2) + t3) + t4)) / po[o]);
                    }
                    if (c->nonslip) {
                        const double th = thickness * dy;
                        r = r - (((((((dt * xnue) * vc) / (th * th)) * ALPHA) * po[o]) * (1. - po[o])) * (1. - po[o]));
                    }
                    v[o] = r;
                }
#pragma omp for
        for (int k = 1; k <= l; ++k)                /* w predictor :335-380 */
            for (int j = 1; j <= n; ++j)
                for (int i = 1; i <= m; ++i) {
                    const size_t o = I3(i, j, k);
                    const double wc = wo[o];
                    const double wE = wo[I3(i + 1, j, k)], wW = wo[I3(i - 1, j, k)];
                    const double wN = wo[I3(i, j + 1, k)], wS = wo[I3(i, j - 1, k)];
                    const double wT = wo[I3(i, j, k + 1)], wB = wo[I3(i, j, k - 1)];
                    double r = wc - ((((dt * uo[o]) * (wE - wW)) / dx) * 0.5);
                    r = r - ((((dt * vo[o]) * (wN - wS)) / dy) * 0.5);
                    r = r - ((((dt * wc) * (wT - wB)) / dz) * 0.5);
                    r = r + ((((dt * xnue) * ((wE - (2. * wc)) + wW)) / dx) / dx);
                    r = r + ((((dt * xnue) * ((wN - (2. * wc)) + wS)) / dy) / dy);
                    r = r + ((((dt * xnue) * ((wT - (2. * wc)) + wB)) / dz) / dz);
                    r = r + ((((dt * (xnue + xlambda)) * (div[I3(i, j, k + 1)] - div[I3(i, j, k - 1)])) / dz) * 0.5);
                    {
                        const double t1 = ((((((((wE - wW) / dx) * 0.5) +
                                                (((uo[I3(i, j, k + 1)] - uo[I3(i, j, k - 1)]) / dz) * 0.5)) * xnue) *
                                              (po[I3(i + 1, j, k)] - po[I3(i - 1, j, k)])) / dx) * 0.5);
                        const double t2 = ((((((((wN - wS) / dy) * 0.5) +
                                                (((vo[I3(i, j, k + 1)] - vo[I3(i, j, k - 1)]) / dz) * 0.5)) * xnue) *
                                              (po[I3(i, j + 1, k)] - po[I3(i, j - 1, k)])) / dy) * 0.5);
                        const double t3 = ((((((((wT - wB) / dz) * 0.5) + (((wT - wB) / dz) * 0.5)) * xnue) *
                                             (po[I3(i, j, k + 1)] - po[I3(i, j, k - 1)])) / dz) * 0.5);
                        const double t4 = ((((div[o] * (po[I3(i, j, k + 1)] - po[I3(i, j, k - 1)])) / dz) * 0.5) * xlambda);
                        r = r + ((dt * (((t1 + t2) + t3) + t4)) / po[o]);
                    }
                    if (c->nonslip) {
                        const double th = thickness * dz;
                        r = r - (((((((dt * xnue) * wc) / (th * th)) * ALPHA) * po[o]) * (1. - po[o])) * (1. - po[o]));
                    }
                    w[o] = r;
                }
#pragma omp for
        for (int k = 1; k <= l; ++k)                /* coefficients + RHS :387-413 */
            for (int j = 1; j <= n; ++j)
                for (int i = 1; i <= m; ++i) {
                    const size_t o = I3(i, j, k);
                    ae[o] = ((dt * fmaxf64(SMALL, (po[I3(i + 1, j, k)] + po[o]) * 0.5)) / dx) / dx;
                    aw[o] = ((dt * fmaxf64(SMALL, (po[o] + po[I3(i - 1, j, k)]) * 0.5)) / dx) / dx;
                    an[o] = ((dt * fmaxf64(SMALL, (po[I3(i, j + 1, k)] + po[o]) * 0.5)) / dy) / dy;
                    as[o] = ((dt * fmaxf64(SMALL, (po[o] + po[I3(i, j - 1, k)]) * 0.5)) / dy) / dy;
                    at[o] = ((dt * fmaxf64(SMALL, (po[I3(i, j, k + 1)] + po[o]) * 0.5)) / dz) / dz;
                    ab[o] = ((dt * fmaxf64(SMALL, (po[o] + po[I3(i, j, k - 1)]) * 0.5)) / dz) / dz;
                    ap[o] = (((((-ae[o]) - aw[o]) - an[o]) - as[o]) - at[o]) - ab[o];
                    bb[o] = (((((((po[I3(i + 1, j, k)] * u[o]) + (po[o] * u[I3(i + 1, j, k)])) * 0.5) -
                                (((po[I3(i - 1, j, k)] * u[o]) + (po[o] * u[I3(i - 1, j, k)])) * 0.5)) * density) / dx) +
                             ((((((po[I3(i, j + 1, k)] * v[o]) + (po[o] * v[I3(i, j + 1, k)])) * 0.5) -
                                (((po[I3(i, j - 1, k)] * v[o]) + (po[o] * v[I3(i, j - 1, k)])) * 0.5)) * density) / dy)) +
                            ((((((po[I3(i, j, k + 1)] * w[o]) + (po[o] * w[I3(i, j, k + 1)])) * 0.5) -
                               (((po[I3(i, j, k - 1)] * w[o]) + (po[o] * w[I3(i, j, k - 1)])) * 0.5)) * density) / dz);
                }
    }
    if (air) {
        boundary_matrix_air(x, phi, p);
    } else {
        for (int k = 1; k <= l; ++k)                /* boundrary_matrix :638-659 */
            for (int j = 1; j <= n; ++j) {
                ae[I3(1, j, k)] = ae[I3(1, j, k)] + aw[I3(1, j, k)];
                aw[I3(1, j, k)] = 0.;
            }
        for (int k = 1; k <= l; ++k)
            for (int j = 1; j <= n; ++j) {
                const size_t o = I3(m, j, k);
                bb[o] = bb[o] + (ae[o] * p[I3(m + 1, j, k)]);
                ae[o] = 0.; aw[o] = 0.; an[o] = 0.; as[o] = 0.; at[o] = 0.; ab[o] = 0.;
            }
    }
    const double err = sor_3d(x, p);
#pragma omp parallel for
    for (int k = 1; k <= l; ++k)                    /* projection main :115-123 */
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= m; ++i) {
                const size_t o = I3(i, j, k);
                u[o] = u[o] - ((((dt / density) * (p[I3(i + 1, j, k)] - p[I3(i - 1, j, k)])) / dx) * 0.5);
                v[o] = v[o] - ((((dt / density) * (p[I3(i, j + 1, k)] - p[I3(i, j - 1, k)])) / dy) * 0.5);
                w[o] = w[o] - ((((dt / density) * (p[I3(i, j, k + 1)] - p[I3(i, j, k - 1)])) / dz) * 0.5);
            }
    if (air) boundary_3d_air(c, phi, u, v, w, p);
    else boundary_3d_uniform(c, u, v, w, p);
    return err;
}

int pfo_step(pfo_ctx *x, const double *phi, double *u, double *v, double *w, double *p,
             int nsteps, double *err) {
    if (!x || !phi || !u || !v || !p || (x->c.dim == 3 && !w) || nsteps < 0) return -1;
    for (int s = 0; s < nsteps; ++s) {
        double e = (x->c.dim == 2) ? step_2d(x, phi, u, v, p) : step_3d(x, phi, u, v, w, p);
        if (err) err[s] = e;
    }
    return 0;
}

/* output_divergent_2d (src/omp_parallel/lib/output.f90:217-224),
 * output_divergent_3d (output.f90:929-940). */
void pfo_divergence(const pfo_case *c, const double *phi, const double *u, const double *v,
                    const double *w, double *out) {
    const int m = c->m, n = c->n, l = c->l, M2 = m + 2, N2 = n + 2;
    const double dx = c->dx, dy = c->dy, dz = c->dz;
    const double *po = phi;
    if (c->dim == 2) {
        for (int i = 1; i <= m; ++i)
            for (int j = 1; j <= n; ++j) {
                const size_t o = I2(i, j);
                out[(size_t)(i - 1) + (size_t)m * (j - 1)] =
                    (((((po[I2(i + 1, j)] * u[o]) + (po[o] * u[I2(i + 1, j)])) / 2.) -
                      (((po[I2(i - 1, j)] * u[o]) + (po[o] * u[I2(i - 1, j)])) / 2.)) / dx) +
                    (((((po[I2(i, j + 1)] * v[o]) + (po[o] * v[I2(i, j + 1)])) / 2.) -
                      (((po[I2(i, j - 1)] * v[o]) + (po[o] * v[I2(i, j - 1)])) / 2.)) / dy);
            }
        return;
    }
#pragma omp parallel for
    for (int k = 1; k <= l; ++k)
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= m; ++i) {
                const size_t o = I3(i, j, k);
                out[(size_t)(i - 1) + (size_t)m * ((size_t)(j - 1) + (size_t)n * (k - 1))] =
                    ((((((po[I3(i + 1, j, k)] * u[o]) + (po[o] * u[I3(i + 1, j, k)])) / 2.) -
                       (((po[I3(i - 1, j, k)] * u[o]) + (po[o] * u[I3(i - 1, j, k)])) / 2.)) / dx) +
                     (((((po[I3(i, j + 1, k)] * v[o]) + (po[o] * v[I3(i, j + 1, k)])) / 2.) -
                       (((po[I3(i, j - 1, k)] * v[o]) + (po[o] * v[I3(i, j - 1, k)])) / 2.)) / dy)) +
                    (((((po[I3(i, j, k + 1)] * w[o]) + (po[o] * w[I3(i, j, k + 1)])) / 2.) -
                      (((po[I3(i, j, k - 1)] * w[o]) + (po[o] * w[I3(i, j, k - 1)])) / 2.)) / dz);
            }
}

/* output_force_log_2d (src/omp_parallel/lib/output.f90:244-305, loop :264-284, i outer)
 * and output_force_log_3d (:1090-1165, loop :1118-1145, k outer): the pressure and the
 * wall-force (viscous) contributions of every interior cell, each term in the Fortran
 * evaluation order, summed serially in loop order.  x**2 is x*x; unary minus applies to
 * the whole product. */
void pfo_force(const pfo_case *c, const double *phi, const double *u, const double *v,
               const double *w, const double *p, double *out) {
    const int m = c->m, n = c->n, l = c->l, M2 = m + 2, N2 = n + 2;
    const double dx = c->dx, dy = c->dy, dz = c->dz, th = c->thickness;
    const double small = 1.e-6, alpha = 32.0, rho = c->density, nu = c->xnue;
    double fpx = 0., fpy = 0., fpz = 0., fvx = 0., fvy = 0., fvz = 0.;
    if (c->dim == 2) {
        for (int i = 1; i <= m; ++i)
            for (int j = 1; j <= n; ++j) {
                const size_t o = I2(i, j);
                const double gx = (phi[I2(i + 1, j)] - phi[I2(i - 1, j)]) * 0.5;
                const double gy = (phi[I2(i, j + 1)] - phi[I2(i, j - 1)]) * 0.5;
                const double na = sqrt((gx * gx) + (gy * gy));
                const double den = na > small ? na : small;
                const double nx = gx / den, ny = gy / den;
                const double ph = phi[o], q = 1.0 - ph;
                const double dpx = -((((((dx * dy) * p[o]) * 2) * ph) * q) / (th * dx) * nx);
                const double dpy = -((((((dx * dy) * p[o]) * 2) * ph) * q) / (th * dy) * ny);
                const double tx = (ph * q) / (th * dx), ty = (ph * q) / (th * dy);
                const double dvx = (((((dx * dy) * alpha) * rho) * nu) * (tx * tx)) * u[o];
                const double dvy = (((((dx * dy) * alpha) * rho) * nu) * (ty * ty)) * v[o];
                fpx = fpx + dpx;
                fpy = fpy + dpy;
                fvx = fvx + dvx;
                fvy = fvy + dvy;
            }
    } else {
        for (int k = 1; k <= l; ++k)
            for (int j = 1; j <= n; ++j)
                for (int i = 1; i <= m; ++i) {
                    const size_t o = I3(i, j, k);
                    const double gx = (phi[I3(i + 1, j, k)] - phi[I3(i - 1, j, k)]) * 0.5;
                    const double gy = (phi[I3(i, j + 1, k)] - phi[I3(i, j - 1, k)]) * 0.5;
                    const double gz = (phi[I3(i, j, k + 1)] - phi[I3(i, j, k - 1)]) * 0.5;
                    const double na = sqrt(((gx * gx) + (gy * gy)) + (gz * gz));
                    const double den = na > small ? na : small;
                    const double nx = gx / den, ny = gy / den, nz = gz / den;
                    const double ph = phi[o], q = 1.0 - ph;
                    const double base = (((((dx * dy) * dz) * p[o]) * 2) * ph) * q;
                    const double dpx = -(base / (th * dx) * nx);
                    const double dpy = -(base / (th * dy) * ny);
                    const double dpz = -(base / (th * dz) * nz);
                    const double tx = (ph * q) / (th * dx), ty = (ph * q) / (th * dy), tz = (ph * q) / (th * dz);
                    const double cv = ((((dx * dy) * dz) * alpha) * rho) * nu;
                    fpx = fpx + dpx;
                    fpy = fpy + dpy;
                    fpz = fpz + dpz;
                    fvx = fvx + (cv * (tx * tx)) * u[o];
                    fvy = fvy + (cv * (ty * ty)) * v[o];
                    fvz = fvz + (cv * (tz * tz)) * w[o];
                }
    }
    out[0] = fpx; out[1] = fpy; out[2] = fpz;
    out[3] = fvx; out[4] = fvy; out[5] = fvz;
}

/* Brute force over the linear-index decode of solve_matrix_vec_omp (3D :493-508; 2D
 * :339-344): visits[cell] counts first-sweep visits (+1) and second-sweep visits (+16). */
int pfo_decode_first_sweep(int dim, int m, int n, int l, int *visits) {
    if (dim == 2) {
        memset(visits, 0, sizeof(int) * (size_t)m * n);
        for (int half = 0; half < 2; ++half)
            for (int kk = half == 0 ? 2 : 1; kk <= m * n; kk += 2) {
                int j = (kk - 1) / m + 1, i = kk - (j - 1) * m;
                if (m % 2 == 0 && j % 2 == 0) i = (half == 0) ? i - 1 : i + 1;
                if (i < 1 || i > m || j < 1 || j > n) return -1;
                visits[(i - 1) + m * (j - 1)] += half == 0 ? 1 : 16;
            }
        return 0;
    }
    memset(visits, 0, sizeof(int) * (size_t)m * n * l);
    const long mnl = (long)m * n * l;
    for (int half = 0; half < 2; ++half)
        for (long ii = half == 0 ? 2 : 1; ii <= mnl; ii += 2) {
            int k = (int)((ii - 1) / ((long)m * n)) + 1;
            int j = (int)(((ii - 1) / m + 1) - (long)(k - 1) * n);
            int i = (int)((ii - (long)(j - 1) * m) - (long)(k - 1) * m * n);
            if (m % 2 != 0 && n % 2 == 0 && k % 2 == 0) {
                if (i % 2 != 0) j = (half == 0) ? j - 1 : j + 1;
                else j = (half == 0) ? j + 1 : j - 1;
            } else if (m % 2 == 0 && (j % 2 + k % 2 == 1)) {
                i = (half == 0) ? i - 1 : i + 1;
            }
            if (i < 1 || i > m || j < 1 || j > n || k < 1 || k > l) return -1;
            visits[(i - 1) + (long)m * ((j - 1) + (long)n * (k - 1))] += half == 0 ? 1 : 16;
        }
    return 0;
}
