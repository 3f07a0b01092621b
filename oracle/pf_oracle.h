/*
 * pf_oracle.h -- CPU restatement of PixelFlow's time step (TEST INFRASTRUCTURE ONLY).
 *
 * This library gives the results the HIP path is compared with, and the CPU baseline timed
 * beside it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path never does. Every function follows the reference's Fortran line by line
 * in arithmetic order (file:line cited at each definition in pf_oracle.c), compiled with
 * -ffp-contract=off so that results match the reference built with -fdefault-real-8.
 *
 * Array convention: every field is a full Fortran-ordered array with ghost layers,
 * index (i, j, k) with 0 <= i <= m+1, 0 <= j <= n+1, 0 <= k <= l+1 at offset
 * i + (m+2)*(j + (n+2)*k).  In 2D the k extent is absent (size (m+2)*(n+2)).
 */
#ifndef PF_ORACLE_H
#define PF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { PFO_BC_UNIFORM = 0, PFO_BC_AIR = 1 };

typedef struct pfo_case {
    int dim;             /* 2 or 3 */
    int m, n, l;         /* interior extents; l = 1 in 2D */
    int bc;              /* PFO_BC_UNIFORM or PFO_BC_AIR (3D only) */
    int nonslip;         /* 0/1: wall-force term of the predictor */
    int iter_max;        /* fixed SOR iterations per step */
    int face[6];         /* air: top, bottom, east, west, north, south codes (0 wall, 1 inlet, 2 outlet) */
    double dx, dy, dz, dt;
    double xnue, xlambda, density, thickness;
    double inlet_velocity, outlet_pressure, aoa;
    double omega;        /* relux_factor */
    int inlet_porosity;  /* 2D backward-step variant: inlet and initial velocity times porosity */
} pfo_case;

typedef struct pfo_ctx pfo_ctx;

pfo_ctx *pfo_create(const pfo_case *c);
void pfo_destroy(pfo_ctx *x);

/* grid_conditions*: clamp raw interior porosity (i fastest) to threshold, fill ghosts. */
void pfo_setup_porosity(const pfo_case *c, const double *raw_interior, double threshold,
                        double *phi);
/* initial_conditions followed by boundary (the state before the first step). */
void pfo_initial(const pfo_case *c, const double *phi, double *u, double *v, double *w,
                 double *p);
/* nsteps time steps in place; err[s] receives the SOR diagnostic of step s (may be NULL). */
int pfo_step(pfo_ctx *x, const double *phi, double *u, double *v, double *w, double *p,
             int nsteps, double *err);
/* output_divergent_*: porosity-weighted flux divergence, interior only, i fastest. */
void pfo_divergence(const pfo_case *c, const double *phi, const double *u, const double *v,
                    const double *w, double *div_interior);
/* output_force_log_2d/3d: out[0..5] = pressure force x, y, z, viscous force x, y, z
 * (2D: z entries 0), summed serially in the reference's loop order. */
void pfo_force(const pfo_case *c, const double *phi, const double *u, const double *v,
               const double *w, const double *p, double *out);
/* Colour the reference's linear-index decode assigns to interior cell (i,j[,k]) in the
 * first sweep: returns 1 if the first sweep updates it, 0 if the second does, -1 if the
 * decode visits it in neither or twice (never for valid extents). Brute force helper. */
int pfo_decode_first_sweep(int dim, int m, int n, int l, int *visits);

#ifdef __cplusplus
}
#endif
#endif
