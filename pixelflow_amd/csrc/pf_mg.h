// pf_mg.h -- multigrid-preconditioned BiCGStab for the pressure correction A e = r.
//
// The reference solves A p = bb with a fixed number of red/black SOR iterations and has no
// convergence exit (ibm_3d_uniform_omp_cpu.f90:460-611).  SOR at omega = 1.7 contracts
// the smooth error by about 1 - 2e-4 per iteration on a 512-cell extent, so a residual
// tolerance of 1e-8 is out of its reach on the large grids.  This solver reaches it: the
// FP64 outer loop (r = bb - A p in FP64, p += e) stays as in the SOR defect correction, and
// the correction equation A e = r is solved to a relative tolerance by BiCGStab,
// right-preconditioned with one multigrid V-cycle per application:
//
// * level 0 is the case's own grid in the colour-split layout; its smoother is the red/black
//   colour sweep of the solve (omega = 1), FP32 for the mixed precision and FP64 otherwise;
// * coarse levels aggregate 2x2x2 cells (2x2 in 2D) with the Galerkin operator of piecewise
//   constant interpolation, P^T A P, built once per porosity field: the coarse coupling
//   across a face is the sum of the fine couplings that cross it, so the solid cells' 1e-6
//   faces and the fluid's faces keep their weights.  Rows the boundary rules decouple (the
//   outlet rows: all six couplings zero, p set by its own row) are left out of the coarse
//   space, so a cell next to them sees them as fixed values on every level;
// * coarse levels are dense arrays with one ghost layer; periodic ghosts are copies (written
//   with the update for even extents, by a refresh pass for odd ones, so every update of a
//   pass reads values from before the pass across a wrap, as the reference's p_old does);
// * the coarsest level (at most MG_COARSEST_CELLS cells where the extents allow) is solved
//   exactly by its inverse, formed on the host when the operators are built, in a one-block
//   kernel (the tail, which can also take the levels above it as phases of that block);
//   red/black sweeps in LDS take its place when that inverse is not available;
// * levels of at most 1024 cells but the coarsest are cycled twice per visit of the level
//   above (a W-cycle over the coarse end).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "pf_kernels.h"
#include "pf_layout.h"

namespace pf {

constexpr int MG_COARSEST_CELLS = 512;   // coarsening stops at this many cells or fewer
constexpr int MG_MAX_LEVELS = 16;
constexpr int MG_NPART = 4096;           // blocks of the reduction kernels (partial sums)
constexpr int MG_TAIL_LEVELS = 8;        // levels the one-block tail of the V-cycle holds at most

// A coarse level: interior extents m x n x l (l = 1 in 2D) with one ghost layer around
// (in z only in 3D), stored colour-split like level 0: ghosted cell (x, y, z) = (I+1, J+1,
// K+KG) lives in colour array (x + y + z) & 1 at slot x / 2 of row (y, z), HX slots per row,
// so a red/black sweep reads the couplings of its own colour only.
struct MgGeom {
    int dim;
    int m, n, l;
    int HX, Y, Z, KG;
    int64_t CS;          // one colour's array: HX * Y * Z
    int64_t N;           // both colours
    int per;             // y (and z in 3D) periodic
    int fuse_y, fuse_z;  // periodic with an even extent: mirrors written by the updates
    int ref_y, ref_z;    // periodic with an odd extent: a refresh pass after each update
};

PF_HD int64_t mg_off(const MgGeom &g, int I, int J, int K) {
    const int x = I + 1, y = J + 1, z = K + g.KG;
    return (int64_t)((x + y + z) & 1) * g.CS + (int64_t)(x >> 1) +
           (int64_t)g.HX * ((int64_t)y + (int64_t)g.Y * (int64_t)z);
}

inline MgGeom mg_make_geom(int dim, int m, int n, int l, int per) {
    MgGeom g;
    g.dim = dim;
    g.m = m;
    g.n = n;
    g.l = (dim == 3) ? l : 1;
    g.HX = (m + 3) / 2;
    g.Y = n + 2;
    g.KG = (dim == 3) ? 1 : 0;
    g.Z = (dim == 3) ? g.l + 2 : 1;
    g.CS = (int64_t)g.HX * g.Y * g.Z;
    g.N = 2 * g.CS;
    g.per = per;
    g.fuse_y = per && (n % 2 == 0);
    g.ref_y = per && (n % 2 == 1);
    g.fuse_z = per && dim == 3 && (g.l % 2 == 0);
    g.ref_z = per && dim == 3 && (g.l % 2 == 1);
    return g;
}

// The coarse extent of a fine extent f (aggregates of two, the last one single when f is
// odd); z is not coarsened in 2D.
inline int mg_coarse_extent(int f) { return (f + 1) / 2; }

template <typename T>
struct MgLevel {
    MgGeom g;
    T *e = nullptr, *r = nullptr;                // correction and right-hand side
    T *ap = nullptr, *cE = nullptr, *cW = nullptr, *cN = nullptr, *cS = nullptr, *cT = nullptr, *cB = nullptr;
    uint8_t *keep = nullptr;                     // 1: the row takes part in the next coarser level
};

// What the solver reads of the case it serves.
struct MgCase {
    PfGeom g;
    StepArgs a;
    double *P = nullptr;             // colour-split FP64 pressure (updated in place)
    const double *BB = nullptr;      // colour-split FP64 right-hand side
    const double *phi = nullptr;     // natural FP64 porosity (the air rows' fluid test)
    const double *PHIs = nullptr;    // colour-split FP64 porosity
    const float *PHI32 = nullptr;    // colour-split FP32 porosity (FP32 solver)
    const float *PHC = nullptr;      // FP64 solver: the porosity coded in 4 bytes, and its table
    const double *PHTAB = nullptr;   //   (used by the level-0 sweeps while `coded`)
    const uint8_t *UNIT = nullptr;   // FP64 solver: the unit-porosity chunk flags, or nullptr
    const uint8_t *UNIT32 = nullptr; // FP32 solver: the same flags of the FP32 porosity, or nullptr
    bool unit_ops = false;           // the level-0 operator kernels use the flags too (set where
                                     // unit-porosity waves hold half the cells or more)
    bool coded = false;
    bool fast = false;               // FP64 faces by the checked constant-divisor division
    int kb32 = 1, kb64 = 1, xw = 64, remap32 = 1, remap64 = 1;   // colour-sweep launch settings
};

// Tuning of the V-cycle, the same for every case (PF_MG_NU, PF_MG_NUC, PF_MG_ALPHA,
// PF_MG_CS and PF_MG_GAMMA in the environment override it when a solver is created).
struct MgParams {
    int nu1 = 2, nu2 = 2;            // level 0: red/black sweeps before and after the coarse correction
    int nuc = 2;                     // coarse levels: sweeps before and after
    double alpha = 1.6;              // over-correction of the piecewise-constant interpolation
    int coarsest_sweeps = 60;        // red/black sweeps of the coarsest level
    int tail_gamma = 2;              // cycles per level inside the one-block tail (1: V, 2: W)
};

// The solver of one case in precision T (float: mixed precision; double: FP64).
template <typename T>
struct MgSolver {
    MgCase c;
    MgParams prm;
    std::vector<MgLevel<T>> lev;     // lev[0]: the case's grid (no coarse arrays); lev[1..]: coarse
    // level-0 vectors, colour-split (2 * HS each): x = the solution e, r = the residual (holds
    // the right-hand side b on entry), the rest BiCGStab's
    T *x = nullptr, *r = nullptr, *rh = nullptr, *p = nullptr, *v = nullptr, *s = nullptr, *t = nullptr,
      *y = nullptr, *z = nullptr;
    double *scal = nullptr;          // device scalars (see pf_mg.hip)
    T *cinv = nullptr;               // the coarsest level's inverse, transposed (n x n)
    bool cinv_ok = false;            // cinv holds the inverse of the current operators
    int tail_q = 0;                  // first level of the one-block tail (0: no tail)
    double *part = nullptr;          // MG_NPART x 4 partial sums
    // the outer loop's reductions: per-wave max |r/ap| and max |e| (nslots each), the
    // reduced residual
    unsigned long long *res_slots = nullptr, *err_slots = nullptr, *res_val = nullptr;
    int nslots = 0;
    double *h_scal = nullptr;        // pinned host copy of the scalars the host tests
    hipEvent_t ev = nullptr;         // marks the end of the work the host waits for next
    void *mem = nullptr;             // the device memory holding all of the above
    bool own_mem = false;            // mem came from hipMalloc (else from the caller)
    int64_t bytes = 0;
    bool built = false;              // coarse operators match the current porosity
    // state of the current step's solve (mg_step_solve)
    int state = 0, cyc = 0, it = 0, vcycles = 0;
    double res = -1., b2 = 0.;
};

// Controls of the pressure solve of one step.
struct MgControl {
    double tol = 1e-8;               // stop when the FP64 max|r/ap| <= tol
    int max_cycles = 50;             // outer (FP64 residual + update) cycles
    double rtol = 1e-3;              // relative residual of each inner solve
    int maxit = 30;                  // BiCGStab iterations per inner solve
};

struct MgStepStats {
    double residual;                 // FP64 max|r/ap| of the final pressure
    int cycles;                      // outer cycles run
    int vcycles;                     // V-cycles applied
};

// Device bytes of a solver for grid g.
int64_t mg_bytes(const PfGeom &g, int bc, bool fp32);
// Carve the arrays from `mem` (mg_bytes of them, zeroed by the caller), or allocate them
// (hipMalloc) when mem is null; false on failure with *why set.
template <typename T>
bool mg_create(MgSolver<T> &S, const MgCase &c, std::string *why, void *mem = nullptr);
template <typename T>
void mg_destroy(MgSolver<T> &S);
// Build the coarse operators from the case's porosity (after every porosity change).
template <typename T>
void mg_setup(MgSolver<T> &S, hipStream_t st);
// The pressure solve of one step for every solver of `runs` together, solver q on
// streams[q % streams.size()]: per solver, cycles of
//   r = bb - A p (FP64, the reference's operator) and max|r/ap|, stop when <= ctl.tol;
//   A e = r from e = 0 by BiCGStab with a V-cycle per application, to ctl.rtol;
//   p += e (FP64), max |e| into err_slots;
// the host waits for one solver at a time while the others' work runs.  out[q] receives
// solver q's statistics.  Returns hipSuccess or the first HIP error.
template <typename T>
hipError_t mg_step_solve(const std::vector<MgSolver<T> *> &runs, const std::vector<hipStream_t> &streams,
                         const MgControl &ctl, std::vector<MgStepStats> &out);

}  // namespace pf
