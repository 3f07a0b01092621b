// pf_kernels.h -- launch interface of the gfx950 kernels (host side, HIP types only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "pf_layout.h"

#ifndef PF_BC_UNIFORM
#define PF_BC_UNIFORM 0
#define PF_BC_AIR 1
#endif

namespace pf {

// Per-case constants of the step, passed by value to every kernel.
struct StepArgs {
    int bc;             // PF_BC_UNIFORM / PF_BC_AIR
    int nonslip;
    int inlet_porosity; // inlet velocity scaled by the inlet cell's porosity (2D backward step)
    int face[6];        // air face codes: top, bottom, east, west, north, south
    double dt, dx, dy, dz;
    double rdx, rdy, rdz;         // RN(1/dx) etc. for the checked constant-divisor path
    double xnue, xlambda, density, thickness;
    double omega, omega_c;        // relux_factor and (1 - relux_factor)
    double u_in, v_in;            // inlet velocity components (host libm cos/sin)
    double inlet_velocity, outlet_pressure;
    // the Poisson faces between two cells of porosity 1, ((dt * 1) / d) / d per axis (0 in
    // z for 2D), and the ap they sum to in solve_p's order: the coefficients of a colour
    // sweep's unit-porosity waves
    double unit_ax, unit_ay, unit_az, unit_ap;
};

struct SweepArgs {
    int colour;         // cells with (i+j+k) & 1 == colour are updated
    int fuse_y, fuse_z; // write the periodic ghost copies in the sweep (even extents)
    int track_err;      // reduce max |dp| into the step's error slot
    int xcd_remap;      // renumber blocks XCD-contiguously
    int gx, gy;         // block grid extents of the sweep (set by the launcher)
    uint32_t mgx, mgy;  // x / gx and x / gy as umulhi(x, m) for every block index x (grid_magic; 0: divide)
    int kb;             // 3D: cells per thread, two planes apart (1, 2 or 4)
    int xw;             // colour slots of one row a wave covers: 64 (one row) or 32 (two rows)
};

void launch_faces(hipStream_t st, const PfGeom &g, const double *phi, double *AW, double *AS, double *AB,
                  double dt, double dx, double dy, double dz);
void launch_div(hipStream_t st, const PfGeom &g, const StepArgs &a, const double *uo, const double *vo,
                const double *wo, double *div);
void launch_predict(hipStream_t st, const PfGeom &g, const StepArgs &a, const double *uo, const double *vo,
                    const double *wo, const double *div, const double *phi, double *u, double *v, double *w);
void launch_rhs(hipStream_t st, const PfGeom &g, const StepArgs &a, const double *u, const double *v,
                const double *w, const double *phi, const double *P, double *BB);
void launch_sor(hipStream_t st, const PfGeom &g, const StepArgs &a, const SweepArgs &s, double *P,
                const double *AW, const double *AS, const double *AB, const double *BB, const double *phi,
                unsigned long long *err_slot);
// PHC, tab: porosity coded in 4 bytes (phi_decode in pf_cells.h); nullptr reads PHIs.
// unit: the unit-porosity chunk flags (launch_unit_flags) or nullptr to read porosity everywhere.
void launch_sor_phi(hipStream_t st, const PfGeom &g, const StepArgs &a, const SweepArgs &s, double *P,
                    const double *PHIs, const double *BB, const double *phi, unsigned long long *err_slot,
                    bool fast, const float *PHC = nullptr, const double *tab = nullptr,
                    const uint8_t *unit = nullptr);
// Unit-porosity chunk flags (unit_index in pf_layout.h) of a colour-split porosity: `unit`
// holds unit_flags_size(g) bytes.
void launch_unit_flags(hipStream_t st, const PfGeom &g, const double *PHIs, uint8_t *unit);
// The same for the FP32 porosity of the mixed-precision sweeps (values that round to 1).
void launch_unit_flags(hipStream_t st, const PfGeom &g, const float *PHI32, uint8_t *unit);
// Cells of both colours that a colour sweep of this shape (xw slots per wave row, kb cells
// per thread) updates in waves that skip porosity, from a host copy of the flags (the block
// order does not change which cells share a wave).
int64_t unit_wave_cells(const PfGeom &g, const std::vector<uint8_t> &unit, int xw, int kb);
// Code a colour-split FP64 porosity (2 * HS values) in 4 bytes per value: `code` receives
// the floats, `tab` the distinct values FP32 cannot hold.  Returns false when there are
// more of those than the code's index holds (`max_tab`), leaving the outputs unspecified.
bool code_porosity(const std::vector<double> &phis, std::vector<float> &code, std::vector<double> &tab,
                   size_t max_tab);
void launch_faces_check_direct(hipStream_t st, const PfGeom &g, const double *phi, const StepArgs &a, int *bad);
void launch_faces_check(hipStream_t st, const PfGeom &g, const double *phi, const double *AW, const double *AS,
                        const double *AB, const StepArgs &a, int *bad);
void launch_split(hipStream_t st, const PfGeom &g, const double *src, double *dst);
void launch_ghost_refresh(hipStream_t st, const PfGeom &g, int do_y, int do_z, double *P);
void launch_project(hipStream_t st, const PfGeom &g, const StepArgs &a, const double *P, double *u, double *v,
                    double *w);
void launch_boundary(hipStream_t st, const PfGeom &g, const StepArgs &a, const double *phi, double *u, double *v,
                     double *w, double *P);
void launch_divergence_out(hipStream_t st, const PfGeom &g, const StepArgs &a, const double *phi, const double *u,
                           const double *v, const double *w, double *out);
void launch_clear_u64(hipStream_t st, unsigned long long *p, int count);
int sor_blocks(const PfGeom &g);
// Temporally blocked 2D FP64 sweeps: K red/black iterations per launch, R owned rows per
// block, from Pin into Pout (two pressure buffers, alternated by the caller).  sor_tb2_ok:
// 2D, y periodic with n even, the held rows distinct and within TB2_LDS_MAX bytes of LDS.
constexpr int64_t TB2_LDS_MAX = 160 * 1024 - 1024;
constexpr int TB2_MAX_ROWS = 256;
constexpr int TB2_THREADS = 1024;
constexpr int TB2_ARRAYS = 14;   // LDS arrays per held cell slot: p, bb and 5 coefficients, per colour
int64_t sor_tb2_lds_bytes(const PfGeom &g, int K, int R);
bool sor_tb2_ok(const PfGeom &g, int bc, int K, int R);
void launch_sor_tb2(hipStream_t st, const PfGeom &g, const StepArgs &a, int first, int K, int R, const double *Pin,
                    double *Pout, const double *PHIs, const double *BB, const double *phi, unsigned long long *slot,
                    bool fast);
int sweep_wave_width(const PfGeom &g);   // 64 or 32 colour slots per wave row
// One red/black iteration in one launch, ping-pong A -> B (pf_fused.h); usable when
// sor_fused_ok (even periodic extents).  B must hold A's non-periodic ghost values.
bool sor_fused_ok(const PfGeom &g, int bc);
void launch_sor_fused(hipStream_t st, const PfGeom &g, const StepArgs &a, const double *A, double *B,
                      const double *F, const double *R, const double *phi, unsigned long long *slot, int track,
                      bool fast);
void launch_sor32_fused(hipStream_t st, const PfGeom &g, const StepArgs &a, const float *A, float *B, const float *F,
                        const float *R, const double *phi, unsigned long long *slot, int track);
// Body force (pressure + wall-force terms): partial holds force_blocks(g)*6 doubles, out 6.
int force_blocks(const PfGeom &g);
void launch_force(hipStream_t st, const PfGeom &g, const StepArgs &a, const double *phi, const double *u,
                  const double *v, const double *w, const double *P, double *partial, double *out);
void launch_copy_probe(hipStream_t st, const double *src, double *dst, int64_t n);
// mixed precision (pf_mixed.hip)
// unit: the unit-porosity chunk flags of the FP64 porosity PHIs, or nullptr.
void launch_residual(hipStream_t st, const PfGeom &g, const StepArgs &a, SweepArgs s, const double *P,
                     const double *PHIs, const double *BB, const double *phi, float *R, float *E,
                     unsigned long long *slots, bool fast, const uint8_t *unit = nullptr);
// unit: the unit-porosity chunk flags of the FP32 porosity (launch_unit_flags), or nullptr.
void launch_sor32(hipStream_t st, const PfGeom &g, const StepArgs &a, SweepArgs s, float *E, const float *F32,
                  const float *R, const double *phi, unsigned long long *err_slot, const uint8_t *unit = nullptr);
void launch_residual(hipStream_t st, const PfGeom &g, const StepArgs &a, SweepArgs s, const double *P,
                     const double *PHIs, const double *BB, const double *phi, double *R, double *E,
                     unsigned long long *slots, bool fast, const uint8_t *unit = nullptr);
void launch_update(hipStream_t st, const PfGeom &g, const float *E, double *P);
void launch_update(hipStream_t st, const PfGeom &g, const double *E, double *P);
void launch_update_track(hipStream_t st, const PfGeom &g, const float *E, double *P, unsigned long long *slots,
                         int nslots);
void launch_update_track(hipStream_t st, const PfGeom &g, const double *E, double *P, unsigned long long *slots,
                         int nslots);
void launch_ghost_refresh32(hipStream_t st, const PfGeom &g, int do_y, int do_z, float *E);
void launch_split32(hipStream_t st, const PfGeom &g, const double *src, float *dst);   // blocks of one colour-sweep launch (= error slots)
void launch_reduce_err(hipStream_t st, unsigned long long *slots, int nslots, unsigned long long *step_err);

}  // namespace pf
