// pf_mg.hip -- multigrid-preconditioned BiCGStab for A e = r (see pf_mg.h).
//
// Kernels, by level:
//   level 0 (the case's grid, colour-split T arrays):  the red/black colour sweeps of the
//     solve with omega = 1 (k_sor32 / k_sor_phi) smooth; k_mg_restrict0 sums the residual
//     of the kept cells of each 2x2x2 aggregate into level 1 (and clears level 1's e);
//     k_mg_prolong0 adds alpha times the aggregate's correction back; k_mg_apply0 is A y
//     with the dot products BiCGStab needs in the same pass;
//   coarse levels (dense, ghosted):  k_mg_smooth (one colour of a red/black sweep),
//     k_mg_restrict, k_mg_prolong, k_mg_refresh (odd periodic extents), k_mg_coarsest (the
//     last level, solved in LDS by one block);
//   set-up:  k_mg_gal0 / k_mg_gal (Galerkin sums P^T A P).
// BiCGStab's scalars stay on the device (k_bicg_scalar*), the host reads one word per
// iteration to test convergence.  Reductions are fixed-order trees, so a solve is
// deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "pf_cells.h"
#include "pf_device.h"
#include "pf_kernels.h"
#include "pf_mg.h"

namespace pf {

template <typename T>
struct MgK {
    T cdx, cdy, cdz;   // FP32 solver: face scale dt / d^2 as the FP32 sweep computes it
};

// ------------------------------------------------------------------------------------
// Level 0: the operator of the colour sweep of precision T at fine cell (i, j, k).
// float: the FP32 sweep's faces (sor32_value); double: the reference's faces (sor_value).
// Returns ap; c receives the six couplings after the boundary rules.
// ------------------------------------------------------------------------------------
template <typename T, int DIM, bool FAST>
__device__ __forceinline__ T op0(const PfGeom &g, const StepArgs &a, const MgK<T> &K, const T *__restrict__ F_,
                                 const double *__restrict__ phi, int i, int j, int k, CoefT<T> &c) {
    const auto F = gmem(F_);
    const T fc = F[pf_soff(g, i, j, k)];
    const T fE = F[pf_soff(g, i + 1, j, k)], fW = F[pf_soff(g, i - 1, j, k)];
    const T fN = F[pf_soff(g, i, j + 1, k)], fS = F[pf_soff(g, i, j - 1, k)];
    T fT = T(0), fB = T(0);
    if (DIM == 3) {
        fT = F[pf_soff(g, i, j, k + 1)];
        fB = F[pf_soff(g, i, j, k - 1)];
    }
    T ap;
    if constexpr (std::is_same<T, float>::value) {
        c.ae = face32(fE, fc, K.cdx);
        c.aw = face32(fc, fW, K.cdx);
        c.an = face32(fN, fc, K.cdy);
        c.as = face32(fc, fS, K.cdy);
        c.at = 0.f;
        c.ab = 0.f;
        if (DIM == 3) {
            c.at = face32(fT, fc, K.cdz);
            c.ab = face32(fc, fB, K.cdz);
        }
        ap = -(c.ae + c.aw + c.an + c.as + c.at + c.ab);
    } else {
        c.ae = face_any<FAST>(fE, fc, a.dt, a.dx, a.rdx);
        c.aw = face_any<FAST>(fc, fW, a.dt, a.dx, a.rdx);
        c.an = face_any<FAST>(fN, fc, a.dt, a.dy, a.rdy);
        c.as = face_any<FAST>(fc, fS, a.dt, a.dy, a.rdy);
        c.at = 0.;
        c.ab = 0.;
        if (DIM == 3) {
            c.at = face_any<FAST>(fT, fc, a.dt, a.dz, a.rdz);
            c.ab = face_any<FAST>(fc, fB, a.dt, a.dz, a.rdz);
            ap = (((((-c.ae) - c.aw) - c.an) - c.as) - c.at) - c.ab;
        } else {
            ap = (((-c.ae) - c.aw) - c.an) - c.as;
        }
    }
    if (is_edge_row(g, a, i, j, k)) row_rules<false, T>(g, a, i, j, k, c, nullptr, nullptr, phi, 0.);
    return ap;
}

// op0 at a cell whose seven porosity operands are all 1 (a unit-porosity chunk): the same
// couplings and ap without the porosity loads.  FP64: the faces of two unit porosities from
// StepArgs (unit_ax/ay/az, unit_ap: the values face_any and op0's sum give for them); FP32:
// face32 of two unit porosities, which the compiler evaluates exactly as for loaded ones.
template <typename T, int DIM>
__device__ __forceinline__ T op0_unit(const PfGeom &g, const StepArgs &a, const MgK<T> &K,
                                      const double *__restrict__ phi, int i, int j, int k, CoefT<T> &c) {
    T ap;
    if constexpr (std::is_same<T, float>::value) {
        c.ae = face32(1.f, 1.f, K.cdx);
        c.aw = face32(1.f, 1.f, K.cdx);
        c.an = face32(1.f, 1.f, K.cdy);
        c.as = face32(1.f, 1.f, K.cdy);
        c.at = 0.f;
        c.ab = 0.f;
        if (DIM == 3) {
            c.at = face32(1.f, 1.f, K.cdz);
            c.ab = face32(1.f, 1.f, K.cdz);
        }
        ap = -(c.ae + c.aw + c.an + c.as + c.at + c.ab);
    } else {
        c.ae = a.unit_ax;
        c.aw = a.unit_ax;
        c.an = a.unit_ay;
        c.as = a.unit_ay;
        c.at = (DIM == 3) ? a.unit_az : 0.;
        c.ab = (DIM == 3) ? a.unit_az : 0.;
        ap = a.unit_ap;
    }
    if (is_edge_row(g, a, i, j, k)) row_rules<false, T>(g, a, i, j, k, c, nullptr, nullptr, phi, 0.);
    return ap;
}

// A row that the boundary rules decouple (all six couplings zero: the outlet rows) takes
// no part in the coarse space.
__device__ __forceinline__ bool dec0(const PfGeom &g, const StepArgs &a, int i, int j, int k, const double *phi) {
    if (!is_edge_row(g, a, i, j, k)) return false;
    if (a.bc == PF_BC_UNIFORM) return i == g.m;
    CoefT<float> c{1.f, 1.f, 1.f, 1.f, 1.f, 1.f};
    row_rules<false, float>(g, a, i, j, k, c, nullptr, nullptr, phi, 0.);
    return c.ae == 0.f && c.aw == 0.f && c.an == 0.f && c.as == 0.f && c.at == 0.f && c.ab == 0.f;
}

// (A y)(i, j, k) at level 0; UNIT: the cell's porosity operands are all 1 (op0_unit).
template <typename T, int DIM, bool FAST, bool UNIT = false>
__device__ __forceinline__ T apply0_at(const PfGeom &g, const StepArgs &a, const MgK<T> &K, const T *F,
                                       const double *phi, const T *__restrict__ y_, int i, int j, int k) {
    CoefT<T> c;
    const T ap = UNIT ? op0_unit<T, DIM>(g, a, K, phi, i, j, k, c) : op0<T, DIM, FAST>(g, a, K, F, phi, i, j, k, c);
    const auto y = gmem(y_);
    T s = ap * y[pf_soff(g, i, j, k)];
    s = s + c.ae * y[pf_soff(g, i + 1, j, k)];
    s = s + c.aw * y[pf_soff(g, i - 1, j, k)];
    s = s + c.an * y[pf_soff(g, i, j + 1, k)];
    s = s + c.as * y[pf_soff(g, i, j - 1, k)];
    if (DIM == 3) {
        s = s + c.at * y[pf_soff(g, i, j, k + 1)];
        s = s + c.ab * y[pf_soff(g, i, j, k - 1)];
    }
    return s;
}

constexpr int PROLONG0_ROWS = 4;   // rows per thread of the level-0 prolongation

// Store v at coarse cell (I,J,K) and at the periodic ghost slots that mirror it.
template <typename T>
__device__ __forceinline__ void mg_store(const MgGeom &g, T *__restrict__ e, int I, int J, int K, T v, bool my,
                                         bool mz) {
    e[mg_off(g, I, J, K)] = v;
    if (my) {
        if (J == 0) e[mg_off(g, I, g.n, K)] = v;
        if (J == g.n - 1) e[mg_off(g, I, -1, K)] = v;
    }
    if (mz && g.dim == 3) {
        if (K == 0) e[mg_off(g, I, J, g.l)] = v;
        if (K == g.l - 1) e[mg_off(g, I, J, -1)] = v;
    }
}

PF_HD void mg_cell(const MgGeom &g, int64_t q, int &I, int &J, int &K) {
    I = (int)(q % g.m);
    const int64_t t = q / g.m;
    J = (int)(t % g.n);
    K = (int)(t / g.n);
}

// ------------------------------------------------------------------------------------
// Level 0 -> 1: r1 = sum over the kept cells of each aggregate of (rhs - A e); e1 = 0.
// Cells of colour `zero_colour` are left out: the pre-smoothing ends with a Gauss-Seidel
// (omega = 1) sweep of that colour, after which their residual is zero but for rounding.
// ------------------------------------------------------------------------------------
// UNIT: `unit` holds the unit-porosity chunk flags of F; a wave whose fine cells all lie in
// flagged chunks takes op0_unit for them.
template <typename T, int DIM, bool FAST, bool UNIT>
__global__ __launch_bounds__(256) void k_mg_restrict0(PfGeom g, StepArgs a, MgK<T> K, const T *F,
                                                      const double *phi, const T *__restrict__ rhs,
                                                      const T *__restrict__ e, MgGeom c1, T *__restrict__ r1,
                                                      T *__restrict__ e1, int zero_colour,
                                                      const uint8_t *__restrict__ unit) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = q < (int64_t)c1.m * c1.n * c1.l;
    int I = 0, J = 0, Kc = 0;
    if (active) mg_cell(c1, q, I, J, Kc);
    // the fine cells this thread sums, and whether all of them lie in unit-porosity chunks
    int ok = 1;
    if (UNIT && active) {
#pragma unroll
        for (int dz = 0; dz < (DIM == 3 ? 2 : 1); ++dz)
#pragma unroll
            for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                for (int dx = 0; dx < 2; ++dx) {
                    const int i = 2 * I + dx + 1, j = 2 * J + dy + 1, k = (DIM == 3) ? 2 * Kc + dz + 1 : 0;
                    if (i > g.m || j > g.n || (DIM == 3 && k > g.l)) continue;
                    const int c = (i + j + k) & 1;
                    if (c == zero_colour) continue;
                    // slot position h + sh - 1 of cell i = 2 h + sh
                    ok &= (int)gmem(unit)[unit_index(g, c, j, k, (i >> 1) + (i & 1) - 1)];
                }
    }
    const bool all_unit = UNIT && __all(ok) != 0;   // the wave's branch
    if (!active) return;
    T sum = T(0);
#pragma unroll
    for (int dz = 0; dz < (DIM == 3 ? 2 : 1); ++dz)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                const int i = 2 * I + dx + 1, j = 2 * J + dy + 1, k = (DIM == 3) ? 2 * Kc + dz + 1 : 0;
                if (i > g.m || j > g.n || (DIM == 3 && k > g.l)) continue;
                if (((i + j + k) & 1) == zero_colour) continue;
                if (dec0(g, a, i, j, k, phi)) continue;
                const T Ae = all_unit ? apply0_at<T, DIM, FAST, true>(g, a, K, F, phi, e, i, j, k)
                                      : apply0_at<T, DIM, FAST>(g, a, K, F, phi, e, i, j, k);
                sum = sum + (gmem(rhs)[pf_soff(g, i, j, k)] - Ae);
            }
    r1[mg_off(c1, I, J, Kc)] = sum;
    mg_store(c1, e1, I, J, Kc, T(0), c1.per, c1.per);
}

// Level 1 -> 0: e += alpha * e1(parent) at every kept fine cell, with the periodic ghost
// copies of even extents (odd extents: the refresh pass follows).
template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_mg_prolong0(PfGeom g, StepArgs a, const double *phi, MgGeom c1,
                                                     const T *__restrict__ e1, T *__restrict__ e, T alpha, int fuse_y,
                                                     int fuse_z) {
    // blockIdx.x: colour (outer) and chunk of 64 slots; lanes on consecutive slots of that
    // colour (i = 2 h + sh), so the wave's loads and stores of e are contiguous.  A thread
    // takes PROLONG0_ROWS rows, 4 apart, and loads all of them before it stores any.
    const int chunks = ((g.m + 1) / 2 + 63) / 64;
    const int cb = (int)blockIdx.x / chunks;
    const int hx = ((int)blockIdx.x - cb * chunks) * 64 + (int)threadIdx.x;
    const int k = (DIM == 3) ? (int)blockIdx.z + 1 : 0;
    T v[PROLONG0_ROWS];
    int64_t so[PROLONG0_ROWS];
    int jj[PROLONG0_ROWS], ii[PROLONG0_ROWS];
    bool act[PROLONG0_ROWS];
#pragma unroll
    for (int q = 0; q < PROLONG0_ROWS; ++q) {
        const int j = (int)blockIdx.y * (4 * PROLONG0_ROWS) + q * 4 + (int)threadIdx.y + 1;
        const int sh = (cb + j + k) & 1;
        const int i = 2 * ((1 - sh) + hx) + sh;
        jj[q] = j;
        ii[q] = i;
        act[q] = i <= g.m && j <= g.n && !dec0(g, a, i, j, k, phi);
        so[q] = 0;
        v[q] = T(0);
        if (!act[q]) continue;
        const T ec = e1[mg_off(c1, (i - 1) >> 1, (j - 1) >> 1, (DIM == 3) ? (k - 1) >> 1 : 0)];
        so[q] = pf_soff(g, i, j, k);
        v[q] = e[so[q]] + alpha * ec;
    }
#pragma unroll
    for (int q = 0; q < PROLONG0_ROWS; ++q) {
        if (!act[q]) continue;
        const int i = ii[q], j = jj[q];
        e[so[q]] = v[q];
        if (fuse_y) {
            if (j == g.n) e[pf_soff(g, i, 0, k)] = v[q];
            if (j == 1) e[pf_soff(g, i, g.n + 1, k)] = v[q];
        }
        if (DIM == 3 && fuse_z) {
            if (k == g.l) e[pf_soff(g, i, j, 0)] = v[q];
            if (k == 1) e[pf_soff(g, i, j, g.l + 1)] = v[q];
        }
    }
}

// out = A y at level 0; MODE 0: partial (w, out); MODE 1: partials (out, w) and (out, out).
// UNIT: `unit` holds the unit-porosity chunk flags of F (unit_index in pf_layout.h); a
// wave whose 64 cells all lie in flagged chunks takes op0_unit.
template <typename T, int DIM, bool FAST, int MODE, bool UNIT>
__global__ __launch_bounds__(256) void k_mg_apply0(PfGeom g, StepArgs a, MgK<T> K, const T *F, const double *phi,
                                                   const T *__restrict__ y, T *__restrict__ out,
                                                   const T *__restrict__ w, double *__restrict__ part,
                                                   const uint8_t *__restrict__ unit) {
    const int rows = g.n * (DIM == 3 ? g.l : 1);
    double d0 = 0., d1 = 0.;
    for (int row = (int)blockIdx.x * 4 + (int)threadIdx.y; row < rows; row += (int)gridDim.x * 4) {
        const int j = row % g.n + 1;
        const int k = (DIM == 3) ? row / g.n + 1 : 0;
        // the row's cells colour by colour, the lanes of a wave on consecutive slots of one
        // colour array (i = 2 h + sh), so every load of the wave is one contiguous run
        for (int c = 0; c < 2; ++c) {
            const int sh = (c + j + k) & 1;
            for (int h0 = 0; h0 < (g.m + 1) / 2; h0 += 64) {
                const int i = 2 * ((1 - sh) + h0 + (int)threadIdx.x) + sh;
                // the lane's chunk is its slot position h + sh - 1 = h0 + lane (the sweeps' own)
                bool all_unit = false;
                if constexpr (UNIT)
                    all_unit = __all(i > g.m || gmem(unit)[unit_index(g, c, j, k, h0 + (int)threadIdx.x)] != 0) != 0;
                if (i > g.m) continue;
                const T Ay = all_unit ? apply0_at<T, DIM, FAST, true>(g, a, K, F, phi, y, i, j, k)
                                      : apply0_at<T, DIM, FAST>(g, a, K, F, phi, y, i, j, k);
                const int64_t so = pf_soff(g, i, j, k);
                out[so] = Ay;
                const double wv = (double)w[so];
                if (MODE == 0) {
                    d0 += wv * (double)Ay;
                } else {
                    d0 += (double)Ay * wv;
                    d1 += (double)Ay * (double)Ay;
                }
            }
        }
    }
    __shared__ double red[2][256];
    const int tid = (int)(threadIdx.x + threadIdx.y * blockDim.x);
    red[0][tid] = d0;
    red[1][tid] = d1;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) {
            red[0][tid] += red[0][tid + off];
            red[1][tid] += red[1][tid + off];
        }
        __syncthreads();
    }
    if (tid == 0) {
        part[(int64_t)blockIdx.x * 4 + 0] = red[0][0];
        part[(int64_t)blockIdx.x * 4 + 1] = red[1][0];
    }
}

// ------------------------------------------------------------------------------------
// Coarse levels.
// ------------------------------------------------------------------------------------
template <typename T>
struct MgOp {
    const T *ap, *cE, *cW, *cN, *cS, *cT, *cB;
};

template <typename T>
__device__ __forceinline__ T mg_ax(const MgGeom &g, const MgOp<T> &A, const T *__restrict__ e, int I, int J, int K,
                                   int64_t o) {
    T s = A.ap[o] * e[o];
    s = s + A.cE[o] * e[mg_off(g, I + 1, J, K)];
    s = s + A.cW[o] * e[mg_off(g, I - 1, J, K)];
    s = s + A.cN[o] * e[mg_off(g, I, J + 1, K)];
    s = s + A.cS[o] * e[mg_off(g, I, J - 1, K)];
    if (g.dim == 3) {
        s = s + A.cT[o] * e[mg_off(g, I, J, K + 1)];
        s = s + A.cB[o] * e[mg_off(g, I, J, K - 1)];
    }
    return s;
}

// One colour of a red/black Gauss-Seidel sweep (omega = 1).  Threads walk the colour's
// cells only, (m + 1) / 2 per row: I = 2 h + ((colour + J + K) & 1).
template <typename T>
__global__ __launch_bounds__(256) void k_mg_smooth(MgGeom g, MgOp<T> A, const T *__restrict__ r, T *__restrict__ e,
                                                   int colour) {
    const uint32_t hm = (uint32_t)(g.m + 1) >> 1;
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (uint64_t)hm * (uint64_t)g.n * (uint64_t)g.l) return;
    const uint64_t row = q / hm;
    const int h = (int)(q - row * hm);
    const int J = (int)(row % (uint64_t)g.n), K = (int)(row / (uint64_t)g.n);
    const int I = 2 * h + ((colour + J + K) & 1);
    if (I >= g.m) return;
    const int64_t o = mg_off(g, I, J, K);
    T s = r[o];
    s = s - A.cE[o] * e[mg_off(g, I + 1, J, K)];
    s = s - A.cW[o] * e[mg_off(g, I - 1, J, K)];
    s = s - A.cN[o] * e[mg_off(g, I, J + 1, K)];
    s = s - A.cS[o] * e[mg_off(g, I, J - 1, K)];
    if (g.dim == 3) {
        s = s - A.cT[o] * e[mg_off(g, I, J, K + 1)];
        s = s - A.cB[o] * e[mg_off(g, I, J, K - 1)];
    }
    mg_store(g, e, I, J, K, s / A.ap[o], g.fuse_y, g.fuse_z);
}

// Periodic ghosts of odd extents from their source cells.
template <typename T>
__global__ void k_mg_refresh(MgGeom g, T *__restrict__ e) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ny = (int64_t)g.m * g.l, nz = (int64_t)g.m * g.n;
    if (g.ref_y && q < ny) {
        const int I = (int)(q % g.m), K = (int)(q / g.m);
        e[mg_off(g, I, -1, K)] = e[mg_off(g, I, g.n - 1, K)];
        e[mg_off(g, I, g.n, K)] = e[mg_off(g, I, 0, K)];
    }
    if (g.ref_z && g.dim == 3 && q < nz) {
        const int I = (int)(q % g.m), J = (int)(q / g.m);
        e[mg_off(g, I, J, -1)] = e[mg_off(g, I, J, g.l - 1)];
        e[mg_off(g, I, J, g.l)] = e[mg_off(g, I, J, 0)];
    }
}

// Level f -> c: rc = sum over the kept cells of each aggregate of (r - A e); ec = 0.
template <typename T>
__global__ __launch_bounds__(256) void k_mg_restrict(MgGeom f, MgOp<T> A, const uint8_t *__restrict__ keep,
                                                     const T *__restrict__ r, const T *__restrict__ e, MgGeom c,
                                                     T *__restrict__ rc, T *__restrict__ ec) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)c.m * c.n * c.l) return;
    int I, J, K;
    mg_cell(c, q, I, J, K);
    T sum = T(0);
    for (int dz = 0; dz < (f.dim == 3 ? 2 : 1); ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int i = 2 * I + dx, j = 2 * J + dy, k = (f.dim == 3) ? 2 * K + dz : 0;
                if (i >= f.m || j >= f.n || k >= f.l) continue;
                const int64_t o = mg_off(f, i, j, k);
                if (!keep[o]) continue;
                sum = sum + (r[o] - mg_ax(f, A, e, i, j, k, o));
            }
    rc[mg_off(c, I, J, K)] = sum;
    mg_store(c, ec, I, J, K, T(0), c.per, c.per);
}

// Level c -> f: e += alpha * ec(parent) at every kept cell of f.
template <typename T>
__global__ __launch_bounds__(256) void k_mg_prolong(MgGeom f, const uint8_t *__restrict__ keep, MgGeom c,
                                                    const T *__restrict__ ec, T *__restrict__ e, T alpha) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)f.m * f.n * f.l) return;
    int I, J, K;
    mg_cell(f, q, I, J, K);
    const int64_t o = mg_off(f, I, J, K);
    if (!keep[o]) return;
    const T v = e[o] + alpha * ec[mg_off(c, I >> 1, J >> 1, (f.dim == 3) ? K >> 1 : 0)];
    mg_store(f, e, I, J, K, v, f.fuse_y, f.fuse_z);
}

// The coarsest level: `sweeps` red/black Gauss-Seidel sweeps from e = 0 in LDS, one block
// of 256 threads, each owning at most MG_CPT cells (their r and couplings in registers).
constexpr int MG_CPT = 4;
constexpr int MG_LDS = 6144;   // ghosted cells of the coarsest level (checked on the host)

template <typename T>
__global__ __launch_bounds__(256) void k_mg_coarsest(MgGeom g, MgOp<T> A, const T *__restrict__ r,
                                                     T *__restrict__ e, int sweeps) {
    __shared__ T se[MG_LDS];
    const int tid = (int)threadIdx.x;
    const int cells = g.m * g.n * g.l;
    for (int q = tid; q < (int)g.N; q += 256) se[q] = T(0);
    int64_t o[MG_CPT];
    int nb[MG_CPT][6];   // the six neighbours' offsets (E, W, N, S, T, B)
    int col[MG_CPT], I[MG_CPT], J[MG_CPT], K[MG_CPT];
    T rr[MG_CPT], ap[MG_CPT], cE[MG_CPT], cW[MG_CPT], cN[MG_CPT], cS[MG_CPT], cT[MG_CPT], cB[MG_CPT];
#pragma unroll
    for (int u = 0; u < MG_CPT; ++u) {
        const int q = tid + 256 * u;
        col[u] = -1;
        if (q < cells) {
            mg_cell(g, q, I[u], J[u], K[u]);
            o[u] = mg_off(g, I[u], J[u], K[u]);
            nb[u][0] = (int)mg_off(g, I[u] + 1, J[u], K[u]);
            nb[u][1] = (int)mg_off(g, I[u] - 1, J[u], K[u]);
            nb[u][2] = (int)mg_off(g, I[u], J[u] + 1, K[u]);
            nb[u][3] = (int)mg_off(g, I[u], J[u] - 1, K[u]);
            nb[u][4] = (g.dim == 3) ? (int)mg_off(g, I[u], J[u], K[u] + 1) : 0;
            nb[u][5] = (g.dim == 3) ? (int)mg_off(g, I[u], J[u], K[u] - 1) : 0;
            col[u] = (I[u] + J[u] + K[u]) & 1;
            rr[u] = r[o[u]];
            ap[u] = A.ap[o[u]];
            cE[u] = A.cE[o[u]];
            cW[u] = A.cW[o[u]];
            cN[u] = A.cN[o[u]];
            cS[u] = A.cS[o[u]];
            cT[u] = (g.dim == 3) ? A.cT[o[u]] : T(0);
            cB[u] = (g.dim == 3) ? A.cB[o[u]] : T(0);
        }
    }
    __syncthreads();
    for (int it = 0; it < sweeps; ++it) {
        for (int c = 0; c < 2; ++c) {
#pragma unroll
            for (int u = 0; u < MG_CPT; ++u) {
                if (col[u] != c) continue;
                const int64_t x = o[u];
                T s = rr[u];
                s = s - cE[u] * se[nb[u][0]];
                s = s - cW[u] * se[nb[u][1]];
                s = s - cN[u] * se[nb[u][2]];
                s = s - cS[u] * se[nb[u][3]];
                if (g.dim == 3) {
                    s = s - cT[u] * se[nb[u][4]];
                    s = s - cB[u] * se[nb[u][5]];
                }
                const T v = s / ap[u];
                se[x] = v;
                if (g.fuse_y) {
                    if (J[u] == 0) se[mg_off(g, I[u], g.n, K[u])] = v;
                    if (J[u] == g.n - 1) se[mg_off(g, I[u], -1, K[u])] = v;
                }
                if (g.fuse_z && g.dim == 3) {
                    if (K[u] == 0) se[mg_off(g, I[u], J[u], g.l)] = v;
                    if (K[u] == g.l - 1) se[mg_off(g, I[u], J[u], -1)] = v;
                }
            }
            __syncthreads();
            if (g.ref_y || g.ref_z) {
                for (int q = tid; q < g.m * (g.l > g.n ? g.l : g.n); q += 256) {
                    const int a0 = q % g.m, b0 = q / g.m;
                    if (g.ref_y && b0 < g.l) {
                        se[mg_off(g, a0, -1, b0)] = se[mg_off(g, a0, g.n - 1, b0)];
                        se[mg_off(g, a0, g.n, b0)] = se[mg_off(g, a0, 0, b0)];
                    }
                    if (g.ref_z && g.dim == 3 && b0 < g.n) {
                        se[mg_off(g, a0, b0, -1)] = se[mg_off(g, a0, b0, g.l - 1)];
                        se[mg_off(g, a0, b0, g.l)] = se[mg_off(g, a0, b0, 0)];
                    }
                }
                __syncthreads();
            }
        }
    }
    for (int q = tid; q < (int)g.N; q += 256) e[q] = se[q];
}

// ------------------------------------------------------------------------------------
// The tail of the V-cycle: every level from the first one of at most MG_TAIL_CELLS cells
// down to the coarsest, in one block of MG_TAIL_THREADS threads.  Each launch of the coarse
// levels is a phase here, separated by block barriers (global memory written by the block
// is visible to the block after the barrier); each thread gathers the operands of up to
// MG_TAIL_U cells before it stores any, so a phase costs about one memory round trip.  The
// coarsest level is solved exactly by its inverse (coarsest_inv, transposed), computed on
// the host when the operators are built.  The operations and their order are those of the
// per-level launches.
// ------------------------------------------------------------------------------------
constexpr int MG_TAIL_THREADS = 1024;
constexpr int MG_TAIL_U = 8;
constexpr int64_t MG_TAIL_CELLS = (int64_t)MG_TAIL_THREADS * MG_TAIL_U;   // the most a tail level may hold
// The tail starts at the first level of at most this many cells (PF_MG_TAIL=n: at most n,
// up to MG_TAIL_CELLS), by default the coarsest level alone: every level above it costs
// less as separate launches over many CUs than as phases of the one block.  With the same
// cycle (and the same bits), mixed / FP64 ms per step with the tail from 8192 / 1024 / 128
// cells: sphere 256x128x128 9.03 / 12.23, 8.24 / 11.21, 6.97 / 9.73; boxes 512x256x256
// 50.3 / 87.6, 49.2 / 86.0, 47.1 / 84.2.
constexpr int64_t MG_TAIL_START_CELLS = MG_COARSEST_CELLS;
// Levels of at most this many cells (but the coarsest) are cycled gamma times per visit of
// the level above (a W-cycle over the coarse end), whether they run as launches or in the
// tail: on the boxes 512x256x256 that takes the V-cycles per step from 25 to 21.5.
constexpr int64_t MG_W_CELLS = 1024;

// Division of a cell index by a level's extent inside the tail: x / d = umulhi(x, M) with
// M = floor(2^32 / d) + 1, exact whenever x * d < 2^32 (tail levels hold at most
// MG_TAIL_CELLS cells, so x and d are below 2^13); M = 0 stands for d = 1.  The host
// checks every quotient a tail level uses against true division (tail_magic_ok).
PF_HD uint32_t tail_magic(uint32_t d) { return d <= 1 ? 0u : (uint32_t)((((uint64_t)1) << 32) / d + 1); }
PF_HD uint32_t tail_div(uint32_t x, uint32_t magic) {
#if defined(__HIP_DEVICE_COMPILE__)
    return magic ? __umulhi(x, magic) : x;
#else
    return magic ? (uint32_t)(((uint64_t)x * magic) >> 32) : x;
#endif
}

template <typename T>
struct MgTailLevel {
    MgGeom g;
    MgOp<T> A;
    const uint8_t *keep;
    T *r, *e;
    uint32_t mm, mh, mn;    // tail_magic of m, (m + 1) / 2 and n
};

template <typename T>
struct MgTailArgs {
    MgTailLevel<T> L[MG_TAIL_LEVELS];
    int nl;                 // levels in L; L[nl - 1] is the coarsest
    int nuc;                // sweeps before and after the coarse correction
    int gamma;              // cycles of the next level per level (1: V, 2: W)
    T alpha;
    const T *cinv;          // coarsest inverse, transposed: cinv[j * n + i] = (A^-1)(i, j)
};

// mg_off in 32 bits (a tail level's arrays hold far fewer than 2^32 elements).
__device__ __forceinline__ uint32_t tail_off(const MgGeom &g, int I, int J, int K) {
    const uint32_t x = (uint32_t)(I + 1), y = (uint32_t)(J + 1), z = (uint32_t)(K + g.KG);
    return ((x + y + z) & 1u) * (uint32_t)g.CS + (x >> 1) + (uint32_t)g.HX * (y + (uint32_t)g.Y * z);
}

// Cell q of a level in natural order (I fastest).
template <typename T>
__device__ __forceinline__ void tail_cell(const MgTailLevel<T> &L, uint32_t q, int &I, int &J, int &K) {
    const uint32_t row = tail_div(q, L.mm);
    I = (int)(q - row * (uint32_t)L.g.m);
    K = (int)tail_div(row, L.mn);
    J = (int)(row - (uint32_t)K * (uint32_t)L.g.n);
}

template <typename T>
__device__ __forceinline__ void tail_store(const MgGeom &g, T *__restrict__ e, int I, int J, int K, T v, bool my,
                                           bool mz) {
    e[tail_off(g, I, J, K)] = v;
    if (my) {
        if (J == 0) e[tail_off(g, I, g.n, K)] = v;
        if (J == g.n - 1) e[tail_off(g, I, -1, K)] = v;
    }
    if (mz && g.dim == 3) {
        if (K == 0) e[tail_off(g, I, J, g.l)] = v;
        if (K == g.l - 1) e[tail_off(g, I, J, -1)] = v;
    }
}

template <typename T>
__device__ __forceinline__ void tail_refresh(const MgTailLevel<T> &L, T *__restrict__ e) {
    const MgGeom &g = L.g;
    if (!(g.ref_y || g.ref_z)) return;
    const uint32_t ny = (uint32_t)g.m * g.l, nz = (uint32_t)g.m * g.n;
    const uint32_t nmax = ny > nz ? ny : nz;
    for (uint32_t q = threadIdx.x; q < nmax; q += MG_TAIL_THREADS) {
        const uint32_t t = tail_div(q, L.mm);
        const int I = (int)(q - t * (uint32_t)g.m);
        if (g.ref_y && q < ny) {
            const int K = (int)t;
            e[tail_off(g, I, -1, K)] = e[tail_off(g, I, g.n - 1, K)];
            e[tail_off(g, I, g.n, K)] = e[tail_off(g, I, 0, K)];
        }
        if (g.ref_z && g.dim == 3 && q < nz) {
            const int J = (int)t;
            e[tail_off(g, I, J, -1)] = e[tail_off(g, I, J, g.l - 1)];
            e[tail_off(g, I, J, g.l)] = e[tail_off(g, I, J, 0)];
        }
    }
    __syncthreads();
}

// One colour of the level's Gauss-Seidel sweep (omega = 1): the threads walk that colour's
// cells only, (m + 1) / 2 slots per row, I = 2 h + ((colour + J + K) & 1).
template <typename T>
__device__ __forceinline__ void tail_smooth(const MgTailLevel<T> &L, int colour) {
    const MgGeom &g = L.g;
    const uint32_t hm = (uint32_t)(g.m + 1) >> 1;
    const uint32_t count = hm * (uint32_t)g.n * (uint32_t)g.l;
    for (uint32_t base = threadIdx.x; base < count; base += (uint32_t)MG_TAIL_THREADS * MG_TAIL_U) {
        T v[MG_TAIL_U];
        int I[MG_TAIL_U], J[MG_TAIL_U], K[MG_TAIL_U];
        bool act[MG_TAIL_U];
#pragma unroll
        for (int u = 0; u < MG_TAIL_U; ++u) {
            const uint32_t x = base + (uint32_t)u * MG_TAIL_THREADS;
            act[u] = false;
            v[u] = T(0);
            if (x >= count) continue;
            const uint32_t row = tail_div(x, L.mh);
            const int h = (int)(x - row * hm);
            K[u] = (int)tail_div(row, L.mn);
            J[u] = (int)(row - (uint32_t)K[u] * (uint32_t)g.n);
            I[u] = 2 * h + ((colour + J[u] + K[u]) & 1);
            if (I[u] >= g.m) continue;
            act[u] = true;
            const uint32_t o = tail_off(g, I[u], J[u], K[u]);
            T s = L.r[o];
            s = s - L.A.cE[o] * L.e[tail_off(g, I[u] + 1, J[u], K[u])];
            s = s - L.A.cW[o] * L.e[tail_off(g, I[u] - 1, J[u], K[u])];
            s = s - L.A.cN[o] * L.e[tail_off(g, I[u], J[u] + 1, K[u])];
            s = s - L.A.cS[o] * L.e[tail_off(g, I[u], J[u] - 1, K[u])];
            if (g.dim == 3) {
                s = s - L.A.cT[o] * L.e[tail_off(g, I[u], J[u], K[u] + 1)];
                s = s - L.A.cB[o] * L.e[tail_off(g, I[u], J[u], K[u] - 1)];
            }
            v[u] = s / L.A.ap[o];
        }
#pragma unroll
        for (int u = 0; u < MG_TAIL_U; ++u)
            if (act[u]) tail_store(g, L.e, I[u], J[u], K[u], v[u], g.fuse_y, g.fuse_z);
    }
    __syncthreads();
    tail_refresh(L, L.e);
}

template <typename T>
__device__ __forceinline__ T tail_ax(const MgGeom &g, const MgOp<T> &A, const T *__restrict__ e, int I, int J, int K,
                                     uint32_t o) {
    T s = A.ap[o] * e[o];
    s = s + A.cE[o] * e[tail_off(g, I + 1, J, K)];
    s = s + A.cW[o] * e[tail_off(g, I - 1, J, K)];
    s = s + A.cN[o] * e[tail_off(g, I, J + 1, K)];
    s = s + A.cS[o] * e[tail_off(g, I, J - 1, K)];
    if (g.dim == 3) {
        s = s + A.cT[o] * e[tail_off(g, I, J, K + 1)];
        s = s + A.cB[o] * e[tail_off(g, I, J, K - 1)];
    }
    return s;
}

template <typename T>
__device__ __forceinline__ void tail_restrict(const MgTailLevel<T> &F, const MgTailLevel<T> &C) {
    const MgGeom &f = F.g, &c = C.g;
    const uint32_t cells = (uint32_t)c.m * c.n * c.l;
    for (uint32_t base = threadIdx.x; base < cells; base += (uint32_t)MG_TAIL_THREADS * MG_TAIL_U) {
        T v[MG_TAIL_U];
        int I[MG_TAIL_U], J[MG_TAIL_U], K[MG_TAIL_U];
#pragma unroll
        for (int u = 0; u < MG_TAIL_U; ++u) {
            const uint32_t q = base + (uint32_t)u * MG_TAIL_THREADS;
            v[u] = T(0);
            if (q >= cells) continue;
            tail_cell(C, q, I[u], J[u], K[u]);
            T sum = T(0);
            for (int dz = 0; dz < (f.dim == 3 ? 2 : 1); ++dz)
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx) {
                        const int i = 2 * I[u] + dx, j = 2 * J[u] + dy, k = (f.dim == 3) ? 2 * K[u] + dz : 0;
                        if (i >= f.m || j >= f.n || k >= f.l) continue;
                        const uint32_t o = tail_off(f, i, j, k);
                        if (!F.keep[o]) continue;
                        sum = sum + (F.r[o] - tail_ax(f, F.A, F.e, i, j, k, o));
                    }
            v[u] = sum;
        }
#pragma unroll
        for (int u = 0; u < MG_TAIL_U; ++u) {
            const uint32_t q = base + (uint32_t)u * MG_TAIL_THREADS;
            if (q >= cells) continue;
            C.r[tail_off(c, I[u], J[u], K[u])] = v[u];
            tail_store(c, C.e, I[u], J[u], K[u], T(0), c.per, c.per);
        }
    }
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ void tail_prolong(const MgTailLevel<T> &F, const MgTailLevel<T> &C, T alpha) {
    const MgGeom &f = F.g, &c = C.g;
    const uint32_t cells = (uint32_t)f.m * f.n * f.l;
    for (uint32_t base = threadIdx.x; base < cells; base += (uint32_t)MG_TAIL_THREADS * MG_TAIL_U) {
        T v[MG_TAIL_U];
        int I[MG_TAIL_U], J[MG_TAIL_U], K[MG_TAIL_U];
        bool act[MG_TAIL_U];
#pragma unroll
        for (int u = 0; u < MG_TAIL_U; ++u) {
            const uint32_t q = base + (uint32_t)u * MG_TAIL_THREADS;
            act[u] = false;
            v[u] = T(0);
            if (q >= cells) continue;
            tail_cell(F, q, I[u], J[u], K[u]);
            const uint32_t o = tail_off(f, I[u], J[u], K[u]);
            if (!F.keep[o]) continue;
            act[u] = true;
            v[u] = F.e[o] + alpha * C.e[tail_off(c, I[u] >> 1, J[u] >> 1, (f.dim == 3) ? K[u] >> 1 : 0)];
        }
#pragma unroll
        for (int u = 0; u < MG_TAIL_U; ++u)
            if (act[u]) tail_store(f, F.e, I[u], J[u], K[u], v[u], f.fuse_y, f.fuse_z);
    }
    __syncthreads();
    tail_refresh(F, F.e);
}

// e = A^-1 r on the coarsest level (n <= MG_COARSEST_CELLS cells).
template <typename T>
__device__ __forceinline__ void tail_coarsest(const MgTailLevel<T> &L, const T *__restrict__ cinv) {
    __shared__ T rl[MG_COARSEST_CELLS];
    const MgGeom &g = L.g;
    const int n = g.m * g.n * g.l;
    for (int q = threadIdx.x; q < n; q += MG_TAIL_THREADS) {
        int I, J, K;
        tail_cell(L, (uint32_t)q, I, J, K);
        rl[q] = L.r[tail_off(g, I, J, K)];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < n; q += MG_TAIL_THREADS) {
        T s = T(0);
        for (int j = 0; j < n; ++j) s = s + cinv[(uint32_t)j * (uint32_t)n + (uint32_t)q] * rl[j];
        int I, J, K;
        tail_cell(L, (uint32_t)q, I, J, K);
        tail_store(g, L.e, I, J, K, s, g.fuse_y, g.fuse_z);
    }
    __syncthreads();
    tail_refresh(L, L.e);
}

// The cycle from tail level Q down: pre-smoothing, restriction, `gamma` cycles of the next
// level (1: V-cycle, 2: W-cycle; a later visit continues from the correction the earlier
// one left), prolongation, post-smoothing.  Unrolled over the level depth at compile time.
template <typename T, int Q>
__device__ __forceinline__ void tail_cycle(const MgTailArgs<T> &t) {
    if constexpr (Q < MG_TAIL_LEVELS) {
        if (Q == t.nl - 1) {
            tail_coarsest(t.L[Q], t.cinv);
            return;
        }
        if constexpr (Q + 1 < MG_TAIL_LEVELS) {
            for (int s = 0; s < t.nuc; ++s) {
                tail_smooth(t.L[Q], 0);
                tail_smooth(t.L[Q], 1);
            }
            tail_restrict(t.L[Q], t.L[Q + 1]);
            const MgGeom &cn = t.L[Q + 1].g;
            const bool w = Q + 2 < t.nl && (int64_t)cn.m * cn.n * cn.l <= MG_W_CELLS;
            for (int c = 0; c < (w ? t.gamma : 1); ++c) tail_cycle<T, Q + 1>(t);
            tail_prolong(t.L[Q], t.L[Q + 1], t.alpha);
            for (int s = 0; s < t.nuc; ++s) {
                tail_smooth(t.L[Q], 1);
                tail_smooth(t.L[Q], 0);
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(MG_TAIL_THREADS) void k_mg_tail(MgTailArgs<T> t) {
    tail_cycle<T, 0>(t);
}

// Every quotient the tail takes on a level (cell index by m, by (m + 1) / 2 and the row by
// n, the refresh's index by m) equals true division.
static bool tail_magic_ok(const MgGeom &g) {
    const uint32_t m = (uint32_t)g.m, n = (uint32_t)g.n, hm = (m + 1) / 2;
    const uint32_t cells = m * n * (uint32_t)g.l;
    const uint32_t mm = tail_magic(m), mh = tail_magic(hm), mn = tail_magic(n);
    const uint32_t rows = n * (uint32_t)g.l;
    const uint32_t span = cells > hm * rows ? cells : hm * rows;
    for (uint32_t x = 0; x < span; ++x) {
        if (tail_div(x, mm) != x / m) return false;
        if (tail_div(x, mh) != x / hm) return false;
    }
    for (uint32_t r = 0; r < rows; ++r)
        if (tail_div(r, mn) != r / n) return false;
    return true;
}

// ------------------------------------------------------------------------------------
// Set-up: Galerkin coarse operators.  For each kept child of an aggregate its diagonal and
// its couplings to kept cells of the same aggregate go to the aggregate's diagonal, and its
// couplings to kept cells of the neighbouring aggregate in a direction go to that
// direction's coupling; couplings to rows left out of the coarse space are dropped.
// ------------------------------------------------------------------------------------
template <typename T>
struct MgOut {
    T *ap, *cE, *cW, *cN, *cS, *cT, *cB;
    uint8_t *keep;
};

template <typename T>
__device__ __forceinline__ void mg_write_row(const MgGeom &c, const MgOut<T> &out, int64_t o, bool any, T sap, T sE,
                                             T sW, T sN, T sS, T sT, T sB) {
    if (!any) {   // an aggregate of left-out rows only: a unit row with zero right-hand side
        out.ap[o] = T(1);
        out.cE[o] = out.cW[o] = out.cN[o] = out.cS[o] = out.cT[o] = out.cB[o] = T(0);
        out.keep[o] = 0;
        return;
    }
    const bool coupled = sE != T(0) || sW != T(0) || sN != T(0) || sS != T(0) || sT != T(0) || sB != T(0);
    out.ap[o] = (sap != T(0)) ? sap : T(1);
    out.cE[o] = sE;
    out.cW[o] = sW;
    out.cN[o] = sN;
    out.cS[o] = sS;
    out.cT[o] = sT;
    out.cB[o] = sB;
    out.keep[o] = coupled ? 1 : 0;
}

template <typename T, int DIM, bool FAST>
__global__ __launch_bounds__(256) void k_mg_gal0(PfGeom g, StepArgs a, MgK<T> K, const T *F, const double *phi,
                                                 MgGeom c, MgOut<T> out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)c.m * c.n * c.l) return;
    int I, J, Kc;
    mg_cell(c, q, I, J, Kc);
    const bool per = (a.bc == PF_BC_UNIFORM);
    bool any = false;
    T sap = T(0), sE = T(0), sW = T(0), sN = T(0), sS = T(0), sT = T(0), sB = T(0);
    for (int dz = 0; dz < (DIM == 3 ? 2 : 1); ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int i = 2 * I + dx + 1, j = 2 * J + dy + 1, k = (DIM == 3) ? 2 * Kc + dz + 1 : 0;
                if (i > g.m || j > g.n || (DIM == 3 && k > g.l)) continue;
                if (dec0(g, a, i, j, k, phi)) continue;
                any = true;
                CoefT<T> cc;
                sap = sap + op0<T, DIM, FAST>(g, a, K, F, phi, i, j, k, cc);
                const T vals[6] = {cc.ae, cc.aw, cc.an, cc.as, cc.at, cc.ab};
                const int di[6] = {1, -1, 0, 0, 0, 0}, dj[6] = {0, 0, 1, -1, 0, 0}, dk[6] = {0, 0, 0, 0, 1, -1};
                for (int d = 0; d < (DIM == 3 ? 6 : 4); ++d) {
                    const T v = vals[d];
                    if (v == T(0)) continue;
                    const int ni = i + di[d];
                    int nj = j + dj[d], nk = k + dk[d];
                    if (per) {
                        if (nj == 0) nj = g.n; else if (nj == g.n + 1) nj = 1;
                        if (DIM == 3) { if (nk == 0) nk = g.l; else if (nk == g.l + 1) nk = 1; }
                    }
                    if (ni < 1 || ni > g.m || nj < 1 || nj > g.n || (DIM == 3 && (nk < 1 || nk > g.l))) continue;
                    if (dec0(g, a, ni, nj, nk, phi)) continue;
                    const bool same = ((ni - 1) >> 1) == I && ((nj - 1) >> 1) == J &&
                                      (DIM != 3 || ((nk - 1) >> 1) == Kc);
                    if (same) sap = sap + v;
                    else if (d == 0) sE = sE + v;
                    else if (d == 1) sW = sW + v;
                    else if (d == 2) sN = sN + v;
                    else if (d == 3) sS = sS + v;
                    else if (d == 4) sT = sT + v;
                    else sB = sB + v;
                }
            }
    mg_write_row(c, out, mg_off(c, I, J, Kc), any, sap, sE, sW, sN, sS, sT, sB);
}

template <typename T>
__global__ __launch_bounds__(256) void k_mg_gal(MgGeom f, MgOp<T> A, const uint8_t *__restrict__ keep, MgGeom c,
                                                MgOut<T> out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)c.m * c.n * c.l) return;
    int I, J, K;
    mg_cell(c, q, I, J, K);
    bool any = false;
    T sap = T(0), sE = T(0), sW = T(0), sN = T(0), sS = T(0), sT = T(0), sB = T(0);
    for (int dz = 0; dz < (f.dim == 3 ? 2 : 1); ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int i = 2 * I + dx, j = 2 * J + dy, k = (f.dim == 3) ? 2 * K + dz : 0;
                if (i >= f.m || j >= f.n || k >= f.l) continue;
                const int64_t o = mg_off(f, i, j, k);
                if (!keep[o]) continue;
                any = true;
                sap = sap + A.ap[o];
                const T vals[6] = {A.cE[o], A.cW[o], A.cN[o], A.cS[o], A.cT[o], A.cB[o]};
                const int di[6] = {1, -1, 0, 0, 0, 0}, dj[6] = {0, 0, 1, -1, 0, 0}, dk[6] = {0, 0, 0, 0, 1, -1};
                for (int d = 0; d < (f.dim == 3 ? 6 : 4); ++d) {
                    const T v = vals[d];
                    if (v == T(0)) continue;
                    const int ni = i + di[d];
                    int nj = j + dj[d], nk = k + dk[d];
                    if (f.per) {
                        if (nj < 0) nj = f.n - 1; else if (nj >= f.n) nj = 0;
                        if (f.dim == 3) { if (nk < 0) nk = f.l - 1; else if (nk >= f.l) nk = 0; }
                    }
                    if (ni < 0 || ni >= f.m || nj < 0 || nj >= f.n || nk < 0 || nk >= f.l) continue;
                    if (!keep[mg_off(f, ni, nj, nk)]) continue;
                    const bool same = (ni >> 1) == I && (nj >> 1) == J && (f.dim != 3 || (nk >> 1) == K);
                    if (same) sap = sap + v;
                    else if (d == 0) sE = sE + v;
                    else if (d == 1) sW = sW + v;
                    else if (d == 2) sN = sN + v;
                    else if (d == 3) sS = sS + v;
                    else if (d == 4) sT = sT + v;
                    else sB = sB + v;
                }
            }
    mg_write_row(c, out, mg_off(c, I, J, K), any, sap, sE, sW, sN, sS, sT, sB);
}

// ------------------------------------------------------------------------------------
// BiCGStab vector kernels over every slot of the colour-split arrays (ghost and padding
// slots of r, rh, p, v, s, t are zero from allocation on and stay zero, so sums over all
// slots are sums over the interior cells), and the scalar updates.
// Scalars: S[0] rho, S[1] alpha, S[2] omega, S[3] beta, S[4] |b|^2, S[5] |r|^2, S[6] stop.
// ------------------------------------------------------------------------------------
__device__ __forceinline__ void block_sum2(double a0, double a1, double *part) {
    __shared__ double red[2][256];
    const int tid = (int)threadIdx.x;
    red[0][tid] = a0;
    red[1][tid] = a1;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) {
            red[0][tid] += red[0][tid + off];
            red[1][tid] += red[1][tid + off];
        }
        __syncthreads();
    }
    if (tid == 0) {
        part[(int64_t)blockIdx.x * 4 + 0] = red[0][0];
        part[(int64_t)blockIdx.x * 4 + 1] = red[1][0];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bicg_init(int64_t n, const T *__restrict__ r, T *__restrict__ rh,
                                                   T *__restrict__ p, T *__restrict__ v, T *__restrict__ x,
                                                   double *part) {
    double d = 0.;
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (int64_t)gridDim.x * 256) {
        const T b = r[q];
        rh[q] = b;
        p[q] = T(0);
        v[q] = T(0);
        x[q] = T(0);
        d += (double)b * (double)b;
    }
    block_sum2(d, 0., part);
}

// The two updates without a reduction take one element per thread (the launch covers n):
// a grid-stride loop over MG_NPART blocks streams at 4.5-5 TB/s where this form reaches a
// plain copy's 6.1-6.5 TB/s.
template <typename T>
__global__ __launch_bounds__(256) void k_bicg_p(int64_t n, const T *__restrict__ r, T *__restrict__ p,
                                                const T *__restrict__ v, const double *S) {
    const T beta = (T)S[3], omega = (T)S[2];
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q < n) p[q] = r[q] + beta * (p[q] - omega * v[q]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_bicg_s(int64_t n, const T *__restrict__ r, const T *__restrict__ v,
                                                T *__restrict__ s, const double *S) {
    const T alpha = (T)S[1];
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q < n) s[q] = r[q] - alpha * v[q];
}

template <typename T>
__global__ __launch_bounds__(256) void k_bicg_xr(int64_t n, T *__restrict__ x, const T *__restrict__ y,
                                                 const T *__restrict__ z, T *__restrict__ r, const T *__restrict__ s,
                                                 const T *__restrict__ t, const T *__restrict__ rh, const double *S,
                                                 double *part) {
    const T alpha = (T)S[1], omega = (T)S[2];
    double d0 = 0., d1 = 0.;
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (int64_t)gridDim.x * 256) {
        x[q] = x[q] + alpha * y[q] + omega * z[q];
        const T rn = s[q] - omega * t[q];
        r[q] = rn;
        d0 += (double)rh[q] * (double)rn;
        d1 += (double)rn * (double)rn;
    }
    block_sum2(d0, d1, part);
}

__device__ __forceinline__ void sum_parts(const double *part, int nb, double &a0, double &a1) {
    __shared__ double red[2][256];
    const int tid = (int)threadIdx.x;
    double s0 = 0., s1 = 0.;
    for (int b = tid; b < nb; b += 256) {
        s0 += part[(int64_t)b * 4 + 0];
        s1 += part[(int64_t)b * 4 + 1];
    }
    red[0][tid] = s0;
    red[1][tid] = s1;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) {
            red[0][tid] += red[0][tid + off];
            red[1][tid] += red[1][tid + off];
        }
        __syncthreads();
    }
    a0 = red[0][0];
    a1 = red[1][0];
}

__device__ __forceinline__ bool finite_nonzero(double x) { return x != 0. && fabs(x) <= 1.7976931348623157e308; }

// stage 0: after k_bicg_init; 1: alpha after A y; 2: omega after A z; 3: beta after x, r.
__global__ __launch_bounds__(256) void k_bicg_scalar(const double *part, int nb, double *S, int stage) {
    double a0, a1;
    sum_parts(part, nb, a0, a1);
    if (threadIdx.x != 0) return;
    if (stage == 0) {
        S[0] = a0;
        S[4] = a0;
        S[5] = a0;
        S[1] = 1.;
        S[2] = 1.;
        S[3] = 0.;
        S[6] = (a0 == 0. || !(a0 <= 1.7976931348623157e308)) ? 1. : 0.;
    } else if (stage == 1) {
        if (finite_nonzero(a0)) {
            S[1] = S[0] / a0;
        } else {
            S[1] = 0.;
            S[6] = 1.;
        }
    } else if (stage == 2) {
        S[2] = (finite_nonzero(a1) && a0 <= 1.7976931348623157e308 && a0 >= -1.7976931348623157e308) ? a0 / a1 : 0.;
    } else {
        const double rho = S[0], omega = S[2];
        if (!finite_nonzero(omega) || !finite_nonzero(rho) || !(fabs(a0) <= 1.7976931348623157e308)) {
            S[3] = 0.;
            S[6] = 1.;
        } else {
            S[3] = (a0 / rho) * (S[1] / omega);
        }
        S[0] = a0;
        S[5] = a1;
    }
}

// One element per thread (the launch covers n).
__global__ void k_mg_zero_t(float *p, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q < n) p[q] = 0.f;
}
__global__ void k_mg_zero_t(double *p, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q < n) p[q] = 0.;
}

// ------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------
static inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }
static inline unsigned flat_blocks(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < MG_NPART ? (b < 1 ? 1 : b) : MG_NPART);
}

static std::vector<MgGeom> mg_levels(const PfGeom &g, int bc) {
    const int per = (bc == PF_BC_UNIFORM);
    std::vector<MgGeom> L;
    L.push_back(mg_make_geom(g.dim, g.m, g.n, g.l, per));
    for (;;) {
        const MgGeom &f = L.back();
        const int64_t cells = (int64_t)f.m * f.n * f.l;
        if (L.size() > 1 && cells <= MG_COARSEST_CELLS) break;
        const int mc = mg_coarse_extent(f.m), nc = mg_coarse_extent(f.n);
        const int lc = (g.dim == 3) ? mg_coarse_extent(f.l) : 1;
        if (mc == f.m && nc == f.n && lc == f.l) break;
        if ((int)L.size() == MG_MAX_LEVELS) break;
        L.push_back(mg_make_geom(g.dim, mc, nc, lc, per));
    }
    return L;
}

static inline int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

static int mg_nslots(const PfGeom &g) { return sor_blocks(g); }

static int64_t mg_bytes_t(const PfGeom &g, int bc, int64_t tsz) {
    const std::vector<MgGeom> L = mg_levels(g, bc);
    int64_t b = 9 * align256(2 * g.HS * tsz);
    for (size_t q = 1; q < L.size(); ++q) b += 9 * align256(L[q].N * tsz) + align256(L[q].N);
    b += align256(16 * 8) + align256((int64_t)MG_NPART * 4 * 8);
    b += 2 * align256((int64_t)mg_nslots(g) * 8) + align256(8);
    if (L.size() > 1) {
        const int64_t n = (int64_t)L.back().m * L.back().n * L.back().l;
        b += align256(n * n * tsz);
    }
    return b;
}

int64_t mg_bytes(const PfGeom &g, int bc, bool fp32) { return mg_bytes_t(g, bc, fp32 ? 4 : 8); }

template <typename T>
bool mg_create(MgSolver<T> &S, const MgCase &c, std::string *why, void *mem) {
    S.c = c;
    const PfGeom &g = c.g;
    const std::vector<MgGeom> L = mg_levels(g, c.a.bc);
    if (L.size() > 1 && L.back().N > MG_LDS) {
        if (why) *why = "multigrid: coarsest level too large for the LDS solve";
        return false;
    }
    if (L.size() > 1 && (int64_t)L.back().m * L.back().n * L.back().l > 256 * MG_CPT) {
        if (why) *why = "multigrid: coarsest level has more cells than one block holds";
        return false;
    }
    {
        const char *v;
        if ((v = std::getenv("PF_MG_NU")) && *v) S.prm.nu1 = S.prm.nu2 = std::max(1, std::atoi(v));
        if ((v = std::getenv("PF_MG_NUC")) && *v) S.prm.nuc = std::max(1, std::atoi(v));
        if ((v = std::getenv("PF_MG_CS")) && *v) S.prm.coarsest_sweeps = std::max(1, std::atoi(v));
        if ((v = std::getenv("PF_MG_ALPHA")) && *v) S.prm.alpha = std::atof(v);
        if ((v = std::getenv("PF_MG_GAMMA")) && *v) S.prm.tail_gamma = std::min(3, std::max(1, std::atoi(v)));
    }
    S.bytes = mg_bytes_t(g, c.a.bc, sizeof(T));
    S.own_mem = (mem == nullptr);
    if (S.own_mem) {
        if (hipMalloc(&S.mem, (size_t)S.bytes) != hipSuccess) {
            S.mem = nullptr;
            if (why) *why = "multigrid: hipMalloc of " + std::to_string(S.bytes) + " bytes failed";
            return false;
        }
        if (hipMemset(S.mem, 0, (size_t)S.bytes) != hipSuccess) {
            mg_destroy(S);
            if (why) *why = "multigrid: hipMemset failed";
            return false;
        }
    } else {
        S.mem = mem;
    }
    if (hipHostMalloc((void **)&S.h_scal, 16 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        S.h_scal = nullptr;
        mg_destroy(S);
        if (why) *why = "multigrid: hipHostMalloc failed";
        return false;
    }
    if (hipEventCreateWithFlags(&S.ev, hipEventDisableTiming) != hipSuccess) {
        S.ev = nullptr;
        mg_destroy(S);
        if (why) *why = "multigrid: hipEventCreate failed";
        return false;
    }
    char *p = (char *)S.mem;
    auto take = [&](int64_t bytes) {
        char *q = p;
        p += align256(bytes);
        return (void *)q;
    };
    const int64_t n0 = 2 * g.HS * (int64_t)sizeof(T);
    T **vecs[9] = {&S.x, &S.r, &S.rh, &S.p, &S.v, &S.s, &S.t, &S.y, &S.z};
    for (T **v : vecs) *v = (T *)take(n0);
    S.lev.assign(L.size(), MgLevel<T>{});
    S.lev[0].g = L[0];
    for (size_t q = 1; q < L.size(); ++q) {
        MgLevel<T> &lv = S.lev[q];
        lv.g = L[q];
        const int64_t nb = L[q].N * (int64_t)sizeof(T);
        T **arr[9] = {&lv.e, &lv.r, &lv.ap, &lv.cE, &lv.cW, &lv.cN, &lv.cS, &lv.cT, &lv.cB};
        for (T **a : arr) *a = (T *)take(nb);
        lv.keep = (uint8_t *)take(L[q].N);
    }
    S.scal = (double *)take(16 * 8);
    S.part = (double *)take((int64_t)MG_NPART * 4 * 8);
    S.nslots = mg_nslots(g);
    S.res_slots = (unsigned long long *)take((int64_t)S.nslots * 8);
    S.err_slots = (unsigned long long *)take((int64_t)S.nslots * 8);
    S.res_val = (unsigned long long *)take(8);
    S.cinv = nullptr;
    S.tail_q = 0;
    if (L.size() > 1) {
        const int64_t n = (int64_t)L.back().m * L.back().n * L.back().l;
        S.cinv = (T *)take(n * n * (int64_t)sizeof(T));
        // the tail starts at the first level of at most MG_TAIL_START_CELLS cells (PF_MG_TAIL=0
        // in the environment keeps one launch per level and step)
        const char *tv = std::getenv("PF_MG_TAIL");
        const bool tail_on = !(tv && *tv && std::atoi(tv) == 0);
        int64_t tail_cells = MG_TAIL_START_CELLS;
        if (tv && *tv && std::atoi(tv) > 1) tail_cells = std::min<int64_t>(MG_TAIL_CELLS, std::atoi(tv));
        for (size_t q = 1; tail_on && q < L.size(); ++q)
            if ((int64_t)L[q].m * L[q].n * L[q].l <= tail_cells) {
                bool ok = L.size() - q <= (size_t)MG_TAIL_LEVELS && n <= MG_COARSEST_CELLS;
                for (size_t k = q; ok && k < L.size(); ++k) ok = tail_magic_ok(L[k]);
                if (ok) S.tail_q = (int)q;
                break;
            }
    }
    S.cinv_ok = false;
    S.built = false;
    return true;
}

template <typename T>
void mg_destroy(MgSolver<T> &S) {
    if (S.mem && S.own_mem) (void)hipFree(S.mem);
    if (S.h_scal) (void)hipHostFree(S.h_scal);
    if (S.ev) (void)hipEventDestroy(S.ev);
    S.mem = nullptr;
    S.h_scal = nullptr;
    S.ev = nullptr;
    S.lev.clear();
    S.built = false;
}

template <typename T>
static MgK<T> mg_k(const StepArgs &a) {
    MgK<T> K;
    K.cdx = (T)(float)(a.dt / (a.dx * a.dx));
    K.cdy = (T)(float)(a.dt / (a.dy * a.dy));
    K.cdz = (T)(float)(a.dt / (a.dz * a.dz));
    return K;
}

template <typename T>
static const T *fine_phi(const MgSolver<T> &S);
template <>
const float *fine_phi<float>(const MgSolver<float> &S) { return S.c.PHI32; }
template <>
const double *fine_phi<double>(const MgSolver<double> &S) { return S.c.PHIs; }
// The unit-porosity chunk flags of that porosity for the level-0 operator kernels, where
// unit-porosity waves hold half the cells or more (nullptr: none).  Measured per step to
// the rule on one box, flags on / off: sphere 256x128x128 FP64 9.23 / 9.36 ms, boxes
// 512x256x256 FP64 71.9 / 73.8 ms, room 192^3 (fewer unit cells) 23.3 / 22.7 ms.
// PF_MG_UNIT=0/1 in the environment overrides the rule.
template <typename T>
static const uint8_t *fine_unit(const MgSolver<T> &S) {
    static const int forced = [] {
        const char *v = std::getenv("PF_MG_UNIT");
        return (v && *v) ? (std::atoi(v) != 0 ? 1 : 0) : -1;
    }();
    if (forced == 0 || (forced < 0 && !S.c.unit_ops)) return nullptr;
    return std::is_same<T, float>::value ? S.c.UNIT32 : S.c.UNIT;
}

template <typename T>
static MgOp<T> op_of(const MgLevel<T> &L) {
    return MgOp<T>{L.ap, L.cE, L.cW, L.cN, L.cS, L.cT, L.cB};
}

// The coarsest operator as a dense matrix (cells in mg_cell order; a periodic neighbour
// wraps, a non-periodic ghost is zero), inverted by Gauss-Jordan elimination with partial
// pivoting in FP64 and stored transposed in S.cinv.  False (the tail is not used) when a
// pivot is below 1e-12 of the largest entry or a HIP call fails.
template <typename T>
static bool coarsest_inverse(MgSolver<T> &S, hipStream_t st) {
    const MgLevel<T> &L = S.lev.back();
    const MgGeom &g = L.g;
    const int n = g.m * g.n * g.l;
    if (hipStreamSynchronize(st) != hipSuccess) return false;
    std::vector<T> h[7];
    const T *src[7] = {L.ap, L.cE, L.cW, L.cN, L.cS, L.cT, L.cB};
    for (int a = 0; a < 7; ++a) {
        h[a].resize((size_t)g.N);
        if (hipMemcpy(h[a].data(), src[a], (size_t)g.N * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) return false;
    }
    auto idx = [&](int I, int J, int K) { return I + g.m * (J + g.n * K); };
    std::vector<double> A((size_t)n * n, 0.0), X((size_t)n * n, 0.0);
    double amax = 0.0;
    for (int q = 0; q < n; ++q) {
        int I, J, K;
        mg_cell(g, q, I, J, K);
        const int64_t o = mg_off(g, I, J, K);
        double *row = &A[(size_t)q * n];
        row[q] += (double)h[0][(size_t)o];
        if (I + 1 < g.m) row[idx(I + 1, J, K)] += (double)h[1][(size_t)o];
        if (I > 0) row[idx(I - 1, J, K)] += (double)h[2][(size_t)o];
        if (J + 1 < g.n) row[idx(I, J + 1, K)] += (double)h[3][(size_t)o];
        else if (g.per) row[idx(I, 0, K)] += (double)h[3][(size_t)o];
        if (J > 0) row[idx(I, J - 1, K)] += (double)h[4][(size_t)o];
        else if (g.per) row[idx(I, g.n - 1, K)] += (double)h[4][(size_t)o];
        if (g.dim == 3) {
            if (K + 1 < g.l) row[idx(I, J, K + 1)] += (double)h[5][(size_t)o];
            else if (g.per) row[idx(I, J, 0)] += (double)h[5][(size_t)o];
            if (K > 0) row[idx(I, J, K - 1)] += (double)h[6][(size_t)o];
            else if (g.per) row[idx(I, J, g.l - 1)] += (double)h[6][(size_t)o];
        }
        X[(size_t)q * n + q] = 1.0;
    }
    for (double v : A) amax = std::max(amax, std::fabs(v));
    if (!(amax > 0.0) || !std::isfinite(amax)) return false;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)piv * n + c])) piv = r;
        const double pv = A[(size_t)piv * n + c];
        if (!(std::fabs(pv) > 1e-12 * amax)) return false;
        if (piv != c)
            for (int k = 0; k < n; ++k) {
                std::swap(A[(size_t)piv * n + k], A[(size_t)c * n + k]);
                std::swap(X[(size_t)piv * n + k], X[(size_t)c * n + k]);
            }
        const double inv = 1.0 / pv;
        for (int k = 0; k < n; ++k) {
            A[(size_t)c * n + k] *= inv;
            X[(size_t)c * n + k] *= inv;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = A[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                A[(size_t)r * n + k] -= f * A[(size_t)c * n + k];
                X[(size_t)r * n + k] -= f * X[(size_t)c * n + k];
            }
        }
    }
    std::vector<T> t((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const double v = X[(size_t)i * n + j];
            if (!std::isfinite(v)) return false;
            t[(size_t)j * n + i] = (T)v;
        }
    return hipMemcpy(S.cinv, t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}

template <typename T>
void mg_setup(MgSolver<T> &S, hipStream_t st) {
    const PfGeom &g = S.c.g;
    const MgK<T> K = mg_k<T>(S.c.a);
    const T *F = fine_phi(S);
    for (size_t q = 1; q < S.lev.size(); ++q) {
        MgLevel<T> &c = S.lev[q];
        const MgOut<T> out{c.ap, c.cE, c.cW, c.cN, c.cS, c.cT, c.cB, c.keep};
        const unsigned nb = blocks_for((int64_t)c.g.m * c.g.n * c.g.l);
        if (q == 1) {
#define PF_GAL0(D, F_) hipLaunchKernelGGL((k_mg_gal0<T, D, F_>), dim3(nb), dim3(256), 0, st, g, S.c.a, K, F, S.c.phi, c.g, out)
            if (g.dim == 3) { if (S.c.fast) PF_GAL0(3, true); else PF_GAL0(3, false); }
            else { if (S.c.fast) PF_GAL0(2, true); else PF_GAL0(2, false); }
#undef PF_GAL0
        } else {
            const MgLevel<T> &f = S.lev[q - 1];
            hipLaunchKernelGGL((k_mg_gal<T>), dim3(nb), dim3(256), 0, st, f.g, op_of(f), (const uint8_t *)f.keep, c.g,
                               out);
        }
    }
    S.cinv_ok = false;
    if (S.tail_q > 0) S.cinv_ok = coarsest_inverse(S, st);
    S.built = true;
}

// One red/black sweep colour of level 0 with omega = 1 (the solve's own colour-sweep kernels).
static void sweep0(const MgSolver<float> &S, hipStream_t st, const StepArgs &a1, int colour, float *e,
                   const float *rhs) {
    const PfGeom &g = S.c.g;
    const MgGeom &l0 = S.lev[0].g;
    SweepArgs sw{};
    sw.colour = colour;
    sw.fuse_y = l0.fuse_y;
    sw.fuse_z = l0.fuse_z;
    sw.track_err = 0;
    sw.xcd_remap = S.c.remap32;
    sw.kb = S.c.kb32;
    sw.xw = S.c.xw;
    launch_sor32(st, g, a1, sw, e, S.c.PHI32, rhs, S.c.phi, nullptr, S.c.UNIT32);
    if (l0.ref_y || l0.ref_z) launch_ghost_refresh32(st, g, l0.ref_y, l0.ref_z, e);
}
static void sweep0(const MgSolver<double> &S, hipStream_t st, const StepArgs &a1, int colour, double *e,
                   const double *rhs) {
    const PfGeom &g = S.c.g;
    const MgGeom &l0 = S.lev[0].g;
    SweepArgs sw{};
    sw.colour = colour;
    sw.fuse_y = l0.fuse_y;
    sw.fuse_z = l0.fuse_z;
    sw.track_err = 0;
    sw.xcd_remap = S.c.remap64;
    sw.kb = S.c.kb64;
    sw.xw = S.c.xw;
    launch_sor_phi(st, g, a1, sw, e, S.c.PHIs, rhs, S.c.phi, nullptr, S.c.fast, S.c.coded ? S.c.PHC : nullptr,
                   S.c.PHTAB, S.c.UNIT);
    if (l0.ref_y || l0.ref_z) launch_ghost_refresh(st, g, l0.ref_y, l0.ref_z, e);
}
static void refresh0(const MgSolver<float> &S, hipStream_t st, float *e) {
    const MgGeom &l0 = S.lev[0].g;
    if (l0.ref_y || l0.ref_z) launch_ghost_refresh32(st, S.c.g, l0.ref_y, l0.ref_z, e);
}
static void refresh0(const MgSolver<double> &S, hipStream_t st, double *e) {
    const MgGeom &l0 = S.lev[0].g;
    if (l0.ref_y || l0.ref_z) launch_ghost_refresh(st, S.c.g, l0.ref_y, l0.ref_z, e);
}

template <typename T>
static void coarse_cycle(MgSolver<T> &S, hipStream_t st, size_t q) {
    MgLevel<T> &L = S.lev[q];
    const MgGeom &g = L.g;
    const int64_t cells = (int64_t)g.m * g.n * g.l;
    if (S.tail_q > 0 && q == (size_t)S.tail_q && S.cinv_ok) {
        MgTailArgs<T> t{};
        t.nl = (int)(S.lev.size() - q);
        for (int k = 0; k < t.nl; ++k) {
            MgLevel<T> &lv = S.lev[q + (size_t)k];
            const uint32_t m = (uint32_t)lv.g.m;
            t.L[k] = MgTailLevel<T>{lv.g, op_of(lv), (const uint8_t *)lv.keep, lv.r, lv.e, tail_magic(m),
                                    tail_magic((m + 1) / 2), tail_magic((uint32_t)lv.g.n)};
        }
        t.nuc = S.prm.nuc;
        t.gamma = S.prm.tail_gamma;
        t.alpha = (T)S.prm.alpha;
        t.cinv = S.cinv;
        hipLaunchKernelGGL((k_mg_tail<T>), dim3(1), dim3(MG_TAIL_THREADS), 0, st, t);
        return;
    }
    if (q + 1 == S.lev.size()) {
        hipLaunchKernelGGL((k_mg_coarsest<T>), dim3(1), dim3(256), 0, st, g, op_of(L), (const T *)L.r, L.e,
                           S.prm.coarsest_sweeps);
        return;
    }
    const unsigned nb = blocks_for(cells);
    const unsigned nbr = blocks_for((int64_t)g.m * (g.l > g.n ? g.l : g.n));
    const unsigned nbs = blocks_for((int64_t)((g.m + 1) / 2) * g.n * g.l);
    auto smooth = [&](int colour) {
        hipLaunchKernelGGL((k_mg_smooth<T>), dim3(nbs), dim3(256), 0, st, g, op_of(L), (const T *)L.r, L.e, colour);
        if (g.ref_y || g.ref_z) hipLaunchKernelGGL((k_mg_refresh<T>), dim3(nbr), dim3(256), 0, st, g, L.e);
    };
    for (int s = 0; s < S.prm.nuc; ++s) {
        smooth(0);
        smooth(1);
    }
    MgLevel<T> &C = S.lev[q + 1];
    hipLaunchKernelGGL((k_mg_restrict<T>), dim3(blocks_for((int64_t)C.g.m * C.g.n * C.g.l)), dim3(256), 0, st, g,
                       op_of(L), (const uint8_t *)L.keep, (const T *)L.r, (const T *)L.e, C.g, C.r, C.e);
    // a later visit continues from the correction the earlier one left
    const int64_t ccells = (int64_t)C.g.m * C.g.n * C.g.l;
    const int visits = (q + 2 < S.lev.size() && ccells <= MG_W_CELLS) ? S.prm.tail_gamma : 1;
    for (int v = 0; v < visits; ++v) coarse_cycle(S, st, q + 1);
    hipLaunchKernelGGL((k_mg_prolong<T>), dim3(nb), dim3(256), 0, st, g, (const uint8_t *)L.keep, C.g,
                       (const T *)C.e, L.e, (T)S.prm.alpha);
    if (g.ref_y || g.ref_z) hipLaunchKernelGGL((k_mg_refresh<T>), dim3(nbr), dim3(256), 0, st, g, L.e);
    for (int s = 0; s < S.prm.nuc; ++s) {
        smooth(1);
        smooth(0);
    }
}

// e = M rhs: one V-cycle from e = 0.
template <typename T>
static void vcycle(MgSolver<T> &S, hipStream_t st, const T *rhs, T *e) {
    const PfGeom &g = S.c.g;
    StepArgs a1 = S.c.a;
    a1.omega = 1.0;
    a1.omega_c = 0.0;
    hipLaunchKernelGGL(k_mg_zero_t, dim3(blocks_for(2 * g.HS)), dim3(256), 0, st, e, 2 * g.HS);
    for (int s = 0; s < S.prm.nu1; ++s) {
        sweep0(S, st, a1, 0, e, rhs);
        sweep0(S, st, a1, 1, e, rhs);
    }
    if (S.lev.size() > 1) {
        MgLevel<T> &C = S.lev[1];
        const MgK<T> K = mg_k<T>(S.c.a);
        const T *F = fine_phi(S);
        const unsigned nb = blocks_for((int64_t)C.g.m * C.g.n * C.g.l);
        // with an odd periodic extent a boundary cell's periodic neighbour has its own colour
        // and is refreshed after the sweep, so its residual is not zero: all cells count
        const int zero_colour = (S.prm.nu1 > 0 && !S.lev[0].g.ref_y && !S.lev[0].g.ref_z) ? 1 : -1;
        const uint8_t *unit = fine_unit(S);
#define PF_R0_U(D, F_, U) hipLaunchKernelGGL((k_mg_restrict0<T, D, F_, U>), dim3(nb), dim3(256), 0, st, g, S.c.a, K, F, \
                                             S.c.phi, rhs, (const T *)e, C.g, C.r, C.e, zero_colour, unit)
#define PF_R0(D, F_) do { if (unit) PF_R0_U(D, F_, true); else PF_R0_U(D, F_, false); } while (0)
        if (g.dim == 3) { if (S.c.fast) PF_R0(3, true); else PF_R0(3, false); }
        else { if (S.c.fast) PF_R0(2, true); else PF_R0(2, false); }
#undef PF_R0
#undef PF_R0_U
        coarse_cycle(S, st, 1);
        const MgGeom &l0 = S.lev[0].g;
        dim3 b(64, 4, 1);
        dim3 gr((unsigned)(2 * (((g.m + 1) / 2 + 63) / 64)), (unsigned)((g.n + 4 * PROLONG0_ROWS - 1) / (4 * PROLONG0_ROWS)),
                (unsigned)(g.dim == 3 ? g.l : 1));
        if (g.dim == 3)
            hipLaunchKernelGGL((k_mg_prolong0<T, 3>), gr, b, 0, st, g, S.c.a, S.c.phi, C.g, (const T *)C.e, e,
                               (T)S.prm.alpha, l0.fuse_y, l0.fuse_z);
        else
            hipLaunchKernelGGL((k_mg_prolong0<T, 2>), gr, b, 0, st, g, S.c.a, S.c.phi, C.g, (const T *)C.e, e,
                               (T)S.prm.alpha, l0.fuse_y, l0.fuse_z);
        refresh0(S, st, e);
    }
    for (int s = 0; s < S.prm.nu2; ++s) {
        sweep0(S, st, a1, 1, e, rhs);
        sweep0(S, st, a1, 0, e, rhs);
    }
}

template <typename T>
static unsigned apply_blocks(const PfGeom &g) {
    const int rows = g.n * (g.dim == 3 ? g.l : 1);
    const int b = (rows + 3) / 4;
    return (unsigned)(b < MG_NPART ? b : MG_NPART);
}

template <typename T, int MODE>
static void apply0(MgSolver<T> &S, hipStream_t st, const T *y, T *out, const T *w) {
    const PfGeom &g = S.c.g;
    const MgK<T> K = mg_k<T>(S.c.a);
    const T *F = fine_phi(S);
    const dim3 b(64, 4, 1), gr(apply_blocks<T>(g));
    // unit-porosity waves for the FP64 operator only: 512x256x256 FP64 360 against 395 us,
    // FP32 245 against 204 us (the flag's load ahead of the operands costs the FP32 form more
    // than the porosity bytes it saves)
    const uint8_t *unit = std::is_same<T, double>::value ? fine_unit(S) : nullptr;
#define PF_AP_U(D, F_, U) hipLaunchKernelGGL((k_mg_apply0<T, D, F_, MODE, U>), gr, b, 0, st, g, S.c.a, K, F, S.c.phi, y, \
                                             out, w, S.part, unit)
#define PF_AP(D, F_) do { if (unit) PF_AP_U(D, F_, true); else PF_AP_U(D, F_, false); } while (0)
    if (g.dim == 3) { if (S.c.fast) PF_AP(3, true); else PF_AP(3, false); }
    else { if (S.c.fast) PF_AP(2, true); else PF_AP(2, false); }
#undef PF_AP
#undef PF_AP_U
}

// ------------------------------------------------------------------------------------
// The step's solve as a state machine per solver.  After each enqueued stage the solver's
// event is recorded; the host waits for one solver's event at a time, reads what that
// stage copied to pinned memory and enqueues the next stage, so several solvers' work is
// in flight on their streams while the host deals with one.
// ------------------------------------------------------------------------------------
enum MgState { MG_RESIDUAL = 0, MG_INIT = 1, MG_ITER = 2, MG_DONE = 3 };

// The FP64 residual r = bb - A p of the reference operator into S.r (precision T) and
// max|r/ap| into h_scal[8]; the colour launches of the mixed path's residual kernel.
template <typename T>
static void enqueue_residual(MgSolver<T> &S, hipStream_t st) {
    const MgCase &c = S.c;
    for (int half = 0; half < 2; ++half) {
        SweepArgs sw{};
        sw.colour = half;
        sw.track_err = 1;
        sw.xcd_remap = c.remap32;
        launch_residual(st, c.g, c.a, sw, c.P, c.PHIs, c.BB, c.phi, S.r, (T *)nullptr, S.res_slots, c.fast,
                        c.unit_ops ? c.UNIT : nullptr);
    }
    launch_reduce_err(st, S.res_slots, S.nslots, S.res_val);
    (void)hipMemcpyAsync(S.h_scal + 8, S.res_val, 8, hipMemcpyDeviceToHost, st);
}

template <typename T>
static void enqueue_init(MgSolver<T> &S, hipStream_t st) {
    const int64_t n = 2 * S.c.g.HS;
    const unsigned fb = flat_blocks(n);
    hipLaunchKernelGGL((k_bicg_init<T>), dim3(fb), dim3(256), 0, st, n, (const T *)S.r, S.rh, S.p, S.v, S.x, S.part);
    hipLaunchKernelGGL(k_bicg_scalar, dim3(1), dim3(256), 0, st, (const double *)S.part, (int)fb, S.scal, 0);
    (void)hipMemcpyAsync(S.h_scal, S.scal, 8 * sizeof(double), hipMemcpyDeviceToHost, st);
}

template <typename T>
static void enqueue_iteration(MgSolver<T> &S, hipStream_t st) {
    const int64_t n = 2 * S.c.g.HS;
    const unsigned fb = flat_blocks(n);
    const int nap = (int)apply_blocks<T>(S.c.g);
    hipLaunchKernelGGL((k_bicg_p<T>), dim3(blocks_for(n)), dim3(256), 0, st, n, (const T *)S.r, S.p, (const T *)S.v,
                       (const double *)S.scal);
    vcycle(S, st, (const T *)S.p, S.y);
    apply0<T, 0>(S, st, S.y, S.v, S.rh);
    hipLaunchKernelGGL(k_bicg_scalar, dim3(1), dim3(256), 0, st, (const double *)S.part, nap, S.scal, 1);
    hipLaunchKernelGGL((k_bicg_s<T>), dim3(blocks_for(n)), dim3(256), 0, st, n, (const T *)S.r, (const T *)S.v, S.s,
                       (const double *)S.scal);
    vcycle(S, st, (const T *)S.s, S.z);
    apply0<T, 1>(S, st, S.z, S.t, S.s);
    hipLaunchKernelGGL(k_bicg_scalar, dim3(1), dim3(256), 0, st, (const double *)S.part, nap, S.scal, 2);
    hipLaunchKernelGGL((k_bicg_xr<T>), dim3(fb), dim3(256), 0, st, n, S.x, (const T *)S.y, (const T *)S.z, S.r,
                       (const T *)S.s, (const T *)S.t, (const T *)S.rh, (const double *)S.scal, S.part);
    hipLaunchKernelGGL(k_bicg_scalar, dim3(1), dim3(256), 0, st, (const double *)S.part, (int)fb, S.scal, 3);
    (void)hipMemcpyAsync(S.h_scal, S.scal, 8 * sizeof(double), hipMemcpyDeviceToHost, st);
    S.vcycles += 2;
}

// p += e and the start of the next cycle.
template <typename T>
static void enqueue_update_and_residual(MgSolver<T> &S, hipStream_t st) {
    launch_update_track(st, S.c.g, S.x, S.c.P, S.err_slots, S.nslots);
    ++S.cyc;
    enqueue_residual(S, st);
    S.state = MG_RESIDUAL;
}

template <typename T>
hipError_t mg_step_solve(const std::vector<MgSolver<T> *> &runs, const std::vector<hipStream_t> &streams,
                         const MgControl &ctl, std::vector<MgStepStats> &out) {
    const size_t nr = runs.size(), ns = streams.size();
    out.assign(nr, MgStepStats{-1., 0, 0});
    if (nr == 0 || ns == 0) return hipSuccess;
    hipError_t err = hipSuccess;
    for (size_t q = 0; q < nr; ++q) {
        MgSolver<T> &S = *runs[q];
        hipStream_t st = streams[q % ns];
        if (!S.built) mg_setup(S, st);
        S.state = MG_RESIDUAL;
        S.cyc = 0;
        S.it = 0;
        S.vcycles = 0;
        S.res = -1.;
        enqueue_residual(S, st);
        if ((err = hipEventRecord(S.ev, st)) != hipSuccess) return err;
    }
    size_t active = nr;
    while (active > 0) {
        for (size_t q = 0; q < nr; ++q) {
            MgSolver<T> &S = *runs[q];
            if (S.state == MG_DONE) continue;
            hipStream_t st = streams[q % ns];
            if ((err = hipEventSynchronize(S.ev)) != hipSuccess) return err;
            if (S.state == MG_RESIDUAL) {
                unsigned long long bits;
                std::memcpy(&bits, S.h_scal + 8, 8);
                std::memcpy(&S.res, &bits, 8);
                // a non-finite residual never converges: stop at once (the caller reports it)
                if (S.res <= ctl.tol || S.cyc >= ctl.max_cycles || !std::isfinite(S.res)) {
                    S.state = MG_DONE;
                    out[q] = MgStepStats{S.res, S.cyc, S.vcycles};
                    --active;
                    continue;
                }
                enqueue_init(S, st);
                S.state = MG_INIT;
            } else if (S.state == MG_INIT) {
                S.b2 = S.h_scal[4];
                if (S.h_scal[6] != 0. || !(S.b2 > 0.)) {
                    enqueue_update_and_residual(S, st);   // e = 0: the residual is re-read
                } else {
                    S.it = 0;
                    enqueue_iteration(S, st);
                    S.state = MG_ITER;
                }
            } else {   // MG_ITER
                ++S.it;
                const double rel = std::sqrt(S.h_scal[5] / S.b2);
                if (!(rel > ctl.rtol) || S.h_scal[6] != 0. || S.it >= ctl.maxit) enqueue_update_and_residual(S, st);
                else enqueue_iteration(S, st);
            }
            if ((err = hipEventRecord(S.ev, st)) != hipSuccess) return err;
        }
    }
    return hipGetLastError();
}

template bool mg_create<float>(MgSolver<float> &, const MgCase &, std::string *, void *);
template bool mg_create<double>(MgSolver<double> &, const MgCase &, std::string *, void *);
template void mg_destroy<float>(MgSolver<float> &);
template void mg_destroy<double>(MgSolver<double> &);
template void mg_setup<float>(MgSolver<float> &, hipStream_t);
template void mg_setup<double>(MgSolver<double> &, hipStream_t);
template hipError_t mg_step_solve<float>(const std::vector<MgSolver<float> *> &, const std::vector<hipStream_t> &,
                                         const MgControl &, std::vector<MgStepStats> &);
template hipError_t mg_step_solve<double>(const std::vector<MgSolver<double> *> &, const std::vector<hipStream_t> &,
                                          const MgControl &, std::vector<MgStepStats> &);

}  // namespace pf
