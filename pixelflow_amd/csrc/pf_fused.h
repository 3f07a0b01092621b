// pf_fused.h -- one red/black SOR iteration (both colours) in one launch.
//
// The two-launch sweep streams p (both colours), porosity (both colours) and the RHS of
// one colour per half-sweep: 6 words per cell and iteration, 2 launches.  Here a block
// owns a tile of TS colour slots x TY rows and marches through a chunk of z planes,
// keeping a 3-plane window in LDS; it updates the first colour (c0) on the tile plus a
// one-cell halo from the old values of the second colour (c1), then c1 on the tile from
// the new c0 values -- exactly the reference's order: every c0 cell sees old c1
// neighbours, every c1 cell new c0 neighbours.  Reads come from buffer A, all writes go to
// buffer B (ping-pong), so blocks never race; halo values are recomputed redundantly
// (identical bits) instead of exchanged.  Traffic per iteration: p read once and written
// once, porosity once, RHS once -- 4 words per cell, 1 launch.
//
// Ghost cells: with periodic y/z (uniform inflow, even n and l) a ghost row/plane holds
// the current value of the interior row/plane it copies, which is what the reference's
// refresh before each half-sweep establishes (ibm_3d_uniform_omp_cpu.f90:463-480); here
// every position is mapped through the wrap, so a ghost position is computed as its
// source cell and written with the new value.  Non-periodic ghosts (x always, all faces
// of the air-conditioning room) are not updated during the SOR: B receives A's value.
// Odd periodic extents (a ghost copy then has the other colour) use the two-launch sweep.
//
// The update of a cell is sor_value / sor32_value (pf_cells.h), the same function the
// two-launch sweep calls, with the same operands: results are identical bit for bit.
//
// Measured on MI355X (256x128x128 sphere): 91 us per FP64 iteration against 36 us for
// the two colour-sweep launches, 56 us against 24 us in FP32.  The halo recomputation
// (c0 on 1.6x the tile's cells), the serial z march and 138 VGPRs cost more than the
// third of the traffic saved, so the two-launch sweep stays the default and this path is
// opt-in (PF_FUSED=1) until the z march is pipelined deeper.
#pragma once
#include "pf_cells.h"

namespace pf {

constexpr int FUSED_TS = 32;   // colour slots per tile row (64 cells of the grid in x)
constexpr int FUSED_TY = 8;    // rows per tile
constexpr int FUSED_KZ = 8;    // z planes per block (3D)

// KIND: 0 = FP64 pressure (sor_value), 1 = FP32 correction (sor32_value).
template <int KIND> struct FusedT { using T = double; };
template <> struct FusedT<1> { using T = float; };

template <typename T>
struct FusedArgs {
    const T *A;       // colour-split values before the iteration (both colours)
    T *B;             // colour-split values after it
    const T *F;       // colour-split porosity (FP64 copy or FP32 copy)
    const T *R;       // colour-split right-hand side (bb, or the FP32 residual)
    const double *phi;  // natural-layout porosity (edge rules of the air room)
    float cdx, cdy, cdz;  // FP32: dt/dx^2 etc.
};

struct FusedGeom {
    int per;          // periodic y (and z in 3D)
    int gx, gy, gz;   // tiles in x (slots), y (rows), z (plane chunks)
};

__device__ __forceinline__ int fused_wrap(int j, int n, int per) {
    if (!per) return j;
    if (j < 1) return j + n;
    if (j > n) return j - n;
    return j;
}

// One tile: interior cells i in [2*x0+1, 2*x0+2*TS] (TS cells of each colour per row), rows
// [j0, j0+TY) and planes [k0, k1] (3D) or the plane k = 0 (2D), clipped to the interior.
// A colour-c cell of row parity s = (c+j+k)&1 sits in slot h = (i-s)/2, so the tile's
// cells of a row occupy slots x0+t+1-s, t in [0, TS).  Windows (slot origin, width):
// c1 values and porosity (x0-1, TS+3) x rows [j0-2, j0+TY+1]; new c0 values and c0
// porosity (x0, TS+1) x rows [j0-1, j0+TY].  Periodic ghost rows/planes of the written
// cells are written too (the wrap's copies); non-periodic ghosts are never written (B
// holds A's ghost values from a copy made before the first iteration of the step).
// Returns the max |new - old| over the tile's updated cells of the tracked colour(s)
// (3D: the second colour; 2D: both, as the reference).
template <int DIM, int KIND, bool FAST>
__device__ double fused_tile(const PfGeom &g, const StepArgs &a, const FusedGeom &fg,
                             const FusedArgs<typename FusedT<KIND>::T> &p, int x0, int j0, int k0, int k1) {
    using T = typename FusedT<KIND>::T;
    constexpr int TS = FUSED_TS, TY = FUSED_TY;
    constexpr int W1 = TS + 3, H1 = TY + 4;     // c1 window
    constexpr int W0 = TS + 1, H0 = TY + 2;     // c0 window
    constexpr int NP = (DIM == 3) ? 3 : 1;      // planes kept
    __shared__ T s_c1[NP][H1][W1];   // old c1 values
    __shared__ T s_f1[NP][H1][W1];   // porosity of c1 cells
    __shared__ T s_c0[NP][H0][W0];   // new c0 values
    __shared__ T s_f0[NP][H0][W0];   // porosity of c0 cells
    const int c0 = (DIM == 3) ? 0 : 1, c1 = 1 - c0;
    const int tid = (int)threadIdx.x;
    const int n = g.n, l = g.l, m = g.m;
    const int per = fg.per;
    const int64_t HPY = (int64_t)g.HP * g.Y;
    double err = 0.;

    auto ring = [](int k) { return ((k % 3) + 3) % 3; };
    // offset (colour c, slot h, row j, plane k) after the periodic wrap; -1 outside the array
    auto offs = [&](int c, int h, int j, int k, int &i, int &jw, int &kw) -> int64_t {
        jw = fused_wrap(j, n, per);
        kw = (DIM == 3) ? fused_wrap(k, l, per) : 0;
        i = 2 * h + ((c + jw + kw) & 1);
        if (h < 0 || h >= g.HP || jw < 0 || jw > n + 1 || kw < 0 || kw > (DIM == 3 ? l + 1 : 0)) return -1;
        return (int64_t)c * g.HS + pf_hoff(g, h, jw, kw);
    };
    auto interior = [&](int i, int jw, int kw) {
        return i >= 1 && i <= m && jw >= 1 && jw <= n && (DIM == 2 || (kw >= 1 && kw <= l));
    };
    // store v for colour c at (h, j, k) in B, and in the periodic ghost copies of that cell
    auto put = [&](int c, int h, int j, int k, T v) {
        T *B = p.B + (int64_t)c * g.HS;
        const int64_t o = pf_hoff(g, h, j, k);
        B[o] = v;
        if (per) {
            if (j == n) B[o - (int64_t)g.HP * n] = v;
            if (j == 1) B[o + (int64_t)g.HP * n] = v;
            if (DIM == 3) {
                if (k == l) B[o - HPY * l] = v;
                if (k == 1) B[o + HPY * l] = v;
            }
        }
    };
    // Fixed element assignment: thread tid handles window elements q = tid + 256*e.
    constexpr int E1 = (W1 * H1 + 255) / 256, E0 = (W0 * H0 + 255) / 256, EC = (TS * TY + 255) / 256;
    // values of the next plane held in registers while the current one is computed
    struct Prefetch {
        T c1[E1], f1[E1], f0[E0], a0[E0], r0[E0], r1[EC];
    };
    // c1 window of plane kc, c0 porosity window, c0 values and RHS of plane k0p, c1 RHS of
    // plane k1p
    auto fetch = [&](Prefetch &f, int kc, int k0p, int k1p) {
#pragma unroll
        for (int e = 0; e < E1; ++e) {
            const int q = tid + 256 * e;
            T c = T(0), fv = T(0);
            if (q < W1 * H1) {
                int i, jw, kw;
                const int64_t o = offs(c1, x0 - 1 + q % W1, j0 - 2 + q / W1, kc, i, jw, kw);
                if (o >= 0) { c = p.A[o]; fv = p.F[o]; }
            }
            f.c1[e] = c;
            f.f1[e] = fv;
        }
#pragma unroll
        for (int e = 0; e < E0; ++e) {
            const int q = tid + 256 * e;
            T fv = T(0), av = T(0), rv = T(0);
            if (q < W0 * H0) {
                int i, jw, kw;
                const int64_t o = offs(c0, x0 + q % W0, j0 - 1 + q / W0, k0p, i, jw, kw);
                if (o >= 0) {
                    fv = p.F[o];
                    av = p.A[o];
                    if (interior(i, jw, kw)) rv = p.R[o];
                }
            }
            f.f0[e] = fv;
            f.a0[e] = av;
            f.r0[e] = rv;
        }
#pragma unroll
        for (int e = 0; e < EC; ++e) {
            const int q = tid + 256 * e;
            T rv = T(0);
            if (q < TS * TY) {
                const int t = q % TS, dy = q / TS, j = j0 + dy;
                const int sp = (c1 + j + k1p) & 1, h = x0 + t + 1 - sp, i = 2 * h + sp;
                if (i <= m && j <= n && k1p >= (DIM == 3 ? 1 : 0) && k1p <= (DIM == 3 ? l : 0))
                    rv = p.R[(int64_t)c1 * g.HS + pf_hoff(g, h, j, k1p)];
            }
            f.r1[e] = rv;
        }
    };
    auto stash_c1 = [&](const Prefetch &f, int kc) {
        const int r = (DIM == 3) ? ring(kc) : 0;
#pragma unroll
        for (int e = 0; e < E1; ++e) {
            const int q = tid + 256 * e;
            if (q < W1 * H1) {
                s_c1[r][q / W1][q % W1] = f.c1[e];
                s_f1[r][q / W1][q % W1] = f.f1[e];
            }
        }
    };
    auto stash_f0 = [&](const Prefetch &f, int k) {
        const int r = (DIM == 3) ? ring(k) : 0;
#pragma unroll
        for (int e = 0; e < E0; ++e) {
            const int q = tid + 256 * e;
            if (q < W0 * H0) s_f0[r][q / W0][q % W0] = f.f0[e];
        }
    };
    // new c0 values of plane k on the c0 window (values and RHS from f); the tile's own cells
    // go to B when `write`
    auto update_c0 = [&](const Prefetch &f, int k, bool write) {
        const int r = (DIM == 3) ? ring(k) : 0;
        const int rb = (DIM == 3) ? ring(k - 1) : 0, rt = (DIM == 3) ? ring(k + 1) : 0;
#pragma unroll
        for (int e = 0; e < E0; ++e) {
            const int q = tid + 256 * e;
            if (q >= W0 * H0) continue;
            const int dx = q % W0, dy = q / W0;
            const int h = x0 + dx, j = j0 - 1 + dy;
            int i, jw, kw;
            const int64_t o = offs(c0, h, j, k, i, jw, kw);
            T v = f.a0[e];
            if (o >= 0 && interior(i, jw, kw)) {
                const int s = (c0 + jw + kw) & 1;
                const int y1 = dy + 1;                       // own row in the c1 window
                const T pW = s_c1[r][y1][dx + s], pE = s_c1[r][y1][dx + s + 1];
                const T pS = s_c1[r][dy][dx + 1], pN = s_c1[r][dy + 2][dx + 1];
                const T fW = s_f1[r][y1][dx + s], fE = s_f1[r][y1][dx + s + 1];
                const T fS = s_f1[r][dy][dx + 1], fN = s_f1[r][dy + 2][dx + 1];
                T pB = T(0), pT = T(0), fB = T(0), fT = T(0);
                if (DIM == 3) {
                    pB = s_c1[rb][y1][dx + 1];
                    pT = s_c1[rt][y1][dx + 1];
                    fB = s_f1[rb][y1][dx + 1];
                    fT = s_f1[rt][y1][dx + 1];
                }
                const T fc = s_f0[r][dy][dx];
                const T old = v;
                if (KIND == 0)
                    v = (T)sor_value<DIM, FAST>(g, a, i, jw, kw, old, pW, pE, pS, pN, pB, pT, fc, fW, fE, fS, fN, fB,
                                                fT, f.r0[e], p.phi);
                else
                    v = (T)sor32_value<DIM>(g, a, i, jw, kw, old, pW, pE, pS, pN, pB, pT, fc, fW, fE, fS, fN, fB, fT,
                                            f.r0[e], p.phi, p.cdx, p.cdy, p.cdz);
                // the tile's own cells: slot x0+t+1-s with t in [0, TS), rows [j0, j0+TY)
                const int t = dx - 1 + s;
                const bool own = t >= 0 && t < TS && dy >= 1 && dy <= TY && j == jw && (DIM == 2 || k == kw);
                if (own) {
                    if (DIM == 2) err = max_keep_nan(err, fabs((double)v - (double)old));
                    if (write) put(c0, h, j, k, v);
                }
            }
            s_c0[r][dy][dx] = v;
        }
    };
    // new c1 values of plane k on the tile (RHS from f), written to B
    auto update_c1 = [&](const Prefetch &f, int k) {
        const int r = (DIM == 3) ? ring(k) : 0;
        const int rb = (DIM == 3) ? ring(k - 1) : 0, rt = (DIM == 3) ? ring(k + 1) : 0;
#pragma unroll
        for (int e = 0; e < EC; ++e) {
            const int q = tid + 256 * e;
            if (q >= TS * TY) continue;
            const int t = q % TS, dy = q / TS;
            const int j = j0 + dy;
            const int s = (c1 + j + k) & 1;       // j, k are interior here: no wrap
            const int h = x0 + t + 1 - s;
            const int i = 2 * h + s;
            if (i > m || j > n || (DIM == 3 && k > l)) continue;
            const int xo = t + 2 - s, yo = dy + 2;      // own cell in the c1 window
            const T old = s_c1[r][yo][xo];
            const int y0 = dy + 1;                      // own row in the c0 window
            const T pW = s_c0[r][y0][t], pE = s_c0[r][y0][t + 1];
            const T pS = s_c0[r][dy][t + 1 - s], pN = s_c0[r][dy + 2][t + 1 - s];
            const T fW = s_f0[r][y0][t], fE = s_f0[r][y0][t + 1];
            const T fS = s_f0[r][dy][t + 1 - s], fN = s_f0[r][dy + 2][t + 1 - s];
            T pB = T(0), pT = T(0), fB = T(0), fT = T(0);
            if (DIM == 3) {
                pB = s_c0[rb][y0][t + 1 - s];
                pT = s_c0[rt][y0][t + 1 - s];
                fB = s_f0[rb][y0][t + 1 - s];
                fT = s_f0[rt][y0][t + 1 - s];
            }
            const T fc = s_f1[r][yo][xo];
            T v;
            if (KIND == 0)
                v = (T)sor_value<DIM, FAST>(g, a, i, j, k, old, pW, pE, pS, pN, pB, pT, fc, fW, fE, fS, fN, fB, fT,
                                            f.r1[e], p.phi);
            else
                v = (T)sor32_value<DIM>(g, a, i, j, k, old, pW, pE, pS, pN, pB, pT, fc, fW, fE, fS, fN, fB, fT,
                                        f.r1[e], p.phi, p.cdx, p.cdy, p.cdz);
            err = max_keep_nan(err, fabs((double)v - (double)old));
            put(c1, h, j, k, v);
        }
    };

    if (DIM == 2) {
        Prefetch f;
        fetch(f, 0, 0, 0);
        stash_c1(f, 0);
        stash_f0(f, 0);
        __syncthreads();
        update_c0(f, 0, true);
        __syncthreads();
        update_c1(f, 0);
        __syncthreads();   // the LDS windows are reused by the next tile of a persistent block
        return err;
    }
    // 3D: c0 of planes k0-1 .. k1+1 (the outer two as halo), c1 of planes k0 .. k1.
    // Iteration kk computes c0(kk) and c1(kk-1) from LDS and registers while the loads of
    // the c1 window of plane kk+2, the c0 porosity/values/RHS of plane kk+1 and the c1 RHS
    // of plane kk are in flight; those go to LDS (into ring slots no longer read) after.
    Prefetch cur, nxt;
    fetch(cur, k0 - 2, k0 - 1, k0 - 2);
    stash_c1(cur, k0 - 2);
    fetch(nxt, k0 - 1, k0 - 1, k0 - 2);
    stash_c1(nxt, k0 - 1);
    stash_f0(cur, k0 - 1);
    fetch(nxt, k0, k0 - 1, k0 - 2);   // c1 window of plane k0 (c0 data of plane k0-1 again)
    stash_c1(nxt, k0);
    __syncthreads();
    for (int kk = k0 - 1; kk <= k1 + 1; ++kk) {
        fetch(nxt, kk + 2, kk + 1, kk);
        update_c0(cur, kk, kk >= k0 && kk <= k1);
        __syncthreads();
        if (kk - 1 >= k0) update_c1(cur, kk - 1);
        __syncthreads();
        stash_c1(nxt, kk + 2);
        stash_f0(nxt, kk + 1);
        cur = nxt;
        __syncthreads();
    }
    return err;
}

// Block max of v (one value per thread, 256 threads) kept as a running max in slot[0]:
// doubles >= 0 order like their bit patterns.
__device__ __forceinline__ void fused_block_max(double v, unsigned long long *slot) {
    unsigned long long e = __double_as_longlong(v);
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o2 = __shfl_xor(e, off, 64);
        e = (o2 > e) ? o2 : e;
    }
    __shared__ unsigned long long red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long b = red[0];
        for (int q = 1; q < 4; ++q) b = (red[q] > b) ? red[q] : b;
        if (b > *slot) *slot = b;
    }
    __syncthreads();
}

// Tiles of a case: x in cells of 2*TS, y rows of TY, z chunks of KZ planes.
__host__ __device__ inline FusedGeom fused_geom(const PfGeom &g, int bc_uniform) {
    FusedGeom f;
    f.per = bc_uniform;
    f.gx = (g.m + 2 * FUSED_TS - 1) / (2 * FUSED_TS);
    f.gy = (g.n + FUSED_TY - 1) / FUSED_TY;
    f.gz = (g.dim == 3) ? (g.l + FUSED_KZ - 1) / FUSED_KZ : 1;
    return f;
}

// The fused sweep needs a ghost copy to have the colour of its source cell: even periodic
// extents (always true without periodic directions).
__host__ __device__ inline bool fused_ok(const PfGeom &g, int bc_uniform) {
    if (!bc_uniform) return true;
    return (g.n % 2 == 0) && (g.dim == 2 || g.l % 2 == 0);
}

// Tile t of a case (0 <= t < gx*gy*gz) -> its cell ranges.
__device__ __forceinline__ void fused_tile_coords(const PfGeom &g, const FusedGeom &f, int t, int &x0, int &j0,
                                                  int &k0, int &k1) {
    const int bx = t % f.gx, r = t / f.gx;
    const int by = r % f.gy, bz = r / f.gy;
    x0 = bx * FUSED_TS;
    j0 = 1 + by * FUSED_TY;
    if (g.dim == 3) {
        k0 = 1 + bz * FUSED_KZ;
        k1 = (k0 + FUSED_KZ - 1 < g.l) ? k0 + FUSED_KZ - 1 : g.l;
    } else {
        k0 = k1 = 0;
    }
}

}  // namespace pf
