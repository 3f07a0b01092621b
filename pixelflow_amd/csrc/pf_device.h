// pf_device.h -- device helpers shared by the FP64 and mixed-precision kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "pf_kernels.h"

#define PF_SMALL 1.e-6
#define PF_ALPHA 32.0

namespace pf {

__device__ __forceinline__ double fmax_ref(double a, double b) { return (a > b) ? a : b; }

// Running max of |dp| (or |r/ap|) over the cells a thread updates.  For finite values it
// is the reference's max; a NaN operand is kept and stays kept (every later compare with
// it is false), so a non-finite pressure reaches the reduction and the step reports it.
__device__ __forceinline__ double max_keep_nan(double e, double d) { return (d > e || d != d) ? d : e; }

// ------------------------------------------------------------------------------------
// boundary_matrix row edits, replayed per cell in the reference's order.
// Uniform: inlet fold at i=1 then outlet row at i=m (3D :636-659, 2D :424-436).
// Air: top, bottom, east, west, north, south (ibm_3d_air_condition_omp_cpu.f90:685-864).
// WITH_BB adds the outlet terms to *bb (p ghosts read from the colour-split P); the top
// outlet assigns bb(i,j,l) = bb(i,j,1) + at*p(i,j,l+1) as written there (:702).
// ------------------------------------------------------------------------------------
template <typename T>
struct CoefT {
    T ae, aw, an, as, at, ab;
};
using Coef = CoefT<double>;

// Global-memory view of a pointer.  Loads and stores through it compile to global_*
// instructions.  Pointers that a kernel reads from memory (the case table of a batch) are
// otherwise generic, and generic accesses compile to flat_* instructions, which count on
// both the vector-memory and the scalar-memory wait counters: every wait for a scalar
// load of the case record then also waits for the tile's outstanding field loads.
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> gmem(T *p) {
    return (gptr<T>)p;
}

template <typename T>
__device__ __forceinline__ void zero6(CoefT<T> &c) { c.ae = T(0); c.aw = T(0); c.an = T(0); c.as = T(0); c.at = T(0); c.ab = T(0); }

__device__ __forceinline__ bool is_edge_row(const PfGeom &g, const StepArgs &a, int i, int j, int k) {
    if (a.bc == PF_BC_UNIFORM) return i == 1 || i == g.m;
    return i == 1 || i == g.m || j == 1 || j == g.n || k == 1 || k == g.l;
}

template <bool WITH_BB, typename T = double>
__device__ void row_rules(const PfGeom &g, const StepArgs &a, int i, int j, int k, CoefT<T> &c,
                          double *bb, const double *P, const double *phi, double bb_k1) {
    if (a.bc == PF_BC_UNIFORM) {
        if (i == 1) { c.ae = c.ae + c.aw; c.aw = T(0); }
        if (i == g.m) {
            if (WITH_BB) *bb = *bb + (c.ae * gmem(P)[pf_soff(g, g.m + 1, j, k)]);
            zero6(c);
        }
        return;
    }
    const int m = g.m, n = g.n, l = g.l;
    const double pcell = gmem(phi)[pf_off(g, i, j, k)];
    const bool fluid = !(pcell < 0.9);
    if (k == l) {
        if (a.face[0] == 0 || a.face[0] == 1 || !fluid) { c.ab = c.ab + c.at; c.at = T(0); }
        else {
            if (WITH_BB) *bb = bb_k1 + (c.at * gmem(P)[pf_soff(g, i, j, l + 1)]);
            zero6(c);
        }
    }
    if (k == 1) {
        if (a.face[1] == 0 || a.face[1] == 1 || !fluid) { c.at = c.at + c.ab; c.ab = T(0); }
        else { if (WITH_BB) *bb = *bb + (c.ab * gmem(P)[pf_soff(g, i, j, 0)]); zero6(c); }
    }
    if (i == m) {
        if (a.face[2] == 0 || a.face[2] == 1 || !fluid) { c.aw = c.aw + c.ae; c.ae = T(0); }
        else { if (WITH_BB) *bb = *bb + (c.ae * gmem(P)[pf_soff(g, m + 1, j, k)]); zero6(c); }
    }
    if (i == 1) {
        if (a.face[3] == 0 || a.face[3] == 1 || !fluid) { c.ae = c.ae + c.aw; c.aw = T(0); }
        else { if (WITH_BB) *bb = *bb + (c.aw * gmem(P)[pf_soff(g, 0, j, k)]); zero6(c); }
    }
    if (j == n) {
        if (a.face[4] == 0 || a.face[4] == 1 || !fluid) { c.as = c.as + c.an; c.an = T(0); }
        else { if (WITH_BB) *bb = *bb + (c.an * gmem(P)[pf_soff(g, i, n + 1, k)]); zero6(c); }
    }
    if (j == 1) {
        if (a.face[5] == 0 || a.face[5] == 1 || !fluid) { c.an = c.an + c.as; c.as = T(0); }
        else { if (WITH_BB) *bb = *bb + (c.as * gmem(P)[pf_soff(g, i, 0, k)]); zero6(c); }
    }
}

// XCD-aware block order.  The hardware hands workgroup L of a 1D grid to XCD L % 8; each
// XCD has its own L2.  Renumbering so that XCD x receives the contiguous range
// [x*B/8, (x+1)*B/8) of logical blocks keeps the rows and planes that neighbouring blocks
// share (the j-1/j+1 rows and k-1/k+1 planes of a stencil) in one L2.
__device__ __forceinline__ int xcd_logical_block(int L, int B) {
    const int q = B >> 3, rem = B & 7;
    const int x = L & 7, r = L >> 3;
    return (x < rem) ? x * (q + 1) + r : rem * (q + 1) + (x - rem) * q + r;
}

// Variant of the sweep that rebuilds the six face coefficients of each updated cell from
// porosity instead of reading them: AW(i) = dt*max(small,(phi(i)+phi(i-1))*0.5)/dx/dx is
// evaluated with the same operands and order as solve_p (3D :390-400, 2D :265-268), so the
// values are the stored ones bit for bit, while a half-sweep streams 3 words per cell
// (p of both colours, porosity of both colours, bb of the colour, p written) instead of 5.
// PHIo / PHIx: porosity in the colour-split layout (own / other colour).
__device__ __forceinline__ double face_coef(double pa, double pb, double dt, double d) {
    return ((dt * fmax_ref(PF_SMALL, (pa + pb) * 0.5)) / d) / d;
}

// a / d for a constant divisor d with rd = RN(1/d) from the host: q0 = a*rd is within one
// ulp of a/d, the residual a - q0*d is exact in an FMA, and one FMA correction gives the
// correctly rounded quotient (Markstein's theorem; 3 instructions instead of the ~10 of a
// general FP64 division).  k_faces_check verifies every face of a case against true
// division before this path is used.
__device__ __forceinline__ double div_const(double a, double d, double rd) {
    const double q0 = a * rd;
    const double e = __builtin_fma(-q0, d, a);
    return __builtin_fma(e, rd, q0);
}
// ------------------------------------------------------------------------------------
// Division by a divisor shared by the whole wave (a grid spacing, the square of a wall
// thickness), bit-identical to the compiler's IEEE division a / d.  gfx950 divides doubles
// in eleven instructions: s = v_div_scale(d) (d itself unless the quotient's exponent is
// near either end of the double range, then d * 2^+-128), the reciprocal r of s from
// v_rcp_f64 and two Newton steps (four FMAs), n = v_div_scale(a), q = n*r,
// e = fma(-s, q, n), v_div_fmas(e, r, q) and v_div_fixup(.., d, a).  DivK holds s and r of
// the divisor as scaled for a unit dividend, computed by that same instruction sequence;
// DivFast recomputes s per dividend and, where it agrees, reuses r: the same instructions
// on the same operands, seven of them per quotient with the check.  A dividend whose s
// differs sets `slow`, and the caller evaluates the cell again with DivIeee.  The check is
// an ordered compare: v_div_scale gives NaN for a zero dividend, and there v_div_fixup
// returns the signed zero whatever its quotient operand holds; a NaN or infinite dividend
// likewise leaves v_div_fixup's result independent of the quotient operand.
// ------------------------------------------------------------------------------------
struct DivK {
    double d, s, r;
};
// A double that every active lane holds, moved to scalar registers (VALU instructions read
// one scalar register pair as an operand, which frees six vector registers per DivK).
__device__ __forceinline__ double wave_uniform(double x) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ DivK divk_make(double d) {
    bool f;
    const double s = __builtin_amdgcn_div_scale(1.0, d, false, &f);
    const double r0 = __builtin_amdgcn_rcp(s);
    const double r1 = __builtin_fma(r0, __builtin_fma(-s, r0, 1.0), r0);
    const double r = __builtin_fma(r1, __builtin_fma(-s, r1, 1.0), r1);
    return DivK{wave_uniform(d), wave_uniform(s), wave_uniform(r)};
}
struct DivFast {
    bool slow = false;
    __device__ __forceinline__ double operator()(double a, const DivK &k) {
        bool f;
        const double s = __builtin_amdgcn_div_scale(a, k.d, false, &f);
        slow = slow || __builtin_islessgreater(s, k.s);
        const double n = __builtin_amdgcn_div_scale(a, k.d, true, &f);
        const double q = n * k.r;
        const double e = __builtin_fma(-k.s, q, n);
        return __builtin_amdgcn_div_fixup(__builtin_amdgcn_div_fmas(e, k.r, q, f), k.d, a);
    }
};
// Zero, produced where the compiler cannot see its value.
__device__ __forceinline__ int opaque_zero() {
    int z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    return z;
}
struct DivIeee {
    __device__ __forceinline__ double operator()(double a, const DivK &k) const { return a / k.d; }
};

__device__ __forceinline__ double face_coef_fast(double pa, double pb, double dt, double d, double rd) {
    return div_const(div_const(dt * fmax_ref(PF_SMALL, (pa + pb) * 0.5), d, rd), d, rd);
}

template <bool FAST>
__device__ __forceinline__ double face_any(double pa, double pb, double dt, double d, double rd) {
    return FAST ? face_coef_fast(pa, pb, dt, d, rd) : face_coef(pa, pb, dt, d);
}

// ------------------------------------------------------------------------------------
// Thread geometry of the colour sweeps.  Block (64, 4): x runs over the colour's slots of
// a row (or, with 32-slot waves, over 32 slots of two rows).  3D: y = (row offset 0..1,
// plane parity 0..1); a block covers 2 (or 4) rows and 2*KB planes and each thread updates
// KB cells of its colour two planes apart (same slot, same x position; independent updates
// whose loads overlap in flight).  2D: y = 4 row groups, KB = 1.  Blocks form a 1D grid
// renumbered XCD-contiguously (xcd_logical_block).
// ------------------------------------------------------------------------------------
struct SweepShape {
    int gx, gy, gz;
    PF_HD int blocks() const { return gx * gy * gz; }
};

// Wave width of the colour sweeps: 64 colour slots of one row, or 32 slots of two rows
// when that leaves fewer lanes idle (rows of 96 slots fill 3/4 of two 64-lane waves but
// all of three 32-slot halves).  Rows of 128 slots keep 64.
PF_HD int sweep_xw(const PfGeom &g) {
    const int S = (g.m + 1) / 2;
    const int w64 = (S + 63) / 64 * 64, w32 = (S + 31) / 32 * 32;
    return (w32 < w64) ? 32 : 64;
}

PF_HD SweepShape sweep_shape(const PfGeom &g, int KB, int xw) {
    SweepShape t;
    xw = (xw == 32) ? 32 : 64;
    const int rpw = 64 / xw;   // rows per wave
    t.gx = ((g.m + 1) / 2 + xw - 1) / xw;
    if (g.dim == 3) {
        t.gy = (g.n + 2 * rpw - 1) / (2 * rpw);
        t.gz = (g.l + 2 * KB - 1) / (2 * KB);
    } else {
        t.gy = (g.n + 4 * rpw - 1) / (4 * rpw);
        t.gz = 1;
    }
    return t;
}

struct SweepThread {
    int j, kbase, sh, h, i;
};

// Block (64, 4): threadIdx.x runs over xw slots of rpw = 64 / xw rows; 3D: threadIdx.y =
// (row pair offset, plane parity), 2D: threadIdx.y = row group.
// Block index / grid extent by multiply-high: m = floor(2^32 / d) + 1 gives
// umulhi(x, m) = x / d exactly for x * d < 2^32 (the error x * (m - 2^32 / d) / 2^32 stays
// below 1 / d).  grid_magic returns 0 (divide instead) for d <= 1 or a grid past that bound.
PF_HD uint32_t grid_magic(uint32_t d, uint32_t blocks) {
    if (d <= 1 || (uint64_t)blocks * d >= ((uint64_t)1 << 32)) return 0u;
    return (uint32_t)((((uint64_t)1) << 32) / d + 1);
}
PF_HD uint32_t grid_div(uint32_t x, uint32_t d, uint32_t m) {
    return m ? (uint32_t)(((uint64_t)x * m) >> 32) : (d <= 1 ? x : x / d);
}
// gx, gy and their multipliers for a grid of `blocks` blocks.
PF_HD void sweep_set_grid(SweepArgs &s, int gx, int gy, int blocks) {
    s.gx = gx;
    s.gy = gy;
    s.mgx = grid_magic((uint32_t)gx, (uint32_t)blocks);
    s.mgy = grid_magic((uint32_t)gy, (uint32_t)blocks);
}

// The cell of thread (tx, ty) of logical block L (host and device).
template <int DIM, int KB>
PF_HD SweepThread sweep_thread_at(const PfGeom &g, const SweepArgs &s, int L, int tx, int ty) {
    const int byz = (int)grid_div((uint32_t)L, (uint32_t)s.gx, s.mgx), bx = L - byz * s.gx;
    const int bz = (int)grid_div((uint32_t)byz, (uint32_t)s.gy, s.mgy), by = byz - bz * s.gy;
    const int rpw = (s.xw == 32) ? 2 : 1;
    const int sub = (s.xw == 32) ? (tx >> 5) : 0;
    const int xl = (s.xw == 32) ? (tx & 31) : tx;
    SweepThread t;
    if (DIM == 3) {
        t.j = by * 2 * rpw + (ty >> 1) * rpw + sub + 1;
        t.kbase = bz * 2 * KB + 1 + (ty & 1);
    } else {
        t.j = by * 4 * rpw + ty * rpw + sub + 1;
        t.kbase = 0;
    }
    t.sh = (s.colour + t.j + t.kbase) & 1;   // the same for kbase + 2q
    t.h = bx * s.xw + xl + 1 - t.sh;
    t.i = 2 * t.h + t.sh;
    return t;
}
template <int DIM, int KB>
__device__ __forceinline__ SweepThread sweep_thread(const PfGeom &g, const SweepArgs &s) {
    const int nb = (int)gridDim.x;
    const int L = s.xcd_remap ? xcd_logical_block((int)blockIdx.x, nb) : (int)blockIdx.x;
    return sweep_thread_at<DIM, KB>(g, s, L, (int)threadIdx.x, (int)threadIdx.y);
}

// Max of 64-bit patterns over the wave, returned in every lane.  Cross-lane moves are DPP
// (data-parallel primitives: a VALU instruction reads a neighbour lane's register), so no
// step involves an LDS round trip; 64-bit shuffles cost six dependent ds_bpermute pairs at
// the end of every wave, which lengthened a colour sweep that reduces by 1.6 us.  The max
// is taken on the high words first, then on the low words of the lanes that hold the
// winning high word: that order is the unsigned 64-bit order.  Each 32-bit max takes six
// steps -- swap pairs and quads (quad_perm), mirror half-rows and rows (max over 16 lanes),
// row_bcast:15 into rows 1 and 3, row_bcast:31 into rows 2 and 3 -- which leave it in lane
// 63.  Every lane of the wave must be active.  Sweeps reduce |dp| as the bit patterns of
// non-negative doubles, which order the same; a NaN pattern (0x7ff8...) given to the
// reduction exceeds them all and is passed on.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned dpp_umax32_step(unsigned x) {
    // old value 0, the identity of the max: lanes a masked row_bcast leaves unwritten keep x
    const unsigned y = (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROW_MASK, 0xF, false);
    return (y > x) ? y : x;
}
__device__ __forceinline__ unsigned wave_umax32(unsigned x) {
    x = dpp_umax32_step<0xB1, 0xF>(x);    // quad_perm [1,0,3,2]
    x = dpp_umax32_step<0x4E, 0xF>(x);    // quad_perm [2,3,0,1]
    x = dpp_umax32_step<0x141, 0xF>(x);   // row_half_mirror
    x = dpp_umax32_step<0x140, 0xF>(x);   // row_mirror
    x = dpp_umax32_step<0x142, 0xA>(x);   // row_bcast:15
    x = dpp_umax32_step<0x143, 0xC>(x);   // row_bcast:31
    return (unsigned)__builtin_amdgcn_readlane((int)x, 63);
}
__device__ __forceinline__ unsigned long long wave_umax(unsigned long long e) {
    const unsigned hi = (unsigned)(e >> 32);
    const unsigned mh = wave_umax32(hi);
    const unsigned ml = wave_umax32(hi == mh ? (unsigned)e : 0u);
    return ((unsigned long long)mh << 32) | ml;
}

// Max of a non-negative double over the wave, folded by lane 0 into the wave's own slot,
// slots[block * waves_per_block + wave], with an atomic max whose result is not read: no
// barrier, and no load of the slot for the wave to wait on before it can exit.  Each wave
// has its own address, so the atomics never contend.
__device__ __forceinline__ void wave_max_to_slot(double v, unsigned long long *slots) {
    const unsigned long long e = wave_umax(__double_as_longlong(v));
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    if ((lane & 63) == 0) {
        const int wpb = (blockDim.x * blockDim.y + 63) >> 6;
        atomicMax(slots + (int64_t)blockIdx.x * wpb + (lane >> 6), e);
    }
}


}  // namespace pf
