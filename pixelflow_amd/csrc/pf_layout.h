// pf_layout.h -- device data layout of one case (shared by host code and kernels).
//
// Natural layout (u, v, w, their *_old copies, porosity, div): one double per cell,
// ghosts included, x fastest, row pitch PX (>= m+2, a multiple of 16 doubles so every row
// starts on a 128-byte line): off(i,j,k) = i + PX*(j + Y*k), Y = n+2, Z = l+2 (Z = 1, k = 0
// in 2D).
//
// Colour-split layout (pressure p, RHS bb, Poisson face coefficients): the cells of one
// red/black colour c = (i+j+k) & 1 of a row sit contiguously, h = i >> 1 being the slot,
// so one SOR half-sweep reads and writes dense, coalesced rows instead of every other
// double.  hoff(h,j,k) = h + HP*(j + Y*k); colour c occupies [c*HS, (c+1)*HS).
// For a colour-c cell (i = 2h+s, s = i & 1) the x neighbours are the other colour's slots
// h-1+s (west) and h+s (east); the y/z neighbours are the other colour's slot h.
//
// Face coefficients: the Poisson matrix of solve_p (ibm_3d_uniform_omp_cpu.f90:390-402)
// is built from one value per cell face, ae(i) == aw(i+1) bit for bit (both evaluate
// dt*max(small,(phi(i+1)+phi(i))*0.5)/dx/dx with the same operands).  AW(i,j,k) stores the
// face between i-1 and i (= aw), AS the one between j-1 and j, AB between k-1 and k, each
// in the colour-split layout of its cell; ae/an/at of a cell are the AW/AS/AB of its
// east/north/top neighbour.  ap is recomputed from the six values in the reference's
// order; the rows that boundary_matrix edits are re-derived from the face codes.
#pragma once
#include <stdint.h>

struct PfGeom {
    int dim;          // 2 or 3
    int m, n, l;      // interior extents (l = 1 in 2D)
    int PX, Y, Z;     // natural pitch and extents with ghosts
    int H, HP;        // colour-split slots per row (ceil((m+2)/2)) and their pitch
    int64_t NS;       // natural field size (doubles)
    int64_t HS;       // one colour's size (doubles)
};

#if defined(__HIPCC__)
#define PF_HD __host__ __device__ __forceinline__
#else
#define PF_HD inline
#endif

PF_HD int64_t pf_off(const PfGeom &g, int i, int j, int k) {
    return (int64_t)i + (int64_t)g.PX * ((int64_t)j + (int64_t)g.Y * k);
}
PF_HD int64_t pf_hoff(const PfGeom &g, int h, int j, int k) {
    return (int64_t)h + (int64_t)g.HP * ((int64_t)j + (int64_t)g.Y * k);
}
// offset of cell (i,j,k) in a colour-split array (both colours in one allocation)
PF_HD int64_t pf_soff(const PfGeom &g, int i, int j, int k) {
    return (int64_t)((i + j + k) & 1) * g.HS + pf_hoff(g, i >> 1, j, k);
}

inline PfGeom pf_make_geom(int dim, int m, int n, int l) {
    PfGeom g;
    g.dim = dim;
    g.m = m;
    g.n = n;
    g.l = (dim == 3) ? l : 1;
    g.PX = ((m + 2) + 15) / 16 * 16;
    g.Y = n + 2;
    g.Z = (dim == 3) ? l + 2 : 1;
    g.H = (m + 3) / 2;
    g.HP = (g.H + 15) / 16 * 16;
    g.NS = (int64_t)g.PX * g.Y * g.Z;
    g.HS = (int64_t)g.HP * g.Y * g.Z;
    return g;
}

// Unit-porosity chunks of the colour sweeps.  One byte per 32 colour slots of a row,
// (colour, plane, row, chunk), is 1 when every cell of that colour the chunk holds
// (1 <= i <= m) and its six neighbours have porosity exactly 1.  A sweep lane's chunk is
// its slot position h + sh - 1 (0 for the first lane of the row) over 32, so a wave of 64
// slots spans two chunks of one row and a wave of 32 slots one chunk of each of its two
// rows.
PF_HD int unit_chunks(const PfGeom &g) { return ((g.m + 1) / 2 + 63) / 64 * 2; }
PF_HD int64_t unit_flags_size(const PfGeom &g) { return (int64_t)2 * g.Z * g.Y * unit_chunks(g); }
PF_HD int64_t unit_index(const PfGeom &g, int c, int j, int k, int x) {
    return (((int64_t)c * g.Z + k) * g.Y + j) * unit_chunks(g) + (x >> 5);
}
