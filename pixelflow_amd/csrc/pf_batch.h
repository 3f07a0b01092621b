// pf_batch.h -- device-side description of a batch of independent cases.
#pragma once
#include <stdint.h>

#include "pf_kernels.h"
#include "pf_layout.h"

namespace pf {

// One case of a batch: geometry, constants and the device addresses of its fields inside
// the batch arena.  Velocities exist twice (U[0], U[1]): the step of parity p predicts into
// U[p] from U[1-p] (the reference copies u to u_old instead).
struct BatchCase {
    PfGeom g;
    StepArgs a;
    double *U[2], *V[2], *W[2];
    double *phi, *div, *dout;
    double *P, *BB, *PHIs;        // colour-split FP64 (2 * HS)
    float *E, *R, *PHI32;         // colour-split FP32 (mixed precision)
    int nat_gx, nat_gy;           // 64 x 4 tiles over (m+2) x (n+2), one plane each
    int face_gx[3];               // 64-wide tile columns of each face shape (xy, yz, xz)
    int swp_gx, swp_gy;           // colour-sweep tiles (sweep_shape with one cell per thread)
    int fuse_y, fuse_z;           // periodic ghost copies written inside the sweep
    int ref_y, ref_z;             // periodic ghost refresh pass needed (odd extents)
    int fast;                     // the checked constant-divisor face division is exact here
    int slot;                     // the host's slot number of the case
};

// Kernel-side view of the batch: the case table, the tile prefix sums (ncases + 1 each)
// and, per tile, its case and its index inside the case (one load instead of a search).
struct BatchView {
    const BatchCase *cases;
    const int *nat_prefix;
    const int *swp_prefix;
    const int2 *nat_tile;
    const int2 *swp_tile;
    const int2 *face_tile[3];     // 64 x 4 tiles over one face: xy (top/bottom), yz (west/east), xz
    const int *active;            // per case: 1 = takes part in the current cycle
    int ncases;
    int nat_tiles, swp_tiles;
    int swp_base, swp_end;        // the tile range a launch covers (a group of whole cases)
    int nat_base, nat_end;
    int face_tiles[3];
};

}  // namespace pf
