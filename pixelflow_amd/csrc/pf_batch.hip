// pf_batch.hip -- many independent cases of uneven size advanced together.
//
// All cases of a batch live in one device arena; a table of BatchCase records gives each
// one's geometry, constants and field addresses.  Every stencil stage of the step is ONE
// launch for the whole batch: the work is cut into tiles (natural-cell tiles for the
// stencil stages, colour-slot tiles for the sweeps, face tiles for the air boundary), the
// tiles of all cases are numbered consecutively, and a per-tile table gives each block its
// case.  Blocks take one tile each (or a persistent grid walks the tile list), after an
// XCD-contiguous renumbering so each XCD's L2 sees neighbouring tiles.  The SOR sweeps run
// per group of consecutive cases sized for the Infinity Cache (dealt to side streams in
// mixed precision), since sweeps do not couple cases.  The per-cell arithmetic is
// pf_cells.h, shared with the single-case path, so a case gives the same bits alone and
// inside any batch.  The launch sequence of one step is captured once into a hipGraph per
// velocity parity and replayed; per-step diagnostics go to a history indexed by a device
// step counter.  Cases can be added and removed between steps (continuous batching): the
// arena has a first-fit allocator, and the tables and graphs are rebuilt when membership
// changes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../include/pixelflow.h"
#include "pf_batch.h"
#include "pf_cells.h"
#include "pf_device.h"
#include "pf_host.h"
#include "pf_kernels.h"
#include "pf_mg.h"

namespace pf {

// Block max of v into slots[t]; all threads of the block call it once per tile.
__device__ __forceinline__ void tile_max_to_slot(double v, unsigned long long *slots, int t) {
    const unsigned long long e = wave_umax(__double_as_longlong(v));
    __shared__ unsigned long long red[4];
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    if ((lane & 63) == 0) red[lane >> 6] = e;
    __syncthreads();
    if (lane == 0) {
        unsigned long long b = red[0];
        for (int q = 1; q < 4; ++q) b = (red[q] > b) ? red[q] : b;
        atomicMax(slots + t, b);   // result unread: the block need not wait for the slot
    }
    __syncthreads();   // red[] is reused by the next tile
}

enum NatOp {
    NAT_DIV = 0,
    NAT_PREDICT = 1,
    NAT_RHS = 2,
    NAT_PROJECT = 3,
    NAT_BOUNDARY_UNIFORM = 4,
    NAT_BOUNDARY_AIR = 5,   // aux = face (0 top, 1 bottom, 2 west, 3 east, 4 north, 5 south)
    NAT_UPDATE = 6,         // p += e (mixed)
    NAT_REFRESH_P = 7,      // periodic ghost refresh for odd extents
    NAT_REFRESH_E = 8,
    NAT_PREDICT2 = 9        // NAT_PREDICT of a 2D batch (chosen by the launcher)
};

template <int OP>
__device__ __forceinline__ void nat_cell(const BatchCase &C, int i, int j, int k, int parity, int aux) {
    {
        const PfGeom &g = C.g;
        const int p = parity, q = 1 - parity;
        if (OP == NAT_DIV) {
            div_cell(g, C.a, C.U[q], C.V[q], C.W[q], C.div, i, j, k);
        } else if (OP == NAT_PREDICT) {
            predict3_cell(g, C.a, C.U[q], C.V[q], C.W[q], C.div, C.phi, C.U[p], C.V[p], C.W[p], i, j, k);
        } else if (OP == NAT_PREDICT2) {
            predict2_cell(g, C.a, C.U[q], C.V[q], C.div, C.phi, C.U[p], C.V[p], i, j);
        } else if (OP == NAT_RHS || OP == NAT_PROJECT) {
            const bool interior = i >= 1 && i <= g.m && j >= 1 && j <= g.n && (g.dim == 2 || (k >= 1 && k <= g.l));
            if (!interior) return;
            if (OP == NAT_RHS) rhs_cell(g, C.a, C.U[p], C.V[p], C.W[p], C.phi, C.P, C.BB, i, j, k);
            else project_cell(g, C.a, C.P, C.U[p], C.V[p], C.W[p], i, j, k);
        } else if (OP == NAT_BOUNDARY_UNIFORM) {
            const bool bnd = (i == 0 || i == 1 || i == g.m + 1 || j == 0 || j == g.n + 1 ||
                              (g.dim == 3 && (k == 0 || k == g.l + 1)));
            if (bnd) boundary_uniform_cell(g, C.a, C.phi, C.U[p], C.V[p], C.W[p], C.P, i, j, k);
        } else if (OP == NAT_BOUNDARY_AIR) {
            int rr = -1, qq = -1;
            if (aux == 0 && k == g.l) { rr = i; qq = j; }
            else if (aux == 1 && k == 1) { rr = i; qq = j; }
            else if (aux == 2 && i == 1) { rr = j; qq = k; }
            else if (aux == 3 && i == g.m) { rr = j; qq = k; }
            else if (aux == 4 && j == g.n) { rr = i; qq = k; }
            else if (aux == 5 && j == 1) { rr = i; qq = k; }
            if (rr >= 0) boundary_air_cell(g, C.a, aux, C.phi, C.U[p], C.V[p], C.W[p], C.P, rr, qq);
        } else if (OP == NAT_UPDATE) {
            const int64_t so = pf_soff(g, i, j, k);
            gmem(C.P)[so] = gmem(C.P)[so] + (double)gmem(C.E)[so];
        } else if (OP == NAT_REFRESH_P || OP == NAT_REFRESH_E) {
            const bool iin = i >= 1 && i <= g.m;
            int sj = -1, sk = -1;
            if (C.ref_y && iin && (j == 0 || j == g.n + 1) && (g.dim == 2 || (k >= 1 && k <= g.l))) {
                sj = (j == 0) ? g.n : 1;
                sk = k;
            } else if (C.ref_z && g.dim == 3 && iin && j >= 1 && j <= g.n && (k == 0 || k == g.l + 1)) {
                sj = j;
                sk = (k == 0) ? g.l : 1;
            }
            if (sj < 0) return;
            if (OP == NAT_REFRESH_P) gmem(C.P)[pf_soff(g, i, j, k)] = gmem(C.P)[pf_soff(g, i, sj, sk)];
            else gmem(C.E)[pf_soff(g, i, j, k)] = gmem(C.E)[pf_soff(g, i, sj, sk)];
        }
    }
}

// Natural-layout tiles are 64 x 4 cells of one plane over i = 1..m, j = 0..n+1; the threads
// at i = 1 and i = m also take the ghost cells i = 0 and i = m + 1 of their row, so that a
// row of 256 cells fills four waves instead of leaving a fifth one with two live lanes.
template <int OP>
__device__ __forceinline__ void nat_tile(const BatchView &v, int t, int parity, int aux) {
    const int2 ti = v.nat_tile[t];
    const int ci = ti.x, local = ti.y;
    if ((OP == NAT_UPDATE || OP == NAT_REFRESH_P || OP == NAT_REFRESH_E) && !v.active[ci]) return;
    const BatchCase &C = v.cases[ci];
    const PfGeom &g = C.g;
    if ((OP == NAT_REFRESH_P || OP == NAT_REFRESH_E) && !C.ref_y && !C.ref_z) return;
    const int tx = local % C.nat_gx, r = local / C.nat_gx;
    const int ty = r % C.nat_gy, tz = r / C.nat_gy;
    const int i = tx * 64 + (int)threadIdx.x + 1, j = ty * 4 + (int)threadIdx.y, k = tz;
    if (i > g.m || j > g.n + 1) return;
    if (i == 1) nat_cell<OP>(C, 0, j, k, parity, aux);
    if (i == g.m) nat_cell<OP>(C, g.m + 1, j, k, parity, aux);
    nat_cell<OP>(C, i, j, k, parity, aux);
}

template <int OP, bool PERSIST>
__global__ __launch_bounds__(256) void k_batch_nat(BatchView v, int parity, int aux) {
    const int G = (int)gridDim.x;
    const int L = xcd_logical_block((int)blockIdx.x, G);
    if (PERSIST) {
        for (int t = v.nat_base + L; t < v.nat_end; t += G) nat_tile<OP>(v, t, parity, aux);
    } else if (v.nat_base + L < v.nat_end) {
        nat_tile<OP>(v, v.nat_base + L, parity, aux);
    }
}

enum SwpKind { SWP_FP64 = 0, SWP_FP32 = 1, SWP_RESIDUAL = 2 };

// Sweep tiles are TX slots wide and 256/TX rows (3D: half the rows, each thread pair
// covering two planes of opposite parity); block = 256 threads, flat.
//
// SEL picks the cases a launch serves: 1 = cases whose face coefficients passed the
// fast-division check (fast path), 0 = the others (IEEE division); FP32 sweeps serve every
// case.  Splitting them keeps the slow path's registers out of the fast kernel.
template <int DIM, int KIND, int TX, int SEL, int KB>
__device__ __forceinline__ void swp_tile(const BatchView &v, int t, int colour, int track,
                                         unsigned long long *slots) {
    constexpr int RY = 256 / TX;
    const int txi = (int)threadIdx.x % TX, tyi = (int)threadIdx.x / TX;
    const int2 ti = v.swp_tile[t];
    const int ci = ti.x, local = ti.y;
    if (!v.active[ci]) return;   // block-uniform: every thread skips the tile together
    const BatchCase &C = v.cases[ci];
    if (KIND != SWP_FP32 && (int)C.fast != SEL) return;
    const PfGeom &g = C.g;
    const int bx = local % C.swp_gx, r = local / C.swp_gx;
    const int by = r % C.swp_gy, bz = r / C.swp_gy;
    int j, kbase;
    if (DIM == 3) {
        j = by * (RY / 2) + (tyi >> 1) + 1;
        kbase = bz * 2 * KB + 1 + (tyi & 1);   // planes kbase, kbase + 2, ... (one parity)
    } else {
        j = by * RY + tyi + 1;
        kbase = 0;
    }
    const int sh = (colour + j + kbase) & 1;
    const int h = bx * TX + txi + 1 - sh;
    const int i = 2 * h + sh;
    double e = 0.;
    if (j <= g.n && i <= g.m) {
        const int64_t own = (int64_t)colour * g.HS, oth = (int64_t)(1 - colour) * g.HS;
        SweepArgs s{};
        s.colour = colour;
        s.fuse_y = C.fuse_y;
        s.fuse_z = C.fuse_z;
        int64_t o[KB];
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            // a plane past the last one stands in as plane l (in bounds) and is skipped
            const int k = (DIM == 3 && kbase + 2 * q > g.l) ? g.l : kbase + 2 * q;
            o[q] = (int64_t)h + (int64_t)g.HP * ((int64_t)j + (int64_t)g.Y * k);
        }
        if (KIND == SWP_FP64) {
            SweepIn<double> x[KB];
#pragma unroll
            for (int q = 0; q < KB; ++q)
                sweep_gather<DIM>(g, sh, o[q], C.P + oth, C.P + own, C.PHIs + own, C.PHIs + oth, C.BB + own, x[q]);
#pragma unroll
            for (int q = 0; q < KB; ++q) {
                const int k = kbase + 2 * q;
                if (DIM == 3 && k > g.l) break;
                const SweepIn<double> &c = x[q];
                const double pn = sor_value<DIM, SEL == 1>(g, C.a, i, j, k, c.c, c.w, c.e, c.s, c.n, c.b, c.t, c.fc,
                                                           c.fw, c.fe, c.fs, c.fn, c.fb, c.ft, c.rhs, C.phi);
                sweep_store<DIM>(g, s, j, k, o[q], pn, C.P + own);
                const double d = fabs(pn - c.c);
                e = max_keep_nan(e, d);
            }
        } else if (KIND == SWP_FP32) {
            const float cdx = (float)(C.a.dt / (C.a.dx * C.a.dx)), cdy = (float)(C.a.dt / (C.a.dy * C.a.dy)),
                        cdz = (float)(C.a.dt / (C.a.dz * C.a.dz));
            SweepIn<float> x[KB];
#pragma unroll
            for (int q = 0; q < KB; ++q)
                sweep_gather<DIM>(g, sh, o[q], C.E + oth, C.E + own, C.PHI32 + own, C.PHI32 + oth, C.R + own, x[q]);
#pragma unroll
            for (int q = 0; q < KB; ++q) {
                const int k = kbase + 2 * q;
                if (DIM == 3 && k > g.l) break;
                const SweepIn<float> &c = x[q];
                const float en = sor32_value<DIM>(g, C.a, i, j, k, c.c, c.w, c.e, c.s, c.n, c.b, c.t, c.fc, c.fw,
                                                  c.fe, c.fs, c.fn, c.fb, c.ft, c.rhs, C.phi, cdx, cdy, cdz);
                sweep_store<DIM>(g, s, j, k, o[q], en, C.E + own);
                const double d = fabs((double)en - (double)c.c);
                e = max_keep_nan(e, d);
            }
        } else {
#pragma unroll
            for (int q = 0; q < KB; ++q) {
                const int k = kbase + 2 * q;
                if (DIM == 3 && k > g.l) break;
                const double d = residual_cell<DIM, SEL == 1>(g, C.a, colour, i, j, k, sh, o[q], C.P, C.PHIs, C.BB,
                                                              C.phi, C.R, C.E);
                e = max_keep_nan(e, d);
            }
        }
    }
    if (track) tile_max_to_slot(e, slots, t);
}

// PERSIST = false: one block per tile, so every load of the case table precedes the
// tile's stores and the compiler keeps the table in scalar registers; true: a fixed grid
// walks the tile list.
template <int DIM, int KIND, int TX, int SEL, bool PERSIST, int KB>
__global__ __launch_bounds__(256) void k_batch_swp(BatchView v, int colour, int track, unsigned long long *slots) {
    const int G = (int)gridDim.x;
    const int L = xcd_logical_block((int)blockIdx.x, G);
    if (PERSIST) {
        for (int t = v.swp_base + L; t < v.swp_end; t += G) swp_tile<DIM, KIND, TX, SEL, KB>(v, t, colour, track, slots);
    } else if (v.swp_base + L < v.swp_end) {
        swp_tile<DIM, KIND, TX, SEL, KB>(v, v.swp_base + L, colour, track, slots);
    }
}

// Air-conditioning boundary of one face for every case: 64 x 4 tiles over the face only
// (top/bottom: (m+2) x (n+2); west/east: (n+2) x (l+2); north/south: (m+2) x (l+2)).
__global__ __launch_bounds__(256) void k_batch_face(BatchView v, int parity, int face) {
    const int shape = face >> 1;
    const int G = (int)gridDim.x;
    const int t = xcd_logical_block((int)blockIdx.x, G);
    if (t >= v.face_tiles[shape]) return;
    const int2 ti = v.face_tile[shape][t];
    const BatchCase &C = v.cases[ti.x];
    const int gx = C.face_gx[shape];
    const int r = (ti.y % gx) * 64 + (int)threadIdx.x, q = (ti.y / gx) * 4 + (int)threadIdx.y;
    boundary_air_cell(C.g, C.a, face, C.phi, C.U[parity], C.V[parity], C.W[parity], C.P, r, q);
}

// Per case: max over its tile slots -> hist[step * stride + slot] (step from *d_step, or
// the fixed row `row` when row >= 0); the slots are cleared.  One block per case.
__global__ void k_batch_reduce(BatchView v, unsigned long long *slots, unsigned long long *hist, const int *d_step,
                               int row, int stride, int only_active) {
    const int ci = (int)blockIdx.x;
    if (ci >= v.ncases) return;
    const int t0 = v.swp_prefix[ci], t1 = v.swp_prefix[ci + 1];
    unsigned long long e = 0ull;
    for (int t = t0 + (int)threadIdx.x; t < t1; t += (int)blockDim.x) {
        const unsigned long long x = slots[t];
        e = (x > e) ? x : e;
        slots[t] = 0ull;
    }
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o2 = __shfl_xor(e, off, 64);
        e = (o2 > e) ? o2 : e;
    }
    __shared__ unsigned long long red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0 && (!only_active || v.active[ci])) {
        unsigned long long b = red[0];
        for (int q = 1; q < 4; ++q) b = (red[q] > b) ? red[q] : b;
        const int r = (row >= 0) ? row : *d_step;
        hist[(int64_t)r * stride + v.cases[ci].slot] = b;
    }
}

// Tolerance mode, after residual evaluation `cyc` of a step (one block): a case stays
// active while its residual (the max over its cells, as double bits in res_hist) exceeds
// tol and fewer than maxc cycles ran; an active case is charged K more iterations and
// one cycle.  Writes the number of still-active cases to *nactive.  Cycle 0 also zeroes
// the step's iteration and cycle counters.
__global__ void k_batch_tol_decide(BatchView v, int *active, const unsigned long long *res_hist, int *it_hist,
                                   int *cy_hist, const int *d_step, int stride, double tol, int cyc, int maxc,
                                   int K, int *nactive) {
    const int64_t row = (int64_t)(*d_step) * stride;
    int count = 0;
    for (int ci = (int)threadIdx.x; ci < v.ncases; ci += (int)blockDim.x) {
        const int slot = v.cases[ci].slot;
        if (cyc == 0) {
            it_hist[row + slot] = 0;
            cy_hist[row + slot] = 0;
        }
        if (!active[ci]) continue;
        const double r = __longlong_as_double((long long)res_hist[row + slot]);
        // a non-finite residual (NaN, or infinity) never converges: the case stops cycling
        if (r <= tol || cyc == maxc || !(fabs(r) <= 1.7976931348623157e308)) {
            active[ci] = 0;
        } else {
            it_hist[row + slot] += K;
            cy_hist[row + slot] += 1;
            ++count;
        }
    }
    __shared__ int red[256];
    red[threadIdx.x] = count;
    __syncthreads();
    for (int off = (int)blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *nactive = red[0];
}

__global__ void k_batch_fill_active(int *active, int ncases, int val) {
    for (int ci = (int)(blockIdx.x * blockDim.x + threadIdx.x); ci < ncases; ci += (int)(gridDim.x * blockDim.x))
        active[ci] = val;
}

__global__ void k_set_int(int *p, int val) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = val;
}
__global__ void k_inc_int(int *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = *p + 1;
}

}  // namespace pf

// ======================================================================================
// Host side
// ======================================================================================
using pf::BatchCase;

struct pf_batch {
    pf_desc kind{};                 // settings shared by every case
    int max_cases = 0;
    hipStream_t st = nullptr;
    char *arena = nullptr;
    int64_t arena_bytes = 0;
    std::map<int64_t, int64_t> free_list;   // offset -> size
    struct Slot {
        bool used = false;
        pf_desc d{};
        int64_t off = 0, bytes = 0;
        BatchCase bc{};
        int64_t steps = 0;
        // tolerance mode with PF_SOLVER_MG: the case's solver, its arrays inside the
        // case's arena range (FP32 for mixed precision, FP64 otherwise)
        pf::MgSolver<float> *mg32 = nullptr;
        pf::MgSolver<double> *mg64 = nullptr;
    };
    std::vector<Slot> slots;
    std::vector<int> order;         // compact case index -> slot
    BatchCase *d_cases = nullptr;
    int *d_nat_prefix = nullptr, *d_swp_prefix = nullptr, *d_active = nullptr, *d_step = nullptr;
    int2 *d_nat_tile = nullptr, *d_swp_tile = nullptr;   // per tile: case, tile within the case
    int2 *d_face_tile[3] = {nullptr, nullptr, nullptr};
    int face_tiles[3] = {0, 0, 0}, face_cap[3] = {0, 0, 0};
    int nat_tiles = 0, swp_tiles = 0, table_cap = 0, tile_cap = 0, nat_cap = 0;
    unsigned long long *d_tile_slots = nullptr;   // per-tile SOR error maxima
    unsigned long long *d_res_slots = nullptr;    // per-tile residual maxima
    unsigned long long *d_err_hist = nullptr, *d_res_hist = nullptr;
    int *d_it_hist = nullptr, *d_cy_hist = nullptr;   // tolerance mode: iterations, cycles
    int *d_nactive = nullptr;
    int hist_rows = 0;
    bool dirty = true;
    bool any_refresh = false;
    bool any_fast = false, any_slow = false;   // cases on the fast / IEEE division path
    int tile_x = 64;                // sweep tile width in slots (32, 64 or 128)
    int swp_kb = 1;                 // 3D sweep tiles: cells per thread along z (1, 2 or 4)
    // The solve runs its sweeps group by group: consecutive cases whose sweep arrays
    // together stay under group_bytes (0: one group) iterate while those arrays stay
    // in the last-level cache, then the next group.
    int64_t group_bytes = 0;        // set at creation (pf_batch_create)
    // Groups may iterate on side streams, forked from and joined back to st, so that one
    // group's sweep launches overlap the launch and drain of another's.
    int nstreams = 1;
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
    hipStream_t lst = nullptr;      // stream of the sweep / refresh launches being enqueued
    struct Group {
        int swp0, swp1, nat0, nat1;
    };
    std::vector<Group> groups;
    int parity = 0;
    pf_batch_config cfg{};
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    // PF_SOLVER_MG: per-step statistics of the last pf_batch_step, kept on the host
    std::vector<double> mg_res;
    std::vector<int> mg_cyc, mg_vc;
    std::string msg;
};

#define PFB_TRY(b, call)                                                                        \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            (b)->msg = std::string(#call) + ": " + hipGetErrorString(e_);                       \
            return PF_EHIP;                                                                     \
        }                                                                                       \
    } while (0)

// For asynchronous copies from or into host vectors: on a failure the stream is drained
// before the early return, so no copy still in flight outlives the vector it uses.
#define PFB_TRY_SYNC(b, call)                                                                   \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            (void)hipStreamSynchronize((b)->st);                                                \
            (b)->msg = std::string(#call) + ": " + hipGetErrorString(e_);                       \
            return PF_EHIP;                                                                     \
        }                                                                                       \
    } while (0)

static thread_local std::string g_batch_create_error;

static int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

// Bytes of one case inside a batch arena (velocity pairs, porosity, div, p/bb/porosity in
// colour-split FP64, FP32 e/r/porosity for mixed precision, the interior output buffer).
static bool mg_mode(const pf_desc &d) { return d.residual_tol > 0 && d.solver == PF_SOLVER_MG; }

static int64_t case_bytes(const pf_desc &d) {
    const PfGeom g = pf_make_geom(d.dim, d.m, d.n, d.l);
    const int nat = (d.dim == 3) ? 8 : 6;
    int64_t b = 0;
    b += nat * align256(8 * g.NS);
    b += 3 * align256(8 * 2 * g.HS);
    if (d.precision == PF_PREC_MIXED) b += 3 * align256(4 * 2 * g.HS);
    b += align256(8 * (int64_t)d.m * d.n * (d.dim == 3 ? d.l : 1));
    if (mg_mode(d)) b += align256(pf::mg_bytes(g, d.bc, d.precision == PF_PREC_MIXED));
    return b;
}

static void slot_release_mg(pf_batch::Slot &S) {
    if (S.mg32) { pf::mg_destroy(*S.mg32); delete S.mg32; S.mg32 = nullptr; }
    if (S.mg64) { pf::mg_destroy(*S.mg64); delete S.mg64; S.mg64 = nullptr; }
}

extern "C" int64_t pf_batch_case_bytes(const pf_desc *d) {
    std::string why;
    if (!pf::desc_valid(d, why)) return -1;
    return case_bytes(*d);
}

static bool same_kind(const pf_desc &a, const pf_desc &b) {
    return a.dim == b.dim && a.bc == b.bc && a.precision == b.precision && a.iter_max == b.iter_max &&
           a.inner_iters == b.inner_iters && a.max_cycles == b.max_cycles && a.residual_tol == b.residual_tol &&
           a.solver == b.solver && a.inner_tol == b.inner_tol;
}

// Wait for every stream of the batch (its own and the two side streams the sweep groups
// and the multigrid solves run on) before tables, histories or graphs are replaced.
static hipError_t drain(pf_batch *b) {
    hipError_t e = b->st ? hipStreamSynchronize(b->st) : hipSuccess;
    for (int q = 0; q < 2 && e == hipSuccess; ++q)
        if (b->side[q]) e = hipStreamSynchronize(b->side[q]);
    return e;
}

static void drop_graphs(pf_batch *b) {
    for (int p = 0; p < 2; ++p) {
        if (b->gexec[p]) (void)hipGraphExecDestroy(b->gexec[p]);
        if (b->graph[p]) (void)hipGraphDestroy(b->graph[p]);
        b->gexec[p] = nullptr;
        b->graph[p] = nullptr;
    }
}

static void destroy_streams(pf_batch *b) {
    for (int q = 0; q < 2; ++q) {
        if (b->ev_join[q]) (void)hipEventDestroy(b->ev_join[q]);
        if (b->side[q]) (void)hipStreamDestroy(b->side[q]);
    }
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->st) (void)hipStreamDestroy(b->st);
}

extern "C" int pf_batch_create(const pf_desc *kind, int max_cases, int64_t arena_bytes, pf_batch **out) {
    if (!out) return PF_EINVAL;
    *out = nullptr;
    std::string why;
    if (!pf::desc_valid(kind, why)) { g_batch_create_error = why; return PF_EINVAL; }
    if (max_cases < 1 || arena_bytes < 4096) { g_batch_create_error = "max_cases >= 1, arena_bytes >= 4096"; return PF_EINVAL; }
    pf_batch *b = new (std::nothrow) pf_batch();
    if (!b) return PF_ENOMEM;
    b->kind = *kind;
    b->max_cases = max_cases;
    b->slots.resize((size_t)max_cases);
    b->cfg.persistent_blocks = 0;
    b->cfg.use_graph = 1;
    b->cfg.inner_iters = kind->inner_iters;
    // Measured on eight served 3D rooms (64^3 to 160x128x96, 10 M cells): FP32 sweep tiles
    // of 32 slots with two cells per thread run 1.25x faster than 64 slots with one;
    // FP64 sweeps lose with more than one cell per thread (register pressure).
    b->cfg.tile_x = (kind->dim == 3) ? 32 : 64;
    b->tile_x = b->cfg.tile_x;
    b->swp_kb = (kind->dim == 3 && kind->precision == PF_PREC_MIXED) ? 2 : 1;
    // Sweep groups against the 256 MB Infinity Cache.  FP64: groups of at most 240 MB on
    // one stream, 1.24x the rate of sweeping all cases in every launch on the 16 served
    // rooms (56 M cells); 320 MB and more lose it, side streams lose 3 %.  Mixed: groups
    // of at most 96 MB dealt to two side streams, 1.03x (16 rooms) and 1.08x (512 2D cases)
    // the rate of 192 MB groups on one stream, itself 1.17x the ungrouped rate.
    const bool mixed = kind->precision == PF_PREC_MIXED;
    b->group_bytes = (int64_t)(mixed ? 96 : 240) << 20;
    b->nstreams = mixed ? 2 : 1;
    {
        const char *gk = std::getenv("PF_BATCH_GROUP_KB");   // 0: every case in every launch
        if (gk && *gk) b->group_bytes = (int64_t)std::max(0, std::atoi(gk)) << 10;
        const char *ns = std::getenv("PF_BATCH_STREAMS");
        if (ns && *ns) b->nstreams = std::max(1, std::min(2, std::atoi(ns)));
        const char *kb = std::getenv("PF_BATCH_KB");
        if (kb && *kb) {
            const int v = std::atoi(kb);
            b->swp_kb = (v == 1 || v == 2 || v == 4) ? v : 1;
        }
    }
    b->arena_bytes = align256(arena_bytes);
    hipError_t e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking);
    for (int q = 0; q < 2 && e == hipSuccess; ++q) e = hipStreamCreateWithFlags(&b->side[q], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
    for (int q = 0; q < 2 && e == hipSuccess; ++q) e = hipEventCreateWithFlags(&b->ev_join[q], hipEventDisableTiming);
    b->lst = b->st;
    if (e == hipSuccess) e = hipMalloc((void **)&b->arena, (size_t)b->arena_bytes);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_step, sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_active, sizeof(int) * (size_t)max_cases);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_nactive, sizeof(int));
    if (e != hipSuccess) {
        g_batch_create_error = std::string("batch allocation: ") + hipGetErrorString(e);
        if (b->arena) (void)hipFree(b->arena);
        if (b->d_step) (void)hipFree(b->d_step);
        if (b->d_active) (void)hipFree(b->d_active);
        if (b->d_nactive) (void)hipFree(b->d_nactive);
        destroy_streams(b);
        delete b;
        return PF_ENOMEM;
    }
    b->free_list[0] = b->arena_bytes;
    *out = b;
    return PF_OK;
}

extern "C" int pf_batch_destroy(pf_batch *b) {
    if (!b) return PF_EINVAL;
    if (b->st) (void)hipStreamSynchronize(b->st);
    for (int q = 0; q < 2; ++q)
        if (b->side[q]) (void)hipStreamSynchronize(b->side[q]);
    for (auto &s : b->slots) slot_release_mg(s);
    drop_graphs(b);
    void *ptrs[] = {b->arena, b->d_cases, b->d_nat_prefix, b->d_swp_prefix, b->d_nat_tile, b->d_swp_tile,
                    b->d_face_tile[0], b->d_face_tile[1], b->d_face_tile[2],
                    b->d_active, b->d_step,
                    b->d_nactive, b->d_tile_slots, b->d_res_slots, b->d_err_hist, b->d_res_hist,
                    b->d_it_hist, b->d_cy_hist};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (int q = 0; q < 2; ++q)
        if (b->side[q]) (void)hipStreamSynchronize(b->side[q]);
    destroy_streams(b);
    delete b;
    return PF_OK;
}

extern "C" const char *pf_batch_last_error(const pf_batch *b) { return b ? b->msg.c_str() : g_batch_create_error.c_str(); }

static int64_t arena_alloc(pf_batch *b, int64_t bytes) {
    for (auto it = b->free_list.begin(); it != b->free_list.end(); ++it) {
        if (it->second >= bytes) {
            const int64_t off = it->first, rest = it->second - bytes;
            b->free_list.erase(it);
            if (rest > 0) b->free_list[off + bytes] = rest;
            return off;
        }
    }
    return -1;
}

static void arena_free(pf_batch *b, int64_t off, int64_t bytes) {
    auto it = b->free_list.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != b->free_list.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        b->free_list.erase(nx);
    }
    if (it != b->free_list.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            b->free_list.erase(it);
        }
    }
}

extern "C" int64_t pf_batch_free_bytes(const pf_batch *b) {
    if (!b) return -1;
    int64_t best = 0;
    for (auto &kv : b->free_list) best = std::max(best, kv.second);
    return best;
}

// Cells per thread of the 3D sweep tiles in the current grid kind (the persistent grid
// walks one-cell tiles).
static int tile_kb(const pf_batch *b) {
    return (b->kind.dim == 3 && b->cfg.persistent_blocks == 0) ? b->swp_kb : 1;
}

static int rebuild_tables(pf_batch *b) {
    PFB_TRY(b, drain(b));   // launches in flight read the tables replaced below
    b->order.clear();
    for (int s = 0; s < b->max_cases; ++s)
        if (b->slots[(size_t)s].used) b->order.push_back(s);
    const int nc = (int)b->order.size();
    std::vector<BatchCase> cs;
    std::vector<int> np(1, 0), sp(1, 0);
    std::vector<int> fp[3] = {std::vector<int>(1, 0), std::vector<int>(1, 0), std::vector<int>(1, 0)};
    b->any_refresh = false;
    b->any_fast = b->any_slow = false;
    for (int s : b->order) {
        BatchCase bc = b->slots[(size_t)s].bc;
        const PfGeom &g = bc.g;
        bc.nat_gx = (g.m + 63) / 64;   // i = 1..m; the x ghosts ride with i = 1 and i = m
        bc.nat_gy = (g.n + 2 + 3) / 4;
        const int TX = b->tile_x, RY = 256 / TX;
        bc.swp_gx = ((g.m + 1) / 2 + TX - 1) / TX;
        bc.swp_gy = (g.dim == 3) ? (g.n + RY / 2 - 1) / (RY / 2) : (g.n + RY - 1) / RY;
        const int kb = tile_kb(b);
        const int swp_gz = (g.dim == 3) ? (g.l + 2 * kb - 1) / (2 * kb) : 1;
        np.push_back(np.back() + bc.nat_gx * bc.nat_gy * g.Z);
        {
            const int X = g.m + 2, Y = g.n + 2, Zf = g.l + 2;
            const int R[3] = {X, Y, X}, Q[3] = {Y, Zf, Zf};
            for (int f = 0; f < 3; ++f) {
                bc.face_gx[f] = (R[f] + 63) / 64;
                fp[f].push_back(fp[f].back() + ((g.dim == 3 && b->kind.bc == PF_BC_AIR) ? bc.face_gx[f] * ((Q[f] + 3) / 4) : 0));
            }
        }
        sp.push_back(sp.back() + bc.swp_gx * bc.swp_gy * swp_gz);
        b->any_refresh = b->any_refresh || bc.ref_y || bc.ref_z;
        b->any_fast = b->any_fast || bc.fast;
        b->any_slow = b->any_slow || !bc.fast;
        b->slots[(size_t)s].bc = bc;
        cs.push_back(bc);
    }
    b->nat_tiles = np.back();
    b->swp_tiles = sp.back();
    b->groups.clear();
    {
        // sweep arrays per slot: unknown (both colours), porosity, right-hand side
        const int64_t per_slot = (b->kind.precision == PF_PREC_MIXED) ? 3 * 4 : 3 * 8;
        int64_t acc = 0;
        int start = 0;
        for (int c = 0; c < nc; ++c) {
            const int64_t bytes = 2 * cs[(size_t)c].g.HS * per_slot;
            if (c > start && b->group_bytes > 0 && acc + bytes > b->group_bytes) {
                b->groups.push_back({sp[(size_t)start], sp[(size_t)c], np[(size_t)start], np[(size_t)c]});
                start = c;
                acc = 0;
            }
            acc += bytes;
        }
        if (nc > 0) b->groups.push_back({sp[(size_t)start], sp[(size_t)nc], np[(size_t)start], np[(size_t)nc]});
    }
    for (int f = 0; f < 3; ++f) b->face_tiles[f] = fp[f].back();
    if (b->table_cap < nc + 1) {
        if (b->d_cases) (void)hipFree(b->d_cases);
        if (b->d_nat_prefix) (void)hipFree(b->d_nat_prefix);
        if (b->d_swp_prefix) (void)hipFree(b->d_swp_prefix);
        b->d_cases = nullptr;
        b->d_nat_prefix = b->d_swp_prefix = nullptr;
        const int cap = std::max(nc + 1, b->max_cases + 1);
        PFB_TRY(b, hipMalloc((void **)&b->d_cases, sizeof(BatchCase) * (size_t)cap));
        PFB_TRY(b, hipMalloc((void **)&b->d_nat_prefix, sizeof(int) * (size_t)cap));
        PFB_TRY(b, hipMalloc((void **)&b->d_swp_prefix, sizeof(int) * (size_t)cap));
        b->table_cap = cap;
    }
    if (b->tile_cap < b->swp_tiles + 1) {
        if (b->d_tile_slots) (void)hipFree(b->d_tile_slots);
        if (b->d_res_slots) (void)hipFree(b->d_res_slots);
        if (b->d_swp_tile) (void)hipFree(b->d_swp_tile);
        b->d_tile_slots = b->d_res_slots = nullptr;
        b->d_swp_tile = nullptr;
        const int cap = b->swp_tiles + 1024;
        PFB_TRY(b, hipMalloc((void **)&b->d_tile_slots, sizeof(unsigned long long) * (size_t)cap));
        PFB_TRY(b, hipMalloc((void **)&b->d_res_slots, sizeof(unsigned long long) * (size_t)cap));
        PFB_TRY(b, hipMalloc((void **)&b->d_swp_tile, sizeof(int2) * (size_t)cap));
        PFB_TRY(b, hipMemsetAsync(b->d_tile_slots, 0, sizeof(unsigned long long) * (size_t)cap, b->st));
        PFB_TRY(b, hipMemsetAsync(b->d_res_slots, 0, sizeof(unsigned long long) * (size_t)cap, b->st));
        b->tile_cap = cap;
    }
    if (b->nat_cap < b->nat_tiles + 1) {
        if (b->d_nat_tile) (void)hipFree(b->d_nat_tile);
        b->d_nat_tile = nullptr;
        const int cap = b->nat_tiles + 4096;
        PFB_TRY(b, hipMalloc((void **)&b->d_nat_tile, sizeof(int2) * (size_t)cap));
        b->nat_cap = cap;
    }
    for (int f = 0; f < 3; ++f) {
        if (b->face_cap[f] < b->face_tiles[f] + 1) {
            if (b->d_face_tile[f]) (void)hipFree(b->d_face_tile[f]);
            b->d_face_tile[f] = nullptr;
            const int cap = b->face_tiles[f] + 1024;
            PFB_TRY(b, hipMalloc((void **)&b->d_face_tile[f], sizeof(int2) * (size_t)cap));
            b->face_cap[f] = cap;
        }
    }
    // host sources of the asynchronous copies live until the synchronisation below
    std::vector<int2> nt((size_t)b->nat_tiles), st((size_t)b->swp_tiles);
    std::vector<int2> ft[3];
    for (int f = 0; f < 3; ++f) {
        ft[f].resize((size_t)b->face_tiles[f]);
        for (int c = 0; c < nc; ++c)
            for (int t = fp[f][(size_t)c]; t < fp[f][(size_t)c + 1]; ++t) ft[f][(size_t)t] = make_int2(c, t - fp[f][(size_t)c]);
        if (b->face_tiles[f] > 0)
            PFB_TRY_SYNC(b, hipMemcpyAsync(b->d_face_tile[f], ft[f].data(), sizeof(int2) * ft[f].size(), hipMemcpyHostToDevice,
                                      b->st));
    }
    std::vector<int> ones((size_t)nc, 1);
    if (nc > 0) {
        for (int c = 0; c < nc; ++c) {
            for (int t = np[(size_t)c]; t < np[(size_t)c + 1]; ++t) nt[(size_t)t] = make_int2(c, t - np[(size_t)c]);
            for (int t = sp[(size_t)c]; t < sp[(size_t)c + 1]; ++t) st[(size_t)t] = make_int2(c, t - sp[(size_t)c]);
        }
        PFB_TRY_SYNC(b, hipMemcpyAsync(b->d_nat_tile, nt.data(), sizeof(int2) * nt.size(), hipMemcpyHostToDevice, b->st));
        PFB_TRY_SYNC(b, hipMemcpyAsync(b->d_swp_tile, st.data(), sizeof(int2) * st.size(), hipMemcpyHostToDevice, b->st));
        PFB_TRY_SYNC(b, hipMemcpyAsync(b->d_cases, cs.data(), sizeof(BatchCase) * (size_t)nc, hipMemcpyHostToDevice, b->st));
        PFB_TRY_SYNC(b, hipMemcpyAsync(b->d_nat_prefix, np.data(), sizeof(int) * np.size(), hipMemcpyHostToDevice, b->st));
        PFB_TRY_SYNC(b, hipMemcpyAsync(b->d_swp_prefix, sp.data(), sizeof(int) * sp.size(), hipMemcpyHostToDevice, b->st));
        PFB_TRY_SYNC(b, hipMemcpyAsync(b->d_active, ones.data(), sizeof(int) * (size_t)nc, hipMemcpyHostToDevice, b->st));
    }
    PFB_TRY(b, hipStreamSynchronize(b->st));
    drop_graphs(b);
    b->dirty = false;
    return PF_OK;
}

static pf::BatchView view(const pf_batch *b) {
    pf::BatchView v;
    v.cases = b->d_cases;
    v.nat_prefix = b->d_nat_prefix;
    v.swp_prefix = b->d_swp_prefix;
    v.nat_tile = b->d_nat_tile;
    v.swp_tile = b->d_swp_tile;
    for (int f = 0; f < 3; ++f) {
        v.face_tile[f] = b->d_face_tile[f];
        v.face_tiles[f] = b->face_tiles[f];
    }
    v.active = b->d_active;
    v.ncases = (int)b->order.size();
    v.nat_tiles = b->nat_tiles;
    v.swp_tiles = b->swp_tiles;
    v.swp_base = 0;
    v.swp_end = b->swp_tiles;
    v.nat_base = 0;
    v.nat_end = b->nat_tiles;
    return v;
}

extern "C" int pf_batch_add(pf_batch *b, const pf_desc *d, const double *raw, double threshold, int *slot_out) {
    if (!b || !d || !raw || !slot_out) return PF_EINVAL;
    std::string why;
    if (!pf::desc_valid(d, why)) { b->msg = why; return PF_EINVAL; }
    if (!same_kind(b->kind, *d)) {
        b->msg = "case differs from the batch kind (dim, bc, precision, iteration settings)";
        return PF_EINVAL;
    }
    int slot = -1;
    for (int s = 0; s < b->max_cases; ++s)
        if (!b->slots[(size_t)s].used) { slot = s; break; }
    if (slot < 0) { b->msg = "batch is full (max_cases)"; return PF_ENOMEM; }
    const int64_t bytes = case_bytes(*d);
    const int64_t off = arena_alloc(b, bytes);
    if (off < 0) { b->msg = "batch arena has no room for the case"; return PF_ENOMEM; }
    pf_batch::Slot &S = b->slots[(size_t)slot];
    S.d = *d;
    S.off = off;
    S.bytes = bytes;
    S.steps = 0;
    BatchCase &bc = S.bc;
    bc = BatchCase{};
    bc.g = pf_make_geom(d->dim, d->m, d->n, d->l);
    bc.a = pf::make_step_args(*d);
    bc.slot = slot;
    const PfGeom &g = bc.g;
    char *p = b->arena + off;
    auto takeD = [&](int64_t n) { double *r = (double *)p; p += align256(8 * n); return r; };
    auto takeF = [&](int64_t n) { float *r = (float *)p; p += align256(4 * n); return r; };
    bc.U[0] = takeD(g.NS); bc.U[1] = takeD(g.NS);
    bc.V[0] = takeD(g.NS); bc.V[1] = takeD(g.NS);
    if (g.dim == 3) { bc.W[0] = takeD(g.NS); bc.W[1] = takeD(g.NS); }
    bc.phi = takeD(g.NS);
    bc.div = takeD(g.NS);
    bc.P = takeD(2 * g.HS);
    bc.BB = takeD(2 * g.HS);
    bc.PHIs = takeD(2 * g.HS);
    if (d->precision == PF_PREC_MIXED) {
        bc.E = takeF(2 * g.HS);
        bc.R = takeF(2 * g.HS);
        bc.PHI32 = takeF(2 * g.HS);
    }
    bc.dout = takeD((int64_t)g.m * g.n * g.l);
    char *mg_mem = mg_mode(*d) ? p : nullptr;   // the solver's arrays, carved below
    const bool periodic = d->bc == PF_BC_UNIFORM;
    bc.fuse_y = periodic && (g.n % 2 == 0);
    bc.fuse_z = periodic && g.dim == 3 && (g.l % 2 == 0);
    bc.ref_y = periodic && !bc.fuse_y;
    bc.ref_z = periodic && g.dim == 3 && !bc.fuse_z;
    PFB_TRY(b, hipMemsetAsync(b->arena + off, 0, (size_t)bytes, b->st));
    // porosity, its colour-split copies, the fast-division check
    std::vector<double> phi;
    pf::host_porosity(g, d->bc, raw, threshold, phi);
    PFB_TRY_SYNC(b, hipMemcpyAsync(bc.phi, phi.data(), phi.size() * 8, hipMemcpyHostToDevice, b->st));
    pf::launch_split(b->st, g, bc.phi, bc.PHIs);
    if (bc.PHI32) pf::launch_split32(b->st, g, bc.phi, bc.PHI32);
    int *bad = nullptr, hbad = 1, zero = 0;
    PFB_TRY(b, hipMalloc((void **)&bad, sizeof(int)));
    hipError_t e = hipMemcpyAsync(bad, &zero, sizeof(int), hipMemcpyHostToDevice, b->st);
    if (e == hipSuccess) {
        pf::launch_faces_check_direct(b->st, g, bc.phi, bc.a, bad);
        e = hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, b->st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->st);
    (void)hipFree(bad);
    PFB_TRY(b, e);
    bc.fast = (hbad == 0);
    if (mg_mem) {
        // the single-case solve's launch settings for this grid (pf_create)
        const int64_t cells = (int64_t)g.m * g.n * (g.dim == 3 ? g.l : 1);
        pf::MgCase mc;
        mc.g = g;
        mc.a = bc.a;
        mc.P = bc.P;
        mc.BB = bc.BB;
        mc.phi = bc.phi;
        mc.PHIs = bc.PHIs;
        mc.PHI32 = bc.PHI32;
        mc.fast = bc.fast;
        mc.kb32 = (cells * 12 <= ((int64_t)128 << 20)) ? 2 : 1;
        mc.kb64 = 1;
        mc.xw = pf::sweep_wave_width(g);
        mc.remap32 = 1;
        mc.remap64 = (cells * 24 > ((int64_t)512 << 20)) ? 0 : 1;
        std::string mwhy;
        bool ok;
        if (d->precision == PF_PREC_MIXED) {
            S.mg32 = new (std::nothrow) pf::MgSolver<float>();
            ok = S.mg32 && pf::mg_create(*S.mg32, mc, &mwhy, mg_mem);
        } else {
            S.mg64 = new (std::nothrow) pf::MgSolver<double>();
            ok = S.mg64 && pf::mg_create(*S.mg64, mc, &mwhy, mg_mem);
        }
        if (!ok) {
            slot_release_mg(S);
            arena_free(b, off, bytes);
            b->msg = mwhy.empty() ? std::string("multigrid solver set-up failed") : mwhy;
            return PF_ENOMEM;
        }
    }
    // initial_conditions + boundary into the buffers the next step reads as "old"
    const int q = 1 - b->parity;
    double u0, v0;
    pf::initial_values(*d, u0, v0);
    std::vector<double> buf;
    pf::fill_interior(g, u0, false, buf);
    if (d->inlet_porosity) pf::scale_interior(g, phi, buf);
    PFB_TRY(b, hipMemcpy(bc.U[q], buf.data(), buf.size() * 8, hipMemcpyHostToDevice));
    pf::fill_interior(g, v0, false, buf);
    if (d->inlet_porosity) pf::scale_interior(g, phi, buf);
    PFB_TRY(b, hipMemcpy(bc.V[q], buf.data(), buf.size() * 8, hipMemcpyHostToDevice));
    if (g.dim == 3) {
        pf::fill_interior(g, 0.0, false, buf);
        PFB_TRY(b, hipMemcpy(bc.W[q], buf.data(), buf.size() * 8, hipMemcpyHostToDevice));
    }
    pf::fill_interior(g, d->outlet_pressure, true, buf);
    PFB_TRY(b, hipMemcpy(bc.P, buf.data(), buf.size() * 8, hipMemcpyHostToDevice));
    pf::launch_boundary(b->st, g, bc.a, bc.phi, bc.U[q], bc.V[q], bc.W[q], bc.P);
    PFB_TRY(b, hipGetLastError());
    PFB_TRY(b, hipStreamSynchronize(b->st));
    S.used = true;
    b->dirty = true;
    *slot_out = slot;
    return PF_OK;
}

extern "C" int pf_batch_remove(pf_batch *b, int slot) {
    if (!b || slot < 0 || slot >= b->max_cases || !b->slots[(size_t)slot].used) return PF_EINVAL;
    PFB_TRY(b, hipStreamSynchronize(b->st));
    pf_batch::Slot &S = b->slots[(size_t)slot];
    PFB_TRY(b, drain(b));
    slot_release_mg(S);
    arena_free(b, S.off, S.bytes);
    S.used = false;
    b->dirty = true;
    return PF_OK;
}

extern "C" int pf_batch_count(const pf_batch *b) {
    if (!b) return PF_EINVAL;
    int n = 0;
    for (auto &s : b->slots) n += s.used ? 1 : 0;
    return n;
}

static int grid_blocks(const pf_batch *b, int tiles) {
    if (tiles <= 0) return 0;
    const int P = b->cfg.persistent_blocks;
    return (P > 0 && P < tiles) ? P : tiles;
}

static void launch_nat(pf_batch *b, int op, int parity, int aux, const pf_batch::Group *grp = nullptr) {
    pf::BatchView v = view(b);
    if (grp) {
        v.nat_base = grp->nat0;
        v.nat_end = grp->nat1;
    }
    const int nb = grid_blocks(b, v.nat_end - v.nat_base);
    if (nb == 0) return;
    dim3 blk(64, 4, 1), gr((unsigned)nb);
#define PFB_NAT(OP)                                                                              \
    do {                                                                                         \
        if (b->cfg.persistent_blocks > 0)                                                        \
            hipLaunchKernelGGL((pf::k_batch_nat<OP, true>), gr, blk, 0, b->lst, v, parity, aux); \
        else                                                                                     \
            hipLaunchKernelGGL((pf::k_batch_nat<OP, false>), gr, blk, 0, b->lst, v, parity, aux); \
    } while (0)
    switch (op) {
        case pf::NAT_DIV: PFB_NAT(pf::NAT_DIV); break;
        case pf::NAT_PREDICT:
            if (b->kind.dim == 3) PFB_NAT(pf::NAT_PREDICT);
            else PFB_NAT(pf::NAT_PREDICT2);
            break;
        case pf::NAT_RHS: PFB_NAT(pf::NAT_RHS); break;
        case pf::NAT_PROJECT: PFB_NAT(pf::NAT_PROJECT); break;
        case pf::NAT_BOUNDARY_UNIFORM: PFB_NAT(pf::NAT_BOUNDARY_UNIFORM); break;
        case pf::NAT_BOUNDARY_AIR: PFB_NAT(pf::NAT_BOUNDARY_AIR); break;
        case pf::NAT_UPDATE: PFB_NAT(pf::NAT_UPDATE); break;
        case pf::NAT_REFRESH_P: PFB_NAT(pf::NAT_REFRESH_P); break;
        default: PFB_NAT(pf::NAT_REFRESH_E); break;
    }
#undef PFB_NAT
}

template <int D, int K, int T, int SEL>
static void launch_swp_t(pf_batch *b, dim3 gr, const pf::BatchView &v, int colour, int track,
                         unsigned long long *slots) {
    // 3D tiles cover 2 * swp_kb planes (one grid kind per batch: the persistent grid keeps
    // one cell per thread)
    if (b->cfg.persistent_blocks > 0)
        hipLaunchKernelGGL((pf::k_batch_swp<D, K, T, SEL, true, 1>), gr, dim3(256), 0, b->lst, v, colour, track, slots);
    else if (D == 3 && b->swp_kb == 2)
        hipLaunchKernelGGL((pf::k_batch_swp<D, K, T, SEL, false, D == 3 ? 2 : 1>), gr, dim3(256), 0, b->lst, v, colour,
                           track, slots);
    else if (D == 3 && b->swp_kb == 4)
        hipLaunchKernelGGL((pf::k_batch_swp<D, K, T, SEL, false, D == 3 ? 4 : 1>), gr, dim3(256), 0, b->lst, v, colour,
                           track, slots);
    else
        hipLaunchKernelGGL((pf::k_batch_swp<D, K, T, SEL, false, 1>), gr, dim3(256), 0, b->lst, v, colour, track, slots);
}

template <int D, int K, int SEL>
static void launch_swp_tx(pf_batch *b, dim3 gr, const pf::BatchView &v, int colour, int track,
                          unsigned long long *slots) {
    if (b->tile_x == 32) launch_swp_t<D, K, 32, SEL>(b, gr, v, colour, track, slots);
    else if (b->tile_x == 128) launch_swp_t<D, K, 128, SEL>(b, gr, v, colour, track, slots);
    else launch_swp_t<D, K, 64, SEL>(b, gr, v, colour, track, slots);
}

template <int D>
static void launch_swp_dim(pf_batch *b, int kind, dim3 gr, const pf::BatchView &v, int colour, int track,
                           unsigned long long *slots) {
    if (kind == pf::SWP_FP32) {
        launch_swp_tx<D, pf::SWP_FP32, 1>(b, gr, v, colour, track, slots);
        return;
    }
    // fast-division cases and the others in separate launches (usually only the first)
    if (kind == pf::SWP_FP64) {
        if (b->any_fast) launch_swp_tx<D, pf::SWP_FP64, 1>(b, gr, v, colour, track, slots);
        if (b->any_slow) launch_swp_tx<D, pf::SWP_FP64, 0>(b, gr, v, colour, track, slots);
    } else {
        if (b->any_fast) launch_swp_tx<D, pf::SWP_RESIDUAL, 1>(b, gr, v, colour, track, slots);
        if (b->any_slow) launch_swp_tx<D, pf::SWP_RESIDUAL, 0>(b, gr, v, colour, track, slots);
    }
}

static void launch_swp(pf_batch *b, int kind, int colour, int track, const pf_batch::Group *grp = nullptr) {
    pf::BatchView v = view(b);
    if (grp) {
        v.swp_base = grp->swp0;
        v.swp_end = grp->swp1;
    }
    const int nb = grid_blocks(b, v.swp_end - v.swp_base);
    if (nb == 0) return;
    const dim3 gr((unsigned)nb);
    unsigned long long *slots = (kind == pf::SWP_RESIDUAL) ? b->d_res_slots : b->d_tile_slots;
    if (b->kind.dim == 3) launch_swp_dim<3>(b, kind, gr, v, colour, track, slots);
    else launch_swp_dim<2>(b, kind, gr, v, colour, track, slots);
}

static void launch_reduce(pf_batch *b, unsigned long long *slots, unsigned long long *hist, int row,
                          int only_active = 0) {
    const pf::BatchView v = view(b);
    if (v.ncases == 0) return;
    hipLaunchKernelGGL(pf::k_batch_reduce, dim3((unsigned)v.ncases), dim3(256), 0, b->st, v, slots, hist, b->d_step,
                       row, b->max_cases, only_active);
}

static void launch_fill_active(pf_batch *b, int val) {
    const pf::BatchView v = view(b);
    if (v.ncases == 0) return;
    hipLaunchKernelGGL(pf::k_batch_fill_active, dim3((unsigned)((v.ncases + 255) / 256)), dim3(256), 0, b->st,
                       b->d_active, v.ncases, val);
}

// `count` red/black iterations of every group.  With side streams, the groups are dealt to
// them in turn after a fork from st and joined back to st afterwards (also inside a graph
// capture, where the fork and join become graph edges).
static void group_sweeps(pf_batch *b, int kind, int count) {
    const int dim = b->kind.dim;
    const int first = (dim == 3) ? 0 : 1;
    auto one = [&](const pf_batch::Group &grp) {
        for (int it = 0; it < count; ++it)
            for (int half = 0; half < 2; ++half) {
                const int colour = half == 0 ? first : 1 - first;
                launch_swp(b, kind, colour, (dim == 2) || (half == 1), &grp);
                if (b->any_refresh) launch_nat(b, kind == pf::SWP_FP64 ? pf::NAT_REFRESH_P : pf::NAT_REFRESH_E, 0, 0, &grp);
            }
    };
    const int S = (b->nstreams > 1 && b->groups.size() > 1) ? b->nstreams : 1;
    if (S == 1) {
        for (const pf_batch::Group &grp : b->groups) one(grp);
        return;
    }
    (void)hipEventRecord(b->ev_fork, b->st);
    for (int q = 0; q < S; ++q) (void)hipStreamWaitEvent(b->side[q], b->ev_fork, 0);
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        b->lst = b->side[gi % (size_t)S];
        one(b->groups[gi]);
    }
    b->lst = b->st;
    for (int q = 0; q < S; ++q) {
        (void)hipEventRecord(b->ev_join[q], b->side[q]);
        (void)hipStreamWaitEvent(b->st, b->ev_join[q], 0);
    }
}

// Iterations of the pressure solve (fixed count): FP64 reference sweeps, or FP64 residual +
// FP32 sweeps + FP64 update cycles.
static void enqueue_solve_fixed(pf_batch *b) {
    const int iters = b->kind.iter_max;
    auto sweeps = [&](int kind, int count) { group_sweeps(b, kind, count); };
    if (b->kind.precision == PF_PREC_FP64) {
        sweeps(pf::SWP_FP64, iters);
        return;
    }
    const int K = b->cfg.inner_iters > 0 ? b->cfg.inner_iters : (iters > 0 ? iters : 1);
    for (int done = 0; done < iters;) {
        const int kk = std::min(K, iters - done);
        launch_swp(b, pf::SWP_RESIDUAL, 0, 0);
        launch_swp(b, pf::SWP_RESIDUAL, 1, 0);
        sweeps(pf::SWP_FP32, kk);
        launch_nat(b, pf::NAT_UPDATE, 0, 0);
        done += kk;
    }
}

static void enqueue_pre(pf_batch *b, int p) {
    launch_nat(b, pf::NAT_DIV, p, 0);
    launch_nat(b, pf::NAT_PREDICT, p, 0);
    launch_nat(b, pf::NAT_RHS, p, 0);
}

static void enqueue_post(pf_batch *b, int p) {
    launch_nat(b, pf::NAT_PROJECT, p, 0);
    if (b->kind.bc == PF_BC_UNIFORM) {
        launch_nat(b, pf::NAT_BOUNDARY_UNIFORM, p, 0);
    } else {
        // faces in the reference's order (top, bottom, west, east, north, south): later
        // faces overwrite the edges they share with earlier ones
        const pf::BatchView v = view(b);
        for (int f = 0; f < 6; ++f) {
            const int n = v.face_tiles[f >> 1];
            if (n > 0) hipLaunchKernelGGL(pf::k_batch_face, dim3((unsigned)n), dim3(64, 4, 1), 0, b->st, v, p, f);
        }
    }
}

static void enqueue_step_fixed(pf_batch *b, int p) {
    enqueue_pre(b, p);
    enqueue_solve_fixed(b);
    launch_reduce(b, b->d_tile_slots, b->d_err_hist, -1);
    if (b->kind.report_residual && b->kind.precision == PF_PREC_MIXED) {
        launch_swp(b, pf::SWP_RESIDUAL, 0, 1);
        launch_swp(b, pf::SWP_RESIDUAL, 1, 1);
        launch_reduce(b, b->d_res_slots, b->d_res_hist, -1);
    }
    enqueue_post(b, p);
    hipLaunchKernelGGL(pf::k_inc_int, dim3(1), dim3(64), 0, b->st, b->d_step);
}

static int ensure_hist(pf_batch *b, int rows) {
    if (b->hist_rows >= rows) return PF_OK;
    PFB_TRY(b, drain(b));
    if (b->d_err_hist) (void)hipFree(b->d_err_hist);
    if (b->d_res_hist) (void)hipFree(b->d_res_hist);
    if (b->d_it_hist) (void)hipFree(b->d_it_hist);
    if (b->d_cy_hist) (void)hipFree(b->d_cy_hist);
    b->d_err_hist = b->d_res_hist = nullptr;
    b->d_it_hist = b->d_cy_hist = nullptr;
    b->hist_rows = 0;
    const size_t n = (size_t)rows * (size_t)b->max_cases;
    PFB_TRY(b, hipMalloc((void **)&b->d_err_hist, n * sizeof(unsigned long long)));
    PFB_TRY(b, hipMalloc((void **)&b->d_res_hist, n * sizeof(unsigned long long)));
    PFB_TRY(b, hipMalloc((void **)&b->d_it_hist, n * sizeof(int)));
    PFB_TRY(b, hipMalloc((void **)&b->d_cy_hist, n * sizeof(int)));
    b->hist_rows = rows;
    drop_graphs(b);   // graphs captured the old history addresses
    return PF_OK;
}

// Tolerance mode: whole steps are captured only while the cycle cap is small (a captured
// step always launches every cycle; the blocks of converged cases exit at once).
static const int kTolGraphMaxCycles = 16;

static int tol_max_cycles(const pf_batch *b) { return b->kind.max_cycles > 0 ? b->kind.max_cycles : 1000; }

static int tol_inner(const pf_batch *b) {
    return b->cfg.inner_iters > 0 ? b->cfg.inner_iters : (b->kind.iter_max > 0 ? b->kind.iter_max : 1);
}

// One residual evaluation + decision (`cyc`), then -- unless it was the last allowed --
// one cycle of K sweeps (FP32 on the error equation + FP64 update, or FP64 sweeps) for the
// cases still active.
static void enqueue_tol_cycle(pf_batch *b, int cyc) {
    const bool mixed = b->kind.precision == PF_PREC_MIXED;
    const int K = tol_inner(b), maxc = tol_max_cycles(b);
    launch_swp(b, pf::SWP_RESIDUAL, 0, 1);
    launch_swp(b, pf::SWP_RESIDUAL, 1, 1);
    launch_reduce(b, b->d_res_slots, b->d_res_hist, -1, 1);
    const pf::BatchView v = view(b);
    hipLaunchKernelGGL(pf::k_batch_tol_decide, dim3(1), dim3(256), 0, b->st, v, b->d_active, b->d_res_hist,
                       b->d_it_hist, b->d_cy_hist, b->d_step, b->max_cases, b->kind.residual_tol, cyc, maxc, K,
                       b->d_nactive);
    if (cyc == maxc) return;
    group_sweeps(b, mixed ? pf::SWP_FP32 : pf::SWP_FP64, K);
    if (mixed) launch_nat(b, pf::NAT_UPDATE, 0, 0);
}

static void enqueue_tol_begin(pf_batch *b, int p) {
    enqueue_pre(b, p);
    launch_fill_active(b, 1);
}

static void enqueue_tol_end(pf_batch *b, int p) {
    launch_fill_active(b, 1);
    launch_reduce(b, b->d_tile_slots, b->d_err_hist, -1);
    enqueue_post(b, p);
    hipLaunchKernelGGL(pf::k_inc_int, dim3(1), dim3(64), 0, b->st, b->d_step);
}

static void enqueue_step_tol_full(pf_batch *b, int p) {
    enqueue_tol_begin(b, p);
    for (int cyc = 0; cyc <= tol_max_cycles(b); ++cyc) enqueue_tol_cycle(b, cyc);
    enqueue_tol_end(b, p);
}

static bool tol_mode(const pf_batch *b) { return b->kind.residual_tol > 0; }

static int capture(pf_batch *b, int p) {
    PFB_TRY(b, hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
    if (tol_mode(b)) enqueue_step_tol_full(b, p);
    else enqueue_step_fixed(b, p);
    hipGraph_t gph = nullptr;
    PFB_TRY(b, hipStreamEndCapture(b->st, &gph));
    b->graph[p] = gph;
    PFB_TRY(b, hipGraphInstantiate(&b->gexec[p], gph, nullptr, nullptr, 0));
    return PF_OK;
}

// Tolerance mode without a graph: the same launches, with the number of active cases read
// back every few cycles so a step ends once every case has converged.
static int step_tolerance_eager(pf_batch *b, int p) {
    enqueue_tol_begin(b, p);
    const int maxc = tol_max_cycles(b);
    const int check_every = 4;
    for (int cyc = 0; cyc <= maxc; ++cyc) {
        enqueue_tol_cycle(b, cyc);
        if (cyc < maxc && (cyc + 1) % check_every == 0) {
            int nact = 0;
            PFB_TRY_SYNC(b, hipMemcpyAsync(&nact, b->d_nactive, sizeof(int), hipMemcpyDeviceToHost, b->st));
            PFB_TRY(b, hipStreamSynchronize(b->st));
            if (nact == 0) break;
        }
    }
    enqueue_tol_end(b, p);
    return PF_OK;
}

// Tolerance mode with PF_SOLVER_MG, step `s` of the call: the batched stages before and
// after the solve, and in between every case's multigrid solve (pf_mg.h), the cases dealt
// to the batch's stream and its two side streams so that one case's host round trips
// overlap the others' work.  Per-case statistics go to the host-side history.
template <typename T>
static int step_mg_t(pf_batch *b, int p, int s, pf::MgSolver<T> *pf_batch::Slot::*member) {
    enqueue_pre(b, p);
    std::vector<pf::MgSolver<T> *> runs;
    std::vector<int> slots;
    for (int slot : b->order) {
        pf::MgSolver<T> *S = b->slots[(size_t)slot].*member;
        if (!S) continue;
        runs.push_back(S);
        slots.push_back(slot);
    }
    std::vector<hipStream_t> streams = {b->st, b->side[0], b->side[1]};
    PFB_TRY(b, hipEventRecord(b->ev_fork, b->st));
    for (int q = 0; q < 2; ++q) PFB_TRY(b, hipStreamWaitEvent(b->side[q], b->ev_fork, 0));
    pf::MgControl ctl;
    ctl.tol = b->kind.residual_tol;
    ctl.max_cycles = b->kind.max_cycles > 0 ? b->kind.max_cycles : 50;
    ctl.rtol = b->kind.inner_tol > 0 ? b->kind.inner_tol : 1e-3;
    ctl.maxit = b->kind.inner_iters > 0 ? b->kind.inner_iters : 30;
    std::vector<pf::MgStepStats> out;
    PFB_TRY(b, pf::mg_step_solve<T>(runs, streams, ctl, out));
    for (size_t q = 0; q < runs.size(); ++q) {
        const int slot = slots[q];
        pf::launch_reduce_err(streams[q % streams.size()], runs[q]->err_slots, runs[q]->nslots,
                              b->d_err_hist + (int64_t)s * b->max_cases + slot);
        b->mg_res[(size_t)s * b->max_cases + slot] = out[q].residual;
        b->mg_cyc[(size_t)s * b->max_cases + slot] = out[q].cycles;
        b->mg_vc[(size_t)s * b->max_cases + slot] = out[q].vcycles;
    }
    for (int q = 0; q < 2; ++q) {
        PFB_TRY(b, hipEventRecord(b->ev_join[q], b->side[q]));
        PFB_TRY(b, hipStreamWaitEvent(b->st, b->ev_join[q], 0));
    }
    enqueue_post(b, p);
    hipLaunchKernelGGL(pf::k_inc_int, dim3(1), dim3(64), 0, b->st, b->d_step);
    return PF_OK;
}

extern "C" int pf_batch_step(pf_batch *b, int nsteps, pf_stats *stats) {
    if (!b || nsteps < 0) return PF_EINVAL;
    if (nsteps == 0) return PF_OK;
    int rc;
    if (b->dirty && (rc = rebuild_tables(b)) != PF_OK) return rc;
    if (b->order.empty()) return PF_OK;
    if ((rc = ensure_hist(b, nsteps)) != PF_OK) return rc;
    const int nc = (int)b->order.size();
    const bool tol = tol_mode(b);
    const bool mg = mg_mode(b->kind);
    const bool graph = !mg && b->cfg.use_graph && (!tol || tol_max_cycles(b) <= kTolGraphMaxCycles);
    if (mg) {
        const size_t n = (size_t)nsteps * b->max_cases;
        b->mg_res.assign(n, -1.0);
        b->mg_cyc.assign(n, 0);
        b->mg_vc.assign(n, 0);
    }
    hipLaunchKernelGGL(pf::k_set_int, dim3(1), dim3(64), 0, b->st, b->d_step, 0);
    for (int s = 0; s < nsteps; ++s) {
        const int p = b->parity;
        if (mg) {
            rc = (b->kind.precision == PF_PREC_MIXED) ? step_mg_t<float>(b, p, s, &pf_batch::Slot::mg32)
                                                      : step_mg_t<double>(b, p, s, &pf_batch::Slot::mg64);
            if (rc != PF_OK) return rc;
        } else if (graph) {
            if (!b->gexec[p] && (rc = capture(b, p)) != PF_OK) return rc;
            PFB_TRY(b, hipGraphLaunch(b->gexec[p], b->st));
        } else if (tol) {
            if ((rc = step_tolerance_eager(b, p)) != PF_OK) return rc;
        } else {
            enqueue_step_fixed(b, p);
        }
        PFB_TRY(b, hipGetLastError());
        b->parity = 1 - p;
    }
    std::vector<unsigned long long> eh((size_t)nsteps * b->max_cases), rh((size_t)nsteps * b->max_cases);
    std::vector<int> ih, ch;
    PFB_TRY_SYNC(b, hipMemcpyAsync(eh.data(), b->d_err_hist, eh.size() * 8, hipMemcpyDeviceToHost, b->st));
    if (!mg && (tol || b->kind.report_residual))
        PFB_TRY_SYNC(b, hipMemcpyAsync(rh.data(), b->d_res_hist, rh.size() * 8, hipMemcpyDeviceToHost, b->st));
    if (mg) {
        ih.assign(eh.size(), 0);
        ch.assign(eh.size(), 0);
        for (size_t q = 0; q < eh.size(); ++q) {
            std::memcpy(&rh[q], &b->mg_res[q], 8);
            ih[q] = b->mg_vc[q];
            ch[q] = b->mg_cyc[q];
        }
    } else if (tol) {
        ih.resize(eh.size());
        ch.resize(eh.size());
        PFB_TRY_SYNC(b, hipMemcpyAsync(ih.data(), b->d_it_hist, ih.size() * sizeof(int), hipMemcpyDeviceToHost, b->st));
        PFB_TRY_SYNC(b, hipMemcpyAsync(ch.data(), b->d_cy_hist, ch.size() * sizeof(int), hipMemcpyDeviceToHost, b->st));
    }
    PFB_TRY(b, hipStreamSynchronize(b->st));
    bool finite = true;
    for (int ci = 0; ci < nc; ++ci) b->slots[(size_t)b->order[(size_t)ci]].steps += nsteps;
    if (stats) {
        const int K = b->cfg.inner_iters > 0 ? b->cfg.inner_iters : (b->kind.iter_max > 0 ? b->kind.iter_max : 1);
        const int fixed_cycles = (b->kind.precision == PF_PREC_MIXED) ? (b->kind.iter_max + K - 1) / K : 0;
        for (int s = 0; s < nsteps; ++s)
            for (int ci = 0; ci < nc; ++ci) {
                const int slot = b->order[(size_t)ci];
                pf_stats &o = stats[(size_t)s * b->max_cases + slot];
                double e, r;
                std::memcpy(&e, &eh[(size_t)s * b->max_cases + slot], 8);
                std::memcpy(&r, &rh[(size_t)s * b->max_cases + slot], 8);
                if (!std::isfinite(e)) finite = false;
                if ((tol || (b->kind.report_residual && b->kind.precision == PF_PREC_MIXED)) && !std::isfinite(r))
                    finite = false;
                o.sor_error = e;
                if (tol) {
                    o.residual = r;
                    o.sor_iters = ih[(size_t)s * b->max_cases + slot];
                    o.cycles = ch[(size_t)s * b->max_cases + slot];
                } else {
                    o.residual = (b->kind.report_residual && b->kind.precision == PF_PREC_MIXED) ? r : -1.0;
                    o.sor_iters = b->kind.iter_max;
                    o.cycles = fixed_cycles;
                }
            }
    }
    if (!finite) { b->msg = "non-finite SOR error or residual in a case: it diverged"; return PF_EDIVERGED; }
    return PF_OK;
}

extern "C" int pf_batch_download(pf_batch *b, int slot, int field, double *host, int64_t ldi, int64_t ldj) {
    if (!b || !host || slot < 0 || slot >= b->max_cases || !b->slots[(size_t)slot].used) return PF_EINVAL;
    const BatchCase &bc = b->slots[(size_t)slot].bc;
    const PfGeom &g = bc.g;
    if (ldi < g.m + 2 || ldj < g.n + 2) { b->msg = "leading dimensions smaller than m+2, n+2"; return PF_EINVAL; }
    const int q = 1 - b->parity;   // the buffers holding the latest velocity
    const double *src = nullptr;
    bool split = false;
    switch (field) {
        case PF_FIELD_U: src = bc.U[q]; break;
        case PF_FIELD_V: src = bc.V[q]; break;
        case PF_FIELD_W: src = (g.dim == 3) ? bc.W[q] : nullptr; break;
        case PF_FIELD_P: src = bc.P; split = true; break;
        case PF_FIELD_POROSITY: src = bc.phi; break;
        default: break;
    }
    if (!src) { b->msg = "unknown field"; return PF_EINVAL; }
    std::vector<double> stage((size_t)(split ? 2 * g.HS : g.NS));
    PFB_TRY(b, hipStreamSynchronize(b->st));
    PFB_TRY(b, hipMemcpy(stage.data(), src, stage.size() * 8, hipMemcpyDeviceToHost));
    pf::to_host(g, split, stage, host, ldi, ldj);
    return PF_OK;
}

extern "C" int pf_batch_divergence(pf_batch *b, int slot, double *out) {
    if (!b || !out || slot < 0 || slot >= b->max_cases || !b->slots[(size_t)slot].used) return PF_EINVAL;
    const BatchCase &bc = b->slots[(size_t)slot].bc;
    const int q = 1 - b->parity;
    pf::launch_divergence_out(b->st, bc.g, bc.a, bc.phi, bc.U[q], bc.V[q], bc.W[q], bc.dout);
    PFB_TRY(b, hipGetLastError());
    const size_t n = (size_t)bc.g.m * bc.g.n * bc.g.l;
    PFB_TRY_SYNC(b, hipMemcpyAsync(out, bc.dout, n * 8, hipMemcpyDeviceToHost, b->st));
    PFB_TRY(b, hipStreamSynchronize(b->st));
    return PF_OK;
}

extern "C" int pf_batch_set_config(pf_batch *b, const pf_batch_config *cfg) {
    if (!b || !cfg || cfg->persistent_blocks < 0 || cfg->inner_iters < 0) return PF_EINVAL;
    const int tx = cfg->tile_x == 0 ? (b->kind.dim == 3 ? 32 : 64) : cfg->tile_x;
    if (tx != 32 && tx != 64 && tx != 128) { b->msg = "tile_x must be 32, 64 or 128"; return PF_EINVAL; }
    PFB_TRY(b, drain(b));
    const int kb_before = tile_kb(b);
    b->cfg = *cfg;
    b->cfg.tile_x = tx;
    if (tx != b->tile_x || tile_kb(b) != kb_before) {   // tile geometry changed
        b->tile_x = tx;
        b->dirty = true;
    }
    drop_graphs(b);
    return PF_OK;
}

extern "C" int pf_batch_get_config(const pf_batch *b, pf_batch_config *cfg) {
    if (!b || !cfg) return PF_EINVAL;
    *cfg = b->cfg;
    return PF_OK;
}

extern "C" int pf_batch_sync(pf_batch *b) {
    if (!b) return PF_EINVAL;
    PFB_TRY(b, hipStreamSynchronize(b->st));
    return PF_OK;
}

extern "C" int64_t pf_batch_case_steps(const pf_batch *b, int slot) {
    if (!b || slot < 0 || slot >= b->max_cases || !b->slots[(size_t)slot].used) return -1;
    return b->slots[(size_t)slot].steps;
}

extern "C" int pf_device_memory(int64_t *free_bytes, int64_t *total_bytes) {
    size_t f = 0, t = 0;
    if (!free_bytes || !total_bytes) return PF_EINVAL;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return PF_EHIP;
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
    return PF_OK;
}
