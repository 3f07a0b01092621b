"""Case inputs in PixelFlow's own formats: controlDict namelists, porosity CSV, generators.

The solver reads two files from the case directory (the executable's CWD):

* ``config/controlDict.txt`` -- seven Fortran namelist groups, read in the order
  physical, file_control, grid_control, porosity_control, calculation_method,
  directory_control, solver_control (reference ``src/omp_parallel/lib/global.f90:47-62``);
* the porosity CSV named by ``csv_file`` -- a header ``m,n,l`` then one row
  ``i, j, k, value`` per interior cell, i fastest (writer
  ``tools/output/output_porosity.py:1-25``, reader ``src/omp_parallel/lib/grid.f90:283-292``).

The generators below build the porosity fields of the benchmark workloads with a seeded
``numpy.random.default_rng``.  Values are rounded exactly as the CSV's ``%.10f`` text
rounds them (:func:`csv_round`), so an in-memory run and a run that goes through the CSV
see bit-identical porosity.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# --------------------------------------------------------------------------------------
# Parameters (namelist variables of global.f90:17-26)
# --------------------------------------------------------------------------------------


@dataclasses.dataclass
class CaseParams:
    """One case: the controlDict variables plus extents and boundary kind.

    ``time`` and ``istep_max`` give ``dt = time / istep_max`` exactly as
    ``grid.f90:52`` derives it; ``dx = width / (m - 1)`` (``grid.f90:50``).
    """

    dim: int
    m: int
    n: int
    l: int = 1
    bc: str = "uniform"  # "uniform" (x inlet/outlet, periodic y[,z]) or "air"
    xnue: float = 0.001
    xlambda: float = 0.0
    density: float = 1.0
    width: float = 1.0
    height: float = 1.0
    depth: float = 1.0
    time: float = 1.0
    inlet_velocity: float = 1.0
    outlet_pressure: float = 0.0
    aoa: float = 0.0
    istep_out: int = 1000000
    istep_max: int = 100
    thickness: float = 1.5
    threshold: float = 1.0e-6
    radius: float = 0.1
    center_x: float = 0.5
    center_y: float = 0.5
    center_z: float = 0.5
    nonslip: bool = True
    output_folder: str = "output"
    csv_file: str = "data/porosity.csv"
    iter_max: int = 100
    relux_factor: float = 1.7
    # air-conditioning face codes (top, bottom, east, west, north, south):
    # 0 wall, 1 inlet where porosity >= 0.9, 2 outlet where porosity >= 0.9.
    # Default = the reference's compile-time choice (ibm_3d_air_condition_omp_cpu.f90:10-15).
    faces: Tuple[int, int, int, int, int, int] = (1, 0, 0, 0, 0, 2)
    # 2D backward-facing-step executable: inlet and initial velocity times porosity
    # (ibm_2d_backstep_omp_cpu.f90:533-534, 615-618)
    inlet_porosity: bool = False

    @property
    def dx(self) -> float:
        return self.width / float(self.m - 1)

    @property
    def dy(self) -> float:
        return self.height / float(self.n - 1)

    @property
    def dz(self) -> float:
        return self.depth / float(self.l - 1) if self.dim == 3 else 0.0

    @property
    def dt(self) -> float:
        return self.time / float(self.istep_max)

    @property
    def cells(self) -> int:
        return self.m * self.n * (self.l if self.dim == 3 else 1)

    def full_shape(self) -> Tuple[int, ...]:
        """numpy (C-order) shape of a field with ghosts: (l+2, n+2, m+2) or (n+2, m+2)."""
        if self.dim == 3:
            return (self.l + 2, self.n + 2, self.m + 2)
        return (self.n + 2, self.m + 2)

    def interior_shape(self) -> Tuple[int, ...]:
        if self.dim == 3:
            return (self.l, self.n, self.m)
        return (self.n, self.m)


# --------------------------------------------------------------------------------------
# controlDict writer / reader
# --------------------------------------------------------------------------------------


def _fnum(x: float) -> str:
    """Shortest decimal that reads back to the same double (Fortran list input is
    correctly rounded, so repr() round-trips)."""
    r = repr(float(x))
    return r.replace("e", "E") if "e" in r else r


def write_control_dict(path: str, p: CaseParams) -> None:
    """Write ``config/controlDict.txt`` with the seven namelist groups of global.f90:47-54."""
    logical = ".true." if p.nonslip else ".false."
    text = f"""!********************************************
&physical
xnue = {_fnum(p.xnue)}
xlambda = {_fnum(p.xlambda)}
density = {_fnum(p.density)}
width   = {_fnum(p.width)}
height  = {_fnum(p.height)}
depth   = {_fnum(p.depth)}
time = {_fnum(p.time)}
inlet_velocity = {_fnum(p.inlet_velocity)}
outlet_pressure = {_fnum(p.outlet_pressure)}
AoA = {_fnum(p.aoa)}
/
&file_control
istep_out = {int(p.istep_out)}
/
&grid_control
istep_max = {int(p.istep_max)}
/
&porosity_control
thickness = {_fnum(p.thickness)}
threshold = {_fnum(p.threshold)}
radius = {_fnum(p.radius)}
center_x = {_fnum(p.center_x)}
center_y = {_fnum(p.center_y)}
center_z = {_fnum(p.center_z)}
/
&calculation_method
nonslip = {logical}
/
&directory_control
output_folder   = "{p.output_folder}"
csv_file        = "{p.csv_file}"
/
&solver_control
iter_max        = {int(p.iter_max)}
relux_factor    = {_fnum(p.relux_factor)}
/
!********************************************
"""
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "w") as f:
        f.write(text)


# --------------------------------------------------------------------------------------
# Porosity CSV
# --------------------------------------------------------------------------------------


def csv_round(x: np.ndarray) -> np.ndarray:
    """Exactly the double that ``float('%.10f' % v)`` yields, vectorised.

    ``s = rint(v * 1e10)`` is the decimal digit string scaled to an integer; dividing two
    exactly representable doubles is correctly rounded, which is what decimal parsing of
    ``s * 10**-10`` gives.  Values whose scaled fraction sits near a rounding tie are
    redone through Python's exact formatter.
    """
    v = np.asarray(x, dtype=np.float64)
    scaled = v * 1e10
    s = np.rint(scaled)
    out = s / 1e10
    frac = np.abs(scaled - np.floor(scaled) - 0.5)
    risky = np.nonzero((frac < 1e-4).ravel())[0]
    if risky.size:
        flat = out.reshape(-1)
        src = v.reshape(-1)
        for idx in risky:
            flat[idx] = float("%.10f" % src[idx])
    out[out == 0.0] = 0.0  # "-0.0000000000" reads back as -0.0; keep +0.0 like the reference generator's inputs
    return out


def write_porosity_csv(path: str, phi_interior: np.ndarray) -> None:
    """Write the CSV schema of tools/output/output_porosity.py:1-25.

    ``phi_interior`` has numpy shape (l, n, m) or (n, m); rows are written i fastest.
    """
    phi = np.asarray(phi_interior, dtype=np.float64)
    if phi.ndim == 2:
        phi = phi[None]
    l, n, m = phi.shape
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    kk, jj, ii = np.meshgrid(np.arange(1, l + 1), np.arange(1, n + 1), np.arange(1, m + 1), indexing="ij")
    with open(path, "w") as f:
        f.write(f"{m},{n},{l}\n")
        cols = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1)
        vals = phi.ravel()
        chunk = 1 << 18
        for s in range(0, vals.size, chunk):
            lines = [f"{a}, {b}, {c}, {v:.10f}\n" for (a, b, c), v in zip(cols[s:s + chunk].tolist(), vals[s:s + chunk].tolist())]
            f.write("".join(lines))


def read_porosity_csv(path: str) -> np.ndarray:
    """Read a porosity CSV (any whitespace/comma layout list-directed input accepts).

    Returns the raw values in numpy shape (l, n, m) for 3D headers with l > 1, else (n, m);
    rows are placed by their (i, j, k) columns, as grid.f90:288-289 does."""
    with open(path) as f:
        head = f.readline().replace(",", " ").split()
        m, n = int(head[0]), int(head[1])
        l = int(head[2]) if len(head) > 2 else 1
        data = np.loadtxt(f, delimiter=",", dtype=np.float64, ndmin=2)
    out = np.zeros((l, n, m), dtype=np.float64)
    i = data[:, 0].astype(np.int64) - 1
    j = data[:, 1].astype(np.int64) - 1
    k = data[:, 2].astype(np.int64) - 1
    out[k, j, i] = data[:, 3]
    return out if l > 1 else out[0]


# --------------------------------------------------------------------------------------
# Generators for the benchmark workloads
# --------------------------------------------------------------------------------------


def tanh_profile(d_cells: np.ndarray, thickness_cells: float) -> np.ndarray:
    """``0.5*tanh(d/thickness)+0.5`` -- the SDF-to-porosity map of
    tools/cylinder/make_porosity_2d_cylinder.py:119-122 (d > 0 in the fluid)."""
    return 0.5 * np.tanh(d_cells / thickness_cells) + 0.5


def voxel_porosity(binary: np.ndarray, thickness_cells: float = 1.5) -> np.ndarray:
    """Binary voxels (1 fluid, 0 solid) -> porosity with a tanh interface, the conversion
    the reference applies to voxel geometry (tools/voxel2poro/voxel2poro.py:189-197 filters
    with a tanh kernel).  Here: signed distance to the interface in cells (half a cell
    outside the last voxel), mapped by ``0.5*tanh(d/thickness)+0.5``."""
    from scipy.ndimage import distance_transform_edt
    b = np.asarray(binary) > 0.5
    if b.all():
        return np.ones(b.shape, dtype=np.float64)
    if not b.any():
        return csv_round(tanh_profile(np.full(b.shape, -1e9), thickness_cells))
    d = np.where(b, distance_transform_edt(b) - 0.5, -(distance_transform_edt(~b) - 0.5))
    return csv_round(tanh_profile(d, thickness_cells))


def cylinder_2d(m: int, n: int, width: float, height: float, center: Tuple[float, float],
                diameter: float, interface_cells: float = 2.0) -> np.ndarray:
    """Circle of the given diameter (solid, porosity -> 0) in a fluid (1); numpy (n, m)."""
    dx = width / (m - 1)
    x = np.arange(m) * dx
    y = np.arange(n) * (height / (n - 1))
    X, Y = np.meshgrid(x, y, indexing="xy")
    d = (np.hypot(X - center[0], Y - center[1]) - 0.5 * diameter) / dx
    return csv_round(tanh_profile(d, interface_cells))


def backstep_2d(m: int, n: int, step_len: int, step_height: int, wall: int = 4,
                interface_cells: float = 2.0) -> np.ndarray:
    """Backward-facing step in a channel, the layout of the reference's test/backstep.zip
    porosity: a solid band ``wall`` cells thick across the middle of y (with the periodic
    y wrap it is the channel's top and bottom wall) that is ``step_height`` cells thicker
    over the first ``step_len`` cells in x; tanh interface; numpy (n, m)."""
    I, J = np.meshgrid(np.arange(m, dtype=np.float64), np.arange(n, dtype=np.float64), indexing="xy")
    y0 = n // 2 - wall // 2 - 0.5            # band rows [y0 + 0.5, y1 - 0.5]
    y1 = y0 + wall
    d_band = np.maximum(y0 - J, J - y1)      # signed distance, positive in the fluid
    dx = I - (step_len - 0.5)
    dy = np.maximum(y0 - J, J - (y1 + step_height))
    d_block = np.hypot(np.maximum(dx, 0.0), np.maximum(dy, 0.0)) + np.minimum(np.maximum(dx, dy), 0.0)
    return csv_round(tanh_profile(np.minimum(d_band, d_block), interface_cells))


def sphere_3d(m: int, n: int, l: int, width: float, height: float, depth: float,
              center: Tuple[float, float, float], diameter: float, interface_cells: float = 2.0) -> np.ndarray:
    """Sphere (solid) in a fluid box; numpy shape (l, n, m)."""
    dx = width / (m - 1)
    x = np.arange(m) * dx
    y = np.arange(n) * (height / (n - 1))
    z = np.arange(l) * (depth / (l - 1))
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    d = (np.sqrt((X - center[0]) ** 2 + (Y - center[1]) ** 2 + (Z - center[2]) ** 2) - 0.5 * diameter) / dx
    return csv_round(tanh_profile(d, interface_cells))


def random_boxes_3d(m: int, n: int, l: int, nboxes: int, side_range: Tuple[int, int], seed: int,
                    clear_faces: bool = False) -> np.ndarray:
    """Binary porosity (1 fluid, 0 solid) with ``nboxes`` axis-aligned boxes, sides uniform
    in ``side_range`` cells, placed uniformly and clipped to the domain; numpy (l, n, m).

    ``clear_faces`` places every box whole inside the region that leaves 16 cells after the
    inlet, the last 35 % of x before the outlet and 4 cells at each periodic y / z face
    free of solid: the reference's explicit step diverges at once where a box meets the
    inlet, and the interface of a box clipped at a periodic face is a jump across the
    wrap."""
    rng = np.random.default_rng(seed)
    phi = np.ones((l, n, m), dtype=np.float64)
    for _ in range(nboxes):
        sx, sy, sz = (int(v) for v in rng.integers(side_range[0], side_range[1] + 1, size=3))
        if clear_faces:
            x_hi, margin = int(0.65 * m), 4
            x0 = int(rng.integers(16, max(17, x_hi - sx + 1)))
            y0 = int(rng.integers(margin, max(margin + 1, n - margin - sy + 1)))
            z0 = int(rng.integers(margin, max(margin + 1, l - margin - sz + 1)))
        else:
            x0 = int(rng.integers(0, m))
            y0 = int(rng.integers(0, n))
            z0 = int(rng.integers(0, l))
        phi[z0:z0 + sz, y0:y0 + sy, x0:x0 + sx] = 0.0
    return phi


def random_circles_2d(m: int, n: int, count: int, radius_range: Tuple[float, float], seed: int,
                      interface_cells: float = 2.0, max_tries: int = 2000, downstream_clear: float = 0.0,
                      max_blockage: float = 1.0, gap: float = 2.0) -> np.ndarray:
    """1..8 style obstacle field: ``count`` non-overlapping circles (radii in cells), tanh
    interface; numpy (n, m).  Circles keep a ``gap``-cell gap and stay clear of the x faces;
    ``downstream_clear`` (a fraction of m) keeps that much of the domain before the outlet
    free of solid, and ``max_blockage`` bounds the fraction of any grid column inside
    circles (the reference's outlet condition and central scheme do not survive a wake
    that reaches the outlet, or a nearly closed channel)."""
    rng = np.random.default_rng(seed)
    circles: List[Tuple[float, float, float]] = []
    blocked = np.zeros(m)
    xs = np.arange(m, dtype=np.float64)
    x_hi = m - max(4.0, downstream_clear * m)
    tries = 0
    while len(circles) < count and tries < max_tries:
        tries += 1
        r = float(rng.uniform(*radius_range))
        r = min(r, 0.45 * min(m, n) - 2.0)
        if r < 1.0:
            break
        cx = float(rng.uniform(r + 4.0, x_hi - r)) if x_hi - 2 * r - 4.0 > 0 else min(m / 2.0, x_hi - r)
        cy = float(rng.uniform(r + 2.0, n - r - 2.0)) if n - 2 * r - 4.0 > 0 else n / 2.0
        if not all(math.hypot(cx - a, cy - b) > r + c + gap for a, b, c in circles):
            continue
        chord = 2.0 * np.sqrt(np.maximum(r * r - (xs - cx) ** 2, 0.0))
        if max_blockage < 1.0 and np.max(blocked + chord) > max_blockage * n:
            continue
        blocked += chord
        circles.append((cx, cy, r))
    I, J = np.meshgrid(np.arange(m, dtype=np.float64), np.arange(n, dtype=np.float64), indexing="xy")
    d = np.full((n, m), np.inf)
    for cx, cy, r in circles:
        d = np.minimum(d, np.hypot(I - cx, J - cy) - r)
    return csv_round(tanh_profile(d, interface_cells))


def room_3d(m: int, n: int, l: int, seed: int, wall_cells: int = 3,
            interface_cells: float = 1.0) -> Tuple[np.ndarray, float]:
    """Closed room for the air-conditioning solver; returns (porosity (l, n, m), inlet speed).

    Walls ``wall_cells`` thick close all six faces; 1-4 openings of 4-16 cells are cut
    through the top wall (the reference's inlet face) and the south wall (its outlet
    face), at least one of each; 0-10 furniture boxes of 4-24 cells stand inside.  The
    voxels become porosity through :func:`voxel_porosity` with a 1-cell interface, which
    keeps the middle of a 4-cell opening above the 0.9 that gates inlet/outlet cells.
    The inlet speed is uniform in [0.5, 2] m/s."""
    rng = np.random.default_rng(seed)
    t = wall_cells
    b = np.zeros((l, n, m), dtype=np.float64)
    b[t:l - t, t:n - t, t:m - t] = 1.0
    for _ in range(int(rng.integers(0, 11))):
        sx, sy, sz = (int(v) for v in rng.integers(4, 25, size=3))
        x0 = int(rng.integers(t, max(t + 1, m - t - sx)))
        y0 = int(rng.integers(t, max(t + 1, n - t - sy)))
        z0 = t
        b[z0:min(z0 + sz, l - t - 2), y0:min(y0 + sy, n - t), x0:min(x0 + sx, m - t)] = 0.0
    nopen = int(rng.integers(1, 5))
    kinds = ["top", "south"] + ["top" if rng.random() < 0.5 else "south" for _ in range(max(0, nopen - 2))]
    for kind in kinds:
        s1 = int(rng.integers(4, 17))
        s2 = int(rng.integers(4, 17))
        x0 = int(rng.integers(t, max(t + 1, m - t - s1)))
        if kind == "top":
            y0 = int(rng.integers(t, max(t + 1, n - t - s2)))
            b[l - t:, y0:y0 + s2, x0:x0 + s1] = 1.0
        else:
            z0 = int(rng.integers(t, max(t + 1, l - t - s2)))
            b[z0:z0 + s2, :t, x0:x0 + s1] = 1.0
    return voxel_porosity(b, interface_cells), float(rng.uniform(0.5, 2.0))


# --------------------------------------------------------------------------------------
# The benchmark workloads
# --------------------------------------------------------------------------------------


def cylinder_case(m: int = 256, n: int = 128, steps: int = 2000, iter_max: int = 100) -> Tuple[CaseParams, np.ndarray]:
    """2D flow past a cylinder: 256x128 over 2.0 m x 1.0 m, D = 0.1 m at (0.5, 0.5), 2-cell
    tanh interface, U = 1, xnue = 1e-3, dt = 2e-4, SOR 100 iterations, omega = 1.7."""
    p = CaseParams(dim=2, m=m, n=n, width=2.0, height=1.0, depth=0.0, xnue=0.001, time=2e-4 * steps,
                   istep_max=steps, inlet_velocity=1.0, iter_max=iter_max, relux_factor=1.7, radius=0.05)
    return p, cylinder_2d(m, n, 2.0, 1.0, (0.5, 0.5), 0.1)


def sphere_case(m: int = 256, n: int = 128, l: int = 128, steps: int = 500, iter_max: int = 100) -> Tuple[CaseParams, np.ndarray]:
    """3D flow around a sphere: 256x128x128 over 2x1x1 m, D = 0.15 m at (0.5, 0.5, 0.5),
    2-cell tanh interface, U = 1, xnue = 1e-3, dt = 1e-4, SOR 100 iterations, omega = 1.7."""
    p = CaseParams(dim=3, m=m, n=n, l=l, width=2.0, height=1.0, depth=1.0, xnue=0.001, time=1e-4 * steps,
                   istep_max=steps, inlet_velocity=1.0, iter_max=iter_max, relux_factor=1.7, radius=0.075)
    return p, sphere_3d(m, n, l, 2.0, 1.0, 1.0, (0.5, 0.5, 0.5), 0.15)


def boxes_case(m: int = 512, n: int = 256, l: int = 256, steps: int = 200, seed: int = 7,
               iter_max: int = 100, binary: bool = False, nboxes: int = 40,
               side_range: Tuple[int, int] = (4, 32), cell_re: Optional[float] = None,
               clear_faces: bool = False) -> Tuple[CaseParams, np.ndarray]:
    """3D mixed-precision case: 512x256x256 over 2x1x1 m, 40 seeded boxes with sides 4-32
    cells, U = 1, xnue = 5e-4, dt = 5e-5.  ``binary=True`` hands the 0/1 voxels to the
    solver as they are (clamped to the threshold, the step diverges within a few steps in
    the reference too: the porosity-gradient terms are divided by porosity); the default
    converts the voxels with the reference's tanh interface (:func:`voxel_porosity`).

    At dx = 2/511 m, xnue = 5e-4 is a cell Reynolds number U dx / xnue of 7.8, where the
    reference's explicit central step diverges on the box interfaces within a dozen steps
    at full size.  ``cell_re`` keeps grid and boxes and sets xnue = U dx / cell_re with dt
    from :func:`stable_dt` instead; ``clear_faces`` keeps the boxes off the inlet, the
    outlet region and the periodic faces (:func:`random_boxes_3d`)."""
    xnue, dt = 0.0005, 5e-5
    if cell_re is not None:
        dx = 2.0 / (m - 1)
        xnue = dx / cell_re
        dt = stable_dt(dx, xnue, 1.0, dim=3)
    p = CaseParams(dim=3, m=m, n=n, l=l, width=2.0, height=1.0, depth=1.0, xnue=xnue, time=dt * steps,
                   istep_max=steps, inlet_velocity=1.0, iter_max=iter_max, relux_factor=1.7)
    vox = random_boxes_3d(m, n, l, nboxes, side_range, seed, clear_faces=clear_faces)
    return p, (vox if binary else voxel_porosity(vox, 1.5))


BATCH_SIZES_2D: Sequence[Tuple[int, int]] = ((64, 64), (128, 64), (256, 128), (384, 192), (512, 256))


def stable_dt(dx: float, xnue: float, velocity: float = 1.0, dim: int = 2, cfl: float = 0.2) -> float:
    """Largest dt the reference's explicit step tolerates with margin: its convection is
    2nd-order central and its update forward Euler (ibm_2d_uniform_omp_cpu.f90:202-208), so
    besides CFL = U*dt/dx <= cfl it needs U^2*dt/xnue <= 1 (central advection under forward
    Euler) and a diffusion number xnue*dt/dx^2 <= 0.2/dim (at 0.4/dim, cases with obstacles
    near each other already diverge)."""
    return min(cfl * dx / velocity, xnue / velocity ** 2, 0.2 / dim * dx * dx / xnue)


BATCH_RE_LENGTH_CELLS = 400.0


def batch_cases_2d(count: int = 512, steps: int = 1000, seed: int = 11, iter_max: int = 60,
                   cfl: float = 0.2) -> List[Tuple[CaseParams, np.ndarray]]:
    """Independent 2D cases: sizes uniform from BATCH_SIZES_2D, 1-8 circles with radii 3-20
    cells (in the upstream 65% of the domain, 4-cell gaps, at most 35% of any column
    blocked), Re uniform in 50-400 via xnue on a length of BATCH_RE_LENGTH_CELLS cells
    (xnue = U * 400 dx / Re, a cell Reynolds number U dx / xnue of 0.125-1), dt from
    :func:`stable_dt` (CFL 0.2 capped by the central-scheme limits), SOR 60 iterations,
    omega = 1.7.

    The reference's explicit central scheme is stable on porosity interfaces only at cell
    Reynolds numbers of about 1 (its cylinder-2d test case runs at 1); with Re on the
    obstacle diameter (3-23 cells at 50-400) the cases diverge within about 100 steps."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(count):
        m, n = BATCH_SIZES_2D[int(rng.integers(0, len(BATCH_SIZES_2D)))]
        width = 2.0
        dx = width / (m - 1)
        height = dx * (n - 1)
        nobs = int(rng.integers(1, 9))
        phi = random_circles_2d(m, n, nobs, (3.0, 20.0), seed=int(rng.integers(0, 2**31 - 1)),
                                downstream_clear=0.35, max_blockage=0.35, gap=4.0)
        re = float(rng.uniform(50.0, 400.0))
        xnue = 1.0 * BATCH_RE_LENGTH_CELLS * dx / re
        dt = stable_dt(dx, xnue, 1.0, 2, cfl)
        p = CaseParams(dim=2, m=m, n=n, width=width, height=height, depth=0.0, xnue=xnue, time=dt * steps,
                       istep_max=steps, inlet_velocity=1.0, iter_max=iter_max, relux_factor=1.7)
        out.append((p, phi))
    return out


ROOM_SIZES_3D: Sequence[Tuple[int, int, int]] = ((64, 64, 64), (96, 96, 96), (128, 128, 128), (160, 128, 96), (192, 192, 192))


ROOM_VISCOSITY_SCALE = 1666.6666666666667  # 1.5e-5 * scale = 0.025, the reference room's xnue


def room_case(seed: int, steps: int = 300, iter_max: int = 100, size: Optional[Tuple[int, int, int]] = None) -> Tuple[CaseParams, np.ndarray]:
    """Served room-ventilation request: grid from ROOM_SIZES_3D, closed room with seeded
    openings and furniture (:func:`room_3d`), inlet speed U[0.5, 2], 1 cm cells,
    ``xnue = 1.5e-5 * scale`` with the scale that gives the reference room project's
    viscosity (0.025, test/room.zip controlDict), ``dt`` = the smaller of CFL 0.2 and a
    diffusion number of 0.1 (explicit central differences need both)."""
    rng = np.random.default_rng(seed)
    if size is None:
        size = ROOM_SIZES_3D[int(rng.integers(0, len(ROOM_SIZES_3D)))]
    m, n, l = size
    phi, uin = room_3d(m, n, l, seed=int(rng.integers(0, 2**31 - 1)))
    dx = 0.01
    xnue = 1.5e-5 * ROOM_VISCOSITY_SCALE
    dt = min(0.2 * dx / uin, 0.1 * dx * dx / xnue)
    p = CaseParams(dim=3, m=m, n=n, l=l, bc="air", width=dx * (m - 1), height=dx * (n - 1), depth=dx * (l - 1),
                   xnue=xnue, time=dt * steps, istep_max=steps, inlet_velocity=uin,
                   iter_max=iter_max, relux_factor=1.7)
    return p, phi


def write_case_dir(root: str, p: CaseParams, phi_interior: np.ndarray) -> None:
    """A complete case directory: config/controlDict.txt + the porosity CSV."""
    write_control_dict(os.path.join(root, "config", "controlDict.txt"), p)
    write_porosity_csv(os.path.join(root, p.csv_file), phi_interior)


def summary(p: CaseParams) -> Dict[str, object]:
    d = dataclasses.asdict(p)
    d.update(dx=p.dx, dy=p.dy, dz=p.dz, dt=p.dt, cells=p.cells)
    return d
