"""Porosity generators of the reference's tools/ directory, as library functions and a CLI.

Each returns the interior porosity in this package's numpy layout, shape (n, m) for 2D
(row j, column i, i fastest), and :func:`write` stores it in the reference's CSV schema
(header ``m,n,1`` then ``i, j, 1, value`` rows with ``%.10f`` values, i fastest:
tools/output/output_porosity.py:1-10).  Cell indices in the formulas are the tools'
0-based loop indices.  The map from a signed distance d (cells, positive in the fluid) to
porosity is the tools' ``0.5*tanh(d/thickness)+0.5``.

    python -m pixelflow_amd.generators boundary-layer --out porosity.csv
    python -m pixelflow_amd.generators channel --H 16 --h 128 --out porosity.csv
    python -m pixelflow_amd.generators channel-thin --out porosity.csv
    python -m pixelflow_amd.generators backstep --H 16 --wall 16 --out porosity.csv
    python -m pixelflow_amd.generators cylinder --m 256 --n 128 --out porosity.csv

The STL, bitmap and viewer tools (tools/stl2poro, tools/voxel2poro, tools/viewer) need
libraries this package does not use (vtk, pyvista, OpenCV, matplotlib) and are not
restated; :func:`pixelflow_amd.cases.voxel_porosity` maps a binary voxel field the way
voxel2poro does.
"""
from __future__ import annotations

import argparse
import sys
from typing import Sequence

import numpy as np

from . import cases


def tanh_porosity(dist: np.ndarray, thickness: float) -> np.ndarray:
    """Porosity of a signed distance (cells, positive in the fluid)."""
    return 0.5 * np.tanh(np.asarray(dist, dtype=np.float64) / thickness) + 0.5


def boundary_layer(m: int = 256, n: int = 256, center: float = 128.0, half_width: float = 16.0,
                   thickness: float = 1.5) -> np.ndarray:
    """A solid band |j - center| < half_width across the whole length
    (tools/boundary-layer/boundary-layer.py:4-10: d = |128 - j| - 16)."""
    j = np.arange(n, dtype=np.float64)
    dist = np.abs(center - j) - half_width
    return np.broadcast_to(tanh_porosity(dist, thickness)[:, None], (n, m)).copy()


def _channel_dist(H: int, h: int):
    m, n = H * 8 + h, H * 4 + h
    cen = n / 2
    j = np.arange(n, dtype=np.float64)
    # fluid between the two walls: below the upper face j = cen + h/2 and above the lower
    # face j = cen - h/2 (tools/channel/channel.py:9-25)
    dist = np.minimum(-(j - h / 2 - cen), j + h / 2 - cen)
    return m, n, dist


def channel(H: int = 16, h: int = 128, thickness: float = 1.5) -> np.ndarray:
    """Plane channel of height h between two walls (tools/channel/channel.py:9-27)."""
    m, n, dist = _channel_dist(H, h)
    return np.broadcast_to(tanh_porosity(dist, thickness)[:, None], (n, m)).copy()


def channel_thin(H: int = 16, h: int = 128, thickness: float = 1.5) -> np.ndarray:
    """The same channel with thin walls: porosity tanh(d/thickness)^2, zero on the wall
    faces only (tools/channel/channel_thin.py:27)."""
    m, n, dist = _channel_dist(H, h)
    return np.broadcast_to((np.tanh(dist / thickness) ** 2)[:, None], (n, m)).copy()


def backstep(H: int = 16, wall_thick: int = 16, thickness: float = 1.5) -> np.ndarray:
    """Backward-facing step (tools/backstep/backstep.py:8-45): a solid wall band of
    wall_thick cells centred at j = n/2 - wall_thick/2 .. n/2 + wall_thick/2 (the channel
    runs through the periodic y boundary), raised by a step of height H on its upper face
    for i <= step_length = 10 H, with a rounded convex corner at the step's foot.  Domain
    m = 50 H + wall_thick, n = 5 H + wall_thick."""
    step = H * 10
    m, n = H * 50 + wall_thick, H * 5 + wall_thick
    cen = n / 2
    i = np.arange(m, dtype=np.float64)[None, :]
    j = np.arange(n, dtype=np.float64)[:, None]
    top = j - wall_thick / 2 - cen            # above the band's upper face
    top_step = top - H                        # above the step's upper face
    up = i - step <= 0                        # the step's streamwise extent
    dist = np.broadcast_to(top, (n, m)).copy()
    # upstream of the step's end the upper face is the step's
    dist = np.where(up, np.minimum(dist, top_step), dist)
    dist = np.where(up & (top_step <= 0), np.maximum(dist, top_step), dist)
    dist = np.where(up & (top_step > 0), np.minimum(dist, top_step), dist)
    # below the band's upper face, upstream: at least minus the distance to the corner
    # (i, j) = (step, wall_thick/2 + cen)
    corner_lo = -np.sqrt((i - step) ** 2 + (j - (wall_thick / 2 + cen)) ** 2)
    dist = np.where((top <= 0) & up, np.maximum(dist, corner_lo), dist)
    # between the band's and the step's upper faces: the step's downstream face i = step
    band = (top_step <= 0) & (top >= 0)
    dist = np.where(band & up, np.maximum(dist, i - step), dist)
    dist = np.where(band & ~up, np.minimum(dist, i - step), dist)
    # above the step's upper face: at most the distance to the step's top corner
    corner_hi = np.sqrt((i - step) ** 2 + (j - (H + wall_thick / 2 + cen)) ** 2)
    dist = np.where(top_step >= 0, np.minimum(dist, corner_hi), dist)
    # the band's lower face
    dist = np.maximum(dist, -(j + wall_thick / 2 - cen))
    return np.minimum(1.0, tanh_porosity(dist, thickness))


def cylinder(m: int = 256, n: int = 128, width: float = 2.0, height: float = 1.0,
             center: Sequence[float] = (0.5, 0.5), diameter: float = 0.1) -> np.ndarray:
    """Circular cylinder with a 2-cell tanh interface (tools/cylinder/
    make_porosity_2d_cylinder.py:119-122), as the benchmark cases use it."""
    return cases.cylinder_2d(m, n, width, height, tuple(center), diameter)


def write(path: str, phi: np.ndarray) -> None:
    """The reference's porosity CSV (tools/output/output_porosity.py:1-25)."""
    cases.write_porosity_csv(path, phi)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="porosity generators of PixelFlow's tools/, CSV output")
    ap.add_argument("kind", choices=["boundary-layer", "channel", "channel-thin", "backstep", "cylinder"])
    ap.add_argument("--out", default="porosity.csv")
    ap.add_argument("--thickness", type=float, default=1.5)
    ap.add_argument("--H", type=int, default=16)
    ap.add_argument("--h", type=int, default=128)
    ap.add_argument("--wall", type=int, default=16)
    ap.add_argument("--m", type=int, default=256)
    ap.add_argument("--n", type=int, default=256)
    args = ap.parse_args(argv)
    if args.kind == "boundary-layer":
        phi = boundary_layer(args.m, args.n, thickness=args.thickness)
    elif args.kind == "channel":
        phi = channel(args.H, args.h, args.thickness)
    elif args.kind == "channel-thin":
        phi = channel_thin(args.H, args.h, args.thickness)
    elif args.kind == "backstep":
        phi = backstep(args.H, args.wall, args.thickness)
    else:
        phi = cylinder(args.m, args.n if args.n != 256 else 128)
    write(args.out, phi)
    n, m = phi.shape
    print(f"wrote {args.out}: {m} x {n}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
