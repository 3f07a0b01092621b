"""Pins the CPU restatement (oracle/pf_oracle.c) before anything is checked against it.

* every fixture written by the reference build (tests/golden/gen_golden.py) is reproduced
  bit for bit: final u, v, w, p, the divergence diagnostic and the per-step SOR error;
* the colour decode of the reference's SOR is brute-forced: each interior cell is visited
  exactly once, and the first sweep is the parity class (i+j+k even in 3D, i+j odd in 2D);
* closed forms: uniform flow through porosity 1 is an exact fixed point of the step.
"""
import hashlib

import numpy as np
import pytest

from conftest import golden_names, load_golden

pyoracle = pytest.importorskip("pyoracle")


def _run(p, g):
    o = pyoracle.OracleRun(p, g["phi_raw"])
    errs = o.step(p.istep_max)
    return o, errs


# the full-length benchmark fixtures store no porosity (their generators rebuild it) and
# are checked below (2D) or on the GPU only (3D: 500 steps is half an hour on the CPU)
LONG = ("g2d_c1_2000", "g3d_c2_500")
FULL = [n for n in golden_names() if not n.startswith("g3d_ref_") and n not in LONG]


@pytest.mark.parametrize("name", FULL)
def test_oracle_matches_reference_bitwise(name):
    p, g = load_golden(name)
    o, errs = _run(p, g)
    for f in ("u", "v", "w", "p"):
        if f in g:
            np.testing.assert_array_equal(o.interior(f), g[f], err_msg=f"{name}:{f}")
    np.testing.assert_array_equal(errs, g["errs"], err_msg=f"{name}:errs")
    np.testing.assert_array_equal(o.divergence(), g["div"], err_msg=f"{name}:div")


def test_oracle_matches_reference_cylinder_2000_steps():
    """The 2D benchmark case over its full 2000 steps (256x128), bit for bit."""
    from pixelflow_amd import cases
    p, g = load_golden("g2d_c1_2000")
    _, phi = cases.cylinder_case(m=p.m, n=p.n, steps=p.istep_max)
    o = pyoracle.OracleRun(p, phi)
    errs = o.step(p.istep_max)
    for f in ("u", "v", "p"):
        np.testing.assert_array_equal(o.interior(f), g[f], err_msg=f)
    np.testing.assert_array_equal(errs, g["errs"])


def test_oracle_matches_reference_room_project():
    """The reference's own room project (test/room.zip, 64^3), compared by SHA-256."""
    p, g = load_golden("g3d_ref_room64")
    o, errs = _run(p, g)
    np.testing.assert_array_equal(errs, g["errs"])
    for f in ("u", "v", "w", "p"):
        a = o.interior(f)
        np.testing.assert_array_equal(a[31], g[f"{f}_plane031"])
        assert hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest() == str(g["sha_" + f]), f


@pytest.mark.parametrize("dim", [2, 3])
def test_colour_decode_is_parity_class(dim):
    for m in range(2, 10):
        for n in range(2, 10):
            for l in (range(2, 8) if dim == 3 else [1]):
                rc, v = pyoracle.decode_first_sweep(dim, m, n, l)
                assert rc == 0
                assert set(np.unique(v)) <= {1, 16}, (m, n, l)
                if dim == 3:
                    k, j, i = np.meshgrid(np.arange(1, l + 1), np.arange(1, n + 1), np.arange(1, m + 1), indexing="ij")
                    first = (i + j + k) % 2 == 0
                else:
                    j, i = np.meshgrid(np.arange(1, n + 1), np.arange(1, m + 1), indexing="ij")
                    first = (i + j) % 2 == 1
                np.testing.assert_array_equal(v == 1, first, err_msg=str((m, n, l)))


@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_flow_is_fixed_point(dim):
    from pixelflow_amd import cases
    if dim == 2:
        p = cases.CaseParams(dim=2, m=20, n=12, width=1.0, height=0.6, time=0.01, istep_max=5, iter_max=7,
                             inlet_velocity=0.7, outlet_pressure=0.0)
        phi = np.ones((12, 20))
    else:
        p = cases.CaseParams(dim=3, m=12, n=8, l=6, width=1.0, height=0.6, depth=0.5, time=0.01, istep_max=5,
                             iter_max=7, inlet_velocity=0.7, outlet_pressure=0.0)
        phi = np.ones((6, 8, 12))
    o = pyoracle.OracleRun(p, phi)
    errs = o.step(5)
    assert np.all(o.interior("u") == 0.7)
    assert np.all(o.interior("v") == 0.0)
    assert np.all(o.interior("p") == 0.0)
    assert np.all(errs == 0.0)
    assert np.all(o.divergence() == 0.0)


def test_fp32_build_is_close_but_not_equal():
    """The shipped FP32 build differs from FP64 at the 1e-7..1e-5 level (reduced precision
    alone cannot meet a 1e-6 field target; defect correction is needed)."""
    p, g = load_golden("g2d_cyl48")
    for f in ("u", "v", "p"):
        a64, a32 = g[f], g[f + "_fp32"].astype(np.float64)
        rel = np.linalg.norm(a32 - a64) / np.linalg.norm(a64)
        assert 1e-9 < rel < 1e-3, (f, rel)


def test_oracle_force_matches_reference_drag_log():
    """output_force_log_2d of the drag executable (run on one thread, so its reduction is
    the serial loop order): every step's pressure and viscous force, bit for bit."""
    p, g = load_golden("g2d_drag")
    o = pyoracle.OracleRun(p, g["phi_raw"])
    for s in range(p.istep_max):
        o.step(1)
        f = o.force()
        np.testing.assert_array_equal(f[[0, 1]], g["force_p"][s], err_msg=f"step {s + 1} Fp")
        np.testing.assert_array_equal(f[[3, 4]], g["force_v"][s], err_msg=f"step {s + 1} Fv")
        assert f[2] == 0.0 and f[5] == 0.0


def test_force_closed_forms():
    """Pressure force of a constant pressure over a mirror-symmetric body is zero in x
    (equal and opposite faces); the viscous force is linear in velocity; without an
    interface (porosity 1) both vanish."""
    from pixelflow_amd import cases
    p = cases.CaseParams(dim=2, m=40, n=20, width=1.0, height=0.487, time=1e-3, istep_max=1, iter_max=5)
    phi = cases.cylinder_2d(40, 20, 1.0, 0.487, (0.5, 0.2435), 0.2)
    o = pyoracle.OracleRun(p, phi)
    o.p_[...] = 2.0
    o.u[...] = 1.0
    o.v[...] = 0.0
    f1 = o.force()
    assert abs(f1[0]) < 1e-12 * np.abs(f1).max() + 1e-300
    o.u[...] = 3.0
    f3 = o.force()
    np.testing.assert_allclose(f3[3], 3.0 * f1[3], rtol=1e-14)
    flat = pyoracle.OracleRun(p, np.ones((20, 40)))
    flat.p_[...] = 5.0
    assert np.all(flat.force() == 0.0)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_force_3d_plane_wall_closed_forms(axis):
    """The 3D branch (output.f90:1090-1165) on a plane interface normal to one axis,
    porosity 0.5*tanh(s/T)+0.5 with T = thickness cells.  Since dphi/ds = (2/T) phi(1-phi),
    the pressure term sums to -p times the wall's area (the integral of dphi/ds is the jump,
    1), and the wall term to alpha*rho*nu*u_b * A * d_axis / (12 T d_b^2) (the integral of
    (dphi/ds)^2 is 1/(3T)).  The cell spacings differ in x, y and z, so a swapped spacing or
    a dropped factor fails by 10% or more; the tolerances are the Riemann sums' own error
    over a 40-cell profile (1e-5 and 1.6e-4)."""
    from pixelflow_amd import cases
    vel, P, T = np.array([0.7, -0.4, 1.1]), 2.5, 1.5
    ext = [10, 8, 6]
    ext[axis] = 40
    m, n, l = ext
    p = cases.CaseParams(dim=3, m=m, n=n, l=l, width=1.0, height=0.6, depth=0.45, xnue=2e-3, density=1.3,
                         thickness=T, time=1e-3, istep_max=1, iter_max=5)
    prof = cases.tanh_profile(np.arange(ext[axis]) - (ext[axis] / 2 + 0.3), T)
    shape = [1, 1, 1]
    shape[2 - axis] = ext[axis]
    o = pyoracle.OracleRun(p, np.broadcast_to(prof.reshape(shape), (l, n, m)).copy())
    d = np.array([o.case.dx, o.case.dy, o.case.dz])
    o.p_[...] = P
    o.u[...], o.v[...], o.w[...] = vel
    f = o.force()
    area = np.prod([ext[a] * d[a] for a in range(3) if a != axis])
    assert abs(f[axis] / (-P * area) - 1) < 5e-5
    assert all(f[a] == 0.0 for a in range(3) if a != axis)      # no gradient across the other axes
    wall = p.density * p.xnue * 32.0 * vel * area * d[axis] / (12 * T * d ** 2)
    np.testing.assert_allclose(f[3:], wall, rtol=5e-4)


def test_force_3d_sphere_symmetry_and_linearity():
    """Constant pressure on a sphere centred in a box that is mirror-symmetric about the
    centre in every axis: the pressure force vanishes in x, y and z (opposite faces cancel);
    the wall term is linear in each velocity component and sees only its own one."""
    from pixelflow_amd import cases
    m, n, l = 24, 20, 16
    p = cases.CaseParams(dim=3, m=m, n=n, l=l, width=1.0, height=0.8, depth=0.6, time=1e-3, istep_max=1, iter_max=5)
    phi = cases.sphere_3d(m, n, l, 1.0, 0.8, 0.6, (0.5, 0.4, 0.3), 0.35)
    o = pyoracle.OracleRun(p, phi)
    o.p_[...] = 3.0
    unit = []
    for b in range(3):
        vel = [0.0, 0.0, 0.0]
        vel[b] = 1.0
        o.u[...], o.v[...], o.w[...] = vel
        f = o.force()
        assert np.all(np.abs(f[:3]) < 1e-12)                  # |each term| is up to about 0.1
        assert f[3 + b] > 0 and all(f[3 + a] == 0.0 for a in range(3) if a != b)
        unit.append(f[3 + b])
    o.u[...], o.v[...], o.w[...] = 2.0, -1.0, 0.5
    np.testing.assert_allclose(o.force()[3:], np.array(unit) * [2.0, -1.0, 0.5], rtol=1e-14)
