"""FP64 parity over the benchmark cases' full step counts, against the reference build.

* 2D cylinder, 256x128, 2000 steps (ibm_2d_uniform_omp_cpu.f90:80-125): the final u, v, p
  of the reference's own OpenMP build (FP64) in full; the GPU path must match bit for bit
  (the written bar, relative L2 <= 1e-12, is checked first so a failure says how far off).
* 3D sphere, 256x128x128, 500 steps (ibm_3d_uniform_omp_cpu.f90:81-132): SHA-256 digests of
  the reference's final u, v, w, p and divergence, its per-step SOR diagnostic, and the mid
  z plane of u and p.
"""
import hashlib

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("tb2", ["0,0", "2,1", "1,3"])
def test_c1_cylinder_2000_steps_matches_reference(tb2, monkeypatch):
    """With the two-launch colour sweeps and with the temporally blocked 2D sweep (2
    iterations per launch and 1 row per block; 1 iteration and 3 rows)."""
    from pixelflow_amd.solver import Solver
    monkeypatch.setenv("PF_SOR_TB2", tb2)
    p, g = load_golden("g2d_c1_2000")
    assert (p.m, p.n, p.istep_max, p.iter_max) == (256, 128, 2000, 100)
    with Solver(p, _c1_porosity(p)) as s:
        st = s.step(p.istep_max)
        for f in ("u", "v", "p"):
            got = s.interior(f)
            assert _rel(got, g[f]) <= 1e-12, (f, _rel(got, g[f]))
            np.testing.assert_array_equal(got, g[f], err_msg=f)
    np.testing.assert_array_equal(st["sor_error"], g["errs"])


def test_c2_sphere_500_steps_matches_reference():
    from pixelflow_amd.solver import Solver
    p, g = load_golden("g3d_c2_500")
    assert (p.m, p.n, p.l, p.istep_max, p.iter_max) == (256, 128, 128, 500, 100)
    from pixelflow_amd import cases
    _, phi = cases.sphere_case(m=p.m, n=p.n, l=p.l, steps=p.istep_max)
    with Solver(p, phi) as s:
        st = s.step(p.istep_max)
        fields = {f: s.interior(f) for f in ("u", "v", "w", "p")}
        fields["div"] = s.divergence()
    for f in ("u", "p"):
        plane = fields[f][64]
        assert _rel(plane, g[f + "_plane064"]) <= 1e-12, (f, _rel(plane, g[f + "_plane064"]))
    for f, a in fields.items():
        assert _sha(a) == str(g["sha_" + f]), f
    np.testing.assert_array_equal(st["sor_error"], g["errs"])


def _c1_porosity(p):
    from pixelflow_amd import cases
    _, phi = cases.cylinder_case(m=p.m, n=p.n, steps=p.istep_max)
    return phi
