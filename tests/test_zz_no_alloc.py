"""pf_force_noalloc (the body force reduced in the library's fixed device scratch) against
pf_force (the same kernels on a buffer the context allocates), through the C-ABI.

Two calls in a row, the second on a smaller grid (fewer reduction blocks than the first
left in the shared scratch), each compared bit for bit with pf_force on the same state;
and the error path for a context that has no fields yet.
"""
import ctypes

import numpy as np
import pytest

from pixelflow_amd import cases

pytestmark = pytest.mark.gpu


def _force_noalloc(s):
    from pixelflow_amd import solver as S
    out = np.zeros(6, dtype=np.float64)
    rc = S.lib().pf_force_noalloc(s._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, S.lib().pf_last_error(s._h).decode()
    return out


def _case(m, n, l, steps):
    p, phi = cases.sphere_case(m=m, n=n, l=l, steps=steps, iter_max=20)
    return p, phi


@pytest.mark.parametrize("first,second", [((96, 48, 48), (32, 16, 16))])
def test_force_noalloc_matches_force_twice_second_smaller(first, second):
    from pixelflow_amd.solver import Solver
    for (m, n, l) in (first, second):
        p, phi = _case(m, n, l, 3)
        with Solver(p, phi) as s:
            s.step(3)
            a = _force_noalloc(s)
            b = s.force()
            np.testing.assert_array_equal(a, b, err_msg=f"{m}x{n}x{l}")
            assert np.all(np.isfinite(a)) and np.any(a != 0.0)


def test_force_noalloc_2d_after_3d():
    from pixelflow_amd.solver import Solver
    p, phi = cases.cylinder_case(m=64, n=32, steps=2, iter_max=10)
    with Solver(p, phi) as s:
        s.step(2)
        np.testing.assert_array_equal(_force_noalloc(s), s.force())


def test_force_noalloc_rejects_null_output():
    from pixelflow_amd import solver as S
    p, phi = _case(32, 16, 16, 1)
    with S.Solver(p, phi) as s:
        assert S.lib().pf_force_noalloc(s._h, None) == -1
