"""Non-finite values are detected wherever a step can produce them.

The sweeps keep a NaN in their running max of |dp| (a finite run folds exactly as the
reference's max), the residual reductions keep it too, and tolerance-mode solves stop at
once on a non-finite residual instead of cycling to their cap.  So a NaN that reaches the
pressure makes pf_step / pf_batch_step return PF_EDIVERGED, a batch never retires such a
case as converged, and the server marks the request failed.

Cases are seeded non-finite in two ways: a NaN uploaded into the velocity of a running
case (pf_upload), and an infinite porosity value in one cell of an added case (the clamp
max(phi, threshold) keeps it, and its faces make the pressure NaN within two steps; the
oracle does the same, checked here on the CPU first).
"""
import math

import numpy as np
import pytest

from pixelflow_amd import cases

pytestmark = pytest.mark.gpu

PF_EDIVERGED = -5


def _sphere(steps=3):
    return cases.sphere_case(m=24, n=12, l=12, steps=steps, iter_max=10)


def _poisoned_2d(m=48, n=24):
    p, phi = cases.cylinder_case(m=m, n=n, steps=6, iter_max=20)
    phi = phi.copy()
    phi[n // 2, m // 3] = np.inf
    return p, phi


def test_oracle_makes_nan_from_infinite_porosity():
    import pyoracle
    p, phi = _poisoned_2d()
    o = pyoracle.OracleRun(p, phi)
    with np.errstate(all="ignore"):
        e = o.step(2)
    assert np.isnan(e[-1]) and np.isnan(o.interior("p")).any()


def _upload_nan_u(s):
    u = s.field("u")
    u[6, 6, 8] = np.nan
    s.set_field("u", u)


def test_fp64_step_reports_nan_velocity():
    from pixelflow_amd.solver import PixelFlowError, Solver
    p, phi = _sphere()
    with Solver(p, phi) as s:
        s.step(1)
        _upload_nan_u(s)
        with pytest.raises(PixelFlowError, match="PF_EDIVERGED"):
            s.step(1)


@pytest.mark.parametrize("tol", [0.0, 1e-8])
def test_mixed_step_reports_nan_velocity_and_stops_cycling(tol):
    from pixelflow_amd import solver as S
    p, phi = _sphere()
    with S.Solver(p, phi, precision=S.PF_PREC_MIXED, inner_iters=5, residual_tol=tol, max_cycles=40) as s:
        s.step(1)
        _upload_nan_u(s)
        st = np.zeros(1, dtype=S.STATS_DTYPE)
        rc = S.lib().pf_step(s._h, 1, S._stats_ptr(st))
        assert rc == PF_EDIVERGED
        if tol > 0:
            # the residual before the first cycle is already NaN: no cycle runs
            assert math.isnan(st["residual"][0]) and st["cycles"][0] == 0


def test_batch_tolerance_mode_never_converges_a_nan_case():
    from pixelflow_amd import solver as S
    good_p, good_phi = cases.cylinder_case(m=48, n=24, steps=6, iter_max=20)
    bad_p, bad_phi = _poisoned_2d()
    kw = dict(precision=S.PF_PREC_MIXED, inner_iters=5, residual_tol=1e-6, max_cycles=8)
    with S.Batch(good_p, max_cases=4, arena_bytes=1 << 26, **kw) as b:
        sg = b.add(good_p, good_phi)
        sb = b.add(bad_p, bad_phi)
        st = b.step(4, raise_diverged=False)
        good = {f: b.interior(sg, f) for f in ("u", "v", "p")}
    bad_res = st["residual"][:, sb]
    bad_err = st["sor_error"][:, sb]
    # the poisoned case goes non-finite and is never reported as converged with a finite residual
    assert not np.all(np.isfinite(bad_res) & np.isfinite(bad_err))
    first = int(np.argmax(~(np.isfinite(bad_res) & np.isfinite(bad_err))))
    assert not (np.isfinite(bad_res[first:]) & (bad_res[first:] <= 1e-6)).any()
    # its companion is unaffected: the same bits as alone
    with S.Solver(good_p, good_phi, **kw) as s:
        ref_st = s.step(4)
        for f in ("u", "v", "p"):
            np.testing.assert_array_equal(good[f], s.interior(f))
    np.testing.assert_array_equal(st["residual"][:, sg], ref_st["residual"])
    # alone, the poisoned case fails the same way
    with S.Solver(bad_p, bad_phi, **kw) as s:
        with pytest.raises(S.PixelFlowError, match="PF_EDIVERGED"):
            s.step(4)


def test_server_fails_a_nan_request_and_finishes_the_other():
    from pixelflow_amd import solver as S
    good_p, good_phi = cases.cylinder_case(m=48, n=24, steps=6, iter_max=20)
    bad_p, bad_phi = _poisoned_2d()
    with S.NativeServer(good_p, max_cases=4, arena_bytes=1 << 26, chunk=2) as srv:
        rg = srv.submit(good_p, good_phi, 6, fields=("u", "v", "p"))
        rb = srv.submit(bad_p, bad_phi, 6, fields=("u", "v", "p"))
        assert srv.wait(rb, 60.0) == S.SERVE_FAILED
        assert "diverged" in srv.error(rb)
        assert srv.wait(rg, 60.0) == S.SERVE_DONE
