"""The library's serving loop (pf_serve_*) on the GPU: each request's result is the same
alone and in the mix.

Requests of uneven size are submitted to a server whose arena holds about three of them, so
later ones wait for memory, join the resident batch between step chunks and leave when
done.  Every result must equal the bits of that case run alone through the single-case path
with the same settings, whatever the timing of admissions.  The autotuner's choice must not
change results either (it only moves launch shapes).  The loop's policy itself (admission
order, chunk sizes, retirement) is checked step by step on the CPU in
tests/test_serve_native_host.py.
"""
import numpy as np
import pytest

from pixelflow_amd import cases

pytestmark = pytest.mark.gpu


def _fields(p):
    return ("u", "v", "w", "p") if p.dim == 3 else ("u", "v", "p")


def _alone(p, phi, steps, **kw):
    from pixelflow_amd.solver import Solver
    with Solver(p, phi, **kw) as s:
        s.step(steps)
        return {f: s.interior(f) for f in _fields(p)}


def _cyl_requests():
    """(params, porosity, steps) of six 2D channels with 1 to 3 cylinders."""
    specs = [(64, 32, 7), (96, 48, 5), (48, 48, 6), (130, 66, 4), (64, 33, 5), (80, 40, 3)]
    out = []
    for rid, (m, n, steps) in enumerate(specs):
        width = 2.0
        dx = width / (m - 1)
        xnue = dx * (1.0 + 0.2 * rid)              # cell Reynolds numbers 1 to 0.5: stable for long runs
        p = cases.CaseParams(dim=2, m=m, n=n, width=width, height=dx * (n - 1), xnue=xnue,
                             time=steps * cases.stable_dt(dx, xnue), istep_max=steps, iter_max=30)
        out.append((p, cases.random_circles_2d(m, n, 1 + rid % 3, (3.0, 8.0), seed=300 + rid), steps))
    return out


def _serve_all(reqs, arena, chunk, prec, **kw):
    """Submit every request, wait for all; returns (per-request fields, times)."""
    from pixelflow_amd.solver import SERVE_DONE, NativeServer
    with NativeServer(reqs[0][0], max_cases=4, arena_bytes=arena, chunk=chunk, precision=prec, **kw) as srv:
        ids = [srv.submit(p, phi, steps) for p, phi, steps in reqs]
        for rid in ids:
            assert srv.wait(rid, 120.0) == SERVE_DONE, srv.error(rid)
        out = [{f: srv.interior(rid, f) for f in _fields(p)} for rid, (p, _, _) in zip(ids, reqs)]
        times = [srv.times(rid) for rid in ids]
        for rid, (_, _, steps) in zip(ids, reqs):
            assert srv.poll(rid)[1] == steps
            srv.release(rid)
        c = srv.counters()
        assert c["queued"] == 0 and c["running"] == 0 and c["steps_run"] > 0
    return out, times


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_served_2d_results_equal_alone(precision):
    from pixelflow_amd.solver import PF_PREC_FP64, PF_PREC_MIXED, batch_case_bytes
    prec = PF_PREC_MIXED if precision == "mixed" else PF_PREC_FP64
    kw = dict(inner_iters=10) if precision == "mixed" else {}
    reqs = _cyl_requests()
    sizes = sorted(batch_case_bytes(p, precision=prec, **kw) for p, _, _ in reqs)
    arena = sizes[-1] + sizes[-2] + sizes[-3]          # about three cases at a time: later ones wait
    got, times = _serve_all(reqs, arena, 2, prec, **kw)
    # at least one request was admitted only after another had finished (memory wait)
    assert any(t[1] >= min(u[2] for j, u in enumerate(times) if j != i) for i, t in enumerate(times))
    for i, ((p, phi, steps), g) in enumerate(zip(reqs, got)):
        ref = _alone(p, phi, steps, precision=prec, **kw)
        for f in _fields(p):
            np.testing.assert_array_equal(g[f], ref[f], err_msg=f"request {i} {f}")


def test_served_rooms_tolerance_mode_equal_alone():
    from pixelflow_amd.solver import PF_PREC_MIXED
    kw = dict(inner_iters=20, residual_tol=1e-6, max_cycles=20)
    reqs = []
    for size, seed, steps in (((24, 24, 20), 1, 4), ((32, 24, 24), 2, 3), ((28, 20, 20), 3, 3)):
        p, phi = cases.room_case(seed, steps=steps, iter_max=20, size=size)
        reqs.append((p, phi, steps))
    got, _ = _serve_all(reqs, 1 << 27, 1, PF_PREC_MIXED, **kw)
    for i, ((p, phi, steps), g) in enumerate(zip(reqs, got)):
        ref = _alone(p, phi, steps, precision=PF_PREC_MIXED, **kw)
        for f in _fields(p):
            np.testing.assert_array_equal(g[f], ref[f], err_msg=f"room {i} {f}")


def test_autotune_picks_a_config_and_keeps_results():
    from pixelflow_amd import autotune
    from pixelflow_amd.solver import Batch
    reqs = _cyl_requests()[:4]
    with Batch(reqs[0][0], max_cases=4, arena_bytes=1 << 27) as b:
        slots = [b.add(p, phi) for p, phi, _ in reqs]
        out = autotune.tune(b, "test-key", 30, False, steps=1, use_cache=False)
        assert len(out["candidates"]) == len(autotune.candidates(30, False))
        assert b.get_config()["tile_x"] == out["best"]["tile_x"]
        done = b.case_steps(slots[0])
        b.step(2)
        for i, ((p, phi, _), s) in enumerate(zip(reqs, slots)):
            ref = _alone(p, phi, done + 2)
            for f in _fields(p):
                np.testing.assert_array_equal(b.interior(s, f), ref[f], err_msg=f"case {i} {f}")


def test_native_server_rejects_case_larger_than_arena():
    from pixelflow_amd.solver import SERVE_DONE, SERVE_FAILED, NativeServer, batch_case_bytes
    reqs = _cyl_requests()
    small = min(reqs, key=lambda r: r[0].cells)
    big = max(reqs, key=lambda r: r[0].cells)
    with NativeServer(small[0], max_cases=2, arena_bytes=batch_case_bytes(small[0]) + 4096, chunk=2) as srv:
        a = srv.submit(big[0], big[1], 2)
        b = srv.submit(small[0], small[1], 2)
        assert srv.wait(a, 60.0) == SERVE_FAILED and "arena" in srv.error(a)
        assert srv.wait(b, 60.0) == SERVE_DONE
