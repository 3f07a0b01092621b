"""Serving loop on the GPU: each request's result is the same alone and in the mix.

Requests of uneven size arrive while others run (a deterministic step clock places the
arrivals), wait for arena memory, join the resident batch between step chunks and leave
when done.  Every result must equal the bits of that case run alone through the
single-case path with the same settings.  The autotuner's choice must not change results
either (it only moves launch shapes).
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
    from pixelflow_amd.serve import Request
    specs = [(64, 32, 7, 0), (96, 48, 5, 0), (48, 48, 6, 2), (130, 66, 4, 3), (64, 33, 5, 4), (80, 40, 3, 9)]
    out = []
    for rid, (m, n, steps, arrival) in enumerate(specs):
        width = 2.0
        dx = width / (m - 1)
        xnue = dx * (1.0 + 0.2 * rid)              # cell Reynolds numbers 1 to 0.5: stable for long runs
        p = cases.CaseParams(dim=2, m=m, n=n, width=width, height=dx * (n - 1), xnue=xnue,
                             time=steps * cases.stable_dt(dx, xnue), istep_max=steps, iter_max=30)
        out.append(Request(rid, p, cases.random_circles_2d(m, n, 1 + rid % 3, (3.0, 8.0), seed=300 + rid), steps,
                           arrival))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_served_2d_results_equal_alone(precision):
    from pixelflow_amd.serve import Server
    from pixelflow_amd.solver import PF_PREC_FP64, PF_PREC_MIXED, batch_case_bytes
    prec = PF_PREC_MIXED if precision == "mixed" else PF_PREC_FP64
    kw = dict(inner_iters=10) if precision == "mixed" else {}
    reqs = _cyl_requests()
    sizes = sorted(batch_case_bytes(r.params, precision=prec, **kw) for r in reqs)
    arena = sizes[-1] + sizes[-2] + sizes[-3]          # about three cases at a time: later ones wait
    srv = Server(reqs[0].params, prec, max_cases=4, arena_bytes=arena, chunk=2, clock="steps", **kw)
    for r in reqs:
        srv.submit(r)
    done = srv.run()
    srv.close()
    assert sorted(r.rid for r in done) == list(range(len(reqs)))
    assert any(r.admitted > r.arrival for r in done), "no request had to wait for memory"
    for r in done:
        ref = _alone(r.params, r.phi, r.steps, precision=prec, **kw)
        for f in _fields(r.params):
            np.testing.assert_array_equal(r.result[f], ref[f], err_msg=f"request {r.rid} {f}")


def test_served_rooms_tolerance_mode_equal_alone():
    from pixelflow_amd.serve import Request, Server
    from pixelflow_amd.solver import PF_PREC_MIXED
    kw = dict(inner_iters=20, residual_tol=1e-6, max_cycles=20)
    specs = [((24, 24, 20), 1, 4, 0), ((32, 24, 24), 2, 3, 1), ((28, 20, 20), 3, 3, 2)]
    reqs = []
    for rid, (size, seed, steps, arrival) in enumerate(specs):
        p, phi = cases.room_case(seed, steps=steps, iter_max=20, size=size)
        reqs.append(Request(rid, p, phi, steps, arrival))
    srv = Server(reqs[0].params, PF_PREC_MIXED, max_cases=4, arena_bytes=1 << 27, chunk=1, clock="steps", **kw)
    for r in reqs:
        srv.submit(r)
    done = srv.run()
    srv.close()
    for r in done:
        ref = _alone(r.params, r.phi, r.steps, precision=PF_PREC_MIXED, **kw)
        for f in _fields(r.params):
            np.testing.assert_array_equal(r.result[f], ref[f], err_msg=f"room {r.rid} {f}")


def test_autotune_picks_a_config_and_keeps_results():
    from pixelflow_amd import autotune
    from pixelflow_amd.solver import Batch
    reqs = _cyl_requests()[:4]
    with Batch(reqs[0].params, max_cases=4, arena_bytes=1 << 27) as b:
        slots = [b.add(r.params, r.phi) for r in reqs]
        out = autotune.tune(b, "test-key", 30, False, steps=1, use_cache=False)
        assert len(out["candidates"]) == len(autotune.candidates(30, False))
        assert b.get_config()["tile_x"] == out["best"]["tile_x"]
        done = b.case_steps(slots[0])
        b.step(2)
        for r, s in zip(reqs, slots):
            ref = _alone(r.params, r.phi, done + 2)
            for f in _fields(r.params):
                np.testing.assert_array_equal(b.interior(s, f), ref[f], err_msg=f"case {r.rid} {f}")


def test_native_server_results_equal_alone():
    """The library's own serving loop (pf_serve_*, worker thread): requests submitted while
    others run, an arena that holds about three cases so later ones wait for memory;
    every result equals the case run alone, bit for bit."""
    import time
    from pixelflow_amd.solver import PF_PREC_MIXED, SERVE_DONE, NativeServer, batch_case_bytes
    kw = dict(inner_iters=10)
    reqs = _cyl_requests()
    sizes = sorted(batch_case_bytes(r.params, precision=PF_PREC_MIXED, **kw) for r in reqs)
    arena = sizes[-1] + sizes[-2] + sizes[-3]
    with NativeServer(reqs[0].params, max_cases=4, arena_bytes=arena, chunk=2, precision=PF_PREC_MIXED, **kw) as srv:
        ids = []
        for r in reqs:
            ids.append(srv.submit(r.params, r.phi, r.steps))
            time.sleep(0.002)
        for rid in ids:
            assert srv.wait(rid, 60.0) == SERVE_DONE, srv.error(rid)
        times = {rid: srv.times(rid) for rid in ids}
        # at least one request was admitted only after another had finished (memory wait)
        assert any(t[1] >= min(u[2] for q, u in times.items() if q != rid) for rid, t in times.items())
        for r, rid in zip(reqs, ids):
            ref = _alone(r.params, r.phi, r.steps, precision=PF_PREC_MIXED, **kw)
            for f in _fields(r.params):
                np.testing.assert_array_equal(srv.interior(rid, f), ref[f], err_msg=f"request {r.rid} {f}")
            srv.release(rid)
        c = srv.counters()
        assert c["queued"] == 0 and c["running"] == 0 and c["steps_run"] > 0


def test_native_server_rejects_case_larger_than_arena():
    from pixelflow_amd.solver import SERVE_DONE, SERVE_FAILED, NativeServer, batch_case_bytes
    reqs = _cyl_requests()
    small = min(reqs, key=lambda r: r.params.cells)
    big = max(reqs, key=lambda r: r.params.cells)
    with NativeServer(small.params, max_cases=2, arena_bytes=batch_case_bytes(small.params) + 4096, chunk=2) as srv:
        a = srv.submit(big.params, big.phi, 2)
        b = srv.submit(small.params, small.phi, 2)
        assert srv.wait(a, 60.0) == SERVE_FAILED and "arena" in srv.error(a)
        assert srv.wait(b, 60.0) == SERVE_DONE
