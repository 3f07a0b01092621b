"""Batched execution: every case of a batch gives the bits it gives alone.

A batch advances cases of uneven size with one launch per stage (persistent tile loops,
hipGraph-captured steps).  The per-cell arithmetic is the single-case path's, so each case
must match a single-case run of it exactly: FP64 (itself bit-identical to the reference),
mixed precision, graph and eager launch, persistent and one-block-per-tile grids, and
cases that join or leave between steps.
"""
import numpy as np
import pytest

from pixelflow_amd import cases

pytestmark = pytest.mark.gpu


def _fields(p):
    return ("u", "v", "w", "p") if p.dim == 3 else ("u", "v", "p")


def _single(p, phi, steps, **kw):
    from pixelflow_amd.solver import Solver
    with Solver(p, phi, **kw) as s:
        st = s.step(steps)
        return {f: s.interior(f) for f in _fields(p)}, st


def _mixed_2d_cases():
    out = []
    for idx, (m, n) in enumerate([(64, 64), (96, 48), (130, 66), (64, 33), (128, 64)]):
        width = 2.0
        dx = width / (m - 1)
        p = cases.CaseParams(dim=2, m=m, n=n, width=width, height=dx * (n - 1), xnue=2e-3 + 1e-3 * idx,
                             time=6 * 0.2 * dx, istep_max=6, iter_max=30)
        out.append((p, cases.random_circles_2d(m, n, 1 + idx % 3, (3.0, 8.0), seed=100 + idx)))
    return out


@pytest.mark.parametrize("tile_x", [32, 64, 128])
@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("persistent", [0, 97])
def test_batch_2d_fp64_matches_single(graph, persistent, tile_x):
    from pixelflow_amd.solver import Batch
    cs = _mixed_2d_cases()
    kind = cs[0][0]
    with Batch(kind, max_cases=8, arena_bytes=1 << 28) as b:
        b.set_config(persistent_blocks=persistent, use_graph=graph, tile_x=tile_x)
        slots = [b.add(p, phi) for p, phi in cs]
        st = b.step(6)
        for (p, phi), s in zip(cs, slots):
            ref, rst = _single(p, phi, 6)
            for f in _fields(p):
                np.testing.assert_array_equal(b.interior(s, f), ref[f], err_msg=f"slot {s} {p.m}x{p.n} {f}")
            np.testing.assert_array_equal(st["sor_error"][:, s], rst["sor_error"])


def test_batch_3d_fp64_matches_single_odd_and_even():
    from pixelflow_amd.solver import Batch
    specs = [(32, 16, 16), (33, 17, 15), (24, 20, 18), (40, 24, 16)]
    cs = []
    for (m, n, l) in specs:
        p = cases.CaseParams(dim=3, m=m, n=n, l=l, width=1.0, height=(n - 1) / (m - 1), depth=(l - 1) / (m - 1),
                             time=4 * 1e-3, istep_max=4, iter_max=25, xlambda=1e-4)
        cs.append((p, cases.sphere_3d(m, n, l, 1.0, p.height, p.depth, (0.3, p.height / 2, p.depth / 2), 0.15)))
    with Batch(cs[0][0], max_cases=4, arena_bytes=1 << 28) as b:
        b.set_config(tile_x=32)
        slots = [b.add(p, phi) for p, phi in cs]
        b.step(4)
        for (p, phi), s in zip(cs, slots):
            ref, _ = _single(p, phi, 4)
            for f in _fields(p):
                np.testing.assert_array_equal(b.interior(s, f), ref[f], err_msg=f"{p.m}x{p.n}x{p.l} {f}")


def test_batch_2d_mixed_matches_single_mixed():
    from pixelflow_amd.solver import PF_PREC_MIXED, Batch
    cs = _mixed_2d_cases()
    with Batch(cs[0][0], max_cases=8, arena_bytes=1 << 28, precision=PF_PREC_MIXED, inner_iters=10) as b:
        b.set_config(persistent_blocks=128, use_graph=True, inner_iters=10)
        slots = [b.add(p, phi) for p, phi in cs]
        b.step(6)
        for (p, phi), s in zip(cs, slots):
            ref, _ = _single(p, phi, 6, precision=PF_PREC_MIXED, inner_iters=10)
            for f in _fields(p):
                np.testing.assert_array_equal(b.interior(s, f), ref[f], err_msg=f"{p.m}x{p.n} {f}")


def test_batch_join_and_leave_between_steps():
    """Continuous batching: a case that joins later and one that leaves early give the
    same results as run alone."""
    from pixelflow_amd.solver import Batch
    cs = _mixed_2d_cases()
    with Batch(cs[0][0], max_cases=4, arena_bytes=1 << 27) as b:
        s0 = b.add(*cs[0])
        s1 = b.add(*cs[1])
        b.step(2)
        b.remove(s1)
        s2 = b.add(*cs[2])           # reuses the freed slot and arena range
        b.step(4)
        ref0, _ = _single(*cs[0], 6)
        ref2, _ = _single(*cs[2], 4)
        assert b.case_steps(s0) == 6 and b.case_steps(s2) == 4
        for f in ("u", "v", "p"):
            np.testing.assert_array_equal(b.interior(s0, f), ref0[f])
            np.testing.assert_array_equal(b.interior(s2, f), ref2[f])


@pytest.mark.parametrize("max_cycles,graph", [(50, True), (50, False), (8, True), (8, False)])
def test_batch_air_tolerance_mode_matches_single(max_cycles, graph):
    """Tolerance mode decides convergence on the device (captured whole when the cycle cap
    is small, else launched with a periodic check); iteration counts and fields must equal
    the single-case run's."""
    from pixelflow_amd.solver import PF_PREC_MIXED, Batch
    specs = [(24, 24, 20, 1), (32, 24, 24, 2)]
    cs = [cases.room_case(seed, steps=3, iter_max=20, size=(m, n, l)) for (m, n, l, seed) in specs]
    kw = dict(precision=PF_PREC_MIXED, inner_iters=20, residual_tol=1e-6, max_cycles=max_cycles)
    with Batch(cs[0][0], max_cases=4, arena_bytes=1 << 27, **kw) as b:
        b.set_config(use_graph=graph, inner_iters=20)
        slots = [b.add(p, phi) for p, phi in cs]
        st = b.step(3)
        for (p, phi), s in zip(cs, slots):
            ref, rst = _single(p, phi, 3, **kw)
            np.testing.assert_array_equal(st["sor_iters"][:, s], rst["sor_iters"])
            np.testing.assert_array_equal(st["cycles"][:, s], rst["cycles"])
            np.testing.assert_array_equal(st["residual"][:, s], rst["residual"])
            for f in _fields(p):
                np.testing.assert_array_equal(b.interior(s, f), ref[f], err_msg=f"{p.m}x{p.n}x{p.l} {f}")


def _odd_even_3d_cases():
    specs = [(32, 16, 16), (33, 17, 15), (24, 20, 18), (40, 24, 16)]
    cs = []
    for (m, n, l) in specs:
        p = cases.CaseParams(dim=3, m=m, n=n, l=l, width=1.0, height=(n - 1) / (m - 1), depth=(l - 1) / (m - 1),
                             time=4 * 1e-3, istep_max=4, iter_max=25, xlambda=1e-4)
        cs.append((p, cases.sphere_3d(m, n, l, 1.0, p.height, p.depth, (0.3, p.height / 2, p.depth / 2), 0.15)))
    return cs


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kb", [2, 4])
def test_batch_3d_cells_per_thread_matches_single(kb, precision, monkeypatch):
    """3D sweep tiles of 2 or 4 cells per thread along z (odd l leaves planes past the last
    one in the last tile row), then a switch to the persistent grid, which re-tiles with
    one cell per thread: every case keeps the bits it gives alone."""
    from pixelflow_amd.solver import PF_PREC_MIXED, Batch
    monkeypatch.setenv("PF_BATCH_KB", str(kb))
    kw = dict(precision=PF_PREC_MIXED, inner_iters=10) if precision == "mixed" else {}
    cs = _odd_even_3d_cases()
    with Batch(cs[0][0], max_cases=4, arena_bytes=1 << 28, **kw) as b:
        b.set_config(tile_x=32, inner_iters=kw.get("inner_iters", 0))
        slots = [b.add(p, phi) for p, phi in cs]
        st = b.step(2)
        b.set_config(tile_x=32, persistent_blocks=64, inner_iters=kw.get("inner_iters", 0))
        st2 = b.step(2)
        for (p, phi), s in zip(cs, slots):
            ref, rst = _single(p, phi, 4, **kw)
            for f in _fields(p):
                np.testing.assert_array_equal(b.interior(s, f), ref[f], err_msg=f"kb{kb} {p.m}x{p.n}x{p.l} {f}")
            np.testing.assert_array_equal(np.concatenate([st["sor_error"][:, s], st2["sor_error"][:, s]]),
                                          rst["sor_error"])


@pytest.mark.parametrize("streams", [1, 2])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_grouped_solve_matches_single(precision, streams, monkeypatch):
    """Solve sweeps run group by group (here every case its own group, 64 KiB cap), on the
    batch's stream or dealt to two side streams: every case keeps the bits it gives alone,
    fixed count and tolerance mode, graph-captured."""
    from pixelflow_amd.solver import PF_PREC_MIXED, Batch
    monkeypatch.setenv("PF_BATCH_GROUP_KB", "64")
    monkeypatch.setenv("PF_BATCH_STREAMS", str(streams))
    kw = dict(precision=PF_PREC_MIXED, inner_iters=10) if precision == "mixed" else {}
    cs = _odd_even_3d_cases()
    with Batch(cs[0][0], max_cases=4, arena_bytes=1 << 28, **kw) as b:
        b.set_config(inner_iters=kw.get("inner_iters", 0))
        slots = [b.add(p, phi) for p, phi in cs]
        st = b.step(3)
        for (p, phi), s in zip(cs, slots):
            ref, rst = _single(p, phi, 3, **kw)
            for f in _fields(p):
                np.testing.assert_array_equal(b.interior(s, f), ref[f], err_msg=f"grouped {p.m}x{p.n}x{p.l} {f}")
            np.testing.assert_array_equal(st["sor_error"][:, s], rst["sor_error"])
    cs2 = _mixed_2d_cases()     # includes odd extents: the periodic ghost refresh runs per group
    with Batch(cs2[0][0], max_cases=8, arena_bytes=1 << 28, **kw) as b:
        b.set_config(inner_iters=kw.get("inner_iters", 0))
        slots = [b.add(p, phi) for p, phi in cs2]
        b.step(3)
        for (p, phi), s in zip(cs2, slots):
            ref, _ = _single(p, phi, 3, **kw)
            for f in _fields(p):
                np.testing.assert_array_equal(b.interior(s, f), ref[f], err_msg=f"grouped 2D {p.m}x{p.n} {f}")
    rooms = [cases.room_case(seed, steps=3, iter_max=20, size=size)
             for seed, size in ((1, (24, 24, 20)), (2, (32, 24, 24)), (3, (20, 28, 22)))]
    tkw = dict(precision=PF_PREC_MIXED if precision == "mixed" else 0, inner_iters=10, residual_tol=1e-6,
               max_cycles=4)
    with Batch(rooms[0][0], max_cases=4, arena_bytes=1 << 27, **tkw) as b:
        b.set_config(inner_iters=10)
        slots = [b.add(p, phi) for p, phi in rooms]
        st = b.step(3)
        for (p, phi), s in zip(rooms, slots):
            ref, rst = _single(p, phi, 3, **tkw)
            np.testing.assert_array_equal(st["sor_iters"][:, s], rst["sor_iters"])
            for f in _fields(p):
                np.testing.assert_array_equal(b.interior(s, f), ref[f], err_msg=f"grouped room {p.m}x{p.n}x{p.l} {f}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_backstep_variant_beside_plain_inlet(precision):
    """The backward-step variant (inlet and initial velocity scaled by porosity) in a batch
    beside the same grid with the plain inlet and a cylinder case: FP64 reproduces the
    reference's backstep fixture bit for bit, and every case gives the bits it gives alone."""
    import dataclasses

    from conftest import load_golden
    from pixelflow_amd.solver import PF_PREC_MIXED, Batch
    p0, g = load_golden("g2d_backstep")
    p1 = dataclasses.replace(p0, inlet_porosity=False)
    m, n = 64, 32
    dx = 2.0 / (m - 1)
    dt = cases.stable_dt(dx, dx)             # cell Reynolds number 1
    p2 = cases.CaseParams(dim=2, m=m, n=n, width=2.0, height=dx * (n - 1), xnue=dx, time=p0.istep_max * dt,
                          istep_max=p0.istep_max, iter_max=p0.iter_max)
    cs = [(p0, g["phi_raw"]), (p1, g["phi_raw"]), (p2, cases.random_circles_2d(m, n, 2, (3.0, 6.0), seed=3))]
    kw = dict(precision=PF_PREC_MIXED, inner_iters=10) if precision == "mixed" else {}
    with Batch(p0, max_cases=4, arena_bytes=1 << 27, **kw) as b:
        b.set_config(inner_iters=kw.get("inner_iters", 0))
        slots = [b.add(p, phi) for p, phi in cs]
        st = b.step(p0.istep_max)
        if precision == "fp64":
            for f in ("u", "v", "p"):
                np.testing.assert_array_equal(b.interior(slots[0], f), g[f], err_msg=f"backstep {f}")
            np.testing.assert_array_equal(st["sor_error"][:, slots[0]], g["errs"])
        for (p, phi), s in zip(cs, slots):
            ref, rst = _single(p, phi, p0.istep_max, **kw)
            for f in _fields(p):
                np.testing.assert_array_equal(b.interior(s, f), ref[f], err_msg=f"slot {s} {f}")
            np.testing.assert_array_equal(st["sor_error"][:, s], rst["sor_error"])


def test_gpu_batched_2d_workload_fp64_leg_bit_identical():
    """The batched 2D workload on a few cases: the FP64 batch of the same cases gives each
    sampled case the bits it gets alone, and no case diverges in the mixed batch."""
    from pixelflow_amd import workloads
    print("[gpu] batched_2d workload, 8 cases", flush=True)
    r = workloads.batched_2d(count=8, steps=20, plain_sample=2, plain_steps=4, accuracy_sample=2, tune_steps=1)
    assert r["diverged_cases"] == 0
    assert r["fp64_batch"]["sample_bit_identical_to_alone"] is True
    assert r["fp64_batch"]["cases_per_s"] > 0 and r["cases_per_s"] > 0
