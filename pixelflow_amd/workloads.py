"""Drivers of the batched and served workloads (used by bench.py).

* :func:`batched_2d` -- many independent 2D cylinder-array cases in one batch, FP32 SOR with
  FP64 defect correction, tuned by the autotuner, against the plain one-case-at-a-time FP64
  path on a sample of the same cases;
* :func:`served_rooms` -- Poisson-arriving 3D room-ventilation requests through the
  continuous-batching server, mixed precision, every step solved to the residual rule by
  the multigrid-preconditioned solve, reporting p50/p99 time-to-solution;
* :func:`rooms_batch` -- the same rooms resident together in one batch (no arrivals), FP64
  and mixed, beside each room run alone: the batch ceiling the server works under.
"""
from __future__ import annotations

import time
from typing import Callable, Dict, Optional

import numpy as np

from . import autotune, cases
from .solver import PF_PREC_FP64, PF_PREC_MIXED, PF_SOLVER_MG, Batch, Solver, batch_case_bytes, device_memory


def _noop(*a):
    pass


def batched_2d(count: int = 512, steps: int = 1000, iter_max: int = 60, seed: int = 11, tune_steps: int = 2,
               plain_sample: int = 16, plain_steps: int = 20, accuracy_sample: int = 5, fp64_batch: bool = True,
               log: Callable = _noop) -> Dict[str, object]:
    cl = cases.batch_cases_2d(count=count, steps=steps, seed=seed, iter_max=iter_max)
    kind = cl[0][0]
    kw = dict(precision=PF_PREC_MIXED, inner_iters=iter_max // 2)
    arena = sum(batch_case_bytes(p, **kw) for p, _ in cl) + (64 << 20)
    total_cells = sum(p.cells for p, _ in cl)
    sample_idx = list(range(0, len(cl), max(1, len(cl) // plain_sample)))[:plain_sample]
    accuracy_idx = list(range(0, len(cl), max(1, len(cl) // accuracy_sample)))[:accuracy_sample]
    key = autotune.size_class_key(2, "uniform", True, iter_max, [p.cells for p, _ in cl])
    log(f"[batch] {count} cases, {total_cells} cells, arena {arena / 2**30:.2f} GiB; autotuning")
    with Batch(kind, max_cases=count, arena_bytes=arena, **kw) as tb:
        for p, phi in cl:
            tb.add(p, phi)
        tuned = autotune.tune(tb, key, iter_max, True, steps=tune_steps, use_cache=False, log=log)
        tb.sync()
    cfg = tuned["best"]
    log(f"[batch] tuned config {cfg}")
    with Batch(kind, max_cases=count, arena_bytes=arena, **kw) as b:
        b.set_config(**cfg)
        slots = [b.add(p, phi) for p, phi in cl]
        b.step(1)                                  # captures the graphs
        b.sync()
        t0 = time.perf_counter()
        st = b.step(steps, raise_diverged=False)
        b.sync()
        wall = time.perf_counter() - t0
        diverged = int(np.sum(~np.all(np.isfinite(st["sor_error"][:, slots]), axis=0)))
        # accuracy on a sample: the batched mixed run vs each case alone in FP64
        accuracy, refs = [], {}
        for i in accuracy_idx:
            s = slots[i]
            p, phi = cl[i]
            with Solver(p, phi) as ref:
                ref.step(steps + 1)
                refs[i] = {f: ref.interior(f) for f in ("u", "v", "p")}
                accuracy.append({"case": f"{p.m}x{p.n}", **{f: float(np.linalg.norm(b.interior(s, f) - refs[i][f])
                                                                       / max(np.linalg.norm(refs[i][f]), 1e-300))
                                                              for f in ("u", "v", "p")}})
    # the same cases batched in FP64: the batching gain alone, every case bit for bit as alone
    fp64 = None
    if fp64_batch:
        kw64 = dict(precision=PF_PREC_FP64)
        arena64 = sum(batch_case_bytes(p, **kw64) for p, _ in cl) + (64 << 20)
        with Batch(kind, max_cases=count, arena_bytes=arena64, **kw64) as b64:
            b64.set_config(tile_x=cfg["tile_x"], persistent_blocks=cfg["persistent_blocks"], use_graph=True)
            slots64 = [b64.add(p, phi) for p, phi in cl]
            b64.step(1)
            b64.sync()
            t0 = time.perf_counter()
            b64.step(steps, raise_diverged=False)
            b64.sync()
            wall64 = time.perf_counter() - t0
            identical = all(np.array_equal(b64.interior(slots64[i], f), refs[i][f])
                            for i in accuracy_idx for f in ("u", "v", "p"))
        fp64 = {"wall_s": wall64, "cases_per_s": count / wall64, "cell_steps_per_s": total_cells * steps / wall64,
                "sample_bit_identical_to_alone": bool(identical), "sample": f"{len(accuracy_idx)} of the cases"}
    cell_steps = total_cells * steps / wall
    # plain GPU path: the same cases, one at a time, FP64 (the reference's arithmetic)
    plain_cells, plain_time = 0, 0.0
    for i in sample_idx:
        p, phi = cl[i]
        with Solver(p, phi) as s:
            s.step(2)
            s.sync()
            t0 = time.perf_counter()
            s.step(plain_steps)
            s.sync()
            plain_time += time.perf_counter() - t0
            plain_cells += p.cells * plain_steps
    plain = plain_cells / plain_time
    return {
        "workload": f"batch2d_{count}_cases_cylinders",
        "cases": count, "total_cells": total_cells, "steps": steps, "iter_max": iter_max,
        "precision": "fp32 SOR + fp64 defect correction (fixed count)",
        "wall_s": wall, "cases_per_s": count / wall,
        "cell_steps_per_s": cell_steps, "sor_cell_updates_per_s": cell_steps * iter_max,
        "plain_fp64_cell_steps_per_s": plain, "speedup_vs_plain_gpu": cell_steps / plain,
        "plain_sample": f"{plain_sample} of the cases, one at a time, {plain_steps} steps each",
        "autotune": {"best": cfg, "candidates": len(tuned.get("candidates", []))},
        "accuracy_vs_fp64_alone": accuracy, "diverged_cases": diverged,
        "fp64_batch": (None if fp64 is None else
                       {**fp64, "speedup_vs_plain_gpu": fp64["cell_steps_per_s"] / plain}),
    }


def _room_requests(n_requests: int, rate_per_s: float, steps: int, iter_max: int, seed: int, sizes=None):
    """The served rooms and their arrival times (seconds from the first)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, size=n_requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    reqs = []
    for r in range(n_requests):
        size = None if sizes is None else sizes[r % len(sizes)]
        reqs.append(cases.room_case(int(rng.integers(0, 2**31 - 1)), steps=steps, iter_max=iter_max, size=size))
    return reqs, arrivals


def _mg_kw(tol: float) -> Dict[str, object]:
    return dict(residual_tol=tol, solver=PF_SOLVER_MG, max_cycles=30)


def served_rooms(n_requests: int = 16, rate_per_s: float = 2.0, steps: int = 300, iter_max: int = 100,
                 tol: float = 1e-9, seed: int = 5, sizes=None, chunk: int = 5,
                 plain_sample: Optional[int] = None, plain_steps: int = 10, accuracy_sample: int = 2,
                 log: Callable = _noop) -> Dict[str, object]:
    """Poisson-arriving room requests through the library's serving loop (pf_serve_*): a
    host thread submits each request at its arrival time while the worker steps the
    resident ones.  Every step solves the pressure to the residual rule max|r/ap| <= tol
    (the multigrid-preconditioned solve, mixed precision: FP32 inner solve, FP64 residual
    and update).  The served rate is set beside the FP64 path under the same rule, each
    room alone (`plain_sample` rooms, `plain_steps` steps each); the `accuracy_sample`
    largest requests keep their fields, compared afterwards with that FP64 path over all
    the steps."""
    from .solver import NativeServer, SERVE_DONE
    reqs, arrivals = _room_requests(n_requests, rate_per_s, steps, iter_max, seed, sizes)
    kind = reqs[0][0]
    kw = _mg_kw(tol)
    by_size = sorted(range(len(reqs)), key=lambda i: (-reqs[i][0].cells, i))
    sample = set(by_size[:accuracy_sample])
    peak = max(batch_case_bytes(p, precision=PF_PREC_MIXED, **kw) for p, _ in reqs)
    log(f"[serve] {n_requests} requests, rate {rate_per_s}/s, {steps} steps each; largest case {peak / 2**20:.0f} MiB")
    free, _total = device_memory()
    with NativeServer(kind, max_cases=min(n_requests, 64), arena_bytes=int(free * 0.3), chunk=chunk,
                      precision=PF_PREC_MIXED, **kw) as srv:
        ids = []
        t0 = time.perf_counter()
        for (p, phi), at in zip(reqs, arrivals):
            delay = at - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            keep = ("u", "v", "w", "p") if len(ids) in sample else ()
            ids.append(srv.submit(p, phi, steps, fields=keep))
        states = [srv.wait(rid) for rid in ids]
        kept = {i: {f: srv.interior(ids[i], f) for f in ("u", "v", "w", "p")}
                for i in sorted(sample) if states[i] == SERVE_DONE}
        wall = time.perf_counter() - t0
        times = [srv.times(rid) for rid in ids]
        polls = [srv.poll(rid) for rid in ids]
        worst = [srv.worst(rid) for rid in ids]
        counters = srv.counters()
        failed = sum(1 for s_ in states if s_ != SERVE_DONE)
    ok = [t for t, s_ in zip(times, states) if s_ == SERVE_DONE]
    tts = np.array([t[2] - t[0] for t in ok])
    wait_s = np.array([t[1] - t[0] for t in ok])
    span = max(t[2] for t in ok) - min(t[0] for t in ok) if ok else 0.0
    upd = sum(p.cells * steps for p, _ in reqs)             # cell-steps
    vcycles = sum(q[2] for q in polls)
    # the FP64 path under the same rule, each room alone
    if plain_sample is None:
        plain_sample = len(reqs)
    plain_upd, plain_time = 0, 0.0
    for p, phi in reqs[:plain_sample]:
        with Solver(p, phi, precision=PF_PREC_FP64, **kw) as s:
            s.step(2)
            s.sync()
            t1 = time.perf_counter()
            s.step(plain_steps)
            s.sync()
            plain_time += time.perf_counter() - t1
            plain_upd += p.cells * plain_steps
    plain = plain_upd / plain_time
    served_busy = upd / max(counters["busy_s"], 1e-9)
    accuracy = []
    for i, fields in kept.items():
        p, phi = reqs[i]
        with Solver(p, phi, precision=PF_PREC_FP64, **kw) as s:
            s.step(steps)
            errs = {f: float(np.linalg.norm(fields[f] - s.interior(f)) / max(np.linalg.norm(s.interior(f)), 1e-300))
                    for f in ("u", "v", "w", "p")}
        accuracy.append({"room": f"{p.m}x{p.n}x{p.l}", **errs})
    n_tts = len(tts)
    return {
        "workload": f"served_rooms_{n_requests}_requests",
        "server": "pf_serve (library worker thread, continuous batching)",
        "rate_per_s": rate_per_s, "steps_per_request": steps, "residual_tol": tol,
        "solver": "multigrid-preconditioned BiCGStab (FP32) in FP64 defect correction, to max|r/ap| <= tol",
        "requests": len(ok), "failed": failed,
        "p50_time_to_solution_s": float(np.percentile(tts, 50)) if n_tts else None,
        "p99_time_to_solution_s": float(np.percentile(tts, 99)) if n_tts else None,
        "max_time_to_solution_s": float(tts.max()) if n_tts else None,
        "latency_samples": n_tts,
        "latency_note": (f"p99 of {n_tts} requests interpolates between the two slowest" if n_tts < 100 else ""),
        "mean_queue_wait_s": float(wait_s.mean()) if len(wait_s) else None,
        "cases_per_s": len(ok) / span if span > 0 else 0.0, "span_s": span, "wall_s": wall,
        "cell_steps_per_s": upd / wall, "busy_s": counters["busy_s"],
        "cell_steps_per_busy_s": served_busy, "fp64_alone_cell_steps_per_s": plain,
        "speedup_vs_fp64_alone_busy": served_busy / plain,
        "fp64_alone_sample": f"{plain_sample} of the rooms, one at a time, FP64 multigrid, same rule, {plain_steps} steps each",
        "worst_residual_any_step": float(max(w["residual"] for w in worst)),
        "vcycles_per_step_mean": vcycles / max(1, n_requests * steps),
        "max_cycles_any_step": int(max(w["cycles"] for w in worst)),
        "sizes": sorted({f"{p.m}x{p.n}x{p.l}" for p, _ in reqs}),
        "field_rel_l2_vs_fp64": accuracy,
    }


def rooms_batch(n_requests: int = 16, steps: int = 4, iter_max: int = 100, tol: float = 1e-9, seed: int = 5,
                rate_per_s: float = 2.0, log: Callable = _noop) -> Dict[str, object]:
    """The served rooms (same generator and seed) resident together in one batch, each step
    solved to max|r/ap| <= tol by the multigrid solve: cell-steps/s of the batch in FP64 and
    in mixed precision, and of each room run alone on the single-case path in both."""
    reqs, _ = _room_requests(n_requests, rate_per_s, steps + 1, iter_max, seed)
    cells = sum(p.cells for p, _ in reqs)
    log(f"[rooms] {n_requests} rooms, {cells} cells, {steps} steps")
    out: Dict[str, object] = {"workload": f"rooms_batch_{n_requests}", "cells": cells, "steps": steps,
                              "residual_tol": tol}
    worst = 0.0
    for label, prec in (("fp64", PF_PREC_FP64), ("mixed", PF_PREC_MIXED)):
        kw = dict(precision=prec, **_mg_kw(tol))
        arena = sum(batch_case_bytes(p, **kw) for p, _ in reqs) + (256 << 20)
        with Batch(reqs[0][0], max_cases=n_requests, arena_bytes=arena, **kw) as b:
            slots = [b.add(p, phi) for p, phi in reqs]
            b.step(1)
            b.sync()
            t0 = time.perf_counter()
            st = b.step(steps)
            dt = time.perf_counter() - t0
        worst = max(worst, float(np.max(st["residual"][:, slots])))
        out[f"batch_{label}_cell_steps_per_s"] = cells * steps / dt
        out[f"batch_{label}_ms_per_step"] = dt * 1e3 / steps
        out[f"batch_{label}_vcycles_per_case_step"] = float(np.mean(st["sor_iters"][:, slots]))
        t_alone = 0.0
        for p, phi in reqs:
            with Solver(p, phi, **kw) as s1:
                s1.step(1)
                s1.sync()
                t0 = time.perf_counter()
                s1.step(steps)
                s1.sync()
                t_alone += time.perf_counter() - t0
        out[f"alone_{label}_cell_steps_per_s"] = cells * steps / t_alone
    out["worst_residual_any_step"] = worst
    out["batch_mixed_vs_alone_fp64"] = out["batch_mixed_cell_steps_per_s"] / out["alone_fp64_cell_steps_per_s"]
    out["batch_mixed_vs_batch_fp64"] = out["batch_mixed_cell_steps_per_s"] / out["batch_fp64_cell_steps_per_s"]
    out["alone_mixed_vs_alone_fp64"] = out["alone_mixed_cell_steps_per_s"] / out["alone_fp64_cell_steps_per_s"]
    return out
