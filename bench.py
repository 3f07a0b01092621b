#!/usr/bin/env python3
"""Benchmark of PixelFlow's time step on one MI355X.

Workload (default): 3D flow around a sphere, 256x128x128 cells over 2x1x1 m, D = 0.15 m,
2-cell tanh interface, U = 1, xnue = 1e-3, dt = 1e-4, 500 steps, 100 red/black SOR
iterations per step with omega = 1.7, FP64 -- the reference's 3D uniform-inflow solver
(src/omp_parallel/ibm_3d_uniform_omp_cpu.f90) reproduced bit for bit.

One "step" is one time step: predictor, Poisson RHS, 100 SOR iterations (200 colour
sweeps), projection and boundary update, all resident in HBM.  The line printed:

  value       cell-steps/s = m*n*l*steps / wall seconds of the timed steps
  roofline    the SOR colour-sweep kernel: algorithmic bytes per launch (3 FP64 words per
              cell: read p, write p of half the cells, read bb of half, read porosity --
              the streamed model 8*(2d+2 + 6K + 2d+1) B per cell-step) / average launch
              time measured with HIP events on the solver's stream; traffic = measured
              HBM bytes per launch from rocprofv3 counters when a profile is committed
  cpu_baseline the reference's own OpenMP build (oracle/_ref, compiled from its sources)
              on the host cores for a bounded number of steps of the same workload, with
              the CPU restatement (oracle/pf_oracle.c) beside it; the restatement alone
              when the reference build is absent; and `single_thread`, the restatement
              on one thread for one step
Further sections: mixed precision on the same workload; the 512x256x256 boxes solved to
the residual rule max|r/ap| <= 1e-8 every step (multigrid-preconditioned, FP64 and mixed,
time to solution per step, and FP64 with one long inner solve per step) with the HBM
roofline of the FP64 colour sweep at that size (20 B per cell there: the porosity is read
coded in 4 bytes); the
2D cylinder benchmark case (256x128, 2000 steps) on the GPU and on the reference's build;
the batched 2D workload, the served 3D rooms and the same rooms resident together.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))

# The CPUs this process may use, read before anything binds a thread.  The in-process CPU
# baseline binds OpenMP threads (OMP_PROC_BIND), which pins the calling thread to one CPU;
# a child started from it afterwards would inherit that single CPU, so every child is given
# this full set back.
try:
    PROCESS_CPUS = sorted(os.sched_getaffinity(0))
except (AttributeError, OSError):
    PROCESS_CPUS = None


def _full_affinity():
    if PROCESS_CPUS:
        os.sched_setaffinity(0, PROCESS_CPUS)
sys.path.insert(0, ROOT)

from pixelflow_amd import cases  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "cell-updates/s + HBM roofline frac (FP64 vs mixed); batch cases/s; p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel: str, workload: str, variant: str = None):
    """HBM bytes per launch from the committed rocprofv3 counter summary, if present
    (`variant`: a sub-record of the summary, such as the kernel with stored FP64 porosity)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if variant:
        d = d.get(variant, {})
    return d.get("bytes_per_launch")


def omp_env(threads: int):
    """OpenMP settings of the CPU baseline: `threads` threads, one per CPU of the process's
    affinity mask, packed.  Measured on the GPU host's 16 CPUs with the reference's 3D build
    (seconds per step): unbound 0.57, close/threads 0.58, spread/threads 0.55, close on an
    explicit list of the 16 CPUs 0.54, close/cores 0.72."""
    return dict(OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="threads")


def host_cpus():
    """CPUs this process may use (affinity mask, bounded by a cgroup CPU quota if one is
    set), and a description of the host's processors."""
    aff = len(PROCESS_CPUS) if PROCESS_CPUS else (os.cpu_count() or 1)
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) // int(per)))
    except (OSError, ValueError):
        pass
    usable = min(aff, quota) if quota else aff
    model, sockets, cores = None, set(), set()
    try:
        phys = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                    sockets.add(v)
                elif k == "core id":
                    cores.add((phys, v))
    except OSError:
        pass
    return {"usable": usable, "affinity": aff, "cgroup_quota_cpus": quota, "model": model,
            "sockets": len(sockets) or None, "physical_cores": len(cores) or None, "logical": os.cpu_count()}


def cpu_baseline_reference(p, phi, steps: int, threads: int, main: str = "ref_ibm_3d_uniform"):
    """Time the reference's own OpenMP build (oracle/_ref/fp64/<main>, compiled from its
    sources with -fdefault-real-8) on the same case for steps + 1 steps.  Its
    output goes to a pseudo-terminal, so the Fortran runtime writes each log line as it
    happens; the time between the first and the last per-step 'SOR iteration' line is
    `steps` whole steps, without the CSV read before them or the output files after (the
    process is stopped once the last step is logged).  Returns (cell-steps/s, seconds), or
    None when the build is absent or the lines came too close together to time (lines read
    together carry one time stamp)."""
    import dataclasses
    import pty
    import select
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_ref", "fp64", main)
    if not os.access(exe, os.X_OK):
        return None
    env = dict(os.environ, **omp_env(threads))
    q = dataclasses.replace(p, istep_max=steps + 1, time=p.dt * (steps + 1), istep_out=10 ** 6)
    stamps = []
    with tempfile.TemporaryDirectory() as root:
        cases.write_case_dir(root, q, phi)
        master, slave = pty.openpty()
        proc = subprocess.Popen([exe], cwd=root, stdin=subprocess.DEVNULL, stdout=slave, stderr=slave, env=env,
                                preexec_fn=_full_affinity)
        os.close(slave)
        buf = b""
        deadline = time.time() + 600
        try:
            while len(stamps) < steps + 1 and time.time() < deadline:
                ready, _, _ = select.select([master], [], [], 5.0)
                if not ready:
                    if proc.poll() is not None:
                        break
                    continue
                try:
                    chunk = os.read(master, 65536)
                except OSError:
                    break
                if not chunk:
                    break
                now = time.perf_counter()
                buf += chunk
                while b"\n" in buf:
                    line, buf = buf.split(b"\n", 1)
                    if b"SOR iteration" in line:
                        stamps.append(now)
        finally:
            if proc.poll() is None:
                proc.kill()
            proc.wait()
            os.close(master)
    if len(stamps) < steps + 1:
        return None
    dt = stamps[-1] - stamps[0]
    if dt <= 0.:
        return None
    return p.cells * steps / dt, dt


def cpu_baseline(p, phi, steps: int, threads: int):
    """Time the oracle (CPU restatement of the reference) for `steps` steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    os.environ.update(omp_env(threads))
    import pyoracle  # noqa: E402  (test infrastructure: the baseline leg only)
    o = pyoracle.OracleRun(p, phi)
    o.step(1)                      # warm the pages and the thread pool
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    o.close()
    return p.cells * steps / dt, dt


def cpu_baseline_one_thread(m: int, n: int, l: int, steps: int):
    """The oracle on one thread, timed in a child process (the OpenMP runtime of this one
    has already fixed its thread count): `steps` steps of the sphere workload of the given
    size, without a warm-up step.  Returns (cell-steps/s, seconds) or None."""
    import subprocess
    code = ("import sys, json, time; sys.path.insert(0, %r); sys.path.insert(0, %r);"
            "from pixelflow_amd import cases; import pyoracle;"
            "p, phi = cases.sphere_case(m=%d, n=%d, l=%d);"
            "o = pyoracle.OracleRun(p, phi); t0 = time.perf_counter(); o.step(%d);"
            "dt = time.perf_counter() - t0; print(json.dumps([p.cells * %d / dt, dt]))"
            % (ROOT, os.path.join(ROOT, "oracle"), m, n, l, steps, steps))
    env = dict(os.environ, OMP_NUM_THREADS="1")
    try:
        out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300,
                             preexec_fn=_full_affinity)
    except subprocess.TimeoutExpired:
        return None
    if out.returncode != 0:
        return None
    v = json.loads(out.stdout.strip().splitlines()[-1])
    return v[0], v[1]


def mixed_comparison(p, phi, steps: int, warmup: int, inner: int, fp64_value: float):
    """Same workload in mixed precision (fixed-count mode: the reference's iter_max), timed,
    then the fields after `steps` steps compared with an FP64 run of the same steps."""
    from pixelflow_amd.solver import PF_PREC_MIXED, Solver
    log(f"[bench] mixed precision: {steps} steps, inner={inner or p.iter_max}")
    b = Solver(p, phi, precision=PF_PREC_MIXED, inner_iters=inner)
    if warmup:
        b.step(warmup)
    b.set_timing(True)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    wall = time.perf_counter() - t0
    sweep_us, _ = b.last_sweep_time_us()
    unit = b.sor_unit_cells()
    a = Solver(p, phi)
    a.step(warmup + steps)
    errs = {}
    for f in ("u", "v", "w", "p"):
        x, y = a.interior(f), b.interior(f)
        errs[f] = float(np.linalg.norm(y - x) / max(np.linalg.norm(x), 1e-300))
    ddiv = float(np.max(np.abs(a.divergence() - b.divergence())))
    a.close()
    b.close()
    v = p.cells * steps / wall
    # mixed bytes per cell-iteration: 2 FP32 half-sweeps of 3 words (less 2 words per cell of
    # a unit-porosity wave over the two: its own porosity and its neighbours'), per cycle an
    # FP64 residual (p 8, bb 8, porosity 8 read, r 4 + e 4 written) and an update (e 4, p 8 + 8)
    K = inner or p.iter_max
    cyc = (p.iter_max + K - 1) // K
    solve_bytes = p.cells * (p.iter_max * 2 * 3 * 4 + cyc * (8 * 3 + 4 + 4 + 4 + 16)) - p.iter_max * 8 * unit
    return {"value": v, "unit": "cell-steps/s", "ms_per_step": wall * 1e3 / steps, "steps": steps,
            "speedup_vs_fp64": v / fp64_value, "sor_cell_updates_per_s": v * p.iter_max,
            "inner_iters": K, "cycles_per_step": cyc, "avg_solve_us_per_iteration_half": sweep_us,
            "solve_roofline_frac": (solve_bytes / (sweep_us * 2 * p.iter_max * 1e-6) / 1e9) / HBM_PEAK_GBS,
            "unit_porosity_cells": unit,
            "field_rel_l2_vs_fp64": errs, "max_abs_divergence_diff_vs_fp64": ddiv,
            "mode": "fixed-count (iter_max FP32 SOR iterations in FP64 defect-correction cycles)"}


def _timed_run(s, steps: int, warmup: int):
    """Warm up, then time `steps` steps with the colour-sweep events on; returns
    (wall seconds, average colour-sweep microseconds, statistics of the timed steps)."""
    if warmup:
        s.step(warmup)
    s.set_timing(True)
    s.sync()
    t0 = time.perf_counter()
    st = s.step(steps)
    s.sync()
    wall = time.perf_counter() - t0
    sweep_us, _ = s.last_sweep_time_us()
    return wall, sweep_us, st


def tolerance_large(steps: int, warmup: int, tol: float):
    """The 512x256x256 boxes workload (40 seeded boxes, sides 4-32 cells, tanh voxel
    interface, U = 1), whose fields exceed the 256 MB last-level cache, in its stable form:
    cell Reynolds number 1 (xnue = U dx, dt from the scheme's stability limits) and boxes
    clear of the inlet, the outlet region and the periodic faces (as specified, with binary
    porosity and xnue = 5e-4, the explicit step goes non-finite after a few steps).  Every
    step solves the pressure to max|r/ap| <= tol: the multigrid-preconditioned solve in
    FP64 and in mixed precision (FP32 inner solve, FP64 residual and update), timed per
    step (time to solution), their fields compared after warmup + steps.  The fixed-count
    FP64 path (iter_max = 100 SOR iterations, the reference's schedule, which leaves the
    residual far above tol) is timed beside them; its colour sweep gives the HBM roofline
    at a size the cache cannot hold."""
    from pixelflow_amd.solver import PF_PREC_FP64, PF_PREC_MIXED, PF_SOLVER_MG, Solver
    p, phi = cases.boxes_case(steps=warmup + steps, cell_re=1.0, clear_faces=True)
    log(f"[bench] boxes3d_{p.m}x{p.n}x{p.l}: {steps} steps to max|r/ap| <= {tol:g}, FP64 and mixed multigrid")
    mg = dict(residual_tol=tol, solver=PF_SOLVER_MG, max_cycles=30)
    res = {"workload": f"boxes3d_{p.m}x{p.n}x{p.l}", "xnue": p.xnue, "dt": p.dt, "cell_reynolds": p.dx / p.xnue,
           "steps": steps, "warmup": warmup, "residual_tol": tol}
    fields = {}
    for label, prec in (("fp64", PF_PREC_FP64), ("mixed", PF_PREC_MIXED)):
        with Solver(p, phi, precision=prec, **mg) as s:
            wall, _, st = _timed_run(s, steps, warmup)
            fields[label] = {f: s.interior(f) for f in ("u", "v", "w", "p")}
            fields[label]["div"] = s.divergence()
        res[label] = {"ms_per_step": wall * 1e3 / steps, "cell_steps_per_s": p.cells * steps / wall,
                      "max_residual": float(np.max(st["residual"])),
                      "vcycles_per_step": float(np.mean(st["sor_iters"])),
                      "cycles_per_step": float(np.mean(st["cycles"]))}
    res["speedup_vs_fp64"] = res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"]
    a, b = fields["fp64"], fields["mixed"]
    res["field_rel_l2_vs_fp64"] = {f: float(np.linalg.norm(b[f] - a[f]) / max(np.linalg.norm(a[f]), 1e-300))
                                   for f in ("u", "v", "w", "p")}
    res["divergence_max_diff_rel_vs_fp64"] = float(np.max(np.abs(b["div"] - a["div"])) /
                                                   max(np.max(np.abs(a["div"])), 1e-300))
    # FP64 with one long inner solve per step (inner tolerance 1e-8): the inner solve is
    # FP64 already, so the outer cycles only restart it; the best FP64 setting measured
    with Solver(p, phi, precision=PF_PREC_FP64, inner_tol=1e-8, inner_iters=100, **mg) as s:
        wall, _, st = _timed_run(s, steps, warmup)
    res["fp64_single_solve"] = {"ms_per_step": wall * 1e3 / steps, "inner_tol": 1e-8,
                                "max_residual": float(np.max(st["residual"])),
                                "vcycles_per_step": float(np.mean(st["sor_iters"])),
                                "cycles_per_step": float(np.mean(st["cycles"]))}
    res["speedup_vs_fp64_single_solve"] = res["fp64_single_solve"]["ms_per_step"] / res["mixed"]["ms_per_step"]
    # the fixed-count FP64 path twice: porosity read as stored FP64 (the library's default
    # here, where unit-porosity waves hold most cells; the roofline figure: 24 B per grid
    # cell and launch, less 8 B per cell in a unit wave) and coded in 4 bytes (PF_PHI_CODE=1:
    # 20 B, less 4 B per such cell)
    runs = {}
    saved = os.environ.get("PF_PHI_CODE")
    try:
        for code in ("0", "1"):
            os.environ["PF_PHI_CODE"] = code
            with Solver(p, phi) as s:
                wall, sweep_us, _ = _timed_run(s, steps, warmup)
                runs[code] = (wall, sweep_us, s.sor_coded(), s.sor_unit_cells())
    finally:
        if saved is None:
            os.environ.pop("PF_PHI_CODE", None)
        else:
            os.environ["PF_PHI_CODE"] = saved
    wall, sweep_us = runs["0"][0], runs["0"][1]
    res["fixed_count_fp64"] = {"ms_per_step": wall * 1e3 / steps, "cell_steps_per_s": p.cells * steps / wall,
                               "iter_max": p.iter_max, "avg_colour_sweep_us": sweep_us,
                               "porosity": "coded in 4 bytes" if runs["0"][2] else "stored FP64",
                               "unit_porosity_cells": runs["0"][3],
                               "ms_per_step_porosity_coded": runs["1"][0] * 1e3 / steps,
                               "avg_colour_sweep_us_porosity_coded": runs["1"][1]}

    def roofline(sweep_us, bytes_per_cell, unit_bytes, unit, kernel, traffic):
        # bytes_per_cell per grid cell and launch, less unit_bytes per cell of either colour
        # whose waves skip the porosity loads
        bytes_ = bytes_per_cell * p.cells - unit_bytes * unit
        achieved = bytes_ / (sweep_us * 1e-6) / 1e9
        out = {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "bytes_per_launch_algorithmic": bytes_,
               "avg_launch_us": sweep_us, "unit_porosity_cells": unit}
        if traffic:
            # the bytes the counters saw (L2 line fetches and writes) per launch time: the
            # rate the memory system actually delivers, to set beside a plain copy's
            out["traffic_GBps"] = traffic / (sweep_us * 1e-6) / 1e9
        return out

    res["sweep_roofline"] = roofline(runs["0"][1], 24, 8, runs["0"][3], "k_sor_phi (red/black colour sweep, FP64)",
                                     measured_traffic("k_sor_phi", f"boxes3d_{p.m}x{p.n}x{p.l}"))
    if runs["1"][2]:
        res["sweep_roofline"]["coded"] = roofline(
            runs["1"][1], 20, 4, runs["1"][3], "k_sor_phi (red/black colour sweep, FP64, porosity coded in 4 bytes)",
            measured_traffic("k_sor_phi", f"boxes3d_{p.m}x{p.n}x{p.l}", "coded_kernel"))
    return res


def c1_cylinder(threads: int, cpu_steps: int):
    """The 2D benchmark case in full: 256x128 cylinder, 2000 steps, FP64, on the GPU (bit
    for bit the reference) and on the reference's OpenMP build on the host cores."""
    from pixelflow_amd.solver import Solver
    p, phi = cases.cylinder_case(steps=2000)
    with Solver(p, phi) as s:
        s.step(5)
        s.sync()
        t0 = time.perf_counter()
        s.step(p.istep_max)
        s.sync()
        wall = time.perf_counter() - t0
    out = {"workload": f"cylinder2d_{p.m}x{p.n}", "steps": p.istep_max, "iter_max": p.iter_max,
           "gpu": {"cell_steps_per_s": p.cells * p.istep_max / wall, "wall_s": wall,
                   "ms_per_step": wall * 1e3 / p.istep_max}}
    if cpu_steps > 0:
        log(f"[bench] C1 on the reference's OpenMP build: {cpu_steps} steps on {threads} threads")
        ref = cpu_baseline_reference(p, phi, cpu_steps, threads, main="ref_ibm_2d_uniform")
        if ref is not None:
            out["cpu_reference"] = {"cell_steps_per_s": ref[0], "wall_s": ref[1], "steps": cpu_steps,
                                    "threads": threads}
            out["gpu_vs_cpu_reference"] = out["gpu"]["cell_steps_per_s"] / ref[0]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=3, help="steps of the CPU baseline sample (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = every CPU this process may use")
    ap.add_argument("--cpu-steps-1t", type=int, default=1,
                    help="steps of the C restatement timed on one thread (0 = skip)")
    ap.add_argument("--cpu-ref-steps", type=int, default=5,
                    help="steps timed on the reference's own OpenMP build, when built (0 = skip)")
    ap.add_argument("--m", type=int, default=256)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--l", type=int, default=128)
    ap.add_argument("--mixed-steps", type=int, default=100,
                    help="steps of the mixed-precision comparison on the same workload (0 = skip)")
    ap.add_argument("--mixed-inner", type=int, default=0, help="FP32 iterations per defect-correction cycle")
    ap.add_argument("--large-steps", type=int, default=8,
                    help="steps of the FP64/mixed comparison on the 512x256x256 boxes workload (0 = skip)")
    ap.add_argument("--tol", type=float, default=1e-9,
                    help="residual rule max|r/ap| <= tol of the 512x256x256 boxes section (<= 1e-8)")
    ap.add_argument("--serve-tol", type=float, default=1e-10,
                    help="residual rule of the served rooms and the rooms batch (<= 1e-8; tighter than the "
                         "boxes' because a closed room's pressure level is fixed only through small outlets)")
    ap.add_argument("--c1-cpu-steps", type=int, default=2000,
                    help="steps of the 2D benchmark case on the reference's build (0 = skip the CPU leg)")
    ap.add_argument("--c1", type=int, default=1, help="run the 2D benchmark case section (0 = skip)")
    ap.add_argument("--batch-cases", type=int, default=512,
                    help="cases of the batched 2D workload (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--serve-requests", type=int, default=16,
                    help="requests of the served 3D-room workload (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=2.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--rooms-batch", type=int, default=16,
                    help="the served rooms resident together in one batch, against each alone (0 = skip)")
    ap.add_argument("--serve-steps", type=int, default=300)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the plain device-copy calibration kernel and print its GB/s")
    args = ap.parse_args(argv)

    if args.copy_probe:
        from pixelflow_amd.solver import probe_copy_gbps
        out = {"copy_probe_GBps": {f"{mb}MB": probe_copy_gbps(mb << 20, 20) for mb in (64, 128, 1024, 4096)}}
        print(json.dumps(out))
        return 0

    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU; multi-GPU runs are not supported",
                          "n_gpus": args.gpus}))
        return 2

    from pixelflow_amd.solver import Solver  # the HIP path; raises if the library is missing

    p, phi = cases.sphere_case(m=args.m, n=args.n, l=args.l, steps=max(args.steps, 1))
    workload = f"sphere3d_{p.m}x{p.n}x{p.l}"
    log(f"[bench] workload {workload} fp64 iter_max={p.iter_max} steps={args.steps} warmup={args.warmup}")
    s = Solver(p, phi)
    if args.warmup:
        s.step(args.warmup)
    s.set_timing(True)
    s.sync()
    t0 = time.perf_counter()
    st = s.step(args.steps)
    s.sync()
    wall = time.perf_counter() - t0
    sweep_us, launches = s.last_sweep_time_us()
    fused = s.sor_fused()
    unit = s.sor_unit_cells()
    s.close()
    log(f"[bench] {args.steps} steps in {wall:.3f} s; colour sweep avg {sweep_us:.2f} us over {launches} sweeps"
        f" ({'fused: one launch per iteration' if fused else 'one launch per colour sweep'})")
    from pixelflow_amd.solver import probe_copy_gbps
    copy_gbps = {"128MB": probe_copy_gbps(128 << 20, 20), "2GB": probe_copy_gbps(2048 << 20, 10)}

    cells = p.cells
    value = cells * args.steps / wall
    if fused:
        # one launch = one red/black iteration: p read and written once, bb and porosity
        # read once (4 FP64 words per cell)
        kernel, launch_us, launch_bytes = "k_sor_fused (one red/black SOR iteration)", 2 * sweep_us, 4 * 8 * cells
    else:
        # one launch = one colour: 3 FP64 words per grid cell, i.e. 6 per updated cell, 2 of
        # them porosity (the cell's own and its share of its neighbours'); a unit-porosity
        # wave skips those 2, which over a launch pair is 1 word per launch for each cell,
        # of either colour, in such waves
        kernel, launch_us, launch_bytes = "k_sor_phi (red/black colour sweep)", sweep_us, 3 * 8 * cells - 8 * unit
    achieved = launch_bytes / (launch_us * 1e-6) / 1e9 if launch_us > 0 else 0.0
    sor_bytes = 4 * 8 * cells if fused else 2 * launch_bytes   # per SOR iteration
    step_bytes = 8 * ((2 * 3 + 2) + (2 * 3 + 1)) * cells + sor_bytes * p.iter_max
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded sphere porosity, the reference's CSV rounding)",
        "config": {"workload": workload, "grid": [p.m, p.n, p.l], "iter_max": p.iter_max, "omega": p.relux_factor,
                   "dt": p.dt, "xnue": p.xnue, "precision": "fp64"},
        "sor_cell_updates_per_s": value * p.iter_max,
        "step_roofline_frac": (step_bytes * args.steps / wall / 1e9) / HBM_PEAK_GBS,
        "final_sor_error": float(st["sor_error"][-1]),
        "device_copy_GBps": copy_gbps,
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(kernel.split()[0], workload),
            "bytes_per_launch_algorithmic": launch_bytes,
            "avg_launch_us": launch_us,
            "unit_porosity_cells": unit,
        },
    }
    if args.mixed_steps > 0:
        res["mixed"] = mixed_comparison(p, phi, args.mixed_steps, args.warmup, args.mixed_inner, value)
    cpus = host_cpus()
    threads = args.cpu_threads or cpus["usable"]
    if args.large_steps > 0:
        res["tolerance_large"] = tolerance_large(args.large_steps, 2, args.tol)
        res["roofline_hbm"] = res["tolerance_large"]["sweep_roofline"]
        log(f"[bench] boxes 512x256x256 to the rule: mixed {res['tolerance_large']['speedup_vs_fp64']:.2f}x FP64; "
            f"FP64 sweep at {res['roofline_hbm']['frac']:.2f} of HBM peak")
    if args.c1:
        res["c1"] = c1_cylinder(threads, args.c1_cpu_steps)
    if args.batch_cases > 0:
        from pixelflow_amd import workloads
        res["batch"] = workloads.batched_2d(count=args.batch_cases, steps=args.batch_steps, log=log)
        log(f"[bench] batch: {res['batch']['cases_per_s']:.2f} cases/s, "
            f"{res['batch']['speedup_vs_plain_gpu']:.2f}x the plain FP64 path")
    if args.serve_requests > 0:
        from pixelflow_amd import workloads
        res["serve"] = workloads.served_rooms(n_requests=args.serve_requests, rate_per_s=args.serve_rate,
                                              steps=args.serve_steps, tol=args.serve_tol, log=log)
        log(f"[bench] serve: p50 {res['serve']['p50_time_to_solution_s']:.2f} s, "
            f"p99 {res['serve']['p99_time_to_solution_s']:.2f} s")
    if args.rooms_batch > 0:
        from pixelflow_amd import workloads
        res["rooms_batch"] = workloads.rooms_batch(n_requests=args.rooms_batch, tol=args.serve_tol, log=log)
        log(f"[bench] rooms batch: mixed {res['rooms_batch']['batch_mixed_vs_alone_fp64']:.2f}x the rooms alone in FP64")
    if args.cpu_steps > 0:
        log(f"[bench] CPU baseline: oracle, {args.cpu_steps} steps on {threads} threads")
        v, dt = cpu_baseline(p, phi, args.cpu_steps, threads)
        port = {"value": v, "unit": "cell-steps/s", "cores": threads, "kind": "port",
                "sample": f"{args.cpu_steps} steps of the same {workload} workload on the C restatement ({dt:.1f} s)"}
        ref = None
        if args.cpu_ref_steps > 0:
            log(f"[bench] CPU baseline: reference OpenMP build, {args.cpu_ref_steps} steps on {threads} threads")
            ref = cpu_baseline_reference(p, phi, args.cpu_ref_steps, threads)
        if ref is not None:
            res["cpu_baseline"] = {"value": ref[0], "unit": "cell-steps/s", "cores": threads, "kind": "reference",
                                   "host": cpus, "omp": omp_env(threads),
                                   "sample": f"{args.cpu_ref_steps} steps of the same {workload} workload on the "
                                             f"reference's OpenMP build, FP64 ({ref[1]:.1f} s, from its per-step log "
                                             f"lines; CSV read and output files excluded)",
                                   "port": port}
        else:
            res["cpu_baseline"] = dict(port, host=cpus, omp=omp_env(threads))
        if args.cpu_steps_1t > 0:
            log(f"[bench] CPU baseline: oracle, {args.cpu_steps_1t} step(s) on one thread")
            one = cpu_baseline_one_thread(args.m, args.n, args.l, args.cpu_steps_1t)
            if one is not None:
                res["cpu_baseline"]["single_thread"] = {
                    "value": one[0], "unit": "cell-steps/s", "cores": 1, "kind": "port",
                    "sample": f"{args.cpu_steps_1t} step(s) of the same {workload} workload on the C restatement, "
                              f"one thread, in a child process ({one[1]:.1f} s)"}
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
