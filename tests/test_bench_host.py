"""bench.py pieces that run without a GPU: the one-GPU guard and the CPU baseline legs (the
reference's own OpenMP build, skipped where that build is absent, and the oracle on one
thread)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["n_gpus"] == 2 and "one GPU" in out["error"]


@pytest.mark.skipif(not os.access(os.path.join(ROOT, "oracle", "_ref", "fp64", "ref_ibm_3d_uniform"), os.X_OK),
                    reason="the reference's OpenMP build is not built here (make -C oracle ref)")
def test_cpu_baseline_times_reference_build_per_step():
    """The reference build runs steps + 1 steps on a small sphere; the baseline is the time
    between its first and last per-step log line, so it covers exactly `steps` steps.  The
    steps take tens of milliseconds, so each log line is read on its own."""
    import bench
    from pixelflow_amd import cases
    p, phi = cases.sphere_case(m=64, n=32, l=32, steps=4, iter_max=50)
    res = bench.cpu_baseline_reference(p, phi, 3, 2)
    assert res is not None
    rate, seconds = res
    assert seconds > 0 and rate == pytest.approx(p.cells * 3 / seconds)


def test_cpu_baseline_one_thread_runs_in_a_child_process():
    """The single-thread baseline runs the oracle in a child process with OMP_NUM_THREADS=1
    (the parent's OpenMP runtime has fixed its thread count) and reports the rate of the
    timed steps."""
    import bench
    res = bench.cpu_baseline_one_thread(32, 16, 16, 1)
    assert res is not None
    rate, seconds = res
    assert seconds > 0 and rate == pytest.approx(32 * 16 * 16 / seconds)


@pytest.mark.skipif(not os.access(os.path.join(ROOT, "oracle", "_ref", "fp64", "ref_ibm_2d_uniform"), os.X_OK),
                    reason="the reference's OpenMP build is not built here (make -C oracle ref)")
def test_cpu_baseline_times_reference_2d_build():
    """The 2D benchmark case's CPU leg uses the 2D executable of the reference build."""
    import bench
    from pixelflow_amd import cases
    p, phi = cases.cylinder_case(m=256, n=128, steps=6, iter_max=100)
    res = bench.cpu_baseline_reference(p, phi, 4, 2, main="ref_ibm_2d_uniform")
    assert res is not None and res[1] > 0 and res[0] == pytest.approx(p.cells * 4 / res[1])


def test_host_cpus_reports_usable_cpus():
    import bench
    h = bench.host_cpus()
    assert 1 <= h["usable"] <= h["affinity"]
