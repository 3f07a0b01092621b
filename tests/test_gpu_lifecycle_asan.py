"""The batch and serving lifecycle under host AddressSanitizer.

`tests/native/batch_lifecycle.cpp`, linked with every library source whose host code is
instrumented (`python -m pixelflow_amd.build --asan`), runs the autotuner's sweep over 36
configurations in mixed precision and FP64 (each a drain, a graph drop, a capture and an
instantiation), continuous batching with the step histories regrown, 3D rooms, multigrid
tolerance-mode batches with their side streams, and a server destroyed with requests in
flight.  Any error code fails the run; any heap misuse in host code stops it with an ASan
report.  Leak detection is off: the HIP runtime keeps allocations until exit on purpose.
"""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "pixelflow_amd", "_build", "asan", "batch_lifecycle")


def test_batch_and_serve_lifecycle_under_host_asan():
    if not os.access(DRIVER, os.X_OK):
        pytest.skip("ASan lifecycle driver not built (python -m pixelflow_amd.build --asan)")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1")
    print("[gpu] host-ASan lifecycle driver", flush=True)
    r = subprocess.run([DRIVER], env=env, capture_output=True, text=True, timeout=300)
    print(r.stdout[-4000:], flush=True)
    assert "AddressSanitizer" not in r.stderr, r.stderr[-8000:]
    assert r.returncode == 0, (r.returncode, r.stderr[-4000:])
    assert "[lifecycle] all sequences done" in r.stdout
