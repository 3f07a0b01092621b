"""The Fortran drop-in executables against the reference executables' own outputs.

tests/golden/out_<name>.tar.gz holds a case directory (controlDict + porosity CSV), every
file the reference executable (built with -fdefault-real-8) wrote there and its standard
output.  The drop-in executable built from pixelflow_amd/fortran runs the same case on the
GPU and must write the same bytes: etc/grid.dat, etc/solution_uvp.dat,
etc/surface_profile.dat, etc/divergent.dat, every VTK snapshot and the final one, and the
same log lines (wall-clock stamps aside).  The drag executable's force lines are sums in a
different order (a device reduction against the reference's serial loop), so they are
compared as numbers, to rounding.
"""
import io
import os
import re
import subprocess
import tarfile

import numpy as np
import pytest

from conftest import ROOT, GOLDEN

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "pixelflow_amd", "_build", "bin")
NAMES = ["2d_uniform", "2d_drag", "2d_backstep", "3d_uniform", "3d_air_condition"]
FORCE = re.compile(r"^\s*(Fp =|Fv =|F  =|Cd =)")


def _load(name):
    with tarfile.open(os.path.join(GOLDEN, f"out_{name}.tar.gz")) as tf:
        files = {m.name: tf.extractfile(m).read() for m in tf.getmembers() if m.isfile()}
    return files


def _numbers(line):
    return np.array([float(t.replace("D", "E")) for t in re.findall(r"[-+]?\d*\.?\d+(?:[EeDd][-+]?\d+)?", line)])


@pytest.mark.parametrize("name", NAMES)
def test_dropin_writes_the_reference_bytes(name, tmp_path):
    exe = os.path.join(BIN, "ibm_" + name)
    assert os.path.exists(exe), f"{exe} not built (python -m pixelflow_amd.build)"
    files = _load(name)
    case = tmp_path / "case"
    for rel, data in files.items():
        if rel.startswith("input/"):
            dst = case / rel[len("input/"):]
            dst.parent.mkdir(parents=True, exist_ok=True)
            dst.write_bytes(data)
    print(f"[gpu] drop-in {name} in {case}", flush=True)
    out = subprocess.run([exe], cwd=case, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]

    expected = {rel[len("output/"):]: data for rel, data in files.items() if rel.startswith("output/")}
    written = {}
    for root, _dirs, fs in os.walk(case):
        for f in fs:
            rel = os.path.relpath(os.path.join(root, f), case)
            if ("input/" + rel) not in files:
                written[rel] = open(os.path.join(root, f), "rb").read()
    assert sorted(written) == sorted(expected)
    for rel, data in expected.items():
        assert written[rel] == data, f"{name}: {rel} differs"

    got = [l for l in out.stdout.splitlines() if "# --- TIME:" not in l]
    ref = files["stdout.txt"].decode().splitlines()
    assert len(got) == len(ref)
    for a, b in zip(got, ref):
        if FORCE.match(b):
            x, y = _numbers(a), _numbers(b)
            assert x.shape == y.shape and np.all(np.abs(x - y) <= 1e-11 * np.abs(y).max() + 1e-300), (a, b)
        else:
            assert a == b
