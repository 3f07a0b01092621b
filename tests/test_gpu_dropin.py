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


def _sections(path):
    secs, cur = {}, None
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s:
                continue
            if re.search(r"[a-df-zA-DF-Z_]", s):
                cur = s
                secs[cur] = []
                continue
            secs[cur].extend(s.split())
    return {k: np.array([float(t) for t in v]) for k, v in secs.items()}


@pytest.mark.parametrize("precision,solver", [("mixed", "multigrid"), ("fp64", "multigrid"), ("mixed", "sor")])
def test_dropin_gpu_control_group_matches_c_abi(tmp_path, precision, solver):
    """A controlDict with the optional &gpu_control group runs the drop-in executable in
    that precision and pressure solve; its etc/solution_uvp.dat holds the fields of the same
    settings run through the C-ABI, bit for bit (list-directed REAL(8) output round-trips)."""
    from pixelflow_amd import cases
    from pixelflow_amd import solver as S
    p, phi = cases.sphere_case(m=32, n=16, l=16, steps=4, iter_max=40)
    cases.write_case_dir(str(tmp_path), p, phi)
    tol, inner = (1e-9, 0) if solver == "multigrid" else (1e-7, 20)
    cd = tmp_path / "config" / "controlDict.txt"
    cd.write_text(cd.read_text() + f"&gpu_control\nprecision = '{precision}'\nsolver = '{solver}'\n"
                  f"residual_tol = {tol!r}\ninner_iters = {inner}\nmax_cycles = 40\n/\n")
    out = subprocess.run([os.path.join(BIN, "ibm_3d_uniform")], cwd=tmp_path, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    assert "# --- GPU control" in out.stdout
    sol = _sections(tmp_path / "etc" / "solution_uvp.dat")
    prec = S.PF_PREC_MIXED if precision == "mixed" else S.PF_PREC_FP64
    kw = dict(precision=prec, residual_tol=tol, inner_iters=inner, max_cycles=40,
              solver=S.PF_SOLVER_MG if solver == "multigrid" else S.PF_SOLVER_SOR)
    with S.Solver(p, phi, **kw) as s:
        st = s.step(p.istep_max)
        ref = {f: s.interior(f) for f in ("u", "v", "w", "p")}
    names = {"u": "velocity u_inst", "v": "velocity v_inst", "w": "velocity w_inst", "p": "pressure P_all"}
    for f, key in names.items():
        np.testing.assert_array_equal(sol[key].reshape(ref[f].shape), ref[f], err_msg=f)
    res = [float(x) for x in re.findall(r"-- residual:\s*(\S+)", out.stdout)]
    np.testing.assert_array_equal(np.array(res), st["residual"])


def test_dropin_without_gpu_control_prints_no_gpu_section(tmp_path):
    from pixelflow_amd import cases
    p, phi = cases.sphere_case(m=24, n=12, l=12, steps=2, iter_max=10)
    cases.write_case_dir(str(tmp_path), p, phi)
    out = subprocess.run([os.path.join(BIN, "ibm_3d_uniform")], cwd=tmp_path, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout[-2000:]
    assert "GPU control" not in out.stdout and "-- residual:" not in out.stdout
