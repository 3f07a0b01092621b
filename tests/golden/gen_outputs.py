#!/usr/bin/env python3
"""Write the drop-in output fixtures (tests/golden/out_<name>.tar.gz) from the reference build.

For each small case this makes a case directory (config/controlDict.txt and the porosity
CSV, the reference's input formats), runs the reference executable built by
``make -C oracle ref`` (oracle/_ref/fp64/ref_*, i.e. -fdefault-real-8) in it, and stores

* ``input/``   -- the case directory as given to the executable;
* ``output/``  -- every file the run wrote (etc/*.dat, <output_folder>/*.vtk);
* ``stdout.txt`` -- its standard output without the wall-clock "# --- TIME:" lines.

tests/test_gpu_dropin.py runs the Fortran drop-in executables on the same inputs and
compares the bytes.  Only where the reference tree and its build exist.

Usage:  python tests/golden/gen_outputs.py [--only NAME]
"""
from __future__ import annotations

import argparse
import io
import os
import shutil
import subprocess
import sys
import tarfile
import tempfile

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
from pixelflow_amd import cases  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "fp64")


def case_list():
    C = cases.CaseParams
    L = []
    p = C(dim=2, m=40, n=20, width=0.39, height=0.19, time=12 * 1e-3, istep_max=12, istep_out=6, iter_max=30,
          xnue=1e-3, radius=0.03, output_folder="cyl")
    L.append(("2d_uniform", "ibm_2d_uniform", p, cases.cylinder_2d(40, 20, 0.39, 0.19, (0.12, 0.095), 0.06), "8"))
    p = C(dim=2, m=40, n=20, width=0.39, height=0.19, time=6 * 1e-3, istep_max=6, istep_out=3, iter_max=30,
          xnue=1e-3, radius=0.03)
    L.append(("2d_drag", "ibm_2d_drag", p, cases.cylinder_2d(40, 20, 0.39, 0.19, (0.12, 0.095), 0.06), "1"))
    p = C(dim=2, m=48, n=24, width=0.94, height=0.46, time=8 * 1e-3, istep_max=8, istep_out=4, iter_max=40,
          xnue=4e-2, inlet_porosity=True)
    L.append(("2d_backstep", "ibm_2d_backstep", p, cases.backstep_2d(48, 24, 12, 4, 4), "8"))
    p = C(dim=3, m=16, n=10, l=8, width=1.0, height=0.6, depth=0.467, time=6 * 1e-3, istep_max=6, istep_out=3,
          iter_max=25, xnue=0.01, radius=0.1)
    L.append(("3d_uniform", "ibm_3d_uniform", p, cases.sphere_3d(16, 10, 8, 1.0, 0.6, 0.467, (0.4, 0.3, 0.23), 0.2),
              "8"))
    p, phi = cases.room_case(3, steps=4, iter_max=30, size=(24, 22, 20))
    import dataclasses
    p = dataclasses.replace(p, istep_out=4)
    L.append(("3d_air_condition", "ibm_3d_air_condition", p, phi, "8"))
    return L


def make_input(work: str, p, phi) -> None:
    cases.write_control_dict(os.path.join(work, "config", "controlDict.txt"), p)
    cases.write_porosity_csv(os.path.join(work, p.csv_file), phi)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default=None)
    args = ap.parse_args()
    for name, exe, p, phi, threads in case_list():
        if args.only and args.only != name:
            continue
        work = tempfile.mkdtemp(prefix="pf_out_")
        try:
            inp = os.path.join(work, "input")
            run = os.path.join(work, "run")
            make_input(inp, p, phi)
            shutil.copytree(inp, run)
            env = dict(os.environ, OMP_NUM_THREADS=threads)
            out = subprocess.run([os.path.join(REF_BIN, "ref_" + exe)], cwd=run, capture_output=True, text=True,
                                 env=env, check=True)
            stdout = "".join(line + "\n" for line in out.stdout.splitlines() if "# --- TIME:" not in line)
            dst = os.path.join(HERE, f"out_{name}.tar.gz")
            with tarfile.open(dst, "w:gz") as tf:
                tf.add(inp, arcname="input")
                for root, _dirs, files in os.walk(run):
                    for f in sorted(files):
                        full = os.path.join(root, f)
                        rel = os.path.relpath(full, run)
                        if os.path.exists(os.path.join(inp, rel)):
                            continue
                        tf.add(full, arcname=os.path.join("output", rel))
                data = stdout.encode()
                info = tarfile.TarInfo("stdout.txt")
                info.size = len(data)
                tf.addfile(info, io.BytesIO(data))
            print("wrote", dst, os.path.getsize(dst), flush=True)
        finally:
            shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    main()
