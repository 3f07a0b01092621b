#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ from the reference build itself.

Runs only where the reference tree and the executables built from it by
``make -C oracle ref`` (oracle/_ref/{fp64,fp32}/ref_*) exist.  For every case it writes a
case directory (controlDict + porosity CSV, in the reference's formats) into a scratch
folder, runs the reference executable there, and stores in ``<name>.npz``:

* ``phi_raw``  -- the porosity values the CSV holds (interior, numpy (l,n,m) / (n,m));
* ``u v w p``  -- final interior fields parsed from ``etc/solution_uvp.dat``
                  (flang's list-directed REAL(8) output round-trips exactly);
* ``div``      -- ``etc/divergent.dat`` (output_divergent_*), interior;
* ``errs``     -- the per-step SOR diagnostic printed as ``p error:``;
* ``params``   -- JSON of the CaseParams.
Large cases store SHA-256 digests of the float64 fields plus a few planes instead.

Usage:  python tests/golden/gen_golden.py [--only NAME]
"""
from __future__ import annotations

import argparse
import dataclasses
import hashlib
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import zipfile

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
from pixelflow_amd import cases  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
REF_BIN = os.path.join(ROOT, "oracle", "_ref")
REFERENCE = os.environ.get("PF_REFERENCE", "/root/reference")

MAINS = {("uniform", 2): "ref_ibm_2d_uniform", ("uniform", 3): "ref_ibm_3d_uniform",
         ("air", 3): "ref_ibm_3d_air_condition"}


def parse_sections(path: str, dtype=np.float64):
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
    return {k: np.array([float(t) for t in v], dtype=dtype) for k, v in secs.items()}


def run_reference(p: cases.CaseParams, phi: np.ndarray, flavour: str = "fp64", csv_text_path: str = None,
                  main: str = None, threads: str = None):
    if main is None:
        main = "ref_ibm_2d_backstep" if p.inlet_porosity else MAINS[(p.bc, p.dim)]
    exe = os.path.join(REF_BIN, flavour, main)
    if not os.path.exists(exe):
        raise SystemExit(f"missing {exe}: run `make -C oracle ref` first")
    work = tempfile.mkdtemp(prefix="pf_golden_")
    try:
        cases.write_control_dict(os.path.join(work, "config", "controlDict.txt"), p)
        if csv_text_path:
            os.makedirs(os.path.join(work, os.path.dirname(p.csv_file)), exist_ok=True)
            shutil.copy(csv_text_path, os.path.join(work, p.csv_file))
        else:
            cases.write_porosity_csv(os.path.join(work, p.csv_file), phi)
        env = dict(os.environ, OMP_NUM_THREADS=threads or os.environ.get("OMP_NUM_THREADS", "8"))
        out = subprocess.run([exe], cwd=work, capture_output=True, text=True, env=env, check=True)
        errs = np.array([float(x) for x in re.findall(r"p error:\s*(\S+)", out.stdout)])
        # output_force_log_2d lines (drag executable): "Fp = fx fy", "Fv = fx fy"
        fp = re.findall(r"Fp =\s*(\S+)\s+(\S+)", out.stdout)
        fv = re.findall(r"Fv =\s*(\S+)\s+(\S+)", out.stdout)
        dt = np.float32 if flavour == "fp32" else np.float64
        sol = parse_sections(os.path.join(work, "etc", "solution_uvp.dat"), dt)
        dv = parse_sections(os.path.join(work, "etc", "divergent.dat"), dt)
        raw = cases.read_porosity_csv(os.path.join(work, p.csv_file))
        sh = p.interior_shape()
        res = {"phi_raw": raw.reshape(sh), "errs": errs.astype(dt),
               "u": sol["velocity u_inst"].reshape(sh), "v": sol["velocity v_inst"].reshape(sh),
               "p": sol["pressure P_all"].reshape(sh), "div": dv["divergent velocity"].reshape(sh)}
        if p.dim == 3:
            res["w"] = sol["velocity w_inst"].reshape(sh)
        if fp:
            res["force_p"] = np.array([[float(a), float(b)] for a, b in fp])
            res["force_v"] = np.array([[float(a), float(b)] for a, b in fv])
        return res, work
    except Exception:
        shutil.rmtree(work, ignore_errors=True)
        raise


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def params_json(p: cases.CaseParams) -> np.ndarray:
    return np.array(json.dumps(dataclasses.asdict(p)))


def save_full(name: str, p: cases.CaseParams, res: dict, extra=None):
    arrs = {k: v for k, v in res.items()}
    arrs["params"] = params_json(p)
    if extra:
        arrs.update(extra)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrs)


def save_hashed(name: str, p: cases.CaseParams, res: dict, planes, keep_phi: bool = True,
                plane_fields=("u", "v", "w", "p", "div")):
    arrs = {"params": params_json(p), "errs": res["errs"]}
    if keep_phi:   # porosity that no generator of this repository rebuilds (a CSV of the reference)
        arrs["phi_raw"] = res["phi_raw"]
    for f in ("u", "v", "w", "p", "div"):
        if f in res:
            arrs["sha_" + f] = np.array(sha(res[f]))
            if f not in plane_fields:
                continue
            for ax, idx in planes:
                arrs[f"{f}_plane{ax}{idx}"] = np.take(res[f], idx, axis=ax)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrs)


def stable_room(m, n, l, seed, steps, iter_max):
    return cases.room_case(seed, steps=steps, iter_max=iter_max, size=(m, n, l))


def case_list():
    C = cases.CaseParams
    L = []
    # G1: 2D 48x24 cylinder, 20 steps, the full 100 SOR iterations.
    p = C(dim=2, m=48, n=24, width=0.47, height=0.23, time=20 * 1e-3, istep_max=20, xnue=1e-3, radius=0.04)
    L.append(("g2d_cyl48", p, cases.cylinder_2d(48, 24, 0.47, 0.23, (0.15, 0.115), 0.08)))
    # Stage isolation on 64x32: iter_max=0 (predictor + RHS + projection + boundary), iter_max=1.
    for it, steps in ((0, 2), (1, 1), (1, 3)):
        p = C(dim=2, m=64, n=32, width=1.0, height=0.49, time=steps * 2e-3, istep_max=steps, iter_max=it)
        L.append((f"g2d_stage_i{it}_s{steps}", p, cases.cylinder_2d(64, 32, 1.0, 0.49, (0.3, 0.24), 0.12)))
    # Odd extents and every optional term switched on (xlambda, AoA, p_out, density, slip).
    p = C(dim=2, m=31, n=17, width=0.6, height=0.32, time=6 * 2e-3, istep_max=6, iter_max=20, xlambda=3e-4,
          aoa=15.0, outlet_pressure=0.25, density=1.2, xnue=4e-3, thickness=2.0, nonslip=False)
    L.append(("g2d_weird_odd", p, cases.cylinder_2d(31, 17, 0.6, 0.32, (0.2, 0.16), 0.1)))
    # The 2D benchmark geometry itself (256x128 cylinder D=0.1 at (0.5,0.5) over 2x1 m), 10 steps.
    p, phi = cases.cylinder_case(steps=10)
    L.append(("g2d_c1_10", p, phi))
    # G3: 3D sphere 32x16x16, 10 steps, 100 SOR iterations.
    p = C(dim=3, m=32, n=16, l=16, width=1.0, height=0.48, depth=0.48, time=10 * 5e-4, istep_max=10, radius=0.06)
    L.append(("g3d_sph32", p, cases.sphere_3d(32, 16, 16, 1.0, 0.48, 0.48, (0.3, 0.24, 0.24), 0.12)))
    # Odd extents: both branches of the colour decode (m odd & n even; m, n odd) and odd wraps.
    for (m, n, l) in ((33, 17, 15), (33, 16, 15), (32, 17, 15)):
        p = C(dim=3, m=m, n=n, l=l, width=1.0, height=0.5, depth=0.45, time=5 * 5e-4, istep_max=5, iter_max=30)
        L.append((f"g3d_odd_{m}_{n}_{l}", p, cases.sphere_3d(m, n, l, 1.0, 0.5, 0.45, (0.3, 0.25, 0.22), 0.13)))
    # Stage isolation 3D.
    for it, steps in ((0, 2), (1, 1)):
        p = C(dim=3, m=24, n=12, l=12, width=1.0, height=0.48, depth=0.48, time=steps * 1e-3, istep_max=steps, iter_max=it)
        L.append((f"g3d_stage_i{it}_s{steps}", p, cases.sphere_3d(24, 12, 12, 1.0, 0.48, 0.48, (0.3, 0.24, 0.24), 0.14)))
    # Every optional term in 3D (AoA uses the /1300 and /360 factors of the 3D code).
    p = C(dim=3, m=20, n=14, l=12, width=1.0, height=0.7, depth=0.6, time=5 * 1e-3, istep_max=5, iter_max=25,
          xlambda=3e-4, aoa=20.0, outlet_pressure=0.3, density=1.3, xnue=0.01, thickness=2.0)
    L.append(("g3d_weird", p, cases.sphere_3d(20, 14, 12, 1.0, 0.7, 0.6, (0.4, 0.3, 0.3), 0.25)))
    # Binary porosity (threshold clamp, 1e-6 faces) as in the boxes workload.
    p = C(dim=3, m=32, n=16, l=16, width=1.0, height=0.48, depth=0.48, time=4 * 5e-4, istep_max=4, iter_max=40, xnue=5e-4)
    L.append(("g3d_boxes", p, cases.random_boxes_3d(32, 16, 16, 6, (3, 8), seed=5)))
    # Backward-facing step (ibm_2d_backstep): inlet and initial velocity times porosity; a
    # channel wall band with a step at the inlet, tanh interface.
    p = C(dim=2, m=48, n=24, width=0.94, height=0.46, time=8 * 1e-3, istep_max=8, iter_max=40, xnue=4e-2,
          inlet_porosity=True)
    L.append(("g2d_backstep", p, cases.backstep_2d(48, 24, 12, 4, 4)))
    # Air-conditioning solver on a synthetic room (stable dt), reference face codes.
    p, phi = stable_room(24, 22, 20, seed=3, steps=5, iter_max=40)
    L.append(("g3d_air_room24", p, phi))
    return L


def reference_room_case():
    """The reference's own air-conditioning project (test/room.zip): room.csv 64^3 with its
    controlDict physics, run for 3 steps at the controlDict's dt (time/istep_max)."""
    z = os.path.join(REFERENCE, "test", "room.zip")
    tmp = tempfile.mkdtemp(prefix="pf_room_")
    with zipfile.ZipFile(z) as zf:
        zf.extract("room/data/room.csv", tmp)
    csv = os.path.join(tmp, "room", "data", "room.csv")
    dt = 1.0 / 2000.0
    p = cases.CaseParams(dim=3, m=64, n=64, l=64, bc="air", xnue=0.025, xlambda=0.0, density=1.0,
                         width=0.63, height=0.63, depth=0.63, time=3 * dt, istep_max=3, inlet_velocity=1.5,
                         outlet_pressure=0.0, aoa=0.0, thickness=1.5, threshold=1.0e-6, radius=0.63,
                         nonslip=True, csv_file="data/room.csv", iter_max=100, relux_factor=1.7)
    return p, csv, tmp


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default=None)
    args = ap.parse_args()
    for name, p, phi in case_list():
        if args.only and args.only != name:
            continue
        res, work = run_reference(p, phi)
        extra = {}
        if name == "g2d_cyl48":
            # FP32 (as-shipped) build of the same case: a sanity reference for reduced precision.
            r32, w32 = run_reference(p, phi, "fp32")
            extra = {k + "_fp32": r32[k] for k in ("u", "v", "p", "errs")}
            shutil.rmtree(w32, ignore_errors=True)
        save_full(name, p, res, extra)
        shutil.rmtree(work, ignore_errors=True)
        print("wrote", name, flush=True)
    if not args.only or args.only == "g2d_drag":
        # Drag executable (ibm_2d_drag): the uniform step plus output_force_log_2d every step,
        # run on one thread so its reduction is the serial loop order.
        C = cases.CaseParams
        p = C(dim=2, m=48, n=24, width=0.47, height=0.23, time=10 * 1e-3, istep_max=10, xnue=1e-3, radius=0.04,
              iter_max=50)
        phi = cases.cylinder_2d(48, 24, 0.47, 0.23, (0.15, 0.115), 0.08)
        res, work = run_reference(p, phi, main="ref_ibm_2d_drag", threads="1")
        save_full("g2d_drag", p, res)
        shutil.rmtree(work, ignore_errors=True)
        print("wrote g2d_drag", flush=True)
    if args.only == "g2d_c1_2000":
        # The 2D benchmark case over its full 2000 steps (256x128, 100 SOR iterations per
        # step): final fields in full, so a run can be held to a relative L2 bar as well as
        # bit for bit.  Long run: only on request.
        p, phi = cases.cylinder_case(steps=2000)
        res, work = run_reference(p, phi)
        save_full("g2d_c1_2000", p, {k: res[k] for k in ("u", "v", "p", "errs")})
        shutil.rmtree(work, ignore_errors=True)
        print("wrote g2d_c1_2000", flush=True)
    if args.only == "g3d_c2_500":
        # The 3D sphere benchmark case over its full 500 steps (256x128x128): digests of the
        # final fields plus the mid z plane of u and p.  About half an hour on 8 cores; only on request.
        p, phi = cases.sphere_case(steps=500)
        res, work = run_reference(p, phi)
        save_hashed("g3d_c2_500", p, res, planes=[(0, 64)], keep_phi=False, plane_fields=("u", "p"))
        shutil.rmtree(work, ignore_errors=True)
        print("wrote g3d_c2_500", flush=True)
    if not args.only or args.only == "g3d_ref_room64":
        p, csv, tmp = reference_room_case()
        res, work = run_reference(p, None, csv_text_path=csv)
        save_hashed("g3d_ref_room64", p, res, planes=[(0, 31), (1, 31), (2, 31)])
        shutil.rmtree(work, ignore_errors=True)
        shutil.rmtree(tmp, ignore_errors=True)
        print("wrote g3d_ref_room64", flush=True)


if __name__ == "__main__":
    main()
