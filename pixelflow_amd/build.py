"""Build the gfx950 HIP library (and the Fortran drop-in executables) in-tree.

    python -m pixelflow_amd.build            # libpixelflow.so + Fortran executables
    python -m pixelflow_amd.build --lib-only

Outputs go to pixelflow_amd/_build/ (git-ignored, shipped with the tree).  Flags:
* ``-ffp-contract=off``: no FMA contraction, so every FP64 expression rounds like the
  reference's Fortran (IEEE division stays correctly rounded without fast-math);
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 instructions.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from typing import List

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
OUT = os.path.join(PKG, "_build")
LIB = os.path.join(OUT, "libpixelflow.so")
ARCH = os.environ.get("PF_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["pf_kernels.hip", "pf_mixed.hip", "pf_api.hip", "pf_batch.hip", "pf_serve.cpp", "pf_noalloc.hip",
               "pf_mg.hip"]
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _run(cmd: List[str], cwd: str = None) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_lib(jobs: int = 4) -> str:
    os.makedirs(OUT, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        objs.append(obj)
        if len(procs) >= jobs:
            for p in procs:
                if p.wait() != 0:
                    raise subprocess.CalledProcessError(p.returncode, "hipcc")
            procs = []
    for p in procs:
        if p.wait() != 0:
            raise subprocess.CalledProcessError(p.returncode, "hipcc")
    tmp = LIB + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-pthread", "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_fortran() -> List[str]:
    """Drop-in executables: Fortran hosts over the C-ABI (see pixelflow_amd/fortran)."""
    if not os.path.isdir(FSRC):
        return []
    flang = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
    if not os.path.exists(flang):
        print("flang not found: skipping the Fortran executables", flush=True)
        return []
    bindir = os.path.join(OUT, "bin")
    moddir = os.path.join(OUT, "mod")
    os.makedirs(bindir, exist_ok=True)
    os.makedirs(moddir, exist_ok=True)
    mods = ["pf_binding.f90", "pf_case_io.f90", "pf_driver.f90"]
    mains = sorted(f for f in os.listdir(FSRC) if f.startswith("pf_main_") and f.endswith(".f90"))
    flags = ["-O2", "-ffp-contract=off", "-J", moddir]
    objs = []
    for f in mods:
        if not os.path.exists(os.path.join(FSRC, f)):
            continue
        obj = os.path.join(moddir, os.path.splitext(f)[0] + ".o")
        _run([flang, *flags, "-c", os.path.join(FSRC, f), "-o", obj])
        objs.append(obj)
    exes = []
    rpath = "-Wl,-rpath,$ORIGIN/.."
    for f in mains:
        exe = os.path.join(bindir, os.path.splitext(f)[0][len("pf_main_"):])
        _run([flang, *flags, os.path.join(FSRC, f), *objs, "-L" + OUT, "-lpixelflow", rpath, "-o", exe])
        exes.append(exe)
    return exes


ASAN_DIR = os.path.join(OUT, "asan")
ASAN_DRIVER = os.path.join(ASAN_DIR, "batch_lifecycle")
# host AddressSanitizer only: each -fsanitize goes to the host compilation (the device code
# is built as in the library)
ASAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]


def build_asan_driver(jobs: int = 4) -> str:
    """tests/native/batch_lifecycle.cpp linked with every library source, host code
    instrumented with AddressSanitizer (a standalone executable: the sanitizer runtime is
    linked in, nothing is preloaded)."""
    os.makedirs(ASAN_DIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES] + [os.path.join(ROOT, "tests", "native", "batch_lifecycle.cpp")]
    objs, procs = [], []
    for src in srcs:
        obj = os.path.join(ASAN_DIR, os.path.splitext(os.path.basename(src))[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") and "native" in src else []
        cmd = [hipcc(), *HIPCC_FLAGS, *ASAN_HOST, "-I" + os.path.join(ROOT, "include"), *lang, "-c", src, "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        objs.append(obj)
        if len(procs) >= jobs:
            for p in procs:
                if p.wait() != 0:
                    raise subprocess.CalledProcessError(p.returncode, "hipcc")
            procs = []
    for p in procs:
        if p.wait() != 0:
            raise subprocess.CalledProcessError(p.returncode, "hipcc")
    _run([hipcc(), f"--offload-arch={ARCH}", *ASAN_HOST, *objs, "-pthread", "-o", ASAN_DRIVER])
    return ASAN_DRIVER


SERVE_FAKE = os.path.join(OUT, "test", "libpfserve_fake.so")


def build_serve_fake() -> str:
    """The library's serving loop (csrc/pf_serve.cpp) linked with a stand-in batch
    (tests/native/serve_fake_batch.cpp): a host-only library for CPU tests of the loop's
    policy.  Built with g++; only the loop's device-selection calls need the HIP runtime."""
    os.makedirs(os.path.dirname(SERVE_FAKE), exist_ok=True)
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    gxx = shutil.which("g++") or "g++"
    tmp = SERVE_FAKE + ".tmp"
    _run([gxx, "-std=c++17", "-O1", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__", f"-I{rocm}/include",
          "-I" + os.path.join(ROOT, "include"), os.path.join(CSRC, "pf_serve.cpp"),
          os.path.join(ROOT, "tests", "native", "serve_fake_batch.cpp"), f"-L{rocm}/lib", "-lamdhip64",
          f"-Wl,-rpath,{rocm}/lib", "-pthread", "-o", tmp])
    os.replace(tmp, SERVE_FAKE)
    return SERVE_FAKE


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("--asan", action="store_true",
                    help="also build the host-AddressSanitizer lifecycle driver (tests/native)")
    args = ap.parse_args(argv)
    build_lib()
    if not args.lib_only:
        build_fortran()
        build_serve_fake()
    if args.asan:
        build_asan_driver()
    return 0


if __name__ == "__main__":
    sys.exit(main())
