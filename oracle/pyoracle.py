"""ctypes wrapper of the C restatement in oracle/pf_oracle.c.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by the
cpu_baseline leg of bench.py, always as the reference results the HIP path is compared
with, or as the timed CPU baseline.  The
product (pixelflow_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libpforacle.so")

_FACE_CODES = 6


class _Case(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("m", ctypes.c_int), ("n", ctypes.c_int), ("l", ctypes.c_int),
        ("bc", ctypes.c_int), ("nonslip", ctypes.c_int), ("iter_max", ctypes.c_int),
        ("face", ctypes.c_int * _FACE_CODES),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double),
        ("xnue", ctypes.c_double), ("xlambda", ctypes.c_double), ("density", ctypes.c_double),
        ("thickness", ctypes.c_double), ("inlet_velocity", ctypes.c_double),
        ("outlet_pressure", ctypes.c_double), ("aoa", ctypes.c_double), ("omega", ctypes.c_double),
        ("inlet_porosity", ctypes.c_int),
    ]


_lib: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.pfo_create.restype = ctypes.c_void_p
        L.pfo_create.argtypes = [ctypes.POINTER(_Case)]
        L.pfo_destroy.argtypes = [ctypes.c_void_p]
        L.pfo_setup_porosity.argtypes = [ctypes.POINTER(_Case), dp, ctypes.c_double, dp]
        L.pfo_initial.argtypes = [ctypes.POINTER(_Case), dp, dp, dp, dp, dp]
        L.pfo_step.restype = ctypes.c_int
        L.pfo_step.argtypes = [ctypes.c_void_p, dp, dp, dp, dp, dp, ctypes.c_int, dp]
        L.pfo_divergence.argtypes = [ctypes.POINTER(_Case), dp, dp, dp, dp, dp]
        L.pfo_force.argtypes = [ctypes.POINTER(_Case), dp, dp, dp, dp, dp, dp]
        L.pfo_decode_first_sweep.restype = ctypes.c_int
        L.pfo_decode_first_sweep.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int)]
        _lib = L
    return _lib


def _dp(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_case(p) -> _Case:
    """Build the C struct from a pixelflow_amd.cases.CaseParams."""
    c = _Case()
    c.dim, c.m, c.n, c.l = p.dim, p.m, p.n, (p.l if p.dim == 3 else 1)
    c.bc = 1 if p.bc == "air" else 0
    c.nonslip = 1 if p.nonslip else 0
    c.iter_max = p.iter_max
    for i, f in enumerate(p.faces):
        c.face[i] = int(f)
    c.dx, c.dy, c.dz, c.dt = p.dx, p.dy, p.dz, p.dt
    c.xnue, c.xlambda, c.density, c.thickness = p.xnue, p.xlambda, p.density, p.thickness
    c.inlet_velocity, c.outlet_pressure, c.aoa = p.inlet_velocity, p.outlet_pressure, p.aoa
    c.omega = p.relux_factor
    c.inlet_porosity = 1 if getattr(p, "inlet_porosity", False) else 0
    return c


class OracleRun:
    """Full state of one case on the CPU restatement (fields with ghost layers)."""

    def __init__(self, p, phi_raw_interior: np.ndarray):
        self.p = p
        self.case = make_case(p)
        shape = p.full_shape()
        self.phi = np.zeros(shape, dtype=np.float64)
        raw = np.ascontiguousarray(phi_raw_interior, dtype=np.float64)
        lib().pfo_setup_porosity(ctypes.byref(self.case), _dp(raw), float(p.threshold), _dp(self.phi))
        self.u = np.zeros(shape, dtype=np.float64)
        self.v = np.zeros(shape, dtype=np.float64)
        self.w = np.zeros(shape, dtype=np.float64) if p.dim == 3 else None
        self.p_ = np.zeros(shape, dtype=np.float64)
        lib().pfo_initial(ctypes.byref(self.case), _dp(self.phi), _dp(self.u), _dp(self.v), _dp(self.w), _dp(self.p_))
        self._ctx = lib().pfo_create(ctypes.byref(self.case))
        if not self._ctx:
            raise ValueError("pfo_create rejected the case")

    def step(self, nsteps: int) -> np.ndarray:
        err = np.zeros(max(nsteps, 1), dtype=np.float64)
        rc = lib().pfo_step(self._ctx, _dp(self.phi), _dp(self.u), _dp(self.v), _dp(self.w), _dp(self.p_),
                            int(nsteps), _dp(err))
        if rc != 0:
            raise RuntimeError("pfo_step failed")
        return err[:nsteps]

    def divergence(self) -> np.ndarray:
        out = np.zeros(self.p.interior_shape(), dtype=np.float64)
        lib().pfo_divergence(ctypes.byref(self.case), _dp(self.phi), _dp(self.u), _dp(self.v), _dp(self.w), _dp(out))
        return out

    def force(self) -> np.ndarray:
        """output_force_log_*: [Fp_x, Fp_y, Fp_z, Fv_x, Fv_y, Fv_z] of the current state."""
        out = np.zeros(6, dtype=np.float64)
        lib().pfo_force(ctypes.byref(self.case), _dp(self.phi), _dp(self.u), _dp(self.v), _dp(self.w),
                        _dp(self.p_), _dp(out))
        return out

    def interior(self, name: str) -> np.ndarray:
        a = {"u": self.u, "v": self.v, "w": self.w, "p": self.p_, "phi": self.phi}[name]
        if self.p.dim == 3:
            return a[1:-1, 1:-1, 1:-1].copy()
        return a[1:-1, 1:-1].copy()

    def close(self) -> None:
        if self._ctx:
            lib().pfo_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def decode_first_sweep(dim: int, m: int, n: int, l: int = 1) -> Tuple[int, np.ndarray]:
    size = m * n * (l if dim == 3 else 1)
    buf = (ctypes.c_int * size)()
    rc = lib().pfo_decode_first_sweep(dim, m, n, l, buf)
    arr = np.frombuffer(buf, dtype=np.int32).copy()
    shape = (l, n, m) if dim == 3 else (n, m)
    return rc, arr.reshape(shape)
