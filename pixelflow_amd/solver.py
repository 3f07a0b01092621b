"""Python host over the C-ABI (include/pixelflow.h): one case resident on the GPU.

Mirrors the reference's time loop: ``Solver(params, porosity)`` does what program main
does before the loop (read_settings / grid_conditions* / initial_conditions / boundary,
e.g. src/omp_parallel/ibm_3d_uniform_omp_cpu.f90:42-71), ``step(n)`` runs n iterations of
the loop body (:81-132 without the VTK output), ``field()`` / ``divergence()`` give what
the output writers read.  The HIP library is required: there is no CPU fallback, and a
missing library raises at import of the first solver.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional

import numpy as np

from .cases import CaseParams

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "_build", "libpixelflow.so")

PF_BC_UNIFORM, PF_BC_AIR = 0, 1
PF_PREC_FP64, PF_PREC_MIXED = 0, 1
PF_SOLVER_SOR, PF_SOLVER_MG = 0, 1
FIELDS = {"u": 0, "v": 1, "w": 2, "p": 3, "porosity": 4}
PF_EDIVERGED = -5
ERRORS = {-1: "PF_EINVAL", -2: "PF_EHIP", -3: "PF_ENOMEM", -4: "PF_ESTATE", -5: "PF_EDIVERGED"}


class PfDesc(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32), ("m", ctypes.c_int32), ("n", ctypes.c_int32), ("l", ctypes.c_int32),
        ("bc", ctypes.c_int32), ("nonslip", ctypes.c_int32), ("iter_max", ctypes.c_int32),
        ("face", ctypes.c_int32 * 6), ("precision", ctypes.c_int32), ("inner_iters", ctypes.c_int32),
        ("max_cycles", ctypes.c_int32), ("report_residual", ctypes.c_int32), ("inlet_porosity", ctypes.c_int32),
        ("solver", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double),
        ("xnue", ctypes.c_double), ("xlambda", ctypes.c_double), ("density", ctypes.c_double),
        ("thickness", ctypes.c_double), ("inlet_velocity", ctypes.c_double),
        ("outlet_pressure", ctypes.c_double), ("aoa", ctypes.c_double), ("omega", ctypes.c_double),
        ("residual_tol", ctypes.c_double), ("inner_tol", ctypes.c_double),
    ]


class PfStats(ctypes.Structure):
    _fields_ = [("sor_error", ctypes.c_double), ("residual", ctypes.c_double),
                ("sor_iters", ctypes.c_int32), ("cycles", ctypes.c_int32)]


class PfBatchConfig(ctypes.Structure):
    _fields_ = [("persistent_blocks", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("inner_iters", ctypes.c_int32), ("tile_x", ctypes.c_int32)]


class PixelFlowError(RuntimeError):
    pass


_lib: Optional[ctypes.CDLL] = None

# Every symbol include/pixelflow.h declares (checked by tests/test_abi.py).
EXPORTS = ("pf_abi_version", "pf_create", "pf_destroy", "pf_last_error", "pf_set_porosity", "pf_initial",
           "pf_upload", "pf_download", "pf_step", "pf_divergence", "pf_force", "pf_force_noalloc", "pf_footprint_bytes", "pf_sync",
           "pf_set_timing", "pf_last_sweep_time_us", "pf_sor_kernel", "pf_sor_fused", "pf_sor_coded", "pf_sor_unit_cells",
           "pf_probe_copy",
           "pf_batch_create", "pf_batch_destroy", "pf_batch_last_error", "pf_batch_case_bytes", "pf_batch_free_bytes",
           "pf_batch_add", "pf_batch_remove", "pf_batch_count", "pf_batch_step", "pf_batch_download",
           "pf_batch_divergence", "pf_batch_set_config", "pf_batch_get_config", "pf_batch_sync", "pf_batch_case_steps",
           "pf_device_memory", "pf_serve_create", "pf_serve_destroy", "pf_serve_batch", "pf_serve_submit",
           "pf_serve_poll", "pf_serve_worst", "pf_serve_wait", "pf_serve_times", "pf_serve_result", "pf_serve_release",
           "pf_serve_counters", "pf_serve_error")



def _bind_serve(L: ctypes.CDLL) -> None:
    """Signatures of the serving calls (pf_serve_*) and the batch calls they rely on."""
    vp, dp, i64 = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int64
    L.pf_batch_last_error.restype = ctypes.c_char_p
    L.pf_batch_last_error.argtypes = [vp]
    L.pf_serve_create.argtypes = [ctypes.POINTER(PfDesc), ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(vp)]
    L.pf_serve_destroy.argtypes = [vp]
    L.pf_serve_batch.restype = vp
    L.pf_serve_batch.argtypes = [vp]
    L.pf_serve_submit.argtypes = [vp, ctypes.POINTER(PfDesc), dp, ctypes.c_double, ctypes.c_int64, ctypes.c_int32,
                                  ctypes.POINTER(ctypes.c_int64)]
    L.pf_serve_poll.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64),
                                ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(PfStats)]
    L.pf_serve_worst.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(PfStats)]
    L.pf_serve_wait.argtypes = [vp, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]
    L.pf_serve_times.argtypes = [vp, ctypes.c_int64, dp, dp, dp]
    L.pf_serve_result.argtypes = [vp, ctypes.c_int64, ctypes.c_int, dp, i64, i64]
    L.pf_serve_release.argtypes = [vp, ctypes.c_int64]
    L.pf_serve_counters.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), dp, ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_int)]
    L.pf_serve_error.restype = ctypes.c_char_p
    L.pf_serve_error.argtypes = [vp, ctypes.c_int64]

def lib() -> ctypes.CDLL:
    """Load libpixelflow.so (build with ``python -m pixelflow_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PixelFlowError(f"HIP library not built: {LIB_PATH} (run `python -m pixelflow_amd.build`)")
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, i64 = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int64
    L.pf_abi_version.restype = ctypes.c_int
    L.pf_create.argtypes = [ctypes.POINTER(PfDesc), ctypes.POINTER(vp)]
    L.pf_destroy.argtypes = [vp]
    L.pf_last_error.restype = ctypes.c_char_p
    L.pf_last_error.argtypes = [vp]
    L.pf_set_porosity.argtypes = [vp, dp, ctypes.c_double]
    L.pf_initial.argtypes = [vp]
    L.pf_upload.argtypes = [vp, ctypes.c_int, dp, i64, i64]
    L.pf_download.argtypes = [vp, ctypes.c_int, dp, i64, i64]
    L.pf_step.argtypes = [vp, ctypes.c_int, ctypes.POINTER(PfStats)]
    L.pf_divergence.argtypes = [vp, dp]
    L.pf_footprint_bytes.restype = ctypes.c_int64
    L.pf_footprint_bytes.argtypes = [ctypes.POINTER(PfDesc)]
    L.pf_sync.argtypes = [vp]
    L.pf_set_timing.argtypes = [vp, ctypes.c_int]
    L.pf_last_sweep_time_us.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_int64)]
    L.pf_sor_kernel.argtypes = [vp]
    L.pf_sor_fused.argtypes = [vp]
    L.pf_sor_coded.argtypes = [vp]
    L.pf_sor_unit_cells.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.pf_probe_copy.argtypes = [ctypes.c_int64, ctypes.c_int, dp]
    L.pf_batch_create.argtypes = [ctypes.POINTER(PfDesc), ctypes.c_int, ctypes.c_int64, ctypes.POINTER(vp)]
    L.pf_batch_destroy.argtypes = [vp]
    L.pf_batch_last_error.restype = ctypes.c_char_p
    L.pf_batch_last_error.argtypes = [vp]
    L.pf_force.argtypes = [vp, dp]
    L.pf_force_noalloc.argtypes = [vp, dp]
    L.pf_batch_case_bytes.restype = ctypes.c_int64
    L.pf_batch_case_bytes.argtypes = [ctypes.POINTER(PfDesc)]
    L.pf_batch_free_bytes.restype = ctypes.c_int64
    L.pf_batch_free_bytes.argtypes = [vp]
    L.pf_batch_add.argtypes = [vp, ctypes.POINTER(PfDesc), dp, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]
    L.pf_batch_remove.argtypes = [vp, ctypes.c_int]
    L.pf_batch_count.argtypes = [vp]
    L.pf_batch_step.argtypes = [vp, ctypes.c_int, ctypes.POINTER(PfStats)]
    L.pf_batch_download.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, i64, i64]
    L.pf_batch_divergence.argtypes = [vp, ctypes.c_int, dp]
    L.pf_batch_set_config.argtypes = [vp, ctypes.POINTER(PfBatchConfig)]
    L.pf_batch_get_config.argtypes = [vp, ctypes.POINTER(PfBatchConfig)]
    L.pf_batch_sync.argtypes = [vp]
    L.pf_batch_case_steps.restype = ctypes.c_int64
    L.pf_batch_case_steps.argtypes = [vp, ctypes.c_int]
    L.pf_device_memory.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    _bind_serve(L)
    _lib = L
    return L


def make_desc(p: CaseParams, precision: int = PF_PREC_FP64, inner_iters: int = 0, max_cycles: int = 0,
              residual_tol: float = 0.0, report_residual: bool = False, solver: int = PF_SOLVER_SOR,
              inner_tol: float = 0.0) -> PfDesc:
    d = PfDesc()
    d.dim, d.m, d.n, d.l = p.dim, p.m, p.n, (p.l if p.dim == 3 else 1)
    d.bc = PF_BC_AIR if p.bc == "air" else PF_BC_UNIFORM
    d.nonslip = 1 if p.nonslip else 0
    d.iter_max = p.iter_max
    for i, f in enumerate(p.faces):
        d.face[i] = int(f)
    d.precision = precision
    d.inner_iters = inner_iters
    d.max_cycles = max_cycles
    d.report_residual = 1 if report_residual else 0
    d.inlet_porosity = 1 if getattr(p, "inlet_porosity", False) else 0
    d.dx, d.dy, d.dz, d.dt = p.dx, p.dy, (p.dz if p.dim == 3 else 1.0), p.dt
    d.xnue, d.xlambda, d.density, d.thickness = p.xnue, p.xlambda, p.density, p.thickness
    d.inlet_velocity, d.outlet_pressure, d.aoa = p.inlet_velocity, p.outlet_pressure, p.aoa
    d.omega = p.relux_factor
    d.residual_tol = residual_tol
    d.solver = int(solver)
    d.inner_tol = float(inner_tol)
    return d


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# pf_stats as a numpy record (same layout), so per-step statistics of large batches are
# filled in place instead of converted struct by struct.
STATS_DTYPE = np.dtype([("sor_error", "<f8"), ("residual", "<f8"), ("sor_iters", "<i4"), ("cycles", "<i4")])
assert STATS_DTYPE.itemsize == ctypes.sizeof(PfStats)


def _stats_ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(PfStats))


def _stats_dict(a: np.ndarray) -> Dict[str, np.ndarray]:
    return {k: np.ascontiguousarray(a[k]) for k in STATS_DTYPE.names}


class Solver:
    """One case on the GPU (FP64 path: reproduces the reference bit for bit)."""

    def __init__(self, p: CaseParams, phi_raw_interior: np.ndarray, precision: int = PF_PREC_FP64, **kw):
        self.p = p
        self.desc = make_desc(p, precision=precision, **kw)
        L = lib()
        h = ctypes.c_void_p()
        rc = L.pf_create(ctypes.byref(self.desc), ctypes.byref(h))
        if rc != 0:
            raise PixelFlowError(f"pf_create: {ERRORS.get(rc, rc)}: {L.pf_last_error(None).decode()}")
        self._h = h
        raw = np.ascontiguousarray(phi_raw_interior, dtype=np.float64).reshape(-1)
        if raw.size != p.cells:
            raise ValueError(f"porosity has {raw.size} values, the grid {p.cells}")
        self._check(L.pf_set_porosity(self._h, _dp(raw), float(p.threshold)), "pf_set_porosity")
        self._check(L.pf_initial(self._h), "pf_initial")
        self.steps_done = 0

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = lib().pf_last_error(self._h).decode()
            raise PixelFlowError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")

    def step(self, nsteps: int = 1) -> Dict[str, np.ndarray]:
        st = np.zeros(max(nsteps, 1), dtype=STATS_DTYPE)
        self._check(lib().pf_step(self._h, int(nsteps), _stats_ptr(st)), "pf_step")
        self.steps_done += nsteps
        return _stats_dict(st[:nsteps])

    def field(self, name: str) -> np.ndarray:
        """Full field with ghosts, numpy shape (l+2, n+2, m+2) or (n+2, m+2)."""
        out = np.zeros(self.p.full_shape(), dtype=np.float64)
        self._check(lib().pf_download(self._h, FIELDS[name], _dp(out), self.p.m + 2, self.p.n + 2), "pf_download")
        return out

    def set_field(self, name: str, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.p.full_shape())
        self._check(lib().pf_upload(self._h, FIELDS[name], _dp(a), self.p.m + 2, self.p.n + 2), "pf_upload")

    def interior(self, name: str) -> np.ndarray:
        a = self.field(name)
        return a[1:-1, 1:-1, 1:-1].copy() if self.p.dim == 3 else a[1:-1, 1:-1].copy()

    def divergence(self) -> np.ndarray:
        out = np.zeros(self.p.interior_shape(), dtype=np.float64)
        self._check(lib().pf_divergence(self._h, _dp(out)), "pf_divergence")
        return out

    def force(self) -> np.ndarray:
        """[Fp_x, Fp_y, Fp_z, Fv_x, Fv_y, Fv_z] on the body for the current fields."""
        out = np.zeros(6, dtype=np.float64)
        self._check(lib().pf_force(self._h, _dp(out)), "pf_force")
        return out

    def sync(self) -> None:
        self._check(lib().pf_sync(self._h), "pf_sync")

    def set_timing(self, on: bool) -> None:
        self._check(lib().pf_set_timing(self._h, 1 if on else 0), "pf_set_timing")

    def last_sweep_time_us(self):
        us = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(lib().pf_last_sweep_time_us(self._h, ctypes.byref(us), ctypes.byref(n)), "pf_last_sweep_time_us")
        return us.value, n.value

    def sor_kernel(self) -> int:
        return int(lib().pf_sor_kernel(self._h))

    def sor_fused(self) -> bool:
        """True when each SOR iteration runs as one fused launch."""
        return bool(lib().pf_sor_fused(self._h))

    def sor_coded(self) -> bool:
        """True when the FP64 sweeps read the porosity coded in 4 bytes per value."""
        return bool(lib().pf_sor_coded(self._h))

    def sor_unit_cells(self) -> int:
        """Cells of both colours whose colour sweeps skip the porosity loads (their wave's
        cells and neighbours all have porosity 1; FP32 sweeps for mixed precision); 0 when the
        unit-porosity flags are off."""
        n = ctypes.c_int64()
        self._check(lib().pf_sor_unit_cells(self._h, ctypes.byref(n)), "pf_sor_unit_cells")
        return int(n.value)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().pf_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def footprint_bytes(p: CaseParams, precision: int = PF_PREC_FP64) -> int:
    d = make_desc(p, precision=precision)
    return int(lib().pf_footprint_bytes(ctypes.byref(d)))


def probe_copy_gbps(nbytes: int, iters: int = 20) -> float:
    """Device-to-device copy bandwidth (read + write bytes per second, GB/s)."""
    out = ctypes.c_double()
    rc = lib().pf_probe_copy(int(nbytes), int(iters), ctypes.byref(out))
    if rc != 0:
        raise PixelFlowError(f"pf_probe_copy: {ERRORS.get(rc, rc)}")
    return out.value


def device_memory():
    """(free, total) device bytes."""
    f, t = ctypes.c_int64(), ctypes.c_int64()
    rc = lib().pf_device_memory(ctypes.byref(f), ctypes.byref(t))
    if rc != 0:
        raise PixelFlowError(f"pf_device_memory: {ERRORS.get(rc, rc)}")
    return f.value, t.value


def batch_case_bytes(p: CaseParams, precision: int = PF_PREC_FP64, **kw) -> int:
    d = make_desc(p, precision=precision, **kw)
    return int(lib().pf_batch_case_bytes(ctypes.byref(d)))


class Batch:
    """Many cases of one kind on the GPU, advanced together (see pf_batch_* in
    include/pixelflow.h).  Cases are added with their porosity and leave when removed."""

    def __init__(self, kind: CaseParams, max_cases: int, arena_bytes: int, precision: int = PF_PREC_FP64, **kw):
        self.kw = dict(precision=precision, **kw)
        self.kind_desc = make_desc(kind, **self.kw)
        self.max_cases = int(max_cases)
        h = ctypes.c_void_p()
        rc = lib().pf_batch_create(ctypes.byref(self.kind_desc), self.max_cases, int(arena_bytes), ctypes.byref(h))
        if rc != 0:
            raise PixelFlowError(f"pf_batch_create: {ERRORS.get(rc, rc)}: {lib().pf_batch_last_error(None).decode()}")
        self._h = h
        self.params: Dict[int, CaseParams] = {}

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            raise PixelFlowError(f"{what}: {ERRORS.get(rc, rc)}: {lib().pf_batch_last_error(self._h).decode()}")

    def add(self, p: CaseParams, phi_raw_interior: np.ndarray) -> int:
        d = make_desc(p, **self.kw)
        raw = np.ascontiguousarray(phi_raw_interior, dtype=np.float64).reshape(-1)
        if raw.size != p.cells:
            raise ValueError("porosity size does not match the grid")
        slot = ctypes.c_int(-1)
        self._check(lib().pf_batch_add(self._h, ctypes.byref(d), _dp(raw), float(p.threshold), ctypes.byref(slot)),
                    "pf_batch_add")
        self.params[slot.value] = p
        return slot.value

    def remove(self, slot: int) -> None:
        self._check(lib().pf_batch_remove(self._h, int(slot)), "pf_batch_remove")
        self.params.pop(slot, None)

    def count(self) -> int:
        return int(lib().pf_batch_count(self._h))

    def free_bytes(self) -> int:
        return int(lib().pf_batch_free_bytes(self._h))

    def case_bytes(self, p: CaseParams) -> int:
        """Arena bytes a case of this batch's settings takes (pf_batch_case_bytes)."""
        return batch_case_bytes(p, **self.kw)

    def step(self, nsteps: int = 1, raise_diverged: bool = True) -> Dict[str, np.ndarray]:
        """Advance every case `nsteps` steps; per-step statistics shaped (nsteps, max_cases).
        A case whose SOR error became non-finite raises, unless `raise_diverged` is False:
        then the statistics come back (the other cases advanced normally) and the caller
        finds the diverged slots by their non-finite ``sor_error``."""
        st = np.zeros((max(nsteps, 1), self.max_cases), dtype=STATS_DTYPE)
        rc = lib().pf_batch_step(self._h, int(nsteps), _stats_ptr(st))
        if not (rc == PF_EDIVERGED and not raise_diverged):
            self._check(rc, "pf_batch_step")
        return _stats_dict(st[:nsteps])

    def field(self, slot: int, name: str) -> np.ndarray:
        p = self.params[slot]
        out = np.zeros(p.full_shape(), dtype=np.float64)
        self._check(lib().pf_batch_download(self._h, int(slot), FIELDS[name], _dp(out), p.m + 2, p.n + 2),
                    "pf_batch_download")
        return out

    def interior(self, slot: int, name: str) -> np.ndarray:
        a = self.field(slot, name)
        return a[1:-1, 1:-1, 1:-1].copy() if self.params[slot].dim == 3 else a[1:-1, 1:-1].copy()

    def divergence(self, slot: int) -> np.ndarray:
        p = self.params[slot]
        out = np.zeros(p.interior_shape(), dtype=np.float64)
        self._check(lib().pf_batch_divergence(self._h, int(slot), _dp(out)), "pf_batch_divergence")
        return out

    def case_steps(self, slot: int) -> int:
        return int(lib().pf_batch_case_steps(self._h, int(slot)))

    def set_config(self, persistent_blocks: int = 0, use_graph: bool = True, inner_iters: int = 0,
                   tile_x: int = 0) -> None:
        c = PfBatchConfig(int(persistent_blocks), 1 if use_graph else 0, int(inner_iters), int(tile_x))
        self._check(lib().pf_batch_set_config(self._h, ctypes.byref(c)), "pf_batch_set_config")

    def get_config(self) -> Dict[str, int]:
        c = PfBatchConfig()
        self._check(lib().pf_batch_get_config(self._h, ctypes.byref(c)), "pf_batch_get_config")
        return {"persistent_blocks": c.persistent_blocks, "use_graph": c.use_graph, "inner_iters": c.inner_iters,
                "tile_x": c.tile_x}

    def sync(self) -> None:
        self._check(lib().pf_batch_sync(self._h), "pf_batch_sync")

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().pf_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SERVE_QUEUED, SERVE_RUNNING, SERVE_DONE, SERVE_FAILED = 0, 1, 2, 3


SERVE_STAND_IN_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "test", "libpfserve_fake.so")


def serve_stand_in() -> ctypes.CDLL:
    """The library's serving loop (pf_serve.cpp) linked with a stand-in batch, for host
    tests of the loop's policy (tests/native/serve_fake_batch.cpp; built by
    ``python -m pixelflow_amd.build``).  Test infrastructure: no GPU work is done."""
    if not os.path.exists(SERVE_STAND_IN_PATH):
        raise PixelFlowError(f"stand-in serving library not built: {SERVE_STAND_IN_PATH}")
    L = ctypes.CDLL(SERVE_STAND_IN_PATH)
    _bind_serve(L)
    L.fake_reset.argtypes = [ctypes.c_int64]
    L.fake_set_gate.argtypes = [ctypes.c_int]
    L.fake_hold_adds.argtypes = [ctypes.c_int]
    L.fake_grant_steps.argtypes = [ctypes.c_int]
    L.fake_wait_step_calls.argtypes = [ctypes.c_int, ctypes.c_double]
    L.fake_wait_blocked.argtypes = [ctypes.c_double]
    L.fake_diverge.argtypes = [ctypes.c_int, ctypes.c_int64]
    L.fake_log.argtypes = [ctypes.c_char_p, ctypes.c_int]
    return L


class NativeServer:
    """The library's continuous-batching server (pf_serve_*): a worker thread inside the
    library admits submitted cases as slots and arena memory free up and advances them
    in shared step chunks; results are kept per request until released."""

    def __init__(self, kind: CaseParams, max_cases: int, arena_bytes: int = 0, chunk: int = 5,
                 precision: int = PF_PREC_FP64, batch_config: Optional[dict] = None, library: ctypes.CDLL = None, **kw):
        """`library`: the library holding pf_serve_* (default libpixelflow.so; the tests pass
        the serving loop built over a stand-in batch, see serve_stand_in)."""
        self.L = library if library is not None else lib()
        self.kw = dict(precision=precision, **kw)
        d = make_desc(kind, **self.kw)
        h = ctypes.c_void_p()
        rc = self.L.pf_serve_create(ctypes.byref(d), int(max_cases), int(arena_bytes), int(chunk), ctypes.byref(h))
        if rc != 0:
            raise PixelFlowError(f"pf_serve_create: {ERRORS.get(rc, rc)}: {self.L.pf_batch_last_error(None).decode()}")
        self._h = h
        self.params: Dict[int, CaseParams] = {}
        if batch_config:
            c = PfBatchConfig(int(batch_config.get("persistent_blocks", 0)), 1 if batch_config.get("use_graph", True) else 0,
                              int(batch_config.get("inner_iters", 0)), int(batch_config.get("tile_x", 0)))
            rc = self.L.pf_batch_set_config(self.L.pf_serve_batch(self._h), ctypes.byref(c))
            if rc != 0:
                raise PixelFlowError(f"pf_batch_set_config: {ERRORS.get(rc, rc)}")

    def _check(self, rc: int, what: str, rid: int = -1) -> None:
        if rc != 0:
            raise PixelFlowError(f"{what}: {ERRORS.get(rc, rc)}: {self.L.pf_serve_error(self._h, rid).decode()}")

    def case_bytes(self, p: CaseParams) -> int:
        return batch_case_bytes(p, **self.kw)

    def submit(self, p: CaseParams, phi_raw_interior: np.ndarray, nsteps: int, fields=("u", "v", "w", "p")) -> int:
        d = make_desc(p, **self.kw)
        raw = np.ascontiguousarray(phi_raw_interior, dtype=np.float64).reshape(-1)
        if raw.size != p.cells:
            raise ValueError("porosity size does not match the grid")
        mask = 0
        for f in fields:
            if not (f == "w" and p.dim == 2):
                mask |= 1 << FIELDS[f]
        rid = ctypes.c_int64()
        self._check(self.L.pf_serve_submit(self._h, ctypes.byref(d), _dp(raw), float(p.threshold), int(nsteps), mask,
                                          ctypes.byref(rid)), "pf_serve_submit")
        self.params[rid.value] = p
        return rid.value

    def poll(self, rid: int):
        """(state, steps done, SOR iterations over them, statistics of the latest step)."""
        st, done, its, last = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64(), PfStats()
        self._check(self.L.pf_serve_poll(self._h, int(rid), ctypes.byref(st), ctypes.byref(done), ctypes.byref(its),
                                        ctypes.byref(last)), "pf_serve_poll", rid)
        return st.value, done.value, its.value, {"sor_error": last.sor_error, "residual": last.residual,
                                                 "sor_iters": last.sor_iters, "cycles": last.cycles}

    def worst(self, rid: int) -> Dict[str, float]:
        """Per field, the maximum of the request's per-step statistics so far."""
        w = PfStats()
        self._check(self.L.pf_serve_worst(self._h, int(rid), ctypes.byref(w)), "pf_serve_worst", rid)
        return {"sor_error": w.sor_error, "residual": w.residual, "sor_iters": w.sor_iters, "cycles": w.cycles}

    def wait(self, rid: int, timeout: float = -1.0) -> int:
        st = ctypes.c_int()
        self._check(self.L.pf_serve_wait(self._h, int(rid), float(timeout), ctypes.byref(st)), "pf_serve_wait", rid)
        return st.value

    def times(self, rid: int):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self.L.pf_serve_times(self._h, int(rid), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
                    "pf_serve_times", rid)
        return a.value, b.value, c.value

    def result(self, rid: int, name: str) -> np.ndarray:
        p = self.params[rid]
        out = np.zeros(p.full_shape(), dtype=np.float64)
        self._check(self.L.pf_serve_result(self._h, int(rid), FIELDS[name], _dp(out), p.m + 2, p.n + 2),
                    "pf_serve_result", rid)
        return out

    def interior(self, rid: int, name: str) -> np.ndarray:
        a = self.result(rid, name)
        return a[1:-1, 1:-1, 1:-1].copy() if self.params[rid].dim == 3 else a[1:-1, 1:-1].copy()

    def error(self, rid: int) -> str:
        return self.L.pf_serve_error(self._h, int(rid)).decode()

    def release(self, rid: int) -> None:
        self._check(self.L.pf_serve_release(self._h, int(rid)), "pf_serve_release", rid)
        self.params.pop(rid, None)

    def counters(self) -> Dict[str, float]:
        steps, busy, q, r = ctypes.c_int64(), ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        self._check(self.L.pf_serve_counters(self._h, ctypes.byref(steps), ctypes.byref(busy), ctypes.byref(q),
                                            ctypes.byref(r)), "pf_serve_counters")
        return {"steps_run": steps.value, "busy_s": busy.value, "queued": q.value, "running": r.value}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.L.pf_serve_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
