"""CPU checks of the C-ABI boundary: the library loads and exports every declared symbol.

No compute call is made (there may be no GPU): only symbol lookup and the pure
descriptor check pf_footprint_bytes, which touches no device.
"""
import os
import re

import pytest

from pixelflow_amd import cases, solver

HEADER = os.path.join(os.path.dirname(__file__), "..", "include", "pixelflow.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_exports_list():
    assert declared_functions() == sorted(solver.EXPORTS)


@pytest.mark.skipif(not os.path.exists(solver.LIB_PATH), reason="HIP library not built")
def test_library_exports_every_declared_symbol():
    L = solver.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    assert L.pf_abi_version() == 3


@pytest.mark.skipif(not os.path.exists(solver.LIB_PATH), reason="HIP library not built")
def test_footprint_and_descriptor_validation():
    p, _ = cases.sphere_case(steps=10)
    fb = solver.footprint_bytes(p)
    # 8 natural FP64 fields + 6 colour-split ones, each >= one value per cell
    assert fb > 14 * 8 * p.cells
    assert fb < 18 * 8 * p.cells
    bad = cases.CaseParams(dim=3, m=2, n=2, l=2)
    assert solver.footprint_bytes(bad) == -1


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(solver, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(solver, "_lib", None)
    with pytest.raises(solver.PixelFlowError):
        solver.lib()


@pytest.mark.skipif(not os.path.exists(solver.LIB_PATH), reason="HIP library not built")
@pytest.mark.parametrize("change,needle", [
    (dict(dim=4), "dim must be 2 or 3"),
    (dict(m=2), "extents too small"),
    (dict(bc=1, dim=2), "air-conditioning solver is 3D"),
    (dict(inlet_porosity=1, dim=3), "inlet_porosity"),
    (dict(precision=7), "precision"),
    (dict(dx=0.0), "dx, dy, dz, dt must be > 0"),
    (dict(iter_max=-1), "iter_max"),
])
def test_invalid_descriptor_is_rejected_before_any_device_call(change, needle):
    """pf_create validates the descriptor first and returns PF_EINVAL with a message
    (pf_last_error(NULL)); nothing aborts and no device is touched."""
    import ctypes
    p, _ = cases.sphere_case(m=16, n=8, l=8, steps=1)
    d = solver.make_desc(p)
    for k, v in change.items():
        if k == "dim" and v == 2:
            d.l = 1
        setattr(d, k, v)
    h = ctypes.c_void_p()
    L = solver.lib()
    assert L.pf_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert needle in L.pf_last_error(None).decode()
    assert L.pf_batch_case_bytes(ctypes.byref(d)) == -1
