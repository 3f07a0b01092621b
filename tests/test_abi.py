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
    assert L.pf_abi_version() == 2


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
