"""Shared test setup: the `gpu` marker, import paths, fixture loading."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def load_golden(name):
    """(CaseParams, dict of arrays) of tests/golden/<name>.npz (no pickles)."""
    from pixelflow_amd import cases
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    params = json.loads(str(d["params"]))
    params["faces"] = tuple(params["faces"])
    p = cases.CaseParams(**params)
    return p, {k: d[k] for k in d.files if k != "params"}


def golden_names(prefix=""):
    return sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz") and f.startswith(prefix))
