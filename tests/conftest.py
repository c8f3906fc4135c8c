import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def golden_cases():
    return sorted(d for d in os.listdir(GOLDEN) if os.path.isfile(os.path.join(GOLDEN, d, "expected.json")))


def load_case(name):
    from meamzilla_amd import formats
    d = os.path.join(GOLDEN, name)
    par = formats.read_param(os.path.join(d, "param"))
    pot = formats.read_potential(os.path.join(d, par["startpot"][0]), embed_extrap=int(par.get("embed_extrap", ["0"])[0]))
    db = formats.read_config(os.path.join(d, par["config"][0]), float(par["energy_weight"][0]),
                             float(par["stress_weight"][0]))
    with open(os.path.join(d, "expected.json")) as fh:
        exp = json.load(fh)
    return pot, db, exp


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    assert a.shape == b.shape, (a.shape, b.shape)
    if a.size == 0:
        return 0.0
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)
