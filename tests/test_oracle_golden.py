"""The oracle against the reference's own numbers (golden fixtures written by
tests/golden/make_golden.py from the reference built in this container)."""
import numpy as np
import pytest

from conftest import golden_cases, load_case, rel_err

import pyoracle
from meamzilla_amd import formats


@pytest.mark.parametrize("case", golden_cases())
def test_oracle_matches_reference(case):
    pot, db, exp = load_case(case)
    o = pyoracle.Oracle(pot, db)
    r = o.eval()
    assert abs(r["F"] - exp["F"]) <= 1e-12 * abs(exp["F"])
    assert rel_err(r["errvec"], exp["errvec"]) < 1e-12
    assert rel_err(r["cell_energy"], exp["cell_energy"]) < 1e-12
    assert rel_err(r["cell_stress"], exp["cell_stress"]) < 1e-12
    assert rel_err(r["forces"], exp["forces"]) < 1e-12
    if formats.is_pair_style(pot.style):
        assert exp["densities"] == []  # the reference computes none for pair styles
    else:
        assert rel_err(o.densities(), exp["densities"]) < 1e-12


@pytest.mark.parametrize("case", golden_cases())
def test_oracle_threaded_matches(case):
    pot, db, exp = load_case(case)
    o = pyoracle.Oracle(pot, db)
    assert abs(o.eval_mt(nthreads=3) - exp["F"]) <= 1e-12 * abs(exp["F"])
