"""The coefficient walk against the coefficient values the reference printed."""
import numpy as np
import pytest

from conftest import golden_cases, load_case


@pytest.mark.parametrize("case", golden_cases())
def test_coeff_slots_match_reference_coefficients(case):
    pot, db, exp = load_case(case)
    slots = pot.coeff_slots()
    assert len(slots) == exp["ncoeff"] == pot.ncoeff()
    kv = pot.knot_vector()
    assert np.array_equal(kv[slots], np.array(exp["coeffs"]))
