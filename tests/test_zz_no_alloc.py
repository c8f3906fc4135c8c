"""mz_objective_noalloc (per-evaluation buffers in the library's static device
array, no allocation at call time) against mz_objective on the same context:
two calls in a row, the second on a smaller database, so the second call runs
on scratch memory that still holds the first call's contents; and a database
too large for the array is rejected with MZ_ERR_CAPACITY."""
import numpy as np
import pytest

from conftest import load_case, rel_err

import pyoracle
from meamzilla_amd import gen
from meamzilla_amd._lib import MeamzError, Objective

pytestmark = pytest.mark.gpu
KEYS = ("errvec", "cell_energy", "cell_stress", "forces")


def _say(*a):
    print(*a, flush=True)


def _same(a, b):
    assert a["F"] == b["F"], (a["F"], b["F"])
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k


def test_noalloc_matches_objective_twice_second_smaller():
    _say("noalloc: 24-cell one-element meam slice, then the meam_spline_c1 fixture")
    w = gen.workload(1, scale=24 / 256)
    gen.add_targets(w, pyoracle.evaluator)
    big = Objective(w.pot, w.db)
    pot, db, exp = load_case("meam_spline_c1")
    small = Objective(pot, db)
    assert small.natoms < big.natoms and small.info.npairs < big.info.npairs

    ref_big = big.compute()
    got_big = big.compute_noalloc()
    _same(got_big, ref_big)
    e = pyoracle.Oracle(w.pot, w.db).eval()
    assert abs(got_big["F"] - e["F"]) <= 1e-10 * abs(e["F"])
    assert rel_err(got_big["forces"], e["forces"]) < 1e-10

    got_small = small.compute_noalloc()
    _same(got_small, small.compute())
    assert abs(got_small["F"] - exp["F"]) <= 1e-10 * abs(exp["F"])
    assert rel_err(got_small["errvec"], exp["errvec"]) < 1e-10

    # outputs off, and a perturbed knot vector on the larger context again
    kv = w.pot.knot_vector() * 1.01
    got = big.compute_noalloc(kv, errvec=False, outputs=False)
    assert got["F"] == big.objective(kv)


def test_noalloc_rejects_database_beyond_scratch():
    _say("noalloc: 60-cell ternary slice exceeds the 64 MiB scratch")
    w = gen.workload(4, scale=60 / 2000)
    obj = Objective(w.pot, w.db)
    assert obj.info.npairs * 24 > 64 << 20
    with pytest.raises(MeamzError, match="scratch"):
        obj.compute_noalloc()
    # the context and the ordinary entry point are unaffected
    assert np.isfinite(obj.objective())
