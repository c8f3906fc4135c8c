"""The full-size workloads (the 256-cell one-element MEAM set and the
2000-cell, 2.3 M-atom ternary set), checked through properties that hold at
any size: F equals the sum of squared residuals, a batched evaluation equals
the plain one, and the adjoint's directional derivative matches a central
difference along a seeded direction.  The one-element set's pair and triplet
counts are compared with the oracle's restatement of the list build.  Targets
stay at zero (the identities do not depend on them)."""
import numpy as np
import pytest

import pyoracle
from meamzilla_amd import gen
from meamzilla_amd._lib import Objective

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("cfg", [1, 4])
def test_full_size_identities(cfg):
    _say("full size", cfg)
    w = gen.workload(cfg)
    obj = Objective(w.pot, w.db)
    _say("pairs", obj.info.npairs, "triplets", obj.info.ntriplets)
    if cfg == 1:
        assert (obj.info.npairs, obj.info.ntriplets) == pyoracle.Oracle(w.pot, w.db).counts()
    kv = w.pot.knot_vector()
    r = obj.compute()
    ev = np.asarray(r["errvec"])
    assert abs(np.dot(ev, ev) - r["F"]) <= 1e-11 * r["F"]
    Fb = obj.batch(np.stack([kv, kv]))
    assert np.all(np.abs(Fb - r["F"]) <= 1e-12 * r["F"])
    F, g = obj.objective_grad(kv)
    assert abs(F - r["F"]) <= 1e-12 * r["F"]
    slots = np.asarray(w.pot.coeff_slots())
    d = np.zeros_like(kv)
    d[slots] = np.random.default_rng(cfg).normal(size=len(slots))
    d /= np.linalg.norm(d)
    h = 1e-5 * max(np.abs(kv[slots]).max(), 1.0)
    fd = (obj.objective(kv + h * d) - obj.objective(kv - h * d)) / (2 * h)
    ad = float(np.dot(g, d))
    _say("directional", ad, fd)
    # 1e-6, the gradient tolerance (measured: 2e-9 one-element, 8e-10 ternary)
    assert abs(fd - ad) <= 1e-6 * max(abs(ad), 1e-300)
