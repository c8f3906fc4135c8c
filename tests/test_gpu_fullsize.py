"""The workloads at full size.  The 16-cell EAM set, the 256-cell
one-element MEAM set, the 512-cell binary GA set and the 1000-cell binary
adjoint set against the oracle in full (its threaded restatement runs on the
host: F, residual vector, per-cell energies and stresses, forces at 1e-10;
pair and triplet counts exact); the 2000-cell ternary set on a 40-cell slice
against the oracle and in full through properties that hold at any size (F
equals the sum of squared residuals, batched equals plain, the adjoint's
directional derivatives along seeded directions match central differences);
and the GA set's 256-candidate population in one batched call against 256
plain evaluations."""
import numpy as np
import pytest

from conftest import rel_err

import pyoracle
from meamzilla_amd import gen
from meamzilla_amd._lib import Objective

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _say(*a):
    print(*a, flush=True)
NTHREADS = 16  # the host-CPU share of one GPU


def _mt_evaluator(pot, db):
    r = pyoracle.Oracle(pot, db).eval_full_mt(nthreads=NTHREADS)
    return r["cell_energy"], r["cell_stress"], r["forces"]


def _check_against_oracle(w, obj, o, kv):
    e = o.eval_full_mt(kv, nthreads=NTHREADS)
    r = obj.compute(kv)
    assert abs(r["F"] - e["F"]) <= TOL * abs(e["F"]), (r["F"], e["F"])
    for k in ("errvec", "cell_energy", "cell_stress", "forces"):
        assert rel_err(r[k], e[k]) < TOL, k


@pytest.mark.parametrize("cfg,scale", [(0, 1.0), (1, 1.0), (2, 1.0), (3, 1.0), (4, 40 / 2000)])
def test_full_workload_matches_oracle(cfg, scale):
    _say("full workload vs oracle", cfg, scale)
    w = gen.workload(cfg, scale=scale)
    gen.add_targets(w, _mt_evaluator)
    o = pyoracle.Oracle(w.pot, w.db)
    obj = Objective(w.pot, w.db)
    _say("cells", w.db.ncells, "atoms", w.db.total_atoms, "pairs", obj.info.npairs, "triplets", obj.info.ntriplets)
    assert (obj.info.npairs, obj.info.ntriplets) == o.counts()
    kv = w.pot.knot_vector()
    _check_against_oracle(w, obj, o, kv)
    # a second knot vector, as an optimizer step would give
    rng = np.random.default_rng(100 + cfg)
    _check_against_oracle(w, obj, o, kv * (1.0 + 0.02 * rng.standard_normal(kv.shape)))


def test_ga_population_256_batched_equals_plain():
    """The binary GA set's defining case: 256 candidate knot vectors (seeded
    base x (1 + U(+-5 %))) scored in one batched call, against 256 plain
    evaluations and, for three of them, the oracle."""
    _say("GA population 256")
    w = gen.workload(2)
    gen.add_targets(w, _mt_evaluator)
    obj = Objective(w.pot, w.db)
    K = w.candidate_knots(256)
    assert K.shape[0] == 256
    Fb = obj.batch(K)
    Fp = np.array([obj.objective(k) for k in K])
    assert np.max(np.abs(Fb - Fp) / np.abs(Fp)) <= 1e-12, np.max(np.abs(Fb - Fp) / np.abs(Fp))
    o = pyoracle.Oracle(w.pot, w.db)
    for n in (0, 101, 255):
        Fo = o.eval_mt(K[n], NTHREADS)
        assert abs(Fb[n] - Fo) <= TOL * abs(Fo)




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
    rng = np.random.default_rng(cfg)
    h = 1e-5 * max(np.abs(kv[slots]).max(), 1.0)
    for n in range(4):
        d = np.zeros_like(kv)
        d[slots] = rng.normal(size=len(slots))
        d /= np.linalg.norm(d)
        fd = (obj.objective(kv + h * d) - obj.objective(kv - h * d)) / (2 * h)
        ad = float(np.dot(g, d))
        _say("directional", n, ad, fd)
        # 1e-6, the gradient tolerance (measured: 2e-9 one-element, 8e-10 ternary)
        assert abs(fd - ad) <= 1e-6 * max(abs(ad), 1e-300)
