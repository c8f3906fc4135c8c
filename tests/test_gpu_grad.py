"""Adjoint gradient: against central differences of the GPU objective, against
the reference's own central-difference gradients (golden fixtures) and against
central differences of the oracle.  Every component is checked against its
own size: |g_i - cd_i| <= 1e-6 |cd_i| + the rounding floor of that central
difference (F(x +- h_i e_i) each carry a rounding error of order 1e3 eps |F|,
sums of up to a million terms, so the quotient is uncertain by about that over
h_i); the gradient of a fixture spans four to seven orders of magnitude, so a
norm-wise bound would leave the small components unchecked."""
import numpy as np
import pytest

from conftest import golden_cases, load_case

import pyoracle
from meamzilla_amd import gen
from meamzilla_amd._lib import Objective

pytestmark = pytest.mark.gpu
GTOL = 1e-6


def _say(*a):
    print(*a, flush=True)


def _cd(fun, kv, idx, h_rel=1e-6, order=2):
    """Central differences with step h = h_rel * max(|k|, 1); order 4 uses the
    five-point stencil, whose O(h^4) truncation error stays below 1e-10 even
    where the squared density penalty makes the third derivative ~1e12."""
    g = np.zeros(len(idx))
    for n, i in enumerate(idx):
        h = h_rel * max(abs(kv[i]), 1.0)
        f = {}
        for m in ((-2, -1, 1, 2) if order == 4 else (-1, 1)):
            k = kv.copy()
            k[i] += m * h
            f[m] = fun(k)
        if order == 4:
            g[n] = (-f[2] + 8 * f[1] - 8 * f[-1] + f[-2]) / (12 * h)
        else:
            g[n] = (f[1] - f[-1]) / (2 * h)
    return g


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _steps(kv, idx, h_rel):
    return h_rel * np.maximum(np.abs(np.asarray(kv)[idx]), 1.0)


def _check(g, cd, h, F, what, tol=GTOL):
    """Component-wise bound; returns the worst ratio of error to bound."""
    floor = 1e3 * np.finfo(np.float64).eps * abs(F) / np.asarray(h)
    ratio = np.abs(np.asarray(g) - cd) / (tol * np.abs(cd) + floor)
    worst = float(np.max(ratio)) if ratio.size else 0.0
    _say("worst component error / bound", what, f"{worst:.3g}")
    assert worst <= 1.0, (what, worst, int(np.argmax(ratio)))
    return worst


@pytest.mark.parametrize("case", golden_cases())
def test_adjoint_matches_gpu_central_differences(case):
    _say("adjoint vs CD", case)
    pot, db, exp = load_case(case)
    obj = Objective(pot, db)
    kv = pot.knot_vector()
    F, g = obj.objective_grad(kv)
    assert abs(F - exp["F"]) <= 1e-10 * abs(exp["F"])
    idx = np.arange(len(kv))
    # In the clamp / extrapolate cases the squared density penalty dominates F
    # and its curvature along the rho knots is ~1e9, so the h = 1e-6 central
    # difference is itself off by ~1e-6; there the difference converges to the
    # adjoint as h shrinks (oracle: 2.5458395e6 at h = 1e-7 vs the adjoint's
    # 2.5458395e6), and h = 1e-7 with the five-point stencil is used.
    penalty_dominated = "clamp" in case or "extrap" in case
    h_rel = 1e-7 if penalty_dominated else 1e-6
    gcd = _cd(obj.objective, kv, idx, h_rel=h_rel, order=4 if penalty_dominated else 2)
    assert _rel(g, gcd) < GTOL, (case, _rel(g, gcd))
    _check(g, gcd, _steps(kv, idx, h_rel), F, case)


@pytest.mark.parametrize("case", [c for c in golden_cases() if load_case(c)[2]["fd_grad"]])
def test_adjoint_matches_reference_fd_gradient(case):
    _say("adjoint vs reference FD", case)
    pot, db, exp = load_case(case)
    obj = Objective(pot, db)
    F, g = obj.objective_grad()
    slots = pot.coeff_slots()
    gc = g[slots]
    ref = np.array(exp["fd_grad"])
    assert _rel(gc, ref) < GTOL, (case, _rel(gc, ref))
    _check(gc, ref, _steps(pot.knot_vector(), slots, exp["fd_h"]), exp["F"], case)


@pytest.mark.parametrize("cfg,scale", [(1, 6 / 256), (2, 10 / 512)])
def test_adjoint_matches_oracle_cd(cfg, scale):
    _say("adjoint vs oracle CD", cfg)
    w = gen.workload(cfg, scale=scale)
    gen.add_targets(w, pyoracle.evaluator)
    o = pyoracle.Oracle(w.pot, w.db)
    obj = Objective(w.pot, w.db)
    kv = w.pot.knot_vector()
    F, g = obj.objective_grad(kv)
    slots = w.pot.coeff_slots()
    gcd = _cd(o.objective, kv, slots)
    assert _rel(g[slots], gcd) < GTOL
    _check(g[slots], gcd, _steps(kv, slots, 1e-6), F, f"oracle cd {cfg}")


def test_adjoint_ternary_matches_central_differences():
    """Ternary potential (three g splines, eight g-adjoint replicas, forward
    tables read from global memory): the adjoint against central differences
    of the GPU objective on a seeded sample of the 252 coefficients."""
    _say("ternary adjoint vs CD")
    w = gen.workload(4, scale=2 / 2000)
    gen.add_targets(w, pyoracle.evaluator)
    obj = Objective(w.pot, w.db)
    kv = w.pot.knot_vector()
    F, g = obj.objective_grad(kv)
    assert abs(F - obj.objective(kv)) <= 1e-12 * abs(F)
    slots = np.asarray(w.pot.coeff_slots())
    idx = np.sort(np.random.default_rng(7).choice(slots, size=48, replace=False))
    gcd = _cd(obj.objective, kv, idx)
    assert _rel(g[idx], gcd) < GTOL, _rel(g[idx], gcd)
    _check(g[idx], gcd, _steps(kv, idx, 1e-6), F, "sampled coefficients")


def test_gmeam_ternary_matches_oracle_and_central_differences():
    """gmeam/spline with three elements (18 g splines chosen per triplet by the
    type of i and the pair of types j, k): objective, residuals and forces
    against the oracle, batched against single evaluation, and the adjoint
    against central differences on a seeded sample of coefficients."""
    _say("gmeam ternary")
    rng = np.random.Generator(np.random.PCG64(91))
    w = gen.workload(2, scale=8 / 512)
    db = w.db.subset([c for c in range(w.db.ncells) if w.db.natoms[c] <= 200][:3])
    db.types = rng.integers(0, 3, size=db.total_atoms).astype(np.int32)
    nk = dict(phi=10, rho=10, U=10, f=10, g=10)
    pot = gen.make_potential(rng, "gmeam/spline", 3, 5.0, nk, (-1.0, 3.0))
    truth = gen.make_potential(rng, "gmeam/spline", 3, 5.0, nk, (-1.0, 3.0))
    e, s, f = pyoracle.evaluator(truth, db)
    db.energy0, db.stress0, db.force0 = e, s, f
    o = pyoracle.Oracle(pot, db)
    obj = Objective(pot, db)
    ref = o.eval()
    r = obj.compute()
    assert abs(r["F"] - ref["F"]) <= 1e-10 * abs(ref["F"])
    for k in ("errvec", "forces"):
        assert _rel(r[k], ref[k]) < 1e-10, k
    kv = pot.knot_vector()
    F, g = obj.objective_grad(kv)
    Fb, Gb = obj.batch(np.stack([kv, kv]), grad=True)
    assert np.allclose(Fb, F, rtol=1e-12) and _rel(Gb[1], g) < 1e-12
    slots = np.asarray(pot.coeff_slots())
    idx = np.sort(np.random.default_rng(5).choice(slots, size=40, replace=False))
    gcd = _cd(obj.objective, kv, idx)
    assert _rel(g[idx], gcd) < GTOL, _rel(g[idx], gcd)
    _check(g[idx], gcd, _steps(kv, idx, 1e-6), F, "sampled coefficients")


@pytest.mark.parametrize("style", ["jmeam2/spline", "jmeam2"])
def test_jmeam2_ternary_adjoint_matches_oracle_central_differences(style):
    """jmeam2 with three elements (two p p q terms beside g): the adjoint
    against central differences of the oracle's objective on a seeded sample
    of coefficients (p, q, p2, q2 included), batched against single."""
    _say("jmeam2 ternary", style)
    rng = np.random.Generator(np.random.PCG64(17))
    w = gen.workload(2, scale=8 / 512)
    db = w.db.subset([c for c in range(w.db.ncells) if w.db.natoms[c] <= 200][:3])
    db.types = rng.integers(0, 3, size=db.total_atoms).astype(np.int32)
    nk = dict(phi=10, rho=10, U=10, f=10, g=9, p=9, q=8)
    pot = gen.make_potential(rng, style, 3, 5.0, nk, (-1.0, 3.0))
    truth = gen.make_potential(rng, style, 3, 5.0, nk, (-1.0, 3.0))
    e, s, f = pyoracle.evaluator(truth, db)
    db.energy0, db.stress0, db.force0 = e, s, f
    o = pyoracle.Oracle(pot, db)
    obj = Objective(pot, db)
    kv = pot.knot_vector()
    F, g = obj.objective_grad(kv)
    assert abs(F - o.eval()["F"]) <= 1e-10 * abs(F)
    Fb, Gb = obj.batch(np.stack([kv, kv]), grad=True)
    assert np.allclose(Fb, F, rtol=1e-12) and _rel(Gb[1], g) < 1e-12
    slots = np.asarray(pot.coeff_slots())
    idx = np.sort(np.random.default_rng(7).choice(slots, size=48, replace=False))
    gcd = _cd(o.objective, kv, idx)
    assert _rel(g[idx], gcd) < GTOL, _rel(g[idx], gcd)


@pytest.mark.parametrize("case", ["eam_spline_c0", "meam_spline_binary", "jmeam_spline_binary"])
def test_torch_gradcheck(case):
    """torch.autograd.gradcheck of the objective through the thin
    torch.autograd.Function over mz_objective_grad: torch's own central
    differences (eps 1e-6) against the adjoint, element by element."""
    import torch
    from meamzilla_amd.autograd import objective
    _say("gradcheck", case)
    pot, db, _ = load_case(case)
    obj = Objective(pot, db)
    kv = torch.tensor(pot.knot_vector(), dtype=torch.float64, requires_grad=True)
    _, g = obj.objective_grad(pot.knot_vector())
    # absolute floor: the difference quotient of F (~1e4) at eps 1e-6 carries
    # ~1e-5 of rounding, and a 1e-7 share of the largest entry
    atol = max(1e-4, 1e-7 * float(np.max(np.abs(g))))
    assert torch.autograd.gradcheck(lambda k: objective(k, obj), (kv,), eps=1e-6, atol=atol, rtol=1e-5)


def test_batched_gradient_matches_single():
    _say("batched gradient")
    w = gen.workload(2, scale=12 / 512)
    gen.add_targets(w, pyoracle.evaluator)
    obj = Objective(w.pot, w.db)
    K = w.candidate_knots(6)
    Fb, Gb = obj.batch(K, grad=True)
    for c in range(len(K)):
        F, g = obj.objective_grad(K[c])
        assert abs(Fb[c] - F) <= 1e-12 * abs(F)
        assert _rel(Gb[c], g) < 1e-12
    for c in range(len(K)):
        assert abs(Fb[c] - obj.objective(K[c])) <= 1e-12 * abs(Fb[c])
