"""Run-to-run reproducibility: the forward path has fixed summation orders
(no atomics), so two evaluations of the same knot vector give bitwise-equal
objective, residual vector, per-cell energies and stresses and forces, plain
and batched.  The gradient's coefficient adjoints are summed per workgroup in
LDS, where waves of one workgroup add in an order that can change between
runs, and then across workgroups in a fixed order; two gradients agree to
1e-13 of the gradient's size and to 1e-12 of every component's own size."""
import numpy as np
import pytest

import pyoracle
from meamzilla_amd import gen
from meamzilla_amd._lib import Objective

pytestmark = pytest.mark.gpu
KEYS = ("errvec", "cell_energy", "cell_stress", "forces")


def _say(*a):
    print(*a, flush=True)


def _close(g1, g2):
    d = np.abs(g1 - g2)
    top = np.max(np.abs(g1))
    assert np.max(d) <= 1e-13 * top, np.max(d) / top
    assert np.all(d <= 1e-12 * np.abs(g1) + 1e-14 * top), np.max(d / (np.abs(g1) + 1e-14 * top))


def _workload(cfg, ncells, total):
    w = gen.workload(cfg, scale=ncells / total)
    gen.add_targets(w, pyoracle.evaluator)
    return w


@pytest.mark.parametrize("cfg,ncells,total", [(1, 32, 256), (4, 3, 2000)])
def test_forward_bitwise_reproducible(cfg, ncells, total):
    _say("determinism forward", cfg, ncells)
    w = _workload(cfg, ncells, total)
    obj = Objective(w.pot, w.db)
    a, b = obj.compute(), obj.compute()
    assert a["F"] == b["F"]
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k
    rng = np.random.Generator(np.random.PCG64(5))
    kv = w.pot.knot_vector() * (1.0 + 0.02 * rng.standard_normal((6, obj.info.nknot)))
    f1, f2 = obj.batch(kv), obj.batch(kv)
    assert np.array_equal(f1, f2)


@pytest.mark.parametrize("cfg,ncells,total", [(1, 32, 256), (3, 40, 1000), (4, 3, 2000)])
def test_gradient_reproducible(cfg, ncells, total):
    _say("determinism gradient", cfg, ncells)
    w = _workload(cfg, ncells, total)
    obj = Objective(w.pot, w.db)
    F1, g1 = obj.objective_grad()
    F2, g2 = obj.objective_grad()
    assert F1 == F2
    _close(g1, g2)
    rng = np.random.Generator(np.random.PCG64(6))
    kv = w.pot.knot_vector() * (1.0 + 0.02 * rng.standard_normal((8, obj.info.nknot)))
    Fa, Ga = obj.batch(kv, grad=True)
    Fb, Gb = obj.batch(kv, grad=True)
    assert np.array_equal(Fa, Fb)
    for a, b in zip(Ga, Gb):
        _close(a, b)
