"""Run-to-run reproducibility: the forward path has fixed summation orders
(no atomics), so two evaluations of the same knot vector give bitwise-equal
objective, residual vector, per-cell energies and stresses and forces, plain
and batched.  The gradient's coefficient adjoints are summed per workgroup
in LDS and then across workgroups in a fixed order; two gradients agree
component by component to round-off (1e-13 of each component's size)."""
import numpy as np
import pytest

import pyoracle
from meamzilla_amd import gen
from meamzilla_amd._lib import Objective

pytestmark = pytest.mark.gpu
KEYS = ("errvec", "cell_energy", "cell_stress", "forces")


def _say(*a):
    print(*a, flush=True)


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
    scale = np.abs(g1) + 1e-12 * np.max(np.abs(g1))
    assert np.all(np.abs(g1 - g2) <= 1e-13 * scale), np.max(np.abs(g1 - g2) / scale)
    rng = np.random.Generator(np.random.PCG64(6))
    kv = w.pot.knot_vector() * (1.0 + 0.02 * rng.standard_normal((8, obj.info.nknot)))
    Fa, Ga = obj.batch(kv, grad=True)
    Fb, Gb = obj.batch(kv, grad=True)
    assert np.array_equal(Fa, Fb)
    sc = np.abs(Ga) + 1e-12 * np.max(np.abs(Ga), axis=1, keepdims=True)
    assert np.all(np.abs(Ga - Gb) <= 1e-13 * sc), np.max(np.abs(Ga - Gb) / sc)
