"""Device pair-list build (mz_create) against the host build
(mz_create_host_lists): the same pairs in the same order, so every result is
bit-identical; the same counts as the oracle's restatement of find_pairs /
find_triplets; the same errors."""
import numpy as np
import pytest

from conftest import golden_cases, load_case

import pyoracle
from meamzilla_amd import gen
from meamzilla_amd._lib import MeamzError, Objective

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _same(pot, db, grad=True):
    """Both builders: equal counts, bit-identical objective, residuals and
    forces, and (where the style has one) the same adjoint gradient."""
    dev = Objective(pot, db)
    host = Objective(pot, db, host_lists=True)
    assert dev.info.list_device == 1 and host.info.list_device == 0
    assert (dev.info.npairs, dev.info.ntriplets, dev.info.nchunks) == \
        (host.info.npairs, host.info.ntriplets, host.info.nchunks)
    a, b = dev.compute(), host.compute()
    for k in ("F", "errvec", "cell_energy", "cell_stress", "forces"):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    if grad:
        # the adjoint sums with unordered atomics: equal to rounding, not bitwise
        Fa, ga = dev.objective_grad()
        Fb, gb = host.objective_grad()
        assert Fa == Fb
        assert np.max(np.abs(ga - gb)) <= 1e-11 * max(np.max(np.abs(gb)), 1e-300)
    return dev, host


@pytest.mark.parametrize("case", golden_cases())
def test_device_lists_match_host_lists_golden(case):
    _say("lists golden", case)
    pot, db, _ = load_case(case)
    _same(pot, db)


@pytest.mark.parametrize("cfg,scale", [(0, 1.0), (1, 24 / 256), (2, 40 / 512), (3, 30 / 1000), (4, 3 / 2000)])
def test_device_lists_match_host_lists_workloads(cfg, scale):
    _say("lists workload", cfg, scale)
    w = gen.workload(cfg, scale=scale)
    gen.add_targets(w, pyoracle.evaluator)
    dev, host = _same(w.pot, w.db, grad=cfg < 4)
    assert (dev.info.npairs, dev.info.ntriplets) == pyoracle.Oracle(w.pot, w.db).counts()
    _say("list build ms: device", round(dev.info.list_ms, 2), "host", round(host.info.list_ms, 2))


def _cells_db(cells):
    from meamzilla_amd.formats import Database
    natoms = np.array([len(c[0]) for c in cells], np.int32)
    box = np.array([c[1] for c in cells], float)
    pos = np.concatenate([np.asarray(c[0], float) for c in cells])
    types = np.concatenate([np.asarray(c[2], np.int32) for c in cells])
    nc, na = len(cells), len(pos)
    return Database(natoms, box, types, pos, np.zeros((na, 3)), np.zeros(nc), np.zeros((nc, 6)),
                    np.ones(nc), np.ones(nc), np.ones(nc), 1.0, 1.0)


def test_device_lists_edge_geometries():
    """Isolated atom, dimer, self-image-only atom in a skewed cell, an atom in
    a cube whose edge is the cutoff (images exactly at r = cutoff) and a pair
    at half the cutoff."""
    w = gen.workload(1, scale=0.01)
    rc = w.pot.max_rcut()
    big = np.diag([20.0, 20.0, 20.0])
    cells = [
        (np.array([[5.0, 5.0, 5.0]]), big, [0]),
        (np.array([[5.0, 5.0, 5.0], [7.9, 5.1, 5.0]]), big, [0, 0]),
        (np.array([[0.3, 0.2, 0.1]]), np.array([[2.9, 0.0, 0.0], [0.2, 3.0, 0.0], [0.1, -0.3, 3.1]]), [0]),
        (np.array([[0.0, 0.0, 0.0]]), np.diag([rc, rc, rc]), [0]),
        (np.array([[0.0, 0.0, 0.0], [rc / 2, 0.0, 0.0]]), np.diag([rc, rc * 1.5, rc * 1.5]), [0, 0]),
    ]
    db = _cells_db(cells)
    gen.add_targets(gen.Workload("edge", w.pot, w.truth, db, w.cutoff, 1, 5), pyoracle.evaluator)
    dev, _ = _same(w.pot, db)
    assert (dev.info.npairs, dev.info.ntriplets) == pyoracle.Oracle(w.pot, db).counts()


def test_device_lists_same_errors():
    """A pair below the first knot of phi: both builders refuse with the same
    message, naming the first offending pair in the reference's order."""
    w = gen.workload(1, scale=0.01)
    big = np.diag([20.0, 20.0, 20.0])
    close = _cells_db([(np.array([[5.0, 5.0, 5.0], [7.9, 5.1, 5.0]]), big, [0, 0]),
                       (np.array([[5.0, 5.0, 5.0], [6.0, 5.0, 5.0], [9.0, 5.0, 5.0]]), big, [0, 0, 0])])
    msgs = []
    for host in (False, True):
        with pytest.raises(MeamzError) as ei:
            Objective(w.pot, close, host_lists=host)
        msgs.append(str(ei.value))
    assert msgs[0] == msgs[1] and "too short" in msgs[0] and "cell=1 atom=0 neigh=1" in msgs[0], msgs
