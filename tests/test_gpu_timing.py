"""Kernel timing: with timing on, calls launch their kernels on the stream
between HIP events; the times are summed over the launches of a call, and
results equal those of the captured graphs that run with timing off."""
import numpy as np
import pytest

import pyoracle
from meamzilla_amd import gen
from meamzilla_amd._lib import Objective

pytestmark = pytest.mark.gpu


def test_timed_graphs_report_kernel_times_and_keep_results():
    print("timing: 16-cell one-element slice", flush=True)
    w = gen.workload(1, scale=16 / 256)
    gen.add_targets(w, pyoracle.evaluator)
    obj = Objective(w.pot, w.db)
    rng = np.random.Generator(np.random.PCG64(9))
    kv = w.pot.knot_vector() * (1.0 + 0.02 * rng.standard_normal((8, obj.info.nknot)))
    F0, G0 = obj.batch(kv, grad=True)
    f0 = obj.batch(kv)
    obj.enable_timing(True)
    F1, G1 = obj.batch(kv, grad=True)
    t = obj.last_timing()
    assert t.launches >= 1
    assert t.atom_ms > 0 and t.adjoint_ms > 0 and t.resid_ms > 0 and t.refresh_ms > 0
    assert t.total_ms >= t.atom_ms + t.adjoint_ms
    f1 = obj.batch(kv)
    tf = obj.last_timing()
    assert tf.launches >= 1 and tf.atom_ms > 0 and tf.adjoint_ms == 0
    obj.enable_timing(False)
    assert np.array_equal(f0, f1)
    assert np.array_equal(F0, F1)
    # coefficient adjoints are summed with LDS atomics whose order between the
    # waves of a workgroup can change from run to run
    for a, b in zip(G0, G1):
        d = np.abs(b - a)
        top = np.max(np.abs(a))
        assert np.max(d) <= 1e-13 * top
        assert np.all(d <= 1e-12 * np.abs(a) + 1e-14 * top)
