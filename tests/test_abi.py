"""The C-ABI library loads on a CPU-only host and exports every function
include/meamz.h declares.  No compute call is made here."""
import ctypes as C

import numpy as np
import pytest

from meamzilla_amd import _lib, formats


def test_library_exports_header_symbols():
    L = _lib.lib()
    names = _lib.header_symbols()
    assert "mz_objective" in names and "mz_objective_batch" in names and "mz_objective_grad" in names
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/meamz.h but not exported"


def test_layout_errors_are_reported_before_device_use():
    L = _lib.lib()
    basis = np.array([3, 3], dtype=np.int32)   # eam/spline needs 3 splines for one type
    nk = np.array([8, 8], dtype=np.int32)
    x0 = np.array([1.8, 1.8])
    xn = np.array([5.0, 5.0])
    bc = np.array([-1, -1, -1, -1], dtype=np.int32)
    lay = _lib._Layout(0, 1, 2, basis.ctypes.data, nk.ctypes.data, x0.ctypes.data, xn.ctypes.data, bc.ctypes.data, 0)
    natoms = np.array([1], dtype=np.int32)
    box = np.eye(3).ravel() * 10
    z = np.zeros(3)
    one = np.ones(1)
    types = np.zeros(1, dtype=np.int32)
    db = _lib._Db(1, natoms.ctypes.data, box.ctypes.data, types.ctypes.data, z.ctypes.data, z.ctypes.data,
                  one.ctypes.data, np.zeros(6).ctypes.data, one.ctypes.data, one.ctypes.data, one.ctypes.data, 1.0, 1.0)
    ctx = C.c_void_p()
    rc = L.mz_create(C.byref(lay), C.byref(db), C.byref(ctx))
    assert rc == -4  # MZ_ERR_LAYOUT
    assert b"splines" in L.mz_last_error()
    assert not ctx.value


def test_knot_vector_layout_matches_formats():
    s = formats.Spline(3, 1.8, 5.0, np.arange(6.0), (0, -1), (0.5, 0.0))
    rec = s.knot_record()
    assert rec.shape == (10,) and rec[6] == 0.5
