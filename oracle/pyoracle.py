"""ctypes front end of the oracle library (oracle/libmeamz_oracle.so).

TEST INFRASTRUCTURE ONLY: the oracle is the checker.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmeamz_oracle.so")
STYLE_CODE = {"eam/spline": 0, "meam/spline": 1, "eam": 2, "meam": 3, "gmeam/spline": 4, "gmeam": 5,
              "pair/spline": 6, "pair": 7, "jmeam/spline": 8, "jmeam": 9, "jmeam2/spline": 10, "jmeam2": 11}


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "libmeamz_oracle.so"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_void_p, C.c_void_p, C.c_char_p]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_counts.argtypes = [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.orc_errvec_size.argtypes = [C.c_void_p]
        L.orc_eval.restype = C.c_double
        L.orc_eval.argtypes = [C.c_void_p, dp, dp, dp, dp, dp]
        L.orc_eval_mt.restype = C.c_double
        L.orc_eval_mt.argtypes = [C.c_void_p, dp, C.c_int]
        L.orc_eval_mt_out.restype = C.c_double
        L.orc_eval_mt_out.argtypes = [C.c_void_p, dp, C.c_int, dp, dp, dp, dp]
        L.orc_densities.argtypes = [C.c_void_p, dp, dp]
        ip = C.POINTER(C.c_int)
        L.orc_spline_refresh.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, ip, dp, dp, dp]
        L.orc_knotshift.restype = C.c_int
        L.orc_knotshift.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, dp]
        L.orc_splint_knot.restype = C.c_double
        L.orc_splint_knot.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, ip, dp, C.c_int, C.c_double, dp]
        L.orc_splint_point.restype = C.c_double
        L.orc_splint_point.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, ip, dp, C.c_double, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


class _Layout(C.Structure):
    _fields_ = [("style", C.c_int), ("ntypes", C.c_int), ("nspl", C.c_int), ("basis", C.c_void_p),
                ("nknots", C.c_void_p), ("x0", C.c_void_p), ("xn", C.c_void_p), ("bc", C.c_void_p),
                ("embed_extrap", C.c_int)]


class _Db(C.Structure):
    _fields_ = [("ncells", C.c_int), ("natoms", C.c_void_p), ("box", C.c_void_p), ("types", C.c_void_p),
                ("pos", C.c_void_p), ("force0", C.c_void_p), ("energy0", C.c_void_p), ("stress0", C.c_void_p),
                ("weight", C.c_void_p), ("eweight", C.c_void_p), ("sweight", C.c_void_p),
                ("energy_weight", C.c_double), ("stress_weight", C.c_double)]


class Oracle:
    """Pair lists are built once; every eval() refreshes the splines from a knot vector."""

    def __init__(self, pot, db):
        L = lib()
        self.pot = pot
        keep = {}
        keep["basis"] = np.array([s.basis for s in pot.splines], dtype=np.int32)
        keep["nknots"] = np.array([s.n for s in pot.splines], dtype=np.int32)
        keep["x0"] = np.array([s.x0 for s in pot.splines], dtype=np.float64)
        keep["xn"] = np.array([s.xn for s in pot.splines], dtype=np.float64)
        keep["bc"] = np.array([b for s in pot.splines for b in s.bc], dtype=np.int32)
        lay = _Layout(STYLE_CODE[pot.style], pot.ntypes, len(pot.splines), keep["basis"].ctypes.data,
                      keep["nknots"].ctypes.data, keep["x0"].ctypes.data, keep["xn"].ctypes.data,
                      keep["bc"].ctypes.data, int(pot.embed_extrap))
        arr = lambda a, t: np.ascontiguousarray(a, dtype=t)
        d = dict(natoms=arr(db.natoms, np.int32), box=arr(db.box, np.float64), types=arr(db.types, np.int32),
                 pos=arr(db.pos, np.float64), force0=arr(db.force0, np.float64),
                 energy0=arr(db.energy0, np.float64), stress0=arr(db.stress0, np.float64),
                 weight=arr(db.weight, np.float64), eweight=arr(db.eweight, np.float64),
                 sweight=arr(db.sweight, np.float64))
        keep.update(d)
        cdb = _Db(db.ncells, *[d[k].ctypes.data for k in ("natoms", "box", "types", "pos", "force0", "energy0",
                                                          "stress0", "weight", "eweight", "sweight")],
                  float(db.energy_weight), float(db.stress_weight))
        err = C.create_string_buffer(256)
        self._ctx = L.orc_create(C.byref(lay), C.byref(cdb), err)
        if not self._ctx:
            raise ValueError(err.value.decode())
        self._keep = keep
        self.ncells = db.ncells
        self.natoms = db.total_atoms
        self.nerr = L.orc_errvec_size(self._ctx)

    def __del__(self):
        if getattr(self, "_ctx", None):
            lib().orc_destroy(self._ctx)
            self._ctx = None

    def counts(self):
        a, b = C.c_longlong(), C.c_longlong()
        lib().orc_counts(self._ctx, C.byref(a), C.byref(b))
        return a.value, b.value

    def eval(self, knots: Optional[np.ndarray] = None) -> dict:
        kv = np.ascontiguousarray(self.pot.knot_vector() if knots is None else knots, dtype=np.float64)
        ev = np.zeros(self.nerr)
        ce = np.zeros(self.ncells)
        cs = np.zeros((self.ncells, 6))
        fr = np.zeros((self.natoms, 3))
        F = lib().orc_eval(self._ctx, _dp(kv), _dp(ev), _dp(ce), _dp(cs), _dp(fr))
        return dict(F=F, errvec=ev, cell_energy=ce, cell_stress=cs, forces=fr)

    def objective(self, knots: Optional[np.ndarray] = None) -> float:
        kv = np.ascontiguousarray(self.pot.knot_vector() if knots is None else knots, dtype=np.float64)
        return lib().orc_eval(self._ctx, _dp(kv), None, None, None, None)

    def eval_mt(self, knots: Optional[np.ndarray] = None, nthreads: int = 1) -> float:
        kv = np.ascontiguousarray(self.pot.knot_vector() if knots is None else knots, dtype=np.float64)
        return lib().orc_eval_mt(self._ctx, _dp(kv), int(nthreads))

    def eval_full_mt(self, knots: Optional[np.ndarray] = None, nthreads: int = 16) -> dict:
        """eval() on nthreads threads (atoms balanced by work, one final sum)."""
        kv = np.ascontiguousarray(self.pot.knot_vector() if knots is None else knots, dtype=np.float64)
        ev = np.zeros(self.nerr)
        ce = np.zeros(self.ncells)
        cs = np.zeros((self.ncells, 6))
        fr = np.zeros((self.natoms, 3))
        F = lib().orc_eval_mt_out(self._ctx, _dp(kv), int(nthreads), _dp(ev), _dp(ce), _dp(cs), _dp(fr))
        return dict(F=F, errvec=ev, cell_energy=ce, cell_stress=cs, forces=fr)

    def densities(self, knots: Optional[np.ndarray] = None) -> np.ndarray:
        kv = np.ascontiguousarray(self.pot.knot_vector() if knots is None else knots, dtype=np.float64)
        out = np.zeros(self.natoms)
        lib().orc_densities(self._ctx, _dp(kv), _dp(out))
        return out


def evaluator(pot, db):
    """gen.add_targets-compatible evaluator (cell energies, stresses, forces)."""
    r = Oracle(pot, db).eval()
    return r["cell_energy"], r["cell_stress"], r["forces"]


def spline_refresh(s):
    L = lib()
    bc = np.array(s.bc, dtype=np.int32)
    rec = s.knot_record()
    ypp = np.zeros(s.n)
    yp4 = np.zeros(s.n)
    L.orc_spline_refresh(s.basis, s.n, s.x0, s.xn, bc.ctypes.data_as(C.POINTER(C.c_int)), _dp(rec), _dp(ypp),
                         _dp(yp4))
    return ypp, yp4


def knotshift(s, r):
    sh = C.c_double(0.0)
    k = lib().orc_knotshift(s.n, s.x0, s.xn, float(r), C.byref(sh))
    return k, sh.value


def splint_knot(s, knot, shift):
    bc = np.array(s.bc, dtype=np.int32)
    rec = s.knot_record()
    g = C.c_double(0.0)
    v = lib().orc_splint_knot(s.basis, s.n, s.x0, s.xn, bc.ctypes.data_as(C.POINTER(C.c_int)), _dp(rec),
                              int(knot), float(shift), C.byref(g))
    return v, g.value


def splint_point(s, r):
    bc = np.array(s.bc, dtype=np.int32)
    rec = s.knot_record()
    g = C.c_double(0.0)
    g2 = C.c_double(0.0)
    v = lib().orc_splint_point(s.basis, s.n, s.x0, s.xn, bc.ctypes.data_as(C.POINTER(C.c_int)), _dp(rec),
                               float(r), C.byref(g), C.byref(g2))
    return v, g.value, g2.value
