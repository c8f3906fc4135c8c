"""ctypes binding of libmeamz.so (the C-ABI in include/meamz.h).

The library is the product path: there is no fallback.  If it is missing or a
call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from typing import Optional

import numpy as np

from .formats import Database, PotentialSpec

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmeamz.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "meamz.h")
STYLE_CODE = {"eam/spline": 0, "meam/spline": 1, "eam": 2, "meam": 3, "gmeam/spline": 4, "gmeam": 5,
              "pair/spline": 6, "pair": 7, "jmeam/spline": 8, "jmeam": 9, "jmeam2/spline": 10, "jmeam2": 11}


class MeamzError(RuntimeError):
    pass


class _Layout(C.Structure):
    _fields_ = [("style", C.c_int), ("ntypes", C.c_int), ("nspl", C.c_int), ("basis", C.c_void_p),
                ("nknots", C.c_void_p), ("x0", C.c_void_p), ("xn", C.c_void_p), ("bc", C.c_void_p),
                ("embed_extrap", C.c_int)]


class _Db(C.Structure):
    _fields_ = [("ncells", C.c_int), ("natoms", C.c_void_p), ("box", C.c_void_p), ("types", C.c_void_p),
                ("pos", C.c_void_p), ("force0", C.c_void_p), ("energy0", C.c_void_p), ("stress0", C.c_void_p),
                ("weight", C.c_void_p), ("eweight", C.c_void_p), ("sweight", C.c_void_p),
                ("energy_weight", C.c_double), ("stress_weight", C.c_double)]


class Info(C.Structure):
    _fields_ = [("ncells", C.c_int), ("natoms", C.c_int), ("nspl", C.c_int), ("nknot", C.c_int),
                ("nerr", C.c_int), ("npairs", C.c_longlong), ("ntriplets", C.c_longlong),
                ("nchunks", C.c_int), ("device", C.c_int), ("list_ms", C.c_double), ("list_device", C.c_int)]


class Tile(C.Structure):
    _fields_ = [("max_atoms", C.c_int), ("max_pairs", C.c_int), ("block", C.c_int), ("cand_tile", C.c_int)]


class Timing(C.Structure):
    _fields_ = [("atom_ms", C.c_double), ("resid_ms", C.c_double), ("refresh_ms", C.c_double),
                ("adjoint_ms", C.c_double), ("total_ms", C.c_double), ("launches", C.c_int)]


_lib = None


def header_symbols() -> list:
    """Function names declared in include/meamz.h."""
    with open(HEADER) as fh:
        txt = fh.read()
    return sorted(set(re.findall(r"\b(mz_[a-z_]+)\s*\(", txt)))


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MeamzError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C meamzilla_amd/csrc")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    L.mz_last_error.restype = C.c_char_p
    L.mz_create.argtypes = [vp, vp, C.POINTER(vp)]
    L.mz_create_host_lists.argtypes = [vp, vp, C.POINTER(vp)]
    L.mz_probe_peaks.argtypes = [dp, dp]
    L.mz_destroy.argtypes = [vp]
    L.mz_info.argtypes = [vp, C.POINTER(Info)]
    L.mz_objective.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.mz_objective_noalloc.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.mz_densities.argtypes = [vp, dp, dp]
    L.mz_objective_grad.argtypes = [vp, dp, dp, dp]
    L.mz_objective_batch.argtypes = [vp, C.c_int, dp, dp, dp]
    L.mz_residuals_batch.argtypes = [vp, C.c_int, dp, dp, dp]
    L.mz_set_tile.argtypes = [vp, C.POINTER(Tile)]
    L.mz_get_tile.argtypes = [vp, C.POINTER(Tile)]
    L.mz_autotune.argtypes = [vp, C.c_int, C.c_int, C.POINTER(Tile)]
    L.mz_autotune_knots.argtypes = [vp, C.c_int, dp, C.c_int, C.c_int, C.POINTER(Tile)]
    L.mz_enable_timing.argtypes = [vp, C.c_int]
    L.mz_last_timing.argtypes = [vp, C.POINTER(Timing)]
    L.mz_spline_refresh.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, dp, dp, dp]
    _lib = L
    return L


def spline_refresh(spline) -> tuple:
    """y'' and y'''' at the knots of one spline (mz_spline_refresh, host only)."""
    rec = np.ascontiguousarray(spline.knot_record(), dtype=np.float64)
    ypp = np.zeros(spline.n)
    yp4 = np.zeros(spline.n)
    _check(lib().mz_spline_refresh(int(spline.basis), int(spline.n), float(spline.x0), float(spline.xn),
                                   int(spline.bc[0] == -1), int(spline.bc[1] == -1), _dp(rec), _dp(ypp), _dp(yp4)))
    return ypp, yp4


def probe_peaks() -> tuple:
    """(FP64 FMA TFLOP/s, HBM copy GB/s) this GPU reaches (mz_probe_peaks)."""
    tf = np.zeros(1)
    gbs = np.zeros(1)
    _check(lib().mz_probe_peaks(_dp(tf), _dp(gbs)))
    return float(tf[0]), float(gbs[0])


def _check(rc: int) -> None:
    if rc != 0:
        msg = lib().mz_last_error()
        raise MeamzError(f"libmeamz error {rc}: {msg.decode() if msg else ''}")


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


class Objective:
    """One database and spline layout resident on the GPU.

    Mirrors the reference's Potential::compute(comm, ErrorVec*) contract:
    ``compute(knots, want_errvec)`` returns the objective and, on request, the
    residual vector; per-cell energies / stresses and forces are returned too.
    ``host_lists`` builds the pair lists on the host (mz_create_host_lists)
    instead of on the device; the lists are the same.
    """

    def __init__(self, pot: PotentialSpec, db: Database, host_lists: bool = False):
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
        ctx = C.c_void_p()
        create = L.mz_create_host_lists if host_lists else L.mz_create
        _check(create(C.byref(lay), C.byref(cdb), C.byref(ctx)))
        self._ctx = ctx
        self.info = Info()
        _check(L.mz_info(ctx, C.byref(self.info)))
        self.ncells = db.ncells
        self.natoms = db.total_atoms

    def close(self):
        if getattr(self, "_ctx", None):
            lib().mz_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        self.close()

    def _kv(self, knots):
        kv = self.pot.knot_vector() if knots is None else knots
        kv = np.ascontiguousarray(kv, dtype=np.float64)
        if kv.shape[-1] != self.info.nknot:
            raise ValueError(f"knot vector has {kv.shape[-1]} entries, layout needs {self.info.nknot}")
        return kv

    def compute(self, knots: Optional[np.ndarray] = None, errvec: bool = True, outputs: bool = True) -> dict:
        kv = self._kv(knots)
        F = np.zeros(1)
        ev = np.zeros(self.info.nerr) if errvec else None
        ce = np.zeros(self.ncells) if outputs else None
        cs = np.zeros((self.ncells, 6)) if outputs else None
        fr = np.zeros((self.natoms, 3)) if outputs else None
        _check(lib().mz_objective(self._ctx, _dp(kv), _dp(F), _dp(ev), _dp(ce), _dp(cs), _dp(fr)))
        return dict(F=float(F[0]), errvec=ev, cell_energy=ce, cell_stress=cs, forces=fr)

    def compute_noalloc(self, knots: Optional[np.ndarray] = None, errvec: bool = True, outputs: bool = True) -> dict:
        """compute() through mz_objective_noalloc: no allocation at call time,
        buffers in the library's static device array."""
        kv = self._kv(knots)
        F = np.zeros(1)
        ev = np.zeros(self.info.nerr) if errvec else None
        ce = np.zeros(self.ncells) if outputs else None
        cs = np.zeros((self.ncells, 6)) if outputs else None
        fr = np.zeros((self.natoms, 3)) if outputs else None
        _check(lib().mz_objective_noalloc(self._ctx, _dp(kv), _dp(F), _dp(ev), _dp(ce), _dp(cs), _dp(fr)))
        return dict(F=float(F[0]), errvec=ev, cell_energy=ce, cell_stress=cs, forces=fr)

    def objective(self, knots: Optional[np.ndarray] = None) -> float:
        kv = self._kv(knots)
        F = np.zeros(1)
        _check(lib().mz_objective(self._ctx, _dp(kv), _dp(F), None, None, None, None))
        return float(F[0])

    def densities(self, knots: Optional[np.ndarray] = None) -> np.ndarray:
        kv = self._kv(knots)
        out = np.zeros(self.natoms)
        _check(lib().mz_densities(self._ctx, _dp(kv), _dp(out)))
        return out

    def objective_grad(self, knots: Optional[np.ndarray] = None):
        kv = self._kv(knots)
        F = np.zeros(1)
        g = np.zeros(self.info.nknot)
        _check(lib().mz_objective_grad(self._ctx, _dp(kv), _dp(F), _dp(g)))
        return float(F[0]), g

    def batch(self, knots: np.ndarray, grad: bool = False):
        kv = self._kv(np.atleast_2d(knots))
        n = kv.shape[0]
        F = np.zeros(n)
        G = np.zeros((n, self.info.nknot)) if grad else None
        _check(lib().mz_objective_batch(self._ctx, n, _dp(kv), _dp(F), _dp(G)))
        return (F, G) if grad else F

    def residuals_batch(self, knots: np.ndarray):
        """(F[n], residual vectors [n, nerr]) of n knot vectors in one call."""
        kv = self._kv(np.atleast_2d(knots))
        n = kv.shape[0]
        F = np.zeros(n)
        E = np.zeros((n, self.info.nerr))
        _check(lib().mz_residuals_batch(self._ctx, n, _dp(kv), _dp(F), _dp(E)))
        return F, E

    def set_tile(self, max_atoms: int, max_pairs: int, block: int, cand_tile: int = 1):
        t = Tile(max_atoms, max_pairs, block, cand_tile)
        _check(lib().mz_set_tile(self._ctx, C.byref(t)))
        _check(lib().mz_info(self._ctx, C.byref(self.info)))

    def get_tile(self) -> Tile:
        t = Tile()
        _check(lib().mz_get_tile(self._ctx, C.byref(t)))
        return t

    def autotune(self, ncand: int = 1, reps: int = 3, grad: bool = False,
                 knots: Optional[np.ndarray] = None) -> Tile:
        """Times tiles with batched calls and keeps the fastest: objective
        calls on zero vectors by default; with knots (ncand vectors) and grad,
        the caller's vectors in objective-and-gradient calls."""
        t = Tile()
        if knots is None and not grad:
            _check(lib().mz_autotune(self._ctx, int(ncand), int(reps), C.byref(t)))
        else:
            kv = None if knots is None else self._kv(np.atleast_2d(knots))
            if kv is not None and kv.shape[0] != ncand:
                raise ValueError(f"{kv.shape[0]} knot vectors for ncand={ncand}")
            _check(lib().mz_autotune_knots(self._ctx, int(ncand), _dp(kv), 1 if grad else 0, int(reps),
                                           C.byref(t)))
        _check(lib().mz_info(self._ctx, C.byref(self.info)))
        return t

    def enable_timing(self, on: bool = True):
        _check(lib().mz_enable_timing(self._ctx, 1 if on else 0))

    def last_timing(self) -> Timing:
        t = Timing()
        _check(lib().mz_last_timing(self._ctx, C.byref(t)))
        return t
