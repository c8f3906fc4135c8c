"""Seeded synthetic workloads for the MEAM/EAM objective.

Five workloads, numbered as they are quoted:

0. plain EAM, one element: 16 FCC 2x2x2 cells (32 atoms, a=4.05) with Gaussian
   displacements (0.1 A) and isotropic strain U(+-2%); cutoff 5.5; natural cubic
   splines phi/rho 10 knots on [1.8,5.5], U 8 knots on [0,2];
1. plain MEAM, one element: 256 cells of 128 atoms, half BCC 4x4x4 (a=3.16,
   displacement sigma U[0.05,0.3]) and half random-uniform at 0.063 /A^3 with
   2.0 A minimum separation; cutoff 5.0; phi/rho/f 12 knots on [1.8,5.0], U 10 on
   [-1,3], g 10 on [-1,1];
2. binary MEAM (12 splines x 12 knots) over 512 cells with log-uniform sizes in
   [8,512] on displaced, triclinically strained BCC/FCC supercells, 50% species A,
   256 candidate vectors = base x (1 + U(+-5%));
3. binary MEAM over 1000 cells of U[64,256] atoms, BCC/FCC/random-dense
   (0.06 /A^3);
4. ternary MEAM (21 splines x 12 knots) over 2000 cells of U[256,2048] atoms,
   cutoff 5.5, random species fractions, 64 candidate vectors.

Targets come from a second ("truth") knot set evaluated by a caller-supplied
evaluator plus Gaussian noise of 1 meV/atom, 0.01 eV/A and 0.1 GPa.  ``scale``
shrinks the number of cells for small test cases; every random draw is taken
from ``numpy.random.Generator(PCG64(seed))`` so a workload is reproducible.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable, List, Optional, Tuple

import numpy as np

from .formats import Database, PotentialSpec, Spline, function_counts, is_meam_style, jmeam_terms

GPA_TO_EV_A3 = 1.0 / 160.21766208
NOISE_E = 1e-3          # eV/atom
NOISE_F = 1e-2          # eV/A
NOISE_S = 0.1 * GPA_TO_EV_A3
MIN_SEP = 1.85          # every pair must stay above the first radial knot (1.8 A)


@dataclasses.dataclass
class Workload:
    name: str
    pot: PotentialSpec          # starting knot set (the vector being scored)
    truth: PotentialSpec        # knot set the targets are computed from
    db: Database                # geometry + targets (targets zero until add_targets)
    cutoff: float
    candidates: int = 1         # candidate vectors for the batched evaluator
    seed: int = 0

    def candidate_knots(self, count: Optional[int] = None, spread: float = 0.05) -> np.ndarray:
        """(count, nknot) knot vectors = base x (1 + U(-spread, spread))."""
        count = self.candidates if count is None else count
        rng = np.random.Generator(np.random.PCG64(self.seed + 7919))
        base = self.pot.knot_vector()
        fac = 1.0 + rng.uniform(-spread, spread, size=(count, base.size))
        out = base[None, :] * fac
        # boundary-condition slots of natural splines carry no meaning; keep them at base
        mask = _bc_slot_mask(self.pot)
        out[:, mask] = base[mask]
        return out


def _bc_slot_mask(pot: PotentialSpec) -> np.ndarray:
    m = []
    for s in pot.splines:
        m += [False] * s.n + [True] * 4
    return np.array(m)


# ----------------------------------------------------------------------------- lattices

def _lattice(kind: str, a: float, reps: Tuple[int, int, int]) -> Tuple[np.ndarray, np.ndarray]:
    if kind == "bcc":
        basis = np.array([[0, 0, 0], [0.5, 0.5, 0.5]])
    elif kind == "fcc":
        basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    else:
        raise ValueError(kind)
    k1, k2, k3 = reps
    grid = np.stack(np.meshgrid(np.arange(k1), np.arange(k2), np.arange(k3), indexing="ij"), -1).reshape(-1, 3)
    frac = (grid[:, None, :] + basis[None, :, :]).reshape(-1, 3) / np.array([k1, k2, k3])
    box = np.diag([k1 * a, k2 * a, k3 * a]).astype(np.float64)
    return frac, box


def _reps_for(m: int) -> Tuple[int, int, int]:
    k3 = max(1, int(round(m ** (1 / 3))))
    k2 = max(1, int(round(math.sqrt(m / k3))))
    k1 = max(1, int(round(m / (k2 * k3))))
    return (k1, k2, k3)


def _min_dist_ok(frac: np.ndarray, box: np.ndarray, dmin: float) -> np.ndarray:
    """Boolean per atom: True if the atom has no periodic neighbour closer than dmin.

    A k-d tree over the atoms and their periodic images (enough image layers
    to cover dmin in every lattice direction) counts the points within dmin
    of each atom; the atom itself is the only one allowed."""
    from scipy.spatial import cKDTree
    n = len(frac)
    heights = np.abs(np.linalg.det(box)) / np.linalg.norm(np.cross(box[[1, 2, 0]], box[[2, 0, 1]]), axis=1)
    reach = [int(math.ceil(dmin / h)) for h in heights]
    shifts = np.stack(np.meshgrid(*[np.arange(-r, r + 1) for r in reach], indexing="ij"), -1).reshape(-1, 3)
    wrapped = frac - np.floor(frac)
    img = (wrapped[None, :, :] + shifts[:, None, :]).reshape(-1, 3)
    # only images within dmin of the cell can be that close to an atom in it
    margin = dmin / heights
    keep = np.all((img > -margin) & (img < 1.0 + margin), axis=1)
    tree = cKDTree(img[keep] @ box)
    # distance < dmin is a violation; query_ball_point counts distance <= r
    r = np.nextafter(dmin, 0.0)
    counts = tree.query_ball_point(wrapped @ box, r=r, return_length=True)
    return counts == 1


def _displace(rng, frac, box, sigma, dmin=MIN_SEP, tries=200):
    inv = np.linalg.inv(box)
    cur = frac + (rng.normal(0.0, sigma, size=frac.shape) @ inv)
    for _ in range(tries):
        ok = _min_dist_ok(cur, box, dmin)
        if ok.all():
            return cur - np.floor(cur)
        bad = ~ok
        cur[bad] = frac[bad] + (rng.normal(0.0, sigma, size=(bad.sum(), 3)) @ inv)
    raise RuntimeError("could not place atoms above the minimum separation")


def _random_dense(rng, n, density, dmin):
    L = (n / density) ** (1.0 / 3.0)
    box = np.diag([L, L, L])
    pts = np.zeros((0, 3))
    tries = 0
    while len(pts) < n:
        tries += 1
        if tries > 200000:
            raise RuntimeError("random packing failed")
        p = rng.uniform(0, 1, size=3)
        if len(pts):
            d = pts - p
            d -= np.rint(d)
            if np.min(np.sum((d * L) ** 2, axis=1)) < dmin * dmin:
                continue
        pts = np.vstack([pts, p])
    return pts, box


def _strain_box(rng, box, amp, triclinic):
    if triclinic:
        e = rng.uniform(-amp, amp, size=(3, 3))
        e = 0.5 * (e + e.T)
    else:
        e = np.eye(3) * rng.uniform(-amp, amp)
    return box @ (np.eye(3) + e)


# ----------------------------------------------------------------------------- splines

def _taper(x, x0, xn):
    return (xn - x) / (xn - x0)


def _radial(rng, x0, xn, n, lo, hi, basis):
    x = np.linspace(x0, xn, n)
    y = rng.uniform(lo, hi, size=n) * _taper(x, x0, xn)
    return Spline(basis, x0, xn, y)


def _flat(rng, x0, xn, n, lo, hi, basis):
    return Spline(basis, x0, xn, rng.uniform(lo, hi, size=n))


def make_potential(rng, style: str, nt: int, rc: float, nk: dict, u_range, basis: int = 3,
                   rho_amp: float = 0.1, g_amp: float = 0.05) -> PotentialSpec:
    counts = function_counts(style, nt)
    spl: List[Spline] = []
    spl += [_radial(rng, 1.8, rc, nk["phi"], -0.2, 0.4, basis) for _ in range(counts[0])]
    if len(counts) == 1:  # pair styles: phi only
        return PotentialSpec(style, nt, spl)
    spl += [_radial(rng, 1.8, rc, nk["rho"], 0.0, rho_amp, basis) for _ in range(counts[1])]
    spl += [_flat(rng, u_range[0], u_range[1], nk["U"], -0.5, 0.5, basis) for _ in range(counts[2])]
    if is_meam_style(style):
        fs = [_radial(rng, 1.8, rc, nk["f"], -1.0, 1.0, basis) for _ in range(counts[3])]
        m = max(np.abs(s.y).max() for s in fs)
        for s in fs:
            s.y = s.y / m          # largest |f| knot is exactly 1
        spl += fs
        spl += [_flat(rng, -1.0, 1.0, nk["g"], -g_amp, g_amp, basis) for _ in range(counts[4])]
    for m in range(jmeam_terms(style)):  # p (largest |knot| exactly 1) and q
        ps = [_radial(rng, 1.8, rc, nk.get("p", nk["f"]), -1.0, 1.0, basis) for _ in range(counts[5 + 2 * m])]
        mp = max(np.abs(s.y).max() for s in ps)
        for s in ps:
            s.y = s.y / mp
        spl += ps
        spl += [_flat(rng, -1.0, 1.0, nk.get("q", nk["g"]), -g_amp, g_amp, basis) for _ in range(counts[6 + 2 * m])]
    return PotentialSpec(style, nt, spl)


# ----------------------------------------------------------------------------- workloads

def _empty_db(cells) -> Database:
    natoms = np.array([len(c[0]) for c in cells], dtype=np.int32)
    box = np.array([c[1] for c in cells])
    pos = np.concatenate([c[0] @ c[1] for c in cells])
    types = np.concatenate([c[2] for c in cells]).astype(np.int32)
    nc, na = len(cells), len(pos)
    return Database(natoms, box, types, pos, np.zeros((na, 3)), np.zeros(nc), np.zeros((nc, 6)),
                    np.ones(nc), np.ones(nc), np.ones(nc), 1.0, 1.0)


def workload(cfg: int, scale: float = 1.0, seed: Optional[int] = None, basis: int = 3) -> Workload:
    """Geometry and knot sets of workload ``cfg`` (targets still zero)."""
    seed = 1000 + cfg if seed is None else seed
    rng = np.random.Generator(np.random.PCG64(seed))
    cells = []
    if cfg == 0:
        ncell = max(1, int(round(16 * scale)))
        for _ in range(ncell):
            frac, box = _lattice("fcc", 4.05, (2, 2, 2))
            box = _strain_box(rng, box, 0.02, triclinic=False)
            cells.append((_displace(rng, frac, box, 0.1), box, np.zeros(len(frac), int)))
        nk = dict(phi=10, rho=10, U=8)
        pot = make_potential(rng, "eam/spline", 1, 5.5, nk, (0.0, 2.0), basis, rho_amp=0.08)
        truth = make_potential(rng, "eam/spline", 1, 5.5, nk, (0.0, 2.0), basis, rho_amp=0.08)
        return Workload("eam-fcc-16x32", pot, truth, _empty_db(cells), 5.5, 1, seed)
    if cfg == 1:
        ncell = max(2, int(round(256 * scale)))
        for c in range(ncell):
            if c < ncell // 2:
                frac, box = _lattice("bcc", 3.16, (4, 4, 4))
                sigma = rng.uniform(0.05, 0.3)
                frac = _displace(rng, frac, box, sigma)
            else:
                frac, box = _random_dense(rng, 128, 0.063, 2.0)
            cells.append((frac, box, np.zeros(len(frac), int)))
        nk = dict(phi=12, rho=12, U=10, f=12, g=10)
        pot = make_potential(rng, "meam/spline", 1, 5.0, nk, (-1.0, 3.0), basis)
        truth = make_potential(rng, "meam/spline", 1, 5.0, nk, (-1.0, 3.0), basis)
        return Workload("meam-256x128", pot, truth, _empty_db(cells), 5.0, 1, seed)
    if cfg in (2, 3, 4):
        nt = 2 if cfg in (2, 3) else 3
        rc = 5.0 if cfg in (2, 3) else 5.5
        ncell = max(2, int(round({2: 512, 3: 1000, 4: 2000}[cfg] * scale)))
        for c in range(ncell):
            if cfg == 2:
                target = math.exp(rng.uniform(math.log(8), math.log(512)))
            elif cfg == 3:
                target = rng.uniform(64, 256)
            else:
                target = rng.uniform(256, 2048)
            kind = rng.choice(["bcc", "fcc", "rnd"] if cfg == 3 else ["bcc", "fcc"])
            if kind == "rnd":
                frac, box = _random_dense(rng, int(round(target)), 0.06, 2.0)
            else:
                apc = 2 if kind == "bcc" else 4
                a = 3.16 if kind == "bcc" else 3.9
                frac, box = _lattice(kind, a, _reps_for(max(1, int(round(target / apc)))))
                box = _strain_box(rng, box, 0.03, triclinic=True)
                frac = _displace(rng, frac, box, rng.uniform(0.05, 0.2))
            if nt == 2:
                types = (rng.uniform(size=len(frac)) >= 0.5).astype(int)
            else:
                p = rng.dirichlet([1.0, 1.0, 1.0])
                types = rng.choice(3, size=len(frac), p=p)
            cells.append((frac, box, types))
        nk = dict(phi=12, rho=12, U=12, f=12, g=12)
        pot = make_potential(rng, "meam/spline", nt, rc, nk, (-1.0, 3.0), basis)
        truth = make_potential(rng, "meam/spline", nt, rc, nk, (-1.0, 3.0), basis)
        name = {2: "binary-meam-ga-512", 3: "binary-meam-adjoint-1000", 4: "ternary-meam-2000"}[cfg]
        cand = {2: 256, 3: 1, 4: 64}[cfg]
        return Workload(name, pot, truth, _empty_db(cells), rc, cand, seed)
    raise ValueError(f"unknown workload {cfg}")


Evaluator = Callable[[PotentialSpec, Database], Tuple[np.ndarray, np.ndarray, np.ndarray]]


def add_targets(w: Workload, evaluate: Evaluator) -> Workload:
    """Fill the database targets from the truth knot set plus noise.

    ``evaluate(pot, db)`` must return (per-atom cell energies (nc,), cell stresses
    (nc, 6) already divided by volume, forces (na, 3)).
    """
    e, s, f = evaluate(w.truth, w.db)
    rng = np.random.Generator(np.random.PCG64(w.seed + 104729))
    db = w.db
    db.energy0 = np.asarray(e) + rng.normal(0.0, NOISE_E, size=db.ncells)
    db.stress0 = np.asarray(s).reshape(-1, 6) + rng.normal(0.0, NOISE_S, size=(db.ncells, 6))
    db.force0 = np.asarray(f).reshape(-1, 3) + rng.normal(0.0, NOISE_F, size=(db.total_atoms, 3))
    return w
