"""Readers and writers for the three text formats the MEAM/EAM fitting program
consumes: the parameter file, the configuration database and the potential
file.  Layouts follow the reference readers:

* parameter file: ``key value...`` lines, blank lines dropped, first match wins
  (src/universe.cpp:138-165, src/input.h:47-77);
* configuration file: ``#N natoms useforce`` / ``#X #Y #Z #E #S #W #EW #SW`` /
  ``#F`` / ``type x y z fx fy fz`` blocks (src/cell.cpp:39-136);
* potential file: comment line, ``#T style ntypes``, then ``#P`` blocks holding
  every function's alloy members as ``#B spline3eq|spline5eq`` splines
  (src/pot_list.cpp:84-142, src/spline3eq.cpp:54-106, src/spline5eq.cpp:57-116).

Numbers are written with 17 significant digits so a round trip is exact.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence

import numpy as np

STYLES = ("eam/spline", "meam/spline", "eam", "meam", "gmeam/spline", "gmeam", "pair/spline", "pair",
          "jmeam/spline", "jmeam", "jmeam2/spline", "jmeam2")


def n_pair_ij(nt: int) -> int:
    return nt * (nt + 1) // 2


def is_meam_style(style: str) -> bool:
    """Styles with the f and g (three-body) functions."""
    return style.startswith(("meam", "gmeam", "jmeam"))


def jmeam_terms(style: str) -> int:
    """Extra p p q energy terms: 1 for jmeam, 2 for jmeam2
    (src/pots/pot_jmeam/pot_jmeam.cpp:30-47, src/pots/pot_jmeam2/pot_jmeam2.cpp:30-46)."""
    return 2 if style.startswith("jmeam2") else (1 if style.startswith("jmeam") else 0)


def is_pair_style(style: str) -> bool:
    """Pair-only styles: phi alone, no density or embedding
    (src/pots/pot_pair/pot_pair.cpp:24-37)."""
    return style.startswith("pair")


def is_gmeam_style(style: str) -> bool:
    """Styles whose g splines are indexed by (type i, types j k) (Triplet_i_jk)."""
    return style.startswith("gmeam")


def function_counts(style: str, nt: int) -> List[int]:
    """Alloy members per function in file order (src/pot_fns.cpp:280-310):
    phi Pair_ij, rho Pair_j, U Pair_i, f Pair_ij, g Triplet_i (gmeam styles:
    Triplet_i_jk, nt * nt (nt + 1) / 2 members, src/pots/pot_gmeam/pot_gmeam.cpp:31-36;
    pair styles: phi only)."""
    if is_pair_style(style):
        return [n_pair_ij(nt)]
    counts = [n_pair_ij(nt), nt, nt]
    if is_meam_style(style):
        counts += [n_pair_ij(nt), nt * n_pair_ij(nt) if is_gmeam_style(style) else nt]
    counts += [n_pair_ij(nt), nt] * jmeam_terms(style)  # p Pair_ij, q Triplet_i (, p2, q2)
    return counts


def pair_ij_index(ti: int, tj: int, nt: int) -> int:
    """Symmetric two-body alloy index (src/pot_fns.cpp:141-143)."""
    if ti <= tj:
        return ti * nt + tj - (ti * (ti + 1)) // 2
    return tj * nt + ti - (tj * (tj + 1)) // 2


@dataclasses.dataclass
class Spline:
    basis: int                      # 3 (spline3eq) or 5 (spline5eq)
    x0: float
    xn: float
    y: np.ndarray                   # knot values, length n
    bc: Sequence[int] = (-1, -1)    # -1 natural, 0 free slope, 1 fixed slope
    grad: Sequence[float] = (0.0, 0.0)
    grad2: Sequence[float] = (0.0, 0.0)
    y_fixed: Optional[np.ndarray] = None

    def __post_init__(self):
        self.y = np.asarray(self.y, dtype=np.float64)
        if self.y_fixed is None:
            self.y_fixed = np.zeros(len(self.y), dtype=np.int32)
        self.y_fixed = np.asarray(self.y_fixed, dtype=np.int32)
        self.bc = tuple(int(b) for b in self.bc)
        self.grad = tuple(float(g) for g in self.grad)
        self.grad2 = tuple(float(g) for g in self.grad2)

    @property
    def n(self) -> int:
        return len(self.y)

    def ncoeff(self) -> int:
        """Adjustable coefficients as counted by read_basis: free knots plus one
        per free (bc == 0) side (src/spline3eq.cpp:77-100)."""
        return int((self.y_fixed == 0).sum()) + sum(1 for b in self.bc if b == 0)

    def knot_record(self) -> np.ndarray:
        """[y_0..y_{n-1}, grad_L, grad_R, grad2_L, grad2_R]"""
        return np.concatenate([self.y, np.array(self.grad + self.grad2)])


@dataclasses.dataclass
class PotentialSpec:
    style: str
    ntypes: int
    splines: List[Spline]
    embed_extrap: int = 0

    def __post_init__(self):
        if self.style not in STYLES:
            raise ValueError(f"unsupported style {self.style}")
        if len(self.splines) != sum(function_counts(self.style, self.ntypes)):
            raise ValueError("wrong number of splines for style/ntypes")

    def knot_vector(self) -> np.ndarray:
        return np.concatenate([s.knot_record() for s in self.splines])

    def set_knot_vector(self, kv: np.ndarray) -> None:
        off = 0
        for s in self.splines:
            n = s.n
            s.y = np.array(kv[off:off + n], dtype=np.float64)
            s.grad = (float(kv[off + n]), float(kv[off + n + 1]))
            s.grad2 = (float(kv[off + n + 2]), float(kv[off + n + 3]))
            off += n + 4

    def max_rcut(self) -> float:
        counts = function_counts(self.style, self.ntypes)
        radial = [0] if is_pair_style(self.style) else [0, 1] + ([3] if is_meam_style(self.style) else [])
        radial += [5, 7][:jmeam_terms(self.style)]
        r = 0.0
        off = 0
        for fi, c in enumerate(counts):
            if fi in radial:
                r = max(r, max(s.xn for s in self.splines[off:off + c]))
            off += c
        return r

    def ncoeff(self) -> int:
        return sum(s.ncoeff() for s in self.splines)

    def coeff_slots(self) -> np.ndarray:
        """Knot-vector index of every adjustable coefficient, in the order the
        reference's Potential::coeff(i) walks them (src/potential.cpp:192-206,
        src/pot_fns.cpp:209-218, src/spline.cpp:252-261): per spline its free
        knots, then its free boundary values -- one slot per free side for cubic
        splines (src/spline3eq.cpp:376-382), slope then second derivative for
        quintic ones (src/spline5eq.cpp:461-471).  The walk is global while each
        quintic free side counts once in ncoeff, so the list is the first
        ncoeff entries of the concatenated walk."""
        walk = []
        off = 0
        for s in self.splines:
            n = s.n
            walk += [off + k for k in range(n) if not s.y_fixed[k]]
            for side in (0, 1):
                if s.bc[side] == 0:
                    walk.append(off + n + side)
                    if s.basis == 5:
                        walk.append(off + n + 2 + side)
            off += n + 4
        return np.array(walk[: self.ncoeff()], dtype=np.int64)

    def copy(self) -> "PotentialSpec":
        return PotentialSpec(self.style, self.ntypes,
                             [dataclasses.replace(s, y=s.y.copy(), y_fixed=s.y_fixed.copy())
                              for s in self.splines], self.embed_extrap)


@dataclasses.dataclass
class Database:
    natoms: np.ndarray          # (ncells,) int
    box: np.ndarray             # (ncells, 3, 3) rows = lattice vectors a0, a1, a2
    types: np.ndarray           # (natoms_total,) int
    pos: np.ndarray             # (natoms_total, 3)
    force0: np.ndarray          # (natoms_total, 3)
    energy0: np.ndarray         # (ncells,) per-atom target energy
    stress0: np.ndarray         # (ncells, 6) xx yy zz xy yz zx
    weight: np.ndarray          # (ncells,)
    eweight: np.ndarray
    sweight: np.ndarray
    energy_weight: float = 1.0
    stress_weight: float = 1.0

    @property
    def ncells(self) -> int:
        return len(self.natoms)

    @property
    def total_atoms(self) -> int:
        return int(self.natoms.sum())

    def cell_offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.natoms)]).astype(np.int64)

    def subset(self, cells: Sequence[int]) -> "Database":
        off = self.cell_offsets()
        idx = np.concatenate([np.arange(off[c], off[c + 1]) for c in cells]) if len(cells) else np.zeros(0, int)
        cells = np.asarray(cells, dtype=np.int64)
        return Database(self.natoms[cells].copy(), self.box[cells].copy(), self.types[idx].copy(),
                        self.pos[idx].copy(), self.force0[idx].copy(), self.energy0[cells].copy(),
                        self.stress0[cells].copy(), self.weight[cells].copy(), self.eweight[cells].copy(),
                        self.sweight[cells].copy(), self.energy_weight, self.stress_weight)


def _g(x: float) -> str:
    return repr(float(x))


def write_potential(path: str, pot: PotentialSpec, comment: str = "# potential") -> None:
    lines = [comment, f"#T {pot.style} {pot.ntypes}", "#P"]
    for s in pot.splines:
        lines.append("#B spline3eq" if s.basis == 3 else "#B spline5eq")
        lines.append(f"{_g(s.x0)} {_g(s.xn)} {s.n}")
        lines.append(f"{s.bc[0]} {s.bc[1]}")
        lines.append(f"{_g(s.grad[0])} {_g(s.grad[1])}")
        if s.basis == 5:
            lines.append(f"{_g(s.grad2[0])} {_g(s.grad2[1])}")
        for fx, yv in zip(s.y_fixed, s.y):
            lines.append(f"{int(fx)} {_g(yv)}")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def _content_lines(path: str) -> List[str]:
    out = []
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\n").replace("\r", "")
            if line.strip():
                out.append(line)
    return out


def read_potential(path: str, embed_extrap: int = 0) -> PotentialSpec:
    """Reads the first #P block of a potential file."""
    lines = _content_lines(path)
    tok = lines[1].split()
    if tok[0] != "#T":
        raise ValueError("potential type line (#T ...) missing")
    style, nt = tok[1], int(tok[2])
    if not lines[2].startswith("#P"):
        raise ValueError("potential line (#P ...) missing")
    ln = 3
    splines = []
    for _ in range(sum(function_counts(style, nt))):
        head = lines[ln].split()
        if head[0] != "#B":
            raise ValueError("potential line (#B ...) missing")
        basis = 3 if head[1] == "spline3eq" else 5 if head[1] == "spline5eq" else None
        if basis is None:
            raise ValueError(f"invalid basis type {head[1]}")
        x0, xn, n = lines[ln + 1].split()[:3]
        bc = [int(v) for v in lines[ln + 2].split()[:2]]
        grad = [float(v) for v in lines[ln + 3].split()[:2]]
        ln += 4
        grad2 = [0.0, 0.0]
        if basis == 5:
            grad2 = [float(v) for v in lines[ln].split()[:2]]
            ln += 1
        n = int(n)
        fixed, y = [], []
        for k in range(n):
            a, b = lines[ln + k].split()[:2]
            fixed.append(int(a))
            y.append(float(b))
        ln += n
        splines.append(Spline(basis, float(x0), float(xn), np.array(y), bc, grad, grad2, np.array(fixed)))
    return PotentialSpec(style, nt, splines, embed_extrap)


def write_config(path: str, db: Database) -> None:
    off = db.cell_offsets()
    out = []
    for c in range(db.ncells):
        out.append(f"#N {int(db.natoms[c])} 1")
        for lab, v in zip(("#X", "#Y", "#Z"), db.box[c]):
            out.append(f"{lab} {_g(v[0])} {_g(v[1])} {_g(v[2])}")
        out.append(f"#E {_g(db.energy0[c])}")
        out.append("#S " + " ".join(_g(v) for v in db.stress0[c]))
        if db.weight[c] != 1.0:
            out.append(f"#W {_g(db.weight[c])}")
        if db.eweight[c] != 1.0:
            out.append(f"#EW {_g(db.eweight[c])}")
        if db.sweight[c] != 1.0:
            out.append(f"#SW {_g(db.sweight[c])}")
        out.append("#F")
        for a in range(off[c], off[c + 1]):
            p, f = db.pos[a], db.force0[a]
            out.append(f"{int(db.types[a])} {_g(p[0])} {_g(p[1])} {_g(p[2])} {_g(f[0])} {_g(f[1])} {_g(f[2])}")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def read_config(path: str, energy_weight: float = 1.0, stress_weight: float = 1.0) -> Database:
    lines = _content_lines(path)
    natoms, box, types, pos, f0, e0, s0, w, ew, sw = [], [], [], [], [], [], [], [], [], []
    ln = 0
    while ln < len(lines):
        tok = lines[ln].split()
        if tok[0] != "#N":
            raise ValueError("number line (#N ...) is missing")
        n = int(tok[1])
        b = np.zeros((3, 3))
        cell = dict(E=0.0, S=np.zeros(6), W=1.0, EW=1.0, SW=1.0)
        ln += 1
        while not lines[ln].startswith("#F"):
            t = lines[ln].split()
            if t[0] in ("#X", "#Y", "#Z"):
                b["#X#Y#Z".index(t[0]) // 2] = [float(v) for v in t[1:4]]
            elif t[0] == "#E":
                cell["E"] = float(t[1])
            elif t[0] == "#S":
                cell["S"] = np.array([float(v) for v in t[1:7]])
            elif t[0] in ("#W", "#EW", "#SW"):
                cell[t[0][1:]] = float(t[1])
            ln += 1
        for k in range(n):
            t = lines[ln + 1 + k].split()
            types.append(int(t[0]))
            pos.append([float(v) for v in t[1:4]])
            f0.append([float(v) for v in t[4:7]])
        ln += n + 1
        natoms.append(n)
        box.append(b)
        e0.append(cell["E"])
        s0.append(cell["S"])
        w.append(cell["W"])
        ew.append(cell["EW"])
        sw.append(cell["SW"])
    return Database(np.array(natoms, dtype=np.int32), np.array(box), np.array(types, dtype=np.int32),
                    np.array(pos).reshape(-1, 3), np.array(f0).reshape(-1, 3), np.array(e0), np.array(s0),
                    np.array(w), np.array(ew), np.array(sw), energy_weight, stress_weight)


def write_param(path: str, config_file: str, pot_file: str, energy_weight: float = 1.0,
                stress_weight: float = 1.0, embed_extrap: int = 0, extra: Sequence[str] = ()) -> None:
    lines = [f"config {config_file}", f"startpot {pot_file}", f"tempfile {pot_file}.tmp",
             f"endpot {pot_file}.end", f"energy_weight {_g(energy_weight)}",
             f"stress_weight {_g(stress_weight)}", f"embed_extrap {int(embed_extrap)}"] + list(extra)
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def read_param(path: str) -> dict:
    out = {}
    for line in _content_lines(path):
        tok = line.split()
        if tok and tok[0] not in out:
            out[tok[0]] = tok[1:]
    return out
