"""Output files of a run: the host spline refresh behind the LAMMPS export
(mz_spline_refresh) against polynomial reproduction and against the LAMMPS
files the reference program wrote for three golden cases
(tests/golden/<case>/ref_out/, made by tests/golden/make_outputs.py), and the
driver's LAMMPS and data files against the reference's files (GPU)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from meamzilla_amd import _lib, formats

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["eam_spline_c0", "meam_spline_q5", "meam_spline_binary", "gmeam_spline_binary", "pair_spline_binary",
         "jmeam2_spline_binary"]
RNG = np.random.Generator(np.random.PCG64(31))


def _spl(basis, x0, xn, y, bc, grad=(0.0, 0.0), grad2=(0.0, 0.0)):
    return formats.Spline(basis, x0, xn, np.asarray(y, float), bc, grad, grad2)


def test_refresh_reproduces_cubic_polynomial():
    p = np.polynomial.Polynomial(RNG.uniform(-1, 1, 4))
    x = np.linspace(0.5, 2.5, 11)
    s = _spl(3, 0.5, 2.5, p(x), (1, 1), (p.deriv()(0.5), p.deriv()(2.5)))
    ypp, yp4 = _lib.spline_refresh(s)
    assert np.max(np.abs(ypp - p.deriv(2)(x))) < 1e-11
    assert not yp4.any()


def test_refresh_natural_ends_have_zero_curvature():
    s = _spl(3, 1.0, 4.0, RNG.uniform(-1, 1, 9), (-1, -1))
    ypp, _ = _lib.spline_refresh(s)
    assert ypp[0] == 0.0 and ypp[-1] == 0.0 and np.abs(ypp[1:-1]).max() > 0


def test_refresh_reproduces_quintic_polynomial():
    p = np.polynomial.Polynomial(RNG.uniform(-1, 1, 6))
    x = np.linspace(-1.0, 1.0, 10)
    d1, d2 = p.deriv(), p.deriv(2)
    s = _spl(5, -1.0, 1.0, p(x), (1, 1), (d1(-1.0), d1(1.0)), (d2(-1.0), d2(1.0)))
    ypp, yp4 = _lib.spline_refresh(s)
    assert np.max(np.abs(ypp - d2(x))) < 1e-9
    assert np.max(np.abs(yp4 - p.deriv(4)(x))) < 1e-7


def _lmp_blocks(path):
    """Parses a LAMMPS export: comment, style, then per spline basis, n,
    end slopes and n knot lines."""
    with open(path) as fh:
        lines = [ln.strip() for ln in fh if ln.strip()]
    out = []
    i = 2
    while i < len(lines):
        basis, n = lines[i], int(lines[i + 1])
        ends = np.array([float(v) for v in lines[i + 2].split()])
        knots = np.array([[float(v) for v in ln.split()] for ln in lines[i + 3:i + 3 + n]])
        out.append((basis, ends, knots))
        i += 3 + n
    return lines[0], lines[1], out


def _slope(s, ypp, yp4, k, t):
    h = (s.xn - s.x0) / (s.n - 1)
    a = 1.0 - t
    d = (s.y[k + 1] - s.y[k]) / h + ((3 * t * t - 1) * ypp[k + 1] - (3 * a * a - 1) * ypp[k]) * h / 6.0
    if s.basis == 5:
        d += ((15 * t ** 4 - 30 * t * t + 7) * yp4[k + 1] - (15 * a ** 4 - 30 * a * a + 7) * yp4[k]) * h ** 3 / 360
    return d


@pytest.mark.parametrize("case", CASES)
def test_refresh_matches_reference_lammps_file(case):
    pot = formats.read_potential(os.path.join(GOLDEN, case, "pot"))
    comment, style, blocks = _lmp_blocks(os.path.join(GOLDEN, case, "ref_out", "out.lmp"))
    assert style == pot.style and len(blocks) == len(pot.splines)
    for s, (basis, ends, knots) in zip(pot.splines, blocks):
        assert basis == ("spline3eq" if s.basis == 3 else "spline5eq") and len(knots) == s.n
        ypp, yp4 = _lib.spline_refresh(s)
        scale = max(np.abs(knots[:, 2]).max(), 1.0)
        assert np.max(np.abs(knots[:, 0] - np.linspace(s.x0, s.xn, s.n))) < 1e-14 * max(abs(s.xn), 1.0)
        assert np.max(np.abs(knots[:, 1] - s.y)) <= 1e-15 * max(np.abs(s.y).max(), 1e-300)  # 16 digits
        assert np.max(np.abs(ypp - knots[:, 2])) < 1e-12 * scale
        if s.basis == 5:
            assert np.max(np.abs(yp4 - knots[:, 3])) < 1e-11 * max(np.abs(knots[:, 3]).max(), 1.0)
        slopes = [_slope(s, ypp, yp4, 0, 0.0), _slope(s, ypp, yp4, s.n - 2, 1.0)]
        assert np.max(np.abs(np.array(slopes) - ends)) < 1e-12 * max(np.abs(ends).max(), 1.0)


def _tokens(path):
    with open(path) as fh:
        return [ln.split() for ln in fh]


def _num(tok):
    try:
        v = float(tok)
    except ValueError:
        return None
    return v if np.isfinite(v) else None  # "inf" columns compare as text


def _same_file(ours, ref):
    """Token by token: text equal, numbers within one unit of the last printed
    digit (both sides round the same double to the same width) or 1e-12
    relative for the 15-digit scientific LAMMPS values."""
    a, b = _tokens(ours), _tokens(ref)
    assert len(a) == len(b), (ours, len(a), len(b))
    for la, lb in zip(a, b):
        assert len(la) == len(lb), (ours, la, lb)
        for ta, tb in zip(la, lb):
            va, vb = _num(ta), _num(tb)
            if va is None or vb is None:
                assert ta == tb, (ours, la, lb)
                continue
            if "e" in tb:
                tol = 1e-12 * max(abs(vb), 1e-300) + 1e-300
            else:
                dec = len(tb.split(".")[1]) if "." in tb else 0
                tol = 1.01 * 10.0 ** (-dec) + 1e-12 * abs(vb)
            assert abs(va - vb) <= tol, (ours, la, lb)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_driver_output_files_match_reference(case, tmp_path):
    driver = os.path.join(os.path.dirname(GOLDEN), "..", "meamzilla_amd", "meamz_amd")
    src = os.path.join(GOLDEN, case)
    for fn in ("config", "pot", "param"):
        shutil.copy(os.path.join(src, fn), tmp_path)
    with open(tmp_path / "param", "a") as fh:
        fh.write("datafile out\nlammpsfile out.lmp\n")
    print("driver outputs", case, flush=True)
    run = subprocess.run([driver, "-p", str(tmp_path / "param")], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr[-2000:]
    for suf in (".energy", ".stress", ".force", ".error", ".punish", ".rho", ".lmp"):
        ref = os.path.join(src, "ref_out", "out" + suf)
        if not os.path.exists(ref):  # pair styles write no densities or constraints
            assert not os.path.exists(tmp_path / ("out" + suf)), suf
            continue
        _same_file(str(tmp_path / ("out" + suf)), ref)
