"""The C++ host driver (meamz_amd): file readers and coefficient walk on the CPU,
end-to-end fits through the C-ABI on the GPU."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, golden_cases, load_case

EXE = os.path.join(ROOT, "meamzilla_amd", "meamz_amd")


@pytest.mark.parametrize("case", golden_cases())
def test_driver_reads_files_and_walks_coefficients(case):
    pot, db, exp = load_case(case)
    out = subprocess.run([EXE, "--check", os.path.join(GOLDEN, case, "param")], capture_output=True, text=True,
                         check=True).stdout
    head, coeffs = out.strip().split("\n")
    assert f"ncoeff {exp['ncoeff']} " in head and f"atoms {db.total_atoms}" in head
    vals = np.array([float(v) for v in coeffs.split()[1:]])
    assert np.array_equal(vals, np.array(exp["coeffs"]))


def _run_fit(tmp_path, case, optstyle, extra=()):
    d = tmp_path / case
    shutil.copytree(os.path.join(GOLDEN, case), d)
    with open(d / "param", "a") as fh:
        fh.write(f"optstyle {optstyle}\n")
        for e in extra:
            fh.write(e + "\n")
    print("driver", case, optstyle, flush=True)
    res = subprocess.run([EXE, "-p", str(d / "param")], capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stderr
    ini = [float(x) for x in re.findall(r"Initial total sum error = ([0-9.eE+-]+)", res.stdout)]
    fin = [float(x) for x in re.findall(r"Final total sum error = ([0-9.eE+-]+)", res.stdout)]
    return ini, fin, res.stdout, d


@pytest.mark.gpu
def test_driver_lbfgs_decreases_objective(tmp_path):
    pot, db, exp = load_case("meam_spline_c1")
    ini, fin, out, d = _run_fit(tmp_path, "meam_spline_c1", "lbfgs", ["max_steps 40"])
    assert abs(ini[0] - exp["F"]) <= 1e-6 * exp["F"]   # printed with 6 decimals
    assert fin[0] < 0.5 * ini[0]
    assert os.path.exists(d / "pot.end")


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["eam_spline_c0", "meam_spline_c1"])
def test_driver_least_squares_fits_every_coefficient(case, tmp_path):
    """optstyle powell runs the Levenberg-Marquardt least squares with every
    coefficient free: the damping keeps the normal equations regular along
    the rho -> a rho (U rescaled) gauge that leaves a Gauss-Newton system
    singular, and each iteration lowers the objective."""
    ini, fin, out, d = _run_fit(tmp_path, case, "powell", ["max_steps 5"])
    rows = [float(f.split()[1]) for f in out.splitlines() if re.match(r"^\s+\d+\s+[0-9.]+e[+-]\d+\s", f)]
    assert rows and all(b < a for a, b in zip([ini[0]] + rows, rows))
    assert fin[0] < 0.5 * ini[0]
    assert os.path.exists(d / "pot.end")


@pytest.mark.gpu
def test_driver_genalg_runs_batched(tmp_path):
    ini, fin, out, d = _run_fit(tmp_path, "eam_spline_c0", "genalg", ["max_steps 3", "pop_size 8", "seed 5"])
    assert "generation 3" in out and fin[0] <= ini[0] * 10


@pytest.mark.gpu
def test_driver_simann_keeps_the_best(tmp_path):
    """One temperature of simulated annealing (2 ncoeff step adjustments of 20
    sweeps of Gaussian-bump moves, scored in batched calls): the returned
    potential is the best seen, so its objective is at most the start's."""
    ini, fin, out, d = _run_fit(tmp_path, "eam_spline_c0", "simann", ["max_steps 1", "seed 3", "anneal_temp 1.0"])
    assert "T=" in out and fin[0] <= ini[0]
    best = float(re.search(r"best F=([0-9.eE+-]+)", out).group(1))
    assert abs(best - fin[0]) <= 1e-6 * fin[0] + 1e-6
    assert os.path.exists(d / "pot.end")


@pytest.mark.gpu
def test_fitrun_lbfgs_and_powell(tmp_path):
    """End-to-end fit run: files in the reference's formats, the C++ driver,
    L-BFGS on the adjoint gradient and the least-squares optimizer on the same
    objective."""
    import io
    import json
    from contextlib import redirect_stdout
    from meamzilla_amd import fitrun
    buf = io.StringIO()
    with redirect_stdout(buf):
        assert fitrun.main(["--workload", "1", "--scale", str(6 / 256), "--lbfgs-steps", "8", "--powell-steps", "1",
                            "--out", str(tmp_path)]) == 0
    r = json.loads(buf.getvalue().strip().splitlines()[-1])
    for style in ("lbfgs", "powell"):
        assert r[style]["final_F"] < r[style]["initial_F"], (style, r[style])
        assert r[style]["calls"] > 0 and r[style]["seconds"] >= 0
    assert len(r["lbfgs"]["trajectory"]) >= 1


@pytest.mark.gpu
def test_fitrun_binary_adjoint_workload_lbfgs_200_beside_least_squares(tmp_path):
    """The 1000-cell binary adjoint workload in full (160 k atoms, 139 free
    coefficients): 200 L-BFGS iterations on the adjoint gradient beside three
    least-squares iterations on batched Jacobians, both in the C++ driver on the
    same files; each lowers the objective by more than half."""
    import io
    import json
    from contextlib import redirect_stdout
    from meamzilla_amd import fitrun
    print("fit run: binary adjoint workload, L-BFGS 200 beside least squares", flush=True)
    buf = io.StringIO()
    with redirect_stdout(buf):
        assert fitrun.main(["--workload", "3", "--lbfgs-steps", "200", "--powell-steps", "3",
                            "--out", str(tmp_path)]) == 0
    r = json.loads(buf.getvalue().strip().splitlines()[-1])
    assert r["cells"] == 1000 and r["atoms"] > 150000
    print({s: (r[s]["initial_F"], r[s]["final_F"], r[s]["calls"], r[s]["seconds"]) for s in ("lbfgs", "powell")},
          flush=True)
    for style in ("lbfgs", "powell"):
        assert r[style]["final_F"] < 0.5 * r[style]["initial_F"], (style, r[style])
    iters = [t["iter"] for t in r["lbfgs"]["trajectory"] if "iter" in t]
    assert max(iters) == 200
