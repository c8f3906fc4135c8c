"""The reference fitting program itself, compiled from its own sources with
the GPU binding of integration/ (its meam/spline and eam/spline potentials
replaced by copies whose Potential::compute and compute_densities call
libmeamz): one evaluation through the reference's virtual compute equals the
reference's CPU results of the golden cases at 1e-10, and the reference's own
Powell optimizer (src/opt_powell.cpp:96-315) runs unchanged on top and lowers
the objective.  The binary is built by __graft_entry__.build() where the
reference tree is present and travels with the repository."""
import json
import os
import re
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT, golden_cases, load_case, rel_err

BIN = os.path.join(ROOT, "integration", "_build", "meamz_gpu")
MPIEXEC = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
TOL = 1e-10

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(BIN), reason="binding not built: it needs the reference tree "
                                                                 "at build time")]


def _style(case):
    pot, _, _ = load_case(case)
    return pot.style


CASES = [c for c in golden_cases() if _style(c) in ("meam/spline", "eam/spline")]


def _run(cwd, *args, timeout=300):
    out = subprocess.run([MPIEXEC, "-n", "1", BIN, "-p", "param", *args], cwd=cwd, capture_output=True, text=True,
                         timeout=timeout)
    assert out.returncode == 0, out.stderr[-2000:] + out.stdout[-2000:]
    return out.stdout


def _copy(case, tmp_path):
    d = os.path.join(str(tmp_path), case)
    shutil.copytree(os.path.join(GOLDEN, case), d, ignore=shutil.ignore_patterns("expected.json", "ref_out"))
    return d


@pytest.mark.parametrize("case", CASES)
def test_reference_compute_through_binding_matches_reference(case, tmp_path):
    print("binding eval", case, flush=True)
    _, _, exp = load_case(case)
    txt = _run(_copy(case, tmp_path), "--eval")
    got = json.loads(txt[txt.index('{"pots"'):])["pots"][0]
    assert abs(got["F"] - exp["F"]) <= TOL * abs(exp["F"])
    for k in ("errvec", "cell_energy", "cell_stress", "forces", "densities"):
        assert rel_err(got[k], exp[k]) < TOL, k


def test_reference_powell_runs_on_the_gpu_objective(tmp_path):
    print("binding powell meam_spline_c1", flush=True)
    d = _copy("meam_spline_c1", tmp_path)
    with open(os.path.join(d, "param"), "a") as fh:
        fh.write("optstyle powell\nmax_steps 10\n")
    out = _run(d)
    ini = float(re.search(r"Initial total sum error = ([0-9.eE+-]+)", out).group(1))
    fin = float(re.search(r"Final total sum error = ([0-9.eE+-]+)", out).group(1))
    print(f"reference Powell on the GPU objective: F {ini} -> {fin}", flush=True)
    assert fin < ini
    # the reference's own throughput line, printed every tenth outer step
    # (src/opt_powell.cpp:294-306)
    steps = [int(m.group(1)) for m in re.finditer(r"P#0 -\s+(\d+)\s+\d+\s+[0-9.eE+-]+", out)]
    if max(steps) >= 10:
        rate = float(re.search(r"calls/sec = ([0-9.eE+-]+)", out).group(1))
        print(f"reference Powell: {rate:.1f} objective calls/s on the GPU objective", flush=True)
    assert os.path.exists(os.path.join(d, "pot.end"))
