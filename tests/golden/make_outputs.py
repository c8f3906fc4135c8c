"""Writes the reference program's output files for a few golden cases.

For each case below, the case's parameter file gains ``datafile out`` and
``lammpsfile out.lmp`` and the reference program (built from its own sources
by oracle/ref/Makefile into oracle/_ref/meamz_ref) runs it with ``--write``,
which calls PotList::write_lmps and PotList::write_data as the program does at
the end of a run (src/meamzilla.cpp:149-150).  The files it writes
(src/spline3eq.cpp:279-295, src/spline5eq.cpp:358-377, src/potential.cpp:497-707,
src/pots/pot_eam/pot_eam.cpp:439-543, src/pots/pot_meam/pot_meam.cpp:444-490)
are committed under tests/golden/<case>/ref_out/; only text output is kept.

Run from the repository root in the build container:
    python tests/golden/make_outputs.py
"""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF = os.path.join(ROOT, "oracle", "_ref", "meamz_ref")
CASES = ["eam_spline_c0", "meam_spline_q5", "meam_spline_binary", "gmeam_spline_binary", "pair_spline_binary",
         "jmeam2_spline_binary"]
SUFFIXES = [".energy", ".stress", ".force", ".error", ".punish", ".rho", ".lmp"]


def main() -> None:
    for case in CASES:
        src = os.path.join(GOLDEN, case)
        with tempfile.TemporaryDirectory() as tmp:
            for fn in ("config", "pot"):
                shutil.copy(os.path.join(src, fn), tmp)
            with open(os.path.join(src, "param")) as fh:
                lines = fh.read().splitlines()
            lines += ["datafile out", "lammpsfile out.lmp"]
            with open(os.path.join(tmp, "param"), "w") as fh:
                fh.write("\n".join(lines) + "\n")
            subprocess.run([REF, "-p", "param", "--write"], cwd=tmp, check=True, capture_output=True)
            dst = os.path.join(src, "ref_out")
            os.makedirs(dst, exist_ok=True)
            for suf in SUFFIXES:
                p = os.path.join(tmp, "out" + suf)
                if os.path.exists(p):
                    shutil.copy(p, os.path.join(dst, "out" + suf))
            print(case, sorted(os.listdir(dst)))


if __name__ == "__main__":
    main()
