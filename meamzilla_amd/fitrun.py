"""Fit runs of the host driver on a generated workload.

    python -m meamzilla_amd.fitrun --workload 3 --lbfgs-steps 200 --powell-steps 2 --out DIR

writes the workload as files in the reference's formats (configuration
database with targets from the truth knot set, start potential with one knot
of every rho and f spline fixed to remove the objective's scale gauges, one
parameter file per optimizer), runs ``meamz_amd -p <param>`` for L-BFGS (adjoint
gradient) and for the least-squares optimizer of optstyle powell
(Levenberg-Marquardt on batched forward-difference Jacobians) on the same
objective, and prints one JSON line with each optimizer's objective
trajectory, calls and wall time.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import time

from . import formats, gen
from ._lib import Objective

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "meamz_amd")


def _evaluator(pot, db):
    r = Objective(pot, db).compute(errvec=False)
    return r["cell_energy"], r["cell_stress"], r["forces"]


def fix_gauge(pot) -> None:
    """Fixes the largest-magnitude knot of every rho and f spline.

    The MEAM objective is invariant under rho -> a rho (with U rescaled in
    its argument) and f -> b f (with g -> g / b^2); anchoring one knot per rho
    and f spline removes those flat directions, which slow L-BFGS and leave a
    Gauss-Newton system singular (the least-squares optimizer's damping
    copes with them, but the fit runs compare optimizers on one objective)."""
    counts = formats.function_counts(pot.style, pot.ntypes)
    off = [sum(counts[:i]) for i in range(len(counts))]
    idx = list(range(off[1], off[1] + counts[1])) if len(counts) > 1 else []
    if len(counts) > 3:
        idx += list(range(off[3], off[3] + counts[3]))
    for i in idx:
        s = pot.splines[i]
        s.y_fixed[int(abs(s.y).argmax())] = 1


def _parse(text: str) -> dict:
    out = {"initial_F": None, "final_F": None, "calls": None, "seconds": None, "trajectory": []}
    m = re.search(r"Initial total sum error = ([0-9.eE+-]+)", text)
    if m:
        out["initial_F"] = float(m.group(1))
    m = re.search(r"Final total sum error = ([0-9.eE+-]+)", text)
    if m:
        out["final_F"] = float(m.group(1))
    m = re.search(r"optimizer \w+: (\d+) objective calls, ([0-9.]+) s", text)
    if m:
        out["calls"], out["seconds"] = int(m.group(1)), float(m.group(2))
    for line in text.splitlines():
        f = line.split()
        # L-BFGS: iter F |g| step seconds; least squares: iter F lambda calls seconds
        if len(f) == 5 and f[0].isdigit():
            out["trajectory"].append({"iter": int(f[0]), "F": float(f[1]), "seconds": float(f[4])})
        # least squares: iter F lambda calls seconds (parsed by the branch above)
        # genetic algorithm: generation g  best F = x
        elif len(f) == 6 and f[0] == "generation":
            out["trajectory"].append({"generation": int(f[1]), "F": float(f[5])})
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", type=int, default=3)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--lbfgs-steps", type=int, default=200)
    ap.add_argument("--powell-steps", type=int, default=2)
    ap.add_argument("--styles", default="lbfgs,powell", help="optimizers to run, comma separated "
                    "(lbfgs, powell, simann, genalg)")
    ap.add_argument("--ga-pop", type=int, default=256, help="genetic algorithm population")
    ap.add_argument("--ga-gens", type=int, default=5, help="genetic algorithm generations")
    ap.add_argument("--ga-init-scale", type=float, default=0.1,
                    help="relative size of the random knot changes of the initial population")
    ap.add_argument("--out", required=True)
    ap.add_argument("--free-gauge", action="store_true",
                    help="keep every knot free (the objective then has flat gauge directions)")
    args = ap.parse_args(argv)
    if not os.path.exists(DRIVER):
        raise SystemExit(f"{DRIVER} is not built")
    os.makedirs(args.out, exist_ok=True)
    t0 = time.time()
    w = gen.workload(args.workload, scale=args.scale)
    gen.add_targets(w, _evaluator)
    if not args.free_gauge:
        fix_gauge(w.pot)
    formats.write_config(os.path.join(args.out, "config"), w.db)
    formats.write_potential(os.path.join(args.out, "pot"), w.pot)
    print(f"workload {w.name}: {w.db.ncells} cells, {w.db.total_atoms} atoms, written in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    result = {"workload": w.name, "cells": w.db.ncells, "atoms": w.db.total_atoms, "ncoeff": w.pot.ncoeff(),
              "gauge": "free" if args.free_gauge else "one knot of every rho and f spline fixed"}
    steps_of = {"lbfgs": args.lbfgs_steps, "powell": args.powell_steps, "simann": args.lbfgs_steps,
                "genalg": args.ga_gens}
    for style in args.styles.split(","):
        steps = steps_of[style]
        extra = [f"optstyle {style}", f"max_steps {steps}"]
        if style == "genalg":
            extra += [f"pop_size {args.ga_pop}", f"init_scale {args.ga_init_scale}", "cross_rate 0.5",
                      "mut_rate 0.05", "seed 1"]
        param = os.path.join(args.out, f"param_{style}")
        formats.write_param(param, "config", "pot", extra=extra)
        ts = time.time()
        run = subprocess.run([DRIVER, "-p", param], capture_output=True, text=True)
        with open(os.path.join(args.out, f"{style}.log"), "w") as fh:
            fh.write(run.stdout + run.stderr)
        if run.returncode != 0:
            raise SystemExit(f"{style} run failed: {run.stderr[-400:]}")
        r = _parse(run.stdout)
        r["wall_seconds"] = time.time() - ts
        result[style] = r
        print(f"{style}: F {r['initial_F']} -> {r['final_F']} in {r['seconds']} s, {r['calls']} calls",
              file=sys.stderr, flush=True)
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
