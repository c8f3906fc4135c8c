#!/usr/bin/env python3
"""Throughput of the MEAM objective (and gradient) on one MI355X.

Workload: 256 periodic cells of 128 atoms (half displaced BCC 4x4x4 at
a=3.16 A, half random-dense at 0.063 /A^3), one-element meam/spline with
phi/rho/f 12 knots on [1.8,5.0], U 10 knots on [-1,3], g 10 knots on [-1,1];
cutoff 5.0 A; targets from a seeded truth knot set plus noise.  Synthetic,
seeded geometry and knots; FP64 throughout.

One step = one objective-and-gradient evaluation of the selected mode:
  adjoint : objective + adjoint gradient of one knot vector (mz_objective_grad)
  batch   : objective + adjoint gradient of --cand knot vectors in one call
  cd      : objective + central-difference gradient (2*ncoeff+1 plain evaluations)
The JSON line reports whole-job evals/s (objective+gradient evaluations per
second), the atom kernel's roofline and a CPU baseline timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from meamzilla_amd import formats, gen  # noqa: E402
from meamzilla_amd._lib import Objective  # noqa: E402

FLOP_PER_TRIPLET = 113      # objective, per j<k triplet (cos, g, density, three-body force/virial)
FLOP_PER_PAIR_MEAM = 148    # objective, per full-list pair (phi, rho, f splines, pair force/virial)
BYTES_PER_PAIR = 36         # j (int32) + r + unit vector (4 x fp64)
BYTES_PER_ATOM = 52         # type + force out + target force
BYTES_PER_CELL = 136        # targets, weights, energy/stress out
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gpu_evaluator(pot, db):
    r = Objective(pot, db).compute(errvec=False)
    return r["cell_energy"], r["cell_stress"], r["forces"]


def cpu_baseline(w, seconds: float) -> dict:
    """The reference program (built from its sources into oracle/_ref) run with
    MPI ranks on this host's cores, timed on the same workload; the oracle's
    threaded restatement if MPI or the build is unavailable."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    ref = os.path.join(ROOT, "oracle", "_ref", "meamz_ref")
    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    if os.path.exists(ref) and os.path.exists(mpiexec):
        tmp = tempfile.mkdtemp(prefix="meamz_cpu_")
        try:
            formats.write_config(os.path.join(tmp, "config"), w.db)
            formats.write_potential(os.path.join(tmp, "pot"), w.pot)
            formats.write_param(os.path.join(tmp, "param"), "config", "pot")

            def run(n):
                t0 = time.time()
                out = subprocess.run([mpiexec, "-n", str(cores), ref, "-p", "param", "--time", str(n)], cwd=tmp,
                                     capture_output=True, text=True, timeout=240)
                if out.returncode != 0:
                    raise RuntimeError(out.stderr[-400:])
                txt = out.stdout
                js = json.loads(txt[txt.index('{"pots"'):])["pots"][0]
                return js["evals_per_s"], time.time() - t0

            r1, _ = run(2)
            n = int(max(2, min(2000, seconds * r1)))
            rate, wall = run(n)
            return dict(value=rate, unit="objective evals/s", cores=cores, kind="reference",
                        sample=f"{n} objective evaluations of the full workload by the reference program "
                               f"(IEEE -O2 build, MPI ranks={cores}, stored triplets), timed inside the program")
        except Exception as e:  # fall back to the restatement
            log("cpu baseline: reference run failed:", e)
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.Oracle(w.pot, w.db)
    kv = w.pot.knot_vector()
    t0 = time.time()
    o.eval_mt(kv, cores)
    one = time.time() - t0
    n = int(max(1, min(500, seconds / max(one, 1e-6))))
    t0 = time.time()
    for _ in range(n):
        o.eval_mt(kv, cores)
    rate = n / (time.time() - t0)
    return dict(value=rate, unit="objective evals/s", cores=cores, kind="port",
                sample=f"{n} objective evaluations of the full workload by the C restatement, {cores} threads")


def _copy_probe(mib: int) -> None:
    """One hipMemcpy device-to-device of a known size: the copy kernel's
    FETCH_SIZE / WRITE_SIZE counters calibrate the traffic of the other kernels."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    hip.hipMemset(a, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy probe: {mib} MiB device-to-device")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", type=int, default=1)
    ap.add_argument("--mode", choices=["cd", "adjoint", "batch"], default="cd")
    ap.add_argument("--cand", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=str, default="", help="max_atoms,max_pairs,block")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (configs shard over GPUs by design "
              f"only, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        _copy_probe(args.copy_probe)
    t0 = time.time()
    w = gen.workload(args.workload)
    gen.add_targets(w, gpu_evaluator)
    log(f"workload {w.name}: {w.db.ncells} cells, {w.db.total_atoms} atoms, built in {time.time() - t0:.1f}s")
    obj = Objective(w.pot, w.db)
    if args.tile:
        a, p, b = (int(x) for x in args.tile.split(","))
        obj.set_tile(a, p, b)
    info = obj.info
    log(f"pairs {info.npairs} triplets {info.ntriplets} chunks {info.nchunks}")
    kv = w.pot.knot_vector()
    ncoeff = w.pot.ncoeff()
    obj.enable_timing(True)

    def plain():
        obj.objective(kv)

    for _ in range(args.warmup):
        plain()
    # kernel durations: HIP events on the context stream around each kernel
    atom_ms = []
    for _ in range(max(args.steps, 10)):
        plain()
        atom_ms.append(obj.last_timing().atom_ms)
    atom_s = float(np.mean(atom_ms)) * 1e-3
    obj.enable_timing(False)  # throughput runs replay the captured graph
    for _ in range(args.warmup):
        plain()
    tp = time.perf_counter()
    nplain = max(args.steps, 10)
    for _ in range(nplain):
        plain()
    plain_dt = (time.perf_counter() - tp) / nplain
    plain_rate = 1.0 / plain_dt

    if args.mode == "cd":
        evals_per_step = 2 * ncoeff + 1
        ts = time.perf_counter()
        for _ in range(args.steps):
            for _ in range(evals_per_step):
                plain()
        dt = (time.perf_counter() - ts) / args.steps
        value = 1.0 / dt
        per_step = 1
        mode_desc = f"objective + central-difference gradient ({evals_per_step} plain evaluations per gradient)"
    elif args.mode == "adjoint":
        for _ in range(args.warmup):
            obj.objective_grad(kv)
        ts = time.perf_counter()
        for _ in range(args.steps):
            obj.objective_grad(kv)
        dt = (time.perf_counter() - ts) / args.steps
        value = 1.0 / dt
        per_step = 1
        mode_desc = "objective + adjoint gradient, one knot vector per call"
    else:
        K = w.candidate_knots(args.cand)
        for _ in range(args.warmup):
            obj.batch(K, grad=True)
        ts = time.perf_counter()
        for _ in range(args.steps):
            obj.batch(K, grad=True)
        dt = (time.perf_counter() - ts) / args.steps
        value = args.cand / dt
        per_step = args.cand
        mode_desc = f"objective + adjoint gradient of {args.cand} knot vectors per batched call"

    P, T, N, Ccells = info.npairs, info.ntriplets, info.natoms, info.ncells
    flop = FLOP_PER_TRIPLET * T + FLOP_PER_PAIR_MEAM * P
    achieved = flop / atom_s / 1e12
    hbm_bytes = BYTES_PER_PAIR * P + BYTES_PER_ATOM * N + BYTES_PER_CELL * Ccells
    out = {
        "metric": "MEAM objective+gradient evals/s, batched vs plain on 1 GPU; triplets/s; %HBM BW",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded geometry, knots and noisy targets)",
        "config": {"workload": w.name, "cells": Ccells, "atoms": N, "pairs": P, "triplets": T,
                   "ncoeff": ncoeff, "mode": args.mode, "candidates_per_step": per_step,
                   "tile": [obj.get_tile().max_atoms, obj.get_tile().max_pairs, obj.get_tile().block]},
        "mode": mode_desc,
        "objective_evals_per_s": plain_rate,
        "pairs_per_s": P * plain_rate,
        "triplets_per_s": T * plain_rate,
        "atom_kernel_ms": atom_s * 1e3,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "k_atoms (fused pair + triplet + embedding + force pass)",
                     "work": f"{FLOP_PER_TRIPLET} FLOP/triplet x {T} + {FLOP_PER_PAIR_MEAM} FLOP/pair x {P}",
                     "note": "FP64 vector work; the fp64 peak is the published 78.6 TFLOP/s"},
        "hbm": {"bytes_per_eval": hbm_bytes, "achieved_GBs": hbm_bytes / atom_s / 1e9,
                "frac": hbm_bytes / atom_s / 1e9 / PEAK_HBM_GBS},
    }
    if not args.no_cpu:
        cb = cpu_baseline(w, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["cpu_baseline"]["objective_grad_cd_evals_per_s"] = cb["value"] / (2 * ncoeff + 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
