#!/usr/bin/env python3
"""Throughput of the MEAM objective and its gradient on one MI355X.

Workload (default): 256 periodic cells of 128 atoms (half displaced BCC
4x4x4 at a=3.16 A, half random-dense at 0.063 /A^3), one-element meam/spline
with phi/rho/f 12 knots on [1.8,5.0], U 10 knots on [-1,3], g 10 knots on
[-1,1]; cutoff 5.0 A; targets from a seeded truth knot set plus noise.
Synthetic, seeded geometry and knots; FP64 throughout.

One step = one call of the selected mode:
  batch   : objective + adjoint gradient of --cand knot vectors in one
            captured launch sequence (mz_objective_batch)       [headline]
  adjoint : objective + adjoint gradient of one knot vector (mz_objective_grad)
  cd      : objective + central-difference gradient of one knot vector
            (2*ncoeff+1 plain evaluations, the per-vector path an optimizer
            without an analytic gradient uses)
`value` is objective+gradient evaluations per second of the selected mode.
The plain objective rate, the plain+CD gradient rate and the single-vector
adjoint rate are always measured beside it, so the line also carries the
batched-vs-plain ratio.  Kernel durations come from HIP events recorded on
the library's stream around each kernel, over K steps of the same mode.
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
from meamzilla_amd._lib import Objective, probe_peaks  # noqa: E402

# Algorithmic work per unit, counted from the kernels' arithmetic (cubic
# splines, knot-style evaluation; one FMA = 2 FLOP, sqrt and divide = 1).
FWD_FLOP_PER_TRIPLET = 59    # k_atoms, per unordered (j,k) neighbour pair of an atom
FWD_FLOP_PER_PAIR = 118      # k_atoms, per full-list pair (geometry, phi/rho/f, force, virial)
ADJ_FLOP_PER_TRIPLET = 80    # k_adjoint, per unordered triplet (g value/slope, both legs, g-coefficient adjoints)
ADJ_FLOP_PER_PAIR = 162      # k_adjoint, per pair (geometry, rho/f, duals, phi/rho/f coefficient adjoints)
# Algorithmic HBM bytes of k_atoms: pair record 8 + stored force vector 24 per
# pair; position 24 + type 4 + pair offset 4 + own-force sum 24 per atom.
FWD_BYTES_PER_PAIR = 32
FWD_BYTES_PER_ATOM = 56
RESID_BYTES_PER_PAIR = 28    # k_resid: pair-force gather (24 B) + incoming index (4 B) per full-list pair
RESID_BYTES_PER_ATOM = 76    # k_resid: own force sum, target force, force out, lambda out, index, type/cell
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gpu_evaluator(pot, db):
    r = Objective(pot, db).compute(errvec=False)
    return r["cell_energy"], r["cell_stress"], r["forces"]


def _cores() -> int:
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return max(1, min(cores, 16))


# The reference stores every triplet (about 88 bytes each); above this many
# triplets the CPU baseline runs on the leading cells of the workload and is
# scaled to the whole by triplet count.
CPU_MAX_TRIPLETS = 50_000_000


def cpu_baseline(w, seconds: float, ncoeff: int, ntriplets: int) -> dict:
    """The reference program (built from its sources into oracle/_ref) run with
    MPI ranks on this host's cores on the same workload; the oracle's threaded
    C restatement when MPI or that build is unavailable.  Its gradient is
    priced as the reference's Powell forms it: ncoeff forward differences,
    i.e. ncoeff+1 objective evaluations per objective+gradient."""
    import copy
    cores = _cores()
    scale, subset_note = 1.0, ""
    if ntriplets > CPU_MAX_TRIPLETS:
        per_cell = ntriplets / w.db.ncells
        m = max(1, int(CPU_MAX_TRIPLETS / per_cell))
        w = copy.copy(w)
        w.db = w.db.subset(range(m))
        sub = Objective(w.pot, w.db).info.ntriplets
        scale = sub / ntriplets
        subset_note = (f"; run on the first {m} of the workload's cells ({sub} of {ntriplets} triplets) and scaled "
                       f"to the whole workload by triplet count")
        log(f"cpu baseline on {m} cells ({sub} triplets)")
    # the reference as shipped (-Ofast -march=corei7 -funroll-all-loops) where
    # built, else its IEEE -O2 build
    ref = os.path.join(ROOT, "oracle", "_ref", "meamz_ref_fast")
    flags = "-Ofast -march=corei7 -funroll-all-loops, the shipped flags"
    if not os.path.exists(ref):
        ref = os.path.join(ROOT, "oracle", "_ref", "meamz_ref")
        flags = "-O2 IEEE build"
    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    if os.path.exists(ref) and os.path.exists(mpiexec):
        tmp = tempfile.mkdtemp(prefix="meamz_cpu_")
        try:
            formats.write_config(os.path.join(tmp, "config"), w.db)
            formats.write_potential(os.path.join(tmp, "pot"), w.pot)
            formats.write_param(os.path.join(tmp, "param"), "config", "pot")

            def run(n):
                out = subprocess.run([mpiexec, "-n", str(cores), ref, "-p", "param", "--time", str(n)], cwd=tmp,
                                     capture_output=True, text=True, timeout=240)
                if out.returncode != 0:
                    raise RuntimeError(out.stderr[-400:])
                txt = out.stdout
                return json.loads(txt[txt.index('{"pots"'):])["pots"][0]["evals_per_s"]

            r1 = run(2)
            n = int(max(2, min(2000, seconds * r1)))
            rate = run(n) * scale
            return dict(value=rate / (ncoeff + 1), unit="objective+gradient evals/s", cores=cores, kind="reference",
                        objective_evals_per_s=rate,
                        sample=f"{n} objective evaluations by the reference program ({flags}, {cores} MPI "
                               f"ranks, stored triplets), timed inside the program; gradient = {ncoeff} forward "
                               f"differences as its Powell forms them" + subset_note)
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
    rate = n / (time.time() - t0) * scale
    return dict(value=rate / (ncoeff + 1), unit="objective+gradient evals/s", cores=cores, kind="port",
                objective_evals_per_s=rate,
                sample=f"{n} objective evaluations by the C restatement, {cores} threads; "
                       f"gradient = {ncoeff} forward differences" + subset_note)


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


def _traffic(kernel: str, workload: str, cand: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE, calibrated against a device copy), if
    one exists for this workload and candidate count."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            tab = json.load(fh)
        e = tab[f"{kernel}|{workload}|{cand}"]
        return e["hbm_bytes_per_launch"], os.path.relpath(path, ROOT)
    except (OSError, KeyError, ValueError):
        return None, None


def _fp64_ops(kernel: str, workload: str, cand: int):
    """Executed FP64 operations per launch of `kernel` from the committed
    rocprofv3 VALU-counter summary, if one exists for this workload."""
    path = os.path.join(ROOT, "profiles", "fp64.json")
    try:
        with open(path) as fh:
            return json.load(fh)[f"{kernel}|{workload}|{cand}"]["fp64_ops_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def _num(x: float):
    return None if x != x else x


def _rate(fn, steps: int) -> float:
    ts = time.perf_counter()
    for _ in range(steps):
        fn()
    return (time.perf_counter() - ts) / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", type=int, default=1)
    ap.add_argument("--mode", choices=["batch", "adjoint", "cd"], default="batch")
    ap.add_argument("--cand", type=int, default=64, help="knot vectors per batched call")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=str, default="", help="max_atoms,max_pairs,block")
    ap.add_argument("--no-side", action="store_true",
                    help="skip the side rates (plain, CD, single adjoint); for profiler runs")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: this build runs on a single GPU "
              f"(how the configuration grid would shard over GPUs is described in the design notes)",
              file=sys.stderr)
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
    P, T, N, Ccells = info.npairs, info.ntriplets, info.natoms, info.ncells
    log(f"pairs {P} triplets {T} chunks {info.nchunks}")
    kv = w.pot.knot_vector()
    ncoeff = w.pot.ncoeff()
    K = w.candidate_knots(args.cand)

    def plain():
        obj.objective(kv)

    def cd():
        for _ in range(2 * ncoeff + 1):
            plain()

    def adjoint():
        obj.objective_grad(kv)

    def batch():
        obj.batch(K, grad=True)

    # side rates (captured graphs, no timing events)
    plain_dt = cd_dt = adj_dt = float("nan")
    if not args.no_side:
        for _ in range(args.warmup):
            plain()
            adjoint()
        plain_dt = _rate(plain, max(3, min(50, int(2.0 / max(_rate(plain, 1), 1e-6)))))
        log(f"plain objective {1.0 / plain_dt:.1f} evals/s")
        cd_dt = _rate(cd, 1 if plain_dt * (2 * ncoeff + 1) > 2.0 else 2)
        log(f"objective + CD gradient {1.0 / cd_dt:.2f} evals/s")
        adj_dt = _rate(adjoint, max(3, min(20, int(2.0 / max(_rate(adjoint, 1), 1e-6)))))
        log(f"objective + adjoint gradient {1.0 / adj_dt:.1f} evals/s")

    step, per_step = {"batch": (batch, args.cand), "adjoint": (adjoint, 1), "cd": (cd, 1)}[args.mode]
    for _ in range(args.warmup):
        step()
    ts = time.perf_counter()
    for _ in range(args.steps):
        step()  # every call returns after its stream has drained
    dt = (time.perf_counter() - ts) / args.steps
    value = per_step / dt
    log(f"{args.mode}: {value:.1f} evals/s over {args.steps} steps")

    # kernel durations: HIP events around each kernel on the library's stream,
    # over the same K steps (stream launches instead of the captured graph)
    obj.enable_timing(True)
    step()
    tm = {"atom_ms": [], "adjoint_ms": [], "resid_ms": [], "refresh_ms": []}
    for _ in range(args.steps):
        if args.mode == "cd":
            plain()
        else:
            step()
        t = obj.last_timing()
        for k in tm:
            tm[k].append(getattr(t, k))
    obj.enable_timing(False)
    ms = {k: float(np.mean(v)) for k, v in tm.items()}
    ncand_launch = per_step if args.mode == "batch" else 1

    fwd_flop = (FWD_FLOP_PER_TRIPLET * T + FWD_FLOP_PER_PAIR * P) * ncand_launch
    adj_flop = (ADJ_FLOP_PER_TRIPLET * T + ADJ_FLOP_PER_PAIR * P) * ncand_launch
    fwd = {"kernel": "k_atoms", "ms": ms["atom_ms"], "flop": fwd_flop,
           "tflops": fwd_flop / (ms["atom_ms"] * 1e-3) / 1e12}
    if args.mode == "cd":
        dom = dict(fwd, work=f"({FWD_FLOP_PER_TRIPLET} FLOP/triplet x {T} + {FWD_FLOP_PER_PAIR} FLOP/pair x {P})")
    else:
        dom = {"kernel": "k_adjoint", "ms": ms["adjoint_ms"], "flop": adj_flop,
               "tflops": adj_flop / (ms["adjoint_ms"] * 1e-3) / 1e12,
               "work": f"({ADJ_FLOP_PER_TRIPLET} FLOP/triplet x {T} + {ADJ_FLOP_PER_PAIR} FLOP/pair x {P}) "
                       f"x {ncand_launch} candidates"}
    traffic, tsrc = _traffic(dom["kernel"], w.name, ncand_launch)
    hbm_fwd = (FWD_BYTES_PER_PAIR * P + FWD_BYTES_PER_ATOM * N) * ncand_launch
    hbm_resid = (RESID_BYTES_PER_PAIR * P + RESID_BYTES_PER_ATOM * N) * ncand_launch
    traffic_resid, _ = _traffic("k_resid", w.name, ncand_launch)
    ex_ops = _fp64_ops(dom["kernel"], w.name, ncand_launch)
    peaks = None
    if not args.no_side:
        tf, gbs = probe_peaks()
        peaks = {"fp64_tflops": tf, "hbm_copy_GBs": gbs,
                 "note": "measured on this GPU: FP64 FMA loop (8 chains per lane) and a 1 GiB stream copy "
                         "(read + write bytes); datasheet peaks are 78.6 TFLOP/s and 8000 GB/s"}
        log(f"measured peaks: FP64 {tf:.1f} TFLOP/s, copy {gbs:.0f} GB/s")
    tile = obj.get_tile()
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
        "config": {"workload": f"{w.name}: {Ccells} cells, {N} atoms, {w.pot.ntypes}-element {w.pot.style}, "
                               f"cutoff {w.cutoff:g} A",
                   "cells": Ccells, "atoms": N, "pairs": P, "triplets": T, "ncoeff": ncoeff, "mode": args.mode,
                   "candidates_per_step": per_step, "tile": [tile.max_atoms, tile.max_pairs, tile.block]},
        "objective_evals_per_s": _num(1.0 / plain_dt),
        "objective_grad_cd_evals_per_s": _num(1.0 / cd_dt),
        "objective_grad_adjoint_evals_per_s": _num(1.0 / adj_dt),
        "batched_vs_plain_cd": _num(value * cd_dt) if args.mode == "batch" else None,
        "pairs_per_s": _num(P / plain_dt),
        "triplets_per_s": _num(T / plain_dt),
        "kernel_ms": ms,
        "roofline": {"bound": "mfma", "achieved": dom["tflops"], "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": dom["tflops"] / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "kernel": dom["kernel"], "work": dom["work"], "launch_ms": dom["ms"],
                     "traffic_source": tsrc,
                     "frac_of_measured_peak": dom["tflops"] / peaks["fp64_tflops"] if peaks else None,
                     "executed_tflops": (ex_ops / (dom["ms"] * 1e-3) / 1e12 if ex_ops else None),
                     "executed_source": "profiles/fp64.json (rocprofv3 SQ_INSTS_VALU_*_F64)" if ex_ops else None,
                     "note": "FP64 vector arithmetic against the published FP64 peak (78.6 TFLOP/s); "
                             "no MFMA is used"},
        "forward_kernel": {"kernel": "k_atoms", "launch_ms": fwd["ms"], "tflops": fwd["tflops"],
                           "fp64_frac": fwd["tflops"] / PEAK_FP64_TFLOPS,
                           "hbm_bytes": hbm_fwd, "hbm_GBs": hbm_fwd / (fwd["ms"] * 1e-3) / 1e9,
                           "hbm_frac": hbm_fwd / (fwd["ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS},
        # the HBM-bound kernel of the path; its time includes the final per-candidate sum
        "force_gather_kernel": {"kernel": "k_resid", "launch_ms": ms["resid_ms"], "hbm_bytes": hbm_resid,
                                "hbm_frac": hbm_resid / (ms["resid_ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS,
                                "traffic": traffic_resid,
                                "traffic_frac": (traffic_resid / (ms["resid_ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS
                                                 if traffic_resid else None)},
        "measured_peaks": peaks,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds, ncoeff, T)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
