#!/usr/bin/env python3
"""Throughput of the MEAM objective and its gradient on one MI355X.

Workload (default, the ternary batched population + adjoint set): 2000
periodic cells with atom counts uniform in [256, 2048] (2.31 M atoms), strained
triclinic BCC/FCC cells with random species fractions, ternary A/B/C
meam/spline (6 phi, 3 rho, 3 U, 3 f, 3 g; 12 knots each, 252 adjustable
coefficients), cutoff 5.5 A; targets from a seeded truth knot set plus noise;
64 candidate knot vectors (seeded base x (1 + U(+-5%))) per step.  Synthetic,
seeded geometry and knots; FP64 throughout.  --workload 1 selects the
one-element 256 x 128-atom meam/spline set, whose figures the default run also
reports under "one_element".

One step = one call of the selected mode:
  batch   : objective + adjoint gradient of --cand knot vectors in one
            captured launch sequence per sub-batch (mz_objective_batch) [headline]
  adjoint : objective + adjoint gradient of one knot vector (mz_objective_grad)
  cd      : objective + central-difference gradient of one knot vector
            (2*ncoeff+1 plain evaluations, the per-vector path an optimizer
            without an analytic gradient uses)
  plain   : objective of one knot vector (mz_objective through its captured
            graph, the latency an annealing trial or a line-search point
            pays); the timed steps replay the graph with timing off, and the
            kernel durations come from a separate pass of the same steps
`value` is objective+gradient evaluations per second of the selected mode.
The plain objective rate, the plain+CD gradient rate and the single-vector
adjoint rate are measured beside it, so the line also carries the
batched-vs-plain ratio.  The K timed steps run with the library's kernel
timing on: each call launches its kernels on the library's stream between HIP
events (instead of replaying its captured graph, whose replays keep no event
timestamps), so `value` and the kernel durations come from the same steps.  A
batch larger than the free device memory holds runs as equal sub-batches, and
per-launch figures divide by the launches of a step.
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
# Algorithmic HBM bytes.  Geometry shared by the candidates of a launch is
# counted once per launch, what each candidate reads or writes once per
# candidate.
# k_atoms: per launch the pair record 8 + its incoming slot 4 per pair,
# position 24 + type 4 + pair offset 4 per atom; per candidate the stored
# force vector 24 (+ leg record G 32 in gradient runs) per pair and the
# own-force sum 24 (+ density and U' 16 in gradient runs) per atom.
FWD_SHARED_BYTES_PER_PAIR = 12
FWD_SHARED_BYTES_PER_ATOM = 32
FWD_BYTES_PER_PAIR = 24
FWD_GRAD_BYTES_PER_PAIR = 32
FWD_BYTES_PER_ATOM = 24
FWD_GRAD_BYTES_PER_ATOM = 16
# k_resid: per candidate the force vectors of the incoming pairs, streamed in
# incoming-index order (24 B per pair), own-force sum in and force out (48 B
# per atom, + lambda out 24 B in gradient runs); per launch target force 24 +
# incoming offset 4 + cell index 4 per atom.
RESID_BYTES_PER_PAIR = 24
RESID_BYTES_PER_ATOM = 48
RESID_GRAD_BYTES_PER_ATOM = 24
RESID_SHARED_BYTES_PER_ATOM = 32
# k_adjoint: per candidate the leg record G (32 B per pair) and the force duals
# and density (32 B per atom); per launch the pair record 8 + owner 4 per pair
# and position 24 + type 4 per atom.
ADJ_BYTES_PER_PAIR = 32
ADJ_BYTES_PER_ATOM = 32
ADJ_SHARED_BYTES_PER_PAIR = 12
ADJ_SHARED_BYTES_PER_ATOM = 28
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gpu_evaluator(pot, db):
    r = Objective(pot, db).compute(errvec=False)
    return r["cell_energy"], r["cell_stress"], r["forces"]


# Worker processes on the host: the CPUs this process may run on, at most 16
# (the host-CPU share of one GPU on the MI355X nodes this runs on, where
# OMP_NUM_THREADS and MAX_JOBS are set to 16).
CPU_SHARE = 16


def _cores() -> int:
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return max(1, min(cores, CPU_SHARE))


def host_cpu() -> dict:
    """Model, logical CPUs, physical cores and usable CPUs of this host."""
    model, phys = "", set()
    try:
        cur = {}
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if ":" not in line:
                    if cur:
                        phys.add((cur.get("physical id", "0"), cur.get("core id", cur.get("processor", ""))))
                    cur = {}
                    continue
                k, v = (t.strip() for t in line.split(":", 1))
                cur[k] = v
                if k == "model name" and not model:
                    model = v
        if cur:
            phys.add((cur.get("physical id", "0"), cur.get("core id", cur.get("processor", ""))))
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "logical_cpus": os.cpu_count(), "physical_cores": len(phys) or None,
            "usable_cpus": usable, "ranks_used": _cores(), "share_per_gpu": CPU_SHARE}


# The reference stores every triplet (about 88 bytes each); above this many
# triplets the CPU baseline runs on the leading cells of the workload and is
# scaled to the whole by triplet count.
CPU_MAX_TRIPLETS = 50_000_000


def cpu_baseline(w, seconds: float, ncoeff: int, ntriplets: int) -> dict:
    """The reference program (built from its sources into oracle/_ref) run with
    MPI ranks on this host's cores on the same workload; the oracle's threaded
    C restatement when MPI or that build is unavailable.  Its objective+gradient
    rate uses the same central-difference scheme as the GPU plain+CD figure:
    2*ncoeff+1 objective evaluations per objective+gradient (the reference's
    own Powell forms ncoeff forward differences, ncoeff+1 evaluations; the
    objective rate is reported too, so either can be formed)."""
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
            return dict(value=rate / (2 * ncoeff + 1), unit="objective+gradient evals/s", cores=cores,
                        kind="reference", objective_evals_per_s=rate, gradient="central differences, 2*ncoeff+1 "
                        "objective evaluations", host=host_cpu(),
                        sample=f"{n} objective evaluations by the reference program ({flags}, {cores} MPI "
                               f"ranks, stored triplets), timed inside the program; gradient = {2 * ncoeff + 1} "
                               f"objective evaluations (central differences, as the GPU plain+CD figure)"
                               + subset_note)
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
    return dict(value=rate / (2 * ncoeff + 1), unit="objective+gradient evals/s", cores=cores, kind="port",
                objective_evals_per_s=rate, gradient="central differences, 2*ncoeff+1 objective evaluations",
                host=host_cpu(),
                sample=f"{n} objective evaluations by the C restatement, {cores} threads; "
                       f"gradient = {2 * ncoeff + 1} objective evaluations (central differences)" + subset_note)


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


def measure(wl: int, steps: int, warmup: int, mode: str, cand: int, side: bool, tile: str = "") -> tuple:
    """One workload: the side rates (plain objective, plain + CD gradient,
    single-vector adjoint), then `steps` timed steps of `mode` with the kernel
    events on; returns (result dict, workload, objective)."""
    t0 = time.time()
    w = gen.workload(wl)
    gen.add_targets(w, gpu_evaluator)
    log(f"workload {w.name}: {w.db.ncells} cells, {w.db.total_atoms} atoms, built in {time.time() - t0:.1f}s")
    obj = Objective(w.pot, w.db)
    tune_s = 0.0
    if tile:
        v = [int(x) for x in tile.split(",")]
        obj.set_tile(v[0], v[1], v[2], v[3] if len(v) > 3 else 2)
    else:
        # the work decomposition the library's autotuner picks for this
        # database (once per database, like the list build; not timed)
        t0 = time.time()
        nt = min(cand, 32)
        kt = w.candidate_knots(nt)
        grad = mode in ("batch", "adjoint")
        # calls per tile: enough for about 0.3 s, so that tiles of a small
        # database (calls of a few ms) are not ranked by noise
        obj.batch(kt, grad=grad)
        t1 = time.perf_counter()
        obj.batch(kt, grad=grad)
        t1 = time.perf_counter() - t1
        reps = int(min(20, max(2, 0.3 / max(t1, 1e-6))))
        obj.autotune(ncand=nt, reps=reps, grad=grad, knots=kt)
        tune_s = time.time() - t0
        t = obj.get_tile()
        log(f"autotuned tile {t.max_atoms},{t.max_pairs},{t.block},{t.cand_tile} in {tune_s:.1f}s")
    info = obj.info
    P, T, N, Ccells = info.npairs, info.ntriplets, info.natoms, info.ncells
    log(f"pairs {P} triplets {T} chunks {info.nchunks}")
    kv = w.pot.knot_vector()
    ncoeff = w.pot.ncoeff()
    K = w.candidate_knots(cand)

    def plain():
        obj.objective(kv)

    def cd():
        for _ in range(2 * ncoeff + 1):
            plain()

    def adjoint():
        obj.objective_grad(kv)

    def batch():
        obj.batch(K, grad=True)

    # side rates (captured graphs, timing off)
    plain_dt = cd_dt = adj_dt = float("nan")
    if side:
        for _ in range(warmup):
            plain()
            adjoint()
        plain_dt = _rate(plain, max(3, min(50, int(2.0 / max(_rate(plain, 1), 1e-6)))))
        log(f"plain objective {1.0 / plain_dt:.1f} evals/s")
        cd_dt = _rate(cd, 1 if plain_dt * (2 * ncoeff + 1) > 2.0 else 2)
        log(f"objective + CD gradient {1.0 / cd_dt:.2f} evals/s")
        adj_dt = _rate(adjoint, max(3, min(20, int(2.0 / max(_rate(adjoint, 1), 1e-6)))))
        log(f"objective + adjoint gradient {1.0 / adj_dt:.1f} evals/s")

    # timed region: K steps of the mode with the HIP events around every
    # kernel on the library's stream
    step, per_step = {"batch": (batch, cand), "adjoint": (adjoint, 1), "cd": (cd, 1), "plain": (plain, 1)}[mode]
    tm = {"atom_ms": [], "adjoint_ms": [], "resid_ms": [], "refresh_ms": [], "total_ms": [], "launches": []}

    def timed_steps(record):
        obj.enable_timing(record)
        for _ in range(warmup):
            step()
        ts = time.perf_counter()
        for _ in range(steps):
            step()  # every call returns after its stream has drained
            if record:
                t = obj.last_timing()
                for k in tm:
                    tm[k].append(getattr(t, k))
        obj.enable_timing(False)
        return (time.perf_counter() - ts) / steps

    if mode == "plain":
        dt = timed_steps(False)  # the captured graph, as an optimizer replays it
        timed_steps(True)
    else:
        dt = timed_steps(True)
    value = per_step / dt
    log(f"{mode}: {value:.1f} evals/s over {steps} steps")
    ms = {k: float(np.mean(v)) for k, v in tm.items() if k != "launches"}  # per step, summed over launches
    launches = int(round(float(np.mean(tm["launches"])))) if mode not in ("cd", "plain") else 1
    cand_step = per_step if mode == "batch" else 1
    cand_launch = cand_step / launches
    grad = mode not in ("cd", "plain")

    fwd_flop = (FWD_FLOP_PER_TRIPLET * T + FWD_FLOP_PER_PAIR * P) * cand_step
    adj_flop = (ADJ_FLOP_PER_TRIPLET * T + ADJ_FLOP_PER_PAIR * P) * cand_step
    fwd_bytes = ((FWD_SHARED_BYTES_PER_PAIR * P + FWD_SHARED_BYTES_PER_ATOM * N) * launches +
                 ((FWD_BYTES_PER_PAIR + (FWD_GRAD_BYTES_PER_PAIR if grad else 0)) * P +
                  (FWD_BYTES_PER_ATOM + (FWD_GRAD_BYTES_PER_ATOM if grad else 0)) * N) * cand_step)
    resid_bytes = (RESID_SHARED_BYTES_PER_ATOM * N * launches +
                   (RESID_BYTES_PER_PAIR * P + (RESID_BYTES_PER_ATOM + (RESID_GRAD_BYTES_PER_ATOM if grad else 0))
                    * N) * cand_step)
    adj_bytes = ((ADJ_SHARED_BYTES_PER_PAIR * P + ADJ_SHARED_BYTES_PER_ATOM * N) * launches +
                 (ADJ_BYTES_PER_PAIR * P + ADJ_BYTES_PER_ATOM * N) * cand_step)
    if not grad:
        dom = {"kernel": "k_atoms", "ms": ms["atom_ms"], "flop": fwd_flop,
               "work": f"({FWD_FLOP_PER_TRIPLET} FLOP/triplet x {T} + {FWD_FLOP_PER_PAIR} FLOP/pair x {P})"}
    else:
        dom = {"kernel": "k_adjoint", "ms": ms["adjoint_ms"], "flop": adj_flop,
               "work": f"({ADJ_FLOP_PER_TRIPLET} FLOP/triplet x {T} + {ADJ_FLOP_PER_PAIR} FLOP/pair x {P}) "
                       f"x {cand_launch:g} candidates per launch"}
    dom["tflops"] = dom["flop"] / (dom["ms"] * 1e-3) / 1e12
    key_cand = int(cand_launch) if cand_launch == int(cand_launch) else None
    traffic, tsrc = _traffic(dom["kernel"], w.name, key_cand)
    traffic_resid, _ = _traffic("k_resid", w.name, key_cand)
    ex_ops = _fp64_ops(dom["kernel"], w.name, key_cand)
    tile_now = obj.get_tile()

    def gbs(b, t_ms):
        return b / (t_ms * 1e-3) / 1e9

    res = {
        "workload": f"{w.name}: {Ccells} cells, {N} atoms, {w.pot.ntypes}-element {w.pot.style}, "
                    f"cutoff {w.cutoff:g} A",
        "name": w.name, "cells": Ccells, "atoms": N, "pairs": P, "triplets": T, "ncoeff": ncoeff, "mode": mode,
        "candidates_per_step": per_step, "launches_per_step": launches, "candidates_per_launch": cand_launch,
        "tile": [tile_now.max_atoms, tile_now.max_pairs, tile_now.block, tile_now.cand_tile],
        "autotune_s": tune_s,
        "value": value, "ms_per_step": dt * 1e3,
        "objective_evals_per_s": _num(1.0 / plain_dt),
        "objective_grad_cd_evals_per_s": _num(1.0 / cd_dt),
        "objective_grad_adjoint_evals_per_s": _num(1.0 / adj_dt),
        "batched_vs_plain_cd": _num(value * cd_dt) if mode == "batch" else None,
        "pairs_per_s": _num(P / plain_dt),
        "triplets_per_s": _num(T / plain_dt),
        "kernel_ms_per_step": ms,
        "roofline": {"bound": "fp64_valu", "achieved": dom["tflops"], "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": dom["tflops"] / PEAK_FP64_TFLOPS,
                     "traffic": traffic, "algorithmic_bytes_per_launch": (adj_bytes if grad else fwd_bytes) / launches,
                     "kernel": dom["kernel"], "work": dom["work"], "launch_ms": dom["ms"] / launches,
                     "traffic_source": tsrc,
                     "executed_tflops": (ex_ops / (dom["ms"] / launches * 1e-3) / 1e12 if ex_ops else None),
                     "executed_source": "profiles/fp64.json (rocprofv3 SQ_INSTS_VALU_*_F64)" if ex_ops else None,
                     "note": "FP64 vector arithmetic (VALU) against the published FP64 peak, 78.6 TFLOP/s; no MFMA "
                             "is used: nothing on this path is a dense contraction"},
        "forward_kernel": {"kernel": "k_atoms", "launch_ms": ms["atom_ms"] / launches,
                           "tflops": fwd_flop / (ms["atom_ms"] * 1e-3) / 1e12,
                           "fp64_frac": fwd_flop / (ms["atom_ms"] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                           "hbm_bytes_per_launch": fwd_bytes / launches, "hbm_GBs": gbs(fwd_bytes, ms["atom_ms"]),
                           "hbm_frac": gbs(fwd_bytes, ms["atom_ms"]) / PEAK_HBM_GBS},
        # the HBM-bound kernel of the path; its time includes the final per-candidate sum
        "force_gather_kernel": {"kernel": "k_resid", "launch_ms": ms["resid_ms"] / launches,
                                "hbm_bytes_per_launch": resid_bytes / launches,
                                "hbm_GBs": gbs(resid_bytes, ms["resid_ms"]),
                                "hbm_frac": gbs(resid_bytes, ms["resid_ms"]) / PEAK_HBM_GBS,
                                "traffic": traffic_resid,
                                "traffic_frac": (gbs(traffic_resid * launches, ms["resid_ms"]) / PEAK_HBM_GBS
                                                 if traffic_resid else None)},
    }
    if grad:
        res["adjoint_kernel"] = {"kernel": "k_adjoint", "hbm_bytes_per_launch": adj_bytes / launches,
                                 "traffic": traffic}
    return res, w, obj


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", type=int, default=4)
    ap.add_argument("--mode", choices=["batch", "adjoint", "cd", "plain"], default="batch")
    ap.add_argument("--cand", type=int, default=64, help="knot vectors per batched call")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=str, default="",
                    help="max_atoms,max_pairs,block[,candidates per workgroup] (default: the tile mz_autotune "
                         "picks for the database)")
    ap.add_argument("--no-side", action="store_true",
                    help="skip the side rates (plain, CD, single adjoint) and the extra workload; for profiler runs")
    ap.add_argument("--no-extra", action="store_true", help="skip the one-element workload figures")
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
    res, w, obj = measure(args.workload, args.steps, args.warmup, args.mode, args.cand, not args.no_side, args.tile)
    peaks = None
    if not args.no_side:
        tf, gb = probe_peaks()
        peaks = {"fp64_tflops": tf, "hbm_copy_GBs": gb,
                 "note": "measured on this GPU: FP64 FMA loop (8 chains per lane) and a 1 GiB stream copy "
                         "(read + write bytes); datasheet peaks are 78.6 TFLOP/s and 8000 GB/s"}
        log(f"measured peaks: FP64 {tf:.1f} TFLOP/s, copy {gb:.0f} GB/s")
        res["roofline"]["frac_of_measured_peak"] = res["roofline"]["achieved"] / tf
    out = {
        "metric": "MEAM objective+gradient evals/s, batched vs plain on 1 GPU; triplets/s; %HBM BW",
        "value": res["value"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded geometry, knots and noisy targets)",
        "config": {k: res[k] for k in ("workload", "cells", "atoms", "pairs", "triplets", "ncoeff", "mode",
                                       "candidates_per_step", "launches_per_step", "tile", "autotune_s")},
    }
    for k in ("objective_evals_per_s", "objective_grad_cd_evals_per_s", "objective_grad_adjoint_evals_per_s",
              "batched_vs_plain_cd", "pairs_per_s", "triplets_per_s", "kernel_ms_per_step", "roofline",
              "forward_kernel", "force_gather_kernel", "adjoint_kernel"):
        if k in res:
            out[k] = res[k]
    out["measured_peaks"] = peaks
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds, res["ncoeff"], res["triplets"])
    del obj
    if not args.no_side and not args.no_extra and args.workload != 1:
        # the one-element 256 x 128-atom set: the same mode, its own figures
        r1, _, o1 = measure(1, 20, 3, args.mode, args.cand, True)
        out["one_element"] = {k: r1[k] for k in ("workload", "tile", "value", "ms_per_step", "candidates_per_step",
                                                 "objective_evals_per_s", "objective_grad_cd_evals_per_s",
                                                 "objective_grad_adjoint_evals_per_s", "batched_vs_plain_cd",
                                                 "kernel_ms_per_step", "roofline", "forward_kernel",
                                                 "force_gather_kernel")}
        del o1
    print(json.dumps(out))


if __name__ == "__main__":
    main()
