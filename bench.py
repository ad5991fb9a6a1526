#!/usr/bin/env python3
"""Benchmark of the MI355X free-wake hot path.

Default workload: the Biot-Savart microbenchmark (1e5 targets, 1e6 helix
filaments, core radius 0.01), one step = one full N-body evaluation of all
targets against all filaments with inputs resident in HBM.  Prints one JSON
line.  See DESIGN.md for the counting conventions.

  python bench.py [--steps K] [--warmup W] [--precision fp64|mixed]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

F_PAIR = 66.0                 # algorithmic flop per pair (FMA = 2, rsqrt = 1)
PEAK_FP64_VALU = 78.6e12      # MI355X vector FP64, FLOP/s
PEAK_FP32_VALU = 157.3e12     # MI355X vector FP32 (packed), FLOP/s
PEAK_HBM = 8.0e12


def helix_workload(n_targets=100_000, n_helix=8, per_helix=125_000, seed=20261202):
    """Targets uniform in [-2,2]^3; filaments joining consecutive nodes of 8
    helices (radius 1, axis through a centre uniform in [-1,1]^2, pitch
    U(0.1,0.3), z in [-2,2]) with N(0, 1e-3^2) node jitter; gam ~ N(0,1),
    core radius 0.01."""
    rng = np.random.default_rng(seed)
    a, b = [], []
    for _ in range(n_helix):
        c = rng.uniform(-1, 1, size=2)
        pitch = rng.uniform(0.1, 0.3)
        ph = rng.uniform(0, 2 * np.pi)
        z = np.linspace(-2.0, 2.0, per_helix + 1)
        th = ph + 2 * np.pi * (z + 2.0) / pitch
        pts = np.stack([c[0] + np.cos(th), c[1] + np.sin(th), z], axis=1)
        pts += 1e-3 * rng.standard_normal(pts.shape)
        a.append(pts[:-1])
        b.append(pts[1:])
    a, b = np.concatenate(a), np.concatenate(b)
    gam = rng.standard_normal(len(a))
    rc = np.full(len(a), 0.01)
    p = rng.uniform(-2, 2, size=(n_targets, 3))
    sample = rng.permutation(n_targets)[:1000]
    return a, b, gam, rc, p, sample


def cpu_baseline(a, b, gam, rc, p, sample, threads):
    """FP64 oracle (plain C, OpenMP over targets) on a bounded sample."""
    from oracle import oracle as orc
    t0 = time.perf_counter()
    v = orc.filament_sum(a, b, gam, rc, p[sample], threads=threads)
    dt = time.perf_counter() - t0
    return len(sample) * len(a) / dt, v, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "sharding is a design in DESIGN.md and is not built.", file=sys.stderr)
        sys.exit(2)

    from volcanor_amd import vlm
    a, b, gam, rc, p, sample = helix_workload()
    ctx = vlm.Context(0)
    prec = vlm.MIXED if args.precision == "mixed" else vlm.FP64
    ctx.set_precision(prec)
    fs = ctx.fset_create(a, b, gam, rc)
    ps, m = ctx.pset_create(p)
    for _ in range(args.warmup):
        ctx.fset_vind(fs, ps)
    ctx.synchronize()
    ctx.profile_enable(True)
    ctx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.fset_vind(fs, ps)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    kms, nlaunch, pairs = ctx.profile_get()
    ctx.profile_enable(False)
    v = ctx.pset_velocity(ps, m)
    pairs_per_step = float(len(a)) * float(m)
    value = pairs_per_step * args.steps / wall
    avg_ms = kms / max(nlaunch, 1)
    flops_launch = F_PAIR * pairs / max(nlaunch, 1)
    achieved = flops_launch / (avg_ms * 1e-3)
    peak = PEAK_FP64_VALU if prec == vlm.FP64 else PEAK_FP32_VALU
    out = {
        "metric": "Biot-Savart pair-interactions/s",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == vlm.FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "Biot-Savart microbenchmark: 1e5 targets x 1e6 helix filaments, "
                               "core radius 0.01, inputs resident in HBM",
                   "targets": int(m), "filaments": int(len(a)),
                   "precision": args.precision},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                     "kernel_ms": avg_ms, "flop_per_pair": F_PAIR},
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        rate, vc, dtc = cpu_baseline(a, b, gam, rc, p, sample, threads)
        rel = np.max(np.linalg.norm(v[sample] - vc, axis=1)) / np.sqrt(np.mean(np.sum(vc**2, 1)))
        out["cpu_baseline"] = {"value": rate, "unit": "pairs/s", "cores": threads, "kind": "port",
                               "sample": f"{len(sample)} targets x {len(a)} filaments FP64 "
                                         f"({dtc:.1f} s)"}
        out["check"] = {"max_err_over_rms_vs_fp64_oracle": float(rel)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
