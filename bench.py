#!/usr/bin/env python3
"""Benchmark of the MI355X free-wake hot path.  Prints one JSON line.

Workloads
  quad (default)    quad-rotor (cases/cfg4_quad.case): 4 three-bladed rotors x
                    1 x 30 panels, alternating rotation, 10 deg steps, AM2
                    free wake, 15-revolution truncated lattice, run by the
                    Fortran host on the GPU.  A step is one free-wake timestep
                    (rigid motion, shed, RHS + circulation solve + loads, AM2
                    predictor-corrector convection with both Biot-Savart
                    passes, rollup).  The run always covers all 370 steps and
                    times the last K (default 20: steps 351-370, the 10th
                    revolution); the steps before build the wake that defines
                    the workload, so --warmup does not shorten the case.
                    Variants: plain FP64 (the value), mixed FP32/FP64, mixed
                    with the learned pass 2, and an FP64 twin with a 1e-9 deg
                    cyclic pitch (the chaos floor of CT).
  rotor2            two-bladed hover rotor (cases/cfg2_rotor.case), 1 x 20
                    panels per blade, 5 deg steps, 720 steps (10
                    revolutions), the last revolution timed; same variants.
  rotorwing         three-bladed rotor in forward flight above an AR-8 wing
                    (cases/cfg5_rotor_wing.case), 1080 steps (15
                    revolutions), the last revolution timed; same variants.
  microbench        1e5 targets x 1e6 helix filaments, one N-body evaluation
                    per step, inputs resident in HBM.
  lu                dense FP64 LU with partial pivoting, N = 2176.
  hbmcopy           1 GiB device copy (HBM counter calibration).

  python bench.py [--workload quad|rotor2|rotorwing|microbench|lu|hbmcopy] [--steps K]
                  [--warmup W]
"""
import argparse
import json
import os
import resource
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

F_PAIR = 66.0                 # algorithmic flop per pair (FMA = 2, rsqrt = 1)
PEAK_FP64_VALU = 78.6e12      # MI355X vector FP64, FLOP/s
PEAK_FP32_VALU = 157.3e12     # MI355X vector FP32 (packed), FLOP/s
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "volcanor")


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
    # the first 1e3 of a seeded permutation are the error sample; the first
    # 1e4 are timed on the host cores as the CPU baseline
    sample = rng.permutation(n_targets)[:10000]
    return a, b, gam, rc, p, sample


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "bs_traffic.json")


def measured_traffic(key):
    """HBM bytes per launch of a kernel, from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs summarised in profiles/ (scaled by a 1 GiB copy)."""
    try:
        return float(json.load(open(TRAFFIC_FILE))["kernels"][key])
    except (OSError, KeyError, ValueError):
        return None


_FMA_STREAM = []


def fma_stream():
    """Sustained FP64 FMA rate of this device (TFLOP/s) and the relative time
    of a v_rsq_f64 stream, measured once per process after the timed work."""
    if not _FMA_STREAM:
        from volcanor_amd import vlm
        _FMA_STREAM.extend(vlm.Context(0).fp64_issue_rate())
    return _FMA_STREAM


def roofline(kernel_ms, launches, pairs, peak, kernel, traffic_key=None):
    avg_ms = kernel_ms / max(launches, 1)
    flops_launch = F_PAIR * pairs / max(launches, 1)
    achieved = flops_launch / (avg_ms * 1e-3) if avg_ms > 0 else 0.0
    out = {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12,
           "unit": "TFLOP/s", "frac": achieved / peak,
           "traffic": measured_traffic(traffic_key) if traffic_key else None,
           "kernel": kernel, "kernel_avg_ms": avg_ms, "launches": int(launches),
           "flop_per_pair": F_PAIR, "pairs_per_launch": pairs / max(launches, 1)}
    if peak == PEAK_FP64_VALU:
        # the published peak assumes every slot an FMA at the peak clock; a
        # plain FMA stream shows what this device sustains
        tf, rsq = fma_stream()
        out["fma_stream_tflops"] = tf
        out["frac_of_fma_stream"] = achieved / 1e12 / tf
        out["rsq_f64_time_over_fma"] = rsq
    return out


def bench_microbench(args):
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
    kms, nl, pairs = ctx.profile_get()
    v = ctx.pset_velocity(ps, m)
    out = {
        "metric": "Biot-Savart pair-interactions/s",
        "value": float(len(a)) * m * args.steps / wall, "unit": "pairs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if prec == vlm.FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "Biot-Savart microbenchmark: 1e5 targets x 1e6 helix "
                               "filaments, core radius 0.01, inputs resident in HBM",
                   "precision": args.precision},
        "roofline": roofline(kms, nl, pairs, PEAK_FP64_VALU if prec == vlm.FP64
                             else PEAK_FP32_VALU,
                             "bs_fp64_kernel<256>" if prec == vlm.FP64 else "bs_mixed_kernel<256>",
                             "micro:bs_fp64_kernel<256>" if prec == vlm.FP64 else None),
    }
    if not args.no_cpu:
        from oracle import oracle as orc
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        vc = orc.filament_sum(a, b, gam, rc, p[sample], threads=threads)
        dtc = time.perf_counter() - t0
        e = slice(0, 1000)
        rms = np.sqrt(np.mean(np.sum(vc[e]**2, 1)))
        out["cpu_baseline"] = {"value": len(sample) * len(a) / dtc, "unit": "pairs/s",
                               "cores": threads, "kind": "port",
                               "sample": f"{len(sample)} targets x {len(a)} filaments, "
                                         f"FP64 oracle ({dtc:.1f} s)"}
        out["check"] = {"max_err_over_rms_vs_fp64_oracle":
                        float(np.max(np.linalg.norm(v[sample[e]] - vc[e], axis=1)) / rms)}
    print(json.dumps(out), flush=True)


def case_variant(case_dir, old, new):
    """Copy of a case directory with one namelist line replaced."""
    tmp = tempfile.mkdtemp(prefix="vlmcase_")
    for f in os.listdir(case_dir):
        if os.path.isfile(os.path.join(case_dir, f)):
            shutil.copy(os.path.join(case_dir, f), tmp)
    g = os.path.join(tmp, "geom01.nml")
    text = open(g).read()
    if old not in text:
        raise ValueError(f"{old!r} not in {g}")
    open(g, "w").write(text.replace(old, new))
    return tmp


def run_reference(case_dir, nsteps, threads):
    """Reference solver (compiled from its own sources) on the first nsteps
    steps of the case, in a scratch directory; returns wall seconds."""
    tmp = tempfile.mkdtemp(prefix="vlmref_")
    try:
        for f in os.listdir(case_dir):
            if os.path.isfile(os.path.join(case_dir, f)):
                shutil.copy(os.path.join(case_dir, f), tmp)
        os.makedirs(os.path.join(tmp, "Results"))
        os.makedirs(os.path.join(tmp, "Restart"))
        cfg = open(os.path.join(tmp, "config.nml")).read()
        lines = [f"nt = {nsteps}" if l.strip().startswith("nt =") else l for l in cfg.splitlines()]
        open(os.path.join(tmp, "config.nml"), "w").write("\n".join(lines) + "\n")
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_SCHEDULE="static",
                   OMP_PROC_BIND="close", OMP_PLACES="cores")

        def unlimited_stack():
            resource.setrlimit(resource.RLIMIT_STACK,
                               (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

        t0 = time.perf_counter()
        subprocess.run([REF_BIN], cwd=tmp, env=env, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, check=True, preexec_fn=unlimited_stack,
                       timeout=600)
        return time.perf_counter() - t0
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


WORKLOADS = {
    # name: case, total steps (10 revolutions), default timed tail, steps per
    # revolution, rotors, description
    "rotor2": ("cfg2_rotor.case", 720, 72, 72, 1,
               "two-bladed hover rotor (cfg2): 1 x 20 panels per blade, -8 deg twist, 8 deg "
               "collective, tip Mach 0.4, 5 deg steps, AM2 free wake with strain and "
               "dissipation, 10 revolutions; timed = last K steps"),
    "quad": ("cfg4_quad.case", 370, 20, 36, 4,
             "quad-rotor (cfg4): 4 three-bladed rotors, 1 x 30 panels per blade, 8 deg "
             "collective, alternating rotation, 10 deg steps, AM2 free wake, 15-revolution "
             "truncated lattice; steps 351-370 (10 revolutions) timed"),
    "rotorwing": ("cfg5_rotor_wing.case", 1080, 72, 72, 2,
                  "rotor above a wing (cfg5): three-bladed rotor, 1 x 30 panels per blade, "
                  "10 deg collective, advance ratio 0.1, over an AR-8 wing of 4 x 40 panels "
                  "0.5 R below the hub, 5 deg steps, AM2 free wakes, 15 revolutions; the last "
                  "revolution timed"),
}
# steps of the case the reference solver runs for the CPU baseline (about
# 20 s on 16 host threads)
CPU_STEPS = {"rotor2": 100, "quad": 22, "rotorwing": 0}
# The reference solver turns NaN from the first step on cases whose coupled
# surfaces have different spanwise panel counts (cfg5: 30 and 40), so no
# reference run is timed for them.
NO_REFERENCE = {"rotorwing"}


def host_cores():
    """Physical cores of this host (lscpu: sockets x cores per socket)."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = dict(l.split(":", 1) for l in out.splitlines() if ":" in l)
        return int(kv["Socket(s)"].strip()) * int(kv["Core(s) per socket"].strip())
    except (OSError, KeyError, ValueError, subprocess.SubprocessError):
        return None


def bench_rotor(args):
    from volcanor_amd import host
    case_name, nt_case, K_def, steps_rev, nrot, desc = WORKLOADS[args.workload]
    case = os.path.join(ROOT, "cases", case_name)
    # The workload is the case's own step range: the run always covers all
    # nt_case steps and times the last K of them (the steps before build the
    # wake whose size defines the workload), so --warmup cannot shorten it.
    K = min(args.steps if args.steps else K_def, nt_case)
    nt = nt_case
    W = nt - K

    runs = {}
    plan = (("fp64", "fp64", False), ("mixed", "mixed", False), ("learned_mixed", "mixed", True),
            ("learned_fp64", "fp64", True))
    want = set(args.variants.split(",")) | ({"learned_mixed"} if args.learned else set())
    plan = tuple(v for v in plan if v[0] == "fp64" or v[0] in want)
    for name, prec, learned in plan:
        t0 = time.perf_counter()
        ct, st = host.run_case(case, precision=prec, nt=nt, tail_steps=K, learned=learned,
                               tol=args.tol, nrotors=nrot, fixed_ploop=args.fixed_ploop,
                               verbose=args.progress)
        runs[name] = dict(ct=ct, st=st, wall=time.perf_counter() - t0)
        print(f"# {name}: run {runs[name]['wall']:.2f} s, timed {st.wall_steps_tail:.3f} s for "
              f"{K} steps, BS kernel {st.bs_kernel_ms:.1f} ms over {st.bs_launches} launches, "
              f"pairs {st.bs_pairs:.3e}", file=sys.stderr, flush=True)

    if not args.no_twin:
        # FP64 twin with a 1e-9 deg cyclic pitch on rotor 1: a blade-
        # asymmetric perturbation of relative size ~1e-10 that measures how
        # far the wake's own dynamics amplify round-off-sized differences.
        tdir = case_variant(case, "thetaC = 0.0", "thetaC = 1.0e-9")
        try:
            t0 = time.perf_counter()
            ct, st = host.run_case(tdir, precision="fp64", nt=nt, tail_steps=K, nrotors=nrot,
                                   fixed_ploop=args.fixed_ploop)
            runs["fp64_twin_thetaC_1e-9"] = dict(ct=ct, st=st, wall=time.perf_counter() - t0)
        finally:
            shutil.rmtree(tdir, ignore_errors=True)
    ref = runs["fp64"]["ct"]
    for r in runs.values():
        r["ct_rel"] = float(np.max(np.abs(r["ct"] - ref) / np.maximum(np.abs(ref), 1e-30)))
        dev = np.abs(r["ct"] - ref) / np.abs(ref).max(axis=1, keepdims=True)
        r["ct_dev_per_rev"] = [float(dev[:, 1 + steps_rev * k: 1 + steps_rev * (k + 1)].max())
                               for k in range((dev.shape[1] - 1) // steps_rev)]
    # CT averaged over the last full revolution, per rotor: the observable
    # that the wake's chaotic amplification of round-off leaves most stable
    last = slice(nt + 1 - steps_rev, nt + 1)
    ref_mean = ref[:, last].mean(axis=1)
    for r in runs.values():
        m = r["ct"][:, last].mean(axis=1)
        r["ct_mean_last_rev"] = [float(x) for x in m]
        r["ct_mean_last_rev_rel_dev"] = float(np.max(np.abs(m - ref_mean) /
                                                     np.maximum(np.abs(ref_mean), 1e-30)))
    # value: the plain FP64 path, every Biot-Savart launch of the timed steps
    prim = "fp64"
    P = runs[prim]["st"]
    value = P.bs_pairs_all / P.wall_steps_tail
    traffic_key = f"{args.workload}_tail{K}:bs_fp64_kernel<256>"
    out = {
        "metric": "Biot-Savart pair-interactions/s",
        "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": P.wall_steps_tail * 1e3 / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "case": f"cases/{case_name}", "timed_steps": K,
                   "timed_step_range": [nt - K + 1, nt], "total_steps": nt,
                   "precision": prim,
                   "predicted_wake": ("every row (--fixed-ploop)" if args.fixed_ploop else
                                      "row 1 only, as the reference (classdef.f90:1552-1560)"),
                   "warmup_note": (f"the {W} steps before the timed ones build the wake that "
                                   f"defines the workload; --warmup "
                                   f"{args.warmup if args.warmup is not None else '(unset)'} "
                                   "does not shorten the case")},
        "roofline": roofline(P.bs_kernel_ms, P.bs_launches, P.bs_pairs, PEAK_FP64_VALU,
                             "bs_fp64_kernel<256>", traffic_key),
        "pairs_timed": {"all_launches": P.bs_pairs_all, "launches": int(P.bs_launches_all),
                        "work_queue_kernel": P.bs_pairs,
                        "work_queue_launches": int(P.bs_launches)},
        "variants": {k: {"s_per_step": r["st"].wall_steps_tail / K,
                         "pairs_per_s": r["st"].bs_pairs_all / r["st"].wall_steps_tail,
                         "ct_mean_last_rev": r["ct_mean_last_rev"],
                         "ct_mean_last_rev_rel_dev_vs_fp64": r["ct_mean_last_rev_rel_dev"],
                         "run_wall_s": r["wall"], "ct_final": [float(x) for x in r["ct"][:, -1]],
                         "ct_max_rel_dev_vs_fp64": r["ct_rel"],
                         "ct_dev_per_rev_vs_fp64": r["ct_dev_per_rev"],
                         "bs_kernel_avg_ms": r["st"].bs_kernel_ms / max(r["st"].bs_launches, 1),
                         "predictor": {"full": r["st"].pred_full,
                                       "learned": r["st"].pred_learned,
                                       "fallbacks": r["st"].pred_fallbacks,
                                       "judged": r["st"].pred_judged,
                                       "mlp_rel_rms": r["st"].pred_mlp_rms,
                                       "extrapolation_rel_rms": r["st"].pred_base_rms}}
                     for k, r in runs.items()},
    }
    out["ct_note"] = ("ct_dev_per_rev_vs_fp64: max |CT - CT_fp64| / max|CT_fp64| per revolution "
                      "(over rotors). The fp64 twin differs from fp64 only by a 1e-9 deg cyclic "
                      "pitch; its growth is the wake's own amplification of round-off-sized "
                      "differences.")
    out["speedup_s_per_step_vs_fp64"] = {
        k: runs["fp64"]["st"].wall_steps_tail / r["st"].wall_steps_tail for k, r in runs.items()}
    for lname in ("learned_mixed", "learned_fp64"):
        if lname not in runs:
            continue
        L = runs[lname]
        out["learned_vs_plain" if lname == "learned_mixed" else "learned_fp64_vs_plain"] = {
            "s_per_step_plain_fp64": runs["fp64"]["st"].wall_steps_tail / K,
            "s_per_step_learned": L["st"].wall_steps_tail / K,
            "speedup": runs["fp64"]["st"].wall_steps_tail / L["st"].wall_steps_tail,
            "tol": args.tol,
            "ct_dev_per_rev": L["ct_dev_per_rev"],
            "fp64_twin_ct_dev_per_rev": runs.get("fp64_twin_thetaC_1e-9", {}).get("ct_dev_per_rev"),
            "revs_within_1e-4": sum(1 for x in L["ct_dev_per_rev"] if x <= 1e-4),
            "ct_mean_last_rev_rel_dev": L["ct_mean_last_rev_rel_dev"],
            "fp64_twin_ct_mean_last_rev_rel_dev":
                runs.get("fp64_twin_thetaC_1e-9", {}).get("ct_mean_last_rev_rel_dev"),
            "fallback_rate": (L["st"].pred_fallbacks / L["st"].pred_judged
                              if L["st"].pred_judged > 0 else None),
            "network_gain_over_extrapolation":
                L["st"].pred_base_rms / L["st"].pred_mlp_rms if L["st"].pred_mlp_rms > 0 else None}
    if not args.no_cpu and args.workload in NO_REFERENCE:
        out["cpu_baseline"] = None
        out["cpu_baseline_note"] = ("the reference solver produces NaN loads on this case "
                                    "(coupled surfaces with different spanwise panel counts), "
                                    "so it is not timed")
    elif not args.no_cpu and os.path.exists(REF_BIN):
        S = args.cpu_steps or CPU_STEPS[args.workload]
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        _, sg = host.run_case(case, precision="fp64", nt=S, tail_steps=S, nrotors=nrot)
        wall_cpu = run_reference(case, S, threads)
        phys = host_cores()
        out["cpu_baseline"] = {
            "value": sg.bs_pairs_all / wall_cpu, "unit": "pairs/s", "cores": threads,
            "kind": "reference", "host_physical_cores": phys,
            "sample": f"reference solver built from its sources, first {S} steps of the same "
                      f"case ({wall_cpu:.1f} s, {threads} OpenMP threads on a host with "
                      f"{phys} physical cores); pairs counted as on the GPU over those steps, "
                      f"every launch ({sg.bs_pairs_all:.3e}); GPU FP64 on the same steps took "
                      f"{sg.wall_steps_tail:.3f} s"}
    print(json.dumps(out), flush=True)


LU_N, LU_SEED = 2176, 20261202


def lu_matrix():
    return np.asfortranarray(np.random.default_rng(LU_SEED).standard_normal((LU_N, LU_N)))


def lu_library_baseline():
    """The same factorisation by the vendor libraries that torch.linalg
    reaches on this device (hipSOLVER/rocSOLVER and MAGMA): a comparison,
    not part of the measured path.  torch carries its own HIP runtime, which
    cannot start in a process where libvlm's has, so this runs in a child
    process (bench.py --workload lulib)."""
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--workload", "lulib"],
                           capture_output=True, text=True, timeout=300)
        return json.loads(r.stdout.strip().splitlines()[-1])
    except (subprocess.SubprocessError, ValueError, IndexError) as ex:
        return {"error": str(ex)[:200]}


def bench_lulib(args, reps=5):
    """Child of lu_library_baseline: events on torch's stream around reps
    factorisations per library, after two untimed ones."""
    import torch
    A = lu_matrix()
    n = A.shape[0]
    At = torch.from_numpy(np.ascontiguousarray(A)).cuda()
    res = {}
    prev = torch.backends.cuda.preferred_linalg_library()
    for name, backend in (("rocsolver", "cusolver"), ("magma", "magma")):
        try:
            torch.backends.cuda.preferred_linalg_library(backend)
            for _ in range(2):
                torch.linalg.lu_factor(At)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                torch.linalg.lu_factor(At)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            res[name] = {"ms": ms, "tflops": 2 * n**3 / 3 / (ms * 1e-3) / 1e12}
        except RuntimeError as ex:
            res[name] = {"error": str(ex)[:200]}
    torch.backends.cuda.preferred_linalg_library(prev)
    res["via"] = "torch.linalg.lu_factor on the same matrix (FP64)"
    print(json.dumps(res), flush=True)


def bench_lu(args):
    """Dense LU of an N = 2176 FP64 matrix (the size of a rotor-wing AIC
    rebuilt every step): device time of the factorisation only."""
    from volcanor_amd import vlm
    A = lu_matrix()
    n = A.shape[0]
    lib = None if args.no_library else lu_library_baseline()
    ctx = vlm.Context(0)
    for _ in range(args.warmup):
        ctx.lu_factor(A)
    ctx.profile_enable(True)
    ctx.profile_reset()
    for _ in range(args.steps):
        LU, piv = ctx.lu_factor(A)
    ms, cnt, flops = ctx.profile_get_lu()
    achieved = flops / (ms * 1e-3)
    out = {"metric": "dense LU factorisation, FP64", "value": achieved / 1e12, "unit": "TFLOP/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms / max(cnt, 1), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"LU with partial pivoting, N={n}, normal random matrix"},
           "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": 78.6,
                        "unit": "TFLOP/s", "frac": achieved / 78.6e12, "traffic": None}}
    if lib is not None:
        out["library_baseline"] = lib
    if not args.no_cpu:
        from scipy.linalg import lu_factor
        t0 = time.perf_counter()
        lu_factor(A)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 2 * n**3 / 3 / dt / 1e12, "unit": "TFLOP/s",
                               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                               "kind": "port", "sample": "scipy/LAPACK dgetrf, same matrix"}
        LUl, pivl = lu_factor(A)
        out["check"] = {"pivots_equal_lapack": bool(np.array_equal(piv, pivl)),
                        "max_rel_diff_vs_lapack": float(np.max(np.abs(LU - LUl)) / np.max(np.abs(LUl)))}
    print(json.dumps(out), flush=True)


def bench_hbmcopy(args):
    """HBM calibration: a plain device-to-device copy of 1 GiB (torch's copy
    kernel).  Its rocprof FETCH_SIZE / WRITE_SIZE per dispatch, against the
    known 1 GiB each way, gives the scale factor for the counters."""
    import torch
    n = 1 << 27                      # float64 elements = 1 GiB
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    print(json.dumps({"metric": "HBM copy bandwidth (read + write)", "unit": "GB/s",
                      "value": 2 * 8 * n * args.steps / wall / 1e9, "bytes_each_way": 8 * n,
                      "steps": args.steps, "ms_per_copy": wall * 1e3 / args.steps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", choices=["rotor2", "quad", "rotorwing", "microbench", "lu", "lulib",
                                           "hbmcopy"], default="quad")
    ap.add_argument("--variants", default="mixed,learned_mixed,learned_fp64",
                    help="comma list of variants run beside fp64 (mixed, learned_mixed, "
                         "learned_fp64)")
    ap.add_argument("--fixed-ploop", action="store_true",
                    help="convect every row of the predicted near wake in AM2 (the reference "
                         "convects row 1 only; with it pass 2 sees pass 1's near wake)")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--learned", action="store_true")
    ap.add_argument("--tol", type=float, default=5e-4,
                    help="judge tolerance of the learned pass: RMS error / RMS speed on the "
                         "judge sample above which the pass is recomputed")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-library", action="store_true",
                    help="lu: skip the rocSOLVER/MAGMA comparison")
    ap.add_argument("--progress", type=int, default=0,
                    help="print a line to stdout every N steps of each run (0: none)")
    ap.add_argument("--no-twin", action="store_true",
                    help="skip the perturbed FP64 twin run (chaos floor of CT)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="steps of the case the reference solver runs (0: per-workload default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); a multi-GPU "
              "sharding of the wake is not built.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "lulib":
        bench_lulib(args)
        return
    if args.workload == "lu":
        args.steps = args.steps or 10
        args.warmup = 2 if args.warmup is None else args.warmup
        bench_lu(args)
    elif args.workload == "hbmcopy":
        args.steps = args.steps or 20
        args.warmup = 3 if args.warmup is None else args.warmup
        bench_hbmcopy(args)
    elif args.workload == "microbench":
        if not args.steps:
            args.steps = 5
        if args.warmup is None:
            args.warmup = 1
        bench_microbench(args)
    else:
        bench_rotor(args)


if __name__ == "__main__":
    main()
