#!/usr/bin/env python3
"""Benchmark: numeric factor + solve of the SPD 27-point operator on a 100^3
grid (n = 1e6) on one MI355X, plain FP64 vs mixed precision (FP32 MFMA Schur
updates on the large fronts + FP64 iterative refinement), at scaled backward
error <= 1e-12.

A step is one numeric factorization + one solve (+ refinement for the mixed
path) of the whole system, values and right-hand side already resident in
HBM. The symbolic analysis (host) and the device plan are built once before
timing, as smf reuses an AnalysisKeep across refactors.

Beside it, the same comparison on an indefinite KKT system of the
configuration-2 generator (300^2 grid, 30,000 constraints, n = 120,000):
threshold-pivoted LDL^T in FP64 vs FP32 MFMA updates + FP64 refinement.

Prints one JSON line."""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "hbm": 8.0e12}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_problem(m, seed=0):
    from opt_solver_amd import generators as g, ordering as o
    A = g.spd27_3d(m, seed=seed)
    perm = o.grid_nd((m, m, m), leaf=64)
    return A, perm


def cpu_baseline(m_cpu, threads):
    """The reference CPU path (oracle/_ref, smf built from its sources with its
    portable BLAS) on the same operator at a smaller grid: factor + solve,
    GFLOP/s by the reference's own flop count."""
    from oracle import refshim as R
    if not R.available():
        return None
    from opt_solver_amd import generators as g
    A = g.spd27_3d(m_cpu, seed=0)
    an = R.RefAnalysis(A, ordering=R.ORDER_METIS, posdef=True)
    x = np.random.default_rng(0).uniform(-1, 1, A[0])
    b = _matvec(A, x)
    t = time.perf_counter()
    rf = R.RefFactor(an, A[3], threads=threads)
    xs = rf.solve(b)
    dt = time.perf_counter() - t
    return {"value": round(rf.flops / dt / 1e9, 3), "unit": "GFLOP/s", "cores": threads, "kind": "reference",
            "sample": f"smf 0.2.3 Solver::factor+solve (METIS, portable reference BLAS) of the same 27-point "
                      f"operator on a {m_cpu}^3 grid (n={A[0]}), {dt:.2f} s, {rf.flops:.3e} flops, "
                      f"fwd err {np.abs(xs - x).max():.1e}"}


def kkt_measure(steps, warmup, gsz=300, m=30000):
    """FP64 vs mixed LDL^T factor + solve (+ refinement) of the KKT system
    of the configuration-2 generator, values and right-hand side resident on
    the device; the library's ordering and analysis, built once."""
    import torch
    from opt_solver_amd import generators as g, ordering as o
    from opt_solver_amd.solver import PREC_FP32, PREC_FP64, Plan, analyse
    A, _ = g.kkt_grid(gsz, m, seed=1)
    n, cp, ri = A[0], A[1], A[2]
    nx = gsz * gsz
    cols = np.repeat(np.arange(n), np.diff(cp))
    sel = (ri >= nx) & (cols < nx)
    order = np.argsort(ri[sel], kind="stable")
    acols = cols[sel][order].reshape(m, 8)
    an = analyse(A, perm=o.kkt_grid_nd(gsz, m, acols, leaf=64), nemin=32)
    plan = Plan(an, posdef=False)
    dev = torch.device("cuda:0")
    x = np.random.default_rng(0).uniform(-1, 1, n)
    vals_d = torch.from_numpy(np.ascontiguousarray(A[3])).to(dev)
    b_d = torch.from_numpy(_matvec(A, x)).to(dev)
    x_d = torch.empty_like(b_d)
    torch.cuda.synchronize()
    out = {}
    for name, prec in (("fp64", PREC_FP64), ("mixed", PREC_FP32)):
        def step():
            fac = plan.factor(values_ptr=vals_d.data_ptr(), u=0.01, precision=prec, mixed_min_front=1024)
            x_d.copy_(b_d)
            torch.cuda.current_stream().synchronize()
            fac.solve_device(x_d.data_ptr(), 1)
            info = {"steps": 0}
            if prec != PREC_FP64:
                info = fac.refine_device(b_d.data_ptr(), x_d.data_ptr(), 1, max_iters=10, tol=1e-12)
            return fac, info
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fac, info = step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        chk = fac.refine_device(b_d.data_ptr(), x_d.data_ptr(), 1, max_iters=0)
        out[name] = {"ms_per_step": round(dt * 1e3, 2), "gflops": round(fac.info.actual_flops / dt / 1e9, 1),
                     "eta": chk["eta"], "refine_steps": info.get("steps", 0), "inertia": list(fac.inertia),
                     "delayed": int(fac.info.delayed_pivots), "fronts_reduced": int(fac.info.fronts_reduced)}
        log("kkt", name, out[name])
    return {"workload": f"kkt_grid {gsz}^2, m={m} (n={n}), u=0.01, geometric ND + cost-model amalgamation",
            "max_front": an.max_front, "flops_per_factor": fac.info.actual_flops, "fp64": out["fp64"],
            "mixed": out["mixed"],
            "speedup_mixed_vs_fp64": round(out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"], 3)}


def batch_systems(nsys):
    """The configuration-3 batch: nsys trajectory-optimisation KKTs with
    horizons T ~ U{10..200}, seeded."""
    from opt_solver_amd import batch
    return batch.ocp_batch(nsys, 10, 200, seed=0)


def batch_measure(steps, warmup, nsys=1024):
    """Factor + solve of nsys independent OCP KKT systems as one forest
    (every level's fronts of all systems in the same launches), values and
    right-hand sides resident on the device; the library's analysis of each
    system, built once."""
    import torch
    from opt_solver_amd import batch
    from opt_solver_amd.solver import Plan, analyse
    t0 = time.perf_counter()
    mats, Ts = batch_systems(nsys)
    U, offs = batch.union_arrays([analyse(A, nemin=16).arrays() for A in mats])
    plan = Plan(U, posdef=False)
    vals = np.concatenate([A[3] for A in mats])
    X = np.concatenate([np.random.default_rng(i).uniform(-1, 1, A[0]) for i, A in enumerate(mats)])
    starts = np.cumsum([0] + [A[0] for A in mats])
    b = np.concatenate([_matvec(A, X[starts[i]:starts[i + 1]]) for i, A in enumerate(mats)])
    dev = torch.device("cuda:0")
    vals_d = torch.from_numpy(vals).to(dev)
    b_d = torch.from_numpy(b).to(dev)
    x_d = torch.empty_like(b_d)
    torch.cuda.synchronize()
    log(f"batch set-up {time.perf_counter() - t0:.1f}s n={U['n']} nsup={U['sn_start'].size}")

    def step():
        fac = plan.factor(values_ptr=vals_d.data_ptr(), u=0.01)
        x_d.copy_(b_d)
        torch.cuda.current_stream().synchronize()
        fac.solve_device(x_d.data_ptr(), 1)
        return fac
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        fac = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / steps
    per = batch.system_stats(fac.front_stats(), offs["sn"])
    inertia_ok = all(tuple(per[i, :3]) == (16 * T + 12, 12 * T, 0) for i, T in enumerate(Ts))
    err = float(np.abs(x_d.cpu().numpy() - X).max())
    kt = plan.kernel_times()
    out = {"workload": f"{nsys} OCP KKTs, T ~ U{{10..200}} (n={U['n']}), u=0.01, one forest",
           "nsup": int(U["sn_start"].size), "levels": int(plan.stats()["levels"]),
           "device_levels": int(kt["device_levels"]), "ms_per_step": round(dt * 1e3, 2),
           "systems_per_s": round(nsys / dt, 1), "flops_per_factor": fac.info.actual_flops,
           "delayed": int(fac.info.delayed_pivots), "inertia_ok": inertia_ok, "fwd_err": err}
    log("batch", out)
    return out


def batch_cpu_baseline(sample, threads):
    """The reference's Solver::factor + solve of the first `sample` systems
    of the batch, one system per call, on `threads` host threads (each call
    single-threaded; the calls run concurrently)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import refshim
    if not refshim.available():
        return None
    mats, _ = batch_systems(sample)
    ans = [refshim.RefAnalysis(A, ordering=refshim.ORDER_AMD, posdef=False) for A in mats]
    rhs = [_matvec(A, np.ones(A[0])) for A in mats]

    def one(i):
        rf = refshim.RefFactor(ans[i], mats[i][3], u=0.01, threads=1)
        rf.solve(rhs[i])
        return rf.flops
    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        fl = sum(ex.map(one, range(sample)))
    dt = time.perf_counter() - t
    return {"value": round(sample / dt, 1), "unit": "systems/s", "cores": threads, "kind": "reference",
            "sample": f"smf 0.2.3 Solver::factor+solve of the first {sample} systems of the batch, one system "
                      f"per call on {threads} threads, AMD (fallback minimum degree), {dt:.2f} s, "
                      f"{fl:.3e} flops"}


def _matvec(A, x):
    n, cp, ri, v = A
    cols = np.repeat(np.arange(n), np.diff(cp))
    y = np.zeros(n)
    np.add.at(y, ri, v * x[cols])
    off = ri != cols
    np.add.at(y, cols[off], v[off] * x[ri[off]])
    return y


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--m", type=int, default=100, help="grid side (n = m^3)")
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--cpu-m", type=int, default=76)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-kkt", action="store_true", help="skip the indefinite KKT comparison")
    ap.add_argument("--no-batch", action="store_true", help="skip the batch of small systems")
    ap.add_argument("--copy-calib", action="store_true",
                    help="run one plain 1 GiB device copy first (scale for the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU path is a design, not built",
              file=sys.stderr)
        sys.exit(2)

    import torch  # device memory for the resident inputs
    from opt_solver_amd.solver import PREC_FP32, PREC_FP64, Plan, analyse

    if not torch.cuda.is_available():
        print("no GPU", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    if args.copy_calib:
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    m = args.m
    t0 = time.perf_counter()
    A, perm = build_problem(m)
    n = A[0]
    an = analyse(A, perm=perm, nemin=32)
    plan = Plan(an, posdef=True)
    ps = plan.stats()
    log(f"analysis+plan {time.perf_counter() - t0:.1f}s nsup={an.nsup} maxfront={an.max_front} "
        f"flops={ps['flops']:.3e} levels={int(ps['levels'])} launches={int(ps['launches'])}")
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, (n, args.nrhs))
    B = np.stack([_matvec(A, X[:, r]) for r in range(args.nrhs)], axis=1)
    vals_d = torch.from_numpy(np.ascontiguousarray(A[3])).to(dev)
    b_d = torch.from_numpy(np.asfortranarray(B).reshape(-1, order="F").copy()).to(dev)
    x_d = torch.empty_like(b_d)
    torch.cuda.synchronize()

    def step(prec):
        fac = plan.factor(values_ptr=vals_d.data_ptr(), precision=prec, mixed_min_front=1024)
        x_d.copy_(b_d)
        torch.cuda.current_stream().synchronize()  # the library runs on its own stream
        fac.solve_device(x_d.data_ptr(), args.nrhs)
        info = {"eta": None, "steps": 0}
        if prec != PREC_FP64:
            info = fac.refine_device(b_d.data_ptr(), x_d.data_ptr(), args.nrhs, max_iters=10, tol=1e-12)
        return fac, info

    results = {}
    for name, prec in (("fp64", PREC_FP64), ("mixed", PREC_FP32)):
        for _ in range(args.warmup):
            step(prec)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            fac, info = step(prec)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        # accuracy (outside the timed region): device refinement's eta of the
        # final x (no further steps)
        chk = fac.refine_device(b_d.data_ptr(), x_d.data_ptr(), args.nrhs, max_iters=0)
        xs = x_d.cpu().numpy().reshape((n, args.nrhs), order="F")
        # timed per-kernel breakdown: one extra instrumented run
        plan.set_timing(True)
        fac2 = plan.factor(values_ptr=vals_d.data_ptr(), precision=prec, mixed_min_front=1024)
        x_d.copy_(b_d)
        torch.cuda.current_stream().synchronize()
        fac2.solve_device(x_d.data_ptr(), args.nrhs)
        kt = plan.kernel_times()
        plan.set_timing(False)
        results[name] = dict(ms_per_step=dt * 1e3, factor_s=fac2.seconds, solve_ms=kt["solve_ms"],
                             eta=chk["eta"], omega=chk["omega"], refine_steps=info.get("steps", 0),
                             fwd_err=float(np.abs(xs - X).max()), kt=kt, flops=fac.flops,
                             fronts_reduced=int(fac.info.fronts_reduced))
        log(name, json.dumps({k: v for k, v in results[name].items() if k != "kt"}), kt)

    flops = results["fp64"]["flops"]
    mixed = results["mixed"]
    fp64 = results["fp64"]
    kt = mixed["kt"]
    syrk_avg_s = kt["syrk_ms"] / max(kt["syrk_launches"], 1) * 1e-3
    syrk_flops_per_launch = kt["syrk_flops"] / max(kt["syrk_launches"], 1)
    achieved = syrk_flops_per_launch / syrk_avg_s if syrk_avg_s > 0 else 0.0
    # HBM bytes per mixed-phase k_syrk launch: the committed summary of the
    # two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) over this
    # workload, scaled by a plain copy's readings
    traffic = None
    tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "k_syrk_spd27_100cube_traffic.txt")
    if m == 100 and args.nrhs == 1 and os.path.exists(tfile):
        for line in open(tfile):
            if line.startswith("traffic_bytes_per_launch_mixed"):
                traffic = int(line.split("=")[1])
    kkt = None
    if not args.no_kkt:
        kkt = kkt_measure(args.steps, args.warmup)
    bat = None
    if not args.no_batch:
        bat = batch_measure(args.steps, args.warmup)
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args.cpu_m, min(os.cpu_count() or 1, 16))
            if bat is not None:
                cpu["batch"] = batch_cpu_baseline(1024, min(os.cpu_count() or 1, 16))
        except Exception as e:  # the baseline is informative only
            log("cpu baseline failed:", e)
    out = {
        "metric": "factor+solve GFLOP/s (mixed FP32-MFMA Schur + FP64 refinement vs FP64), SPD 27-pt 100^3, "
                  "bwd err <= 1e-12",
        "value": round(flops / (mixed["ms_per_step"] * 1e-3) / 1e9, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic (seeded 27-point operator, b = A x*, x* ~ U[-1,1])",
        "config": {"workload": f"spd27_{m}^3", "n": n, "nrhs": args.nrhs, "ordering": "geometric ND",
                   "nsup": an.nsup, "max_front": an.max_front, "flops_per_factor": flops,
                   "mixed_policy": "FP32 Schur on fronts >= 1024 rows, FP64 IR"},
        "mixed": {"ms_per_step": round(mixed["ms_per_step"], 3), "factor_s": round(mixed["factor_s"], 4),
                  "solve_ms": round(mixed["solve_ms"], 3), "eta": mixed["eta"], "omega": mixed["omega"],
                  "refine_steps": mixed["refine_steps"], "fronts_reduced": mixed["fronts_reduced"]},
        "fp64": {"ms_per_step": round(fp64["ms_per_step"], 3), "factor_s": round(fp64["factor_s"], 4),
                 "solve_ms": round(fp64["solve_ms"], 3), "eta": fp64["eta"], "omega": fp64["omega"],
                 "gflops": round(flops / (fp64["ms_per_step"] * 1e-3) / 1e9, 2)},
        "speedup_mixed_vs_fp64": round(fp64["ms_per_step"] / mixed["ms_per_step"], 3),
        "roofline": {"bound": "mfma", "kernel": "k_syrk (Schur update)", "achieved": achieved / 1e12,
                     "peak": PEAK["fp32"] / 1e12, "unit": "TFLOP/s", "frac": achieved / PEAK["fp32"],
                     "traffic": traffic},
        "cpu_baseline": cpu,
        "kkt": kkt,
        "batch": bat,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
