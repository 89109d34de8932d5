#!/usr/bin/env python3
"""Benchmark of the smf numeric phase on one MI355X: factor + solve seconds
and GFLOP/s, learned mixed precision vs plain FP64, at scaled backward error
<= 1e-12.

Headline (configuration 4): the SPD 27-point operator on a 100^3 grid
(n = 1e6) with the reference's own METIS ordering (a committed fixture), so
the elimination tree is the configuration's. A step is one numeric
factorization plus one solve (+ FP64 refinement to eta <= 1e-12 for the
reduced-precision paths), values and right-hand side resident in HBM; the
analysis and the device plan are built once, as smf reuses an AnalysisKeep.
Paths: plain FP64; static FP32, split BF16 and BF16 Schur updates on fronts
of >= 1024 rows; and the learned per-supernode selector (precision AUTO).

Beside it: 16 right-hand sides; the configuration-2 KKT (300^2, m = 30000,
the reference's METIS ordering); an interior-point refactor sequence of
configuration-5 KKTs (40 refactors, one analysis, per-refactor inertia and
backward error); the configuration-3 batch of 1024 small KKTs; the
configuration-1 Laplacian; and the reference CPU path timed on the host
cores (oracle/_ref).

Prints one JSON line."""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16": 2.5e15, "hbm": 8.0e12}
MIXED_MIN_FRONT = 1024


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TUNED = None  # the autotuner's choices for this device (opt_solver_amd/tuned), set in main


def tuned(plan, posdef):
    """Apply this device's tuned plan parameters, if the autotuner has
    persisted any (`python -m opt_solver_amd.autotune`)."""
    if TUNED:
        from opt_solver_amd import autotune
        autotune.apply_tuned(plan, TUNED, posdef)
    return plan


def metis_perm(name):
    """The reference's METIS ordering of a benchmark matrix (committed
    fixture written by tests/golden/make_golden.py), or None."""
    f = os.path.join(HERE, "tests", "golden", f"metis_{name}.npz")
    if not os.path.exists(f):
        return None
    z = np.load(f)
    return np.cumsum(z["perm_delta"].astype(np.int64)).astype(np.int32)


def _matvec(A, x):
    n, cp, ri, v = A
    cols = np.repeat(np.arange(n), np.diff(cp))
    y = np.zeros(n)
    np.add.at(y, ri, v * x[cols])
    off = ri != cols
    np.add.at(y, cols[off], v[off] * x[ri[off]])
    return y


def _kkt_perm(gsz, m, A):
    from opt_solver_amd import ordering as o
    n, cp, ri = A[0], A[1], A[2]
    nx = gsz * gsz
    cols = np.repeat(np.arange(n), np.diff(cp))
    sel = (ri >= nx) & (cols < nx)
    order = np.argsort(ri[sel], kind="stable")
    return o.kkt_grid_nd(gsz, m, cols[sel][order].reshape(m, 8), leaf=64)


class Runner:
    """Timed factor + solve (+ refinement) steps of one plan on resident
    data."""

    def __init__(self, plan, vals_d, b_d, nrhs, flops, posdef):
        import torch
        self.torch = torch
        self.plan, self.vals_d, self.b_d, self.nrhs, self.flops = plan, vals_d, b_d, nrhs, flops
        self.x_d = torch.empty_like(b_d)
        self.posdef = posdef

    def step(self, prec, values_ptr=None, u=0.01):
        from opt_solver_amd.solver import PREC_FP64
        fac = self.plan.factor(values_ptr=values_ptr or self.vals_d.data_ptr(), precision=prec,
                               mixed_min_front=MIXED_MIN_FRONT, u=u)
        self.x_d.copy_(self.b_d)
        self.torch.cuda.current_stream().synchronize()  # the library runs on its own stream
        fac.solve_device(self.x_d.data_ptr(), self.nrhs)
        info = {"steps": 0, "fallback": False}
        if prec != PREC_FP64:
            info = fac.refine_device(self.b_d.data_ptr(), self.x_d.data_ptr(), self.nrhs, max_iters=30, tol=1e-12,
                                     max_steps=40)
        return fac, info

    def measure(self, name, prec, steps, warmup, X=None, timed_run=True):
        from opt_solver_amd.solver import PREC_FP64
        torch = self.torch
        for _ in range(warmup):
            self.step(prec)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fac, info = self.step(prec)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        chk = fac.refine_device(self.b_d.data_ptr(), self.x_d.data_ptr(), self.nrhs, max_iters=0)
        pr = fac.precisions()
        out = {"ms_per_step": round(dt * 1e3, 3), "gflops": round(self.flops / dt / 1e9, 1), "eta": chk["eta"],
               "omega": chk["omega"], "refine_steps": info.get("steps", 0),
               "inertia": list(fac.inertia), "fronts_by_precision": {str(int(k)): int(v) for k, v in
                                                                     zip(*np.unique(pr, return_counts=True))}}
        if X is not None:
            n = self.plan.n
            xs = self.x_d.cpu().numpy().reshape((n, self.nrhs), order="F")
            out["fwd_err"] = float(np.abs(xs - X).max())
        if timed_run:
            # one instrumented run: per-kernel-family device times
            self.plan.set_timing(True)
            fac2 = self.plan.factor(values_ptr=self.vals_d.data_ptr(), precision=prec,
                                    mixed_min_front=MIXED_MIN_FRONT)
            self.x_d.copy_(self.b_d)
            torch.cuda.current_stream().synchronize()
            fac2.solve_device(self.x_d.data_ptr(), self.nrhs)
            kt = self.plan.kernel_times()
            self.plan.set_timing(False)
            out["factor_ms"] = round(fac2.seconds * 1e3, 2)
            out["kt"] = kt
            out["prec_counts"] = pr
        if prec == PREC_FP64:
            out.pop("refine_steps")
        log(name, json.dumps({k: v for k, v in out.items() if k not in ("kt", "prec_counts")}))
        return out


def c4_measure(args, torch):
    from opt_solver_amd import generators as g, ordering as o
    from opt_solver_amd.solver import PREC_AUTO, PREC_BF16, PREC_BF16X3, PREC_FP32, PREC_FP64, Plan, analyse
    dev = torch.device("cuda:0")
    m = args.m
    t0 = time.perf_counter()
    A = g.spd27_3d(m, seed=0)
    perm = metis_perm(f"spd27_{m}") if args.ordering == "metis" else None
    ordering = "METIS (the reference's, fixture)" if perm is not None else "geometric nested dissection"
    if perm is None:
        perm = o.grid_nd((m, m, m), leaf=64)
    an = analyse(A, perm=perm, nemin=32)
    plan = tuned(Plan(an, posdef=True), True)
    ps = plan.stats()
    log(f"c4 analysis+plan {time.perf_counter() - t0:.1f}s nsup={an.nsup} maxfront={an.max_front} "
        f"flops={ps['flops']:.3e} levels={int(ps['levels'])} launches={int(ps['launches'])}")
    n = A[0]
    res = {}
    for nrhs in ([1, 16] if not args.no_rhs16 else [1]):
        rng = np.random.default_rng(0)
        X = rng.uniform(-1, 1, (n, nrhs))
        B = np.stack([_matvec(A, X[:, r]) for r in range(nrhs)], axis=1)
        vals_d = torch.from_numpy(np.ascontiguousarray(A[3])).to(dev)
        b_d = torch.from_numpy(np.asfortranarray(B).reshape(-1, order="F").copy()).to(dev)
        torch.cuda.synchronize()
        run = Runner(plan, vals_d, b_d, nrhs, ps["flops"], True)
        legs = ([("fp64", PREC_FP64), ("fp32", PREC_FP32), ("bf16x3", PREC_BF16X3), ("bf16", PREC_BF16),
                 ("learned", PREC_AUTO)] if nrhs == 1 else [("fp64", PREC_FP64), ("learned", PREC_AUTO)])
        steps = args.steps if nrhs == 1 else max(1, args.steps // 2)
        res[nrhs] = {name: run.measure(f"c4 nrhs={nrhs} {name}", prec, steps, args.warmup, X=X,
                                       timed_run=(nrhs == 1))
                     for name, prec in legs}
    return an, ps, ordering, res


def kkt_measure(args, torch):
    """Configuration 2's KKT (300^2 grid, 30000 constraints, u = 0.01):
    FP64, FP32 and learned steps on the library's ordering (geometric
    nested dissection of the grid and constraints + cost-model
    amalgamation); and one FP64 factor + solve on the reference's METIS
    tree, whose tall fronts (up to 38k rows over a few thousand eligible
    columns) stall the threshold search of many panels."""
    from opt_solver_amd import generators as g
    from opt_solver_amd.solver import PREC_AUTO, PREC_FP32, PREC_FP64, Plan, analyse
    gsz, m = 300, 30000
    A, _ = g.kkt_grid(gsz, m, seed=1)
    dev = torch.device("cuda:0")
    x = np.random.default_rng(0).uniform(-1, 1, A[0])
    vals_d = torch.from_numpy(np.ascontiguousarray(A[3])).to(dev)
    b_d = torch.from_numpy(_matvec(A, x)).to(dev)
    torch.cuda.synchronize()
    an = analyse(A, perm=_kkt_perm(gsz, m, A), nemin=32)
    plan = tuned(Plan(an, posdef=False), False)
    run = Runner(plan, vals_d, b_d, 1, an.flops, False)
    out = {}
    for name, prec in (("fp64", PREC_FP64), ("fp32", PREC_FP32), ("learned", PREC_AUTO)):
        out[name] = run.measure(f"kkt {name}", prec, args.steps, args.warmup, timed_run=False)
    res = {"workload": f"kkt_grid {gsz}^2, m={m} (n={A[0]}), u=0.01, geometric ND + cost-model amalgamation",
           "nsup": an.nsup, "max_front": an.max_front, "flops_per_factor_delay_free": an.flops, **out,
           "speedup_learned_vs_fp64": round(out["fp64"]["ms_per_step"] / out["learned"]["ms_per_step"], 3)}
    perm = metis_perm("kkt_300")
    if perm is not None and not args.no_kkt_metis:
        del plan, run
        anm = analyse(A, perm=perm, nemin=32)
        planm = tuned(Plan(anm, posdef=False), False)
        runm = Runner(planm, vals_d, b_d, 1, anm.flops, False)
        r = runm.measure("kkt metis fp64", PREC_FP64, 1, 0, timed_run=False)
        res["metis_fp64"] = {"workload": "the same KKT on the reference's METIS ordering (fixture)",
                             "nsup": anm.nsup, "max_front": anm.max_front,
                             "flops_per_factor_delay_free": anm.flops, **r}
    return res


def seq_measure(args, torch):
    """Interior-point refactor sequence (configuration 5's generator): KKT
    [[H + Sigma_k, A^T], [A, -1e-8 I]] on a G^2 grid with m = 0.32 G^2
    constraints, log10 Sigma_k = log10 Sigma_0 + k s; one analysis, 40
    refactors per path; every refactor: factor + solve (+ refinement),
    inertia checked against (n_x, m, 0) and the scaled backward error."""
    from opt_solver_amd import generators as g
    from opt_solver_amd.solver import PREC_AUTO, PREC_FP32, PREC_FP64, Plan, analyse
    G = args.seq_grid
    m = int(round(0.32 * G * G))
    nx = G * G
    A0, s0 = g.kkt_grid(G, m, seed=5)
    an = analyse(A0, perm=_kkt_perm(G, m, A0), nemin=32)
    plan = tuned(Plan(an, posdef=False), False)
    dev = torch.device("cuda:0")
    nref = args.seq_refactors
    vals = [torch.from_numpy(g.kkt_grid_refactor_values(A0, g.kkt_sigma_sequence(s0, k, seed=6), nx)).to(dev)
            for k in range(nref)]
    x = np.random.default_rng(1).uniform(-1, 1, A0[0])
    bs = [torch.from_numpy(_matvec((A0[0], A0[1], A0[2], v.cpu().numpy()), x)).to(dev) for v in vals]
    torch.cuda.synchronize()
    out = {}
    paths = (("fp64", PREC_FP64), ("fp32", PREC_FP32), ("learned", PREC_AUTO))
    for name, prec in [pp for pp in paths if pp[0] in args.seq_paths.split(",")]:
        run = Runner(plan, vals[0], bs[0], 1, an.flops, False)
        run.step(prec)  # warm-up
        times, etas, steps, flops, ok = [], [], [], 0.0, True
        for k in range(nref):
            run.vals_d, run.b_d = vals[k], bs[k]
            torch.cuda.synchronize()
            t = time.perf_counter()
            fac, info = run.step(prec)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
            chk = fac.refine_device(bs[k].data_ptr(), run.x_d.data_ptr(), 1, max_iters=0)
            etas.append(chk["eta"])
            steps.append(info.get("steps", 0))
            flops += fac.flops
            ok = ok and tuple(fac.inertia) == (nx, m, 0)
            if (k + 1) % 10 == 0:
                log("seq", name, f"refactor {k + 1}/{nref}: {times[-1] * 1e3:.1f} ms, eta {etas[-1]:.2e}")
        tot = sum(times)
        out[name] = {"total_s": round(tot, 3), "ms_per_refactor": round(tot / nref * 1e3, 2),
                     "gflops": round(flops / tot / 1e9, 1), "inertia_ok_every_refactor": ok,
                     "eta_max": max(etas), "refine_steps_max": max(steps)}
        log("seq", name, out[name])
    return {"workload": f"kkt_grid {G}^2, m={m} (n={A0[0]}), {nref} refactors of Sigma_k, u=0.01, "
                        f"geometric ND of the grid + constraints", "nsup": an.nsup, "max_front": an.max_front,
            "flops_per_factor_delay_free": an.flops, **out,
            "speedup_learned_vs_fp64": (round(out["fp64"]["total_s"] / out["learned"]["total_s"], 3)
                                        if "fp64" in out and "learned" in out else None)}


def batch_systems(nsys):
    from opt_solver_amd import batch
    return batch.ocp_batch(nsys, 10, 200, seed=0)


def batch_measure(steps, warmup, nsys=1024):
    """Configuration 3: nsys independent OCP KKT systems as one forest
    (every level's fronts of all systems in the same launches), values and
    right-hand sides resident on the device; the library's analysis of each
    system, built once."""
    import torch
    from opt_solver_amd import batch
    from opt_solver_amd.solver import Plan, analyse
    t0 = time.perf_counter()
    mats, Ts = batch_systems(nsys)
    U, offs = batch.union_arrays([analyse(A, nemin=16).arrays() for A in mats])
    plan = tuned(Plan(U, posdef=False), False)
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


def c1_measure(steps, warmup):
    """Configuration 1: FP64 Cholesky of the 7-point Laplacian on 20^3
    (latency-bound), the library's geometric nested dissection."""
    import torch
    from opt_solver_amd import generators as g, ordering as o
    from opt_solver_amd.solver import PREC_FP64, Plan, analyse
    A = g.laplace7_3d(20)
    an = analyse(A, perm=o.grid_nd((20, 20, 20), leaf=32), nemin=32)
    plan = tuned(Plan(an, posdef=True), True)
    dev = torch.device("cuda:0")
    X = np.random.default_rng(0).uniform(-1, 1, (A[0], 1))
    vals_d = torch.from_numpy(A[3]).to(dev)
    b_d = torch.from_numpy(_matvec(A, X[:, 0])).to(dev)
    run = Runner(plan, vals_d, b_d, 1, an.flops, True)
    r = run.measure("c1 fp64", PREC_FP64, steps, warmup, X=X, timed_run=False)
    return {"workload": "laplace7 20^3 (n=8000), FP64 Cholesky, geometric ND", "nsup": an.nsup,
            "flops_per_factor": an.flops, **r}


def cpu_baseline(m_cpu, threads):
    """The reference CPU path (oracle/_ref, smf built from its sources with
    its portable BLAS) on the same operator at a smaller grid: factor +
    solve, GFLOP/s by the reference's own flop count."""
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


def batch_cpu_baseline(sample, threads):
    """The reference's Solver::factor + solve of the first `sample` systems
    of the batch, one system per call, on `threads` host threads."""
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


def autotune_workloads(args):
    """The autotuner on this benchmark's workloads: configuration 4 on the
    reference's METIS ordering (the timed one), the grid KKT on 300^2 with
    the geometric ordering, the 1024-system batch."""
    from opt_solver_amd import autotune, batch, generators as g, ordering as o
    from opt_solver_amd.solver import analyse
    wl = {}
    A = g.spd27_3d(args.m, seed=0)
    perm = metis_perm(f"spd27_{args.m}") if args.ordering == "metis" else None
    if perm is None:
        perm = o.grid_nd((args.m, args.m, args.m), leaf=64)
    # the right-hand side of the timed legs (c4_measure: x from default_rng(0))
    wl["spd"] = autotune.workload(A, analyse(A, perm=perm, nemin=32).arrays(), True, seed=0)
    gsz, m = 300, 30000
    K, _ = g.kkt_grid(gsz, m, seed=1)
    wl["kkt"] = autotune.workload(K, analyse(K, perm=_kkt_perm(gsz, m, K), nemin=32).arrays(), False, seed=0)
    mats, _ = batch_systems(1024)
    U, _ = batch.union_arrays([analyse(Ai, nemin=16).arrays() for Ai in mats])
    wl["batch"] = autotune.workload((int(U["n"]), None, None, np.concatenate([Ai[3] for Ai in mats])), U, False)
    autotune.autotune(workloads=wl, log=log)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--m", type=int, default=100, help="grid side of configuration 4 (n = m^3)")
    ap.add_argument("--ordering", choices=("metis", "geometric"), default="metis")
    ap.add_argument("--cpu-m", type=int, default=76)
    ap.add_argument("--seq-grid", type=int, default=450,
                    help="grid side of the refactor sequence (450: the largest that fits one GPU)")
    ap.add_argument("--seq-refactors", type=int, default=40)
    ap.add_argument("--seq-paths", default="fp64,fp32,learned", help="paths of the refactor sequence")
    ap.add_argument("--only-seq", action="store_true", help="run the refactor sequence only (no SPD legs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rhs16", action="store_true")
    ap.add_argument("--no-kkt", action="store_true")
    ap.add_argument("--no-kkt-metis", action="store_true")
    ap.add_argument("--no-seq", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the plan parameters on this benchmark's workloads (configuration 4 on the "
                         "reference's METIS ordering, the grid KKT on 300^2, the 1024-system batch), write "
                         "opt_solver_amd/tuned/<device>.json and exit")
    ap.add_argument("--copy-calib", action="store_true",
                    help="run one plain 1 GiB device copy first (scale for the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU path is a design, not built",
              file=sys.stderr)
        sys.exit(2)
    import torch  # device memory for the resident inputs
    if not torch.cuda.is_available():
        print("no GPU", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    global TUNED
    from opt_solver_amd import autotune
    if args.autotune:
        autotune_workloads(args)
        return
    TUNED = autotune.load_tuned(autotune.device_name())
    if TUNED:
        log("tuned plan parameters:", json.dumps(TUNED["knobs"]))
    if args.copy_calib:
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.only_seq:
        # the refactor sequence alone (a long run at a larger grid): one JSON
        # line with the sequence's own numbers
        seq = seq_measure(args, torch)
        tot = seq["learned"]["total_s"] if "learned" in seq else seq["fp64"]["total_s"]
        print(json.dumps({"metric": "factor+solve s (40-refactor IPM sequence)", "value": round(tot, 3),
                          "unit": "s", "n_gpus": 1, "higher_is_better": False, "data": "synthetic",
                          "config": {"workload": seq["workload"]}, "ipm_sequence": seq}))
        return
    an, ps, ordering, res = c4_measure(args, torch)
    r1 = res[1]
    learned, fp64 = r1["learned"], r1["fp64"]
    flops = ps["flops"]
    # roofline of the dominant kernel (the Schur update) in the learned step
    kt = learned["kt"]
    syrk_avg_s = kt["syrk_ms"] / max(kt["syrk_launches"], 1) * 1e-3
    achieved = (kt["syrk_flops"] / max(kt["syrk_launches"], 1)) / syrk_avg_s if syrk_avg_s > 0 else 0.0
    pr = learned["prec_counts"]
    # the peak of the reduced precision the selector gave most fronts (BF16
    # MFMA for split and plain BF16)
    red = pr[pr != 0]
    top = int(np.bincount(red).argmax()) if red.size else 0
    peak = {0: PEAK["fp64"], 1: PEAK["fp32"], 2: PEAK["bf16"], 4: PEAK["bf16"]}[top]
    traffic = None
    tfile = os.path.join(HERE, "profiles", "k_syrk_spd27_100cube_traffic.txt")
    if args.m == 100 and os.path.exists(tfile):
        for line in open(tfile):
            if line.startswith("traffic_bytes_per_launch_mixed"):
                traffic = int(line.split("=")[1])
    extra = {}
    if not args.no_kkt:
        extra["kkt"] = kkt_measure(args, torch)
    if not args.no_seq:
        extra["ipm_sequence"] = seq_measure(args, torch)
    if not args.no_batch:
        extra["batch"] = batch_measure(args.steps, args.warmup)
    if not args.no_c1:
        extra["c1"] = c1_measure(max(args.steps, 5), args.warmup)
    cpu = None
    if not args.no_cpu:
        try:
            threads = min(os.cpu_count() or 1, 16)
            cpu = cpu_baseline(args.cpu_m, threads)
            if cpu is not None and "batch" in extra:
                cpu["batch"] = batch_cpu_baseline(1024, threads)
        except Exception as e:  # the baseline is informative only
            log("cpu baseline failed:", e)

    def strip(d):
        return {k: v for k, v in d.items() if k not in ("kt", "prec_counts")}
    out = {
        "metric": "factor+solve s and GFLOP/s: learned-mixed vs FP64 GPU at bwd err <=1e-12",
        "value": round(flops / (learned["ms_per_step"] * 1e-3) / 1e9, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": learned["ms_per_step"],
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64 pivots/D/residuals; Schur updates per front f64|f32|bf16x3|bf16 (learned selector)",
        "data": "synthetic (seeded 27-point operator, b = A x*, x* ~ U[-1,1]); selector weights trained on "
                "smaller operators (opt_solver_amd/selector_weights.json)",
        "config": {"workload": f"spd27_{args.m}^3 (configuration 4), {ordering}", "n": int(an.n), "nrhs": 1,
                   "nsup": an.nsup, "max_front": an.max_front, "flops_per_factor": flops,
                   "mixed_min_front": MIXED_MIN_FRONT, "selector_tol": 1e-4},
        "learned": strip(learned),
        "fp64": strip(fp64),
        "static": {k: strip(r1[k]) for k in ("fp32", "bf16x3", "bf16")},
        "speedup_learned_vs_fp64": round(fp64["ms_per_step"] / learned["ms_per_step"], 3),
        "nrhs16": {k: strip(v) for k, v in res.get(16, {}).items()},
        "roofline": {"bound": "mfma", "kernel": "k_syrk / k_syrk_red (Schur update), learned step",
                     "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic,
                     "traffic_source": "profiles/k_syrk_spd27_100cube_traffic.txt: rocprofv3 FETCH_SIZE x 2.00 "
                                       "+ WRITE_SIZE per Schur launch pair of the learned step (METIS tree), "
                                       "calibrated on a 1 GiB device copy; not measured in this run"},
        "cpu_baseline": cpu,
        "tuned": ({"file": os.path.relpath(autotune.tuned_path(TUNED["device"]), HERE), "knobs": TUNED["knobs"]}
                  if TUNED else None),
        **extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
