"""Autotuner of the per-plan kernel parameters, by front-size bucket.

The numeric phase has one tunable per bucket of fronts (and per SPD
schedule stage); `smfgpu_plan_tune` sets them on a plan. This module times
factor + solve on representative workloads of each bucket on the current
device, walks the buckets' parameters one at a time over candidate values
(coordinate descent, every other parameter at its best value so far), and
persists the choices for the device as a small JSON file under
`opt_solver_amd/tuned/`. `load_tuned` / `apply_tuned` read it back; a plan
without them keeps the library's defaults.

Every parameter selects between schedules or tile shapes that compute the
same factorization (the same pivots; sums in another order at most, checked
by the parity tests), so tuning changes time only.

Usage (on the GPU):  python -m opt_solver_amd.autotune [--quick]
"""
from __future__ import annotations

import json
import os
import re
import statistics
import sys
import time

import numpy as np

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED_DIR = os.path.join(HERE, "tuned")

# (bucket, knob name, candidate values, library default, workloads it acts on)
KNOBS = [
    ("LDL^T device levels, fronts of <= 48 rows: one wavefront (1) or one workgroup (0) per front",
     "LDLT_DEVICE_WAVE", [1, 0], 1, ("batch", "kkt")),
    ("LDL^T device levels: largest symbolic front laid out on the device (rows)",
     "LDLT_DEVICE_FRONTS", [95, 64, 48], 95, ("batch", "kkt")),
    ("LDL^T fronts from this many rows take the multi-workgroup path (below: one workgroup)",
     "LDLT_BIG_MIN", [256, 192, 384, 512], 256, ("kkt", "kkt_bf16x3")),
    ("LDL^T large fronts: workgroups of the MFMA update launches",
     "LDLT_UPDATE_WG", [1024, 512, 2048], 1024, ("kkt", "kkt_bf16x3")),
    ("LDL^T large fronts on bf16 planes: outer block (columns per flush)",
     "LDLT_OUTER_WIDE", [1024, 768, 512], 1024, ("kkt_bf16x3",)),
    ("SPD fronts, FP64: pivot columns per trailing Schur update",
     "PANEL_OUTER", [512, 256, 1024], 512, ("spd",)),
    ("SPD fronts, reduced precision: pivot columns per trailing Schur update",
     "PANEL_OUTER_REDUCED", [256, 512], 256, ("spd_bf16x3",)),
    ("SPD contribution blocks: pivot columns per update",
     "CB_SEG", [4096, 2048, 8192], 4096, ("spd", "spd_bf16x3")),
    ("SPD Schur tiles: 16-column chunks with global loads in flight",
     "SYRK_AHEAD", [2, 1], 2, ("spd",)),
    ("SPD assembly: largest LDS column chunk (rows)",
     "ASM_CHUNK", [2048, 1024, 4096], 2048, ("spd",)),
    ("SPD assembly: levels of fronts up to this many rows take a wavefront per column",
     "ASM_WAVE_ROWS", [2048, 1024, 3072], 2048, ("spd",)),
]
# the parameters of SPD plans (the others are set on LDL^T plans)
SPD_KNOBS = ("PANEL_OUTER", "PANEL_OUTER_REDUCED", "CB_SEG", "SYRK_AHEAD", "ASM_CHUNK", "ASM_WAVE_ROWS")


def device_key(name):
    """File-name key of a device name ("AMD Instinct MI355X (gfx950)" ->
    "AMD_Instinct_MI355X_gfx950")."""
    return re.sub(r"[^A-Za-z0-9]+", "_", name).strip("_") or "device"


def tuned_path(device_name):
    return os.path.join(TUNED_DIR, device_key(device_name) + ".json")


def load_tuned(device_name):
    """The persisted choices for this device, or None."""
    p = tuned_path(device_name)
    if not os.path.exists(p):
        return None
    with open(p) as fh:
        d = json.load(fh)
    if d.get("device") != device_name or not isinstance(d.get("knobs"), dict):
        return None
    return d


def apply_tuned(plan, tuned, posdef):
    """Set the tuned parameters on a plan (the SPD ones on SPD plans, the
    LDL^T ones on indefinite plans); returns what was set."""
    done = {}
    if not tuned:
        return done
    for name, value in tuned["knobs"].items():
        spd_knob = name in SPD_KNOBS
        if spd_knob != bool(posdef):
            continue
        plan.tune(getattr(_lib, "TUNE_" + name), int(value))
        done[name] = int(value)
    return done


def _kkt_grid_perm(gsz, m, A):
    """Geometric nested dissection of a grid KKT, each multiplier in the
    separator of the smallest box holding its columns (the ordering the
    benchmark times)."""
    from . import ordering as o
    n, cp, ri = A[0], A[1], A[2]
    nx = gsz * gsz
    cols = np.repeat(np.arange(n), np.diff(cp))
    sel = (ri >= nx) & (cols < nx)
    order = np.argsort(ri[sel], kind="stable")
    return o.kkt_grid_nd(gsz, m, cols[sel][order].reshape(m, 8), leaf=64)


def _sym_matvec(A, x):
    """y = A x for a symmetric matrix given by its lower triangle (CSC)."""
    n, cp, ri, v = A
    cols = np.repeat(np.arange(n), np.diff(cp))
    y = np.zeros(n)
    np.add.at(y, ri, v * x[cols])
    off = ri != cols
    np.add.at(y, cols[off], v[off] * x[ri[off]])
    return y


def workload(A, arrays, posdef, vals=None, seed=1):
    """A tuning workload: the plan's analysis arrays, and values and a
    right-hand side b = A x (x ~ U[-1, 1] drawn from default_rng(seed))
    resident on the device (b random when only the values are given, as for
    a batch). The number of refinement steps a reduced-precision factor
    needs can depend on b near the 1e-12 target: a benchmark tunes with its
    own b."""
    import torch
    dev = torch.device("cuda:0")
    vals = A[3] if vals is None else vals
    n = A[0]
    x = np.random.default_rng(seed).uniform(-1, 1, n)
    b = _sym_matvec(A, x) if A[1] is not None else x
    w = dict(A=A, arrays=arrays, posdef=posdef)
    w["vals"] = torch.from_numpy(np.ascontiguousarray(vals)).to(dev)
    w["b"] = torch.from_numpy(b).to(dev)
    w["x"] = torch.empty_like(w["b"])
    return w


def _workloads(quick):
    """Plans' arrays and resident values of the tuning workloads: an SPD
    27-point operator, a grid KKT (configuration 2's generator) in FP64 and
    split BF16, and a batch of OCP KKTs (small fronts on device levels)."""
    from . import batch as bt
    from . import generators as g, ordering as o
    from .solver import analyse
    out = {}
    m = 32 if quick else 100
    A = g.spd27_3d(m, seed=0)
    out["spd"] = workload(A, analyse(A, perm=o.grid_nd((m, m, m), leaf=64), nemin=32).arrays(), True)
    G = 96 if quick else 300
    mk = int(round(0.32 * G * G))
    K, _ = g.kkt_grid(G, mk, seed=3)
    out["kkt"] = workload(K, analyse(K, perm=_kkt_grid_perm(G, mk, K), nemin=32).arrays(), False)
    mats, _ = bt.ocp_batch(32 if quick else 1024, 10, 200, seed=0)
    U, _ = bt.union_arrays([analyse(Ai, nemin=16).arrays() for Ai in mats])
    vals = np.concatenate([Ai[3] for Ai in mats])
    out["batch"] = workload((int(U["n"]), None, None, vals), U, False)
    return out


def _time_one(plan, w, prec, reps):
    """Median seconds of factor + solve (device-resident data), with the
    refinement to a backward error of 1e-12 for a reduced precision (a
    parameter that makes the factor cheaper but less accurate can cost a
    refinement step more than it saves)."""
    import torch
    from .solver import PREC_FP64
    ts = []
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t = time.perf_counter()
        fac = plan.factor(values_ptr=w["vals"].data_ptr(), precision=prec, mixed_min_front=1024)
        w["x"].copy_(w["b"])
        torch.cuda.current_stream().synchronize()
        fac.solve_device(w["x"].data_ptr(), 1)
        ok = True
        if prec != PREC_FP64:
            ri = fac.refine_device(w["b"].data_ptr(), w["x"].data_ptr(), 1, max_iters=30, tol=1e-12, max_steps=40)
            ok = ri["reached"] and ri["eta"] <= 1e-12
        torch.cuda.synchronize()
        if not ok:
            return float("inf")  # the value misses the accuracy target: never chosen
        ts.append(time.perf_counter() - t)
        if fac.status != 0:
            raise RuntimeError(f"tuning factorization failed (status {fac.status})")
    return statistics.median(ts[1:])


def device_name():
    """The current device as the tuned files name it: its marketing name
    and its gfx architecture (the former alone can be generic)."""
    import torch
    pr = torch.cuda.get_device_properties(0)
    arch = getattr(pr, "gcnArchName", "").split(":")[0]
    return f"{pr.name} ({arch})" if arch else pr.name


def autotune(quick=False, reps=3, log=print, out_path=None, workloads=None):
    """Tune every parameter of KNOBS on this device and write the choices.
    workloads: {"spd": ..., "kkt": ..., "batch": ...} built by `workload`
    (bench.py --autotune passes the benchmark's own); by default the
    27-point operator on 100^3 (geometric nested dissection), the grid KKT
    on 300^2 and 1024 OCP KKTs (quick: small ones, for a test)."""
    from .solver import PREC_BF16X3, PREC_FP64, Plan
    name = device_name()
    wl = workloads if workloads is not None else _workloads(quick)
    # (workload key, precision) of each objective term
    terms = {"spd": ("spd", PREC_FP64), "spd_bf16x3": ("spd", PREC_BF16X3), "kkt": ("kkt", PREC_FP64),
             "kkt_bf16x3": ("kkt", PREC_BF16X3), "batch": ("batch", PREC_FP64)}
    plans = {k: Plan(w["arrays"], posdef=w["posdef"]) for k, w in wl.items()}
    best = {k[1]: k[3] for k in KNOBS}
    trials = []

    def objective(used):
        tot = 0.0
        for t in used:
            wk, prec = terms[t]
            if wk not in wl:
                continue
            p = plans[wk]
            for kname, v in best.items():
                spd_knob = kname in SPD_KNOBS
                if spd_knob == wl[wk]["posdef"]:
                    p.tune(getattr(_lib, "TUNE_" + kname), int(v))
            tot += _time_one(p, wl[wk], prec, reps)
        return tot

    for bucket, kname, cands, default, used in KNOBS:
        scores = {}
        for v in cands:
            keep = best[kname]
            best[kname] = v
            scores[v] = objective(used)
            best[kname] = keep
            trials.append({"knob": kname, "value": v, "seconds": round(scores[v], 6)})
        choice = min(scores, key=scores.get)
        # keep the default unless another value is better by more than 1%
        if scores[choice] > 0.99 * scores.get(default, float("inf")):
            choice = default
        best[kname] = choice
        log(f"autotune {kname}: {{{', '.join(f'{v}: {s * 1e3:.1f} ms' for v, s in scores.items())}}} -> {choice}")
    res = {"device": name, "knobs": best,
           "buckets": {k[1]: k[0] for k in KNOBS},
           "workloads": {k: {"n": int(w["A"][0]), "nsup": int(len(w["arrays"]["sn_start"]))} for k, w in wl.items()},
           "objective": "median factor + solve (+ refinement to 1e-12 for split BF16) seconds over the "
                        "workloads each parameter acts on",
           "trials": trials, "quick": bool(quick)}
    path = out_path or tuned_path(name)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as fh:
        json.dump(res, fh, indent=1)
    log(f"autotune: wrote {path}")
    return res


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(HERE))
    autotune(quick="--quick" in sys.argv)
