"""Train the learned per-supernode precision selector and write its weights
(opt_solver_amd/selector_weights.json).

Training pairs come from the solver itself: each training operator is
factored in FP64 with feature collection on, and every front of at least
`--min-front` rows with a contribution block gives one pair: the features
the device computed from the assembled front before factoring it, and the
measured log10 relative error of its Schur update run on the MFMA tiles in
FP32, split BF16 and BF16 against FP64 (smfgpu_selector_pairs). The MLP is
trained with the library's HIP training kernels (smfgpu_selector_train).

Operators: the SPD 27-point operator (diagonal = |row sum| + shift, shifts
from 1 down to 1e-9), the 7-point Laplacian with shifts, and the same with
a symmetric diagonal scaling (random, or the equilibrating one flagged as
scaled), on grids of 12^3 to 32^3 with geometric nested dissection; and
interior-point KKT refactor sequences (LDL^T, u = 0.01) on 2D grids, whose
large fronts (the multi-workgroup path) give the pairs.

Needs a GPU: python -m opt_solver_amd.train_selector [--out FILE]."""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from opt_solver_amd import generators as g  # noqa: E402
from opt_solver_amd import ordering as o  # noqa: E402
from opt_solver_amd.solver import SELECTOR_WEIGHTS_FILE, Plan, analyse, selector_train  # noqa: E402


def shifted(A, shift, base):
    """A with its diagonal moved by shift - base (the generators' diagonal
    is |row sum| + base)."""
    n, cp, ri, v = A
    v = v.copy()
    v[cp[:-1]] += shift - base
    return (n, cp, ri, v)


def scaled(A, s):
    """S A S for the diagonal S = diag(s)."""
    n, cp, ri, v = A
    cols = np.repeat(np.arange(n), np.diff(cp))
    return (n, cp, ri, v * s[ri] * s[cols])


def operators(sizes, seed0=0):
    rng = np.random.default_rng(seed0)
    for m in sizes:
        perm = o.grid_nd((m, m, m), leaf=16 if m < 24 else 32)
        for shift in (1.0, 1e-3, 1e-6, 1e-9):
            A = shifted(g.spd27_3d(m, seed=int(rng.integers(1 << 30))), shift, 1e-3)
            yield f"spd27 {m}^3 shift {shift:g}", A, perm, False
            d = A[3][A[1][:-1]]
            yield f"spd27 {m}^3 shift {shift:g} equilibrated", scaled(A, 1.0 / np.sqrt(d)), perm, True
            yield (f"spd27 {m}^3 shift {shift:g} random scaling",
                   scaled(A, 10.0 ** rng.uniform(-3, 3, A[0])), perm, False)
        for shift in (1.0, 1e-4, 1e-8):
            L = g.laplace7_3d(m)
            yield f"laplace7 {m}^3 shift {shift:g}", shifted(L, shift, 0.0), perm, False


def kkt_operators(grids, refactors=(0, 13, 26, 39)):
    """Interior-point KKT sequences: [[H + Sigma_k, A^T], [A, -1e-8 I]] on
    g x g grids with m = 0.32 g^2 constraints (the 500^2 / 80k ratio),
    Sigma_k moving as kkt_sigma_sequence; the library's ordering."""
    for gsz in grids:
        m = int(0.32 * gsz * gsz)
        A0, s0 = g.kkt_grid(gsz, m, seed=gsz)
        nx = gsz * gsz
        cols = np.repeat(np.arange(A0[0]), np.diff(A0[1]))
        sel = (A0[2] >= nx) & (cols < nx)
        order = np.argsort(A0[2][sel], kind="stable")
        perm = o.kkt_grid_nd(gsz, m, cols[sel][order].reshape(m, 8), leaf=64)
        for k in refactors:
            v = g.kkt_grid_refactor_values(A0, g.kkt_sigma_sequence(s0, k, seed=gsz + 1), nx)
            yield f"kkt {gsz}^2 m={m} refactor {k}", (A0[0], A0[1], A0[2], v), perm, False


def _pairs(name, A, perm, sc, posdef, min_front, log, plans):
    key = (id(perm), posdef, A[2].size, hash(A[2].tobytes()))
    if key not in plans:  # one analysis per pattern, as smf reuses an AnalysisKeep
        plans.clear()
        plans[key] = Plan(analyse(A, perm=perm, nemin=32), posdef=posdef)
    plan = plans[key]
    fac = plan.factor(A[3], collect=True, mixed_min_front=min_front, scaled=sc, u=0.01)
    if fac.status != 0:
        log(f"  {name}: status {fac.status}, skipped")
        return None
    X, Y, _ = fac.selector_pairs()
    log(f"  {name}: {len(X)} pairs, labels fp32 {Y[:, 0].min():.1f}..{Y[:, 0].max():.1f} "
        f"bf16x3 {Y[:, 1].min():.1f}..{Y[:, 1].max():.1f} bf16 {Y[:, 2].min():.1f}..{Y[:, 2].max():.1f}"
        if len(X) else f"  {name}: no pairs")
    return X, Y


def collect(sizes, min_front, grids=(), kkt_min_front=256, log=print):
    Xs, Ys, names = [], [], []
    plans = {}
    for name, A, perm, sc in operators(sizes):
        r = _pairs(name, A, perm, sc, True, min_front, log, plans)
        if r is not None and len(r[0]):
            Xs.append(r[0])
            Ys.append(r[1])
            names += [name] * len(r[0])
    for name, A, perm, sc in kkt_operators(grids):
        r = _pairs(name, A, perm, sc, False, kkt_min_front, log, plans)
        if r is not None and len(r[0]):
            Xs.append(r[0])
            Ys.append(r[1])
            names += [name] * len(r[0])
    return np.concatenate(Xs), np.concatenate(Ys), names


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=SELECTOR_WEIGHTS_FILE)
    ap.add_argument("--sizes", default="12,16,20,24,28,32")
    ap.add_argument("--min-front", type=int, default=64)
    ap.add_argument("--kkt-grids", default="60,90,120")
    ap.add_argument("--kkt-min-front", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    t0 = time.time()
    sizes = [int(x) for x in args.sizes.split(",")]
    grids = [int(x) for x in args.kkt_grids.split(",") if x]
    X, Y, names = collect(sizes, args.min_front, grids, args.kkt_min_front)
    print(f"pairs {len(X)} in {time.time() - t0:.1f}s", flush=True)
    # held-out check: every fifth pair
    idx = np.arange(len(X))
    test = idx % 5 == 4
    w, loss = selector_train(X[~test], Y[~test], epochs=args.epochs, lr=args.lr, seed=args.seed)
    from opt_solver_amd.solver import selector_predict
    pt = selector_predict(w, X[test])
    rmse = np.sqrt(((pt - Y[test]) ** 2).mean(0))
    print(f"train loss {loss[0]:.3f} -> {loss[-1]:.4f}; held-out RMSE (log10) fp32 {rmse[0]:.3f} "
          f"bf16x3 {rmse[1]:.3f} bf16 {rmse[2]:.3f}", flush=True)
    # final weights on all pairs, started from the held-out run's
    w, loss = selector_train(X, Y, epochs=args.epochs // 2, lr=args.lr / 3, weights=w)
    meta = {"pairs": int(len(X)), "operators": sorted(set(names)), "min_front": args.min_front,
            "kkt_min_front": args.kkt_min_front,
            "epochs": args.epochs, "lr": args.lr, "final_loss": float(loss[-1]),
            "heldout_rmse_log10": [float(r) for r in rmse],
            "features": ["log2 f", "log2 p", "q/f", "log10 max/min column norm", "log10 min |a_jj|/||a_j||",
                         "mean log10 |a_jj|/||a_j||", "log10 max/min |a_jj|", "log10 max column norm / max|A|",
                         "delayed columns / p", "scaled"],
            "outputs": ["log10 rel err FP32", "log10 rel err split BF16", "log10 rel err BF16"]}
    with open(args.out, "w") as fh:
        json.dump({"weights": [float(x) for x in w], "meta": meta}, fh)
    print(f"wrote {args.out} ({time.time() - t0:.1f}s)", flush=True)


if __name__ == "__main__":
    main()
