"""The learned per-supernode precision selector on the GPU: its inference
and training kernels against the numpy restatement, its training pairs
(features and the measured Schur-update errors of fronts factored both
ways) against closed forms, and factorizations with precision AUTO refined
to the mixed-precision gate."""
import numpy as np
import pytest

from tests import selector_np as S
from tests.conftest import gpu_available
from tests.helpers import matvec

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from opt_solver_amd import generators as g  # noqa: E402
from opt_solver_amd import ordering as o  # noqa: E402
from opt_solver_amd.solver import (PREC_AUTO, PREC_BF16, PREC_BF16X3, PREC_FP32, PREC_FP64, Plan,  # noqa: E402
                                   analyse, selector_init, selector_predict, selector_train)


def _rand_weights(seed):
    rng = np.random.default_rng(seed)
    w = selector_init(seed).astype(np.float64)
    w[S.NP:S.NP + S.NF] = rng.normal(0, 1, S.NF)
    w[S.NP + S.NF:] = rng.uniform(0.5, 2.0, S.NF)
    w[:S.NP] += rng.normal(0, 0.05, S.NP)
    return w.astype(np.float32)


def test_inference_kernel_matches_numpy():
    print("case selector predict", flush=True)
    w = _rand_weights(1)
    X = np.random.default_rng(2).normal(0, 2, (1000, S.NF)).astype(np.float32)
    Yg = selector_predict(w, X)
    Yn, _ = S.forward(w, X)
    assert np.abs(Yg - Yn).max() <= 1e-5 * max(1.0, np.abs(Yn).max())


def test_training_kernels_match_numpy():
    """Adam steps on the device against numpy's (float32 sums on the
    device, so agreement to ~1e-5 relative after a few steps)."""
    print("case selector train vs numpy", flush=True)
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, (700, S.NF)).astype(np.float32)
    Y = rng.normal(0, 1, (700, S.NO)).astype(np.float32)
    M = (rng.uniform(0, 1, (700, S.NO)) > 0.2).astype(np.float32)
    w0 = _rand_weights(4)
    wg, lg = selector_train(X, Y, M, epochs=5, lr=1e-3, weights=w0)
    # the device call sets the input scaling from X first; numpy starts there too
    w0n = w0.astype(np.float64)
    w0n[S.NP:] = wg[S.NP:]
    wn, ln = S.adam(w0n, X, Y, M, steps=5, lr=1e-3)
    assert np.allclose(lg, ln, rtol=1e-4, atol=0)
    assert np.abs(wg - wn).max() <= 2e-5


def _spd27(m, seed=5, shift=None):
    A = g.spd27_3d(m, seed=seed)
    if shift is not None:  # diagonal = |row sum| + shift instead of + 1e-3
        n, cp, ri, v = A
        v = v.copy()
        v[cp[:-1]] += shift - 1e-3
        A = (n, cp, ri, v)
    an = analyse(A, perm=o.grid_nd((m, m, m), leaf=16))
    return A, an


def test_pairs_labels_against_numpy():
    """Pairs of an FP64 factorization: one per front of at least
    mixed_min_front rows with a contribution block; labels ordered FP32 <
    split BF16 < BF16 by their unit roundoffs; the FP32 and BF16 labels of
    each front against numpy's rounded-operand product and the contribution
    block S = A22 - A21 A11^-1 A12 of the front's subtree, in closed form."""
    print("case selector pairs", flush=True)
    m = 10
    A, an = _spd27(m)
    plan = Plan(an, posdef=True)
    fac = plan.factor(A[3], collect=True, mixed_min_front=64)
    X, Y, fr = fac.selector_pairs()
    arr = an.arrays()
    f_of = np.diff(arr["front_row_ptr"])
    p_of = arr["sn_end"] - arr["sn_start"]
    want = [s for s in range(an.nsup) if f_of[s] >= 64 and f_of[s] > p_of[s]]
    assert list(fr) == want and len(fr) > 3
    assert np.all(np.isfinite(X)) and np.all(np.isfinite(Y))
    assert np.all(Y[:, 0] < Y[:, 1]) and np.all(Y[:, 1] < Y[:, 2])
    assert np.all(X[:, 0] == np.log2(f_of[fr]).astype(np.float32))
    # numpy: Z = L21 L21^T from the exported factor, rounded operands
    ex = fac.export()
    import torch
    dense = g.csc_to_dense(*A)
    perm = arr["perm"]
    P = dense[np.ix_(perm, perm)]
    rows = ex["row_indices"]
    for k, s in enumerate(fr[:6]):
        ro, fo, p, f = ex["steps"][s]
        Ls = ex["factor_values"][fo:fo + f * p].reshape(p, f).T
        L21 = Ls[p:, :]
        Z = L21 @ L21.T
        # subtree of s: supernodes in [first descendant, s]
        sub = [s]
        par = arr["sn_parent"]
        for t in range(s - 1, -1, -1):
            u = t
            while u != -1 and u < s:
                u = par[u]
            if u == s:
                sub.append(t)
        E = np.concatenate([np.arange(arr["sn_start"][t], arr["sn_end"][t]) for t in sorted(sub)])
        R = rows[ro + p:ro + f]
        Sblk = -P[np.ix_(R, E)] @ np.linalg.solve(P[np.ix_(E, E)], P[np.ix_(E, R)])
        smax = np.abs(Sblk).max()
        for col, dt in ((0, torch.float32), (2, torch.bfloat16)):
            Lr = torch.from_numpy(L21).to(torch.float32).to(dt).to(torch.float64).numpy()
            ez = np.abs(Lr @ Lr.T - Z).max() / smax
            assert abs(np.log10(ez) - Y[k, col]) <= 0.7, (s, col, np.log10(ez), Y[k, col])


def _train_on(plans_values, mmf=64):
    Xs, Ys = [], []
    for plan, vals in plans_values:
        fac = plan.factor(vals, collect=True, mixed_min_front=mmf)
        X, Y, _ = fac.selector_pairs()
        Xs.append(X)
        Ys.append(Y)
    X = np.concatenate(Xs)
    Y = np.concatenate(Ys)
    w, loss = selector_train(X, Y, epochs=1500, lr=3e-3)
    return w, loss, X, Y


def test_selector_trains_and_auto_factor_refines(oracle_lib):
    """Train on pairs from a few operators, then an AUTO factorization of
    another: the precisions the device picked equal the numpy policy on the
    device features, reduced fronts exist, the inertia is (n, 0, 0) and
    refinement reaches a scaled backward error <= 1e-12."""
    print("case selector train + auto", flush=True)
    train = []
    for m, shift, seed in ((10, None, 1), (12, 1.0, 2), (12, 1e-6, 3), (14, None, 4)):
        A, an = _spd27(m, seed=seed, shift=shift)
        train.append((Plan(an, posdef=True), A[3]))
    w, loss, X, Y = _train_on(train)
    print(f"  pairs {len(X)} loss {loss[0]:.3f} -> {loss[-1]:.4f}", flush=True)
    assert loss[-1] < 0.2 * loss[0]
    A, an = _spd27(14, seed=9)
    plan = Plan(an, posdef=True)
    plan.selector_set(w)
    tol = 1e-4
    ref = plan.factor(A[3], collect=True, mixed_min_front=64)
    Xt, _, frt = ref.selector_pairs()
    fac = plan.factor(A[3], precision=PREC_AUTO, mixed_min_front=64, selector_tol=tol)
    pr = fac.precisions()
    want = S.choose(S.forward(w, Xt)[0], tol)
    agree = np.mean(pr[frt] == want)
    assert agree >= 0.95, agree  # float32 on both sides: ties at the threshold may differ
    assert fac.info.fronts_reduced > 0 and fac.inertia == (A[0], 0, 0)
    assert set(np.unique(pr)) <= {PREC_FP64, PREC_FP32, PREC_BF16, PREC_BF16X3}
    x = np.random.default_rng(3).uniform(-1, 1, A[0])
    b = matvec(A, x)
    xr, st = fac.refine(b, fac.solve(b), max_iters=20, tol=1e-14)
    eta, _ = oracle_lib.backward_error(A, b, xr)
    print(f"  AUTO: {dict(zip(*np.unique(pr, return_counts=True)))} refine steps {st['steps']} eta {eta:.1e}",
          flush=True)
    assert eta <= 1e-12


def _kkt_seq(gsz=40, m=320, ks=(0, 10, 20, 30, 35)):
    from opt_solver_amd._lib import TUNE_LDLT_BIG_MIN
    A0, s0 = g.kkt_grid(gsz, m, seed=5)
    nx = gsz * gsz
    vals = {k: g.kkt_grid_refactor_values(A0, g.kkt_sigma_sequence(s0, k, seed=6), nx) for k in ks}
    plan = Plan(analyse(A0, nemin=16).arrays(), posdef=False)
    plan.tune(TUNE_LDLT_BIG_MIN, 96)  # fronts of 96+ rows take the multi-workgroup (MFMA) path
    return A0, vals, plan, (nx, m, 0)


def test_ldlt_pairs_leave_the_factor_unchanged_and_auto_refines(oracle_lib):
    """KKT refactors: a collecting FP64 factorization (each large front's
    contribution-block update also run in FP32, split BF16 and BF16 from the
    same square, then put back) gives the plain FP64 factor bit for bit;
    its pairs have finite features and labels mostly ordered by unit
    roundoff; a selector trained on them runs an AUTO factorization of a
    later refactor with Sylvester's inertia, refined to <= 1e-12."""
    print("case selector ldlt", flush=True)
    A0, vals, plan, inertia = _kkt_seq()
    Xs, Ys = [], []
    for k in (0, 10, 20, 30):
        ref = plan.factor(vals[k], u=0.01).export()
        fac = plan.factor(vals[k], u=0.01, collect=True, mixed_min_front=96)
        ex = fac.export()
        for key in ("steps", "row_indices", "pivot_types", "factor_values"):
            assert np.array_equal(ex[key], ref[key]), (k, key)
        X, Y, fr = fac.selector_pairs()
        assert len(fr) > 3 and np.all(np.isfinite(X)) and np.all(np.isfinite(Y))
        assert np.mean((Y[:, 0] < Y[:, 1]) & (Y[:, 1] < Y[:, 2])) >= 0.9
        Xs.append(X)
        Ys.append(Y)
    w, loss = selector_train(np.concatenate(Xs), np.concatenate(Ys), epochs=1500, lr=3e-3)
    assert loss[-1] < 0.3 * loss[0]
    plan.selector_set(w)
    v = vals[35]
    A = (A0[0], A0[1], A0[2], v)
    fac = plan.factor(v, u=0.01, precision=PREC_AUTO, mixed_min_front=96, selector_tol=1e-4)
    pr = fac.precisions()
    print(f"  AUTO ldlt: {dict(zip(*np.unique(pr, return_counts=True)))}", flush=True)
    assert fac.status == 0 and fac.inertia == inertia
    x = np.random.default_rng(7).uniform(-1, 1, A0[0])
    b = matvec(A, x)
    xr, st = fac.refine(b, fac.solve(b), max_iters=20, tol=1e-14)
    eta, _ = oracle_lib.backward_error(A, b, xr)
    assert eta <= 1e-12, (eta, st)


def _all_bf16_weights():
    """Weights whose every prediction is 1e-10: the selector then gives
    every candidate front BF16."""
    w = selector_init(0)
    o = S.NH * S.NF + S.NH + S.NH * S.NH + S.NH  # W3 starts here
    w[o:o + S.NO * S.NH] = 0.0
    w[o + S.NO * S.NH:o + S.NO * S.NH + S.NO] = -10.0
    return w


@pytest.mark.parametrize("kind", ["spd", "kkt"])
def test_fallback_to_fp64_when_refinement_stalls(oracle_lib, kind):
    """smfgpu_factor_solve with the selector forced to BF16 on every
    candidate front and no refinement step allowed: the BF16 factor's solve
    misses eta <= 1e-14, refinement stops there, and the plain FP64 factor
    takes over and meets it; with a budget BF16 can meet, no fallback."""
    print("case selector fallback", kind, flush=True)
    if kind == "spd":
        A, an = _spd27(12, seed=11)
        plan = Plan(an, posdef=True)
        v, mmf, inertia = A[3], 64, (A[0], 0, 0)
    else:
        A0, vals, plan, inertia = _kkt_seq(ks=(39,))
        v, mmf = vals[39], 96
        A = (A0[0], A0[1], A0[2], v)
    plan.selector_set(_all_bf16_weights())
    x = np.random.default_rng(8).uniform(-1, 1, A[0])
    b = matvec(A, x)
    fac, xs, ri = plan.factor_solve(v, b, precision=PREC_AUTO, mixed_min_front=mmf, max_iters=0, max_steps=0,
                                    tol=1e-14)
    eta, _ = oracle_lib.backward_error(A, b, xs)
    print(f"  tight budget: fallback {ri['fallback']} eta {eta:.1e} reduced {fac.info.fronts_reduced}", flush=True)
    assert ri["fallback"] and ri["reached"] and fac.info.fronts_reduced == 0 and fac.inertia == inertia
    assert eta <= 1e-14
    fac, xs, ri = plan.factor_solve(v, b, precision=PREC_AUTO, mixed_min_front=mmf, max_iters=30, max_steps=40,
                                    tol=1e-12)
    eta, _ = oracle_lib.backward_error(A, b, xs)
    pr = fac.precisions()
    print(f"  budget: fallback {ri['fallback']} steps {ri['steps']} eta {eta:.1e}", flush=True)
    assert not ri["fallback"] and np.all(pr[pr != PREC_FP64] == PREC_BF16) and np.any(pr == PREC_BF16)
    assert eta <= 1e-12
