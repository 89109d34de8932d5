"""The interior-point refactor sequence of configuration 5 at reduced size:
one analysis and one device plan, 40 numeric refactorizations of a grid KKT
whose barrier diagonal moves as log10 Sigma_k = log10 Sigma_0 + k s
(kkt_sigma_sequence), as smf's Solver::factor(ak, values, ...) is called by
an IPM (IpmSession::factor, solver/src/ipm_session.cpp:147-243).

FP64: every refactor against the oracle (pivot types, row order, delays,
2x2 count, flops, L and D within 1e-12) and Sylvester's inertia (n_x, m, 0).
Reduced-precision MFMA updates: the same inertia on every refactor and FP64
refinement to a scaled backward error <= 1e-12."""
import numpy as np
import pytest

from tests.conftest import gpu_available
from tests.helpers import matvec

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from opt_solver_amd import generators as g  # noqa: E402
from opt_solver_amd._lib import TUNE_LDLT_BIG_MIN  # noqa: E402
from opt_solver_amd.solver import PREC_BF16X3, PREC_FP32, PREC_FP64, Plan, analyse  # noqa: E402

GRID, M, NREF = 40, 320, 40


def _sequence():
    A0, sigma0 = g.kkt_grid(GRID, M, seed=5)
    nx = GRID * GRID
    vals = [g.kkt_grid_refactor_values(A0, g.kkt_sigma_sequence(sigma0, k, seed=6), nx) for k in range(NREF)]
    from oracle import refshim
    if refshim.available():
        arrays = refshim.RefAnalysis(A0, ordering=refshim.ORDER_METIS, posdef=False).plan_arrays()
    else:
        arrays = analyse(A0, nemin=16).arrays()
    return A0, vals, arrays


def _block_err(ex, eo):
    err = 0.0
    for s in range(ex["steps"].shape[0]):
        _, fo, p, f = eo["steps"][s]
        a = ex["factor_values"][fo:fo + f * p]
        b = eo["factor_values"][fo:fo + f * p]
        if a.size:
            err = max(err, np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    return err


def test_ipm_sequence_fp64_against_oracle(oracle_lib):
    A0, vals, arrays = _sequence()
    plan = Plan(arrays, posdef=False)
    plan.tune(TUNE_LDLT_BIG_MIN, 96)  # fronts of 96+ rows take the multi-workgroup path
    inertia = (GRID * GRID, M, 0)
    worst = 0.0
    delays = []
    for k, v in enumerate(vals):
        print("case ipm fp64 refactor", k, flush=True)
        fac = plan.factor(v, u=0.01)
        orc = oracle_lib.Factor(arrays, v, posdef=False)
        assert fac.status == orc.status == 0
        assert fac.inertia == orc.inertia == inertia, k
        assert fac.info.delayed_pivots == orc.delayed and fac.info.two_by_two_pivots == orc.two_by_two, k
        assert fac.flops == orc.flops
        ex, eo = fac.export(), orc.export()
        assert np.array_equal(ex["steps"], eo["steps"])
        assert np.array_equal(ex["row_indices"], eo["row_indices"])
        assert np.array_equal(ex["pivot_types"], eo["pivot_types"])
        err = _block_err(ex, eo)
        assert err <= 1e-12, (k, err)
        worst = max(worst, err)
        delays.append(fac.info.delayed_pivots)
        if k % 13 == 0:
            A = (A0[0], A0[1], A0[2], v)
            x = np.random.default_rng(k).uniform(-1, 1, A0[0])
            b = matvec(A, x)
            eta, _ = oracle_lib.backward_error(A, b, fac.solve(b))
            assert eta <= 1e-14, (k, eta)
    print(f"  40 refactors: worst L/D rel err {worst:.1e}, delays {min(delays)}..{max(delays)}", flush=True)


@pytest.mark.parametrize("prec", [PREC_FP32, PREC_BF16X3])
def test_ipm_sequence_reduced(oracle_lib, prec):
    A0, vals, arrays = _sequence()
    plan = Plan(arrays, posdef=False)
    plan.tune(TUNE_LDLT_BIG_MIN, 96)
    inertia = (GRID * GRID, M, 0)
    steps = []
    for k, v in enumerate(vals):
        print("case ipm reduced prec", prec, "refactor", k, flush=True)
        fac = plan.factor(v, u=0.01, precision=prec, mixed_min_front=0)
        assert fac.status == 0 and fac.inertia == inertia, k
        assert fac.info.fronts_reduced > 0
        A = (A0[0], A0[1], A0[2], v)
        x = np.random.default_rng(100 + k).uniform(-1, 1, A0[0])
        b = matvec(A, x)
        xr, st = fac.refine(b, fac.solve(b), max_iters=10, tol=1e-13, max_steps=30)
        eta, om = oracle_lib.backward_error(A, b, xr)
        assert eta <= 1e-12, (k, eta, st)
        steps.append(st["steps"])
    print(f"  40 refactors: refinement steps {min(steps)}..{max(steps)}", flush=True)
