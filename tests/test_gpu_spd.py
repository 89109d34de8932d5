"""GPU parity of the SPD (Cholesky) numeric phase through the C ABI."""
import os

import numpy as np
import pytest

from tests.conftest import gpu_available
from tests.helpers import matvec
from tests.test_golden import load_mf

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from opt_solver_amd import generators as g  # noqa: E402
from opt_solver_amd import ordering as o  # noqa: E402
from opt_solver_amd._lib import TUNE_CB_SEG, TUNE_PANEL_OUTER, TUNE_SOLVE_FUSE_ROWS  # noqa: E402
from opt_solver_amd.solver import PREC_BF16, PREC_BF16X3, PREC_FP32, Plan, SmfGpuError, analyse  # noqa: E402


def block_rel_err(e1, e2):
    err = 0.0
    for s in range(e1["steps"].shape[0]):
        ro, fo, p, f = e1["steps"][s]
        a = e1["factor_values"][fo:fo + f * p]
        b = e2["factor_values"][fo:fo + f * p]
        err = max(err, np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    return err


def test_golden_laplace_reference_analysis():
    """The reference's own analysis of the 7-point Laplacian 8^3: identical
    steps and row indices, L blocks within 1e-12 of the reference's."""
    z, a, A = load_mf("laplace7_8")
    print("case golden laplace7_8", flush=True)
    plan = Plan(a, posdef=True)
    fac = plan.factor(A[3])
    assert fac.status == 0
    ex = fac.export()
    assert np.array_equal(ex["steps"], z["steps"])
    assert np.array_equal(ex["row_indices"], z["row_indices"])
    gold = dict(steps=z["steps"], factor_values=z["factor_values"])
    assert block_rel_err(ex, gold) < 1e-12
    assert fac.flops == z["stats"][6]
    x = fac.solve(z["b"])
    assert np.abs(x - z["x_ref"]).max() < 1e-12


def test_c1_against_oracle_and_reference(oracle_lib):
    """Config 1 (7-point 20^3) on the reference's METIS analysis when the
    reference library is built, else on the library's own nested
    dissection; factor vs the oracle, solve and backward error."""
    print("case c1 laplace7_20", flush=True)
    A = g.laplace7_3d(20)
    from oracle import refshim
    if refshim.available():
        arrays = refshim.RefAnalysis(A, ordering=refshim.ORDER_METIS, posdef=True).plan_arrays()
    else:
        arrays = analyse(A, perm=o.grid_nd((20, 20, 20), leaf=64)).arrays()
    plan = Plan(arrays, posdef=True)
    fac = plan.factor(A[3])
    orc = oracle_lib.Factor(arrays, A[3], posdef=True)
    ex, eo = fac.export(), orc.export()
    assert np.array_equal(ex["steps"], eo["steps"]) and np.array_equal(ex["row_indices"], eo["row_indices"])
    assert block_rel_err(ex, eo) < 1e-12
    assert fac.inertia == (A[0], 0, 0) and fac.flops == orc.flops
    x = np.random.default_rng(0).uniform(-1, 1, A[0])
    b = matvec(A, x)
    xs = fac.solve(b)
    eta, om = oracle_lib.backward_error(A, b, xs)
    assert eta <= 1e-14
    assert np.abs(xs - x).max() < 1e-12


def test_spd27_own_analysis_multi_rhs(oracle_lib):
    m = 14
    print("case spd27_14 16 rhs", flush=True)
    A = g.spd27_3d(m, seed=3)
    an = analyse(A, perm=o.grid_nd((m, m, m), leaf=32), nemin=16)
    arrays = an.arrays()
    plan = Plan(an, posdef=True)
    fac = plan.factor(A[3])
    orc = oracle_lib.Factor(arrays, A[3], posdef=True)
    assert block_rel_err(fac.export(), orc.export()) < 1e-12
    X = np.random.default_rng(1).uniform(-1, 1, (A[0], 16))
    B = matvec(A, X)
    XS = fac.solve(B)
    eta, om = oracle_lib.backward_error(A, B, XS)
    assert eta <= 1e-14
    # the same as a loop over single right-hand sides
    for r in (0, 7, 15):
        assert np.abs(fac.solve(B[:, r]) - XS[:, r]).max() < 1e-13


def test_large_front_stepped_solve(oracle_lib):
    """spd27 40^3 has fronts of several pivot blocks (2000+ rows, the rows
    below the pivots past the fused-update limit): 1, 3, 4, 5 and 16
    right-hand sides, and the fused-update limit at 0 / default / all,
    against the known solution and the reference backward-error formula."""
    m = 40
    print("case spd27_40 stepped solve", flush=True)
    A = g.spd27_3d(m, seed=4)
    an = analyse(A, perm=o.grid_nd((m, m, m), leaf=64), nemin=32)
    assert an.max_front >= 2000
    plan = Plan(an, posdef=True)
    fac = plan.factor(A[3])
    X = np.random.default_rng(4).uniform(-1, 1, (A[0], 16))
    B = matvec(A, X)
    for fuse in (512, 0, 1 << 20):
        plan.tune(TUNE_SOLVE_FUSE_ROWS, fuse)
        for nr in (1, 3, 4, 5, 16):
            print(f"  fuse {fuse} nrhs {nr}", flush=True)
            XS = fac.solve(B[:, :nr] if nr > 1 else B[:, 0])
            XS = XS.reshape(A[0], nr)
            eta, om = oracle_lib.backward_error(A, B[:, :nr], XS)
            assert eta <= 1e-14, (fuse, nr, eta)
            assert np.abs(XS - X[:, :nr]).max() < 1e-10


@pytest.mark.parametrize("prec", [PREC_FP32, PREC_BF16X3, PREC_BF16])
def test_mixed_precision_refinement(oracle_lib, prec):
    """Reduced-precision Schur updates on every front (FP32, split BF16, or
    BF16 MFMA, FP32 accumulation), FP64 refinement against A: scaled
    backward error <= 1e-12 (the mixed-precision gate), inertia unchanged,
    and the factor close to the FP64 one by the precision's unit roundoff."""
    m = 16
    print("case mixed spd27_16 prec", prec, flush=True)
    A = g.spd27_3d(m, seed=5)
    an = analyse(A, perm=o.grid_nd((m, m, m), leaf=32))
    plan = Plan(an, posdef=True)
    ref = plan.factor(A[3]).export()
    fac = plan.factor(A[3], precision=prec, mixed_min_front=0)
    assert fac.status == 0 and fac.inertia == (A[0], 0, 0)
    assert fac.info.fronts_reduced == an.nsup
    err = block_rel_err(fac.export(), ref)
    # FP32: 2^-24 per operand rounding; split BF16: about 2^-16; BF16: 2^-9
    lo, hi = {PREC_FP32: (1e-9, 1e-5), PREC_BF16X3: (1e-7, 1e-3), PREC_BF16: (1e-4, 0.2)}[prec]
    print(f"  factor rel diff vs FP64 {err:.2e}", flush=True)
    assert lo < err < hi
    x = np.random.default_rng(2).uniform(-1, 1, A[0])
    b = matvec(A, x)
    x0 = fac.solve(b)
    eta0, _ = oracle_lib.backward_error(A, b, x0)
    xr, st = fac.refine(b, x0, max_iters=30, tol=1e-14)
    eta, om = oracle_lib.backward_error(A, b, xr)
    print(f"  eta0 {eta0:.2e} refined {eta:.2e} in {st['steps']} steps", flush=True)
    assert eta0 > 1e-12  # the reduced Schur updates are visible before refinement
    assert eta <= 1e-12 and st["steps"] >= 1
    # the device's double-double residual and reductions against the
    # oracle's long-double formula on the same x
    print(f"  device eta {st['eta']:.6e} oracle {eta:.6e}", flush=True)
    assert abs(st["eta"] - eta) <= 1e-3 * eta + 1e-30


def test_single_launch_sweeps_match_stepped():
    """The single-launch sweeps (one workgroup per item, ordered by tickets
    and per-step counters) against the per-step launches on spd27 40^3:
    1, 4 and 5 right-hand sides, every solve job, within 1e-14 (the
    forward updates of shared rows are atomic adds in either order)."""
    from opt_solver_amd._lib import TUNE_SOLVE_PERSISTENT
    m = 40
    print("case spd27_40 single-launch sweeps", flush=True)
    A = g.spd27_3d(m, seed=4)
    an = analyse(A, perm=o.grid_nd((m, m, m), leaf=64), nemin=32)
    plan = Plan(an, posdef=True)
    fac = plan.factor(A[3])
    B = np.random.default_rng(6).uniform(-1, 1, (A[0], 5))
    for nr in (1, 4, 5):
        for job in (0, 1, 2, 3, 4):
            b = B[:, :nr] if nr > 1 else B[:, 0]
            plan.tune(TUNE_SOLVE_PERSISTENT, 0)
            ref = fac.solve(b, job=job)
            plan.tune(TUNE_SOLVE_PERSISTENT, 1)
            got = fac.solve(b, job=job)
            err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300)
            assert err <= 1e-14, (nr, job, err)
    plan.tune(TUNE_SOLVE_PERSISTENT, 0)


def test_not_positive_definite_detected():
    print("case not posdef", flush=True)
    A = g.laplace7_3d(6, shift=7.0)  # diagonal -1: indefinite
    an = analyse(A)
    plan = Plan(an, posdef=True)
    fac = plan.factor(A[3])
    assert fac.status == 1
    with pytest.raises(SmfGpuError):
        fac.solve(np.ones(A[0]))


def test_non_finite_value_rejected():
    print("case non-finite", flush=True)
    A = g.laplace7_3d(4)
    vals = A[3].copy()
    vals[5] = np.nan
    plan = Plan(analyse(A), posdef=True)
    with pytest.raises(SmfGpuError):
        plan.factor(vals)


def test_refactor_same_plan():
    """Two factorizations on one plan (the refactor path): the second
    supersedes the first; results are those of its own values."""
    print("case refactor", flush=True)
    A = g.spd27_3d(10, seed=1)
    plan = Plan(analyse(A), posdef=True)
    f1 = plan.factor(A[3])
    x = np.random.default_rng(3).uniform(-1, 1, A[0])
    A2 = (A[0], A[1], A[2], A[3] * 2.0)
    f2 = plan.factor(A2[3])
    assert np.abs(f2.solve(matvec(A2, x)) - x).max() < 1e-12
    with pytest.raises(SmfGpuError):
        f1.solve(x)


def test_outer_panel_width_invariance():
    """The trailing-update width (64, 256, the default 512, 1024) and the
    contribution-block segment (64, 100, unbounded, the default 4096)
    change only the order of the Schur sums: factors agree to 1e-13
    relative."""
    print("case outer panel widths", flush=True)
    m = 24
    A = g.spd27_3d(m, seed=8)
    plan = Plan(analyse(A, perm=o.grid_nd((m, m, m), leaf=32), nemin=16), posdef=True)
    ref = plan.factor(A[3]).export()
    x = np.random.default_rng(8).uniform(-1, 1, A[0])
    for w in (64, 512, 1024, 256):
        plan.tune(TUNE_PANEL_OUTER, w)
        fac = plan.factor(A[3])
        assert block_rel_err(fac.export(), ref) < 1e-13, w
        assert np.abs(fac.solve(matvec(A, x)) - x).max() < 1e-12, w
    for seg in (64, 100, 1 << 20, 4096):
        plan.tune(TUNE_CB_SEG, seg)
        fac = plan.factor(A[3])
        assert block_rel_err(fac.export(), ref) < 1e-13, seg
        assert np.abs(fac.solve(matvec(A, x)) - x).max() < 1e-12, seg
    with pytest.raises(SmfGpuError):
        plan.tune(TUNE_PANEL_OUTER, 100)


def test_replayed_factor_matches_eager():
    """The SPD launch sequence is captured once and replayed on refactors.
    Three replays with different values (host, then device memory), with
    solves between them, against the eager sequence (timing on) on the same
    values: identical factors bit for bit, and correct solves."""
    import torch
    print("case graph replay", flush=True)
    m = 14
    A = g.spd27_3d(m, seed=5)
    plan = Plan(analyse(A, perm=o.grid_nd((m, m, m), leaf=32), nemin=16), posdef=True)
    x = np.random.default_rng(5).uniform(-1, 1, A[0])
    for rep, scale in enumerate((1.0, 3.0, 0.25)):
        vals = A[3] * scale
        plan.set_timing(True)
        eager = plan.factor(vals).export()
        plan.set_timing(False)
        if rep < 2:
            fac = plan.factor(vals)
        else:
            vd = torch.from_numpy(vals).to("cuda:0")
            torch.cuda.synchronize()
            fac = plan.factor(values_ptr=vd.data_ptr())
        got = fac.export()
        assert np.array_equal(got["factor_values"], eager["factor_values"]), rep
        Ak = (A[0], A[1], A[2], vals)
        assert np.abs(fac.solve(matvec(Ak, x)) - x).max() < 1e-12, rep


def test_full_size_c4_property():
    """1M-row 27-point system at full size: FP64 solve to eta <= 1e-14 and
    mixed + refinement to <= 1e-12 (size-independent properties)."""
    if os.environ.get("SMFGPU_SKIP_BIG"):
        pytest.skip("big case disabled")
    print("case c4 full size", flush=True)
    m = 100
    A = g.spd27_3d(m, seed=0)
    an = analyse(A, perm=o.grid_nd((m, m, m), leaf=64), nemin=32)
    plan = Plan(an, posdef=True)
    x = np.random.default_rng(0).uniform(-1, 1, A[0])
    b = matvec(A, x)
    fac = plan.factor(A[3])
    assert fac.status == 0
    xs = fac.solve(b)
    _, st = fac.refine(b, xs, max_iters=0)
    assert st["eta"] <= 1e-14
    fac = plan.factor(A[3], precision=PREC_FP32, mixed_min_front=1024)
    xr, st = fac.refine(b, fac.solve(b), max_iters=10, tol=1e-14)
    assert st["eta"] <= 1e-12
