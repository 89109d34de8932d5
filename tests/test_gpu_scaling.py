"""GPU parity of symmetric scaling (smf's UserSupplied and Equilibration
scaling, factor_indef.cpp:508-559, factor_posdef.cpp:405-421, solve.cpp:
1022-1067, scaling.cpp:57-119) through the C ABI: the device equilibration
against the restatement and the reference, the scaled factorization against
the oracle's factorization of S A S (the reference's own identity, pinned in
tests/test_scaling.py), scaled solves against the reference's, and the
refinement of a scaled factor against the unscaled matrix."""
import numpy as np
import pytest

from tests.conftest import gpu_available
from tests.helpers import matvec
from tests.test_gpu_ldlt import compare_factor
from tests.test_gpu_spd import block_rel_err

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

import torch  # noqa: E402

from opt_solver_amd import generators as g  # noqa: E402
from opt_solver_amd import ordering as o  # noqa: E402
from opt_solver_amd.solver import PREC_BF16X3, Plan, SmfGpuError, analyse  # noqa: E402
from oracle import scaling as osc  # noqa: E402


def _case(name):
    if name == "kkt_30":
        A, _ = g.kkt_grid(30, 180, seed=30)
        return A, False
    if name == "ocp_T60":
        return g.ocp_kkt(60, seed=60), False
    if name == "spd27_12":
        return g.spd27_3d(12, seed=2), True
    raise KeyError(name)


def _arrays(A, posdef, amd=False):
    from oracle import refshim
    if refshim.available():
        order = refshim.ORDER_AMD if amd else refshim.ORDER_METIS
        return refshim.RefAnalysis(A, ordering=order, posdef=posdef).plan_arrays(), order
    if posdef:
        m = round(A[0] ** (1 / 3))
        return analyse(A, perm=o.grid_nd((m, m, m), leaf=16), nemin=16).arrays(), None
    return analyse(A, nemin=16).arrays(), None


@pytest.mark.parametrize("name", ["kkt_30", "ocp_T60", "spd27_12"])
def test_equilibration_on_device(name):
    """smfgpu_equilibration_scale equals the restated Equilibration scale to
    the bit (and the reference's, when built), from host and device values."""
    from oracle import refshim
    A, posdef = _case(name)
    print("case equilibration", name, flush=True)
    arrays, order = _arrays(A, posdef, amd=name.startswith("ocp"))
    plan = Plan(arrays, posdef=posdef)
    s = plan.equilibration_scale(A[3])
    s_orc = osc.equilibration_scale(*A, iters=3)
    assert np.array_equal(s, s_orc)
    vd = torch.from_numpy(np.ascontiguousarray(A[3])).cuda()
    assert np.array_equal(plan.equilibration_scale(values_ptr=vd.data_ptr()), s_orc)
    assert np.array_equal(plan.equilibration_scale(A[3], iters=1), osc.equilibration_scale(*A, iters=1))
    if refshim.available():
        an = refshim.RefAnalysis(A, ordering=order, posdef=posdef)
        rf = refshim.RefFactor(an, A[3], u=0.01, equilibrate=True)
        assert np.array_equal(s, rf.scale)
    bad = np.array(A[3], copy=True)
    bad[3] = np.nan
    with pytest.raises(SmfGpuError):
        plan.equilibration_scale(bad)


@pytest.mark.parametrize("name,kind", [("kkt_30", "random"), ("kkt_30", "equil"), ("ocp_T60", "random"),
                                       ("spd27_12", "random")])
def test_scaled_factor_against_oracle(oracle_lib, name, kind):
    """The factor of A with scale S equals the oracle's factor of the
    restated S A S: steps, rows, pivot types, inertia, delays and 2x2
    counts exactly, L and D within 1e-12; the full solve equals
    S (S A S)^-1 S b; the forward and backward jobs equal the reference's
    scaled jobs; the refinement against the unscaled A reaches 1e-14."""
    from oracle import refshim
    A, posdef = _case(name)
    n = A[0]
    print("case scaled", name, kind, flush=True)
    arrays, order = _arrays(A, posdef, amd=name.startswith("ocp"))
    s = (10.0 ** np.random.default_rng(11).uniform(-2, 2, n) if kind == "random"
         else osc.equilibration_scale(*A, iters=3))
    plan = Plan(arrays, posdef=posdef)
    fac = plan.factor(A[3], u=0.01, scale=s)
    sas = osc.scale_values(*A, s)
    orc = oracle_lib.Factor(arrays, sas, posdef=posdef)
    assert fac.status == orc.status == 0
    assert fac.inertia == orc.inertia
    if not posdef:
        assert fac.info.delayed_pivots == orc.delayed and fac.info.two_by_two_pivots == orc.two_by_two
    ex, eo = fac.export(), orc.export()
    if posdef:
        assert np.array_equal(ex["steps"], eo["steps"]) and np.array_equal(ex["row_indices"], eo["row_indices"])
        err = block_rel_err(ex, eo)
        assert err < 1e-12
    else:
        err = compare_factor(ex, eo, 1e-12)
    b = np.random.default_rng(12).uniform(-1, 1, (n, 3))
    x = fac.solve(b)
    x_orc = osc.scaled_solution(orc.solve, b, s)
    assert np.abs(x - x_orc).max() <= 1e-9 * max(np.abs(x_orc).max(), 1.0)
    eta, _ = oracle_lib.backward_error(A, b[:, 0], x[:, 0])
    if refshim.available():
        an = refshim.RefAnalysis(A, ordering=order, posdef=posdef)
        rs = refshim.RefFactor(an, A[3], u=0.01, scale=s)
        if not posdef:  # the reference's Cholesky path leaves the inertia counts at zero
            assert fac.inertia == rs.inertia
        for job in (1, 3):
            xj, xr = fac.solve(b, job=job), rs.solve_job(b, job)
            assert np.abs(xj - xr).max() <= 1e-9 * max(np.abs(xr).max(), 1.0), job
    xr, st = fac.refine(b[:, 0], x[:, 0], max_iters=5, tol=1e-15)
    eta2, _ = oracle_lib.backward_error(A, b[:, 0], xr)
    print(f"  L err {err:.1e} eta {eta:.1e} refined {eta2:.1e} max|SAS| {fac.info.max_abs:.3e}")
    assert eta2 <= 1e-14
    assert fac.info.max_abs == np.abs(sas).max()


def test_scaled_large_fronts_and_refactor(oracle_lib):
    """A grid KKT whose fronts take the multi-workgroup path: scaled FP64
    against the oracle; split BF16 refined to 1e-12; then an unscaled
    factorization of the same plan equals a fresh plan's (the scaled state
    does not leak), and the scaled one again repeats bit for bit."""
    A, _ = g.kkt_grid(48, 460, seed=48)
    n = A[0]
    print("case scaled large fronts kkt_48", flush=True)
    arrays, _ = _arrays(A, False)
    s = osc.equilibration_scale(*A, iters=3)
    plan = Plan(arrays, posdef=False)
    plan.tune(5, 96)  # fronts from 96 rows take the multi-workgroup path
    fac = plan.factor(A[3], u=0.01, scale=s)
    assert fac.info.max_front >= 96
    orc = oracle_lib.Factor(arrays, osc.scale_values(*A, s), posdef=False)
    assert fac.inertia == orc.inertia and fac.info.delayed_pivots == orc.delayed
    e1 = fac.export()
    compare_factor(e1, orc.export(), 1e-12)
    b = matvec(A, np.random.default_rng(3).uniform(-1, 1, n))
    vd = torch.from_numpy(np.ascontiguousarray(A[3])).cuda()
    bd = torch.from_numpy(b).cuda()
    xd = torch.empty_like(bd)
    fb = plan.factor(values_ptr=vd.data_ptr(), precision=PREC_BF16X3, mixed_min_front=96, scale=s)
    xd.copy_(bd)
    torch.cuda.synchronize()
    fb.solve_device(xd.data_ptr(), 1)
    ri = fb.refine_device(bd.data_ptr(), xd.data_ptr(), 1, max_iters=30, tol=1e-12, max_steps=40)
    assert ri["reached"] and ri["eta"] <= 1e-12
    fu = plan.factor(A[3], u=0.01)
    fresh = Plan(arrays, posdef=False)
    fresh.tune(5, 96)
    compare_factor(fu.export(), fresh.factor(A[3], u=0.01).export(), 0.0)
    x = fu.solve(b)
    assert oracle_lib.backward_error(A, b, x)[0] <= 1e-13
    fac2 = plan.factor(A[3], u=0.01, scale=s)
    compare_factor(fac2.export(), e1, 0.0)


def test_scaled_spd_graph_replay(oracle_lib):
    """SPD factorizations replay a captured launch sequence: scaled,
    unscaled and scaled again on one plan each equal their own first run."""
    A, _ = _case("spd27_12")
    arrays, _ = _arrays(A, True)
    s = 10.0 ** np.random.default_rng(4).uniform(-1, 1, A[0])
    plan = Plan(arrays, posdef=True)
    e_s = plan.factor(A[3], scale=s).export()["factor_values"]
    e_u = plan.factor(A[3]).export()["factor_values"]
    for _ in range(2):
        assert np.array_equal(plan.factor(A[3], scale=s).export()["factor_values"], e_s)
        assert np.array_equal(plan.factor(A[3]).export()["factor_values"], e_u)
    orc = oracle_lib.Factor(arrays, osc.scale_values(*A, s), posdef=True)
    ex = plan.factor(A[3], scale=s).export()
    assert block_rel_err(ex, orc.export()) < 1e-12


def test_invalid_scale():
    """A non-positive or non-finite scale entry is refused (host and device
    scale), as validate_positive_scale refuses it."""
    A, _ = _case("ocp_T60")
    arrays, _ = _arrays(A, False)
    plan = Plan(arrays, posdef=False)
    for bad in (0.0, -1.0, np.inf, np.nan):
        s = np.ones(A[0])
        s[A[0] // 2] = bad
        with pytest.raises(SmfGpuError):
            plan.factor(A[3], scale=s)
        sd = torch.from_numpy(s).cuda()
        with pytest.raises(SmfGpuError):
            plan.factor(A[3], scale_ptr=sd.data_ptr())
    fac = plan.factor(A[3], scale=np.ones(A[0]))
    plain = plan.factor(A[3])
    assert fac.status == plain.status == 0
