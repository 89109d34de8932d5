"""Batched independent trajectory-optimisation KKTs on the GPU, factored
as one forest (every level's fronts of all systems in the same launches):
per-system inertia, delayed pivots and 2x2 pivots against the reference's
one-system-at-a-time factor, the whole factor bit for bit against the
oracle, solves against the known solutions. The levels of small fronts run
laid out on the device (default), partly (device fronts capped at 28 rows,
so levels whose delays grow a front past that go to the host layout), not
at all, and on the device with 256 threads for every front instead of one
wavefront for the small ones: the same factor each way."""
import numpy as np
import pytest

from tests.conftest import gpu_available
from tests.helpers import matvec

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from opt_solver_amd import batch  # noqa: E402
from opt_solver_amd._lib import TUNE_LDLT_DEVICE_FRONTS, TUNE_LDLT_DEVICE_WAVE  # noqa: E402
from opt_solver_amd.solver import Plan, analyse  # noqa: E402


@pytest.mark.parametrize("dev_fronts,wave", [(95, 1), (28, 1), (0, 1), (95, 0)])
def test_ocp_batch_per_system_parity(oracle_lib, dev_fronts, wave):
    from oracle import refshim
    nsys = 32
    print(f"case ocp batch {nsys} device fronts {dev_fronts} one-wave class {wave}", flush=True)
    mats, Ts = batch.ocp_batch(nsys, 10, 60, seed=21)
    ref = refshim.available()
    if ref:
        ans = [refshim.RefAnalysis(A, ordering=refshim.ORDER_AMD, posdef=False) for A in mats]
        arrs = [an.plan_arrays() for an in ans]
    else:
        arrs = [analyse(A, nemin=16).arrays() for A in mats]
    U, offs = batch.union_arrays(arrs)
    vals = np.concatenate([A[3] for A in mats])
    plan = Plan(U, posdef=False)
    plan.tune(TUNE_LDLT_DEVICE_FRONTS, dev_fronts)
    plan.tune(TUNE_LDLT_DEVICE_WAVE, wave)
    fac = plan.factor(vals, u=0.01)
    nlev = plan.stats()["levels"]
    dev = plan.kernel_times()["device_levels"]
    print(f"  levels {nlev:.0f} on the device {dev:.0f}", flush=True)
    if dev_fronts == 95:
        assert dev == nlev
    elif dev_fronts == 0:
        assert dev == 0
    else:
        assert 0 < dev < nlev
    per = batch.system_stats(fac.front_stats(), offs["sn"])
    for i, T in enumerate(Ts):
        assert tuple(per[i, :3]) == (16 * T + 12, 12 * T, 0), i
    if ref:
        for i, (an, A) in enumerate(zip(ans, mats)):
            rf = refshim.RefFactor(an, A[3], u=0.01)
            assert tuple(per[i, :3]) == rf.inertia, i
            assert per[i, 3] == rf.two_by_two and per[i, 4] == rf.delayed, (i, per[i], rf.two_by_two, rf.delayed)
    orc = oracle_lib.Factor(U, vals, posdef=False)
    eg, eo = fac.export(), orc.export()
    assert np.array_equal(eg["steps"], eo["steps"])
    assert np.array_equal(eg["row_indices"], eo["row_indices"])
    assert np.array_equal(eg["pivot_types"], eo["pivot_types"])
    assert np.array_equal(eg["factor_values"], eo["factor_values"])
    X = [np.random.default_rng(i).uniform(-1, 1, A[0]) for i, A in enumerate(mats)]
    b = np.concatenate([matvec(A, x) for A, x in zip(mats, X)])
    xs = fac.solve(b)
    assert np.abs(xs - np.concatenate(X)).max() < 1e-8
    print("  delays", int(per[:, 4].sum()), "2x2", int(per[:, 3].sum()), flush=True)


def test_full_batch_host_layout_large_launches():
    """The 1024-system batch on the reference's analysis with the host layout
    (device levels off): its first level's assembly has more than 2^22
    columns, so the launch goes out in pieces (a dispatch counts its
    work-items in 32 bits). Every system's inertia and the solve are checked."""
    from oracle import refshim
    if not refshim.available():
        pytest.skip("reference library not built")
    nsys = 1024
    print(f"case ocp batch {nsys}, host layout", flush=True)
    mats, Ts = batch.ocp_batch(nsys, 10, 200, seed=0)
    arrs = [refshim.RefAnalysis(A, ordering=refshim.ORDER_AMD, posdef=False).plan_arrays() for A in mats]
    U, offs = batch.union_arrays(arrs)
    vals = np.concatenate([A[3] for A in mats])
    plan = Plan(U, posdef=False)
    plan.tune(TUNE_LDLT_DEVICE_FRONTS, 0)
    fac = plan.factor(vals, u=0.01)
    per = batch.system_stats(fac.front_stats(), offs["sn"])
    for i, T in enumerate(Ts):
        assert tuple(per[i, :3]) == (16 * T + 12, 12 * T, 0), i
    X = np.concatenate([np.random.default_rng(i).uniform(-1, 1, A[0]) for i, A in enumerate(mats)])
    starts = np.cumsum([0] + [A[0] for A in mats])
    b = np.concatenate([matvec(A, X[starts[i]:starts[i + 1]]) for i, A in enumerate(mats)])
    xs = fac.solve(b)
    print(f"  n {U['n']} delays {int(per[:, 4].sum())} err {np.abs(xs - X).max():.1e}", flush=True)
    assert np.abs(xs - X).max() < 1e-8
