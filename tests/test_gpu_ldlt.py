"""GPU parity of the indefinite (LDL^T, threshold 1x1/2x2 pivoting with
delayed pivots) numeric phase through the C ABI."""
import numpy as np
import pytest

from tests.conftest import gpu_available
from tests.helpers import matvec
from tests.test_golden import load_mf

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from opt_solver_amd import generators as g  # noqa: E402
from opt_solver_amd.solver import Plan, SmfGpuError, analyse  # noqa: E402


def compare_factor(ex, eo, tol):
    assert np.array_equal(ex["steps"], eo["steps"])
    assert np.array_equal(ex["row_indices"], eo["row_indices"])
    assert np.array_equal(ex["pivot_types"], eo["pivot_types"])
    err = 0.0
    for s in range(ex["steps"].shape[0]):
        ro, fo, p, f = ex["steps"][s]
        a = ex["factor_values"][fo:fo + f * p]
        b = eo["factor_values"][fo:fo + f * p]
        if a.size:
            err = max(err, np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    assert err <= tol, err
    return err


@pytest.mark.parametrize("name", ["ocp_T10", "kkt_12"])
def test_golden_indefinite(name):
    """The reference's own analysis and factor (tests/golden): identical
    steps, row order, pivot types, inertia, delays and 2x2 count; every
    front is below the blocked size, so L and D are bitwise the
    reference's."""
    z, a, A = load_mf(name)
    print("case golden", name, flush=True)
    plan = Plan(a, posdef=False)
    fac = plan.factor(A[3], u=0.01)
    st = z["stats"]
    assert fac.status == int(st[0])
    assert fac.inertia == (int(st[1]), int(st[2]), int(st[3]))
    assert fac.info.delayed_pivots == int(st[4]) and fac.info.two_by_two_pivots == int(st[5])
    assert fac.flops == st[6]
    ex = fac.export()
    gold = dict(steps=z["steps"], row_indices=z["row_indices"], pivot_types=z["pivot_types"],
                factor_values=z["factor_values"])
    err = compare_factor(ex, gold, 0.0)
    x = fac.solve(z["b"])
    assert np.abs(x - z["x_ref"]).max() < 1e-10
    print("  bitwise, err", err)


def _arrays_for(A, ref_order):
    from oracle import refshim
    if refshim.available():
        return refshim.RefAnalysis(A, ordering=ref_order, posdef=False).plan_arrays()
    return analyse(A, nemin=16).arrays()


@pytest.mark.parametrize("case", ["ocp_T60", "ocp_T200", "kkt_30", "kkt_48"])
def test_indefinite_against_oracle(oracle_lib, case):
    """OCP KKTs (many delays, 2x2 pivots) and grid KKTs with fronts past the
    blocked size, on the reference's analysis when it is built: the GPU
    factor against the oracle's, inertia against Sylvester's law."""
    from oracle import refshim
    print("case", case, flush=True)
    if case.startswith("ocp"):
        T = int(case[5:])
        A = g.ocp_kkt(T, seed=T)
        inertia = (16 * T + 12, 12 * T, 0)
        order = refshim.ORDER_AMD
    else:
        gsz = int(case[4:])
        m = gsz * gsz // 5
        A, _ = g.kkt_grid(gsz, m, seed=gsz)
        inertia = (gsz * gsz, m, 0)
        order = refshim.ORDER_METIS
    arrays = _arrays_for(A, order)
    plan = Plan(arrays, posdef=False)
    fac = plan.factor(A[3], u=0.01)
    orc = oracle_lib.Factor(arrays, A[3], posdef=False)
    assert fac.status == orc.status == 0
    assert fac.inertia == orc.inertia == inertia
    assert fac.info.delayed_pivots == orc.delayed and fac.info.two_by_two_pivots == orc.two_by_two
    assert fac.flops == orc.flops
    err = compare_factor(fac.export(), orc.export(), 1e-12)
    x = np.random.default_rng(1).uniform(-1, 1, A[0])
    b = matvec(A, x)
    xs = fac.solve(b)
    eta, om = oracle_lib.backward_error(A, b, xs)
    xr, st = fac.refine(b, xs, max_iters=5, tol=1e-15)
    eta2, _ = oracle_lib.backward_error(A, b, xr)
    print(f"  max front {fac.info.max_front} delays {fac.info.delayed_pivots} 2x2 {fac.info.two_by_two_pivots} "
          f"L err {err:.1e} eta {eta:.1e} refined {eta2:.1e}")
    assert eta2 <= 1e-14


def test_singular_reported(oracle_lib):
    """A saddle point with a rank-deficient constraint block: zero pivots at
    the root, status Singular, as the oracle."""
    print("case singular", flush=True)
    n = 6
    rows = [0, 1, 2, 3, 4, 5, 3, 4, 5, 5]
    cols = [0, 1, 2, 3, 4, 5, 0, 0, 1, 2]
    vals = [2.0, 2.0, 2.0, 0.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0]
    A = g.coo_lower_to_csc(n, rows, cols, vals)
    an = analyse(A, nemin=8)
    plan = Plan(an, posdef=False)
    fac = plan.factor(A[3])
    orc = oracle_lib.Factor(an.arrays(), A[3], posdef=False)
    assert fac.status == orc.status == 2
    assert fac.inertia == orc.inertia
    with pytest.raises(SmfGpuError):
        fac.solve(np.ones(n))


def test_inertia_limit():
    """max_negative_pivots (multifrontal_serial.cpp:1619-1627): a KKT with
    12 T negative eigenvalues factors with the limit at 12 T and stops with
    InertiaExceeded below it (the factor is then unusable, as the
    reference's, factor_indef.cpp:682-715)."""
    print("case inertia limit", flush=True)
    T = 20
    A = g.ocp_kkt(T, seed=3)
    plan = Plan(analyse(A, nemin=16), posdef=False)
    ok = plan.factor(A[3], max_negative_pivots=12 * T)
    assert ok.status == 0 and ok.inertia == (16 * T + 12, 12 * T, 0)
    stop = plan.factor(A[3], max_negative_pivots=12 * T - 1)
    assert stop.status == 4
    assert stop.inertia[1] > 12 * T - 1
    with pytest.raises(SmfGpuError):
        stop.solve(np.ones(A[0]))
    again = plan.factor(A[3], max_negative_pivots=-1)
    assert again.status == 0 and again.inertia == (16 * T + 12, 12 * T, 0)


def _large_path_case(case):
    from oracle import refshim
    if case.startswith("kkt_grid"):
        gsz = int(case.split("_")[2])
        m = gsz * gsz // 5
        A, _ = g.kkt_grid(gsz, m, seed=gsz)
        return A, analyse(A, nemin=16).arrays(), (gsz * gsz, m, 0)
    if case.startswith("ocp"):
        T = int(case[5:])
        A = g.ocp_kkt(T, seed=T)
        return A, _arrays_for(A, refshim.ORDER_AMD), (16 * T + 12, 12 * T, 0)
    gsz = int(case[4:])
    m = gsz * gsz // 5
    A, _ = g.kkt_grid(gsz, m, seed=gsz)
    return A, _arrays_for(A, refshim.ORDER_METIS), (gsz * gsz, m, 0)


@pytest.mark.parametrize("mode", ["group", "panel", "one_wg"])
@pytest.mark.parametrize("case", ["kkt_30", "kkt_48", "ocp_T200", "kkt_grid_60"])
def test_large_front_path_against_oracle(oracle_lib, case, mode):
    """Every blocked front (f >= 96) on the multi-workgroup path with
    speculative panel groups ("group") or one panel at a time ("panel"), or
    on the one-workgroup blocked kernel ("one_wg"): the same pivot sequence,
    delays, 2x2 pivots and inertia as the oracle's blocked kernel, L and D
    within 1e-12 per block (the deferred updates past a panel group run on
    MFMA in another summation order), solves to 1e-14."""
    from opt_solver_amd._lib import TUNE_LDLT_BIG_MIN, TUNE_LDLT_GROUP
    print("case large path", case, mode, flush=True)
    A, arrays, inertia = _large_path_case(case)
    plan = Plan(arrays, posdef=False)
    plan.tune(TUNE_LDLT_BIG_MIN, 1 << 30 if mode == "one_wg" else 96)
    plan.tune(TUNE_LDLT_GROUP, 1 if mode == "group" else 0)
    fac = plan.factor(A[3], u=0.01)
    orc = oracle_lib.Factor(arrays, A[3], posdef=False)
    assert fac.inertia == orc.inertia == inertia
    assert fac.info.delayed_pivots == orc.delayed and fac.info.two_by_two_pivots == orc.two_by_two
    # stalled panels, partner pulls, rotated columns, second-pass searches
    stalls = tuple(int(v) for v in fac.stall_stats().sum(axis=0))
    assert stalls == orc.stall_stats, (stalls, orc.stall_stats)
    err = compare_factor(fac.export(), orc.export(), 1e-12)
    print(f"  L/D block rel err {err:.1e} stall counters {stalls}", flush=True)
    x = np.random.default_rng(7).uniform(-1, 1, A[0])
    b = matvec(A, x)
    xs = fac.solve(b)
    eta, _ = oracle_lib.backward_error(A, b, xs)
    assert eta <= 1e-14, eta


def test_large_front_default_threshold(oracle_lib):
    """A grid KKT whose fronts of 256 rows and more (the largest 3016) take
    the large-front path at the default threshold: Sylvester inertia and a
    backward error at FP64 level (too large for the CPU oracle's dense
    kernel in a test)."""
    print("case kkt_grid_120", flush=True)
    gsz = 120
    m = gsz * gsz // 5
    A, _ = g.kkt_grid(gsz, m, seed=gsz)
    an = analyse(A, nemin=16)
    assert an.max_front >= 2048
    plan = Plan(an, posdef=False)
    fac = plan.factor(A[3], u=0.01)
    assert fac.status == 0 and fac.inertia == (gsz * gsz, m, 0)
    x = np.random.default_rng(9).uniform(-1, 1, A[0])
    b = matvec(A, x)
    xs = fac.solve(b)
    eta, _ = oracle_lib.backward_error(A, b, xs)
    assert eta <= 1e-13, eta


@pytest.mark.parametrize("prec", ["bf16x3", "bf16"])
def test_large_front_planes_bitwise(oracle_lib, prec):
    """The large fronts' outer-block updates in BF16 and split BF16 read
    bf16 copies of their operands written once per update (the held-back
    pivots' flush past each outer block, the contribution block's chunks)
    in 256 x 256 tiles: with 512-column outer blocks, the same factor, bit
    for bit, as rounding the FP64 operands inside every 128 x 128 tile (the
    same roundings, chunks and MFMA order per accumulator), on a grid KKT
    whose largest fronts (3016 rows) have several outer blocks and
    contribution blocks. With the default 1024-column outer blocks: the
    FP64 factor's inertia, and refinement to eta <= 1e-12."""
    from opt_solver_amd._lib import TUNE_LDLT_OUTER_WIDE, TUNE_LDLT_PLANES
    from opt_solver_amd.solver import PREC_BF16, PREC_BF16X3
    print("case planes", prec, flush=True)
    gsz = 120
    m = gsz * gsz // 5
    A, _ = g.kkt_grid(gsz, m, seed=gsz)
    an = analyse(A, nemin=16)
    plan = Plan(an, posdef=False)
    code = {"bf16x3": PREC_BF16X3, "bf16": PREC_BF16}[prec]
    ex = {}
    plan.tune(TUNE_LDLT_OUTER_WIDE, 512)
    for planes in (1, 0):
        plan.tune(TUNE_LDLT_PLANES, planes)
        fac = plan.factor(A[3], u=0.01, precision=code, mixed_min_front=512)
        assert fac.status == 0 and fac.info.fronts_reduced > 0
        ex[planes] = (fac.export(), fac.inertia, fac.info.delayed_pivots)
    assert ex[0][1:] == ex[1][1:]
    for key in ("steps", "row_indices", "pivot_types", "factor_values"):
        assert np.array_equal(ex[0][0][key], ex[1][0][key]), key
    print(f"  reduced fronts {fac.info.fronts_reduced} inertia {ex[1][1]}", flush=True)
    plan.tune(TUNE_LDLT_PLANES, 1)
    plan.tune(TUNE_LDLT_OUTER_WIDE, 1024)
    fac = plan.factor(A[3], u=0.01, precision=code, mixed_min_front=512)
    assert fac.status == 0 and fac.inertia == (gsz * gsz, m, 0) == ex[1][1]
    x = np.random.default_rng(9).uniform(-1, 1, A[0])
    b = matvec(A, x)
    xr, st = fac.refine(b, fac.solve(b), max_iters=30, tol=1e-12, max_steps=40)
    eta, _ = oracle_lib.backward_error(A, b, xr)
    print(f"  outer 1024: refinement steps {st['steps']} eta {eta:.1e}", flush=True)
    assert eta <= 1e-12, (eta, st)


@pytest.mark.parametrize("case", ["ocp_T200", "kkt_grid_60"])
def test_blocked_solve_multi_rhs(oracle_lib, case):
    """The blocked LDL^T sweeps (64-pivot steps, the rows past a front's
    pivots by atomic add) with 1, 5, 16 and 21 right-hand sides (21: two
    passes of 16): against the known solutions, the reference backward-error
    formula, and column by column against the single-RHS solve. ocp_T200
    has delayed and 2x2 pivots, kkt_grid_60 fronts of several blocks."""
    print("case blocked solve", case, flush=True)
    A, arrays, _ = _large_path_case(case)
    plan = Plan(arrays, posdef=False)
    fac = plan.factor(A[3], u=0.01)
    assert fac.status == 0
    print(f"  delays {fac.info.delayed_pivots} 2x2 {fac.info.two_by_two_pivots}", flush=True)
    if case == "ocp_T200":
        assert fac.info.delayed_pivots > 0 and fac.info.two_by_two_pivots > 0
    X = np.random.default_rng(11).uniform(-1, 1, (A[0], 21))
    B = matvec(A, X)
    for nr in (1, 5, 16, 21):
        XS = fac.solve(B[:, :nr] if nr > 1 else B[:, 0]).reshape(A[0], nr)
        eta, _ = oracle_lib.backward_error(A, B[:, :nr], XS)
        assert eta <= 1e-14, (nr, eta)
        assert np.abs(XS - X[:, :nr]).max() < 1e-8, nr
        if nr == 21:
            for r in (0, 15, 16, 20):
                assert np.abs(fac.solve(B[:, r]) - XS[:, r]).max() <= 1e-12 * max(1.0, np.abs(XS[:, r]).max())


@pytest.mark.parametrize("prec", [1, 4, 2])
@pytest.mark.parametrize("case", ["kkt_grid_60", "kkt_48"])
def test_mixed_precision_ldlt(oracle_lib, case, prec):
    """FP32 / split-BF16 / BF16 MFMA updates on every large front (the pivot
    logic in FP64), then FP64 refinement against A: scaled backward error
    <= 1e-12 (the mixed gate) and the Sylvester inertia."""
    from opt_solver_amd._lib import TUNE_LDLT_BIG_MIN
    print("case mixed ldlt", case, "prec", prec, flush=True)
    A, arrays, inertia = _large_path_case(case)
    plan = Plan(arrays, posdef=False)
    plan.tune(TUNE_LDLT_BIG_MIN, 96)
    fac = plan.factor(A[3], u=0.01, precision=prec, mixed_min_front=0)
    assert fac.info.fronts_reduced > 0
    assert fac.inertia == inertia
    x = np.random.default_rng(13).uniform(-1, 1, A[0])
    b = matvec(A, x)
    x0 = fac.solve(b)
    xr, st = fac.refine(b, x0, max_iters=30, tol=1e-14)
    eta, _ = oracle_lib.backward_error(A, b, xr)
    print(f"  reduced {fac.info.fronts_reduced} refine steps {st['steps']} eta {eta:.2e}", flush=True)
    assert eta <= 1e-12, eta


@pytest.mark.parametrize("case", ["ocp_T60", "ocp_T200", "kkt_30", "kkt_48"])
def test_tiny_u_against_reference(oracle_lib, case):
    """IPOPT's pivot tolerance u = 1e-8 on the reference's own analysis: the
    GPU factor against the compiled reference's (steps, row order, pivot
    types, inertia, delays, 2x2 pivots; L and D within 1e-12 per block) and
    against the oracle's."""
    from oracle import refshim
    if not refshim.available():
        pytest.skip("reference library not built")
    print("case tiny u sparse", case, flush=True)
    if case.startswith("ocp"):
        T = int(case[5:])
        A = g.ocp_kkt(T, seed=T)
        order = refshim.ORDER_AMD
    else:
        gsz = int(case[4:])
        A, _ = g.kkt_grid(gsz, gsz * gsz // 5, seed=gsz)
        order = refshim.ORDER_METIS
    an = refshim.RefAnalysis(A, ordering=order, posdef=False)
    arrays = an.plan_arrays()
    ref = refshim.RefFactor(an, A[3], u=1e-8)
    plan = Plan(arrays, posdef=False)
    fac = plan.factor(A[3], u=1e-8)
    orc = oracle_lib.Factor(arrays, A[3], posdef=False, u=1e-8)
    assert fac.status == ref.status == 0
    assert fac.inertia == ref.inertia == orc.inertia
    assert fac.info.delayed_pivots == ref.delayed == orc.delayed
    assert fac.info.two_by_two_pivots == ref.two_by_two == orc.two_by_two
    ex = fac.export()
    e_ref = compare_factor(ex, ref.export(), 1e-12)
    e_orc = compare_factor(ex, orc.export(), 1e-12)
    print(f"  delays {ref.delayed} 2x2 {ref.two_by_two} L/D vs reference {e_ref:.1e} vs oracle {e_orc:.1e}",
          flush=True)


def test_limits_against_reference():
    """max_negative_pivots and the work limit against the compiled
    reference's own stopped runs (one thread, serial postorder): the same
    status and, for every limit, the partial counts the reference reports
    (inertia for an inertia stop; delays, 2x2 pivots, front work and flops
    for both)."""
    from oracle import refshim
    from opt_solver_amd._lib import STATUS_INERTIA_EXCEEDED, STATUS_MAX_PIVOT_DELAYS
    if not refshim.available():
        pytest.skip("reference library not built")
    print("case limits vs reference", flush=True)
    T = 200
    A = g.ocp_kkt(T, seed=T)
    an = refshim.RefAnalysis(A, ordering=refshim.ORDER_AMD, posdef=False)
    assert np.array_equal(an.postorder, np.arange(an.nsup))  # supernode order is the reference's order
    plan = Plan(an.plan_arrays(), posdef=False)
    full = refshim.RefFactor(an, A[3], u=0.01)
    total = full.front_work
    outcomes = set()
    for lim in (12 * T, 12 * T - 1, 6 * T, 100, 0):
        ref = refshim.RefFactor(an, A[3], u=0.01, max_negative_pivots=lim)
        fac = plan.factor(A[3], u=0.01, max_negative_pivots=lim)
        print(f"  negatives <= {lim}: status {fac.status} / {ref.status} inertia {fac.inertia} / {ref.inertia}",
              flush=True)
        assert fac.status == ref.status
        assert fac.inertia == ref.inertia
        assert fac.info.delayed_pivots == ref.delayed and fac.info.two_by_two_pivots == ref.two_by_two
        assert fac.info.front_work == ref.front_work and fac.flops == ref.flops
        outcomes.add(fac.status)
    for lim in (total + 0.5, total - 1, total / 2, 1000.0, 0.0):
        ref = refshim.RefFactor(an, A[3], u=0.01, work_limit=lim)
        fac = plan.factor(A[3], u=0.01, work_limit=lim)
        print(f"  work <= {lim}: status {fac.status} / {ref.status} work {fac.info.front_work} / {ref.front_work}",
              flush=True)
        assert fac.status == ref.status
        assert fac.info.delayed_pivots == ref.delayed and fac.info.two_by_two_pivots == ref.two_by_two
        assert fac.info.front_work == ref.front_work and fac.flops == ref.flops
        if fac.status == 0:
            assert fac.inertia == ref.inertia
        outcomes.add(fac.status)
    assert {0, STATUS_INERTIA_EXCEEDED, STATUS_MAX_PIVOT_DELAYS} <= outcomes


def test_work_limit(oracle_lib):
    """Work limit (multifrontal_serial.cpp:653-671): a run stops, with
    MaxPivotDelays and an unusable factor, exactly when its total front
    work sum(num_elig * f^2) (realized sizes, delays included) exceeds
    floor(work_limit); the reported work then passes the limit. With the
    limit at the total, or off, the factor is the unlimited one."""
    from opt_solver_amd._lib import STATUS_MAX_PIVOT_DELAYS
    print("case work limit", flush=True)
    A, arrays, inertia = _large_path_case("ocp_T200")
    plan = Plan(arrays, posdef=False)
    full = plan.factor(A[3], u=0.01)
    total = full.info.front_work
    assert full.status == 0 and full.info.delayed_pivots > 0 and total == int(total)
    ref = full.export()
    at = plan.factor(A[3], u=0.01, work_limit=total + 0.5)
    assert at.status == 0 and at.inertia == inertia
    assert np.array_equal(at.export()["factor_values"], ref["factor_values"])
    for lim in (total - 1, total / 2, 0.0):
        st = plan.factor(A[3], u=0.01, work_limit=lim)
        assert st.status == STATUS_MAX_PIVOT_DELAYS, lim
        assert st.info.front_work > lim
        with pytest.raises(SmfGpuError):
            st.solve(np.ones(A[0]))


def test_multichunk_assembly_against_oracle(oracle_lib):
    """Fronts of more than 2048 rows (the LDL^T assembly's LDS chunk): every
    child column crosses several chunks, each starting at the previous
    chunk's end. Against the oracle, and three refactorizations bit for bit
    equal to each other (an assembly that lost a chunk's entries now and
    then would show here)."""
    print("case kkt_110 multi-chunk assembly", flush=True)
    A, _ = g.kkt_grid(110, 2400, seed=110)
    from oracle import refshim
    arrays = _arrays_for(A, refshim.ORDER_METIS if refshim.available() else None)
    assert int(np.diff(arrays["front_row_ptr"]).max()) > 2048
    plan = Plan(arrays, posdef=False)
    fac = plan.factor(A[3], u=0.01)
    orc = oracle_lib.Factor(arrays, A[3], posdef=False)
    assert fac.status == orc.status == 0
    assert fac.inertia == orc.inertia == (110 * 110, 2400, 0)
    assert fac.info.delayed_pivots == orc.delayed and fac.info.two_by_two_pivots == orc.two_by_two
    e0 = fac.export()
    err = compare_factor(e0, orc.export(), 1e-12)
    for _ in range(3):
        compare_factor(plan.factor(A[3], u=0.01).export(), e0, 0.0)
    x = np.random.default_rng(2).uniform(-1, 1, A[0])
    b = matvec(A, x)
    f2 = plan.factor(A[3], u=0.01)
    xs = f2.solve(b)
    xr, _ = f2.refine(b, xs, max_iters=5, tol=1e-15)
    eta, _ = oracle_lib.backward_error(A, b, xr)
    print(f"  max front {fac.info.max_front} L err {err:.1e} refined eta {eta:.1e}")
    assert eta <= 1e-14
