"""Static pivoting on the GPU against the compiled reference (same
analysis, same Control): static_pivot_tol (a column with no acceptable
pivot is forced as a 1x1 with |d| raised, ldlt_kernel.cpp:285-320), its
in-order, as-zero and zero-stays variants, and fallback static pivoting on
roots and over-delayed fronts (multifrontal_serial.cpp:562, 615-629).
Inertia, delays, 2x2 and static pivot counts, flops and pivot growth
exactly; factor values to rounding. The matrices keep pivot decisions away
from rounding-level ties: where the choice between candidates rests on
rounding noise (the near-zero Schur complement of a numerically singular
KKT), the blocked kernels, which agree with the reference to rounding and
not bit for bit, may pick another row."""
import numpy as np
import pytest

from tests.conftest import gpu_available
from tests.helpers import matvec

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from opt_solver_amd import generators as g  # noqa: E402
from opt_solver_amd._lib import STATIC_AS_ZERO, STATIC_IN_ORDER, STATIC_ZERO_STAYS_ZERO  # noqa: E402
from opt_solver_amd.solver import Plan  # noqa: E402

CASES = [
    ("ocp_T60", dict(static_pivot_tol=1e-8)),
    ("ocp_T60", dict(static_pivot_tol=1e-8, static_flags=STATIC_IN_ORDER)),
    ("ocp_T30", dict(static_pivot_tol=1e-6, static_flags=STATIC_AS_ZERO)),
    ("kkt_48", dict(static_pivot_tol=1e-8)),
    ("ocp_T60", dict(fallback_static_pivot_tol=1e-8, u=0.1)),
    ("zrow_20", dict(fallback_static_pivot_tol=1e-8)),
    ("zrow_20", dict(static_pivot_tol=1e-10, static_flags=STATIC_ZERO_STAYS_ZERO)),
]


def _matrix(case):
    from oracle import refshim
    if case.startswith("ocp"):
        return g.ocp_kkt(int(case[5:]), seed=4), refshim.ORDER_AMD
    gsz = int(case.split("_")[1])
    if case.startswith("kkt"):
        return g.kkt_grid(gsz, gsz * gsz // 5, seed=gsz)[0], refshim.ORDER_METIS
    # a grid KKT without regularisation whose first 6 constraint rows are
    # zero: exact zero pivots that end at the root (singular)
    n, cp, ri, v = g.kkt_grid(gsz, gsz * gsz // 5, seed=gsz, delta=0.0)[0]
    v = v.copy()
    nx = gsz * gsz
    col = np.repeat(np.arange(n), np.diff(cp))
    dead = ((ri >= nx) & (ri < nx + 6)) | ((col >= nx) & (col < nx + 6))
    v[dead] = 0.0
    return (n, cp, ri, v), refshim.ORDER_METIS


@pytest.mark.parametrize("case,kw", CASES)
def test_static_pivoting_against_reference(case, kw):
    from oracle import refshim
    if not refshim.available():
        pytest.skip("reference library not built")
    print("case", case, kw, flush=True)
    A, order = _matrix(case)
    kw = dict(kw)
    u = kw.pop("u", 0.01)
    an = refshim.RefAnalysis(A, ordering=order, posdef=False)
    rf = refshim.RefFactor(an, A[3], u=u, **kw)
    plan = Plan(an.plan_arrays(), posdef=False)
    fac = plan.factor(A[3], u=u, **kw)
    inf = fac.info
    print(f"  status {fac.status}/{rf.status} inertia {fac.inertia}/{rf.inertia} static {inf.static_pivots}/"
          f"{rf.static_pivots} delays {inf.delayed_pivots}/{rf.delayed} growth {inf.pivot_growth:.6e}/"
          f"{rf.pivot_growth:.6e} max front {inf.max_front}", flush=True)
    assert fac.status == rf.status
    assert fac.inertia == rf.inertia
    assert inf.static_pivots == rf.static_pivots
    assert inf.delayed_pivots == rf.delayed and inf.two_by_two_pivots == rf.two_by_two
    assert fac.flops == rf.flops
    assert inf.pivot_growth == pytest.approx(rf.pivot_growth, rel=1e-9, abs=0)
    ex, eo = fac.export(), rf.export()
    assert np.array_equal(ex["steps"], eo["steps"])
    assert np.array_equal(ex["row_indices"], eo["row_indices"])
    assert np.array_equal(ex["pivot_types"], eo["pivot_types"])
    fv, fo = ex["factor_values"], eo["factor_values"]
    err = np.abs(fv - fo).max() / np.abs(fo).max()
    print(f"  L/D rel err {err:.1e}", flush=True)
    assert err <= 1e-12
    if case.startswith("ocp") and fac.status == 0 and fac.inertia[2] == 0:
        # a perturbed factor: refinement recovers the solution
        x = np.random.default_rng(1).uniform(-1, 1, A[0])
        b = matvec(A, x)
        xs, st = fac.refine(b, fac.solve(b), max_iters=10, tol=1e-14)
        print(f"  refined eta {st['eta']:.1e} in {st['steps']} steps", flush=True)
        assert st["eta"] <= 1e-12
