"""Symmetric scaling (smf's ScalingMethod::UserSupplied and Equilibration):
the numpy restatement (oracle/scaling.py) pinned against the reference
built from its sources and against closed forms, and the reference's own
identity that a scaled factorization is the plain factorization of S A S
(the GPU tests of the scaling rely on it)."""
import numpy as np
import pytest

from opt_solver_amd import generators as g
from oracle import scaling as osc


def _cases():
    A1, _ = g.kkt_grid(14, 40, seed=3)
    A2 = g.ocp_kkt(12, seed=5)
    A3 = g.spd27_3d(6, seed=1)
    return {"kkt_14": (A1, False), "ocp_T12": (A2, False), "spd27_6": (A3, True)}


def test_equilibration_diagonal_closed_form():
    """A diagonal matrix: the first pass gives s_i = 1 / sqrt|a_ii| (a zero
    row keeps 1), and the later passes change nothing once every |s a s| is
    within rounding of 1."""
    d = np.array([4.0, -9.0, 0.0, 2.5e-7, 1e100])
    n = d.size
    cp = np.arange(n + 1, dtype=np.int32)
    ri = np.arange(n, dtype=np.int32)
    s1 = osc.equilibration_scale(n, cp, ri, d, iters=1)
    for i in (0, 1, 3, 4):
        assert s1[i] == 1.0 / np.sqrt(abs(d[i]))
    assert s1[2] == 1.0
    assert np.allclose(s1, [0.5, 1.0 / 3.0, 1.0, 1.0 / np.sqrt(2.5e-7), 1e-50], rtol=1e-15)
    s3 = osc.equilibration_scale(n, cp, ri, d, iters=3)
    sd = np.abs(d) * s3 * s3
    assert np.allclose(sd[d != 0], 1.0, rtol=4e-16)


def test_equilibration_of_unit_rows_is_identity():
    """Rows whose largest |entry| is 1 already: s stays 1 in every pass."""
    A, _ = g.kkt_grid(10, 20, seed=2)
    n, cp, ri, v = A
    r = ri.astype(np.int64)
    c = np.repeat(np.arange(n), np.diff(cp))
    v1 = np.sign(v) * np.where(r == c, 1.0, 0.5)
    v1[v == 0] = 0.0
    s = osc.equilibration_scale(n, cp, ri, v1, iters=3)
    rows = np.zeros(n)
    np.maximum.at(rows, r, np.abs(v1))
    np.maximum.at(rows, c, np.abs(v1))
    assert np.array_equal(s[rows == 1.0], np.ones(int((rows == 1.0).sum())))


def test_scale_values_order():
    """a * (s_r * s_c), not (a * s_r) * s_c: the two differ here."""
    a, s0, s1 = 0.8716638489288124, 0.7566899017869496, 0.13022701777491794
    assert a * (s1 * s0) != (a * s1) * s0
    cp = np.array([0, 2, 3], np.int32)
    ri = np.array([0, 1, 1], np.int32)
    v = np.array([1.0, a, 2.0])
    out = osc.scale_values(2, cp, ri, v, np.array([s0, s1]))
    assert out[1] == a * (s1 * s0)


@pytest.mark.parametrize("name", ["kkt_14", "ocp_T12", "spd27_6"])
def test_equilibration_against_reference(ref_lib, name):
    """The restated Equilibration scale equals the reference's (the scale its
    Equilibration factorization holds) to the bit."""
    A, posdef = _cases()[name]
    an = ref_lib.RefAnalysis(A, ordering=ref_lib.ORDER_METIS, posdef=posdef)
    rf = ref_lib.RefFactor(an, A[3], u=0.01, equilibrate=True)
    assert rf.status == 0 and rf.scale is not None
    s = osc.equilibration_scale(*A, iters=3)
    assert np.array_equal(s, rf.scale)


@pytest.mark.parametrize("name", ["kkt_14", "ocp_T12", "spd27_6"])
def test_user_scale_is_factor_of_sas_in_reference(ref_lib, name):
    """The reference's UserSupplied factorization equals its unscaled
    factorization of the restated S A S bit for bit (steps, rows, pivots,
    L and D, inertia, delays), and its full solve equals S (S A S)^-1 S b
    bit for bit; the forward and backward jobs scale b before and x after."""
    A, posdef = _cases()[name]
    n = A[0]
    s = 10.0 ** np.random.default_rng(7).uniform(-2, 2, n)
    an = ref_lib.RefAnalysis(A, ordering=ref_lib.ORDER_METIS, posdef=posdef)
    rs = ref_lib.RefFactor(an, A[3], u=0.01, scale=s)
    rp = ref_lib.RefFactor(an, osc.scale_values(*A, s), u=0.01)
    assert rs.status == rp.status == 0
    assert np.array_equal(rs.scale, s)
    assert (rs.inertia, rs.delayed, rs.two_by_two) == (rp.inertia, rp.delayed, rp.two_by_two)
    es, ep = rs.export(), rp.export()
    for k in ("steps", "row_indices", "pivot_types", "factor_values"):
        assert np.array_equal(es[k], ep[k]), k
    b = np.random.default_rng(8).uniform(-1, 1, (n, 2))
    assert np.array_equal(rs.solve(b), osc.scaled_solution(rp.solve, b, s))
    # forward job: S b, then the permuting forward sweep; backward: x then S
    assert np.array_equal(rs.solve_job(b, 1), rp.solve_job(b * s[:, None], 1))
    assert np.array_equal(rs.solve_job(b, 3), rp.solve_job(b, 3) * s[:, None])
