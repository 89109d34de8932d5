"""The oracle (oracle/oracle.c) pinned against the reference's known-answer
tests, closed forms and, when it is built, the reference itself."""
import numpy as np
import pytest

from tests.helpers import reconstruct_front


def _lower(M):
    return np.tril(np.asarray(M, dtype=np.float64))


# --- known answers of solver/tests/test_inertia_dense.cpp:6-60 and
# test_ldlt_dense.cpp:18-48 (restated on the front kernel) ---------------
@pytest.mark.parametrize("A,inertia", [
    ([[4, 2], [2, 3]], (2, 0, 0)),                       # LDLT.SPD2x2
    ([[-4, -2], [-2, -3]], (0, 2, 0)),                    # LDLT.AllNegative2x2
    ([[0, 1], [1, 0]], (1, 1, 0)),                        # AddInertia2x2_Saddle
    ([[2, 0], [0, 3]], (2, 0, 0)),                        # BothPos
    ([[-2, 0], [0, -3]], (0, 2, 0)),                      # BothNeg
    ([[4, 2, 1], [2, 5, 2], [1, 2, 6]], (3, 0, 0)),       # SPD3x3_Inertia
    ([[2, 1, 0], [1, 0, 1], [0, 1, 2]], (2, 1, 0)),       # LDLT.Indefinite3x3 (det -4)
])
def test_inertia_known_answers(oracle_lib, A, inertia):
    A = np.array(A, dtype=float)
    _, _, tags, st = oracle_lib.ldlt_front(_lower(A), A.shape[0], u=0.01)
    assert st["p"] == A.shape[0]
    assert (st["pos"], st["neg"], st["zero"]) == inertia
    # Sylvester: the same counts as the eigenvalues
    ev = np.linalg.eigvalsh(A)
    assert (int((ev > 0).sum()), int((ev < 0).sum())) == inertia[:2]


@pytest.mark.parametrize("scale", [1e-100, 1e100])
def test_inertia_extreme_scaling(oracle_lib, scale):
    A = np.array([[0.0, scale], [scale, 0.0]])
    # eff_small as factor_indef sets it (factor_indef.cpp:633-636): small *
    # max|A| when max|A| < small
    eff = 1e-20 * scale if scale < 1e-20 else 1e-20
    _, _, tags, st = oracle_lib.ldlt_front(_lower(A), 2, u=0.01, eff_small=eff)
    assert (st["pos"], st["neg"], st["zero"]) == (1, 1, 0)
    assert list(tags) == [2, -1]


def test_tiny_u_prefers_2x2_over_near_null_partner(oracle_lib):
    # test_inertia_dense.cpp: TinyUPrefers2x2OverNearNullPartner
    A = np.array([[0.0, 1.0, 0.0], [1.0, 2e-6, 100.0], [0.0, 100.0, 1.0]])
    _, _, tags, st = oracle_lib.ldlt_front(_lower(A), 3, u=1e-8)
    assert tags[0] == 2
    assert (st["pos"], st["neg"], st["zero"]) == (2, 1, 0)


def _weak_front(d00, a21):
    A = np.zeros((4, 4))
    A[0, 0] = d00
    A[1, 0] = 1.0
    A[2, 0] = 0.1
    A[3, 0] = 0.05
    A[1, 1] = 1e-3
    A[2, 1] = a21
    A[3, 1] = 0.1
    A[2, 2] = 4.0
    A[3, 2] = 0.3
    A[3, 3] = -5.0
    return A


@pytest.mark.parametrize("d00,a21,u,first", [
    (5e-8, 0.2, 1e-8, 2),    # weak 1x1 gives way to the stable 2x2
    (5e-8, 3e7, 1e-8, 1),    # 2x2 growth larger: the weak 1x1 stays
    (0.05, 0.2, 0.01, 1),    # not weak
    (0.05, 0.2, 1e-8, 1),
    (5e-8, 0.2, 0.01, 2),    # u = 0.01: the weak diagonal fails outright
])
def test_weak_one_by_one_rule(oracle_lib, d00, a21, u, first):
    # test_ldlt_dense.cpp:232-272
    _, _, tags, _ = oracle_lib.ldlt_front(_weak_front(d00, a21), 4, u=u)
    assert tags[0] == first


def _random_front(rng, f, diag_scale=0.3):
    M = rng.standard_normal((f, f))
    M = M + M.T
    M[np.arange(f), np.arange(f)] *= rng.uniform(0, diag_scale, f)
    return M


@pytest.mark.parametrize("f,ne,nb", [(12, 12, 16), (30, 20, 16), (80, 50, 16), (130, 90, 16), (160, 160, 8)])
def test_front_reconstruction_and_schur(oracle_lib, f, ne, nb):
    """P^T A P = L D L^T on the eliminated block and the Schur complement of
    the rest equals A22 - A21 A11^-1 A12 (dense LAPACK): closed forms that a
    dropped term, wrong sign or transposed operand fails."""
    rng = np.random.default_rng(f * 7 + ne)
    A = _random_front(rng, f)
    Fo, rows, tags, st = oracle_lib.ldlt_front(_lower(A), ne, nb=nb)
    p = st["p"]
    P = A[np.ix_(rows, rows)]
    L, D = reconstruct_front(Fo, tags, p)
    # eliminated columns
    R = L @ D @ L.T
    scale = np.abs(A).max()
    assert np.abs((P - R)[:, :p]).max() / scale < 1e-11
    # Schur complement
    if p < f:
        S = P[p:, p:] - P[p:, :p] @ np.linalg.solve(P[:p, :p], P[:p, p:])
        Sk = np.tril(Fo[p:, p:])
        assert np.abs(np.tril(S) - Sk).max() / max(np.abs(S).max(), 1.0) < 1e-9
    # inertia of the eliminated block (Sylvester)
    ev = np.linalg.eigvalsh(P[:p, :p])
    assert (st["pos"], st["neg"]) == (int((ev > 0).sum()), int((ev < 0).sum()))


def test_stalled_panel_pulls_partners(oracle_lib):
    """test_ldlt_dense.cpp:274-414: constraint columns with zero diagonal
    coupled only to far variable columns; every eligible column is accepted
    with the block's inertia."""
    m, nv, q, kfar = 40, 120, 40, 100
    e = m + nv
    f = e + q
    rng = np.random.default_rng(5)
    A = np.zeros((f, f))
    B = rng.standard_normal((nv, nv))
    A[m:e, m:e] = B @ B.T + nv * np.eye(nv)
    for c in range(m):
        v = m + (c + kfar - m) % nv if c + kfar >= m else m + c
        v = m + ((c + kfar) % nv)
        A[v, c] = A[c, v] = 1.0 + rng.uniform()
    A[e:, :e] = 0.01 * rng.standard_normal((q, e))
    A[:e, e:] = A[e:, :e].T
    A[e:, e:] = np.eye(q)
    _, _, tags, st = oracle_lib.ldlt_front(_lower(A), e, nb=16)
    assert st["p"] == e
    ev = np.linalg.eigvalsh(A[:e, :e])
    assert (st["pos"], st["neg"], st["zero"]) == (int((ev > 0).sum()), int((ev < 0).sum()), 0)


def test_stall_counters_by_construction(oracle_lib):
    """The blocked path's counters against what a front's structure fixes:
    a diagonally dominant front never stalls; in the far-partner front every
    stalled panel is rescued by pulls (no rotation, no second-pass search);
    columns with no entry inside the eligible block have no partner, so
    each is rotated and finally delayed after a full search."""
    rng = np.random.default_rng(11)
    D = _random_front(rng, 160) + 400.0 * np.eye(160)
    _, _, _, st = oracle_lib.ldlt_front(_lower(D), 160, nb=16)
    assert (st["stalls"], st["pulls"], st["rotated"], st["searches"]) == (0, 0, 0, 0)
    m, nv, q, kfar = 40, 120, 40, 100
    e = m + nv
    rng = np.random.default_rng(5)
    A = np.zeros((e + q, e + q))
    B = rng.standard_normal((nv, nv))
    A[m:e, m:e] = B @ B.T + nv * np.eye(nv)
    for c in range(m):
        v = m + ((c + kfar) % nv)
        A[v, c] = A[c, v] = 1.0 + rng.uniform()
    A[e:, e:] = np.eye(q)
    _, _, _, st = oracle_lib.ldlt_front(_lower(A), e, nb=16)
    assert st["p"] == e and st["stalls"] > 0 and st["pulls"] > 0
    assert st["rotated"] == 0 and st["searches"] == 0
    f, ne = 200, 150
    rng = np.random.default_rng(3)
    M = rng.standard_normal((f, f))
    M = M + M.T
    M[np.arange(f), np.arange(f)] = rng.uniform(5, 10, f)
    z = list(range(0, ne, 7))
    for c in z:
        M[c, :ne] = 0.0
        M[:ne, c] = 0.0
    _, rows, _, st = oracle_lib.ldlt_front(_lower(M), ne, nb=16)
    assert st["p"] == ne - len(z) and sorted(rows[st["p"]:ne]) == z
    assert st["pulls"] == 0 and st["rotated"] >= len(z) and st["searches"] >= 1


@pytest.mark.parametrize("f,ne", [(20, 12), (60, 60), (130, 90), (200, 150)])
def test_dense_front_problem(oracle_lib, f, ne):
    """tests.helpers.dense_front_problem: the multifrontal oracle on the
    one- or two-supernode problem stores exactly what the front kernel
    leaves in the first front (rows, pivot types, L and D bitwise)."""
    from tests.helpers import dense_front_problem, stored_columns
    M = _random_front(np.random.default_rng(f), f)
    A, arrays = dense_front_problem(M, ne)
    ex = oracle_lib.Factor(arrays, A[3], posdef=False).export()
    Fo, rows, tags, st = oracle_lib.ldlt_front(_lower(M), ne, nb=16)
    ro, fo, p, ff = ex["steps"][0]
    p_store = ne if ne == f else st["p"]
    tags[st["p"]:p_store] = 0
    assert ff == f and p == p_store
    assert np.array_equal(ex["row_indices"][:f], rows)
    assert np.array_equal(ex["pivot_types"][:p], tags[:p])
    L = ex["factor_values"][fo:fo + f * p].reshape(p, f).T
    assert np.array_equal(L, stored_columns(Fo, tags, p_store))


def test_backward_error_closed_form(oracle_lib):
    # A = [[2, 1], [1, 3]] (lower CSC), x = [1, 1], b = A x + [1e-3, 0]
    A = (2, np.array([0, 2, 3], np.int32), np.array([0, 1, 1], np.int32), np.array([2.0, 1.0, 3.0]))
    x = np.array([1.0, 1.0])
    b = np.array([3.0 + 1e-3, 4.0])
    eta, om = oracle_lib.backward_error(A, b, x)
    # ||r|| = 1e-3, ||A||_inf = 4, ||x|| = 1, ||b|| = max(3.001, 4) = 4
    assert abs(eta - 1e-3 / (4 * 1 + 4)) < 1e-15
    # omega = max_i |r_i| / ((|A||x|)_i + |b_i|) = 1e-3 / (3 + 3.001)
    assert abs(om - 1e-3 / (3 + 3.001)) < 1e-15


# --- the oracle against the compiled reference -------------------------
@pytest.mark.parametrize("f,ne", [(20, 12), (60, 60), (90, 70), (140, 100), (200, 150)])
def test_front_kernel_matches_reference(oracle_lib, ref_lib, f, ne):
    rng = np.random.default_rng(f + ne)
    for _ in range(4):
        A = _random_front(rng, f)
        Fr, rr, tr, sr = ref_lib.ldlt_front(_lower(A), ne, nb=16)
        Fo, ro, to, so = oracle_lib.ldlt_front(_lower(A), ne, nb=16)
        assert np.array_equal(rr, ro) and np.array_equal(tr, to) and sr["p"] == so["p"]
        assert (sr["pos"], sr["neg"], sr["zero"]) == (so["pos"], so["neg"], so["zero"])
        assert sr["flops"] == so["flops"]
        d = np.abs(np.tril(Fr) - np.tril(Fo)).max() / np.abs(Fr).max()
        if f < 96:
            assert d == 0.0  # the unblocked path is bitwise the reference's
        else:
            assert d < 1e-10


def _cases(g, R):
    return [
        ("laplace7_12", g.laplace7_3d(12), True, R.ORDER_METIS),
        ("ocp_T10", g.ocp_kkt(10, seed=0), False, R.ORDER_AMD),
        ("ocp_T40", g.ocp_kkt(40, seed=3), False, R.ORDER_AMD),
        ("kkt_20", g.kkt_grid(20, 60, seed=1)[0], False, R.ORDER_METIS),
    ]


def test_multifrontal_matches_reference(oracle_lib, ref_lib):
    from opt_solver_amd import generators as g
    for name, A, posdef, order in _cases(g, ref_lib):
        an = ref_lib.RefAnalysis(A, ordering=order, posdef=posdef)
        rf = ref_lib.RefFactor(an, A[3])
        of = oracle_lib.Factor(an.plan_arrays(), A[3], posdef=posdef)
        ex, eo = rf.export(), of.export()
        assert rf.status == of.status == 0, name
        assert np.array_equal(ex["steps"], eo["steps"]), name
        assert np.array_equal(ex["row_indices"], eo["row_indices"]), name
        if not posdef:
            assert np.array_equal(ex["pivot_types"], eo["pivot_types"]), name
            assert rf.inertia == of.inertia and rf.delayed == of.delayed and rf.two_by_two == of.two_by_two
        assert rf.flops == of.flops
        d = np.abs(ex["factor_values"] - eo["factor_values"]).max() / np.abs(ex["factor_values"]).max()
        assert d < 1e-12, (name, d)


def test_kkt_inertia_sylvester(oracle_lib):
    """[[H, A^T], [A, -dI]] with H SPD and A of full row rank has inertia
    (n_x, m, 0); the OCP KKT (16T+12, 12T, 0)."""
    from opt_solver_amd import generators as g
    from opt_solver_amd.solver import analyse  # host analysis only: no GPU needed
    cases = [(g.kkt_grid(16, 40, seed=2)[0], (256, 40, 0)), (g.ocp_kkt(12, seed=1), (16 * 12 + 12, 12 * 12, 0))]
    for A, inertia in cases:
        an = analyse(A, perm=None, nemin=8)
        of = oracle_lib.Factor(an.arrays(), A[3], posdef=False)
        assert of.inertia == inertia


def test_reference_solve_jobs_compose(ref_lib):
    """The reference's partial solve jobs (SolveJob 1..4), which the GPU job
    tests compare with: forward then diagonal-and-backward, and forward,
    diagonal, backward (division by d instead of the D^-1 table) give the
    full solve to rounding (the full solve of several right-hand sides takes
    its own route); for a Cholesky factor the diagonal job is the identity."""
    from opt_solver_amd import generators as g
    for name, A, posdef, order in _cases(g, ref_lib):
        an = ref_lib.RefAnalysis(A, ordering=order, posdef=posdef)
        rf = ref_lib.RefFactor(an, A[3])
        b = np.random.default_rng(2).uniform(-1, 1, (A[0], 2))
        full = rf.solve(b)
        y = rf.solve_job(b, 1)
        db = rf.solve_job(y, 4)
        assert np.abs(db - full).max() <= 1e-13 * np.abs(full).max(), name
        d = rf.solve_job(y, 2)
        if posdef:
            assert np.array_equal(d, y), name
        x = rf.solve_job(d, 3)
        assert np.abs(x - full).max() <= 1e-12 * np.abs(full).max(), name


def test_backward_error_against_exact_arithmetic(oracle_lib):
    """eta_inf and omega of the oracle (the reference's long-double residual,
    refinement.cpp:39-130) against the same formulas in exact rational
    arithmetic, at a backward error near the roundoff level where a residual
    evaluated in plain double would be off by tens of percent."""
    from fractions import Fraction
    from opt_solver_amd import generators as g
    A = g.spd27_3d(5, seed=5)
    n, cp, ri, v = A
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, n)
    cols = np.repeat(np.arange(n), np.diff(cp))
    vf = [Fraction(float(t)) for t in v]
    xf = [Fraction(float(t)) for t in x]
    ax = [Fraction(0)] * n
    absax = [Fraction(0)] * n
    rowabs = np.zeros(n)
    for k in range(len(v)):
        i, j = ri[k], cols[k]
        ax[i] += vf[k] * xf[j]
        absax[i] += abs(vf[k] * xf[j])
        rowabs[i] += abs(v[k])
        if i != j:
            ax[j] += vf[k] * xf[i]
            absax[j] += abs(vf[k] * xf[i])
            rowabs[j] += abs(v[k])
    b = np.array([float(t) for t in ax])
    b[::5] = np.nextafter(b[::5], np.inf)  # one-ulp perturbations
    r = [Fraction(float(b[i])) - ax[i] for i in range(n)]
    eta = float(max(abs(t) for t in r)) / (rowabs.max() * np.abs(x).max() + np.abs(b).max())
    om = max(float(abs(r[i]) / (absax[i] + abs(Fraction(float(b[i]))))) for i in range(n))
    e2, o2 = oracle_lib.backward_error(A, b, x)
    assert eta < 1e-15
    assert abs(e2 - eta) <= 1e-3 * eta and abs(o2 - om) <= 1e-3 * om
