"""The reference's dense front-kernel cases on the HIP kernels: each front
as a one- or two-supernode sparse problem through smfgpu_factor, on the
small-front kernels and, for fronts of 96 rows and more, on the
one-workgroup blocked kernel and the multi-workgroup large-front path with
and without speculative panel groups. Every case is checked against the
oracle's front kernel (oracle.c, pinned to the reference by
tests/test_oracle.py and tests/test_golden.py) and, where the reference
states one, its known answer."""
import os

import numpy as np
import pytest

from tests.conftest import gpu_available
from tests.helpers import dense_front_problem, stored_columns

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from opt_solver_amd.solver import Plan  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODES = ["one_wg", "panel", "group"]


def _plan(arrays, mode):
    from opt_solver_amd._lib import TUNE_LDLT_BIG_MIN, TUNE_LDLT_GROUP
    plan = Plan(arrays, posdef=False)
    if mode in MODES:
        plan.tune(TUNE_LDLT_BIG_MIN, 1 << 30 if mode == "one_wg" else 96)
        plan.tune(TUNE_LDLT_GROUP, 1 if mode == "group" else 0)
    return plan


def run_front(oracle_lib, M, ne, u=0.01, mode="default", tol=None):
    """Factor M (ne eligible columns) on the GPU; check the first front
    against the oracle's front kernel: pivot order, tags, inertia, delays,
    stall counters, and the stored L and D (bitwise below the blocked size,
    else within tol, default 1e-12 of max|L|). Returns (gpu factor, oracle
    stats, oracle tags)."""
    M = np.asarray(M, dtype=np.float64)
    f = M.shape[0]
    A, arrays = dense_front_problem(M, ne)
    vmax = np.abs(A[3]).max()
    eff = 1e-20 * vmax if 0.0 < vmax < 1e-20 else 1e-20
    Fo, rows, tags, st = oracle_lib.ldlt_front(np.tril(M), ne, u=u, eff_small=eff, nb=16)
    # as the multifrontal store keeps a root: unaccepted eligible columns
    # become zero pivots
    root = ne == f
    p_store = ne if root else st["p"]
    tags = tags.copy()
    if root:
        tags[st["p"]:ne] = 0
    # "host": a limit that is never reached keeps the small fronts off the
    # device-level layout (the per-level path with its own small kernel)
    kw = dict(max_negative_pivots=f) if mode == "host" else {}
    plan = _plan(arrays, mode)
    fac = plan.factor(A[3], u=u, **kw)
    fs = fac.front_stats()[0]
    assert list(fs[:4]) == [f, st["p"], st["pos"], st["neg"]]
    assert fs[5] == int((tags[:p_store] == 2).sum())
    if not root:
        assert fs[4] == st["zero"] and fs[6] == ne - st["p"]
    ss = fac.stall_stats()[0]
    assert tuple(ss) == (st["stalls"], st["pulls"], st["rotated"], st["searches"]), (tuple(ss), st)
    ex = fac.export()
    ro, fo, p, ff = ex["steps"][0]
    assert ff == f and p == p_store
    assert np.array_equal(ex["row_indices"][:f], rows)
    assert np.array_equal(ex["pivot_types"][:p], tags[:p])
    Lg = ex["factor_values"][fo:fo + f * p].reshape(p, f).T
    Lo = stored_columns(Fo, tags, p_store)
    err = np.abs(Lg - Lo).max() / max(np.abs(Lo).max(), 1e-300) if p else 0.0
    if f < 96:
        assert err == 0.0, err
    else:
        assert err <= (1e-12 if tol is None else tol), err
    print(f"  f {f} ne {ne} u {u:g} mode {mode}: p {st['p']} inertia {st['pos']},{st['neg']},{st['zero']} "
          f"stalls {tuple(ss)} L/D err {err:.1e}", flush=True)
    return fac, st, tags


# --- known answers: test_inertia_dense.cpp:6-60, test_ldlt_dense.cpp:18-48
KATS = [
    ([[4, 2], [2, 3]], (2, 0, 0)),
    ([[-4, -2], [-2, -3]], (0, 2, 0)),
    ([[0, 1], [1, 0]], (1, 1, 0)),
    ([[2, 0], [0, 3]], (2, 0, 0)),
    ([[-2, 0], [0, -3]], (0, 2, 0)),
    ([[4, 2, 1], [2, 5, 2], [1, 2, 6]], (3, 0, 0)),
    ([[2, 1, 0], [1, 0, 1], [0, 1, 2]], (2, 1, 0)),
]


@pytest.mark.parametrize("mode", ["default", "host"])
@pytest.mark.parametrize("idx", range(len(KATS)))
def test_known_answers(oracle_lib, idx, mode):
    M, inertia = KATS[idx]
    print("case kat", idx, mode, flush=True)
    fac, st, _ = run_front(oracle_lib, M, len(M), mode=mode)
    assert fac.inertia == inertia
    ev = np.linalg.eigvalsh(np.array(M, float))
    assert (int((ev > 0).sum()), int((ev < 0).sum())) == inertia[:2]


def _tile(block, count, rng=None, coupling=0.0):
    """count copies of block on the diagonal (a front past the blocked
    size), optionally with weak random coupling between the copies."""
    b = np.asarray(block, float)
    k = b.shape[0]
    M = np.kron(np.eye(count), b)
    if coupling and rng is not None:
        C = coupling * rng.standard_normal(M.shape)
        C = np.tril(C, -k)
        M = M + C + C.T
    return M


@pytest.mark.parametrize("scale", [1e-100, 1e100])
@pytest.mark.parametrize("mode", ["default"] + MODES)
def test_extreme_scaling(oracle_lib, scale, mode):
    """A saddle [[0, s], [s, 0]] at s = 1e-100 and 1e100 (eff_small scales
    with max|A| below small_pivot, factor_indef.cpp:633-636): one 2x2 pivot
    each, inertia (1, 1, 0); as 64 coupled copies on the blocked paths."""
    print("case scaling", scale, mode, flush=True)
    S = [[0.0, scale], [scale, 0.0]]
    if mode == "default":
        fac, _, tags = run_front(oracle_lib, S, 2)
        assert fac.inertia == (1, 1, 0) and list(tags) == [2, -1]
        return
    M = _tile(S, 64, np.random.default_rng(3), 1e-3 * scale)
    fac, st, tags = run_front(oracle_lib, M, 128, mode=mode)
    ev = np.linalg.eigvalsh(M / scale)
    assert fac.inertia == (int((ev > 0).sum()), int((ev < 0).sum()), 0) == (64, 64, 0)


TINY_U = [[0.0, 1.0, 0.0], [1.0, 2e-6, 100.0], [0.0, 100.0, 1.0]]


@pytest.mark.parametrize("mode", ["default"] + MODES)
def test_tiny_u_prefers_2x2(oracle_lib, mode):
    """test_inertia_dense.cpp TinyUPrefers2x2OverNearNullPartner at
    u = 1e-8: the first pivot is the 2x2, inertia (2, 1, 0); 43 copies on
    the blocked paths."""
    print("case tiny u", mode, flush=True)
    if mode == "default":
        fac, _, tags = run_front(oracle_lib, TINY_U, 3, u=1e-8)
        assert tags[0] == 2 and fac.inertia == (2, 1, 0)
        return
    fac, st, tags = run_front(oracle_lib, _tile(TINY_U, 43), 129, u=1e-8, mode=mode)
    assert tags[0] == 2 and fac.inertia == (86, 43, 0)


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
    return A + np.tril(A, -1).T


@pytest.mark.parametrize("mode", ["default"] + MODES)
@pytest.mark.parametrize("d00,a21,u,first", [
    (5e-8, 0.2, 1e-8, 2),    # weak 1x1 gives way to the stable 2x2
    (5e-8, 3e7, 1e-8, 1),    # 2x2 growth larger: the weak 1x1 stays
    (0.05, 0.2, 0.01, 1),    # not weak
    (0.05, 0.2, 1e-8, 1),
    (5e-8, 0.2, 0.01, 2),    # u = 0.01: the weak diagonal fails outright
])
def test_weak_one_by_one_rule(oracle_lib, d00, a21, u, first, mode):
    """test_ldlt_dense.cpp:232-272: the weak-1x1 rule (|d| below 1e-6 of
    its column) against the 2x2 alternative, on the 4x4 front and as 32
    copies on the blocked paths."""
    print("case weak 1x1", d00, a21, u, mode, flush=True)
    W = _weak_front(d00, a21)
    if mode == "default":
        _, _, tags = run_front(oracle_lib, W, 4, u=u)
        assert tags[0] == first
        return
    _, _, tags = run_front(oracle_lib, _tile(W, 32), 128, u=u, mode=mode)
    assert tags[0] == first


def _stalled_front():
    """test_ldlt_dense.cpp:274-414 (as restated in tests/test_oracle.py):
    constraint columns with zero diagonal coupled only to far variable
    columns."""
    m, nv, q, kfar = 40, 120, 40, 100
    e = m + nv
    f = e + q
    rng = np.random.default_rng(5)
    A = np.zeros((f, f))
    B = rng.standard_normal((nv, nv))
    A[m:e, m:e] = B @ B.T + nv * np.eye(nv)
    for c in range(m):
        v = m + ((c + kfar) % nv)
        A[v, c] = A[c, v] = 1.0 + rng.uniform()
    A[e:, :e] = 0.01 * rng.standard_normal((q, e))
    A[:e, e:] = A[e:, :e].T
    A[e:, e:] = np.eye(q)
    return A, e


@pytest.mark.parametrize("mode", MODES)
def test_stalled_panels_pull_partners(oracle_lib, mode):
    """Stalled first-pass panels pull far partners (pulls > 0) and every
    eligible column is accepted with the block's inertia."""
    print("case stalled", mode, flush=True)
    A, e = _stalled_front()
    fac, st, _ = run_front(oracle_lib, A, e, mode=mode)
    ss = fac.stall_stats()[0]
    assert ss[0] > 0 and ss[1] > 0
    assert st["p"] == e
    ev = np.linalg.eigvalsh(A[:e, :e])
    assert (st["pos"], st["neg"], st["zero"]) == (int((ev > 0).sum()), int((ev < 0).sum()), 0)


def _rotation_front(z):
    """Eligible columns z with no entry inside the eligible block (zero
    diagonal, coupled only to the contribution rows): no partner exists, so
    their panels rotate them to the end and the second pass delays them
    after a full search. The other pivots are diagonally dominant (|d| in
    [30, 40]), so L stays within a few tens and the blocked paths' other
    summation order keeps L and D within 1e-12 of the oracle's."""
    f, ne = 200, 150
    rng = np.random.default_rng(3)
    M = rng.standard_normal((f, f))
    M = M + M.T
    M[np.arange(f), np.arange(f)] = rng.uniform(30, 40, f) * np.sign(rng.standard_normal(f))
    for c in z:
        M[c, :ne] = 0.0
        M[:ne, c] = 0.0
    return M, ne


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("zi", [0, 1])
def test_rotation_and_second_pass(oracle_lib, zi, mode):
    print("case rotation", zi, mode, flush=True)
    z = [range(5, 13), list(range(0, 150, 7))][zi]
    M, ne = _rotation_front(z)
    fac, st, _ = run_front(oracle_lib, M, ne, mode=mode)
    ss = fac.stall_stats()[0]
    assert ss[2] > 0 and ss[3] > 0
    assert st["p"] < ne and fac.info.delayed_pivots >= ne - st["p"]


def _golden_fronts():
    z = np.load(os.path.join(GOLD, "front_kernel.npz"))
    out = []
    idx = 0
    while f"in{idx}" in z.files:
        f, ne, u = z[f"meta{idx}"]
        out.append((idx, int(f), int(ne), float(u)))
        idx += 1
    return out


GOLDEN = _golden_fronts()


@pytest.mark.parametrize("mode", ["default"] + MODES)
@pytest.mark.parametrize("idx", [g[0] for g in GOLDEN])
def test_front_kernel_golden(oracle_lib, idx, mode):
    """tests/golden/front_kernel.npz (written by the compiled reference's
    ldlt_factor_front): pivot order, tags, p and inertia equal, L and D
    bitwise below the blocked size. Above it the reference sums its deferred
    updates through BLAS, so the oracle already differs from it by rounding
    (1.2e-12 of max|L| on the 120-row front); the GPU stays within 1e-12 of
    the oracle (run_front) and so within that distance plus 1e-12 of the
    reference."""
    _, f, ne, u = GOLDEN[idx]
    if mode != "default" and f < 96:
        pytest.skip("below the blocked size: one kernel")
    print("case golden front", idx, f, ne, u, mode, flush=True)
    z = np.load(os.path.join(GOLD, "front_kernel.npz"))
    Fin = np.asarray(z[f"in{idx}"], float)
    M = np.tril(Fin) + np.tril(Fin, -1).T
    fac, st, tags = run_front(oracle_lib, M, ne, u=u, mode=mode)
    s = z[f"stats{idx}"]
    assert [st["p"], st["pos"], st["neg"], st["zero"]] == [int(v) for v in s[:4]]
    ex = fac.export()
    ro, fo, p, ff = ex["steps"][0]
    assert np.array_equal(ex["row_indices"][:f], z[f"rows{idx}"])
    gt = np.array(z[f"tags{idx}"], np.int8)
    if ne == f:
        gt[int(s[0]):ne] = 0
    assert np.array_equal(ex["pivot_types"][:p], gt[:p])
    Lg = ex["factor_values"][fo:fo + f * p].reshape(p, f).T
    Lr = stored_columns(z[f"out{idx}"], gt, p)
    scale = max(np.abs(Lr).max(), 1e-300)
    err = np.abs(Lg - Lr).max() / scale
    Fo, _, _, _ = oracle_lib.ldlt_front(Fin, ne, u=u, nb=16)
    err_orc = np.abs(stored_columns(Fo, gt, p) - Lr).max() / scale
    print(f"  vs reference: GPU {err:.1e}, oracle {err_orc:.1e}", flush=True)
    if f < 96:
        assert err == 0.0 and err_orc == 0.0, (err, err_orc)
    else:
        assert err <= err_orc + 1e-12, (err, err_orc)
