"""The library's host symbolic analysis against brute-force symbolic
factorization (every front holds its columns' exact structure; fundamental
supernodes with nemin=1 hold nothing more)."""
import numpy as np
import pytest

from opt_solver_amd import generators as g
from opt_solver_amd import ordering as o
from opt_solver_amd.solver import analyse


def symbolic_pattern(D, perm):
    B = D[np.ix_(perm, perm)] != 0
    n = B.shape[0]
    S = np.tril(B).copy()
    for k in range(n):
        rows = np.nonzero(S[k + 1:, k])[0] + k + 1
        for i in rows:
            S[rows[rows >= i], i] = True
    return S


@pytest.mark.parametrize("case", ["lap6_nd", "spd27_5", "spd27_5_nd", "ocp4", "kkt8"])
def test_fronts_cover_structure(case):
    if case == "lap6_nd":
        A, perm, nemin = g.laplace7_3d(6), o.grid_nd((6, 6, 6), leaf=8), 1
    elif case == "spd27_5":
        A, perm, nemin = g.spd27_3d(5), None, 1
    elif case == "spd27_5_nd":
        A, perm, nemin = g.spd27_3d(5), o.grid_nd((5, 5, 5), leaf=10), 8
    elif case == "ocp4":
        A, perm, nemin = g.ocp_kkt(4), None, 4
    else:
        (A, _), perm, nemin = g.kkt_grid(8, 20, seed=3), None, 4
    an = analyse(A, perm=perm, nemin=nemin)
    a = an.arrays()
    n = A[0]
    P = a["perm"]
    assert sorted(P.tolist()) == list(range(n))
    S = symbolic_pattern(g.csc_to_dense(*A), P)
    extra = 0
    for s in range(a["sn_start"].size):
        fr = a["front_row_idx"][a["front_row_ptr"][s]:a["front_row_ptr"][s + 1]]
        assert np.all(np.diff(fr) > 0)
        assert np.array_equal(fr[:a["sn_end"][s] - a["sn_start"][s]], np.arange(a["sn_start"][s], a["sn_end"][s]))
        par = a["sn_parent"][s]
        assert par == -1 or par > s
        for j in range(a["sn_start"][s], a["sn_end"][s]):
            need = set((np.nonzero(S[j:, j])[0] + j).tolist())
            have = set(fr[fr >= j].tolist())
            assert need <= have
            extra += len(have - need)
    if nemin == 1:
        assert extra == 0
    # the entry map sends every input entry to a slot holding its (row, col)
    for e in range(A[2].size):
        j = np.searchsorted(A[1], e, side="right") - 1
        i = A[2][e]
        sl = a["orig_to_slot"][e]
        ip = np.empty(n, np.int64)
        ip[P] = np.arange(n)
        r, c = max(ip[i], ip[j]), min(ip[i], ip[j])
        assert a["prow_idx"][sl] == r and a["pcol_ptr"][c] <= sl < a["pcol_ptr"][c + 1]


def test_geometric_nd_is_permutation():
    p = o.grid_nd((7, 5, 9), leaf=16)
    assert sorted(p.tolist()) == list(range(7 * 5 * 9))
    (A, _), = [g.kkt_grid(10, 25, seed=4)]
    rng = np.random.default_rng(4)
    cols = np.stack([rng.choice(100, 8, replace=False) for _ in range(25)])
    pk = o.kkt_grid_nd(10, 25, cols, leaf=16)
    assert sorted(pk.tolist()) == list(range(125))
