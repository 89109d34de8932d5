"""Batched independent systems as one forest (opt_solver_amd.batch): the
union's factor is the systems' factors side by side. Checked with the oracle
on the CPU: totals are the sums of the per-system results, and every
supernode's stored block equals the one of its own system's factor."""
import numpy as np
import pytest

from opt_solver_amd import batch
from opt_solver_amd.solver import analyse
from tests.helpers import matvec


def _systems(nsys, tmin, tmax, seed):
    mats, Ts = batch.ocp_batch(nsys, tmin, tmax, seed=seed)
    arrs = [analyse(A, nemin=16).arrays() for A in mats]
    return mats, Ts, arrs


def test_union_offsets_and_shapes():
    mats, Ts, arrs = _systems(5, 10, 30, seed=11)
    U, offs = batch.union_arrays(arrs)
    assert U["n"] == sum(A[0] for A in mats) == offs["col"][-1]
    assert U["sn_start"].size == offs["sn"][-1] == sum(a["sn_start"].size for a in arrs)
    assert U["orig_to_slot"].size == offs["val"][-1] == sum(A[3].size for A in mats)
    # each system's columns are a contiguous range covered by its supernodes
    for i in range(len(mats)):
        s0, s1 = offs["sn"][i], offs["sn"][i + 1]
        assert U["sn_start"][s0] == offs["col"][i] and U["sn_end"][s1 - 1] == offs["col"][i + 1]
        par = U["sn_parent"][s0:s1]
        assert np.all((par == -1) | ((par > np.arange(s0, s1)) & (par < s1)))


def test_union_factor_is_per_system_factor(oracle_lib):
    mats, Ts, arrs = _systems(6, 10, 25, seed=7)
    U, offs = batch.union_arrays(arrs)
    vals = np.concatenate([A[3] for A in mats])
    fu = oracle_lib.Factor(U, vals, posdef=False)
    singles = [oracle_lib.Factor(a, A[3], posdef=False) for a, A in zip(arrs, mats)]
    assert fu.inertia == tuple(sum(f.inertia[k] for f in singles) for k in range(3))
    assert fu.delayed == sum(f.delayed for f in singles)
    assert fu.two_by_two == sum(f.two_by_two for f in singles)
    assert fu.flops == pytest.approx(sum(f.flops for f in singles), rel=0, abs=0)
    eu = fu.export()
    for i, (f, T) in enumerate(zip(singles, Ts)):
        assert f.inertia == (16 * T + 12, 12 * T, 0)
        e = f.export()
        for s in range(e["steps"].shape[0]):
            ro, fo, p, fr = e["steps"][s]
            rou, fou, pu, fru = eu["steps"][offs["sn"][i] + s]
            assert (p, fr) == (pu, fru)
            assert np.array_equal(e["factor_values"][fo:fo + fr * p], eu["factor_values"][fou:fou + fr * p])
            assert np.array_equal(e["row_indices"][ro:ro + fr] + offs["col"][i], eu["row_indices"][rou:rou + fr])
    # solve of the union = the systems' solves
    X = [np.random.default_rng(i).uniform(-1, 1, A[0]) for i, A in enumerate(mats)]
    b = np.concatenate([matvec(A, x) for A, x in zip(mats, X)])
    xu = fu.solve(b)
    assert np.abs(xu - np.concatenate(X)).max() < 1e-9
