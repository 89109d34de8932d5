"""The oracle against the committed golden fixtures written by the compiled
reference (tests/golden/make_golden.py)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_mf(name):
    z = np.load(os.path.join(GOLD, f"mf_{name}.npz"))
    a = {k[3:]: z[k] for k in z.files if k.startswith("an_")}
    a["n"] = int(z["an_n"])
    A = (int(z["an_n"]), z["col_ptr"], z["row_idx"], z["values"])
    return z, a, A


def test_front_kernel_golden(oracle_lib):
    z = np.load(os.path.join(GOLD, "front_kernel.npz"))
    idx = 0
    while f"in{idx}" in z.files:
        f, ne, u = z[f"meta{idx}"]
        Fo, rows, tags, st = oracle_lib.ldlt_front(z[f"in{idx}"], int(ne), u=float(u), nb=16)
        s = z[f"stats{idx}"]
        assert np.array_equal(rows, z[f"rows{idx}"]) and np.array_equal(tags, z[f"tags{idx}"])
        assert [st["p"], st["pos"], st["neg"], st["zero"]] == [int(v) for v in s[:4]]
        assert st["flops"] == s[4]
        d = np.abs(np.tril(Fo) - z[f"out{idx}"]).max() / np.abs(z[f"out{idx}"]).max()
        assert d == 0.0 if f < 96 else d < 1e-10
        idx += 1
    assert idx >= 6


@pytest.mark.parametrize("name", ["laplace7_8", "ocp_T10", "kkt_12"])
def test_multifrontal_golden(oracle_lib, name):
    z, a, A = load_mf(name)
    posdef = bool(int(z["posdef"]))
    of = oracle_lib.Factor(a, A[3], posdef=posdef)
    eo = of.export()
    st = z["stats"]
    assert of.status == int(st[0])
    if not posdef:
        assert of.inertia == (int(st[1]), int(st[2]), int(st[3]))
        assert of.delayed == int(st[4]) and of.two_by_two == int(st[5])
        assert np.array_equal(eo["pivot_types"], z["pivot_types"])
    assert of.flops == st[6]
    assert np.array_equal(eo["steps"], z["steps"])
    assert np.array_equal(eo["row_indices"], z["row_indices"])
    d = np.abs(eo["factor_values"] - z["factor_values"]).max() / np.abs(z["factor_values"]).max()
    assert d < 1e-13
    x = of.solve(z["b"])
    assert np.abs(x - z["x_ref"]).max() < 1e-10
