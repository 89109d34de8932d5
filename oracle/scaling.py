"""Numpy restatement of smf's symmetric scaling, the checker of the device
scaling (test infrastructure: only tests/ use it).

A matrix is (n, col_ptr, row_idx, values): the lower triangle (row >= col)
in CSC, as smf's cleaned matrix. Every product is taken in the reference's
order, so results agree with it to the bit (numpy's float64 multiply,
divide and sqrt are IEEE correctly rounded, as the host's are).

Pinned by tests/test_scaling.py: against the reference built from its
sources (oracle/_ref: Equilibration and UserSupplied factorizations), and by
closed forms (a diagonal matrix's first pass gives 1/sqrt|a_ii|; one pass on
a matrix whose rows' largest entries are already 1 leaves s = 1).
"""
from __future__ import annotations

import numpy as np


def _rows_cols(n, col_ptr, row_idx):
    cols = np.repeat(np.arange(n, dtype=np.int64), np.diff(np.asarray(col_ptr, np.int64)))
    return np.asarray(row_idx, np.int64), cols


def scale_values(n, col_ptr, row_idx, values, s):
    """S A S entry by entry: a_rc * (s_r * s_c) (factor_indef.cpp:530-546;
    apply_scale_to_matrix, scaling.cpp:93-101)."""
    r, c = _rows_cols(n, col_ptr, row_idx)
    s = np.asarray(s, np.float64)
    return np.asarray(values, np.float64) * (s[r] * s[c])


def equilibration_scale(n, col_ptr, row_idx, values, iters=3):
    """compute_equilibration_scale (scaling.cpp:57-91): from s = 1, each pass
    takes row_max_i = max over the symmetric row i of |a| * s_r * s_c (in
    that order, with the pass's starting s), then s_i /= sqrt(row_max_i)
    where row_max_i > 0."""
    r, c = _rows_cols(n, col_ptr, row_idx)
    a = np.abs(np.asarray(values, np.float64))
    s = np.ones(n)
    for _ in range(iters):
        val = a * s[r] * s[c]
        rm = np.zeros(n)
        np.maximum.at(rm, r, val)
        np.maximum.at(rm, c, val)  # the diagonal's second visit changes no maximum
        pos = rm > 0.0
        s = s.copy()
        s[pos] = s[pos] / np.sqrt(rm[pos])
    return s


def scaled_solution(solve, b, s):
    """The full solve of a scaled factor (solve.cpp:1022-1047):
    x = S (S A S)^-1 (S b), with solve(y) = (S A S)^-1 y."""
    s = np.asarray(s, np.float64)
    b = np.asarray(b, np.float64)
    sb = b * (s if b.ndim == 1 else s[:, None])
    y = solve(sb)
    return y * (s if y.ndim == 1 else s[:, None])
