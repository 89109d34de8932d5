"""Seeded synthetic matrices of the five benchmark workloads.

Every generator returns a lower-triangular CSC matrix ``(n, col_ptr,
row_idx, values)`` with int32 indices, rows sorted inside each column and
the diagonal first, the input format of ``smf_analyse``
(solver/include/smf/smf_c.h:139-154 of the reference: 0-based lower CSC).

Workloads:
  * ``laplace7_3d(m)``        SPD 7-point Laplacian on an m^3 grid, diagonal 6,
                              off-diagonal -1; the same matrix, entry order
                              included, as the reference's ``grid_3d``
                              (solver/tests/test_parallel_determinism.cpp:341-361).
  * ``kkt_grid(g, m, seed)``  indefinite KKT [[H, A^T], [A, -1e-8 I]], H the
                              2D 5-point Laplacian on a g x g grid plus a
                              log-uniform diagonal Sigma in [1e-6, 1e6], A an
                              m x g^2 matrix with 8 distinct random columns
                              per row and N(0,1) values.
  * ``ocp_kkt(T, seed)``      trajectory-optimisation KKT with 12 states and 4
                              controls per stage (n = 28 T + 12).
  * ``spd27_3d(m, seed)``     SPD 27-point operator on an m^3 grid,
                              off-diagonals -U(0.5, 1.5), diagonal = |row sum|
                              + 1e-3.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "laplace7_3d",
    "spd27_3d",
    "kkt_grid",
    "kkt_sigma_sequence",
    "kkt_grid_refactor_values",
    "ocp_kkt",
    "coo_lower_to_csc",
    "csc_to_dense",
    "grid_coords_3d",
]


def coo_lower_to_csc(n, rows, cols, vals):
    """Lower COO (row >= col, no duplicates) -> sorted lower CSC (int32)."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64)
    if rows.size and np.any(rows < cols):
        raise ValueError("entries above the diagonal")
    order = np.lexsort((rows, cols))
    rows = rows[order]
    cols = cols[order]
    vals = vals[order]
    counts = np.bincount(cols, minlength=n)
    col_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=col_ptr[1:])
    if col_ptr[-1] >= 2**31:
        raise ValueError("more than 2^31 - 1 entries")
    return n, col_ptr.astype(np.int32), rows.astype(np.int32), vals


def csc_to_dense(n, col_ptr, row_idx, values):
    """Full symmetric dense matrix of a lower CSC (small n only)."""
    A = np.zeros((n, n))
    for j in range(n):
        for k in range(col_ptr[j], col_ptr[j + 1]):
            i = row_idx[k]
            A[i, j] += values[k]
            if i != j:
                A[j, i] += values[k]
    return A


def laplace7_3d(m: int, shift: float = 0.0):
    """7-point Laplacian on an m^3 grid (diag 6 - shift, off -1), lower CSC in
    the entry order of the reference's grid_3d: for each node j = (z*m+y)*m+x,
    the diagonal, then +x, +y, +z neighbours."""
    n = m * m * m
    j = np.arange(n, dtype=np.int64)
    x = j % m
    y = (j // m) % m
    z = j // (m * m)
    has = np.stack([np.ones(n, bool), x + 1 < m, y + 1 < m, z + 1 < m])
    off = np.array([0, 1, m, m * m], dtype=np.int64)
    val = np.array([6.0 - shift, -1.0, -1.0, -1.0])
    counts = has.sum(axis=0)
    col_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=col_ptr[1:])
    # entries column by column in the order diag, +x, +y, +z
    sel = has.T  # n x 4
    rows = (j[:, None] + off[None, :])[sel]
    vals = np.broadcast_to(val, (n, 4))[sel]
    return n, col_ptr.astype(np.int32), rows.astype(np.int32), vals.astype(np.float64)


def grid_coords_3d(m: int):
    j = np.arange(m * m * m, dtype=np.int64)
    return np.stack([j % m, (j // m) % m, j // (m * m)], axis=1)


def spd27_3d(m: int, seed: int = 0):
    """27-point SPD operator on an m^3 grid: every off-diagonal -U(0.5, 1.5)
    (one draw per unordered pair), diagonal = sum of |off-diagonals| of its
    row + 1e-3 (strictly diagonally dominant, hence SPD)."""
    rng = np.random.default_rng(seed)
    n = m * m * m
    j = np.arange(n, dtype=np.int64)
    x = j % m
    y = (j // m) % m
    z = j // (m * m)
    # the 13 "forward" neighbour offsets (i > j in lexicographic (z, y, x))
    offs = []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                d = (dz * m + dy) * m + dx
                if (dz, dy, dx) > (0, 0, 0):
                    offs.append((dx, dy, dz, d))
    rows = [j]
    cols = [j]
    vals = [None]
    rsum = np.zeros(n)
    for dx, dy, dz, d in offs:
        ok = (x + dx >= 0) & (x + dx < m) & (y + dy >= 0) & (y + dy < m) & (z + dz >= 0) & (z + dz < m)
        c = j[ok]
        r = c + d
        v = -rng.uniform(0.5, 1.5, size=c.size)
        rows.append(r)
        cols.append(c)
        vals.append(v)
        np.add.at(rsum, c, np.abs(v))
        np.add.at(rsum, r, np.abs(v))
    vals[0] = rsum + 1e-3
    return coo_lower_to_csc(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def _laplace5_2d_lower(g: int):
    n = g * g
    j = np.arange(n, dtype=np.int64)
    x = j % g
    y = j // g
    r1 = j[x + 1 < g]
    r2 = j[y + 1 < g]
    rows = np.concatenate([r1 + 1, r2 + g])
    cols = np.concatenate([r1, r2])
    return n, rows, cols


def kkt_grid(g: int, m: int, seed: int = 1, sigma=None, delta: float = 1e-8):
    """KKT [[H, A^T], [A, -delta I]], H = 5-point Laplacian (diag 4 + Sigma,
    off -1) on a g x g grid, Sigma_i = 10^U(-6, 6); A is m x g^2 with 8 distinct
    uniformly random columns per row and N(0,1) values.

    Returns (matrix, sigma) so a refactor sequence can reuse the pattern."""
    rng = np.random.default_rng(seed)
    nx = g * g
    if sigma is None:
        sigma = 10.0 ** rng.uniform(-6.0, 6.0, size=nx)
    # A: 8 distinct columns per row
    cols = np.empty((m, 8), dtype=np.int64)
    for i in range(m):
        cols[i] = rng.choice(nx, size=8, replace=False)
    avals = rng.standard_normal(size=(m, 8))
    _, hr, hc = _laplace5_2d_lower(g)
    d = np.arange(nx, dtype=np.int64)
    arow = (nx + np.repeat(np.arange(m, dtype=np.int64), 8))
    acol = cols.reshape(-1)
    n = nx + m
    rows = np.concatenate([d, hr, arow, nx + np.arange(m)])
    colsa = np.concatenate([d, hc, acol, nx + np.arange(m)])
    vals = np.concatenate([4.0 + sigma, -np.ones(hr.size), avals.reshape(-1), -delta * np.ones(m)])
    return coo_lower_to_csc(n, rows, colsa, vals), sigma


def kkt_sigma_sequence(sigma0, k: int, seed: int = 2):
    """Sigma of refactor k: log10 Sigma_{i,k} = log10 Sigma_{i,0} + k s_i,
    s_i ~ U(-0.3, 0.3), clipped to [1e-10, 1e10]."""
    rng = np.random.default_rng(seed)
    s = rng.uniform(-0.3, 0.3, size=np.asarray(sigma0).size)
    return np.clip(10.0 ** (np.log10(sigma0) + k * s), 1e-10, 1e10)


def kkt_grid_refactor_values(A, sigma, nx: int):
    """Values of a kkt_grid matrix A (same pattern and entry order) with the
    barrier diagonal Sigma replaced: H_ii = 4 + sigma_i for the first nx
    columns (the diagonal is each column's first entry)."""
    n, cp, ri, v = A
    first = cp[:nx]
    if not np.array_equal(ri[first], np.arange(nx)):
        raise ValueError("not a kkt_grid matrix: a column's first entry is not its diagonal")
    out = np.array(v, dtype=np.float64, copy=True)
    out[first] = 4.0 + np.asarray(sigma, dtype=np.float64)
    return out


def ocp_kkt(T: int, seed: int = 0, reg: float = 1e-9):
    """Trajectory-optimisation KKT, stages t = 0..T-1 ordered
    [x_t (12), u_t (4), lambda_t (12)], then x_T (12); n = 28 T + 12.

    Stage Hessian of (x_t, u_t): M M^T + 0.1 I with M ~ N(0,1) 16 x 16 (x_T:
    12 x 12). Dynamics rows lambda_t: x_{t+1} - A_t x_t - B_t u_t with
    A_t = I + 0.1 N(0,1), B_t = 0.1 N(0,1); -reg on the lambda diagonal.
    Inertia (16 T + 12, 12 T, 0)."""
    rng = np.random.default_rng(seed)
    nx, nu = 12, 4
    stage = nx + nu + nx
    n = stage * T + nx
    R, C, V = [], [], []

    def add_block(r0, c0, B, lower_only):
        rr, cc = np.nonzero(np.ones_like(B, dtype=bool))
        vv = B[rr, cc]
        rr = rr + r0
        cc = cc + c0
        if lower_only:
            keep = rr >= cc
            rr, cc, vv = rr[keep], cc[keep], vv[keep]
        R.append(rr)
        C.append(cc)
        V.append(vv)

    for t in range(T):
        base = stage * t
        M = rng.standard_normal((nx + nu, nx + nu))
        H = M @ M.T + 0.1 * np.eye(nx + nu)
        add_block(base, base, H, True)
        A = np.eye(nx) + 0.1 * rng.standard_normal((nx, nx))
        B = 0.1 * rng.standard_normal((nx, nu))
        lam = base + nx + nu
        # constraint row lambda_t: -A x_t - B u_t + x_{t+1}
        add_block(lam, base, -A, False)
        add_block(lam, base + nx, -B, False)
        add_block(lam, lam, -reg * np.eye(nx), True)
        xn = stage * (t + 1)
        # entry (x_{t+1}, lambda_t) lies below the diagonal: row x_{t+1} > lambda_t
        add_block(xn, lam, np.eye(nx), False)
    M = rng.standard_normal((nx, nx))
    add_block(stage * T, stage * T, M @ M.T + 0.1 * np.eye(nx), True)
    rows = np.concatenate(R)
    cols = np.concatenate(C)
    vals = np.concatenate(V)
    nz = vals != 0.0
    keep = nz | (rows == cols)
    return coo_lower_to_csc(n, rows[keep], cols[keep], vals[keep])
