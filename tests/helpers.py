"""Shared test helpers (dense reconstruction, matrix-vector products)."""
import numpy as np


def full_dense(A):
    from opt_solver_amd.generators import csc_to_dense
    return csc_to_dense(*A)


def matvec(A, x):
    n, cp, ri, v = A
    cols = np.repeat(np.arange(n), np.diff(cp))
    y = np.zeros(n) if x.ndim == 1 else np.zeros((n, x.shape[1]))
    np.add.at(y, ri, (v * x[cols].T).T if x.ndim > 1 else v * x[cols])
    off = ri != cols
    np.add.at(y, cols[off], (v[off] * x[ri[off]].T).T if x.ndim > 1 else v[off] * x[ri[off]])
    return y


def reconstruct_front(Fout, tags, p):
    """(L, D) of the first p columns from a factored front in the kernel's
    output layout (L below the pivot blocks, D on the diagonal, 2x2
    off-diagonal at (c+1, c))."""
    f = Fout.shape[0]
    L = np.eye(f)
    D = np.zeros((p, p))
    c = 0
    while c < p:
        if tags[c] == 2:
            D[c, c] = Fout[c, c]
            D[c + 1, c] = D[c, c + 1] = Fout[c + 1, c]
            D[c + 1, c + 1] = Fout[c + 1, c + 1]
            L[c + 2:, c] = Fout[c + 2:, c]
            L[c + 2:, c + 1] = Fout[c + 2:, c + 1]
            c += 2
        else:
            D[c, c] = Fout[c, c]
            L[c + 1:, c] = Fout[c + 1:, c]
            c += 1
    return L[:, :p], D


def dense_front_problem(M, ne):
    """A dense symmetric f x f matrix as a sparse problem whose first front
    is the whole matrix with ne eligible columns: one supernode (ne = f) or
    two (columns [0, ne), then [ne, f) as its parent). Returns the lower
    CSC tuple (every lower entry stored) and the analysis arrays (identity
    ordering and postorder)."""
    M = np.asarray(M, dtype=np.float64)
    f = M.shape[0]
    cp = np.zeros(f + 1, np.int32)
    cp[1:] = np.cumsum(np.arange(f, 0, -1))
    ri = np.concatenate([np.arange(j, f, dtype=np.int32) for j in range(f)])
    vals = np.concatenate([M[j:, j] for j in range(f)])
    two = ne < f
    i32 = lambda v: np.asarray(v, np.int32)  # noqa: E731
    arrays = dict(
        n=f, sn_start=i32([0, ne] if two else [0]), sn_end=i32([ne, f] if two else [f]),
        sn_parent=i32([1, -1] if two else [-1]), child_ptr=i32([0, 0, 1] if two else [0, 0]),
        child_idx=i32([0] if two else []), front_row_ptr=i32([0, f, 2 * f - ne] if two else [0, f]),
        front_row_idx=np.concatenate([np.arange(f), np.arange(ne, f)]).astype(np.int32) if two
        else np.arange(f, dtype=np.int32),
        pcol_ptr=cp, prow_idx=ri, orig_to_slot=np.arange(ri.size, dtype=np.int32),
        perm=np.arange(f, dtype=np.int32), postorder=np.arange(2 if two else 1, dtype=np.int32))
    return (f, cp, ri, vals), arrays


def stored_columns(Fo, tags, p_store):
    """The first p_store columns of a factored front as the factor store
    keeps them (multifrontal_serial.cpp's store step): L below the pivot
    blocks, D on the diagonal, a 2x2 pivot's off-diagonal entry moved above
    the diagonal of its second column, unaccepted (tag 0) columns zero."""
    L = np.tril(np.asarray(Fo))[:, :p_store].copy()
    for c in range(p_store):
        if tags[c] == 0:
            L[:, c] = 0.0
        elif tags[c] == 2:
            d10 = L[c + 1, c]
            L[c + 1, c] = 0.0
            L[c, c + 1] = d10
    return L
