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
