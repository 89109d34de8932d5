"""smfgpu_solve_noalloc (the device solve whose right-hand-side work block is
a fixed device array) against smfgpu_solve on device memory, through the C
ABI: two calls in a row, the second on a smaller system, for a Cholesky and
an LDL^T factor and every solve job; the no-allocation call leaves the plan's
device buffers as they were."""
import ctypes

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu
if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

import torch  # noqa: E402

from opt_solver_amd import _lib  # noqa: E402
from opt_solver_amd import generators as g  # noqa: E402
from opt_solver_amd import ordering as o  # noqa: E402
from opt_solver_amd.solver import Plan, analyse  # noqa: E402

# the forward sweeps add into ancestor rows with atomic adds, so two solves
# of the same factor agree to the last bits only up to summation order
TOL = 1e-13


def _system(kind, size):
    if kind == "spd":
        A = g.spd27_3d(size, seed=3)
        an = analyse(A, perm=o.grid_nd((size, size, size), leaf=16))
        return A, Plan(an, posdef=True)
    A, _ = g.kkt_grid(size, size * size // 3, seed=4)
    return A, Plan(analyse(A), posdef=False)


def _solve_pair(fac, B, job):
    """(x from smfgpu_solve_noalloc, x from smfgpu_solve), both on device
    copies of B (n x nrhs, column-major, ldb = n + 3)."""
    L = _lib.lib()
    n, nrhs = B.shape
    ldb = n + 3
    host = np.zeros((ldb, nrhs), order="F")
    host[:n] = B
    d1 = torch.from_numpy(host.reshape(-1, order="F").copy()).cuda()
    d2 = d1.clone()
    torch.cuda.synchronize()
    bytes0 = fac.plan.stats()["device_bytes"]
    _lib.check(L.smfgpu_solve_noalloc(fac.h, ctypes.c_void_p(d1.data_ptr()), nrhs, ldb, job), "solve_noalloc")
    assert fac.plan.stats()["device_bytes"] == bytes0  # nothing allocated
    _lib.check(L.smfgpu_solve(fac.h, ctypes.c_void_p(d2.data_ptr()), nrhs, ldb, job, _lib.MEM_DEVICE), "solve")
    torch.cuda.synchronize()
    x1 = d1.cpu().numpy().reshape((ldb, nrhs), order="F")
    x2 = d2.cpu().numpy().reshape((ldb, nrhs), order="F")
    # rows past n (the leading-dimension padding) are untouched
    assert np.array_equal(x1[n:], host[n:]) and np.array_equal(x2[n:], host[n:])
    return x1[:n], x2[:n]


@pytest.mark.parametrize("kind,sizes,nrhs", [("spd", (14, 9), (6, 1)), ("ldlt", (24, 12), (17, 2))])
def test_noalloc_matches_solve(kind, sizes, nrhs):
    for size, nr in zip(sizes, nrhs):  # second call: a smaller system, fewer right-hand sides
        print("case noalloc", kind, "size", size, "nrhs", nr, flush=True)
        A, plan = _system(kind, size)
        fac = plan.factor(A[3], u=0.01) if kind == "ldlt" else plan.factor(A[3])
        assert fac.status == 0
        n = A[0]
        rng = np.random.default_rng(size)
        B = rng.uniform(-1, 1, (n, nr))
        for job in (0, 1, 2, 3, 4):
            x1, x2 = _solve_pair(fac, B, job)
            scale = max(1.0, np.abs(x2).max())
            assert np.abs(x1 - x2).max() <= TOL * scale, (kind, size, job)
        # the full solve solves A x = b
        x1, _ = _solve_pair(fac, B, 0)
        from helpers import matvec
        r = matvec(A, x1) - B
        assert np.abs(r).max() <= 1e-10 * max(1.0, np.abs(x1).max())
        del fac, plan
