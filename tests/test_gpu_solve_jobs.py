"""Partial solve jobs (smf::SolveJob, solver/include/smf/types.hpp:250-256;
solve_core, solver/src/solve.cpp:1050-1068) through the C ABI against the
reference's own jobs on the same analysis: forward (result in the permuted
ordering), diagonal only (the reference's division and scaled 2x2 solve),
backward, diagonal then backward. The compositions give the full solve."""
import numpy as np
import pytest

from tests.conftest import gpu_available
from tests.helpers import matvec

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from opt_solver_amd import generators as g  # noqa: E402
from opt_solver_amd.solver import (SOLVE_BACKWARD, SOLVE_DIAG, SOLVE_DIAG_BACK, SOLVE_FORWARD,  # noqa: E402
                                   SOLVE_FULL, Plan)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("case", ["ocp_T60", "kkt_30", "lap3_12"])
@pytest.mark.parametrize("nrhs", [1, 3])
def test_solve_jobs_against_reference(case, nrhs):
    from oracle import refshim
    if not refshim.available():
        pytest.skip("reference library not built")
    print("case", case, "nrhs", nrhs, flush=True)
    posdef = case.startswith("lap3")
    if case.startswith("ocp"):
        A = g.ocp_kkt(int(case[5:]), seed=3)
        order = refshim.ORDER_AMD
    elif case.startswith("kkt"):
        gsz = int(case[4:])
        A, _ = g.kkt_grid(gsz, gsz * gsz // 5, seed=gsz)
        order = refshim.ORDER_METIS
    else:
        A = g.laplace7_3d(int(case[5:]))
        order = refshim.ORDER_METIS
    an = refshim.RefAnalysis(A, ordering=order, posdef=posdef)
    rf = refshim.RefFactor(an, A[3], u=0.01)
    plan = Plan(an.plan_arrays(), posdef=posdef)
    fac = plan.factor(A[3], u=0.01)
    assert fac.status == 0 and rf.status == 0
    rng = np.random.default_rng(5)
    b = rng.uniform(-1, 1, (A[0], nrhs)) if nrhs > 1 else rng.uniform(-1, 1, A[0])
    full = fac.solve(b, SOLVE_FULL)
    # each job on the same input as the reference's
    yg, yr = fac.solve(b, SOLVE_FORWARD), rf.solve_job(b, 1)
    e_fwd = _rel(yg, yr)
    dg, dr = fac.solve(yr, SOLVE_DIAG), rf.solve_job(yr, 2)
    e_diag = _rel(dg, dr)
    xg, xr = fac.solve(dr, SOLVE_BACKWARD), rf.solve_job(dr, 3)
    e_bwd = _rel(xg, xr)
    zg, zr = fac.solve(yr, SOLVE_DIAG_BACK), rf.solve_job(yr, 4)
    e_db = _rel(zg, zr)
    print(f"  forward {e_fwd:.1e} diag {e_diag:.1e} backward {e_bwd:.1e} diag+back {e_db:.1e}", flush=True)
    assert e_fwd <= 1e-12 and e_diag <= 1e-13 and e_bwd <= 1e-12 and e_db <= 1e-12
    if posdef:
        assert np.array_equal(dg, yr)  # D = I: the diagonal job copies
    # compositions of the device's own jobs
    comp1 = fac.solve(fac.solve(fac.solve(b, SOLVE_FORWARD), SOLVE_DIAG), SOLVE_BACKWARD)
    comp2 = fac.solve(fac.solve(b, SOLVE_FORWARD), SOLVE_DIAG_BACK)
    assert _rel(comp1, full) <= 1e-11 and _rel(comp2, full) <= 1e-12
    x1 = full if nrhs == 1 else full[:, 0]
    b1 = b if nrhs == 1 else b[:, 0]
    assert np.abs(matvec(A, x1) - b1).max() <= 1e-8 * max(1.0, np.abs(x1).max())


def test_solve_job_out_of_range():
    A = g.laplace7_3d(5)
    from opt_solver_amd.solver import SmfGpuError, analyse
    plan = Plan(analyse(A, nemin=16).arrays(), posdef=True)
    fac = plan.factor(A[3])
    with pytest.raises(SmfGpuError):
        fac.solve(np.ones(A[0]), 5)


@pytest.mark.parametrize("case", ["ocp_T60", "lap3_12"])
def test_leading_dimension(case):
    """b with ldb > n (solve.cpp's multi-RHS layout): every job returns the
    same columns as with ldb = n and leaves rows [n, ldb) untouched, for host
    arrays, device arrays and the solve that allocates nothing."""
    import ctypes

    import torch
    from opt_solver_amd._lib import lib
    from opt_solver_amd.solver import analyse
    print("case ldb", case, flush=True)
    posdef = case.startswith("lap3")
    A = g.laplace7_3d(12) if posdef else g.ocp_kkt(60, seed=3)
    n, nrhs, ldb = A[0], 3, A[0] + 37
    plan = Plan(analyse(A, nemin=16).arrays(), posdef=posdef)
    fac = plan.factor(A[3], u=0.01)
    assert fac.status == 0
    b = np.random.default_rng(9).uniform(-1, 1, (n, nrhs))
    pad = np.full((ldb, nrhs), 7.25)
    pad[:n] = b
    for job in (SOLVE_FULL, SOLVE_FORWARD, SOLVE_DIAG, SOLVE_BACKWARD, SOLVE_DIAG_BACK):
        ref = fac.solve(b, job)
        out = fac.solve(pad, job)
        assert np.all(out[n:] == 7.25), job
        assert _rel(out[:n], ref) <= 1e-13, job
        bd = torch.from_numpy(pad.reshape(-1, order="F").copy()).cuda()
        torch.cuda.synchronize()
        fac.solve_device(bd.data_ptr(), nrhs, ldb=ldb, job=job)
        od = bd.cpu().numpy().reshape((ldb, nrhs), order="F")
        assert np.all(od[n:] == 7.25), job
        assert _rel(od[:n], ref) <= 1e-13, job
        bd = torch.from_numpy(pad.reshape(-1, order="F").copy()).cuda()
        torch.cuda.synchronize()
        rc = lib().smfgpu_solve_noalloc(fac.h, ctypes.c_void_p(bd.data_ptr()), nrhs, ldb, int(job))
        assert rc == 0, lib().smfgpu_error_string()
        od = bd.cpu().numpy().reshape((ldb, nrhs), order="F")
        assert np.all(od[n:] == 7.25), job
        assert _rel(od[:n], ref) <= 1e-13, job
