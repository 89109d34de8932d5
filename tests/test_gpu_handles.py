"""Handle lifetimes through the C ABI: factor handles that outlive their
plan (the adapter keeps the plan with the analysis and the factors apart,
so they are freed independently) report SMFGPU_ERR_STATE on every call and
free cleanly, whether they were the plan's live factor or superseded."""
import ctypes

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from opt_solver_amd import _lib  # noqa: E402
from opt_solver_amd import generators as g  # noqa: E402
from opt_solver_amd.solver import Plan, analyse  # noqa: E402

ERR_STATE = -5


@pytest.mark.parametrize("posdef", [True, False])
def test_factors_outlive_their_plan(posdef):
    print("case handles posdef", posdef, flush=True)
    L = _lib.lib()
    A = g.laplace7_3d(6) if posdef else g.ocp_kkt(12, seed=2)
    plan = Plan(analyse(A, nemin=8), posdef=posdef)
    old = plan.factor(A[3])
    live = plan.factor(A[3])
    assert live.status == 0
    b = np.ones(A[0])
    # the superseded handle is refused while the plan lives
    assert L.smfgpu_solve(old.h, b.ctypes.data_as(ctypes.c_void_p), 1, A[0], 0, _lib.MEM_HOST) == ERR_STATE
    h = plan.h
    L.smfgpu_plan_free(ctypes.byref(h))
    plan.h = None
    sizes = np.zeros(4, np.int64)
    out = np.zeros(8 * plan.nsup, np.int32)
    for f in (old, live):
        assert L.smfgpu_solve(f.h, b.ctypes.data_as(ctypes.c_void_p), 1, A[0], 0, _lib.MEM_HOST) == ERR_STATE
        assert L.smfgpu_solve_noalloc(f.h, b.ctypes.data_as(ctypes.c_void_p), 1, A[0], 0) == ERR_STATE
        assert L.smfgpu_export_factor(f.h, sizes.ctypes.data_as(ctypes.c_void_p), None, None, None, None) == ERR_STATE
        assert L.smfgpu_factor_front_stats(f.h, out.ctypes.data_as(ctypes.c_void_p)) == ERR_STATE
        assert b"plan was freed" in L.smfgpu_error_string()
        L.smfgpu_factor_free(ctypes.byref(f.h))
        assert not f.h
        f.h = None
