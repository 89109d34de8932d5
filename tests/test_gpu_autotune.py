"""The autotuner on the device (a quick pass over small workloads): a
choice for every parameter among its candidates, the file it writes, and a
plan tuned with the choices factoring an indefinite KKT with the untuned
plan's pivots and inertia and a refined backward error at FP64 level."""
import json

import numpy as np
import pytest

from tests.conftest import gpu_available
from tests.helpers import matvec

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from opt_solver_amd import autotune as at  # noqa: E402
from opt_solver_amd import generators as g  # noqa: E402
from opt_solver_amd.solver import Plan, analyse  # noqa: E402


def test_quick_autotune_and_tuned_plan(oracle_lib, tmp_path):
    print("case autotune quick", flush=True)
    out = tmp_path / "tuned.json"
    res = at.autotune(quick=True, reps=1, log=lambda *a: print(*a, flush=True), out_path=str(out))
    saved = json.loads(out.read_text())
    assert saved["knobs"] == res["knobs"]
    for bucket, name, cands, default, used in at.KNOBS:
        assert res["knobs"][name] in cands, name
    assert len(res["trials"]) == sum(len(k[2]) for k in at.KNOBS)
    gsz = 60
    m = gsz * gsz // 5
    A, _ = g.kkt_grid(gsz, m, seed=gsz)
    arrays = analyse(A, nemin=16).arrays()
    plain = Plan(arrays, posdef=False).factor(A[3], u=0.01)
    tp = Plan(arrays, posdef=False)
    at.apply_tuned(tp, res, posdef=False)
    fac = tp.factor(A[3], u=0.01)
    assert fac.inertia == plain.inertia == (gsz * gsz, m, 0)
    ea, eb = fac.export(), plain.export()
    for key in ("steps", "row_indices", "pivot_types"):
        assert np.array_equal(ea[key], eb[key]), key
    x = np.random.default_rng(4).uniform(-1, 1, A[0])
    b = matvec(A, x)
    eta, _ = oracle_lib.backward_error(A, b, fac.solve(b))
    assert eta <= 1e-13, eta
