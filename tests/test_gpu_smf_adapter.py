"""smf's own C API (solver/src/smf_c.cpp, on the CPU) against the same calls
routed through the GPU adapter smf would add (tests/smf_adapter, the code of
INTEGRATION.md compiled against the reference's headers and objects by
oracle/Makefile.ref and linked to libsmfgpu): configuration 1 (the 7-point
Laplacian on a 20^3 grid, Cholesky) and two grid KKTs of configuration 2's
generator, one and three right-hand sides, plus a refactorization on the
adapter's cached plan."""
import json
import os
import subprocess

import pytest

from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                   "smf_adapter_check")


@pytest.mark.skipif(not gpu_available(), reason="no GPU")
def test_smf_c_api_through_adapter():
    if not os.path.exists(BIN):
        pytest.skip("adapter check not built (needs the reference sources at build time)")
    print("case smf C API through the adapter", flush=True)
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    rows = [json.loads(line) for line in p.stdout.splitlines() if line.startswith("{")]
    for r in rows:
        print(" ", r, flush=True)
    assert len(rows) == 3, p.stderr[-2000:]
    for r in rows:
        assert r["rc_cpu"] == 0 and r["rc_gpu"] == 0, r
        assert r["eta_gpu"] <= 1e-14 and r["eta_cpu"] <= 1e-13, r
        assert r["rel_diff"] <= 1e-8 and r["refactor_eta"] <= 1e-14 and r["refactor_rel_diff"] <= 1e-10, r
        assert r["ok"], r
    assert rows[0]["inertia_gpu"] == [8000, 0, 0]
    assert rows[1]["inertia_gpu"] == rows[1]["inertia_cpu"] == [900, 180, 0]
    assert rows[2]["inertia_gpu"] == rows[2]["inertia_cpu"] == [3600, 1152, 0]
    assert p.returncode == 0
