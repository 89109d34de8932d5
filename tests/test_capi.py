"""libsmfgpu.so loads (without a GPU) and exports every symbol that
include/smfgpu.h declares; host-only entry points work without a device."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "smfgpu.h")).read()
    return sorted(set(re.findall(r"SMFGPU_API\s+[\w\s\*]+?\b(smfgpu_\w+)\s*\(", txt)))


def test_header_symbols_exported():
    from opt_solver_amd import _lib
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib.EXPORTED)


def test_error_paths_without_gpu():
    from opt_solver_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.smfgpu_analyse(3, None, None, None, 0, ctypes.byref(h))
    assert rc == -1 and b"bad matrix" in L.smfgpu_error_string()
    cp = np.array([0, 1, 2], np.int32)
    ri = np.array([0, 1], np.int32)
    bad = np.array([0, 0], np.int32)
    rc = L.smfgpu_analyse(2, cp.ctypes.data_as(ctypes.c_void_p), ri.ctypes.data_as(ctypes.c_void_p),
                          bad.ctypes.data_as(ctypes.c_void_p), 0, ctypes.byref(h))
    assert rc == -1 and b"permutation" in L.smfgpu_error_string()
    opt = _lib.FactorOptions()
    assert L.smfgpu_default_options(ctypes.byref(opt)) == 0
    assert opt.pivot_u == 0.01 and opt.small_pivot == 1e-20 and opt.panel == 16
