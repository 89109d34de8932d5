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


def test_struct_layout_matches_header(tmp_path):
    """The ctypes mirrors of the option, info and analysis structs have the
    header's size and field offsets (compiled by the C compiler here)."""
    import subprocess
    from opt_solver_amd import _lib
    structs = {"smfgpu_factor_options": _lib.FactorOptions, "smfgpu_factor_info": _lib.FactorInfo,
               "smfgpu_analysis_desc": _lib.AnalysisDesc}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "smfgpu.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'    printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        cname, what, val = line.split()
        py = structs[cname]
        got = ctypes.sizeof(py) if what == "size" else getattr(py, what).offset
        assert got == int(val), (cname, what, got, val)
