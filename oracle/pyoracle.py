"""ctypes access to oracle/liboracle.so, the C restatement of smf's numeric
phase (oracle/oracle.c). Test infrastructure: only tests/, __graft_entry__
.smoke() and bench.py's CPU-baseline leg use it, as the checker."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None
_P = ctypes.c_void_p


class _Params(ctypes.Structure):
    _fields_ = [("u", ctypes.c_double), ("eff_small", ctypes.c_double), ("small_pivot", ctypes.c_double),
                ("nb", ctypes.c_int), ("posdef", ctypes.c_int)]


class _Result(ctypes.Structure):
    _fields_ = [("p", ctypes.c_int), ("pos", ctypes.c_int), ("neg", ctypes.c_int), ("zero", ctypes.c_int),
                ("flops", ctypes.c_double), ("stalls", ctypes.c_int), ("pulls", ctypes.c_int),
                ("rotated", ctypes.c_int), ("searches", ctypes.c_int)]


class _Analysis(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("nsup", ctypes.c_int32)] + \
               [(k, _P) for k in ("sn_start", "sn_end", "sn_parent", "child_ptr", "child_idx", "front_row_ptr",
                                  "front_row_idx", "pcol_ptr", "prow_idx", "orig_to_slot", "perm")] + \
               [("n_orig", ctypes.c_int64)]


class _Factor(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("nsup", ctypes.c_int32), ("status", ctypes.c_int32),
                ("pos", ctypes.c_int32), ("neg", ctypes.c_int32), ("zero", ctypes.c_int32),
                ("delayed", ctypes.c_int32), ("two_by_two", ctypes.c_int32), ("max_front", ctypes.c_int32),
                ("flops", ctypes.c_double), ("front_work", ctypes.c_double), ("max_abs", ctypes.c_double),
                ("steps", _P), ("L", _P), ("nL", ctypes.c_int64), ("nrows", ctypes.c_int64), ("rows", _P),
                ("tags", _P), ("ntags", ctypes.c_int64), ("posdef", ctypes.c_int32),
                ("stalls", ctypes.c_int32), ("pulls", ctypes.c_int32), ("rotated", ctypes.c_int32),
                ("searches", ctypes.c_int32)]


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_PATH)
        L.orc_ldlt_front.restype = ctypes.c_int
        L.orc_ldlt_front.argtypes = [_P, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P]
        L.orc_factor_run.restype = ctypes.c_int
        L.orc_factor_run.argtypes = [_P, _P, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, _P]
        L.orc_factor_free.argtypes = [_P]
        L.orc_solve.argtypes = [_P, _P, _P, ctypes.c_int]
        L.orc_backward_error.argtypes = [ctypes.c_int32, _P, _P, _P, _P, _P, ctypes.c_int, _P, _P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def ldlt_front(F, num_elig, u=0.01, eff_small=1e-20, small_pivot=1e-20, nb=16, posdef=False, rows=None):
    """Restated ldlt_factor_front on a copy of F; returns (F, rows, tags, stats)."""
    f = F.shape[0]
    Fc = np.array(F, dtype=np.float64, order="F", copy=True)
    r = np.arange(f, dtype=np.int32) if rows is None else np.array(rows, dtype=np.int32, copy=True)
    tags = np.zeros(max(f, 1), np.int8)
    prm = _Params(u, eff_small, small_pivot, nb, 1 if posdef else 0)
    res = _Result()
    lib().orc_ldlt_front(_p(Fc), f, num_elig, _p(r), ctypes.byref(prm), _p(tags), ctypes.byref(res))
    return Fc, r, tags[:f], dict(p=res.p, pos=res.pos, neg=res.neg, zero=res.zero, flops=res.flops,
                                 stalls=res.stalls, pulls=res.pulls, rotated=res.rotated, searches=res.searches)


class Factor:
    """Restated multifrontal factorization over an analysis (dict of arrays)."""

    def __init__(self, a, values, posdef, u=0.01, small_pivot=1e-20, nb=16):
        self._a = {k: np.ascontiguousarray(v, np.int32) if isinstance(v, np.ndarray) else v for k, v in a.items()}
        A = _Analysis()
        A.n = int(a["n"])
        A.nsup = int(self._a["sn_start"].size)
        for k in ("sn_start", "sn_end", "sn_parent", "child_ptr", "child_idx", "front_row_ptr", "front_row_idx",
                  "pcol_ptr", "prow_idx", "orig_to_slot", "perm"):
            setattr(A, k, _p(self._a[k]) if self._a[k].size else None)
        A.n_orig = int(self._a["orig_to_slot"].size)
        self._A = A
        self._vals = np.ascontiguousarray(values, np.float64)
        self.f = _Factor()
        self.posdef = posdef
        lib().orc_factor_run(ctypes.byref(A), _p(self._vals), u, small_pivot, nb, 1 if posdef else 0,
                             ctypes.byref(self.f))
        f = self.f
        self.status = int(f.status)
        self.inertia = (int(f.pos), int(f.neg), int(f.zero))
        self.delayed = int(f.delayed)
        self.two_by_two = int(f.two_by_two)
        self.flops = float(f.flops)
        self.max_front = int(f.max_front)
        self.stall_stats = (int(f.stalls), int(f.pulls), int(f.rotated), int(f.searches))

    def export(self):
        f = self.f
        ns = int(f.nsup)

        def arr(ptr, count, ct, dt):
            if count == 0 or not ptr:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), shape=(count,)).copy()

        steps = arr(f.steps, 4 * ns, ctypes.c_int64, np.int64).reshape(ns, 4)
        return dict(steps=steps, factor_values=arr(f.L, int(f.nL), ctypes.c_double, np.float64),
                    row_indices=arr(f.rows, int(f.nrows), ctypes.c_int32, np.int32),
                    pivot_types=arr(f.tags, int(f.ntags), ctypes.c_int8, np.int8))

    def solve(self, b):
        b = np.array(b, dtype=np.float64, order="F", copy=True)
        nrhs = 1 if b.ndim == 1 else b.shape[1]
        lib().orc_solve(ctypes.byref(self.f), _p(self._a["perm"]), _p(b), nrhs)
        return b

    def __del__(self):
        if getattr(self, "f", None) is not None and lib is not None:
            try:
                lib().orc_factor_free(ctypes.byref(self.f))
            except Exception:
                pass


def backward_error(A, b, x):
    """(eta_inf, omega) by the reference formula (refinement.cpp:39-130)."""
    n, cp, ri, v = A
    cp = np.ascontiguousarray(cp, np.int32)
    ri = np.ascontiguousarray(ri, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    b = np.asfortranarray(b, dtype=np.float64)
    x = np.asfortranarray(x, dtype=np.float64)
    nrhs = 1 if b.ndim == 1 else b.shape[1]
    eta = ctypes.c_double()
    om = ctypes.c_double()
    lib().orc_backward_error(int(n), _p(cp), _p(ri), _p(v), _p(b), _p(x), nrhs, ctypes.byref(eta), ctypes.byref(om))
    return float(eta.value), float(om.value)
