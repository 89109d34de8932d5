"""ctypes access to oracle/_ref/libsmfref.so: the reference smf library built
from its own sources (oracle/Makefile.ref) plus oracle/ref_shim.cpp.

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker. ``available()`` is False when
the library was not built (no reference sources on this machine)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libsmfref.so")
_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_L = ctypes.c_int64


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_PATH)
        L.ref_analyse.restype = _P
        L.ref_analyse.argtypes = [_I, _P, _P, _P, _I, _P, _I, _P]
        L.ref_analysis_sizes.argtypes = [_P, _P]
        L.ref_analysis_export.argtypes = [_P] + [_P] * 12
        L.ref_analysis_free.argtypes = [_P]
        L.ref_factor.restype = _P
        L.ref_factor.argtypes = [_P, _P, _L, _D, _I, _P]
        L.ref_factor_ex.restype = _P
        L.ref_factor_ex.argtypes = [_P, _P, _L, _D, _I, _D, _D, _I, _P]
        L.ref_factor_scaled.restype = _P
        L.ref_factor_scaled.argtypes = [_P, _P, _L, _D, _I, _P, _P, _P]
        L.ref_factor_limits.restype = _P
        L.ref_factor_limits.argtypes = [_P, _D, _I, _D, _P]
        L.ref_factor_sizes.argtypes = [_P, _P]
        L.ref_factor_export.argtypes = [_P, _P, _P, _P, _P, _P]
        L.ref_pivot_types_size.restype = _L
        L.ref_pivot_types_size.argtypes = [_P]
        L.ref_solve.restype = _I
        L.ref_solve.argtypes = [_P, _P, _I, _I, _P]
        L.ref_solve_job.restype = _I
        L.ref_solve_job.argtypes = [_P, _P, _I, _I, _I]
        L.ref_factor_free.argtypes = [_P]
        L.ref_ldlt_front.restype = _I
        L.ref_ldlt_front.argtypes = [_P, _I, _I, _P, _D, _D, _D, _I, _I, _D, _P, _P]
        L.ref_posdef_front.restype = _I
        L.ref_posdef_front.argtypes = [_P, _I, _I, _P, _P]
        _lib = L
    return _lib


ORDER_AMD = 2
ORDER_METIS = 3
ORDER_USER = 4


class RefAnalysis:
    """The reference's AnalysisKeep, exported as plain arrays."""

    def __init__(self, A, ordering=ORDER_METIS, perm=None, posdef=False):
        n, cp, ri, vals = A
        self.n = n
        self._keep = (np.ascontiguousarray(cp, np.int32), np.ascontiguousarray(ri, np.int32),
                      np.ascontiguousarray(vals, np.float64))
        up = None if perm is None else np.ascontiguousarray(perm, np.int32)
        sec = np.zeros(1)
        self.h = lib().ref_analyse(n, _ptr(self._keep[0]), _ptr(self._keep[1]), _ptr(self._keep[2]),
                                   ordering, _ptr(up), 1 if posdef else 0, _ptr(sec))
        if not self.h:
            raise RuntimeError("reference analyse failed")
        self.seconds = float(sec[0])
        self.posdef = posdef
        sz = np.zeros(8, np.int64)
        lib().ref_analysis_sizes(self.h, _ptr(sz))
        n_, ns, nch, nfr, pnnz, onnz = (int(v) for v in sz[:6])
        self.nsup = ns
        self.sn_start = np.zeros(ns, np.int32)
        self.sn_end = np.zeros(ns, np.int32)
        self.sn_parent = np.zeros(ns, np.int32)
        self.child_ptr = np.zeros(ns + 1, np.int32)
        self.child_idx = np.zeros(max(nch, 1), np.int32)[:nch]
        self.front_row_ptr = np.zeros(ns + 1, np.int32)
        self.front_row_idx = np.zeros(nfr, np.int32)
        self.postorder = np.zeros(ns, np.int32)
        self.perm = np.zeros(n, np.int32)
        self.pcol_ptr = np.zeros(n + 1, np.int32)
        self.prow_idx = np.zeros(pnnz, np.int32)
        self.orig_to_slot = np.zeros(onnz, np.int32)
        lib().ref_analysis_export(self.h, _ptr(self.sn_start), _ptr(self.sn_end), _ptr(self.sn_parent),
                                  _ptr(self.child_ptr), _ptr(self.child_idx) if nch else None,
                                  _ptr(self.front_row_ptr), _ptr(self.front_row_idx),
                                  _ptr(self.postorder), _ptr(self.perm), _ptr(self.pcol_ptr),
                                  _ptr(self.prow_idx), _ptr(self.orig_to_slot))

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_analysis_free(self.h)
            self.h = None

    def plan_arrays(self):
        """The fields a numeric phase reads (AnalysisKeep, analysis.hpp:25-144)."""
        return dict(n=self.n, sn_start=self.sn_start, sn_end=self.sn_end, sn_parent=self.sn_parent,
                    child_ptr=self.child_ptr, child_idx=self.child_idx,
                    front_row_ptr=self.front_row_ptr, front_row_idx=self.front_row_idx,
                    postorder=self.postorder, perm=self.perm, pcol_ptr=self.pcol_ptr,
                    prow_idx=self.prow_idx, orig_to_slot=self.orig_to_slot)


class RefFactor:
    STATUS = {0: "Success", 1: "NotPositiveDefinite", 2: "Singular", 3: "MaxPivotDelays", 4: "InertiaExceeded"}

    def __init__(self, an: RefAnalysis, values, u=0.01, threads=1, static_pivot_tol=0.0,
                 fallback_static_pivot_tol=0.0, static_flags=0, max_negative_pivots=None, work_limit=None,
                 scale=None, equilibrate=False):
        """max_negative_pivots / work_limit: the early stops, on the values the
        analysis was built from (values must be those). scale (n doubles):
        ScalingMethod::UserSupplied; equilibrate: ScalingMethod::Equilibration;
        self.scale then holds the factor's scale."""
        self.an = an
        vals = np.ascontiguousarray(values, np.float64)
        st = np.zeros(16)
        self.scale = None
        if scale is not None or equilibrate:
            us = None if scale is None else np.ascontiguousarray(scale, np.float64)
            out = np.zeros(an.n)
            self.h = lib().ref_factor_scaled(an.h, _ptr(vals), vals.size, u, 1 if us is not None else 2,
                                             None if us is None else _ptr(us), _ptr(out), _ptr(st))
            self.scale = out if int(st[13]) == an.n else None
        elif max_negative_pivots is not None or work_limit is not None:
            self.h = lib().ref_factor_limits(an.h, u, -1 if max_negative_pivots is None else int(max_negative_pivots),
                                             -1.0 if work_limit is None else float(work_limit), _ptr(st))
        else:
            self.h = lib().ref_factor_ex(an.h, _ptr(vals), vals.size, u, threads, static_pivot_tol,
                                         fallback_static_pivot_tol, static_flags, _ptr(st))
        self.pivot_growth = float(st[12])
        self.status = int(st[0])
        self.error = int(st[1])
        self.inertia = (int(st[2]), int(st[3]), int(st[4]))
        self.delayed = int(st[5])
        self.two_by_two = int(st[6])
        self.flops = float(st[7])
        self.front_work = float(st[8])
        self.seconds = float(st[9])
        self.static_pivots = int(st[10])
        self.max_front = int(st[11])

    def export(self):
        sz = np.zeros(4, np.int64)
        lib().ref_factor_sizes(self.h, _ptr(sz))
        ns, nf, nr = int(sz[0]), int(sz[1]), int(sz[2])
        steps = np.zeros((ns, 4), np.int64)
        fv = np.zeros(nf)
        rows = np.zeros(nr, np.int32)
        npt = int(lib().ref_pivot_types_size(self.h))
        pt = np.zeros(max(npt, 1), np.int8)
        ssn = np.zeros(ns, np.int32)
        lib().ref_factor_export(self.h, _ptr(steps), _ptr(fv), _ptr(rows), _ptr(pt), _ptr(ssn))
        return dict(steps=steps, factor_values=fv, row_indices=rows, pivot_types=pt[:npt],
                    step_supernode=ssn)

    def solve(self, b):
        b = np.array(b, dtype=np.float64, order="F", copy=True)
        nrhs = 1 if b.ndim == 1 else b.shape[1]
        sec = np.zeros(1)
        rc = lib().ref_solve(self.h, _ptr(b), self.an.n, nrhs, _ptr(sec))
        if rc != 0:
            raise RuntimeError(f"reference solve returned {rc}")
        self.solve_seconds = float(sec[0])
        return b

    def solve_job(self, b, job):
        """The reference's partial solve jobs (SolveJob 1..4) on a copy of b."""
        b = np.array(b, dtype=np.float64, order="F", copy=True)
        nrhs = 1 if b.ndim == 1 else b.shape[1]
        rc = lib().ref_solve_job(self.h, _ptr(b), self.an.n, nrhs, int(job))
        if rc != 0:
            raise RuntimeError(f"reference solve job {job} returned {rc}")
        return b

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_factor_free(self.h)
            self.h = None


def ldlt_front(F, num_elig, u=0.01, eff_small=1e-20, small_pivot=1e-20, nb=16, posdef=False,
               static_pivot=0.0, rows=None):
    """The reference's dense front kernel (ldlt_factor_front) on a copy of F
    (f x f, lower triangle read). Returns (F, rows, tags, stats dict)."""
    f = F.shape[0]
    Fc = np.array(F, dtype=np.float64, order="F", copy=True)
    r = np.arange(f, dtype=np.int32) if rows is None else np.array(rows, dtype=np.int32, copy=True)
    tags = np.zeros(f, np.int8)
    st = np.zeros(6)
    lib().ref_ldlt_front(_ptr(Fc), f, num_elig, _ptr(r), u, eff_small, small_pivot, nb,
                         1 if posdef else 0, static_pivot, _ptr(tags), _ptr(st))
    return Fc, r, tags, dict(p=int(st[0]), pos=int(st[1]), neg=int(st[2]), zero=int(st[3]),
                             flops=float(st[4]), static=int(st[5]))
