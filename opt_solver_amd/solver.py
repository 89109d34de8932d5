"""Host-side mirror of smf's analyse / factor / solve split over the smfgpu
C ABI (include/smfgpu.h).

Mirrors ``smf::Solver`` (solver/include/smf/solver.hpp:60-127 of the
reference): ``analyse`` gives an analysis (the AnalysisKeep fields the
numeric phase reads), ``Plan`` is the per-analysis device set-up, ``factor``
runs the numeric factorization on the GPU and returns a ``Factor`` whose
``solve`` / ``refine`` run on the GPU as well. Status values are
smf::FactorStatus's; failures raise ``SmfGpuError``."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import (MEM_DEVICE, MEM_HOST, PREC_AUTO, PREC_BF16, PREC_BF16X3, PREC_FP32, PREC_FP64, SmfGpuError,
                   check, lib)

# smf::SolveJob (solver/include/smf/types.hpp:250-256)
SOLVE_FULL, SOLVE_FORWARD, SOLVE_DIAG, SOLVE_BACKWARD, SOLVE_DIAG_BACK = 0, 1, 2, 3, 4

__all__ = ["analyse", "Analysis", "Plan", "Factor", "SmfGpuError", "PREC_FP64", "PREC_FP32", "PREC_BF16",
           "PREC_BF16X3", "PREC_AUTO"]


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Analysis:
    """Result of the host symbolic analysis (``smfgpu_analyse``)."""

    def __init__(self, A, perm=None, nemin=32):
        n, cp, ri, _ = A
        self.n = int(n)
        cp = _i32(cp)
        ri = _i32(ri)
        pm = None if perm is None else _i32(perm)
        h = ctypes.c_void_p()
        check(lib().smfgpu_analyse(self.n, _p(cp), _p(ri), _p(pm), nemin, ctypes.byref(h)), "analyse")
        self.h = h
        st = np.zeros(6)
        check(lib().smfgpu_analysis_stats(self.h, _p(st)), "analysis_stats")
        self.nsup, self.front_rows, self.factor_entries, self.max_front, self.flops, self.depth = (
            int(st[0]), int(st[1]), int(st[2]), int(st[3]), float(st[4]), int(st[5]))
        self.desc = _lib.AnalysisDesc()
        check(lib().smfgpu_analysis_describe(self.h, ctypes.byref(self.desc)), "analysis_describe")

    def arrays(self):
        """Copies of the AnalysisKeep-style arrays."""
        d = self.desc
        ns, n = d.nsup, d.n

        def arr(ptr, count, dt=np.int32):
            if count == 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_int32)), shape=(count,)).copy()

        cp = arr(d.child_ptr, ns + 1)
        fp = arr(d.front_row_ptr, ns + 1)
        pc = arr(d.pcol_ptr, n + 1)
        return dict(n=n, sn_start=arr(d.sn_start, ns), sn_end=arr(d.sn_end, ns), sn_parent=arr(d.sn_parent, ns),
                    child_ptr=cp, child_idx=arr(d.child_idx, int(cp[-1])), front_row_ptr=fp,
                    front_row_idx=arr(d.front_row_idx, int(fp[-1])), pcol_ptr=pc,
                    prow_idx=arr(d.prow_idx, int(pc[-1])), orig_to_slot=arr(d.orig_to_slot, int(d.n_orig)),
                    perm=arr(d.perm, n), postorder=np.arange(ns, dtype=np.int32))

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().smfgpu_analysis_free(ctypes.byref(self.h))
            except Exception:  # interpreter shutdown: the library is going away anyway
                pass
            self.h = None


def analyse(A, perm=None, nemin=32) -> Analysis:
    return Analysis(A, perm=perm, nemin=nemin)


class Plan:
    """Device plan of one analysis (``smfgpu_plan_create``)."""

    def __init__(self, analysis, posdef=True):
        if isinstance(analysis, Analysis):
            self._owner = analysis
            desc = analysis.desc
            self.n = analysis.n
        else:  # dict of AnalysisKeep arrays (e.g. exported from smf's own analysis)
            a = {k: (np.ascontiguousarray(v, np.int32) if isinstance(v, np.ndarray) else v)
                 for k, v in analysis.items()}
            self._owner = a
            desc = _lib.AnalysisDesc()
            desc.n = int(a["n"])
            desc.nsup = int(a["sn_start"].size)
            for k in ("sn_start", "sn_end", "sn_parent", "child_ptr", "child_idx", "front_row_ptr",
                      "front_row_idx", "pcol_ptr", "prow_idx", "orig_to_slot", "perm"):
                setattr(desc, k, _p(a[k]) if a[k].size else None)
            desc.n_orig = int(a["orig_to_slot"].size)
            self.n = desc.n
        self._desc = desc
        self.posdef = posdef
        h = ctypes.c_void_p()
        check(lib().smfgpu_plan_create(ctypes.byref(desc), 1 if posdef else 0, ctypes.byref(h)), "plan_create")
        self.h = h
        self.n_orig = int(desc.n_orig)
        self.nsup = int(desc.nsup)

    def stats(self):
        st = np.zeros(8)
        check(lib().smfgpu_plan_stats(self.h, _p(st)), "plan_stats")
        keys = ("device_bytes", "levels", "launches", "flops", "factor_entries", "cb_entries", "asm_bytes",
                "solve_bytes_per_rhs")
        return dict(zip(keys, st.tolist()))

    def set_timing(self, on=True):
        check(lib().smfgpu_set_timing(self.h, 1 if on else 0), "set_timing")

    def kernel_times(self):
        st = np.zeros(16)
        check(lib().smfgpu_kernel_times(self.h, _p(st)), "kernel_times")
        keys = ("syrk_ms", "syrk_launches", "syrk_flops", "asm_ms", "asm_bytes", "solve_ms", "solve_bytes",
                "fwd_ms", "bwd_ms", "fwd_launches", "bwd_launches", "device_levels", "dev_run_ms",
                "dev_readback_ms", "stats_ms", "solve_lists_ms")
        return dict(zip(keys, st[:16].tolist()))

    def launch_profile(self):
        """Per launch of the last timed SPD factorization: kind, level, items,
        Schur flops, ms (-1: not timed)."""
        n = int(lib().smfgpu_launch_profile(self.h, None, 0))
        out = np.zeros((max(n, 1), 5))
        lib().smfgpu_launch_profile(self.h, _p(out), n)
        return out[:n]

    def tune(self, key, value):
        check(lib().smfgpu_plan_tune(self.h, key, int(value)), "plan_tune")

    def selector_set(self, weights):
        """Load the learned precision selector's weights (SELECTOR_WEIGHTS
        floats) into the plan; PREC_AUTO factorizations use them."""
        w = np.ascontiguousarray(weights, np.float32)
        if w.size != _lib.SELECTOR_WEIGHTS:
            raise ValueError(f"expected {_lib.SELECTOR_WEIGHTS} weights, got {w.size}")
        check(lib().smfgpu_selector_set(self.h, _p(w)), "selector_set")
        self.selector_loaded = True

    def factor(self, values=None, values_ptr=None, precision=PREC_FP64, mixed_min_front=1024, u=0.01,
               small_pivot=1e-20, panel=16, max_negative_pivots=-1, work_limit=None, static_pivot_tol=0.0,
               fallback_static_pivot_tol=0.0, static_flags=0, selector_tol=1e-4, selector_mask=0, scaled=False,
               collect=False, scale=None, scale_ptr=None) -> "Factor":
        """scale (n host doubles) or scale_ptr (a device pointer to them): the
        symmetric scaling S of the original rows; the factor is then of S A S
        and its solves return A^-1 b (smf's ScalingMethod::UserSupplied)."""
        if precision == PREC_AUTO and not getattr(self, "selector_loaded", False):
            w = default_selector_weights()
            if w is not None:
                self.selector_set(w)
        opt, vp, keep = self._options(values, values_ptr, precision, mixed_min_front, u, small_pivot, panel,
                                      max_negative_pivots, work_limit, static_pivot_tol, fallback_static_pivot_tol,
                                      static_flags, selector_tol, selector_mask, scaled, collect)
        skeep = self._scale_option(opt, scale, scale_ptr)
        info = _lib.FactorInfo()
        h = ctypes.c_void_p()
        check(lib().smfgpu_factor(self.h, vp, self.n_orig, ctypes.byref(opt), ctypes.byref(h), ctypes.byref(info)),
              "factor")
        return Factor(self, h, info)

    def factor_solve(self, values, b, precision=PREC_FP64, mixed_min_front=1024, u=0.01, selector_tol=1e-4,
                     selector_mask=0, max_iters=10, tol=1e-14, max_steps=None, omega_target=None, method=None,
                     **kw):
        """smfgpu_factor_solve: factor, solve and refine (host arrays); a
        reduced-precision factor whose refinement stalls is replaced by the
        plain FP64 one. Returns (factor, x, refine info)."""
        if precision == PREC_AUTO and not getattr(self, "selector_loaded", False):
            w = default_selector_weights()
            if w is not None:
                self.selector_set(w)
        opt, vp, keep = self._options(values, None, precision, mixed_min_front, u, kw.get("small_pivot", 1e-20),
                                      kw.get("panel", 16), -1, None, kw.get("static_pivot_tol", 0.0),
                                      kw.get("fallback_static_pivot_tol", 0.0), kw.get("static_flags", 0),
                                      selector_tol, selector_mask, kw.get("scaled", False), False)
        skeep = self._scale_option(opt, kw.get("scale"), None)
        b = np.array(b, dtype=np.float64, order="F", copy=True)
        nrhs = 1 if b.ndim == 1 else b.shape[1]
        x = np.zeros_like(b)
        ro = Factor._refine_opts(max_iters, tol, method, max_steps, omega_target, 20)
        info = _lib.FactorInfo()
        rinf = _lib.RefineInfo()
        h = ctypes.c_void_p()
        check(lib().smfgpu_factor_solve(self.h, vp, self.n_orig, ctypes.byref(opt), _p(b), _p(x), nrhs, self.n,
                                        ctypes.byref(ro), MEM_HOST, ctypes.byref(h), ctypes.byref(info),
                                        ctypes.byref(rinf)), "factor_solve")
        rd = Factor._refine_dict(rinf)
        rd["fallback"] = bool(rinf.fallback)
        return Factor(self, h, info), x, rd

    def _scale_option(self, opt, scale, scale_ptr):
        if scale_ptr is not None:
            opt.scale = ctypes.c_void_p(scale_ptr)
            opt.scale_mem = MEM_DEVICE
            return None
        if scale is None:
            return None
        s = np.ascontiguousarray(scale, np.float64)
        if s.size != self.n:
            raise ValueError(f"expected {self.n} scale factors, got {s.size}")
        opt.scale = s.ctypes.data_as(ctypes.c_void_p)
        opt.scale_mem = MEM_HOST
        return s

    def equilibration_scale(self, values=None, values_ptr=None, iters=3):
        """smf's Equilibration scaling (compute_equilibration_scale,
        solver/src/scaling.cpp:57-101) of the given values, computed on the
        device: n factors in the original ordering."""
        s = np.zeros(self.n)
        if values_ptr is not None:
            vp, mem = ctypes.c_void_p(values_ptr), MEM_DEVICE
        else:
            v = np.ascontiguousarray(values, np.float64)
            if v.size != self.n_orig:
                raise ValueError(f"expected {self.n_orig} values, got {v.size}")
            vp, mem = _p(v), MEM_HOST
        check(lib().smfgpu_equilibration_scale(self.h, vp, self.n_orig, mem, int(iters), _p(s), MEM_HOST),
              "equilibration_scale")
        return s

    def _options(self, values, values_ptr, precision, mixed_min_front, u, small_pivot, panel, max_negative_pivots,
                 work_limit, static_pivot_tol, fallback_static_pivot_tol, static_flags, selector_tol, selector_mask,
                 scaled, collect):
        opt = _lib.FactorOptions()
        lib().smfgpu_default_options(ctypes.byref(opt))
        opt.posdef = 1 if self.posdef else 0
        opt.pivot_u = u
        opt.small_pivot = small_pivot
        opt.panel = panel
        opt.precision = precision
        opt.mixed_min_front = mixed_min_front
        opt.max_negative_pivots = max_negative_pivots
        opt.static_pivot_tol = static_pivot_tol
        opt.fallback_static_pivot_tol = fallback_static_pivot_tol
        opt.static_flags = static_flags
        opt.selector_tol = selector_tol
        opt.selector_mask = selector_mask
        opt.scaled = 1 if scaled else 0
        opt.collect = 1 if collect else 0
        if work_limit is not None:
            opt.work_limit_on = 1
            opt.work_limit = float(work_limit)
        if values_ptr is not None:
            opt.value_mem = MEM_DEVICE
            vp = ctypes.c_void_p(values_ptr)
            keep = None
        else:
            keep = np.ascontiguousarray(values, np.float64)
            if keep.size != self.n_orig:
                raise ValueError(f"expected {self.n_orig} values, got {keep.size}")
            opt.value_mem = MEM_HOST
            vp = _p(keep)
        return opt, vp, keep

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().smfgpu_plan_free(ctypes.byref(self.h))
            except Exception:  # interpreter shutdown: the library is going away anyway
                pass
            self.h = None


class Factor:
    def __init__(self, plan: Plan, h, info):
        self.plan = plan
        self.h = h
        self.info = info
        self.status = int(info.status)
        self.inertia = (int(info.num_positive), int(info.num_negative), int(info.num_zero))
        self.flops = float(info.actual_flops)
        self.seconds = float(info.factor_seconds)

    def solve(self, b, job=SOLVE_FULL):
        """job: SOLVE_FULL, SOLVE_FORWARD (result in the permuted ordering),
        SOLVE_DIAG (permuted in and out), SOLVE_BACKWARD and SOLVE_DIAG_BACK
        (permuted in, original out), as smf's SolveJob. b is n x nrhs, or
        ldb x nrhs with ldb > n (rows [n, ldb) are left as they are)."""
        b = np.array(b, dtype=np.float64, order="F", copy=True)
        nrhs = 1 if b.ndim == 1 else b.shape[1]
        ldb = b.shape[0]
        check(lib().smfgpu_solve(self.h, _p(b), nrhs, ldb, int(job), MEM_HOST), "solve")
        return b

    def precisions(self):
        """Per supernode: the precision code (PREC_*) of its Schur updates."""
        out = np.zeros(self.plan.nsup, np.int8)
        check(lib().smfgpu_factor_precisions(self.h, _p(out)), "factor_precisions")
        return out

    def selector_pairs(self):
        """Training pairs of a collect=True FP64 factorization: features X
        (k x 10), labels Y (k x 3: log10 relative Schur-update error in FP32,
        split BF16, BF16) and the fronts' supernodes."""
        k = lib().smfgpu_selector_pairs(self.h, None, None, None, 0)
        if k < 0:
            check(k, "selector_pairs")
        X = np.zeros((max(k, 1), _lib.SELECTOR_FEATURES), np.float32)
        Y = np.zeros((max(k, 1), _lib.SELECTOR_OUTPUTS), np.float32)
        fr = np.zeros(max(k, 1), np.int32)
        got = lib().smfgpu_selector_pairs(self.h, _p(X), _p(Y), _p(fr), k)
        if got < 0:
            check(got, "selector_pairs")
        return X[:got], Y[:got], fr[:got]

    def front_stats(self):
        """Per supernode: f, p, positive, negative, zero, 2x2, delayed, eligible."""
        out = np.zeros((self.plan.nsup, 8), np.int32)
        check(lib().smfgpu_factor_front_stats(self.h, _p(out)), "front_stats")
        return out

    def stall_stats(self):
        """Per supernode: stalled first-pass panels, partner pairs pulled,
        columns rotated, second-pass full searches (blocked fronts)."""
        out = np.zeros((self.plan.nsup, 4), np.int32)
        check(lib().smfgpu_factor_stall_stats(self.h, _p(out)), "stall_stats")
        return out

    def solve_device(self, ptr, nrhs, ldb=None, job=SOLVE_FULL):
        check(lib().smfgpu_solve(self.h, ctypes.c_void_p(ptr), nrhs, ldb or self.plan.n, int(job), MEM_DEVICE),
              "solve")

    def refine(self, b, x, max_iters=10, tol=1e-14, method=None, max_steps=None, omega_target=None,
               restart=20, values=None):
        """Refine x against A (the factored values, or `values`): smf's
        refine with Control::refine_method = method (default Auto),
        refinement_max_iters = max_iters, refine_max_steps = max_steps
        (default max_iters), targets eta_inf <= tol and omega <= omega_target
        (default: none). Returns (x, info)."""
        b = np.array(b, dtype=np.float64, order="F", copy=True)
        x = np.array(x, dtype=np.float64, order="F", copy=True)
        nrhs = 1 if b.ndim == 1 else b.shape[1]
        o = self._refine_opts(max_iters, tol, method, max_steps, omega_target, restart)
        keep = None
        if values is not None:
            keep = np.ascontiguousarray(values, np.float64)
            o.value_mem = MEM_HOST
        inf = _lib.RefineInfo()
        n = self.plan.n
        check(lib().smfgpu_refine_ex(self.h, _p(keep), _p(b), n, _p(x), n, nrhs, ctypes.byref(o), MEM_HOST,
                                     ctypes.byref(inf)), "refine")
        return x, self._refine_dict(inf)

    def refine_device(self, b_ptr, x_ptr, nrhs, max_iters=10, tol=1e-14, method=None, max_steps=None,
                      omega_target=None, restart=20):
        o = self._refine_opts(max_iters, tol, method, max_steps, omega_target, restart)
        inf = _lib.RefineInfo()
        n = self.plan.n
        check(lib().smfgpu_refine_ex(self.h, None, ctypes.c_void_p(b_ptr), n, ctypes.c_void_p(x_ptr), n, nrhs,
                                     ctypes.byref(o), MEM_DEVICE, ctypes.byref(inf)), "refine")
        return self._refine_dict(inf)

    @staticmethod
    def _refine_opts(max_iters, tol, method, max_steps, omega_target, restart):
        o = _lib.RefineOptions()
        lib().smfgpu_refine_default_options(ctypes.byref(o))
        o.method = _lib.REFINE_AUTO if method is None else int(method)
        o.max_iters = int(max_iters)
        o.max_steps = int(max_iters if max_steps is None else max_steps)
        o.restart = int(restart)
        o.eta_target = float(tol) if tol > 0 else 1e-300
        o.omega_target = 1.7e308 if omega_target is None else float(omega_target)
        return o

    @staticmethod
    def _refine_dict(inf):
        return dict(eta=float(inf.eta), omega=float(inf.omega), eta_start=float(inf.eta_start),
                    omega_start=float(inf.omega_start), steps=int(inf.steps),
                    method="fgmres" if inf.method_used == _lib.REFINE_FGMRES else "classical",
                    reached=bool(inf.reached), stalled=bool(inf.stalled))

    def export(self):
        """The FactorKeep layout: steps (row_offset, factor_offset, p, f per
        supernode), factor values, row indices, pivot types."""
        sz = np.zeros(4, np.int64)
        check(lib().smfgpu_export_factor(self.h, _p(sz), None, None, None, None), "export sizes")
        ns, nf, nr, n = (int(v) for v in sz)
        steps = np.zeros((ns, 4), np.int64)
        fv = np.zeros(nf)
        rows = np.zeros(nr, np.int32)
        pt = np.zeros(n, np.int8)
        check(lib().smfgpu_export_factor(self.h, _p(sz), _p(steps), _p(fv), _p(rows), _p(pt)), "export")
        return dict(steps=steps, factor_values=fv, row_indices=rows, pivot_types=pt)

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().smfgpu_factor_free(ctypes.byref(self.h))
            except Exception:  # interpreter shutdown: the library is going away anyway
                pass
            self.h = None


# ---- learned precision selector -------------------------------------------
import json as _json  # noqa: E402
import os as _os  # noqa: E402

SELECTOR_WEIGHTS_FILE = _os.path.join(_os.path.dirname(_os.path.abspath(__file__)), "selector_weights.json")


def default_selector_weights():
    """The weights trained by opt_solver_amd/train_selector.py, or None."""
    if not _os.path.exists(SELECTOR_WEIGHTS_FILE):
        return None
    with open(SELECTOR_WEIGHTS_FILE) as fh:
        return np.asarray(_json.load(fh)["weights"], np.float32)


def selector_init(seed=0):
    """Glorot-uniform start weights (biases 0, identity input scaling)."""
    rng = np.random.default_rng(seed)
    nf, nh, no = _lib.SELECTOR_FEATURES, 32, _lib.SELECTOR_OUTPUTS

    def glorot(o, i):
        a = np.sqrt(6.0 / (i + o))
        return rng.uniform(-a, a, (o, i))
    parts = [glorot(nh, nf).ravel(), np.zeros(nh), glorot(nh, nh).ravel(), np.zeros(nh), glorot(no, nh).ravel(),
             np.zeros(no), np.zeros(nf), np.ones(nf)]
    w = np.concatenate(parts).astype(np.float32)
    assert w.size == _lib.SELECTOR_WEIGHTS
    return w


def selector_train(X, Y, M=None, epochs=2000, lr=3e-3, weights=None, seed=0):
    """Adam training of the selector MLP on the device (the library's HIP
    kernels). Returns (weights, per-epoch loss)."""
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.float32)
    Mk = None if M is None else np.ascontiguousarray(M, np.float32)
    w = selector_init(seed) if weights is None else np.array(weights, np.float32, copy=True)
    loss = np.zeros(max(epochs, 1), np.float32)
    check(lib().smfgpu_selector_train(_p(X), _p(Y), _p(Mk), X.shape[0], int(epochs), float(lr), _p(w), _p(loss)),
          "selector_train")
    return w, loss[:epochs]


def selector_predict(weights, X):
    """The selector's inference kernel over feature rows: (n x 3) predicted
    log10 relative errors."""
    w = np.ascontiguousarray(weights, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    Y = np.zeros((X.shape[0], _lib.SELECTOR_OUTPUTS), np.float32)
    check(lib().smfgpu_selector_predict(_p(w), _p(X), X.shape[0], _p(Y)), "selector_predict")
    return Y
