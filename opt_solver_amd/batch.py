"""Many independent systems in one factorization: the union of their
analyses is one forest, so every level's fronts of all systems go through
the same launches (the batched trajectory-optimisation workload: 1024 KKT
systems of uneven size). Each system keeps its own analysis (ordering,
supernodes, front rows); the union only renumbers them.

The reference factors such systems one call at a time
(``Solver::factor``, solver/include/smf/solver.hpp:88); the per-system
results here are those calls' results: inertia, delayed pivots and 2x2
pivots per system from the per-supernode statistics."""
from __future__ import annotations

import numpy as np

from . import generators as g

_FIELDS = ("sn_start", "sn_end", "sn_parent", "child_ptr", "child_idx", "front_row_ptr", "front_row_idx",
           "pcol_ptr", "prow_idx", "orig_to_slot", "perm")


def union_arrays(systems):
    """Concatenate AnalysisKeep-style array dicts (one per system) into one
    forest. Returns (arrays, offsets) with offsets = dict of per-system
    starts: 'sn' (supernodes), 'col' (columns), 'val' (input entries)."""
    out = {k: [] for k in _FIELDS}
    sn0 = col0 = val0 = slot0 = cp0 = fr0 = pc0 = 0
    offs = {"sn": [], "col": [], "val": []}
    for a in systems:
        n = int(a["n"])
        ns = int(np.asarray(a["sn_start"]).size)
        offs["sn"].append(sn0)
        offs["col"].append(col0)
        offs["val"].append(val0)
        out["sn_start"].append(np.asarray(a["sn_start"], np.int64) + col0)
        out["sn_end"].append(np.asarray(a["sn_end"], np.int64) + col0)
        par = np.asarray(a["sn_parent"], np.int64)
        out["sn_parent"].append(np.where(par < 0, -1, par + sn0))
        cp = np.asarray(a["child_ptr"], np.int64)
        out["child_ptr"].append(cp[:-1] + cp0)
        out["child_idx"].append(np.asarray(a["child_idx"], np.int64) + sn0)
        fp = np.asarray(a["front_row_ptr"], np.int64)
        out["front_row_ptr"].append(fp[:-1] + fr0)
        out["front_row_idx"].append(np.asarray(a["front_row_idx"], np.int64) + col0)
        pc = np.asarray(a["pcol_ptr"], np.int64)
        out["pcol_ptr"].append(pc[:-1] + pc0)
        out["prow_idx"].append(np.asarray(a["prow_idx"], np.int64) + col0)
        o2s = np.asarray(a["orig_to_slot"], np.int64)
        out["orig_to_slot"].append(np.where(o2s < 0, -1, o2s + slot0))
        out["perm"].append(np.asarray(a["perm"], np.int64) + col0)
        sn0 += ns
        col0 += n
        val0 += o2s.size
        slot0 += int(pc[-1])
        cp0 += int(cp[-1])
        fr0 += int(fp[-1])
        pc0 += int(pc[-1])
    out["child_ptr"].append(np.array([cp0]))
    out["front_row_ptr"].append(np.array([fr0]))
    out["pcol_ptr"].append(np.array([pc0]))
    arrays = {k: np.concatenate(v).astype(np.int32) for k, v in out.items()}
    for k in ("pcol_ptr", "child_ptr", "front_row_ptr"):
        if int(arrays[k][-1]) != int(np.concatenate(out[k])[-1]):
            raise OverflowError(f"{k} exceeds int32")
    arrays["n"] = col0
    arrays["postorder"] = np.arange(sn0, dtype=np.int32)
    offs = {k: np.array(v + [tot], np.int64) for (k, v), tot in zip(offs.items(), (sn0, col0, val0))}
    return arrays, offs


def system_stats(front_stats, sn_offsets):
    """Per system: (positive, negative, zero, 2x2 pivots, delayed pivots)
    summed over its supernodes (columns 2..6 of the per-front statistics)."""
    fs = np.asarray(front_stats, np.int64)
    cs = np.vstack([np.zeros((1, fs.shape[1]), np.int64), np.cumsum(fs, axis=0)])
    per = cs[sn_offsets[1:]] - cs[sn_offsets[:-1]]
    return per[:, 2:7]


def ocp_batch(nsys, tmin=10, tmax=200, seed=0):
    """nsys trajectory-optimisation KKTs with horizons T ~ U{tmin..tmax}
    (12 states, 4 controls per stage); system i is ocp_kkt(T_i, seed + i)."""
    rng = np.random.default_rng(seed)
    Ts = rng.integers(tmin, tmax + 1, size=nsys)
    return [g.ocp_kkt(int(T), seed=seed + i) for i, T in enumerate(Ts)], Ts
