"""Writes the golden fixtures of this directory from the reference smf
library compiled from its sources (oracle/_ref, built by oracle/Makefile.ref).

Fixtures (numpy .npz, no pickles):
  front_kernel.npz   dense fronts (unblocked and blocked sizes): input lower
                     triangle, num_elig, u -> reference output front, row
                     permutation, tags, pivots, inertia, flops
  mf_<case>.npz      small sparse systems: the reference's own analysis
                     (AnalysisKeep arrays), values, and its factorization in
                     the FactorKeep layout, inertia, delays, 2x2 count, flops,
                     and a right-hand side with the reference's solution
  metis_<case>.npz   the reference's METIS ordering of a benchmark matrix
                     (first differences), so the benchmark factors the
                     configuration's own elimination tree
Run: python tests/golden/make_golden.py [metis]"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from opt_solver_amd import generators as g  # noqa: E402
from oracle import refshim as R  # noqa: E402


def front_cases():
    rng = np.random.default_rng(2024)
    out = {}
    cases = [(16, 10, 0.01), (24, 24, 0.01), (48, 30, 0.1), (64, 64, 1e-8), (100, 70, 0.01), (120, 120, 0.01)]
    for idx, (f, ne, u) in enumerate(cases):
        M = rng.standard_normal((f, f))
        M = M + M.T
        M[np.arange(f), np.arange(f)] *= rng.uniform(0, 0.3, f)
        A = np.tril(M)
        Fo, rows, tags, st = R.ldlt_front(A, ne, u=u, nb=16)
        out[f"in{idx}"] = A
        out[f"meta{idx}"] = np.array([f, ne, u])
        out[f"out{idx}"] = np.tril(Fo)
        out[f"rows{idx}"] = rows
        out[f"tags{idx}"] = tags
        out[f"stats{idx}"] = np.array([st["p"], st["pos"], st["neg"], st["zero"], st["flops"]])
    np.savez_compressed(os.path.join(HERE, "front_kernel.npz"), **out)


def mf_case(name, A, posdef, ordering):
    an = R.RefAnalysis(A, ordering=ordering, posdef=posdef)
    rf = R.RefFactor(an, A[3])
    ex = rf.export()
    x = np.random.default_rng(7).uniform(-1, 1, A[0])
    n, cp, ri, v = A
    cols = np.repeat(np.arange(n), np.diff(cp))
    b = np.zeros(n)
    np.add.at(b, ri, v * x[cols])
    off = ri != cols
    np.add.at(b, cols[off], v[off] * x[ri[off]])
    xs = rf.solve(b)
    pa = an.plan_arrays()
    np.savez_compressed(
        os.path.join(HERE, f"mf_{name}.npz"),
        col_ptr=cp, row_idx=ri, values=v, posdef=np.array(int(posdef)),
        **{f"an_{k}": np.asarray(val) for k, val in pa.items() if k != "n"}, an_n=np.array(pa["n"]),
        steps=ex["steps"], factor_values=ex["factor_values"], row_indices=ex["row_indices"],
        pivot_types=ex["pivot_types"],
        stats=np.array([rf.status, rf.inertia[0], rf.inertia[1], rf.inertia[2], rf.delayed, rf.two_by_two, rf.flops]),
        b=b, x_ref=xs)


def metis_perms():
    """The reference's METIS orderings (its own analysis, ordering_metis.cpp
    with the vendored METIS 5.2.1) of the configuration-4 operator (27-point
    100^3, seed 0) and the configuration-2 KKT (300^2, m = 30000, seed 1),
    stored as first differences (perm = cumsum(perm_delta))."""
    for name, A, posdef in (("metis_spd27_100", g.spd27_3d(100, seed=0), True),
                            ("metis_kkt_300", g.kkt_grid(300, 30000, seed=1)[0], False)):
        an = R.RefAnalysis(A, ordering=R.ORDER_METIS, posdef=posdef)
        perm = np.asarray(an.perm, np.int64)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"),
                            perm_delta=np.diff(perm, prepend=0).astype(np.int32), n=np.array(A[0]))


if __name__ == "__main__":
    if not R.available():
        sys.exit("oracle/_ref/libsmfref.so is not built")
    if len(sys.argv) > 1 and sys.argv[1] == "metis":
        metis_perms()
        sys.exit(0)
    metis_perms()
    front_cases()
    mf_case("laplace7_8", g.laplace7_3d(8), True, R.ORDER_METIS)
    mf_case("ocp_T10", g.ocp_kkt(10, seed=0), False, R.ORDER_AMD)
    mf_case("kkt_12", g.kkt_grid(12, 30, seed=1)[0], False, R.ORDER_METIS)
    print("written", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
