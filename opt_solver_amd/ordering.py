"""Fill-reducing orderings for the generated workloads.

smf orders with METIS nested dissection or AMD inside its analysis
(solver/src/ordering_metis.cpp, ordering_amd.cpp of the reference); that
analysis stays smf's. For the library's own analysis of the benchmark
matrices, whose grid geometry is known, this module gives a geometric
nested dissection: recursive bisection of the box by a one-node-thick plane
across its longest side (a separator for 5-, 7- and 27-point stencils), down
to leaves of at most ``leaf`` nodes ordered lexicographically. Nodes off the
grid (KKT multipliers) join the separator of the smallest dissection box
holding all of their grid neighbours, so the ordering stays a nested
dissection of the whole graph.

All orderings are ``perm[new] = old``."""
from __future__ import annotations

import numpy as np

__all__ = ["grid_nd", "kkt_grid_nd", "identity"]


def identity(n):
    return np.arange(n, dtype=np.int32)


def _nd_boxes(dims, leaf):
    """Dissection tree of the box [0, dims): list of (lo, hi, sep_axis, sep_pos)
    in postorder, a separator plane of each internal box last."""
    out = []

    def rec(lo, hi):
        size = int(np.prod([h - l for l, h in zip(lo, hi)]))
        ext = [h - l for l, h in zip(lo, hi)]
        ax = int(np.argmax(ext))
        if size <= leaf or ext[ax] < 3:
            out.append((tuple(lo), tuple(hi), -1, -1))
            return
        mid = lo[ax] + ext[ax] // 2
        hi1 = list(hi)
        hi1[ax] = mid
        lo2 = list(lo)
        lo2[ax] = mid + 1
        rec(lo, hi1)
        rec(lo2, hi)
        out.append((tuple(lo), tuple(hi), ax, mid))

    rec([0] * len(dims), list(dims))
    return out


def _box_nodes(lo, hi, ax, pos, dims):
    """Grid node ids (x fastest) of box [lo, hi), restricted to the plane
    coordinate[ax] == pos when ax >= 0."""
    rngs = []
    for d in range(len(dims)):
        if d == ax:
            rngs.append(np.array([pos]))
        else:
            rngs.append(np.arange(lo[d], hi[d]))
    # ids = x + dims0 * (y + dims1 * z)
    grids = np.meshgrid(*rngs[::-1], indexing="ij")  # z, y, x order -> x fastest
    coords = [g.reshape(-1) for g in grids[::-1]]
    ids = np.zeros(coords[0].size, dtype=np.int64)
    mul = 1
    for d in range(len(dims)):
        ids += coords[d] * mul
        mul *= dims[d]
    return ids


def grid_nd(dims, leaf=64):
    """Geometric nested dissection of a grid with node id x + d0 (y + d1 z)."""
    boxes = _nd_boxes(dims, leaf)
    parts = [_box_nodes(lo, hi, ax, pos, dims) for lo, hi, ax, pos in boxes]
    perm = np.concatenate(parts).astype(np.int32)
    assert perm.size == int(np.prod(dims))
    return perm


def kkt_grid_nd(g, m, A_cols, leaf=64):
    """Nested dissection of the KKT graph of ``generators.kkt_grid``: the g x g
    grid part by geometric dissection; multiplier row i (node g*g + i), whose
    grid neighbours are A_cols[i], joins the separator of the smallest
    dissection box containing all its neighbours (the leaf itself when they
    share one)."""
    dims = (g, g)
    boxes = _nd_boxes(dims, leaf)
    nb = len(boxes)
    # box of each grid node in the postordered tree + parent links
    owner = np.full(g * g, -1, dtype=np.int64)
    parent = np.full(nb, -1, dtype=np.int64)
    stack = []
    for b, (lo, hi, ax, pos) in enumerate(boxes):
        if ax < 0:
            stack.append(b)
            owner[_box_nodes(lo, hi, -1, -1, dims)] = b
        else:
            c2 = stack.pop()
            c1 = stack.pop()
            parent[c1] = b
            parent[c2] = b
            stack.append(b)
            owner[_box_nodes(lo, hi, ax, pos, dims)] = b
    depth = np.zeros(nb, dtype=np.int64)
    for b in range(nb - 1, -1, -1):
        if parent[b] >= 0:
            depth[b] = depth[parent[b]] + 1

    def lca(a, b):
        while a != b:
            if depth[a] > depth[b]:
                a = parent[a]
            else:
                b = parent[b]
        return a

    A_cols = np.asarray(A_cols)
    extra_owner = np.empty(A_cols.shape[0], dtype=np.int64)
    for i in range(A_cols.shape[0]):
        o = owner[A_cols[i, 0]]
        for c in A_cols[i, 1:]:
            o = lca(o, owner[c])
        extra_owner[i] = o
    order = []
    ext_by_box = [[] for _ in range(nb)]
    for i in np.argsort(extra_owner, kind="stable"):
        ext_by_box[extra_owner[i]].append(g * g + i)
    for b, (lo, hi, ax, pos) in enumerate(boxes):
        order.append(_box_nodes(lo, hi, ax, pos, dims) if ax >= 0 else _box_nodes(lo, hi, -1, -1, dims))
        if ext_by_box[b]:
            order.append(np.asarray(ext_by_box[b], dtype=np.int64))
    perm = np.concatenate(order).astype(np.int32)
    assert perm.size == g * g + A_cols.shape[0]
    return perm
