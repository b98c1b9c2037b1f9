"""Shared test helpers: the oracle pipeline computed independently from the raw input."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def oracle_tree(x32, y32, m32=None, mode="quadtree", cap=16, bits=32, depth=None):
    """Bounds -> codes -> stable sort -> tree -> FP64 moments, all on the CPU oracle."""
    x = np.asarray(x32, np.float32).astype(np.float64)
    y = np.asarray(y32, np.float32).astype(np.float64)
    m = np.ones_like(x) if m32 is None else np.asarray(m32, np.float32).astype(np.float64)
    minx, miny, maxx, maxy = x.min(), y.min(), x.max(), y.max()
    if mode == "quadtree":
        side = O.quadtree_side(minx, miny, maxx, maxy)
        codes = O.morton64(x, y, minx, miny, side, bits)
    elif mode == "gpu_ref":  # the reference GPU step's tree, depth 6 unless given
        if depth is None:
            depth = 6
        x32, y32 = np.asarray(x32, np.float32), np.asarray(y32, np.float32)
        side = float(max(np.float32(x32.max() - x32.min()), np.float32(y32.max() - y32.min())))
        codes = O.gpuref_codes32(x32, y32, depth).astype(np.uint64)
    else:
        side = max(maxx + (-minx), maxy + (-miny))
        codes = O.morton32(x, y, minx, miny, side).astype(np.uint64)
    perm = O.stable_sort(codes)
    cs = codes[perm]
    if mode == "quadtree":
        tree = O.adaptive_tree(cs, bits=bits, leaf_cap=cap)
        pp = (tree["nchild"] <= 1).astype(np.uint8)
    else:
        if depth is None:
            depth = O.get_depth(cap, x.size)
        tree = O.fixed_preorder(cs.astype(np.uint32), depth)
        pp = (tree["level"] == depth).astype(np.uint8)
    xs, ys, ms = x[perm], y[perm], m[perm]
    mom = O.node_moments(xs, ys, ms, tree)
    return dict(x=x, y=y, m=m, xs=xs, ys=ys, ms=ms, codes=codes, perm=perm, codes_sorted=cs,
                tree=tree, pp=pp, mom=mom, side=side, depth=depth,
                bounds=(minx, miny, maxx, maxy))


def oracle_forces(t, kernel, theta, mode="quadtree", G=1.0, eps=0.0, nthreads=8):
    """FP64 Barnes-Hut on the oracle tree; returns (ax, ay, tot) in the caller's order."""
    sem = O.SEM_QUADTREE if mode == "quadtree" else O.SEM_MORTON
    ax, ay, tot, cnt = O.bh_traverse(t["xs"], t["ys"], t["ms"], t["tree"], t["mom"], t["side"],
                                     sem, kernel, G=G, eps2=eps * eps, theta=theta, pp=t["pp"],
                                     nthreads=nthreads)
    n = t["x"].size
    out = np.empty((3, n))
    out[0, t["perm"]] = ax
    out[1, t["perm"]] = ay
    out[2, t["perm"]] = tot
    return out[0], out[1], out[2], cnt


def rel_err(ax, ay, bx, by):
    """Per-particle |F_a - F_b| / |F_b|."""
    num = np.hypot(np.asarray(ax, np.float64) - bx, np.asarray(ay, np.float64) - by)
    den = np.hypot(bx, by)
    return num / np.where(den > 0, den, 1.0)
