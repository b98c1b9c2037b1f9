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


def exact_walk(ctx, x32, y32, m32, kernel, theta, eps=0.0, G=1.0, sinks=None, t64=None,
               nthreads=8):
    """The oracle's walk over the GPU's own exported node records with the traversal's
    single-precision acceptance expression (orc_bh_traverse_f32), for comparison with the
    forces ctx computed at `theta` on its last build.

    Returns a dict in the caller's particle order (or per caller-order sink): ax, ay, tot (FP64
    sums of the selected interactions), abs_x, abs_y (sums of |contribution|), and pn, pp (the
    interaction counts of those sinks) and, when t64 (an oracle_tree of the same input) is
    given, `mismatch`: the decisions on the walk that the reference's FP64 rule would take the
    other way."""
    _, perm = ctx.morton()
    g = ctx.tree()
    _, side = ctx.bounds()
    n = perm.size
    x32 = np.asarray(x32, np.float32)
    y32 = np.asarray(y32, np.float32)
    m = np.ones(n, np.float32) if m32 is None else np.asarray(m32, np.float32)
    tree = dict(level=g["level"], first=g["first"], count=g["count"], skip=g["skip"],
                pp=((g["flags"] >> 1) & 1).astype(np.uint8), cx=g["cx"], cy=g["cy"],
                mass=g["mass"])
    where = np.empty(n, np.int64)
    where[perm] = np.arange(n)
    ss = None if sinks is None else where[np.asarray(sinks, np.int64)]
    e32 = np.float32(eps)
    eps2 = float(e32 * e32)
    sem = O.SEM_QUADTREE if ctx.tree_mode == "quadtree" else O.SEM_MORTON
    mom64 = None if t64 is None else t64["mom"]
    ax, ay, tot, bx, by, cnt = O.bh_traverse_f32(
        x32[perm], y32[perm], m[perm], tree, np.float32(side), O.inv_theta2_f32(theta),
        O.KERNEL_GRAVITY if kernel == "gravity" else O.KERNEL_TSNE, G=G, eps2=eps2, sinks=ss,
        mom64=mom64, side64=side if t64 is None else t64["side"], theta=theta, semantics=sem,
        nthreads=nthreads)
    out = dict(pn=int(cnt[0]), pp=int(cnt[1]), mismatch=int(cnt[2]) if t64 is not None else None)
    if sinks is None:
        for k, v in (("ax", ax), ("ay", ay), ("tot", tot), ("abs_x", bx), ("abs_y", by)):
            a = np.empty(n)
            a[perm] = v
            out[k] = a
    else:
        out.update(ax=ax, ay=ay, tot=tot, abs_x=bx, abs_y=by)
    return out


def walk_error(ax, ay, w):
    """Per-particle error of GPU forces against exact_walk's sums, each component normalised by
    the sum of the absolute values of its contributions (a bound that does not grow where the
    contributions cancel)."""
    ex = np.abs(np.asarray(ax, np.float64) - w["ax"]) / np.where(w["abs_x"] > 0, w["abs_x"], 1.0)
    ey = np.abs(np.asarray(ay, np.float64) - w["ay"]) / np.where(w["abs_y"] > 0, w["abs_y"], 1.0)
    return np.maximum(ex, ey)


# the bar of the exact-walk comparison: FP32 interactions with compensated accumulation against
# FP64 sums of the same interactions, per component relative to the sum of |contributions|
WALK_TOL = 1e-5


def check_exact_walk(ctx, x, y, m, kernel, theta, eps, ax, ay, tot, t64=None):
    """GPU forces and counters against the oracle's walk over the GPU's own node records with
    the traversal's single-precision acceptance: equal particle-node and particle-particle
    counts, every particle's error within WALK_TOL of its |contribution| sums, the total within
    the same bar. Returns the oracle walk (with the FP64-rule disagreement count when t64 is
    given)."""
    w = exact_walk(ctx, x, y, m, kernel, theta, eps=eps, t64=t64)
    pn, pp = ctx.counters()
    assert (int(pn), int(pp)) == (w["pn"], w["pp"]), ((pn, pp), (w["pn"], w["pp"]))
    e = walk_error(ax, ay, w)
    assert e.max() <= WALK_TOL, (float(e.max()), int(np.argmax(e)))
    assert abs(tot - w["tot"].sum()) <= WALK_TOL * np.abs(w["tot"]).sum()
    return w


def batched_export(ctx):
    """(offsets, sort permutation, exported tree) of a batched context, for exact_walk_system."""
    _, off = ctx.batch_info()
    _, perm = ctx.morton()
    return off, perm, ctx.tree()


def exact_walk_system(ctx, s, x32, y32, m32, kernel, theta, sinks, eps=0.0, G=1.0, R=4,
                      nthreads=8, export=None):
    """exact_walk for system s of a batched context: the oracle's single-precision-decision walk
    over that system's subtree of the GPU's node records (levels counted from the system root at
    level R, side = the system's own QuadTree root square), for caller-order sink indices local
    to the system. x32, y32, m32: the system's own particles; export: batched_export(ctx), taken
    once for many systems. Returns (ax, ay, abs_x, abs_y) per sink."""
    off, perm, g = batched_export(ctx) if export is None else export
    lo, hi = int(off[s]), int(off[s + 1])
    root = int(np.flatnonzero((g["first"] == lo) & (g["level"] == R))[0])
    end = int(g["skip"][root])
    sl = slice(root, end)
    tree = dict(level=g["level"][sl] - R, first=g["first"][sl] - lo, count=g["count"][sl],
                skip=np.minimum(g["skip"][sl], end) - root,
                pp=((g["flags"][sl] >> 1) & 1).astype(np.uint8), cx=g["cx"][sl], cy=g["cy"][sl],
                mass=g["mass"][sl])
    local = perm[lo:hi].astype(np.int64) - lo  # system-local caller index at each sorted slot
    x32 = np.asarray(x32, np.float32)
    y32 = np.asarray(y32, np.float32)
    m = np.ones(hi - lo, np.float32) if m32 is None else np.asarray(m32, np.float32)
    xd, yd = x32.astype(np.float64), y32.astype(np.float64)
    side = O.quadtree_side(xd.min(), yd.min(), xd.max(), yd.max())
    where = np.empty(hi - lo, np.int64)
    where[local] = np.arange(hi - lo)
    e32 = np.float32(eps)
    ax, ay, _, bx, by, _ = O.bh_traverse_f32(
        x32[local], y32[local], m[local], tree, np.float32(side), O.inv_theta2_f32(theta),
        O.KERNEL_GRAVITY if kernel == "gravity" else O.KERNEL_TSNE, G=G, eps2=float(e32 * e32),
        sinks=where[np.asarray(sinks, np.int64)], nthreads=nthreads)
    return ax, ay, bx, by
