"""ctypes wrapper of the CPU oracle (liboracle.so) and of the compiled reference (_ref/libref_bh.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, to check results / as the CPU baseline. The product path never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_bh.so")

KERNEL_GRAVITY = 0
KERNEL_TSNE = 1
SEM_QUADTREE = 0
SEM_MORTON = 1

_lib = None
_ref = None

_d = C.POINTER(C.c_double)
_u32 = C.POINTER(C.c_uint32)
_u64 = C.POINTER(C.c_uint64)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)
_u8 = C.POINTER(C.c_uint8)
_vp = C.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        I64, I32, D = C.c_int64, C.c_int, C.c_double
        sig = {
            "orc_morton32": (None, [_vp, _vp, I64, D, D, D, _vp]),
            "orc_gpuref_codes32": (None, [_vp, _vp, I64, I32, _vp]),
            "orc_morton64": (None, [_vp, _vp, I64, D, D, D, I32, _vp]),
            "orc_quantize_axis": (None, [_vp, I64, D, D, I32, _vp]),
            "orc_stable_sort_u64": (None, [_vp, I64, _vp]),
            "orc_get_depth": (I32, [I32, I64]),
            "orc_fixed_tree_slots": (I64, [I32]),
            "orc_fixed_tree": (None, [_vp, _vp, I64, I32, D, D, D, D] + [_vp] * 10),
            "orc_adaptive_tree": (I64, [_vp, I64, I32, I32, I64] + [_vp] * 5),
            "orc_fixed_preorder": (I64, [_vp, I64, I32, I64] + [_vp] * 5),
            "orc_node_moments": (None, [_vp, _vp, _vp, I64, _vp, _vp, _vp, _vp, _vp, _vp]),
            "orc_bh_traverse": (None, [_vp, _vp, _vp, I64] + [_vp] * 8 + [I64, D, I32, I32, D, D, D,
                                                                          _vp, I64, _vp, _vp, _vp,
                                                                          _vp, I32]),
            "orc_bh_traverse_f32": (None, [_vp, _vp, _vp, I64] + [_vp] * 8 + [I64, C.c_float,
                                                                          C.c_float, I32, D, D,
                                                                          _vp, _vp, D, D, I32,
                                                                          _vp, I64] + [_vp] * 6
                                    + [I32]),
            "orc_direct": (None, [_vp, _vp, _vp, I64, I32, D, D, _vp, I64, _vp, _vp, _vp, I32]),
            "orc_cell_features": (None, [_vp, _vp, _vp, I64] + [_vp] * 4 + [I32, D, D, D]
                                  + [_vp] * 4),
            "orc_learned_basis": (None, [D, D, D, _vp]),
            "orc_pair_label": (None, [_vp, _vp, _vp, C.c_uint32, C.c_uint32, D, D, D, D, D, D,
                                      I64, D, _vp]),
            "orc_learned_loss_grad": (D, [_vp, _vp, I64, I32, I32] + [_vp] * 12),
            "orc_bh_traverse_learned": (None, [_vp, _vp, _vp, I64] + [_vp] * 8
                                        + [I64, C.c_float, C.c_float, D, D, _vp, C.c_float, _vp,
                                           I64, _vp, _vp, _vp]),
            "orc_quadtree_build": (_vp, [_vp, _vp, I64, I32]),
            "orc_quadtree_free": (None, [_vp]),
            "orc_quadtree_node_count": (I64, [_vp]),
            "orc_quadtree_export": (None, [_vp] + [_vp] * 10),
            "orc_quadtree_occupants": (None, [_vp, _vp, _vp]),
            "orc_quadtree_forces": (D, [_vp, _vp, _vp, I32, D, D, D, D, _vp, _vp, _vp, I32]),
            "orc_leapfrog": (None, [_vp] * 5 + [I64, I32, D, D, D, D, I32, I32, I32, I32]),
            "orc_tsne_theta": (D, [D, D, I32]),
            "orc_tsne_attraction": (None, [_vp, _vp, I64, _vp, _vp, _vp, D, _vp, _vp]),
            "orc_tsne_kl": (D, [_vp, _vp, I64, _vp, _vp, _vp, I32]),
            "orc_tsne_z": (D, [_vp, _vp, I64, I32]),
            "orc_tsne_step": (D, [_vp, _vp, _vp, _vp, I64, _vp, _vp, _vp, D, D, D, D, _vp, _vp,
                                  _vp, _vp, I32]),
            "orc_mlp_forward": (None, [_vp, I64, I32, I32, I32] + [_vp] * 7),
            "orc_mlp_loss_grad": (D, [_vp, _vp, I64, I32, I32, I32] + [_vp] * 12),
            "orc_adam": (None, [_vp, _vp, _vp, _vp, I64, D, D, D, D, I64]),
            "orc_fmm_sym": (D, [_vp, _vp, I64, I32] + [_vp] * 8 + [D, _vp, _vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"compiled reference missing: {REF_PATH} (run `make -C oracle ref`)")
        L = C.CDLL(REF_PATH)
        I64, I32, D = C.c_int64, C.c_int, C.c_double
        L.ref_bh_solve.restype = D
        L.ref_bh_solve.argtypes = [_vp, _vp, I64, I32, D, D, D, D, I32, _vp, _vp, _vp, _vp]
        L.ref_quadtree_export.restype = I64
        L.ref_quadtree_export.argtypes = [_vp, _vp, I64, I32] + [_vp] * 10
        L.ref_naive_tsne.restype = D
        L.ref_naive_tsne.argtypes = [_vp, _vp, I64, _vp, _vp]
        _ref = L
    return _ref


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------------------------------------
def morton32(x, y, minx, miny, L):
    x, y = _f64(x), _f64(y)
    out = np.empty(x.size, np.uint32)
    lib().orc_morton32(_p(x), _p(y), x.size, float(minx), float(miny), float(L), _p(out))
    return out


def gpuref_codes32(x32, y32, depth):
    """Paths of the reference GPU step's partition (nBodySolverGpuBH.cu:78-200) as 32-bit codes."""
    x = np.ascontiguousarray(x32, np.float32)
    y = np.ascontiguousarray(y32, np.float32)
    out = np.empty(x.size, np.uint32)
    lib().orc_gpuref_codes32(_p(x), _p(y), x.size, int(depth), _p(out))
    return out


def morton64(x, y, minx, miny, L, bits=32):
    x, y = _f64(x), _f64(y)
    out = np.empty(x.size, np.uint64)
    lib().orc_morton64(_p(x), _p(y), x.size, float(minx), float(miny), float(L), int(bits), _p(out))
    return out


def quantize_axis(v, vmin, L, bits):
    v = _f64(v)
    out = np.empty(v.size, np.uint64)
    lib().orc_quantize_axis(_p(v), v.size, float(vmin), float(L), int(bits), _p(out))
    return out


def stable_sort(keys):
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    perm = np.empty(k.size, np.uint32)
    lib().orc_stable_sort_u64(_p(k), k.size, _p(perm))
    return perm


def get_depth(max_children, n):
    return int(lib().orc_get_depth(int(max_children), int(n)))


def quadtree_side(minx, miny, maxx, maxy):
    """Root side of the reference QuadTree: ((min + largest) + 1e-4) - min (quadtree.h:45-47)."""
    largest = max(maxx - minx, maxy - miny)
    return ((minx + largest) + 0.0001) - minx


def fixed_tree(xs, ys, depth, minx, miny, maxx, maxy):
    xs, ys = _f64(xs), _f64(ys)
    slots = int(lib().orc_fixed_tree_slots(int(depth)))
    out = dict(
        first_child=np.empty(slots, np.uint32), first_particle=np.empty(slots, np.uint32),
        count=np.empty(slots, np.uint32), bb_cx=np.empty(slots), bb_cy=np.empty(slots),
        bb_len=np.empty(slots), com_x=np.empty(slots), com_y=np.empty(slots), m0=np.empty(slots),
        m2=np.empty(3 * slots))
    lib().orc_fixed_tree(_p(xs), _p(ys), xs.size, int(depth), float(minx), float(miny), float(maxx),
                         float(maxy), *[_p(out[k]) for k in ("first_child", "first_particle", "count",
                                                             "bb_cx", "bb_cy", "bb_len", "com_x",
                                                             "com_y", "m0", "m2")])
    out["m2"] = out["m2"].reshape(slots, 3)
    return out


def fmm_sym(xs, ys, tree, depth, theta):
    """Symmetric dual-tree FMM with the t-SNE kernels (nBodySolverFMM_SYM_MORTON.h) over a
    fixed_tree(): (repulsive derivative x, y per sorted point, total Z), FP64."""
    xs, ys = _f64(xs), _f64(ys)
    ax, ay = np.zeros(xs.size), np.zeros(xs.size)
    m2 = np.ascontiguousarray(tree["m2"], dtype=np.float64).reshape(-1)
    z = lib().orc_fmm_sym(_p(xs), _p(ys), xs.size, int(depth), _p(tree["first_child"]),
                          _p(tree["first_particle"]), _p(tree["count"]), _p(tree["bb_len"]),
                          _p(tree["com_x"]), _p(tree["com_y"]), _p(tree["m0"]), _p(m2),
                          float(theta), _p(ax), _p(ay))
    return ax, ay, z


def _pre_arrays(cap):
    return dict(level=np.empty(cap, np.int32), first=np.empty(cap, np.uint32),
                count=np.empty(cap, np.uint32), skip=np.empty(cap, np.uint32),
                nchild=np.empty(cap, np.uint32))


def adaptive_tree(codes_sorted, bits=32, leaf_cap=16):
    c = np.ascontiguousarray(codes_sorted, dtype=np.uint64)
    cap = 3 * c.size + 64
    a = _pre_arrays(cap)
    nn = lib().orc_adaptive_tree(_p(c), c.size, int(bits), int(leaf_cap), cap, _p(a["level"]),
                                 _p(a["first"]), _p(a["count"]), _p(a["skip"]), _p(a["nchild"]))
    if nn < 0:
        raise RuntimeError("node capacity exceeded")
    return {k: v[:nn].copy() for k, v in a.items()}


def fixed_preorder(codes32_sorted, depth):
    c = np.ascontiguousarray(codes32_sorted, dtype=np.uint32)
    cap = int(lib().orc_fixed_tree_slots(int(depth))) + 64
    cap = min(cap, c.size * (depth + 1) + 64)
    a = _pre_arrays(cap)
    nn = lib().orc_fixed_preorder(_p(c), c.size, int(depth), cap, _p(a["level"]), _p(a["first"]),
                                  _p(a["count"]), _p(a["skip"]), _p(a["nchild"]))
    if nn < 0:
        raise RuntimeError("node capacity exceeded")
    return {k: v[:nn].copy() for k, v in a.items()}


def node_moments(xs, ys, ms, tree, with_m2=False):
    xs, ys = _f64(xs), _f64(ys)
    ms = None if ms is None else _f64(ms)
    nn = tree["first"].size
    cx, cy, mass = np.empty(nn), np.empty(nn), np.empty(nn)
    m2 = np.empty(3 * nn) if with_m2 else None
    lib().orc_node_moments(_p(xs), _p(ys), _p(ms), nn, _p(tree["first"]), _p(tree["count"]),
                           _p(cx), _p(cy), _p(mass), _p(m2))
    out = dict(cx=cx, cy=cy, mass=mass)
    if with_m2:
        out["m2"] = m2.reshape(nn, 3)
    return out


def bh_traverse(xs, ys, ms, tree, mom, root_side, semantics, kernel, G=1.0, eps2=0.0, theta=0.5,
                sinks=None, pp=None, nthreads=8):
    xs, ys = _f64(xs), _f64(ys)
    ms = None if ms is None else _f64(ms)
    if pp is None:
        pp = (tree["nchild"] <= 1).astype(np.uint8)
    pp = np.ascontiguousarray(pp, dtype=np.uint8)
    ns = xs.size if sinks is None else len(sinks)
    sk = None if sinks is None else np.ascontiguousarray(sinks, dtype=np.int64)
    ax, ay, tot = np.empty(ns), np.empty(ns), np.empty(ns)
    cnt = np.zeros(2, np.uint64)
    lib().orc_bh_traverse(_p(xs), _p(ys), _p(ms), xs.size, _p(tree["level"]), _p(tree["first"]),
                          _p(tree["count"]), _p(tree["skip"]), _p(pp), _p(mom["cx"]),
                          _p(mom["cy"]), _p(mom["mass"]), tree["first"].size, float(root_side),
                          int(semantics), int(kernel), float(G), float(eps2), float(theta), _p(sk),
                          ns, _p(ax), _p(ay), _p(tot), _p(cnt), int(nthreads))
    return ax, ay, tot, cnt


def bh_traverse_f32(xs, ys, ms, tree, root_side, inv_theta2, kernel, G=1.0, eps2=0.0, sinks=None,
                    mom64=None, side64=0.0, theta=0.5, semantics=SEM_QUADTREE, nthreads=8):
    """Walk over given single-precision node records (tree: level, first, count, skip, pp, cx,
    cy, mass) with the single-precision acceptance expression d2 > l2 * inv_theta2; FP64
    interaction sums. Returns ax, ay, tot, abs_x, abs_y (indexed like sinks) and the counters
    [pn, pp, mismatches against the FP64 rule evaluated with mom64 (cx, cy) and side64]."""
    xs = np.ascontiguousarray(xs, np.float32)
    ys = np.ascontiguousarray(ys, np.float32)
    ms = None if ms is None else np.ascontiguousarray(ms, np.float32)
    lev = np.ascontiguousarray(tree["level"], np.int32)
    fi = np.ascontiguousarray(tree["first"], np.uint32)
    co = np.ascontiguousarray(tree["count"], np.uint32)
    sk_ = np.ascontiguousarray(tree["skip"], np.uint32)
    pp = np.ascontiguousarray(tree["pp"], np.uint8)
    cx = np.ascontiguousarray(tree["cx"], np.float32)
    cy = np.ascontiguousarray(tree["cy"], np.float32)
    mass = np.ascontiguousarray(tree["mass"], np.float32)
    ns = xs.size if sinks is None else len(sinks)
    sk = None if sinks is None else np.ascontiguousarray(sinks, dtype=np.int64)
    ax, ay, tot, bx, by = (np.empty(ns) for _ in range(5))
    cnt = np.zeros(3, np.uint64)
    c64x = None if mom64 is None else _f64(mom64["cx"])
    c64y = None if mom64 is None else _f64(mom64["cy"])
    lib().orc_bh_traverse_f32(_p(xs), _p(ys), _p(ms), xs.size, _p(lev), _p(fi), _p(co), _p(sk_),
                              _p(pp), _p(cx), _p(cy), _p(mass), lev.size, float(np.float32(root_side)),
                              float(np.float32(inv_theta2)), int(kernel), float(G), float(eps2),
                              _p(c64x), _p(c64y), float(side64), float(theta), int(semantics),
                              _p(sk), ns, _p(ax), _p(ay), _p(tot), _p(bx), _p(by), _p(cnt),
                              int(nthreads))
    return ax, ay, tot, bx, by, cnt


def inv_theta2_f32(theta):
    """1/theta^2 as the traversal receives it: computed in FP64, rounded to single precision."""
    with np.errstate(divide="ignore"):
        return np.float32(1.0 / (float(theta) * float(theta))) if theta > 0 else np.float32(np.inf)


def direct(x, y, m, kernel, G=1.0, eps2=0.0, sinks=None, nthreads=8):
    x, y = _f64(x), _f64(y)
    m = None if m is None else _f64(m)
    ns = x.size if sinks is None else len(sinks)
    sk = None if sinks is None else np.ascontiguousarray(sinks, dtype=np.int64)
    ax, ay, tot = np.empty(ns), np.empty(ns), np.empty(ns)
    lib().orc_direct(_p(x), _p(y), _p(m), x.size, int(kernel), float(G), float(eps2), _p(sk), ns,
                     _p(ax), _p(ay), _p(tot), int(nthreads))
    return ax, ay, tot


class QuadTree:
    """Restated reference QuadTree (quadtree.h) + traverseBH (nBodySolverBH.h)."""

    def __init__(self, x, y, max_children=16):
        self.x, self.y = _f64(x), _f64(y)
        self.h = lib().orc_quadtree_build(_p(self.x), _p(self.y), self.x.size, int(max_children))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_quadtree_free(self.h)
            self.h = None

    def export(self):
        nn = int(lib().orc_quadtree_node_count(self.h))
        d = dict(depth=np.empty(nn, np.int32), count=np.empty(nn, np.uint32),
                 nchild=np.empty(nn, np.uint32), lowx=np.empty(nn), lowy=np.empty(nn),
                 highx=np.empty(nn), highy=np.empty(nn), mass=np.empty(nn), comx=np.empty(nn),
                 comy=np.empty(nn))
        lib().orc_quadtree_export(self.h, *[_p(d[k]) for k in ("depth", "count", "nchild", "lowx",
                                                               "lowy", "highx", "highy", "mass",
                                                               "comx", "comy")])
        return d

    def occupants(self):
        nn = int(lib().orc_quadtree_node_count(self.h))
        nf = np.empty(nn, np.uint32)
        occ = np.empty(self.x.size, np.uint32)
        lib().orc_quadtree_occupants(self.h, _p(nf), _p(occ))
        return nf, occ

    def forces(self, kernel, G=1.0, m=1.0, eps2=0.0, theta=0.5, nthreads=1):
        ax, ay = np.zeros(self.x.size), np.zeros(self.x.size)
        cnt = np.zeros(2, np.uint64)
        tot = lib().orc_quadtree_forces(self.h, _p(self.x), _p(self.y), int(kernel), float(G),
                                        float(m), float(eps2), float(theta), _p(ax), _p(ay),
                                        _p(cnt), int(nthreads))
        return ax, ay, tot, cnt


def leapfrog(x, y, vx, vy, m, steps, dt, G, eps2, theta, semantics=SEM_QUADTREE, leaf_cap=16,
             bits=32, nthreads=8):
    x, y, vx, vy = (_f64(a).copy() for a in (x, y, vx, vy))
    m = None if m is None else _f64(m)
    lib().orc_leapfrog(_p(x), _p(y), _p(vx), _p(vy), _p(m), x.size, int(steps), float(dt), float(G),
                       float(eps2), float(theta), int(semantics), int(leaf_cap), int(bits),
                       int(nthreads))
    return x, y, vx, vy


def tsne_theta(min_theta, max_theta, iteration):
    return float(lib().orc_tsne_theta(float(min_theta), float(max_theta), int(iteration)))


def tsne_attraction(x, y, row_ptr, col, val, exag):
    x, y = _f64(x), _f64(y)
    rp = np.ascontiguousarray(row_ptr, np.uint32)
    cl = np.ascontiguousarray(col, np.uint32)
    vl = _f64(val)
    dx, dy = np.zeros(x.size), np.zeros(x.size)
    lib().orc_tsne_attraction(_p(x), _p(y), x.size, _p(rp), _p(cl), _p(vl), float(exag), _p(dx),
                              _p(dy))
    return dx, dy


def tsne_z(x, y, nthreads=8):
    x, y = _f64(x), _f64(y)
    return float(lib().orc_tsne_z(_p(x), _p(y), x.size, int(nthreads)))


def tsne_kl(x, y, row_ptr, col, val, nthreads=8):
    x, y = _f64(x), _f64(y)
    rp = np.ascontiguousarray(row_ptr, np.uint32)
    cl = np.ascontiguousarray(col, np.uint32)
    vl = _f64(val)
    return float(lib().orc_tsne_kl(_p(x), _p(y), x.size, _p(rp), _p(cl), _p(vl), int(nthreads)))


def tsne_step(x, y, px, py, row_ptr, col, val, theta, exag, lr, mu, nthreads=8):
    """One reference t-SNE iteration -> (dx, dy, x_next, y_next, Z)."""
    x, y, px, py = _f64(x), _f64(y), _f64(px), _f64(py)
    rp = np.ascontiguousarray(row_ptr, np.uint32)
    cl = np.ascontiguousarray(col, np.uint32)
    vl = _f64(val)
    dx, dy, nx, ny = (np.empty(x.size) for _ in range(4))
    z = lib().orc_tsne_step(_p(x), _p(y), _p(px), _p(py), x.size, _p(rp), _p(cl), _p(vl),
                            float(theta), float(exag), float(lr), float(mu), _p(dx), _p(dy),
                            _p(nx), _p(ny), int(nthreads))
    return dx, dy, nx, ny, float(z)


def mlp_forward(X, params):
    X = _f64(X)
    B, din = X.shape
    W1, b1, W2, b2, W3, b3 = (_f64(params[k]) for k in ("W1", "b1", "W2", "b2", "W3", "b3"))
    h, dout = W1.shape[1], W3.shape[1]
    Y = np.empty((B, dout))
    lib().orc_mlp_forward(_p(X), B, din, h, dout, _p(W1), _p(b1), _p(W2), _p(b2), _p(W3), _p(b3),
                          _p(Y))
    return Y


def mlp_loss_grad(X, T, params):
    X, T = _f64(X), _f64(T)
    B, din = X.shape
    W1, b1, W2, b2, W3, b3 = (_f64(params[k]) for k in ("W1", "b1", "W2", "b2", "W3", "b3"))
    h, dout = W1.shape[1], W3.shape[1]
    g = {k: np.empty_like(v) for k, v in (("W1", W1), ("b1", b1), ("W2", W2), ("b2", b2),
                                          ("W3", W3), ("b3", b3))}
    loss = lib().orc_mlp_loss_grad(_p(X), _p(T), B, din, h, dout, _p(W1), _p(b1), _p(W2), _p(b2),
                                   _p(W3), _p(b3), _p(g["W1"]), _p(g["b1"]), _p(g["W2"]),
                                   _p(g["b2"]), _p(g["W3"]), _p(g["b3"]))
    return float(loss), g


def adam(p, g, m1, m2, lr, beta1, beta2, eps, t):
    for a in (p, m1, m2):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    g = _f64(g)
    lib().orc_adam(_p(p), _p(g), _p(m1), _p(m2), p.size, float(lr), float(beta1), float(beta2),
                   float(eps), int(t))


# ---------------------------------------------------------------------------------------------
def ref_bh_solve(x, y, kernel, G=1.0, m=1.0, eps2=0.0, theta=0.5, max_children=16):
    x, y = _f64(x), _f64(y)
    ax, ay = np.empty(x.size), np.empty(x.size)
    tb, ts = C.c_double(0), C.c_double(0)
    tot = ref().ref_bh_solve(_p(x), _p(y), x.size, int(kernel), float(G), float(m), float(eps2),
                             float(theta), int(max_children), _p(ax), _p(ay), C.byref(tb),
                             C.byref(ts))
    return ax, ay, tot, tb.value, ts.value


def ref_quadtree(x, y, max_children=16):
    x, y = _f64(x), _f64(y)
    R = ref()
    nn = int(R.ref_quadtree_export(_p(x), _p(y), x.size, int(max_children), *([None] * 10)))
    d = dict(depth=np.empty(nn, np.int32), count=np.empty(nn, np.uint32),
             nchild=np.empty(nn, np.uint32), mass=np.empty(nn), comx=np.empty(nn),
             comy=np.empty(nn), lowx=np.empty(nn), lowy=np.empty(nn), highx=np.empty(nn),
             highy=np.empty(nn))
    R.ref_quadtree_export(_p(x), _p(y), x.size, int(max_children),
                          *[_p(d[k]) for k in ("depth", "count", "nchild", "mass", "comx", "comy",
                                               "lowx", "lowy", "highx", "highy")])
    return d


def ref_naive_tsne(x, y):
    x, y = _f64(x), _f64(y)
    ax, ay = np.empty(x.size), np.empty(x.size)
    tot = ref().ref_naive_tsne(_p(x), _p(y), x.size, _p(ax), _p(ay))
    return ax, ay, tot


# ---------------------------------------------------------------------------------------------
# learned far-field correction
# ---------------------------------------------------------------------------------------------
CELL_FEATURES = 10
LEARNED_TERMS = 7


def cell_features(xs, ys, ms, tree, codes_sorted, bits, minx, miny, root_side, mom):
    """FP64 features of every node (see orc_cell_features): [nnodes, 10]."""
    xs, ys = _f64(xs), _f64(ys)
    ms = None if ms is None else _f64(ms)
    nn = tree["first"].size
    f = np.empty((nn, CELL_FEATURES))
    lib().orc_cell_features(_p(xs), _p(ys), _p(ms), nn,
                            _p(np.ascontiguousarray(tree["level"], np.int32)),
                            _p(np.ascontiguousarray(tree["first"], np.uint32)),
                            _p(np.ascontiguousarray(tree["count"], np.uint32)),
                            _p(np.ascontiguousarray(codes_sorted, np.uint64)), int(bits),
                            float(minx), float(miny), float(root_side), _p(_f64(mom["cx"])),
                            _p(_f64(mom["cy"])), _p(_f64(mom["mass"])), _p(f))
    return f


def learned_basis(ux, uy, q):
    b = np.empty(2 * LEARNED_TERMS)
    lib().orc_learned_basis(float(ux), float(uy), float(q), _p(b))
    return b


def pair_label(xs, ys, ms, first, count, cx, cy, mass, l, px, py, self_index=-1, eps2=0.0):
    """(label x, label y, w_x, w_y, q, scale) of one (sink, cell) pair."""
    xs, ys = _f64(xs), _f64(ys)
    ms = None if ms is None else _f64(ms)
    out = np.empty(6)
    lib().orc_pair_label(_p(xs), _p(ys), _p(ms), int(first), int(count), float(cx), float(cy),
                         float(mass), float(l), float(px), float(py), int(self_index), float(eps2),
                         _p(out))
    return out


def learned_loss_grad(X, T, params, terms=LEARNED_TERMS):
    X, T = _f64(X), _f64(T)
    B = X.shape[0]
    W1, b1, W2, b2, W3, b3 = (_f64(params[k]) for k in ("W1", "b1", "W2", "b2", "W3", "b3"))
    g = {k: np.empty_like(v) for k, v in (("W1", W1), ("b1", b1), ("W2", W2), ("b2", b2),
                                          ("W3", W3), ("b3", b3))}
    loss = lib().orc_learned_loss_grad(_p(X), _p(T), B, W1.shape[1], int(terms), _p(W1), _p(b1),
                                       _p(W2), _p(b2), _p(W3), _p(b3), _p(g["W1"]), _p(g["b1"]),
                                       _p(g["W2"]), _p(g["b2"]), _p(g["W3"]), _p(g["b3"]))
    return float(loss), g


def bh_traverse_learned(xs, ys, ms, tree, root_side, inv_theta2, coef, inv_band2=0.0, G=1.0,
                        eps2=0.0, sinks=None):
    """Single-precision-decision gravity walk plus the learned correction with per-node
    coefficients coef [nnodes, 14]; returns (ax, ay, counters [pn, pp, corrected])."""
    xs = np.ascontiguousarray(xs, np.float32)
    ys = np.ascontiguousarray(ys, np.float32)
    ms = None if ms is None else np.ascontiguousarray(ms, np.float32)
    ns = xs.size if sinks is None else len(sinks)
    sk = None if sinks is None else np.ascontiguousarray(sinks, dtype=np.int64)
    ax, ay = np.empty(ns), np.empty(ns)
    cnt = np.zeros(3, np.uint64)
    arr = [np.ascontiguousarray(tree["level"], np.int32),
           np.ascontiguousarray(tree["first"], np.uint32),
           np.ascontiguousarray(tree["count"], np.uint32),
           np.ascontiguousarray(tree["skip"], np.uint32),
           np.ascontiguousarray(tree["pp"], np.uint8),
           np.ascontiguousarray(tree["cx"], np.float32),
           np.ascontiguousarray(tree["cy"], np.float32),
           np.ascontiguousarray(tree["mass"], np.float32)]
    lib().orc_bh_traverse_learned(_p(xs), _p(ys), _p(ms), xs.size, *[_p(a) for a in arr],
                                  arr[0].size, float(np.float32(root_side)),
                                  float(np.float32(inv_theta2)), float(G), float(eps2),
                                  _p(_f64(coef).reshape(-1)), float(np.float32(inv_band2)), _p(sk),
                                  ns, _p(ax), _p(ay), _p(cnt))
    return ax, ay, cnt


def quadrupole_coefficients(features):
    """The coefficients that make the learned correction the exact quadrupole term: c_1 =
    0.75 tr S, c_2 = 0.375 Q, c_3 = 1.875 conj(Q), Q = S_xx - S_yy + 2i S_xy (S as normalised
    in the features), the rest 0. [nnodes, 14]."""
    f = np.atleast_2d(features)
    c = np.zeros((f.shape[0], 2 * LEARNED_TERMS))
    sxx, sxy, syy = f[:, 0], f[:, 1], f[:, 2]
    qr, qi = sxx - syy, 2.0 * sxy
    c[:, 0] = 0.75 * (sxx + syy)
    c[:, 2], c[:, 3] = 0.375 * qr, 0.375 * qi
    c[:, 4], c[:, 5] = 1.875 * qr, -1.875 * qi
    return c


# ---------------------------------------------------------------------------------------------
# t-SNE input affinities (NumPy, FP64): the reference computes P offline with openTSNE 1.x
# (MnistLoader.ipynb:539-547: exact PerplexityBasedNN, k = 3 * perplexity, and
# joint_probabilities_nn) and reads it as a Matrix Market file (tsne.h:94-98); openTSNE is not
# importable here, so its published algorithm is restated: exact Euclidean kNN, a per-point
# bisection on the Gaussian precision beta to the perplexity's entropy, P = (P_c + P_c^T) / (2n).
# Pinned in tests/test_tsne_host.py: kNN against a full sort, the bisection against the
# two-point closed form and the entropy property, the symmetrisation against a dense matrix.
# ---------------------------------------------------------------------------------------------
def knn_exact(X, k, rows=None, chunk=256):
    """Exact k nearest neighbours (self excluded) of `rows` (all by default) by FP64 Euclidean
    distance: candidates from the FP64 Gram identity (the 4k nearest), final distances from the
    differences, ascending, ties by index. Returns (idx int64 [r, k], d2 float64 [r, k])."""
    X64 = np.asarray(X, np.float64)
    n = X64.shape[0]
    if k >= n:
        raise ValueError(f"k={k} needs more than {n} points")
    rows = np.arange(n) if rows is None else np.asarray(rows, np.int64)
    sq = (X64 * X64).sum(1)
    m = min(n - 1, 4 * k)
    idx = np.empty((rows.size, k), np.int64)
    d2 = np.empty((rows.size, k))
    for s in range(0, rows.size, chunk):
        r = rows[s:s + chunk]
        G = sq[r, None] + sq[None, :] - 2.0 * (X64[r] @ X64.T)
        G[np.arange(r.size), r] = np.inf
        cand = np.argpartition(G, m - 1, axis=1)[:, :m]
        diff = X64[r, None, :] - X64[cand]
        dc = np.einsum("ijk,ijk->ij", diff, diff)
        order = np.lexsort((cand, dc), axis=1)[:, :k]
        idx[s:s + r.size] = np.take_along_axis(cand, order, 1)
        d2[s:s + r.size] = np.take_along_axis(dc, order, 1)
    return idx, d2


def conditional_probabilities(d2, perplexity, tol=1e-5, max_iter=200):
    """Row-wise p_{j|i} = exp(-beta_i (d2_ij - d2_i0)) / sum_j, beta_i bisected (doubling /
    halving until bracketed, then midpoints) until the entropy (natural log) is within tol of
    log(perplexity). d2: float64 [n, k], ascending per row. Returns (P [n, k], beta [n])."""
    import math

    d2 = np.asarray(d2, dtype=np.float64)
    n = d2.shape[0]
    target = math.log(perplexity)
    dr = d2 - d2[:, :1]
    beta = np.ones(n)
    lo = np.full(n, -np.inf)
    hi = np.full(n, np.inf)
    done = np.zeros(n, bool)
    P = None
    for _ in range(max_iter):
        E = np.exp(-beta[:, None] * dr)
        S = E.sum(1)
        P = E / S[:, None]
        H = np.log(S) + beta * (dr * P).sum(1)
        diff = H - target
        done |= np.abs(diff) < tol
        if done.all():
            break
        up = (diff > 0) & ~done  # entropy too high: sharpen
        dn = (diff <= 0) & ~done
        lo = np.where(up, beta, lo)
        hi = np.where(dn, beta, hi)
        beta = np.where(up, np.where(np.isinf(hi), beta * 2.0, 0.5 * (beta + hi)), beta)
        beta = np.where(dn, np.where(np.isinf(lo), beta * 0.5, 0.5 * (beta + lo)), beta)
    return P, beta


def symmetrize(idx, Pc):
    """P = (P_c + P_c^T) / (2 n) from neighbour lists, as CSR sorted by (row, col): each value
    divided by 2n in FP64, equal (row, col) entries summed in FP64, rounded to FP32."""
    n, k = idx.shape
    rows = np.repeat(np.arange(n, dtype=np.int64), k)
    cols = np.asarray(idx).reshape(-1).astype(np.int64)
    vals = np.asarray(Pc, np.float64).reshape(-1)
    r = np.concatenate([rows, cols])
    c = np.concatenate([cols, rows])
    v = np.concatenate([vals, vals]) / (2.0 * n)
    key = r * n + c
    order = np.argsort(key, kind="stable")
    key, v = key[order], v[order]
    uniq, start = np.unique(key, return_index=True)
    vsum = np.add.reduceat(v, start)
    ur, uc = uniq // n, uniq % n
    row_ptr = np.zeros(n + 1, np.uint32)
    np.cumsum(np.bincount(ur, minlength=n), out=row_ptr[1:])
    return row_ptr, uc.astype(np.uint32), vsum.astype(np.float32)


def joint_probabilities(X, perplexity=30.0, k=90):
    """The whole CPU pipeline: (row_ptr, col, val) of the symmetrised P."""
    idx, d2 = knn_exact(X, k)
    Pc, _ = conditional_probabilities(d2, perplexity)
    return symmetrize(idx, Pc)
