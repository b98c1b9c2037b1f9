"""Host-side mirror of the reference's solver plug-in interface over the C ABI.

`NBodySolverHipBH` keeps the shape of `NBodySolver<T>` (src/nbodysolvers/cpu/nBodySolver.h:9-23):
`updateTree(points, minPos, maxPos)`, `solveNbody(points)` (returns `total`, which the C++
interface passes by reference), `getNodesBufferData(level)`, and the public knobs
`maxChildren`, `theta`, `cell_size`. `Points` stands in for `std::vector<TsnePoint2D>`
(positions and derivatives in FP64, particles/tsnePoint2D.h:6-32).

`BHContext` is the thin owner of one `bh_ctx` used by drivers, tests and the benchmark.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import BHConfig, check

_vp = C.c_void_p


def _ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):  # torch tensor (device or host)
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(_vp)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


class BHContext:
    """Owner of one solver context of the C ABI (include/bh.h)."""

    def __init__(self, n_max, kernel="gravity", tree_mode="quadtree", max_children=16,
                 fixed_depth=0, code_bits=32, G=1.0, softening=0.0, theta=0.5, sizes=None):
        """One system of up to n_max particles, or, with `sizes`, len(sizes) independent
        systems of those sizes stored one after the other (bh_create_batched)."""
        L = _lib.load()
        cfg = BHConfig()
        check(L.bh_config_default(C.byref(cfg)))
        cfg.n_max = int(n_max)
        cfg.kernel = {"gravity": _lib.KERNEL_GRAVITY, "tsne": _lib.KERNEL_TSNE}[kernel]
        cfg.tree_mode = {"quadtree": _lib.TREE_QUADTREE,
                         "morton_fixed": _lib.TREE_MORTON_FIXED,
                         "gpu_ref": _lib.TREE_GPU_REF}[tree_mode]
        cfg.max_children = int(max_children)
        cfg.fixed_depth = int(fixed_depth)
        cfg.code_bits = int(code_bits)
        cfg.G = float(G)
        cfg.softening = float(softening)
        cfg.theta = float(theta)
        self.cfg = cfg
        self.kernel = kernel
        self.tree_mode = tree_mode
        self.n = 0
        h = C.c_void_p()
        if sizes is None:
            st = L.bh_create(C.byref(h), C.byref(cfg))
        else:
            sz = np.ascontiguousarray(sizes, np.uint32)
            st = L.bh_create_batched(C.byref(h), C.byref(cfg), _ptr(sz), int(sz.size))
        self._h = h.value
        if st != _lib.BH_OK:
            msg = L.bh_last_error(self._h).decode() if self._h else ""
            if self._h:
                L.bh_destroy(self._h)
                self._h = None
            raise _lib.BHError(st, msg)

    # ------------------------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            _lib.load().bh_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _c(self, st):
        check(st, self._h)

    @property
    def handle(self):
        return self._h

    # ------------------------------------------------------------------------------------
    def set_particles(self, x, y, m=None, vx=None, vy=None):
        """Upload particles (numpy arrays on the host, or torch tensors on the device)."""
        L = _lib.load()
        if hasattr(x, "is_cuda") and x.is_cuda:
            n = int(x.numel())
            self._keep = (x, y, m, vx, vy)
            self._c(L.bh_set_particles_device(self._h, n, _ptr(x), _ptr(y), _ptr(m), _ptr(vx),
                                              _ptr(vy)))
        else:
            x, y, m, vx, vy = (_f32(a) for a in (x, y, m, vx, vy))
            n = int(x.size)
            self._c(L.bh_set_particles(self._h, n, _ptr(x), _ptr(y), _ptr(m), _ptr(vx), _ptr(vy)))
        self.n = n

    def get_particles(self):
        x, y, vx, vy = (np.empty(self.n, np.float32) for _ in range(4))
        self._c(_lib.load().bh_get_particles(self._h, _ptr(x), _ptr(y), _ptr(vx), _ptr(vy)))
        return x, y, vx, vy

    def build(self, bounds=None):
        b = None if bounds is None else np.ascontiguousarray(bounds, dtype=np.float64)
        self._c(_lib.load().bh_build(self._h, _ptr(b)))

    def forces(self, theta, want_total=True):
        ax, ay = np.empty(self.n, np.float32), np.empty(self.n, np.float32)
        tot = C.c_double(0.0)
        self._c(_lib.load().bh_forces(self._h, float(theta), _ptr(ax), _ptr(ay),
                                      C.byref(tot) if want_total else None))
        return ax, ay, tot.value

    def forces_device(self, theta, ax, ay, want_total=False):
        tot = C.c_double(0.0)
        self._c(_lib.load().bh_forces_device(self._h, float(theta), _ptr(ax), _ptr(ay),
                                             C.byref(tot) if want_total else None))
        return tot.value

    def forces_device_only(self, theta):
        """Traversal only: results stay on the device (used for timing)."""
        self._c(_lib.load().bh_forces_device(self._h, float(theta), None, None, None))

    def leapfrog_graph(self, steps, dt, theta):
        """Leapfrog steps replayed from captured HIP graphs (same results as leapfrog)."""
        self._c(_lib.load().bh_leapfrog_graph(self._h, int(steps), float(dt), float(theta)))

    def batch_info(self):
        """(number of systems, offsets[systems + 1]); (0, None) for a one-system context."""
        ns = C.c_uint32(0)
        self._c(_lib.load().bh_batch_info(self._h, C.byref(ns), None))
        if ns.value == 0:
            return 0, None
        off = np.empty(ns.value + 1, np.uint32)
        self._c(_lib.load().bh_batch_info(self._h, C.byref(ns), _ptr(off)))
        return ns.value, off

    def learned_enable(self, on=True):
        """Keep cell second moments in the builds (the learned correction needs them)."""
        self._c(_lib.load().bh_learned_enable(self._h, int(on)))

    def forces_learned(self, mlp, theta, band=0.0, want_total=False):
        """Forces with the learned correction on accepted cells with l/r >= band."""
        ax, ay = np.empty(self.n, np.float32), np.empty(self.n, np.float32)
        tot = C.c_double(0)
        self._c(_lib.load().bh_forces_learned(self._h, mlp.handle, float(theta), float(band),
                                              _ptr(ax), _ptr(ay),
                                              C.byref(tot) if want_total else None))
        return ax, ay, tot.value

    def learned_export(self, mlp=None):
        """The current tree's cell features [n_nodes x 10] and, with a network, the walk's
        coefficient records [n_nodes x 16] (bh_learned_export); (features, coef or None)."""
        nn, _, _ = self.tree_info()
        f = np.empty((nn, 10), np.float32)
        c = np.empty((nn, 16), np.float32) if mlp is not None else None
        self._c(_lib.load().bh_learned_export(self._h, mlp.handle if mlp is not None else None,
                                              _ptr(f), _ptr(c) if c is not None else None, nn))
        return f, c

    def leapfrog(self, steps, dt, theta):
        self._c(_lib.load().bh_leapfrog(self._h, int(steps), float(dt), float(theta)))

    def energy(self, theta):
        """FP64 kinetic and potential energy (the walk's potential at theta) and momentum of the
        current state (bh_energy): dict(kinetic, potential, total, px, py)."""
        k, w, px, py = (C.c_double(0.0) for _ in range(4))
        self._c(_lib.load().bh_energy(self._h, float(theta), C.byref(k), C.byref(w), C.byref(px),
                                      C.byref(py)))
        return dict(kinetic=k.value, potential=w.value, total=k.value + w.value, px=px.value,
                    py=py.value)

    def direct_fp64(self, sinks):
        """FP64 direct sum on the GPU for caller-order sink indices -> (ax, ay, total)."""
        sk = np.ascontiguousarray(sinks, dtype=np.uint32)
        ax, ay, tot = (np.empty(sk.size, np.float64) for _ in range(3))
        self._c(_lib.load().bh_direct_fp64(self._h, _ptr(sk), sk.size, _ptr(ax), _ptr(ay),
                                           _ptr(tot)))
        return ax, ay, tot

    def bounds(self):
        b = np.empty(4, np.float64)
        side = C.c_double(0)
        self._c(_lib.load().bh_get_bounds(self._h, _ptr(b), C.byref(side)))
        return b, side.value

    def quantizer(self):
        """(M, es, R) of the device-prepared quantiser of the last build (bh_get_quantizer)."""
        M, R, es = C.c_uint64(0), C.c_uint64(0), C.c_int32(0)
        self._c(_lib.load().bh_get_quantizer(self._h, C.byref(M), C.byref(es), C.byref(R)))
        return M.value, es.value, R.value

    def morton(self):
        codes = np.empty(self.n, np.uint64)
        perm = np.empty(self.n, np.uint32)
        self._c(_lib.load().bh_get_morton(self._h, _ptr(codes), _ptr(perm)))
        return codes, perm

    def tree_info(self):
        nn, ml, of = C.c_uint32(0), C.c_uint32(0), C.c_uint32(0)
        self._c(_lib.load().bh_tree_info(self._h, C.byref(nn), C.byref(ml), C.byref(of)))
        return nn.value, ml.value, of.value

    def tree(self):
        nn, _, _ = self.tree_info()
        t = dict(level=np.empty(nn, np.int32), first=np.empty(nn, np.uint32),
                 count=np.empty(nn, np.uint32), skip=np.empty(nn, np.uint32),
                 flags=np.empty(nn, np.uint8), cx=np.empty(nn, np.float32),
                 cy=np.empty(nn, np.float32), mass=np.empty(nn, np.float32))
        self._c(_lib.load().bh_get_tree(self._h, *[_ptr(t[k]) for k in (
            "level", "first", "count", "skip", "flags", "cx", "cy", "mass")]))
        return t

    def tree_slots(self):
        depth, ns = C.c_int32(0), C.c_uint64(0)
        L = _lib.load()
        self._c(L.bh_get_tree_slots(self._h, C.byref(depth), C.byref(ns), None, None, None))
        a, b, c = (np.empty(ns.value, np.uint32) for _ in range(3))
        self._c(L.bh_get_tree_slots(self._h, None, None, _ptr(a), _ptr(b), _ptr(c)))
        return depth.value, dict(first_child=a, first_particle=b, count=c)

    def counters(self):
        pn, pp = C.c_uint64(0), C.c_uint64(0)
        self._c(_lib.load().bh_counters(self._h, C.byref(pn), C.byref(pp)))
        return pn.value, pp.value

    def profile_enable(self, on=True):
        self._c(_lib.load().bh_profile_enable(self._h, 1 if on else 0))

    def profile_phases(self, names):
        """Time only the named phases (names from PHASE_NAMES); the others record no events."""
        mask = 0
        for name in names:
            mask |= 1 << _lib.PHASE_NAMES.index(name)
        self._c(_lib.load().bh_profile_phases(self._h, mask))

    def profile_read(self, reset=True):
        ms = np.zeros(_lib.NUM_PHASES, np.float64)
        la = np.zeros(_lib.NUM_PHASES, np.uint64)
        self._c(_lib.load().bh_profile_read(self._h, _ptr(ms), _ptr(la), 1 if reset else 0))
        return {name: (float(ms[i]), int(la[i])) for i, name in enumerate(_lib.PHASE_NAMES)}

    def set_variant(self, variant):
        self._c(_lib.load().bh_set_traverse_variant(self._h, int(variant)))
        self._variant = int(variant)

    @property
    def variant(self):
        """Traversal variant in use (the library starts with 4: two sinks per lane)."""
        return getattr(self, "_variant", 4)

    def set_sort_spin_limit(self, polls):
        """Look-back polls before the radix sort gives up (default 2^22); a small value only
        exercises the error path (bh_set_sort_spin_limit)."""
        self._c(_lib.load().bh_set_sort_spin_limit(self._h, int(polls)))

    SOLVERS = {"bh": 0, "fmm_sym": 1}

    def set_solver(self, solver):
        """'bh' (Barnes-Hut walk) or 'fmm_sym' (symmetric dual-tree FMM, t-SNE kernels,
        morton_fixed tree); see bh_set_solver."""
        self._c(_lib.load().bh_set_solver(self._h, self.SOLVERS[solver]))
        self._solver = solver

    def order_by_cost(self):
        """Batched contexts: heaviest work entries of the last walk first (bh_batch_order_by_cost)."""
        self._c(_lib.load().bh_batch_order_by_cost(self._h))

    def set_priority(self, on=True):
        """Progress-ordered issue priority of the walk's waves (bh_set_walk_priority)."""
        self._c(_lib.load().bh_set_walk_priority(self._h, 1 if on else 0))

    @property
    def priority(self):
        v = C.c_int(0)
        self._c(_lib.load().bh_get_walk_priority(self._h, C.byref(v)))
        return bool(v.value)

    def set_split(self, split):
        """Waves per sink group in the traversal (1..8); see bh_set_traverse_split."""
        self._c(_lib.load().bh_set_traverse_split(self._h, int(split)))
        self._split = int(split)

    @property
    def split(self):
        return getattr(self, "_split", 1)

    @staticmethod
    def variant_count():
        return int(_lib.load().bh_traverse_variant_count())

    def synchronize(self):
        self._c(_lib.load().bh_synchronize(self._h))


# ----------------------------------------------------------------------------------------
@dataclass
class Points:
    """Stand-in for std::vector<TsnePoint2D>: FP64 positions and derivatives, N x 2."""
    position: np.ndarray
    derivative: np.ndarray = field(default=None)

    def __post_init__(self):
        self.position = np.ascontiguousarray(self.position, dtype=np.float64).reshape(-1, 2)
        if self.derivative is None:
            self.derivative = np.zeros_like(self.position)

    def __len__(self):
        return self.position.shape[0]


class NBodySolverHipBH:
    """NBodySolver<T> (nBodySolver.h:9-23) implemented by the HIP Barnes-Hut step.

    kernel "tsne" reproduces NBodySolverBH with TSNEBHPNKernel/TSNEBHPPKernel
    (nBodySolverBH.h:94-116); kernel "gravity" uses the Plummer-softened monopole.
    """

    def __init__(self, kernel="tsne", max_children=16, theta=0.5, n_max=1 << 16,
                 tree_mode="quadtree", G=1.0, softening=0.0, code_bits=32):
        self.maxChildren = int(max_children)
        self.theta = float(theta)
        self.cell_size = 0.0
        self._kw = dict(kernel=kernel, tree_mode=tree_mode, max_children=max_children, G=G,
                        softening=softening, code_bits=code_bits)
        self._ctx = BHContext(n_max, **self._kw)

    @property
    def context(self) -> BHContext:
        return self._ctx

    def updateTree(self, points: Points, minPos=None, maxPos=None):
        """nBodySolver.h:18. Positions are rounded to FP32 on upload; the bounds (minPos,
        maxPos) are honoured in morton_fixed mode and ignored in quadtree mode, as in the
        reference (nBodySolverBH.h:43-46, nBodySolverFMM_SYM_MORTON.h:79-95)."""
        n = len(points)
        if n > self._ctx.cfg.n_max:
            self._ctx.close()
            self._ctx = BHContext(max(n, 2 * self._ctx.cfg.n_max), **self._kw)
        p = points.position
        self._ctx.set_particles(p[:, 0], p[:, 1])
        bounds = None
        if minPos is not None and maxPos is not None and self._kw["tree_mode"] == "morton_fixed":
            bounds = [minPos[0], minPos[1], maxPos[0], maxPos[1]]
        self._ctx.build(bounds)

    def solveNbody(self, points: Points) -> float:
        """nBodySolver.h:17: adds the interaction sums to points.derivative (the caller zeroes
        it first, tsne.h:281-285) and returns `total` (reset to 0 each call,
        nBodySolverBH.h:35)."""
        ax, ay, total = self._ctx.forces(self.theta, want_total=True)
        points.derivative[:, 0] += ax
        points.derivative[:, 1] += ay
        return total

    def getNodesBufferData(self, level: int):
        """nBodySolver.h:19 / quadtree.h:174-208: 8 vertices (4 edges) per cell of `level`
        (-1 = all levels), colour colors[min(level+1, 6)]. Returns (positions (K,2) float32,
        colours (K,3) float32)."""
        colors = np.array([[1, 1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [0, 1, 1],
                           [1, 0, 1]], np.float32)
        t = self._ctx.tree()
        codes, _ = self._ctx.morton()
        b, side = self._ctx.bounds()
        B = int(self._kw["code_bits"]) if self._kw["tree_mode"] == "quadtree" else 16
        sel = np.ones(t["level"].size, bool) if level == -1 else (t["level"] == level)
        lv = t["level"][sel].astype(np.int64)
        c = codes[t["first"][sel]]
        ix = np.zeros(lv.size, np.int64)
        iy = np.zeros(lv.size, np.int64)
        for bit in range(B):
            ix |= ((c >> np.uint64(2 * bit)) & np.uint64(1)).astype(np.int64) << bit
            iy |= ((c >> np.uint64(2 * bit + 1)) & np.uint64(1)).astype(np.int64) << bit
        cell = side / np.exp2(lv.astype(np.float64))
        x0, y0 = b[0], b[1]
        if self._kw["tree_mode"] == "gpu_ref":  # root centred on the bounding box
            x0 = float(np.float32(np.float32(b[0]) + np.float32(b[2])) * np.float32(0.5)) - 0.5 * side
            y0 = float(np.float32(np.float32(b[1]) + np.float32(b[3])) * np.float32(0.5)) - 0.5 * side
        lx = x0 + (ix >> (B - lv)) * cell
        ly = y0 + (iy >> (B - lv)) * cell
        hx, hy = lx + cell, ly + cell
        v = np.stack([lx, ly, hx, ly, lx, ly, lx, hy, lx, hy, hx, hy, hx, ly, hx, hy], 1)
        col = colors[min(level + 1, 6)]
        pos = v.reshape(-1, 2).astype(np.float32)
        return pos, np.tile(col, (pos.shape[0], 1))
