"""Batched systems (bh_create_batched) and graph-captured leapfrog steps (bh_leapfrog_graph).

A batched context builds one QuadTree per system with the same launches and traverses all
systems in one launch. Each system must come out exactly as it does alone in a one-system
context with the same code precision (24 bits per axis): same sort order, same codes, same
forces bit for bit (same waves, same walks, same summation order), and the FP64 direct sum
must stay inside each system. Graph replays must equal eager steps bit for bit.
"""
import numpy as np
import pytest

from masterthesislucas_amd import BHContext
from masterthesislucas_amd._lib import BHError
from masterthesislucas_amd.workloads import plummer, uniform_disk

pytestmark = pytest.mark.gpu

EPS = float(np.float32(1e-3))
SIZES = [3000, 1, 17, 25000, 16, 9000, 100, 2]


def _systems(sizes, seed=40):
    out = []
    for k, n in enumerate(sizes):
        if k % 2 == 0:
            p = uniform_disk(n, seed=seed + k)
        else:
            p = plummer(n, a=0.1, seed=seed + k)
        out.append(p)
    return out


def _concat(ps):
    return [np.concatenate([getattr(p, f) for p in ps]) for f in ("x", "y", "m", "vx", "vy")]


@pytest.mark.parametrize("variant", [0, 4, 8])
def test_batched_equals_each_system_alone(gpu, variant):
    ps = _systems(SIZES)
    x, y, m, vx, vy = _concat(ps)
    b = BHContext(len(x), kernel="gravity", softening=EPS, sizes=SIZES)
    b.set_variant(variant)
    b.set_particles(x, y, m, vx, vy)
    b.build()
    ax, ay, _ = b.forces(0.5, want_total=False)
    codes, perm = b.morton()
    pn_b, pp_b = b.counters()
    ns, off = b.batch_info()
    assert ns == len(SIZES) and list(np.diff(off)) == SIZES
    pn = pp = 0
    for s, p in enumerate(ps):
        a = BHContext(len(p.x), kernel="gravity", softening=EPS, code_bits=24)
        a.set_variant(variant)
        a.set_particles(p.x, p.y, p.m)
        a.build()
        sx, sy, _ = a.forces(0.5, want_total=False)
        sc, sp = a.morton()
        lo, hi = off[s], off[s + 1]
        assert np.array_equal(perm[lo:hi] - lo, sp)
        assert np.array_equal(codes[lo:hi] & np.uint64((1 << 48) - 1), sc)
        assert np.all((codes[lo:hi] >> np.uint64(48)) == s)
        assert np.array_equal(ax[lo:hi], sx) and np.array_equal(ay[lo:hi], sy), s
        c = a.counters()
        pn += c[0]
        pp += c[1]
        a.close()
    assert (pn_b, pp_b) == (pn, pp)
    b.close()


def test_many_systems_equal_each_system_alone(gpu):
    """The 256-system workload (11.1 M particles): its trees defer more top nodes than the
    upward pass keeps on chip, so they are settled through global memory, bucketed by level.
    Forces and interaction counts of sampled systems are bit-identical to each system alone
    (24 code bits, as the batched keys have)."""
    from masterthesislucas_amd.workloads import batched_systems
    ps = batched_systems()
    sizes = [len(q.x) for q in ps]
    x, y, m, vx, vy = _concat(ps)
    b = BHContext(len(x), kernel="gravity", softening=EPS, sizes=sizes)
    b.set_particles(x, y, m, vx, vy)
    b.build()
    ax, ay, _ = b.forces(0.5, want_total=False)
    ns, off = b.batch_info()
    assert ns == 256
    b.close()
    for s in (0, 1, 128, 255, int(np.argmax(sizes))):
        p = ps[s]
        a = BHContext(len(p.x), kernel="gravity", softening=EPS, code_bits=24)
        a.set_particles(p.x, p.y, p.m)
        a.build()
        sx, sy, _ = a.forces(0.5, want_total=False)
        lo, hi = off[s], off[s + 1]
        assert np.array_equal(ax[lo:hi].view(np.uint32), sx.view(np.uint32)), s
        assert np.array_equal(ay[lo:hi].view(np.uint32), sy.view(np.uint32)), s
        a.close()


def test_batched_direct_sum_stays_in_system(gpu):
    ps = _systems([4000, 3000, 50])
    x, y, m, _, _ = _concat(ps)
    b = BHContext(len(x), kernel="gravity", softening=EPS, sizes=[4000, 3000, 50])
    b.set_particles(x, y, m)
    b.build()
    sinks = np.array([0, 1234, 3999, 4000, 6999, 7000, 7049], np.uint32)
    bx, by, bt = b.direct_fp64(sinks)
    off = [0, 4000, 7000, 7050]
    for s, p in enumerate(ps):
        a = BHContext(len(p.x), kernel="gravity", softening=EPS, code_bits=24)
        a.set_particles(p.x, p.y, p.m)
        a.build()
        sel = (sinks >= off[s]) & (sinks < off[s + 1])
        rx, ry, rt = a.direct_fp64(sinks[sel] - off[s])
        assert np.allclose(bx[sel], rx, rtol=1e-12, atol=0)
        assert np.allclose(by[sel], ry, rtol=1e-12, atol=0)
        a.close()
    b.close()


def test_batched_rejects_bad_sizes(gpu):
    with pytest.raises(BHError):
        BHContext(10, kernel="gravity", sizes=[5, 0, 5])
    b = BHContext(10, kernel="gravity", sizes=[5, 5])
    with pytest.raises(BHError):
        b.set_particles(np.zeros(9, np.float32), np.zeros(9, np.float32))
    b.close()


@pytest.mark.parametrize("batched", [False, True])
def test_graph_steps_equal_eager_steps(gpu, batched):
    if batched:
        ps = _systems([5000, 700, 12000])
        x, y, m, vx, vy = _concat(ps)
        kw = dict(sizes=[5000, 700, 12000])
    else:
        p = plummer(20000, a=0.1, seed=50)
        x, y, m, vx, vy = p.x, p.y, p.m, p.vx, p.vy
        vx = vx + 0.1 * y
        vy = vy - 0.1 * x
        kw = {}
    ctxs = []
    for _ in range(2):
        c = BHContext(len(x), kernel="gravity", softening=EPS, **kw)
        c.set_particles(x, y, m, vx, vy)
        ctxs.append(c)
    eager, graph = ctxs
    # replays interleaved with host reads; an odd step count exercises both captured parities
    for steps, theta in ((1, 0.5), (2, 0.5), (3, 0.5), (1, 0.7)):
        eager.leapfrog(steps, 1e-4, theta)
        graph.leapfrog_graph(steps, 1e-4, theta)
        ge, gg = eager.get_particles(), graph.get_particles()
        for a, b in zip(ge, gg):
            assert np.array_equal(a, b), (steps, theta)
    for c in ctxs:
        c.close()


def test_order_by_cost(gpu):
    """bh_batch_order_by_cost: the work table in descending cost of the last walk, the entries
    of at least 16x the mean walked as 8 slices each (the one-child chains above untruncated
    Plummer clusters make such entries). Counts equal the system-order walk's, forces agree up
    to the slices' summation order, graph-replayed steps equal eager steps, a split > 1 restores
    the system order, and the call is refused where it does not apply."""
    from masterthesislucas_amd.workloads import batched_systems
    from tests.helpers import rel_err
    ps = batched_systems(12, seed=73)
    sizes = [len(q.x) for q in ps]
    cat = [np.concatenate([getattr(q, f) for q in ps]) for f in ("x", "y", "m", "vx", "vy")]
    n = int(sum(sizes))
    ctx = BHContext(n, kernel="gravity", softening=EPS, sizes=sizes)
    ctx.set_particles(*cat[:3])
    ctx.build()
    with pytest.raises(BHError):
        ctx.order_by_cost()  # no walk yet
    ax0, ay0, t0 = ctx.forces(0.5)
    c0 = ctx.counters()
    ctx.order_by_cost()
    ax1, ay1, t1 = ctx.forces(0.5)
    assert ctx.counters() == c0
    r = rel_err(ax1, ay1, ax0.astype(np.float64), ay0.astype(np.float64))
    assert np.median(r) == 0.0 and np.quantile(r, 0.999) <= 1e-4 and r.max() <= 1e-3, (
        np.quantile(r, 0.999), r.max())
    assert abs(t1 - t0) <= 1e-6 * abs(t0)
    ctx.order_by_cost()  # again, from the reordered table's counters: the same table
    ax2, ay2, _ = ctx.forces(0.5)
    assert np.array_equal(ax1.view(np.uint32), ax2.view(np.uint32))
    # the slices without totals (unmasked particle-particle form) add up to the same forces
    ax4, ay4, _ = ctx.forces(0.5, want_total=False)
    assert np.array_equal(ax4.view(np.uint32), ax2.view(np.uint32))
    assert np.array_equal(ay4.view(np.uint32), ay2.view(np.uint32))
    ctx.set_split(2)  # back to system order, each entry walked by 2 waves
    ax3, _, _ = ctx.forces(0.5)
    assert ctx.counters() == c0
    ctx.close()
    eager = BHContext(n, kernel="gravity", softening=EPS, sizes=sizes)
    graph = BHContext(n, kernel="gravity", softening=EPS, sizes=sizes)
    for c in (eager, graph):
        c.set_particles(*cat)
        c.build()
        c.forces(0.5)
        c.order_by_cost()
    eager.leapfrog(3, 1e-4, 0.5)
    graph.leapfrog_graph(3, 1e-4, 0.5)
    xe, ye, _, _ = eager.get_particles()
    xg, yg, _, _ = graph.get_particles()
    assert np.array_equal(xe, xg) and np.array_equal(ye, yg)
    eager.close()
    graph.close()
    one = BHContext(1000, kernel="gravity", softening=EPS)
    one.set_particles(cat[0][:1000], cat[1][:1000], cat[2][:1000])
    one.build()
    one.forces(0.5)
    with pytest.raises(BHError):
        one.order_by_cost()  # not batched
    one.close()
