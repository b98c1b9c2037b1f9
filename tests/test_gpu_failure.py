"""Failure paths of the build, forced from the host so that every kernel exits normally.

Non-finite positions: the bounds pass flags them and every call that builds returns
BH_ERR_ARG (the reference would build a meaningless tree without a word, quadtree.h:29-53).
Radix-sort look-back give-up: the wait bound is set to 0 polls from the host, so tiles that find
an earlier tile not yet published stop waiting at once; the later passes then do nothing, the
gather falls back to the input order, and the call returns BH_ERR_HIP. Afterwards, with the
default bound restored, the same context builds the bit-exact order again.
"""
import numpy as np
import pytest

from masterthesislucas_amd import BHContext, _lib
from masterthesislucas_amd._lib import BHError
from masterthesislucas_amd.workloads import two_plummer_disks, uniform_disk
from oracle import oracle as O
from tests.helpers import oracle_tree

pytestmark = pytest.mark.gpu

EPS = float(np.float32(1e-3))


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
@pytest.mark.parametrize("mode", ["quadtree", "morton_fixed", "gpu_ref"])
def test_nonfinite_position_is_rejected(gpu, bad, mode):
    p = uniform_disk(5000, seed=31)
    x = p.x.copy()
    x[1234] = bad
    ctx = BHContext(p.n, kernel="gravity", tree_mode=mode, softening=EPS)
    ctx.set_particles(x, p.y, p.m)
    with pytest.raises(BHError) as e:
        ctx.build()
    assert e.value.status == _lib.BH_ERR_ARG and "finite" in str(e.value)
    with pytest.raises(BHError) as e:  # the tree is marked unbuilt
        ctx.forces(0.5)
    assert e.value.status == _lib.BH_ERR_STATE
    ctx.set_particles(p.x, p.y, p.m)  # a finite upload builds again
    ctx.build()
    ax, ay, _ = ctx.forces(0.5)
    assert np.all(np.isfinite(ax)) and np.all(np.isfinite(ay))
    ctx.close()


def test_nonfinite_y_in_leapfrog_and_batched(gpu):
    p = uniform_disk(3000, seed=32)
    y = p.y.copy()
    y[7] = np.nan
    ctx = BHContext(p.n, kernel="gravity", softening=EPS)
    ctx.set_particles(p.x, y, p.m)
    with pytest.raises(BHError) as e:
        ctx.leapfrog(2, 1e-4, 0.5)
    assert e.value.status == _lib.BH_ERR_ARG
    ctx.close()
    sizes = [1000, 2000]
    ctx = BHContext(sum(sizes), kernel="gravity", softening=EPS, sizes=sizes)
    x = np.concatenate([p.x[:1000], p.x[:2000]])
    yy = np.concatenate([p.y[:1000], p.y[:2000]])
    x[1500] = np.inf  # in the second system
    ctx.set_particles(x, yy)
    with pytest.raises(BHError) as e:
        ctx.build()
    assert e.value.status == _lib.BH_ERR_ARG
    ctx.close()


def test_sort_give_up_is_reported_then_recovers(gpu):
    n = 1 << 20
    p = two_plummer_disks(n, seed=33)
    ctx = BHContext(n, kernel="gravity", softening=EPS)
    ctx.set_particles(p.x, p.y, p.m)
    ctx.set_sort_spin_limit(0)
    raised = 0
    for _ in range(3):
        try:
            ctx.build()
        except BHError as e:
            assert e.status == _lib.BH_ERR_HIP and "look-back" in str(e)
            raised += 1
    # 256 tiles per pass started together: some tile always finds its predecessor unpublished
    assert raised >= 1
    with pytest.raises(BHError) as e:
        ctx.forces(0.5)
    assert e.value.status == _lib.BH_ERR_STATE
    ctx.set_sort_spin_limit(1 << 22)
    ctx.build()
    codes, perm = ctx.morton()
    t = oracle_tree(p.x, p.y, p.m, mode="quadtree")
    assert np.array_equal(perm, t["perm"])
    assert np.array_equal(codes, t["codes_sorted"])
    ctx.close()
