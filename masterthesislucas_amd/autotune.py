"""Autotuner for the Barnes-Hut traversal kernel.

The traversal has a few compile-time shapes (see bh_set_traverse_variant in include/bh.h):
whether monopoles are computed only when a lane accepts or always (masked), and how many
source particles each scalar fetch brings into the particle-particle loop. For small N the walk
can also be split over several waves per sink group (bh_set_traverse_split), which multiplies the
waves of a launch at the price of re-reading the path to each wave's tree range. Which is fastest
depends on the particle distribution, N and theta, so the tuner times every variant on the
caller's own context (already holding the workload) with the library's HIP-event phase timers
and keeps the fastest. Choices are remembered per (kernel, tree mode, log2 N, theta) in a JSON
table next to this file so later runs can skip the search.
"""
from __future__ import annotations

import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuned_variants.json")


def _key(ctx, theta):
    return f"{ctx.kernel}|{ctx.tree_mode}|n{int(math.log2(max(ctx.n, 1)))}|t{theta:.3f}"


def load_table(path=TABLE):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def lookup(ctx, theta, path=TABLE):
    return load_table(path).get(_key(ctx, theta))


def time_variant(ctx, variant, theta, repeats=3):
    """Average traversal milliseconds of `variant` on the current tree."""
    ctx.set_variant(variant)
    ctx.forces_device_only(theta)  # warm the code object and caches
    ctx.profile_enable(True)
    ctx.profile_read(reset=True)
    for _ in range(repeats):
        ctx.forces_device_only(theta)
    ms, launches = ctx.profile_read(reset=True)["traversal"]
    ctx.profile_enable(False)
    return ms / max(launches, 1)


def tune(ctx, theta, repeats=3, save=False, path=TABLE):
    """Time all variants on ctx's current tree (ctx must have been built), select the fastest,
    and return (best_variant, {variant: ms}).

    The variants are timed with the walk's progress-ordered issue priority on
    (bh_set_walk_priority); the winner is timed once more without it, under the key
    "<variant>:no-priority", and the context keeps whichever was faster. The priority pays when
    all waves of a launch are resident at once and can cost a few percent when they are not."""
    times = {}
    ctx.set_priority(True)
    for v in range(ctx.variant_count()):
        times[v] = time_variant(ctx, v, theta, repeats)
    best = min(times, key=times.get)
    ctx.set_priority(False)
    off = time_variant(ctx, best, theta, repeats)
    times[f"{best}:no-priority"] = off
    ctx.set_priority(off >= times[best])
    ctx.set_variant(best)
    if save:
        table = load_table(path)
        table[_key(ctx, theta)] = {"variant": best, "priority": ctx.priority,
                                   "ms": {str(k): v for k, v in times.items()}}
        with open(path, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    return best, times


def tune_split(ctx, theta, splits=(1, 2, 4, 8), variants=None, repeats=3):
    """Time every (variant, split) pair on ctx's current tree, select the fastest and return
    ((variant, split), {"variant/split": ms})."""
    variants = range(ctx.variant_count()) if variants is None else variants
    times = {}
    for s in splits:
        ctx.set_split(s)
        for v in variants:
            times[(v, s)] = time_variant(ctx, v, theta, repeats)
    best = min(times, key=times.get)
    ctx.set_variant(best[0])
    ctx.set_split(best[1])
    return best, {f"{v}/{s}": ms for (v, s), ms in times.items()}
