#!/usr/bin/env python3
"""Benchmark of the Barnes-Hut force step on one MI355X.

Workload: two Plummer disks (scale radius 0.1, 524,288 particles each) centred at (+-0.5, 0) on
counter-rotating circular orbits, N = 1,048,576 FP32 particles, G = 1, Plummer softening 1e-3,
QuadTree tree (leaf capacity 16), theta = 0.5, kick-drift-kick leapfrog with dt = 1e-4. One
step = bounds, Morton codes, radix sort, gather, tree build, upward pass, traversal and the
leapfrog update, with the particle state resident in HBM. Metric: force evaluations per second
(particles x steps / wall time).

Prints one JSON line. --gpus above 1 is not supported: the benchmark runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_INTERACTION = {"gravity": 14, "tsne": 15}
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(p, theta, eps, n):
    """The reference CPU Barnes-Hut (NBodySolverBH + QuadTree, compiled from the reference
    sources into oracle/_ref) timed on this host, one core: one updateTree + solveNbody step on
    the benchmark's initial state. Falls back to the oracle's C restatement if the compiled
    reference is absent."""
    from oracle import oracle as O

    x = p.x.astype(np.float64)
    y = p.y.astype(np.float64)
    if O.ref_available():
        _, _, _, tb, ts = O.ref_bh_solve(x, y, O.KERNEL_GRAVITY, G=1.0, m=1.0 / n, eps2=eps * eps,
                                         theta=theta, max_children=16)
        kind = "reference"
        secs = tb + ts
    else:
        t0 = time.perf_counter()
        qt = O.QuadTree(x, y, 16)
        qt.forces(O.KERNEL_GRAVITY, G=1.0, m=1.0 / n, eps2=eps * eps, theta=theta, nthreads=1)
        secs = time.perf_counter() - t0
        kind = "port"
    # all host cores (up to 16, this host's share of the machine): the oracle's restatement of
    # the same QuadTree + traversal, threaded over sinks
    cores = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    qt = O.QuadTree(x, y, 16)
    qt.forces(O.KERNEL_GRAVITY, G=1.0, m=1.0 / n, eps2=eps * eps, theta=theta, nthreads=cores)
    secs_all = time.perf_counter() - t0
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": n / secs, "unit": "force evals/s", "cores": 1, "kind": kind,
            "seconds": secs, "host_cpu": model, "host_nproc": os.cpu_count(),
            "sample": f"one updateTree + solveNbody step (QuadTree, leaf capacity 16, "
                      f"theta={theta}) over all {n} particles of the benchmark's initial state, "
                      f"FP64, single thread",
            "all_cores": {"value": n / secs_all, "cores": cores, "kind": "port",
                          "seconds": secs_all,
                          "sample": "the same step, tree build serial, traversal threaded over "
                                    "sinks"}}


def learned_comparison(args, eps):
    """Learned far-field correction vs plain Barnes-Hut: a 2D Plummer disk (a = 1),
    N = args.learned_n. The network is trained on pairs labelled from a second realisation of
    the distribution (different seed), collected at each learned theta of the sweep, and
    evaluated on the first realisation. Errors are medians of the relative force error against
    the FP64 direct sum on 4,096 seeded sinks. A step is build + forces: for plain,
    N / (build + walk); for learned, N / (build + cell features + per-cell inference + corrected
    walk), all from the library's HIP-event phase timers."""
    from masterthesislucas_amd import BHContext
    from masterthesislucas_amd.autotune import time_variant
    from masterthesislucas_amd.learned import FarFieldMLP
    from masterthesislucas_amd.workloads import SEED_PLUMMER, plummer, sample_indices

    n = args.learned_n
    thetas = sorted({float(t) for t in args.learned_thetas.split(",")} | {args.learned_theta})
    ctx = BHContext(n, kernel="gravity", tree_mode="quadtree", softening=eps)
    ctx.learned_enable(True)
    t0 = time.perf_counter()
    log(f"[bench] learned: training system n={n}")
    tr = plummer(n, a=1.0, seed=SEED_PLUMMER + 1)
    ctx.set_particles(tr.x, tr.y, tr.m)
    ctx.build()
    net = FarFieldMLP(max_samples=args.learned_pairs + 4096 * 1024, seed=SEED_PLUMMER)
    net.terms = args.learned_terms
    added, rnd = 0, 0
    for k, th in enumerate(thetas):  # an equal share of the pair budget at every theta
        goal = args.learned_pairs * (k + 1) // len(thetas)
        while added < goal and rnd < 256:
            sinks = sample_indices(n, 4096, seed=1000 + rnd).astype(np.uint32)
            added += net.collect(ctx, th, sinks, band=args.learned_band, max_pairs_per_sink=1024)
            rnd += 1
    t_collect = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = net.train(epochs=args.learned_epochs, batch=4096, lr=2e-3, seed=7)
    t_train = time.perf_counter() - t0
    log(f"[bench] learned: {added} pairs ({t_collect:.1f} s), trained {args.learned_epochs} "
        f"epochs in {t_train:.1f} s, loss {loss:.3e}")
    ev = plummer(n, a=1.0, seed=SEED_PLUMMER)
    ctx.set_particles(ev.x, ev.y, ev.m)
    ctx.profile_enable(True)
    ctx.profile_read(reset=True)
    for _ in range(3):
        ctx.build()
    ph = ctx.profile_read(reset=True)
    build_ms = sum(ph[k][0] for k in ("bounds+morton", "sort", "gather", "tree", "upward")) / 3
    sample = sample_indices(n, 4096).astype(np.uint32)
    dx, dy, _ = ctx.direct_fp64(sample)
    dn = np.hypot(dx, dy)

    def med(ax, ay):
        return float(np.median(np.hypot(ax[sample] - dx, ay[sample] - dy) / dn))

    plain = []
    for th in (0.3, 0.4, 0.5, 0.6, 0.7, 0.9, 1.0):
        ms = time_variant(ctx, ctx.variant, th, repeats=3)
        e = med(*ctx.forces(th, want_total=False)[:2])
        plain.append({"theta": th, "traversal_ms": ms, "median_rel": e,
                      "force_evals_per_s": n / ((ms + build_ms) / 1e3)})

    def plain_at(err):
        """plain theta and rate with median error err (linear interpolation in the sweep)"""
        for a, b in zip(plain, plain[1:]):
            if a["median_rel"] <= err <= b["median_rel"]:
                f = (err - a["median_rel"]) / max(b["median_rel"] - a["median_rel"], 1e-30)
                return {"theta": a["theta"] + f * (b["theta"] - a["theta"]),
                        "force_evals_per_s": a["force_evals_per_s"]
                        + f * (b["force_evals_per_s"] - a["force_evals_per_s"])}
        return None

    learned = []
    for th in thetas:
        ctx.forces_learned(net, th, band=args.learned_band)  # warm
        ctx.profile_enable(True)
        ctx.profile_read(reset=True)
        for _ in range(3):  # a learned step: build, cell features, per-cell inference, walk
            ctx.build()
            ctx.forces_learned(net, th, band=args.learned_band)
        ph = ctx.profile_read(reset=True)
        ctx.profile_enable(False)
        step_ms = sum(v[0] for v in ph.values()) / 3
        e = med(*ctx.forces_learned(net, th, band=args.learned_band)[:2])
        rate = n / (step_ms / 1e3)
        eq = plain_at(e)
        learned.append({"theta": th, "step_ms": step_ms,
                        "walk_ms": ph["traversal"][0] / 3,
                        "median_rel": e, "force_evals_per_s": rate,
                        "plain_at_equal_error": eq,
                        "speedup_vs_plain_at_equal_error":
                            rate / eq["force_evals_per_s"] if eq else None})
    p05 = next(r for r in plain if r["theta"] == 0.5)
    head = next(r for r in learned if r["theta"] == args.learned_theta)
    # the learned theta whose error equals plain theta = 0.5's (interpolated in the sweep)
    match = None
    for a, b in zip(learned, learned[1:]):
        if a["median_rel"] <= p05["median_rel"] <= b["median_rel"]:
            f = (p05["median_rel"] - a["median_rel"]) / max(b["median_rel"] - a["median_rel"],
                                                            1e-30)
            rate = a["force_evals_per_s"] + f * (b["force_evals_per_s"] - a["force_evals_per_s"])
            match = {"theta": a["theta"] + f * (b["theta"] - a["theta"]),
                     "force_evals_per_s": rate,
                     "speedup_vs_plain_theta_0.5": rate / p05["force_evals_per_s"]}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = learned_cpu_baseline(ev, eps)
        log(f"[bench] learned workload CPU baseline: {cpu['seconds']:.2f} s")
    out = {"workload": f"2D Plummer disk a=1.0, N={n}, eps={eps}",
           "theta_learned": args.learned_theta, "terms": net.terms, "band": args.learned_band,
           "training_pairs": added, "training_thetas": thetas,
           "training_epochs": args.learned_epochs, "train_loss": loss,
           "collect_s": t_collect, "train_s": t_train, "build_ms": build_ms,
           "learned_traversal_ms": head["walk_ms"], "learned_step_ms": head["step_ms"],
           "learned_median_rel": head["median_rel"],
           "learned_force_evals_per_s": head["force_evals_per_s"],
           "plain": plain, "learned_sweep": learned,
           "plain_at_equal_error": head["plain_at_equal_error"],
           "speedup_vs_plain_at_equal_error": head["speedup_vs_plain_at_equal_error"],
           "speedup_vs_plain_theta_0.5": head["force_evals_per_s"] / p05["force_evals_per_s"],
           "error_ratio_vs_plain_theta_0.5": head["median_rel"] / p05["median_rel"],
           "learned_at_plain_theta_0.5_error": match,
           "cpu_baseline": cpu}
    net.close()
    ctx.close()
    return out


def learned_cpu_baseline(p, eps, theta=0.5):
    """Plain Barnes-Hut at theta on the learned workload's evaluation system on this host: the
    oracle's C restatement of the reference QuadTree and walk (FP64), tree build serial and
    traversal threaded over sinks on all host cores (up to 16)."""
    from oracle import oracle as O

    x = p.x.astype(np.float64)
    y = p.y.astype(np.float64)
    n = x.size
    cores = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    qt = O.QuadTree(x, y, 16)
    t1 = time.perf_counter()
    qt.forces(O.KERNEL_GRAVITY, G=1.0, m=1.0 / n, eps2=eps * eps, theta=theta, nthreads=cores)
    t2 = time.perf_counter()
    return {"value": n / (t2 - t0), "unit": "force evals/s", "cores": cores, "kind": "port",
            "seconds": t2 - t0, "tree_s": t1 - t0, "walk_s": t2 - t1,
            "sample": f"one plain step (QuadTree build + walk, theta={theta}) over all {n} "
                      f"particles of the evaluation system, FP64"}


def batched_comparison(args, eps, theta, dt, rmax=0.0, variant=None, split=None):
    """Batched systems: `args.batched_systems` independent systems with N log-uniform in
    [1,024, 262,144] (even: uniform disks, odd: Plummer a = 0.1, untruncated unless rmax > 0),
    all in one context (one set of build launches, one traversal launch per step). Leapfrog
    steps timed eagerly and replayed from captured HIP graphs; the traversal variant is
    autotuned on the batched tree unless given."""
    from masterthesislucas_amd import BHContext
    from masterthesislucas_amd.autotune import tune
    from masterthesislucas_amd.workloads import batched_systems

    ps = batched_systems(args.batched_systems, plummer_rmax=rmax)
    sizes = [len(p.x) for p in ps]
    cat = [np.concatenate([getattr(p, f) for p in ps]) for f in ("x", "y", "m", "vx", "vy")]
    n = int(sum(sizes))
    log(f"[bench] batched: {len(sizes)} systems, {n} particles")
    ctx = BHContext(n, kernel="gravity", softening=eps, sizes=sizes)
    ctx.set_particles(*cat)
    ctx.build()
    times = {}
    if variant is None:
        from masterthesislucas_amd.autotune import time_variant

        # the variants are timed on cost-ordered work tables, as the steps will run: one walk
        # and bh_batch_order_by_cost per table (variants 0, 4, 8: 1, 2, 4 sinks per lane)
        for v0 in (0, 4, 8):
            ctx.set_variant(v0)
            ctx.forces_device_only(theta)
            ctx.order_by_cost()
        variant, times = tune(ctx, theta)
        # then the split (each work entry walked by F waves, each taking a slice of every
        # particle-particle cell) for the chosen variant
        best_ms = min(times[variant], times.get(f"{variant}:no-priority", times[variant]))
        split = 1
        for F in (2, 4, 8):
            ctx.set_split(F)
            ms = time_variant(ctx, variant, theta)
            times[f"{variant}/split{F}"] = ms
            if ms < best_ms:
                split, best_ms = F, ms
    else:
        ctx.set_variant(variant)
    ctx.set_split(1)  # (a split above 1 restored the system-order tables)
    # walk the heaviest work entries first (bh_batch_order_by_cost, from one walk's counters)
    ctx.forces_device_only(theta)
    ctx.order_by_cost()
    ctx.set_split(split or 1)
    ctx.set_particles(*cat)
    K = args.batched_steps
    ctx.leapfrog(2, dt, theta)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.leapfrog(K, dt, theta)
    ctx.synchronize()
    eager = (time.perf_counter() - t0) / K
    ctx.leapfrog_graph(2, dt, theta)  # capture + warm
    t0 = time.perf_counter()
    ctx.leapfrog_graph(K, dt, theta)
    graph = (time.perf_counter() - t0) / K
    pn, pp = ctx.counters()
    prio = ctx.priority
    ctx.close()
    cpu = None
    if not args.no_cpu_baseline and rmax == 0.0:
        cpu = batched_cpu_baseline(ps, eps, theta)
        log(f"[bench] batched CPU baseline: {cpu['seconds']:.2f} s for {cpu['particles']} "
            f"particles")
    plum = f"Plummer a=0.1 truncated at R={rmax}" if rmax > 0 else "Plummer a=0.1 untruncated"
    return {"workload": f"{len(sizes)} systems, N log-uniform in [1024, 262144] (seeded), "
                        f"even uniform disk / odd {plum}, eps={eps}, theta={theta}",
            "systems": len(sizes), "particles": n, "n_min": int(min(sizes)),
            "n_max": int(max(sizes)), "steps": K, "traverse_variant": variant,
            "traverse_split": split or 1, "walk_priority": prio, "ordered_by_cost": True,
            "autotune_ms": {str(k): round(v, 4) for k, v in times.items()},
            "eager_ms_per_step": eager * 1e3, "graph_ms_per_step": graph * 1e3,
            "particle_steps_per_s": n / graph, "interactions_per_particle": (pn + pp) / n,
            "cpu_baseline": cpu}


def batched_cpu_baseline(ps, eps, theta, budget=1_000_000):
    """Plain Barnes-Hut on this host for the batched workload, bounded: the first systems in
    the seeded order until `budget` particles, each with the oracle's C restatement of the
    reference QuadTree and walk (FP64, tree build serial, walk threaded over all host cores up
    to 16); particle-steps/s over the sample."""
    from oracle import oracle as O

    cores = max(1, min(16, os.cpu_count() or 1))
    tot, secs, used = 0, 0.0, 0
    for p in ps:
        if tot >= budget:
            break
        x, y = p.x.astype(np.float64), p.y.astype(np.float64)
        t0 = time.perf_counter()
        qt = O.QuadTree(x, y, 16)
        qt.forces(O.KERNEL_GRAVITY, G=1.0, m=1.0 / p.n, eps2=eps * eps, theta=theta,
                  nthreads=cores)
        secs += time.perf_counter() - t0
        tot += p.n
        used += 1
    return {"value": tot / secs, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "seconds": secs, "particles": tot, "systems": used,
            "sample": f"one plain step (QuadTree build + walk, theta={theta}) of the first {used} "
                      f"systems ({tot} particles), FP64"}


def tsne_run(args):
    """t-SNE Barnes-Hut gradient descent, N = args.tsne_n: a seeded 10-component isotropic
    Gaussian mixture in 784-D (means ~ N(0, 4 I), sigma 1), perplexity 30, exact k = 90 kNN,
    symmetrised P in FP32, embedding N(0, 1e-4 I), early exaggeration 12 (48 with the gradient's
    factor 4) for 250 of args.tsne_iters iterations then 4, theta 0.5, KL in FP64. The reference
    publishes about 0.23 s per iteration (tree + repulsion, one CPU core, MNIST N = 70,000)."""
    import torch

    from masterthesislucas_amd import tsne as T
    from masterthesislucas_amd.workloads import SEED_TSNE

    n = args.tsne_n
    t0 = time.perf_counter()
    X, _ = T.gaussian_mixture(n, dim=784, components=10, seed=SEED_TSNE)
    rp, col, val = T.joint_probabilities(X, perplexity=30.0, k=90, device="cuda")
    del X
    torch.cuda.empty_cache()
    t_p = time.perf_counter() - t0
    y0 = T.initial_embedding(n, seed=SEED_TSNE, std=1e-2)
    run = T.TSNE(rp, col, val, y0, min_theta=0.5, max_theta=0.5, exaggeration=(48.0, 4.0))
    kl0, _ = run.costFunction()
    # warm, then tune the walk variant and split per phase of the run: on the trees of two
    # snapshots inside the early-exaggeration phase (iterations 60 and 180: compact clusters,
    # crowded leaves) for the first `switch` iterations, and of three later ones (260, mid-run,
    # end) for the rest; the walks change with the embedding's spread, so each choice minimises
    # the summed time of its phase's snapshots
    switch = min(250, args.tsne_iters)
    if args.tsne_split > 0:
        run.run(args.tsne_iters)
        best = (args.variant if args.variant >= 0 else 4, args.tsne_split)
        best_early, tuning, tuning_early = best, None, None
    else:
        from masterthesislucas_amd.autotune import tune_split
        late = {min(260, args.tsne_iters), args.tsne_iters // 2, args.tsne_iters}
        early = {min(60, switch), min(180, switch)} - late if switch > 0 else set()
        tuning, tuning_early, done = {}, {}, 0
        for stop in sorted(early | late):
            run.run(stop - done)
            done = stop
            _, times = tune_split(run.ctx, 0.5, splits=(1, 2, 4, 8))
            acc = tuning_early if stop in early else tuning
            for key, ms in times.items():
                acc[key] = acc.get(key, 0.0) + ms
        v, s = min(tuning, key=tuning.get).split("/")
        best = (int(v), int(s))
        if tuning_early:
            v, s = min(tuning_early, key=tuning_early.get).split("/")
            best_early = (int(v), int(s))
        else:
            best_early = best
    log(f"[bench] t-SNE walk: variant {best_early[0]}, split {best_early[1]} for the first "
        f"{switch} iterations, then variant {best[0]}, split {best[1]}")
    run.close()
    def make():
        r = T.TSNE(rp, col, val, y0, min_theta=0.5, max_theta=0.5, exaggeration=(48.0, 4.0))
        r.ctx.set_split(max(best_early[1], best[1]))  # the split buffer allocated up front
        r.ctx.set_variant(best_early[0])
        r.ctx.set_split(best_early[1])
        return r

    def drive(r):
        r.run(switch)
        r.ctx.set_variant(best[0])
        r.ctx.set_split(best[1])
        r.run(args.tsne_iters - switch)

    run = make()
    t0 = time.perf_counter()
    drive(run)
    wall = time.perf_counter() - t0
    pn, pp = run.ctx.counters()  # last iteration's repulsion (before the KL rebuilds the tree)
    kl1, z = run.costFunction()
    out = {"workload": f"Gaussian mixture 10 x 784-D, N={n}, perplexity 30, kNN k=90, theta 0.5",
           "n": n, "nnz_P": int(col.size), "p_matrix_s": t_p, "iterations": args.tsne_iters,
           "ms_per_iteration": 1e3 * wall / args.tsne_iters,
           "iterations_per_s": args.tsne_iters / wall,
           "interactions_per_point": (pn + pp) / n,
           "traversal_variant": best[0], "traversal_split": best[1],
           "early_phase": {"iterations": switch, "traversal_variant": best_early[0],
                           "traversal_split": best_early[1],
                           "tuning_ms": ({k: round(v, 4) for k, v in tuning_early.items()}
                                         if tuning_early else None)},
           "tuning_ms": ({k: round(v, 4) for k, v in tuning.items()} if tuning else None),
           "kl_initial": kl0, "kl_final": kl1, "z_final": z,
           "reference_cpu_s_per_iteration_published": 0.23}
    y_end = run.embeddedPoints
    run.close()
    ph = replay_phases(make, args.tsne_iters, drive)
    rep_ms = sum(ph[k][0] for k in ("bounds+morton", "sort", "gather", "tree", "upward",
                                     "traversal")) / args.tsne_iters
    out["repulsion_ms_per_iteration"] = rep_ms
    out["phase_ms_per_iteration"] = {k: round(v[0] / args.tsne_iters, 5) for k, v in ph.items()}
    out["repulsion_force_evals_per_s"] = n / (rep_ms / 1e3)
    # against a figure published for other hardware and other data (MNIST), not measured here
    out["ratio_vs_published_cpu_figure_other_data"] = 0.23 / (rep_ms / 1e3)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = tsne_cpu_baseline(T, rp, col, val, y_end)
        out["gpu_over_cpu_repulsion_final_embedding"] = (
            out["cpu_baseline"]["repulsion_s"] * 1e3 / out["cpu_baseline"]["gpu_repulsion_ms"])
        log(f"[bench] t-SNE CPU baseline: repulsion {out['cpu_baseline']['repulsion_s']:.3f} s "
            f"(reference, one core) vs GPU {out['cpu_baseline']['gpu_repulsion_ms']:.3f} ms")
    if args.tsne_nmae_iters > 0:
        out["nmae_trajectory"] = tsne_nmae_trajectory(args, T, rp, col, val, y0)
    if not args.no_tsne_fmm:
        out["fmm_sym"] = tsne_fmm_run(args, T, rp, col, val, y0, y_end)
    out["gpu_reference_step"] = tsne_gpu_reference_run(args, T, rp, col, val, y0)
    return out


def tsne_cpu_baseline(T, rp, col, val, y_end, theta=0.5):
    """The t-SNE gradient on this host at the run's final embedding: the repulsion by the
    reference's own CPU NBodySolverBH with its t-SNE kernels (TSNEBHPNKernel / TSNEBHPPKernel,
    compiled from the reference sources into oracle/_ref; the oracle's C restatement where that
    is absent), one core, and the sparse attraction by the oracle's restatement, one core. The
    GPU side is the library's tree build and repulsion walk on the same embedding."""
    from masterthesislucas_amd import BHContext
    from oracle import oracle as O

    x = np.ascontiguousarray(y_end[:, 0]).astype(np.float64)
    y = np.ascontiguousarray(y_end[:, 1]).astype(np.float64)
    n = x.size
    if O.ref_available():
        _, _, _, tb, ts = O.ref_bh_solve(x, y, O.KERNEL_TSNE, theta=theta, max_children=16)
        rep_s, kind = tb + ts, "reference"
    else:
        t0 = time.perf_counter()
        O.QuadTree(x, y, 16).forces(O.KERNEL_TSNE, theta=theta, nthreads=1)
        rep_s, kind = time.perf_counter() - t0, "port"
    t0 = time.perf_counter()
    O.tsne_attraction(x, y, rp, col, val, 1.0)
    att_s = time.perf_counter() - t0
    ctx = BHContext(n, kernel="tsne", tree_mode="quadtree", max_children=16)
    ctx.set_particles(y_end[:, 0].astype(np.float32), y_end[:, 1].astype(np.float32))
    ctx.build()
    ctx.forces_device_only(theta)
    ctx.profile_enable(True)
    ctx.profile_read(reset=True)
    for _ in range(5):
        ctx.build()
        ctx.forces_device_only(theta)
    ph = ctx.profile_read(reset=True)
    ctx.close()
    gpu_ms = sum(v[0] for v in ph.values()) / 5
    return {"value": n / (rep_s + att_s), "unit": "t-SNE gradient evals/s", "cores": 1,
            "kind": kind, "repulsion_s": rep_s, "attraction_s": att_s, "attraction_kind": "port",
            "gpu_repulsion_ms": gpu_ms,
            "sample": f"one gradient (tree build + Barnes-Hut repulsion at theta={theta}, "
                      f"sparse attraction) at the final embedding of the {n}-point run, FP64"}


def tsne_nmae_trajectory(args, T, rp, col, val, y0, theta=0.5):
    """The harness' accuracy protocol (tsneTests.h:781-1085) over args.tsne_nmae_iters
    iterations: the embedding is advanced with a near-exact gradient (the walk at theta 0.02,
    FP32 with compensated sums), and at every iteration the repulsion at `theta` (tree build +
    walk, timed) is compared with the FP64 direct sum over every point by the NMAE of
    tsneTests.h:143-165; the mean NMAE and mean time over the iterations are reported, and the
    advancing gradient's own NMAE is recorded as a check."""
    import torch

    from masterthesislucas_amd import BHContext

    n = args.tsne_n
    adv = T.TSNE(rp, col, val, y0, min_theta=0.02, max_theta=0.02, exaggeration=(48.0, 4.0))
    ev = BHContext(n, kernel="tsne", tree_mode="quadtree", max_children=16)
    nm, nm_adv, tms = [], [], []
    for j in range(args.tsne_nmae_iters):
        yj = adv.embeddedPoints
        ev.set_particles(np.ascontiguousarray(yj[:, 0]), np.ascontiguousarray(yj[:, 1]))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev.build()
        ev.forces_device_only(theta)
        ev.synchronize()
        tms.append(time.perf_counter() - t0)
        nm.append(T.repulsion_nmae(ev, theta)[0])
        if j % 25 == 0:
            nm_adv.append(T.repulsion_nmae(ev, 0.02)[0])
        adv.timeStep()
    adv.close()
    ev.close()
    log(f"[bench] t-SNE NMAE trajectory: mean {np.mean(nm):.4g} over {len(nm)} iterations")
    return {"iterations": len(nm), "theta": theta, "mean_nmae": float(np.mean(nm)),
            "nmae_first": float(nm[0]), "nmae_last": float(nm[-1]),
            "mean_tree_and_repulsion_ms": 1e3 * float(np.mean(tms)),
            "advancing_walk_theta": 0.02, "advancing_walk_max_nmae": float(np.max(nm_adv)),
            "protocol": "tsneTests.h:781-1085 (naive-advanced, NMAE per iteration averaged)"}


def replay_phases(make, iters, drive=None):
    """Per-phase times of a second run identical to a timed one (the walk, the FMM and every
    reduction sum in a fixed order, so it retraces the same embeddings) with every phase
    bracketed by HIP events; the timed run records none, since the events delay each next
    kernel's start. `drive(run)` runs the iterations as the timed run did (default: all of
    them in one call)."""
    run = make()
    run.ctx.profile_enable(True)
    run.ctx.profile_read(reset=True)
    if drive is None:
        run.run(iters)
    else:
        drive(run)
    ph = run.ctx.profile_read(reset=True)
    run.close()
    return ph


def tsne_gpu_reference_run(args, T, rp, col, val, y0):
    """The reference's GPU t-SNE step (TsneGpu + NBodySolverGpuBH: complete depth-6 tree with
    the centred root, theta 1, learning rate 1000, momentum 0.2, 1/Z, no exaggeration) on the
    same P and start, args.tsne_iters iterations."""
    run = T.TsneGpu(rp, col, val, y0)
    run.run(5)
    run.close()
    run = T.TsneGpu(rp, col, val, y0)
    t0 = time.perf_counter()
    run.run(args.tsne_iters)
    wall = time.perf_counter() - t0
    kl1, _ = run.costFunction()
    run.close()
    ph = replay_phases(lambda: T.TsneGpu(rp, col, val, y0), args.tsne_iters)
    log(f"[bench] t-SNE (reference GPU step): {1e3 * wall / args.tsne_iters:.3f} ms/iter")
    return {"workload": "same P and start, depth-6 centred tree, theta 1, lr 1000, momentum 0.2",
            "ms_per_iteration": 1e3 * wall / args.tsne_iters,
            "phase_ms_per_iteration": {k: round(v[0] / args.tsne_iters, 5) for k, v in ph.items()},
            "kl_final": kl1}


def tsne_fmm_run(args, T, rp, col, val, y0, y_end):
    """The same t-SNE run with the symmetric dual-tree FMM repulsion (fixed-depth Morton tree,
    theta 0.5), then both repulsion solvers on the Barnes-Hut run's final embedding: walk time
    (tree built beforehand), median relative error against the FP64 direct sum over a seeded
    sample of sinks, and the harness' NMAE over every point (each side scaled by 4/Z of its own
    solver, tsneTests.h:143-165)."""
    from masterthesislucas_amd import BHContext
    from masterthesislucas_amd.workloads import sample_indices

    n = args.tsne_n
    run = T.TSNE(rp, col, val, y0, min_theta=0.5, max_theta=0.5, exaggeration=(48.0, 4.0),
                 solver="fmm_sym")
    run.run(5)
    run.close()
    def make():
        return T.TSNE(rp, col, val, y0, min_theta=0.5, max_theta=0.5, exaggeration=(48.0, 4.0),
                      solver="fmm_sym")

    run = make()
    t0 = time.perf_counter()
    run.run(args.tsne_iters)
    wall = time.perf_counter() - t0
    kl1, _ = run.costFunction()
    run.close()
    ph = replay_phases(make, args.tsne_iters)
    rep_ms = sum(ph[k][0] for k in ("bounds+morton", "sort", "gather", "tree", "upward",
                                     "traversal")) / args.tsne_iters
    x = np.ascontiguousarray(y_end[:, 0])
    y = np.ascontiguousarray(y_end[:, 1])
    sample = sample_indices(n, args.error_sample)
    walk, err, nmae, ref = {}, {}, {}, None
    for name, mode in (("bh", "quadtree"), ("fmm_sym", "morton_fixed")):
        ctx = BHContext(n, kernel="tsne", tree_mode=mode)
        if name != "bh":
            ctx.set_solver(name)
        ctx.set_particles(x, y)
        ctx.build()
        nmae[name] = T.repulsion_nmae(ctx, 0.5)[0]
        ax, ay, _ = ctx.forces(0.5)
        if ref is None:
            ref = ctx.direct_fp64(sample)
        for _ in range(3):
            ctx.forces_device_only(0.5)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(20):
            ctx.forces_device_only(0.5)
        ctx.synchronize()
        walk[name] = (time.perf_counter() - t0) / 20 * 1e3
        dx, dy = ref[0], ref[1]
        e = np.hypot(ax[sample] - dx, ay[sample] - dy) / np.maximum(np.hypot(dx, dy), 1e-300)
        err[name] = float(np.median(e))
        ctx.close()
    cpu = None
    if not args.no_cpu_baseline:
        cpu = fmm_cpu_baseline(x, y)
        log(f"[bench] FMM CPU restatement: {cpu['seconds']:.2f} s per evaluation")
    log(f"[bench] t-SNE FMM: {1e3 * wall / args.tsne_iters:.3f} ms/iter, walk "
        f"{walk['fmm_sym']:.3f} ms vs BH {walk['bh']:.3f} ms, error {err['fmm_sym']:.2e} vs "
        f"{err['bh']:.2e}, NMAE {nmae['fmm_sym']:.2e} vs {nmae['bh']:.2e}")
    return {"workload": "same P and start, symmetric dual-tree FMM repulsion, theta 0.5",
            "ms_per_iteration": 1e3 * wall / args.tsne_iters,
            "repulsion_ms_per_iteration": rep_ms,
            "phase_ms_per_iteration": {k: round(v[0] / args.tsne_iters, 5) for k, v in ph.items()},
            "kl_final": kl1,
            "final_embedding_walk_ms": {k: round(v, 4) for k, v in walk.items()},
            "final_embedding_median_rel_error": err,
            "final_embedding_nmae": nmae,
            "cpu_baseline": cpu}


def fmm_cpu_baseline(x32, y32, theta=0.5):
    """The oracle's FP64 restatement of NBodySolverFMM_SYM_MORTON (the reference's Morton solvers
    need boost and cannot be compiled here) timed on this host, one thread: Morton codes, sort,
    fixed-depth tree and the FMM evaluation on the given points."""
    from oracle import oracle as O

    x = np.asarray(x32, np.float32).astype(np.float64)
    y = np.asarray(y32, np.float32).astype(np.float64)
    t0 = time.perf_counter()
    minx, miny, maxx, maxy = x.min(), y.min(), x.max(), y.max()
    side = max(maxx + (-minx), maxy + (-miny))
    perm = O.stable_sort(O.morton32(x, y, minx, miny, side).astype(np.uint64))
    xs, ys = x[perm], y[perm]
    depth = O.get_depth(16, x.size)
    tree = O.fixed_tree(xs, ys, depth, minx, miny, maxx, maxy)
    O.fmm_sym(xs, ys, tree, depth, theta)
    secs = time.perf_counter() - t0
    return {"value": x.size / secs, "unit": "force evals/s", "cores": 1, "kind": "port",
            "seconds": secs,
            "sample": f"one FMM evaluation (theta={theta}) on the {x.size}-point final embedding, "
                      f"FP64, single thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--dt", type=float, default=1e-4)
    ap.add_argument("--code-bits", type=int, default=32)
    ap.add_argument("--error-sample", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="run a plain device copy of this many MiB first (PMC traffic calibration)")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the theta sweep (traversal time and error at 0.3/0.5/0.7/1.0)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-correction comparison (Plummer a=1, --learned-n)")
    ap.add_argument("--learned-n", type=int, default=1 << 22)
    ap.add_argument("--learned-theta", type=float, default=0.9)
    ap.add_argument("--learned-band", type=float, default=0.0)
    ap.add_argument("--learned-pairs", type=int, default=2_000_000)
    ap.add_argument("--learned-epochs", type=int, default=8)
    ap.add_argument("--learned-thetas", default="0.9,1.0,1.1",
                    help="learned thetas of the sweep (training pairs are collected at each)")
    ap.add_argument("--learned-terms", type=int, default=7, choices=(3, 7))
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-systems run (256 systems, graph-captured steps)")
    ap.add_argument("--batched-systems", type=int, default=256)
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--batched-variant", type=int, default=-1,
                    help="walk variant of the batched runs (default: autotuned), e.g. to profile one")
    ap.add_argument("--no-batched-truncated", action="store_true",
                    help="skip the batched run with the Plummer systems truncated at R = 3")
    ap.add_argument("--no-tsne", action="store_true",
                    help="skip the t-SNE run (70,000 points, 1,000 iterations)")
    ap.add_argument("--no-tsne-fmm", action="store_true",
                    help="skip the t-SNE run with the symmetric dual-tree FMM repulsion")
    ap.add_argument("--tsne-n", type=int, default=70000)
    ap.add_argument("--tsne-iters", type=int, default=1000)
    ap.add_argument("--tsne-nmae-iters", type=int, default=100,
                    help="iterations of the NMAE protocol along a near-exact trajectory (0: skip)")
    ap.add_argument("--tsne-split", type=int, default=0,
                    help="waves per sink group in the t-SNE walk; 0 = tune (variant and split)")
    ap.add_argument("--variant", type=int, default=-1,
                    help="traversal variant; -1 = autotune on the workload before timing")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs the Barnes-Hut step on one GPU only; --gpus {args.gpus} "
              f"(WORLD_SIZE={world}) is not supported", file=sys.stderr)
        sys.exit(2)

    import torch

    from masterthesislucas_amd import BHContext
    from masterthesislucas_amd.workloads import sample_indices, two_plummer_disks

    assert torch.cuda.is_available(), "no GPU visible"
    if args.copy_mb:
        src = torch.ones(args.copy_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {args.copy_mb} MiB done")
        del src, dst
    n, theta, eps, dt = args.n, args.theta, 1e-3, args.dt
    log(f"[bench] generating two Plummer disks n={n}")
    p = two_plummer_disks(n)
    ctx = BHContext(n, kernel="gravity", tree_mode="quadtree", max_children=16,
                    code_bits=args.code_bits, G=1.0, softening=eps)
    ctx.set_particles(p.x, p.y, p.m, p.vx, p.vy)
    tuning = None
    if args.variant >= 0:
        ctx.set_variant(args.variant)
        variant = args.variant
    else:
        from masterthesislucas_amd.autotune import tune

        ctx.build()
        variant, times = tune(ctx, theta)
        tuning = {str(k): round(v, 4) for k, v in times.items()}
        log(f"[bench] autotune: variant {variant} ({tuning} ms)")
        ctx.set_particles(p.x, p.y, p.m, p.vx, p.vy)
    log(f"[bench] warmup {args.warmup} steps")
    ctx.leapfrog(args.warmup, dt, theta)
    # inside the timed steps only the traversal is bracketed by HIP events (its average launch
    # is the roofline's denominator): events around every phase delay each next kernel's start
    ctx.profile_phases(["traversal"])
    ctx.profile_read(reset=True)
    torch.cuda.synchronize()
    ctx.synchronize()
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    ctx.leapfrog(args.steps, dt, theta)
    ctx.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    trav_ms, trav_launches = ctx.profile_read(reset=True)["traversal"]
    wall = t1 - t0
    pn, pp = ctx.counters()  # interactions of the last step's force evaluation
    log(f"[bench] {wall:.4f} s for {args.steps} steps; pn={pn} pp={pp}")
    # per-phase breakdown: as many further steps with every phase timed
    ctx.profile_enable(True)
    ctx.leapfrog(args.steps, dt, theta)
    phases = ctx.profile_read(reset=True)
    ctx.profile_enable(False)

    trav_avg_s = trav_ms / 1e3 / max(trav_launches, 1)
    flop = FLOP_PER_INTERACTION["gravity"] * (pn + pp)
    achieved = flop / trav_avg_s / 1e12
    nodes, max_level, overfull = ctx.tree_info()

    # Barnes-Hut error of the final state against the FP64 direct sum on sampled sinks
    err = None
    sweep = []
    if args.error_sample > 0:
        sample = sample_indices(n, args.error_sample).astype(np.uint32)
        dx, dy, _ = ctx.direct_fp64(sample)
        dnorm = np.hypot(dx, dy)

        def bh_error(th):
            ax, ay, _ = ctx.forces(th, want_total=False)
            num = np.hypot(ax[sample].astype(np.float64) - dx, ay[sample].astype(np.float64) - dy)
            rel = num / dnorm
            return {"median_rel": float(np.median(rel)), "p99_rel": float(np.quantile(rel, 0.99))}

        err = {"sample": int(sample.size), **bh_error(theta)}
        if not args.no_sweep:
            from masterthesislucas_amd.autotune import time_variant

            for th in (0.3, 0.5, 0.7, 1.0):
                ms = time_variant(ctx, variant, th, repeats=3)
                cpn, cpp = ctx.counters()
                sweep.append({"theta": th, "traversal_ms": ms,
                              "interactions_per_particle": (cpn + cpp) / n,
                              "force_evals_per_s_traversal_only": n / (ms / 1e3),
                              **bh_error(th)})
            ctx.set_variant(variant)

    result = {
        "metric": "BH force evals/sec at N=1M, theta=0.5",
        "value": n * args.steps / wall,
        "unit": "force evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "two Plummer disks (a=0.1, 2 x 524,288) at (+-0.5,0), "
                               "counter-rotating; G=1, eps=1e-3; QuadTree leaf capacity 16; "
                               "leapfrog KDK dt=1e-4, full rebuild every step",
                   "n": n, "theta": theta, "tree": "quadtree", "code_bits": args.code_bits,
                   "parallelism": "single GPU"},
        "interactions_per_s": (pn + pp) * args.steps / wall,
        "particle_steps_per_s": n * args.steps / wall,
        "interactions_per_particle": (pn + pp) / n,
        "tree": {"nodes": nodes, "max_level": max_level, "overfull_leaves": overfull},
        "traverse_variant": variant,
        "walk_priority": ctx.priority,
        "autotune_ms": tuning,
        "phase_ms_per_step": {k: v[0] / args.steps for k, v in phases.items()},
        "roofline": {"bound": "valu", "pipe": "FP32 VALU (traversal)", "kernel": "k_traverse",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     # packed FP32 is off in this build: the issue-limited VALU peak is half
                     "frac_of_non_packed_peak": achieved / (PEAK_FP32_TFLOPS / 2),
                     "flop_per_launch": flop, "avg_launch_ms": trav_avg_s * 1e3},
        "bh_error_vs_fp64_direct": err,
        "theta_sweep": sweep,
    }
    traffic_file = os.path.join(ROOT, "profiles", "traverse_traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            result["roofline"]["traffic"] = json.load(f).get("bytes_per_launch")
    ctx.close()
    if not args.no_batched:
        bv = args.batched_variant if args.batched_variant >= 0 else None
        result["batched"] = batched_comparison(args, eps, theta, dt, variant=bv)
        if not args.no_batched_truncated:
            result["batched"]["truncated_r3"] = batched_comparison(args, eps, theta, dt, rmax=3.0,
                                                                   variant=bv)
    if not args.no_learned:
        result["learned"] = learned_comparison(args, eps)
    if not args.no_tsne:
        result["tsne"] = tsne_run(args)
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (reference BH, one core)")
        result["cpu_baseline"] = cpu_baseline(p, theta, eps, n)
        result["gpu_over_cpu"] = result["value"] / result["cpu_baseline"]["value"]
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
