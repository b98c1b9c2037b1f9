"""Seeded synthetic inputs of the Barnes-Hut workloads (FP32 particles, SoA).

All generators use numpy's PCG64 stream with the seeds the reference's own drivers use
(main.cpp:170, nBodySolverGpuBH.cu:565), so CPU and GPU runs see identical inputs.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

SEED_DISK = 216308
SEED_TWO_DISKS = 592340823
SEED_TSNE = 4523
SEED_BATCHED = 296343
SEED_PLUMMER = 1952732


@dataclass
class Particles:
    x: np.ndarray
    y: np.ndarray
    m: np.ndarray
    vx: np.ndarray
    vy: np.ndarray

    @property
    def n(self) -> int:
        return int(self.x.size)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def uniform_disk(n: int, seed: int = SEED_DISK, radius: float = 1.0) -> Particles:
    """Uniform unit disk by rejection sampling from the enclosing square (the sampling rule of
    gravitysim.h:182-196), zero velocity, masses 1/N."""
    rng = np.random.default_rng(seed)
    out = np.empty((0, 2))
    while out.shape[0] < n:
        p = rng.uniform(-radius, radius, size=(2 * (n - out.shape[0]) + 64, 2))
        p = p[(p * p).sum(1) < radius * radius]
        out = np.concatenate([out, p])
    out = out[:n]
    z = np.zeros(n, np.float32)
    return Particles(_f32(out[:, 0]), _f32(out[:, 1]), np.full(n, 1.0 / n, np.float32), z, z.copy())


def plummer_radii(rng, n: int, a: float, rmax: float = 0.0) -> np.ndarray:
    """Radii of a 2D Plummer disk, surface density ~ (1 + R^2/a^2)^-2, whose enclosed mass
    fraction is R^2/(R^2+a^2): R = a sqrt(u/(1-u)), u ~ U[0,1). rmax > 0 truncates at rmax
    (u ~ U[0, umax) with umax = rmax^2 / (rmax^2 + a^2)); rmax = 0: not truncated."""
    umax = rmax * rmax / (rmax * rmax + a * a) if rmax > 0 else 1.0
    u = rng.random(n) * umax
    return a * np.sqrt(u / (1.0 - u))


def two_plummer_disks(n: int = 1 << 20, a: float = 0.1, sep: float = 0.5, G: float = 1.0,
                      eps: float = 1e-3, seed: int = SEED_TWO_DISKS) -> Particles:
    """Two Plummer disks (scale radius a, n/2 particles each) centred at (+-sep, 0), total mass 1,
    each rotating on circular orbits about its own centre in opposite senses with
    v_c = sqrt(G M_enc(R) R^2 / (R^2 + eps^2)^(3/2)), M_enc = (1/2) R^2/(R^2+a^2)."""
    rng = np.random.default_rng(seed)
    h = n // 2
    xs, ys, vxs, vys = [], [], [], []
    for k, (cx, sense) in enumerate(((-sep, 1.0), (sep, -1.0))):
        nk = h if k == 0 else n - h
        R = plummer_radii(rng, nk, a)
        phi = rng.random(nk) * 2.0 * np.pi
        menc = 0.5 * R * R / (R * R + a * a)
        vc = np.sqrt(G * menc * R * R / (R * R + eps * eps) ** 1.5)
        xs.append(cx + R * np.cos(phi))
        ys.append(R * np.sin(phi))
        vxs.append(-sense * vc * np.sin(phi))
        vys.append(sense * vc * np.cos(phi))
    x, y = np.concatenate(xs), np.concatenate(ys)
    vx, vy = np.concatenate(vxs), np.concatenate(vys)
    return Particles(_f32(x), _f32(y), np.full(n, 1.0 / n, np.float32), _f32(vx), _f32(vy))


def plummer(n: int = 1 << 22, a: float = 1.0, seed: int = SEED_PLUMMER,
            rmax: float = 0.0) -> Particles:
    """One 2D Plummer disk of scale radius a centred at the origin (truncated at rmax when
    rmax > 0), zero velocity, masses 1/N."""
    rng = np.random.default_rng(seed)
    R = plummer_radii(rng, n, a, rmax)
    phi = rng.random(n) * 2.0 * np.pi
    z = np.zeros(n, np.float32)
    return Particles(_f32(R * np.cos(phi)), _f32(R * np.sin(phi)), np.full(n, 1.0 / n, np.float32),
                     z, z.copy())


def batched_sizes(systems: int = 256, nmin: int = 1024, nmax: int = 262144,
                  seed: int = SEED_BATCHED) -> np.ndarray:
    """N per system, log-uniform in [nmin, nmax]: floor(exp(U(ln nmin, ln nmax)))."""
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax), systems))).astype(np.int64)


def batched_systems(systems: int = 256, seed: int = SEED_BATCHED, plummer_rmax: float = 0.0):
    """Independent systems: even-numbered ones uniform disks, odd-numbered Plummer (a = 0.1),
    untruncated by default (plummer_rmax = 0) or truncated at R = plummer_rmax. Untruncated, a
    few far outliers per system set the QuadTree root square; the cluster then sits under a
    chain of one-child cells, which the reference rule (a cell with at most one child is summed
    particle by particle over all its occupants, nBodySolverBH.h:68) turns into large
    particle-particle sums: for the 256-system workload about 2,960 particle-particle
    interactions per particle against 136 truncated at R = 3 (30 a)."""
    sizes = batched_sizes(systems, seed=seed)
    out = []
    for k, nk in enumerate(sizes):
        if k % 2 == 0:
            out.append(uniform_disk(int(nk), seed=seed + 1 + k))
        else:
            p = plummer(int(nk), a=0.1, seed=seed + 1 + k, rmax=plummer_rmax)
            out.append(p)
    return out


def sample_indices(n: int, k: int = 4096, seed: int = 12345) -> np.ndarray:
    """Seeded sample of k distinct particle indices for the FP64 direct-sum oracle."""
    rng = np.random.default_rng(seed)
    k = min(k, n)
    return np.sort(rng.choice(n, size=k, replace=False)).astype(np.int64)
