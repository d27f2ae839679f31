"""Seeded initial conditions for the benchmark workloads (G = 1, mean mass 1).

exp_disk      one exponential disk: a central body of mass 1e4 at the origin,
              radii ~ Exp(scale 1.0) resampled above 5.0, uniform angles,
              masses LogNormal(0, 0.5), circular velocities from the enclosed
              mass plus 5 % Gaussian dispersion per component
two_disks     two such disks at x = -8 and +8 approaching at relative speed
              0.5, the second rotated in-plane by 30 degrees
uniform_disk  uniform in a disk of radius 10, masses U[0.5, 1.5], at rest
ensemble      many exp_disk systems, N log-uniform in [256, 65536]
"""
from __future__ import annotations

import numpy as np


def exp_disk(n: int, seed: int, *, center=(0.0, 0.0), bulk=(0.0, 0.0), rot_deg: float = 0.0,
             central_mass: float = 1e4, scale: float = 1.0, rmax: float = 5.0,
             sigma_m: float = 0.5, disp: float = 0.05, G: float = 1.0):
    rng = np.random.default_rng(seed)
    nd = n - 1
    r = rng.exponential(scale, nd)
    bad = r > rmax
    while bad.any():
        r[bad] = rng.exponential(scale, int(bad.sum()))
        bad = r > rmax
    phi = rng.uniform(0.0, 2.0 * np.pi, nd)
    m = rng.lognormal(0.0, sigma_m, nd)
    order = np.argsort(r, kind="stable")
    enc = np.empty(nd)
    enc[order] = central_mass + np.concatenate(([0.0], np.cumsum(m[order])[:-1]))
    vc = np.sqrt(G * enc / r)
    c, s = np.cos(phi), np.sin(phi)
    pos = np.stack([r * c, r * s], axis=1)
    vel = np.stack([-s * vc, c * vc], axis=1) + disp * vc[:, None] * rng.standard_normal((nd, 2))
    pos = np.vstack([[0.0, 0.0], pos])
    vel = np.vstack([[0.0, 0.0], vel])
    mass = np.concatenate(([central_mass], m))
    if rot_deg:
        a = np.deg2rad(rot_deg)
        R = np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]])
        pos = pos @ R.T
        vel = vel @ R.T
    pos = pos + np.asarray(center, dtype=np.float64)
    vel = vel + np.asarray(bulk, dtype=np.float64)
    return np.ascontiguousarray(mass), np.ascontiguousarray(pos), np.ascontiguousarray(vel)


def two_disks(n: int, seed: int):
    h = n // 2
    m1, p1, v1 = exp_disk(h, seed, center=(-8.0, 0.0), bulk=(0.25, 0.0))
    m2, p2, v2 = exp_disk(n - h, seed + 1, center=(8.0, 0.0), bulk=(-0.25, 0.0), rot_deg=30.0)
    return (np.concatenate([m1, m2]), np.vstack([p1, p2]), np.vstack([v1, v2]))


def uniform_disk(n: int, seed: int, radius: float = 10.0):
    rng = np.random.default_rng(seed)
    r = radius * np.sqrt(rng.uniform(0.0, 1.0, n))
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    pos = np.stack([r * np.cos(phi), r * np.sin(phi)], axis=1)
    mass = rng.uniform(0.5, 1.5, n)
    return mass, np.ascontiguousarray(pos), np.zeros((n, 2))


def ensemble_sizes(nsys: int, seed: int, lo: int = 256, hi: int = 65536):
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), nsys))).astype(np.int64)


def ensemble(nsys: int, seed: int):
    sizes = ensemble_sizes(nsys, seed)
    systems = [exp_disk(int(k), seed * 1_000_003 + i) for i, k in enumerate(sizes)]
    return sizes, systems


def sample_targets(n: int, m: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    m = min(m, n)
    return np.sort(rng.choice(n, size=m, replace=False)).astype(np.int64)


def force_errors(acc, ref):
    """RMS and max relative error of acc against ref, both (m, 2)."""
    acc = np.asarray(acc, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    rel = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return float(np.sqrt(np.mean(rel ** 2))), float(np.max(rel))
