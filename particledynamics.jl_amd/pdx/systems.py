"""Seeded synthetic systems of the benchmark configurations (numpy, FP64 construction
rounded once to FP32, so every consumer sees identical FP32 inputs). Coordinates are
centred in [-L/2, L/2)."""
from __future__ import annotations

import numpy as np


def _wrap(r, L):
    return r - L * np.floor(r / L + 0.5)


def fcc_lattice(N=4000, density=0.8442, jitter=0.05, seed=0):
    """FCC lattice of N = 4 n^3 atoms at reduced density, Gaussian jitter (the LJ fluid config)."""
    n = int(round((N / 4) ** (1.0 / 3.0)))
    if 4 * n ** 3 != N:
        raise ValueError("N must be 4 n^3")
    L = (N / density) ** (1.0 / 3.0)
    a = L / n
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    pos = ((g[:, None, :] + basis[None, :, :]).reshape(-1, 3) + 0.25) * a - L / 2
    rng = np.random.default_rng(seed)
    pos = pos + rng.normal(0.0, jitter, size=pos.shape)
    pos = _wrap(pos, L).astype(np.float32)
    return dict(x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(), box=(L, L, L), N=N)


def rock_salt(n=40, a=2.82, jitter=0.1, seed=0, vacancy_fraction=0.0):
    """n^3 rock-salt sites, charges +1/-1 by parity, masses 22.99/35.45 (the ionic config)."""
    L = n * a
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    q = np.where(g.sum(1) % 2 == 0, 1.0, -1.0)
    pos = (g + 0.25) * a - L / 2
    rng = np.random.default_rng(seed)
    keep = np.ones(len(q), dtype=bool)
    if vacancy_fraction > 0:
        k = int(round(vacancy_fraction * len(q) / 2))
        cat = np.flatnonzero(q > 0)
        an = np.flatnonzero(q < 0)
        keep[rng.choice(cat, k, replace=False)] = False
        keep[rng.choice(an, k, replace=False)] = False
    pos, q = pos[keep], q[keep]
    if jitter > 0:
        pos = pos + rng.normal(0.0, jitter, size=pos.shape)
    pos = _wrap(pos, L).astype(np.float32)
    mass = np.where(q > 0, 22.99, 35.45)
    return dict(x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(), q=q.astype(np.float32),
                mass=mass, box=(L, L, L), N=len(q))


def random_cloud(N, box, seed=1, neutral=True):
    """Uniform random positions and N(0,1) charges, neutralised (test/test_electrostatics.jl:16-21)."""
    rng = np.random.default_rng(seed)
    L = np.asarray(box, dtype=np.float64)
    R = rng.random((N, 3)) * L - L / 2
    q = rng.standard_normal(N)
    if neutral:
        q -= q.sum() / N
    return R, q


def maxwell_velocities(N, temperature, mass=1.0, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.normal(0.0, np.sqrt(temperature / mass), size=(N, 3))
    v -= v.mean(0)
    return v.astype(np.float32)
