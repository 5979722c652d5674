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


def _fcc_counts(n):
    """FCC cell counts (mx, my, mz) of the lattice with the fewest sites 4 mx my mz >= n among those
    whose counts differ by at most one: n sites drawn from it sit at nearly the target density
    (not on a lattice up to twice as dense as a cube of m^3 cells would give), and its cells in the
    cubic box are nearly cubic, so no axis is compressed (a 3 x 6 x 6 lattice would put nearest
    neighbours at 0.9 of the isotropic spacing)."""
    c = (n / 4.0) ** (1.0 / 3.0)
    lo, hi = max(1, int(np.floor(c)) - 1), int(np.ceil(c)) + 1
    best = None
    for mx in range(lo, hi + 1):
        for my in range(lo, hi + 1):
            for mz in range(lo, hi + 1):
                M = 4 * mx * my * mz
                spread = max(mx, my, mz) - min(mx, my, mz)
                if M >= n and spread <= 1 and (best is None or M < best[0]):
                    best = (M, (mx, my, mz))
    return best[1]


def replica_batch(nrep=512, seed=0, n_range=(256, 8192), density_range=(0.70, 0.90), temperature=0.72,
                  models=("wca", "lj"), jitter=0.02):
    """Seeded batch of independent LJ / WCA boxes (the batched-replica config): per replica,
    n log-uniform in ``n_range``, density uniform in ``density_range``, the model drawn
    uniformly from ``models`` (LJ: eps = sigma = 1, rc = 2.5, energy shifted; WCA: eps =
    sigma = 1), atoms on a seeded random subset of the sites of the FCC lattice with the fewest
    sites that holds them with nearly cubic cells (at most ~25 % more sites than atoms), Gaussian
    jitter (0.02: with 0.05 the closest pairs of the 512 boxes start at 0.78 sigma and even an
    FP64 integration drifts by ~1e-2 per atom over 1000 steps; at 0.02 they start above 0.9
    sigma and drift below 1e-3), Maxwell velocities at ``temperature`` with zero net momentum.
    Returns specs (for ``ReplicaBatch``), concatenated FP32 x, y, z, vx, vy, vz and the offsets."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(n_range[0]), np.log(n_range[1])
    specs, xs, vs = [], [], []
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    for k in range(nrep):
        n = int(min(n_range[1], max(n_range[0], round(np.exp(rng.uniform(lo, hi))))))
        rho = float(rng.uniform(*density_range))
        model = models[int(rng.integers(len(models)))]
        Lb = (n / rho) ** (1.0 / 3.0)
        mx, my, mz = _fcc_counts(n)
        g = np.stack(np.meshgrid(np.arange(mx), np.arange(my), np.arange(mz), indexing="ij"), -1).reshape(-1, 3)
        cell = Lb / np.array([mx, my, mz], dtype=np.float64)
        sites = ((g[:, None, :] + basis[None, :, :]).reshape(-1, 3) + 0.25) * cell - Lb / 2
        sites = sites[np.sort(rng.choice(len(sites), n, replace=False))]
        pos = _wrap(sites + rng.normal(0.0, jitter, size=sites.shape), Lb)
        v = rng.normal(0.0, np.sqrt(temperature), size=(n, 3))
        v -= v.mean(0)
        spec = {"model": model, "n": n, "box": (Lb, Lb, Lb), "epsilon": 1.0, "sigma": 1.0, "density": rho}
        if model == "lj":
            spec.update(cutoff=2.5, shift=True)
        specs.append(spec)
        xs.append(pos)
        vs.append(v)
    X = np.concatenate(xs).astype(np.float32)
    V = np.concatenate(vs).astype(np.float32)
    offsets = np.concatenate([[0], np.cumsum([s["n"] for s in specs])]).astype(np.int64)
    # FP32 rounding may carry a coordinate onto +L/2: fold it back into [-L/2, L/2)
    for k, s in enumerate(specs):
        Lb = np.float32(s["box"][0])
        c = X[offsets[k]:offsets[k + 1]]
        c[c >= Lb / 2] -= Lb
        c[c < -Lb / 2] += Lb
    return dict(specs=specs, x=X[:, 0].copy(), y=X[:, 1].copy(), z=X[:, 2].copy(), vx=V[:, 0].copy(),
                vy=V[:, 1].copy(), vz=V[:, 2].copy(), offsets=offsets, N=int(offsets[-1]))


def kg_chains(nchains, length=100, density=0.85, bond=0.97, seed=0, max_cos=-0.66):
    """Random-walk bead-spring chains (the Kremer-Grest melt's starting point): chain starts
    uniform in the box, steps of length ``bond`` in uniform random directions, redrawn while
    the angle to the previous step has cos < ``max_cos`` (so beads i and i+2 stay > 0.8
    apart), box side (N / density)^(1/3). Beads of a chain are consecutive ids. Returns x, y,
    z (FP32, centred), the bonds (nb, 2) and the box."""
    rng = np.random.default_rng(seed)
    N = nchains * length
    L = (N / density) ** (1.0 / 3.0)
    pos = np.empty((nchains, length, 3))
    pos[:, 0] = rng.random((nchains, 3)) * L - L / 2
    prev = None
    for s in range(1, length):
        v = rng.normal(size=(nchains, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        if prev is not None:
            bad = (v * prev).sum(1) < max_cos
            while bad.any():
                w = rng.normal(size=(int(bad.sum()), 3))
                v[bad] = w / np.linalg.norm(w, axis=1, keepdims=True)
                bad = (v * prev).sum(1) < max_cos
        pos[:, s] = pos[:, s - 1] + bond * v
        prev = v
    P = _wrap(pos.reshape(-1, 3), L).astype(np.float32)
    for d in range(3):  # FP32 rounding may leave a coordinate a hair outside [-L/2, L/2)
        c = P[:, d]
        c[c >= np.float32(L / 2)] -= np.float32(L)
        c[c < -np.float32(L / 2)] += np.float32(L)
    ids = np.arange(N).reshape(nchains, length)
    bonds = np.stack([ids[:, :-1].reshape(-1), ids[:, 1:].reshape(-1)], 1).astype(np.int32)
    return dict(x=P[:, 0].copy(), y=P[:, 1].copy(), z=P[:, 2].copy(), bonds=bonds, box=(L, L, L), N=N)
