"""Push-off of a random-walk bead-spring melt with the library's own kernels: capped steepest
descent on WCA + FENE (every bead moves at most ``step`` along its force per iteration), so
the overlaps of the random walks open up before a precision study or a run."""
from __future__ import annotations

from .engine import ForceEngine, default_capacity
from .forces import FENEBonds, SoftRepulsive, WCA


def wrap_centred(x, L):
    """Coordinates into [-L/2, L/2] in FP32: the floor shift can land a hair outside when
    x / L + 1/2 rounds, so the result is folded once more."""
    import torch
    x = x - L * torch.floor(x / L + 0.5)
    x = torch.where(x < -0.5 * L, x + L, x)
    return torch.where(x > 0.5 * L, x - L, x)


def push_off(sysd, iterations=300, step=0.01, k=30.0, R0=1.5, skin=0.4, soft_fraction=0.7, max_bond=1.2):
    """Relax ``sysd`` (from ``systems.kg_chains``) in place on the device: first
    ``soft_fraction`` of the iterations with a bounded harmonic repulsion (rc = sigma, its
    strength ramped from 5 to 100) so overlaps open up, then WCA with a quarter of the step.
    After every step bonds longer than ``max_bond`` are shortened symmetrically to it (each
    bead is the first member of at most one bond and the second of at most one, as in linear
    chains), so no bond leaves the FENE domain however crowded a bead is. Returns the device
    coordinates (x, y, z) and the bonds."""
    import numpy as np
    import torch
    N, box = sysd["N"], sysd["box"]
    X, Y, Z = (torch.from_numpy(sysd[c]).cuda() for c in ("x", "y", "z"))
    pairs = np.asarray(sysd["bonds"], dtype=np.int64)
    if len(np.unique(pairs[:, 0])) != len(pairs) or len(np.unique(pairs[:, 1])) != len(pairs):
        raise ValueError("push_off expects linear chains (each bead first / second in at most one bond)")
    bonds = FENEBonds(N, sysd["bonds"], k=k, R0=R0)
    ba = torch.from_numpy(pairs[:, 0]).cuda()
    bb = torch.from_numpy(pairs[:, 1]).cuda()
    fx, fy, fz = (torch.empty(N, device="cuda") for _ in range(3))
    Lt = torch.tensor(box, dtype=torch.float32, device="cuda")
    n_soft = int(iterations * soft_fraction)
    # random walks pile beads up locally: rows get twice the uniform-density capacity
    cap = 2 * default_capacity(N, box, 2 ** (1 / 6) + skin)
    soft = [ForceEngine(N, box, SoftRepulsive(epsilon=eps, sigma=1.0), skin=skin, bonds=bonds, cap=cap)
            for eps in (5.0, 20.0, 100.0)]
    wca = ForceEngine(N, box, WCA(epsilon=1.0, sigma=1.0), skin=skin, bonds=bonds, cap=cap)
    for it in range(iterations):
        if it < n_soft:
            eng, h = soft[min(2, 3 * it // max(n_soft, 1))], step
        else:
            eng, h = wca, 0.25 * step
        eng.evaluate(X, Y, Z, None, fx, fy, fz, None, None, energy=False)
        # direction of the force and a step of h min(1, |F|), safe for huge or infinite WCA
        # forces of near-coincident beads
        F = torch.nan_to_num(torch.stack([fx, fy, fz], 1), nan=0.0, posinf=3e38, neginf=-3e38)
        s = F.abs().amax(dim=1, keepdim=True).clamp_min(1e-30)
        Fn = F / s
        nrm = Fn.norm(dim=1, keepdim=True).clamp_min(1e-30)
        P = torch.stack([X, Y, Z], 1) + Fn / nrm * (h * torch.clamp(s * nrm, max=1.0))
        d = P[bb] - P[ba]
        d -= Lt * torch.floor(d / Lt + 0.5)
        r = d.norm(dim=1, keepdim=True)
        C = 0.5 * (r - max_bond).clamp_min(0.0) / r.clamp_min(1e-12) * d
        shift = torch.zeros_like(P)
        shift[ba] += C
        shift2 = torch.zeros_like(P)
        shift2[bb] += C
        P += shift - shift2
        X.copy_(wrap_centred(P[:, 0], Lt[0])); Y.copy_(wrap_centred(P[:, 1], Lt[1]))
        Z.copy_(wrap_centred(P[:, 2], Lt[2]))
    torch.cuda.synchronize()
    bonds.status()
    for e in soft + [wca]:
        e.nl.status()
    for c, t in zip(("x", "y", "z"), (X, Y, Z)):
        sysd[c] = t.cpu().numpy()
    return X, Y, Z, bonds
