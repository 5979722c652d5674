"""Push-off of a random-walk bead-spring melt with the library's own kernels: capped steepest
descent on WCA + FENE (every bead moves at most ``step`` along its force per iteration), so
the overlaps of the random walks open up before a precision study or a run."""
from __future__ import annotations

from .engine import ForceEngine
from .forces import FENEBonds, SoftRepulsive, WCA


def push_off(sysd, iterations=300, step=0.01, k=30.0, R0=1.5, skin=0.4, soft_fraction=0.7):
    """Relax ``sysd`` (from ``systems.kg_chains``) in place on the device: first
    ``soft_fraction`` of the iterations with a bounded harmonic repulsion (rc = sigma, its
    strength ramped from 5 to 100) so the FENE bonds hold while overlaps open up, then WCA
    with a quarter of the step. Returns the device coordinates (x, y, z) and the bonds."""
    import torch
    N, box = sysd["N"], sysd["box"]
    X, Y, Z = (torch.from_numpy(sysd[c]).cuda() for c in ("x", "y", "z"))
    bonds = FENEBonds(N, sysd["bonds"], k=k, R0=R0)
    fx, fy, fz = (torch.empty(N, device="cuda") for _ in range(3))
    Lt = torch.tensor(box, dtype=torch.float32, device="cuda")
    n_soft = int(iterations * soft_fraction)
    soft = [ForceEngine(N, box, SoftRepulsive(epsilon=eps, sigma=1.0), skin=skin, bonds=bonds)
            for eps in (5.0, 20.0, 100.0)]
    wca = ForceEngine(N, box, WCA(epsilon=1.0, sigma=1.0), skin=skin, bonds=bonds)
    for it in range(iterations):
        if it < n_soft:
            eng, h = soft[min(2, 3 * it // max(n_soft, 1))], step
        else:
            eng, h = wca, 0.25 * step
        eng.evaluate(X, Y, Z, None, fx, fy, fz, None, None, energy=False)
        F = torch.stack([fx, fy, fz], 1)
        mag = F.norm(dim=1, keepdim=True).clamp_min(1.0)
        M = F * (h / mag)
        for c, t in enumerate((X, Y, Z)):
            t += M[:, c]
            t -= Lt[c] * torch.floor(t / Lt[c] + 0.5)
    torch.cuda.synchronize()
    bonds.status()
    for e in soft + [wca]:
        e.nl.status()
    for c, t in zip(("x", "y", "z"), (X, Y, Z)):
        sysd[c] = t.cpu().numpy()
    return X, Y, Z, bonds
