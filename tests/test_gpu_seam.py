"""Pairs across the periodic seam (test/test_force_precision_gpu.jl:20-75): ten LJ particles
in a 149.38 box, most of them in pairs that straddle a box face, evaluated on the ELL rows,
the all-pairs rows and the cluster tiles, at the initial positions, after a small move and
after a move that wraps particles across the faces. Forces, energy and virial against the FP64
direct sums within the reference's FP32 bound (2e-5 relative), with and without energies.
Inputs must be centred ([-L/2, L/2)): the reference's "positive image" variant is refused by
this ABI (PDX_ERR_COORDINATE at the status readback), so only the centred image runs."""
import numpy as np
import pytest

import oracle
import pdx
from conftest import dev

pytestmark = pytest.mark.gpu

POS = np.array([[-73.96331, 74.49271, -30.13671, -28.84318, 2.14, 2.18, 40.2, 40.25, -5.1309, -3.8938],
                [0.2, 0.24, 25.25, 25.28, -74.13124, 73.93639, 30.14, 30.2, -5.6216, -5.6417],
                [0.3, 0.35, 25.32, 25.4, 35.17, 35.2, -74.04009, 74.11345, -5.2937, -5.2491]])
L = float(np.float32(149.38015821857212))


def _phases():
    R = POS.astype(np.float32)
    yield "initial", R.copy()
    ids = np.arange(1, 11, dtype=np.float64)
    R = R + np.stack([(0.017 * np.sin(ids * (0.73 + 0.11 * d))).astype(np.float32) for d in (1, 2, 3)])
    yield "small_move", R.copy()
    R = R + np.float32(0.33)
    R = (R - np.float32(L) * np.floor(R / np.float32(L) + np.float32(0.5))).astype(np.float32)
    yield "crossing", R.copy()


def _engine(backend):
    import torch
    lj = pdx.LennardJones(1.0, 1.0, 2.5)
    if backend == "allpairs":
        nl = pdx.DeviceNeighborWorkspace(10, (L, L, L), 2.5, 0.4, 64, allpairs=True)
        pf = pdx.PairForce(lj)
        tot = torch.zeros(7, dtype=torch.float64, device="cuda")

        def ev(X, Y, Z, f, E, V, energy):
            nl.update(X, Y, Z, None)
            pf.compute(nl, *f, E, V, tot if energy else None, energy=energy)
            torch.cuda.synchronize()
            nl.status()
            return tot.cpu().numpy()
        return ev
    kw = dict(backend="cluster", prune_buffer=0.05) if backend == "cluster" else {}
    eng = pdx.ForceEngine(10, (L, L, L), lj, skin=0.4, cap=64, **kw)

    def ev(X, Y, Z, f, E, V, energy):
        eng.evaluate(X, Y, Z, None, *f, E, V, energy=energy)
        torch.cuda.synchronize()
        t = eng.totals() if energy else None
        return None if t is None else np.array([t["energy"]] + list(t["virial"]))
    return ev


@pytest.mark.parametrize("backend", ["ell", "allpairs", "cluster"])
def test_seam_pairs_against_fp64(cuda, backend):
    import torch
    ev = _engine(backend)
    tol = 2e-5
    for phase, R in _phases():
        X, Y, Z = (dev(R[d]) for d in range(3))
        f = [torch.empty(10, device="cuda") for _ in range(3)]
        E = torch.empty(10, device="cuda")
        V = torch.empty(6, 10, device="cuda")
        tot = ev(X, Y, Z, f, E, V, True)
        Rd = R.astype(np.float64)
        ref = oracle.pair_direct(Rd[0], Rd[1], Rd[2], (L, L, L), kind="lj", cutoff=2.5)
        F = torch.stack(f, 1).double().cpu().numpy()
        nF = np.linalg.norm(ref["F"])
        assert np.linalg.norm(F - ref["F"]) / nF < tol, (backend, phase)
        assert np.linalg.norm(F.sum(0)) / nF < tol, (backend, phase)
        Eref = ref["energy"].sum()
        assert abs(tot[0] - Eref) / max(abs(Eref), 1.0) < tol, (backend, phase)
        Wref = ref["virial"].sum(0)
        assert np.linalg.norm(tot[1:] - Wref) / max(np.linalg.norm(Wref), 1.0) < tol, (backend, phase)
        f2 = [torch.empty(10, device="cuda") for _ in range(3)]
        ev(X, Y, Z, f2, None, None, False)
        F2 = torch.stack(f2, 1).double().cpu().numpy()
        assert np.linalg.norm(F2 - ref["F"]) / nF < tol, (backend, phase)


def test_uncentred_coordinates_are_refused(cuda):
    """Coordinates in [0, L) (the reference's positive image) latch a coordinate error that
    the status readback raises; the list stays unusable until valid coordinates come."""
    import torch
    nl = pdx.DeviceNeighborWorkspace(10, (L, L, L), 2.5, 0.4, 64)
    R = POS.astype(np.float32) + np.float32(L / 2)
    X, Y, Z = (dev(R[d]) for d in range(3))
    nl.update(X, Y, Z, None)
    with pytest.raises(pdx.ArgumentError):
        nl.status()
