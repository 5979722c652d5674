"""Full force evaluation (list check, pairs, mesh): overlapped streams and hipGraph replay
give bitwise the results of the serial eager path (every kernel is deterministic)."""
import numpy as np
import pytest

import pdx
from conftest import dev
from pdx import systems

pytestmark = pytest.mark.gpu


def _setup(seed=0, traversal="rows"):
    import torch
    s = systems.rock_salt(10, 2.82, 0.1, seed=seed)
    spec = pdx.PME(coulomb_prefactor=14.399645, real_cutoff=6.0, real_space_tolerance=1e-5, mesh_spacing=1.0,
                   order=5)
    eng = pdx.ForceEngine(s["N"], s["box"], pdx.LennardJones(0.01, 2.8, 6.0), electrostatics=spec,
                          charges=s["q"], skin=0.5, traversal=traversal)
    X, Y, Z, Q = (dev(s[k]) for k in ("x", "y", "z", "q"))
    bufs = [torch.empty(s["N"], device="cuda") for _ in range(4)] + [torch.empty(6, s["N"], device="cuda")]
    return s, eng, (X, Y, Z, Q), bufs


def _snap(eng, bufs):
    import torch
    torch.cuda.synchronize()
    tot = eng.totals()
    return [b.clone() for b in bufs], tot["energy"], list(tot["virial"])


@pytest.mark.parametrize("traversal", ["rows", "cells"])
def test_overlap_and_graph_match_serial(cuda, traversal):
    import torch
    s, eng, P, bufs = _setup(traversal=traversal)
    eng.evaluate(*P, *bufs, energy=True, overlap=False)
    ref, Eref, Wref = _snap(eng, bufs)
    eng.evaluate(*P, *bufs, energy=True, overlap=True)
    got, E, W = _snap(eng, bufs)
    for a, b in zip(got, ref):
        assert torch.equal(a, b)
    assert E == Eref and W == Wref
    eng.capture(*P, *bufs, energy=True)
    # replays interleaved with unrelated eager work, each checked against the eager result
    s2, eng2, P2, bufs2 = _setup(seed=1)
    for k in range(3):
        for b in bufs:
            b.fill_(float("nan"))
        eng.replay()
        got, E, W = _snap(eng, bufs)
        for a, b in zip(got, ref):
            assert torch.equal(a, b), f"replay {k}"
        assert E == Eref and W == Wref
        eng2.evaluate(*P2, *bufs2, energy=True)
        torch.cuda.synchronize()


def test_energy_total_consistent(cuda):
    s, eng, P, bufs = _setup()
    eng.evaluate(*P, *bufs, energy=True)
    got, E, W = _snap(eng, bufs)
    # per-particle energies (pair halves + self + mesh halves) sum to the FP64 total
    assert abs(got[3].double().sum().item() - E) < 1e-5 * abs(E)
    assert np.isfinite(W).all()
