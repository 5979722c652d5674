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


def _fresh(s, traversal="rows"):
    import torch
    spec = pdx.PME(coulomb_prefactor=14.399645, real_cutoff=6.0, real_space_tolerance=1e-5, mesh_spacing=1.0,
                   order=5)
    eng = pdx.ForceEngine(s["N"], s["box"], pdx.LennardJones(0.01, 2.8, 6.0), electrostatics=spec,
                          charges=s["q"], skin=0.5, traversal=traversal)
    bufs = [torch.empty(s["N"], device="cuda") for _ in range(4)] + [torch.empty(6, s["N"], device="cuda")]
    return eng, bufs


def _unrelated(seed):
    # another jittered crystal, translated by several lattice spacings: every ion is far from
    # where the previous build binned it, and no two ions come close
    s = systems.rock_salt(10, 2.82, 0.1, seed=seed)
    L = s["box"][0]
    for k, d in zip("xyz", (9.3, -7.1, 11.4)):
        r = s[k].astype(np.float64) + d
        s[k] = (r - L * np.floor(r / L + 0.5)).astype(np.float32)
    return s


@pytest.mark.parametrize("graph", [False, True])
def test_engine_on_unrelated_configuration(cuda, graph):
    """One engine evaluated on configuration A, then on an unrelated configuration B, gives B's
    forces, energy and virial: the spread from the cells of A's build (valid only while every
    particle moved less than the skin since it) is replaced by the per-particle spread and gather
    for that evaluation. Compared with a fresh engine on B."""
    import torch
    a = systems.rock_salt(10, 2.82, 0.1, seed=0)
    b = _unrelated(7)
    eng, bufs = _fresh(a)
    PA = [dev(a[k]) for k in ("x", "y", "z", "q")]
    PB = [dev(b[k]) for k in ("x", "y", "z", "q")]
    work = [t.clone() for t in PA]
    eng.evaluate(*work, *bufs, energy=True)
    eng.evaluate(*work, *bufs, energy=True)  # the second check takes the snapshot of A's build
    if graph:
        eng.capture(*work, *bufs, energy=True)
        eng.replay()
    torch.cuda.synchronize()
    for t, u in zip(work, PB):
        t.copy_(u)
    for rep in range(2):  # B with A's cells (stale snapshot), then B with B's own snapshot
        if graph:
            eng.replay()
        else:
            eng.evaluate(*work, *bufs, energy=True)
        got, E, W = _snap(eng, bufs)
        ref_eng, ref_bufs = _fresh(b)
        for _ in range(rep + 1):
            ref_eng.evaluate(*PB, *ref_bufs, energy=True)
        ref, Eref, Wref = _snap(ref_eng, ref_bufs)
        F = torch.stack(got[:3]).double()
        Fr = torch.stack(ref[:3]).double()
        err = ((F - Fr).norm() / Fr.norm()).item()
        assert err < 1e-6, f"evaluation {rep}: force rel. difference {err:.3e}"
        assert abs(E - Eref) < 1e-7 * abs(Eref), (rep, E, Eref)
        for w, wr in zip(W, Wref):
            assert abs(w - wr) < 1e-6 * max(abs(wr), abs(Eref)), (rep, W, Wref)
        if rep == 1:  # both spread from a valid snapshot of B's build: the same kernels, bitwise
            for x, y in zip(got, ref):
                assert torch.equal(x, y)


@pytest.mark.parametrize("system", ["rock_salt", "melt"])
def test_spatial_reorder_is_label_independent(cuda, system):
    """Spatial reordering (src/engine/Ordering.jl:19-79): a configuration and a random relabelling
    of it, each reordered by its engine, end in the same storage order bit for bit (cells, then
    position inside a cell), and then give bitwise equal rows, forces, energies and virials; the
    tags map every slot back to its original particle."""
    import torch
    if system == "rock_salt":
        s = systems.rock_salt(12, 2.82, 0.1, seed=2)
        N = s["N"]
        arrays = {k: s[k] for k in ("x", "y", "z", "q")}

        def engine(q):
            spec = pdx.PME(coulomb_prefactor=14.399645, real_cutoff=6.0, real_space_tolerance=1e-5,
                           mesh_spacing=1.0, order=5)
            return pdx.ForceEngine(N, s["box"], pdx.LennardJones(0.01, 2.8, 6.0), electrostatics=spec, charges=q,
                                   skin=0.5)
    else:
        from pdx import melt
        s = systems.kg_chains(60, length=30, density=0.85, seed=5)
        melt.push_off(s, iterations=100)
        N = s["N"]
        arrays = {k: s[k] for k in ("x", "y", "z")}
        arrays["q"] = None

        def engine(q, bonds):
            return pdx.ForceEngine(N, s["box"], pdx.WCA(), skin=0.4, bonds=pdx.FENEBonds(N, bonds),
                                   accumulation="fixed64")
    rng = np.random.default_rng(9)
    sigma = rng.permutation(N)
    inv = np.argsort(sigma)
    results = []
    for relabel in (False, True):
        idx = sigma if relabel else np.arange(N)
        P = [dev(arrays[k][idx]) for k in ("x", "y", "z")]
        Q = dev(arrays["q"][idx]) if arrays["q"] is not None else None
        tags = torch.from_numpy(idx.astype(np.int64)).cuda()  # slot -> original id
        if system == "rock_salt":
            eng = engine(arrays["q"][idx])
        else:
            bonds = inv[s["bonds"]] if relabel else s["bonds"]
            eng = engine(None, bonds)
        f = [torch.empty(N, device="cuda") for _ in range(4)]
        V = torch.empty(6, N, device="cuda")
        eng.evaluate(*P, Q, *f, V, energy=True)
        eng.reorder(*P, Q, tags=tags)
        eng.evaluate(*P, Q, *f, V, energy=True)
        eng.evaluate(*P, Q, *f, V, energy=True)
        tot = eng.totals()
        rows = eng.nl.rows_host()
        results.append(([t.clone() for t in P], tags.clone(), [t.clone() for t in f], V.clone(), tot["energy"],
                        list(tot["virial"]), rows))
    (Pa, ta, fa, Va, Ea, Wa, ra), (Pb, tb, fb, Vb, Eb, Wb, rb) = results
    for a, b in zip(Pa, Pb):
        assert torch.equal(a, b)
    assert torch.equal(ta, tb)  # both slots hold the same original particle
    assert ra == rb
    for a, b in zip(fa, fb):
        assert torch.equal(a, b)
    assert torch.equal(Va, Vb) and Ea == Eb and Wa == Wb
    # tags: the slot's coordinates are the tagged particle's
    x0 = arrays["x"][ta.cpu().numpy()]
    assert np.array_equal(Pa[0].cpu().numpy(), x0)
