"""Cluster-pair engine (src/clusterpairs): forces, energies and virials against the FP64 direct
sums and the ELL-row path (the reference's own check, test/test_cluster_pairs.jl:55-95), list
coverage (every pair inside the cutoff in exactly one tile), the build / prune decisions
(test/test_cluster_pruning_gpu.jl), exceptions, capacity recovery, determinism and graph
replay."""
import math

import numpy as np
import pytest

import oracle
import pdx
from conftest import dev
from pdx import systems
from pdx.clusters import ClusterPairEngine

pytestmark = pytest.mark.gpu


def _relrms(a, b):
    return math.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())


def _host(fx, fy, fz, E=None, V=None, tot=None):
    import torch
    F = torch.stack([fx, fy, fz], 1).double().cpu().numpy()
    return (F, None if E is None else E.double().cpu().numpy(), None if V is None else V.double().cpu().numpy().T,
            None if tot is None else tot.cpu().numpy())


def _cluster_eval(sysd, pf, skin=0.4, prune_buffer=None, q=None, capacity=0, waves=0):
    X, Y, Z = dev(sysd["x"]), dev(sysd["y"]), dev(sysd["z"])
    Q = dev(q) if q is not None else None
    eng = ClusterPairEngine(len(sysd["x"]), sysd["box"], pf, skin, prune_buffer, capacity=capacity, waves=waves)
    out = eng.compute(X, Y, Z, Q, energy=True)
    eng.status()
    return _host(*out), eng


def _rows_eval(sysd, pf, cutoff, skin=0.4, q=None):
    import torch
    N = len(sysd["x"])
    X, Y, Z = dev(sysd["x"]), dev(sysd["y"]), dev(sysd["z"])
    Q = dev(q) if q is not None else None
    cap = pdx.default_capacity(N, sysd["box"], cutoff + skin)
    ws = pdx.DeviceNeighborWorkspace(N, sysd["box"], cutoff, skin, cap)
    ws.update(X, Y, Z, Q)
    fx, fy, fz, E = (torch.empty(N, device="cuda") for _ in range(4))
    V = torch.empty(6, N, device="cuda")
    tot = torch.zeros(7, dtype=torch.float64, device="cuda")
    pf.compute(ws, fx, fy, fz, E, V, tot, energy=True)
    ws.status()
    return _host(fx, fy, fz, E, V, tot)


def _lattice(shape, box, drop, seed):
    """Jittered simple-cubic sites in a centred box with `drop` random sites removed."""
    rng = np.random.default_rng(seed)
    n = np.array(shape)
    g = np.stack(np.meshgrid(*[np.arange(k) for k in shape], indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    P = (g + 0.5) / n * np.array(box) - np.array(box) / 2 + rng.normal(0, 0.03, g.shape)
    keep = np.sort(rng.permutation(len(P))[: len(P) - drop])
    P = P[keep]
    P -= np.array(box) * np.round(P / np.array(box))
    return dict(x=P[:, 0].astype(np.float32), y=P[:, 1].astype(np.float32), z=P[:, 2].astype(np.float32),
                box=tuple(float(b) for b in box), N=len(P))


@pytest.mark.parametrize("kind", ["lj", "wca"])
def test_fcc_against_direct_sum(cuda, kind):
    """The LJ fluid of 4000 atoms: force rel. RMS < 2e-5, energy and virial rel. < 1e-5
    (test/test_device_engine_pairs_gpu.jl:124-125), and the row path's results."""
    s = systems.fcc_lattice(4000, 0.8442, 0.05, seed=2)
    model = pdx.LennardJones(1.0, 1.0, 2.5) if kind == "lj" else pdx.WCA(1.0, 1.0)
    pf = pdx.PairForce(model)
    (F, E, V, tot), eng = _cluster_eval(s, pf)
    ref = oracle.pair_direct(s["x"], s["y"], s["z"], s["box"], kind=kind, cutoff=2.5)
    err = _relrms(F, ref["F"])
    st = eng.status()
    print(f"[clusters] {kind} relrms={err:.3e} {st}", flush=True)
    assert err < 2e-5
    Eref = ref["energy"].sum()
    assert abs(tot[0] - Eref) / abs(Eref) < 1e-5
    assert abs(E.sum() - Eref) / abs(Eref) < 1e-5
    Wref = ref["virial"].sum(0)
    assert np.abs(tot[1:] - Wref).max() / abs(Wref[:3].sum()) < 1e-5
    assert np.allclose(E, ref["energy"], rtol=1e-4, atol=1e-5)
    assert np.allclose(V, ref["virial"], rtol=1e-3, atol=1e-4)
    Fr, Er, Vr, tr = _rows_eval(s, pf, model.cutoff)
    assert _relrms(F, Fr) < 1e-6
    assert np.allclose(tot, tr, rtol=1e-6, atol=1e-6 * abs(tr[0]))
    assert st.builds == 1 and st.prunes == 1 and st.error == 0


def test_uneven_box_and_cluster_sizes(cuda):
    """2501 atoms (not a multiple of 8) in a 5.3 x 5.6 x 6.1 box, epsilon 0.8, sigma 0.32,
    cutoff 1.0 (the system of test/test_cluster_pairs.jl:68-80, on a jittered lattice so no pair
    overlaps), every wave count."""
    s = _lattice((13, 14, 14), (5.3, 5.6, 6.1), 47, seed=7)
    assert s["N"] == 2501
    pf = pdx.PairForce(pdx.LennardJones(0.8, 0.32, 1.0))
    ref = oracle.pair_direct(s["x"], s["y"], s["z"], s["box"], kind="lj", sigma=0.32, epsilon=0.8, cutoff=1.0)
    base = None
    for waves in (1, 2, 4, 8):
        (F, E, V, tot), eng = _cluster_eval(s, pf, skin=0.2, waves=waves)
        assert _relrms(F, ref["F"]) < 2e-5, waves
        assert abs(tot[0] - ref["energy"].sum()) < 1e-5 * abs(ref["energy"].sum())
        if base is None:
            base = F
        else:
            assert _relrms(F, base) < 1e-6
    assert eng.info["waves"] == 8
    # entries split over 1..8 workgroups per super-cluster: same pairs, sums in another order
    X, Y, Z = dev(s["x"]), dev(s["y"]), dev(s["z"])
    for splits in (1, 2, 8):
        eng.set_launch(4, splits)
        F = _host(*eng.compute(X, Y, Z, None, energy=True)[:3])[0]
        assert eng.info["splits"] == splits
        assert _relrms(F, base) < 1e-6, splits
        st = eng.status()
        assert st.builds == 2 + (splits > 1) * (1 if splits == 2 else 2)


def test_list_covers_every_pair_once(cuda):
    """Slots are a permutation, and every pair within rlist at the build is in exactly one tile
    of the outer list (mask bit of its i-cluster, j-cluster index >= i-cluster index)."""
    s = _lattice((13, 14, 14), (5.3, 5.6, 6.1), 47, seed=9)
    N, box = s["N"], np.array(s["box"])
    rc, skin = 1.0, 0.2
    pf = pdx.PairForce(pdx.LennardJones(0.8, 0.32, rc))
    (_, _, _, _), eng = _cluster_eval(s, pf, skin=skin)
    perm = eng.perm().cpu().numpy()
    slot_of = eng.slot_of().cpu().numpy()
    used = perm[perm >= 0]
    assert np.array_equal(np.sort(used), np.arange(N))
    assert np.array_equal(perm[slot_of], np.arange(N))
    ne, ent = (t.cpu().numpy() for t in eng.outer())
    tiles = set()
    NS = eng.status().super_clusters
    for S in range(NS):
        for e in ent[S, : ne[S]]:
            cj, m = int(e) >> 8, int(e) & 255
            for c in range(8):
                if m >> c & 1:
                    key = (S * 8 + c, cj)
                    assert key not in tiles
                    tiles.add(key)
    P = np.stack([s["x"], s["y"], s["z"]], 1).astype(np.float64)
    d = P[:, None, :] - P[None, :, :]
    d -= box * np.round(d / box)
    r2 = (d ** 2).sum(-1)
    ii, jj = np.nonzero(np.triu(r2 < (rc + skin) ** 2 * 0.999, 1))
    ci, cj = slot_of[ii] // 8, slot_of[jj] // 8
    lo, hi = np.minimum(ci, cj), np.maximum(ci, cj)
    missing = [(a, b) for a, b in zip(lo, hi) if (a, b) not in tiles]
    print(f"[clusters] {len(ii)} pairs within rlist, {len(tiles)} tiles, missing {len(missing)}", flush=True)
    assert not missing


def test_rock_salt_coulomb_matches_rows(cuda):
    """LJ + real-space Ewald on rock salt: the cluster tiles give the row path's forces, energies,
    virials and FP64 totals up to the FP32 summation order; two runs give the same bits."""
    s = systems.rock_salt(12, 2.82, 0.1, seed=5)
    coul = pdx.RealSpaceCoulomb(C=14.399645, cutoff=8.0, alpha=0.39)
    pf = pdx.PairForce(pdx.LennardJones(0.01, 2.8, 8.0), coulomb=coul)
    (F1, E1, V1, t1) = _rows_eval(s, pf, 8.0, skin=0.5, q=s["q"])
    (F2, E2, V2, t2), eng = _cluster_eval(s, pf, skin=0.5, q=s["q"])
    (F3, E3, V3, t3), _ = _cluster_eval(s, pf, skin=0.5, q=s["q"])
    print(f"[clusters] coulomb rel rms vs rows {_relrms(F2, F1):.2e} {eng.status()}", flush=True)
    assert _relrms(F2, F1) < 1e-6
    assert np.allclose(t2, t1, rtol=1e-6, atol=1e-6 * abs(t1[0]))
    assert np.allclose(E2, E1, rtol=1e-5, atol=1e-6) and np.allclose(V2, V1, rtol=1e-4, atol=1e-4)
    assert np.array_equal(F2, F3) and np.array_equal(t2, t3) and np.array_equal(E2, E3)


def test_exceptions_match_rows(cuda):
    """Scaled and excluded Coulomb pairs (normalize_exceptions, Spec.jl:200-224): the scatter's
    exception pass gives the row kernel's results."""
    s = systems.rock_salt(10, 2.82, 0.1, seed=3)
    N = s["N"]
    rng = np.random.default_rng(5)
    exc = []
    for i in rng.choice(N - 1, 60, replace=False):
        exc.append((int(i) + 1, int(i) + 2, float(rng.choice([0.0, 0.5, 0.8333]))))
    coul = pdx.RealSpaceCoulomb(C=14.399645, cutoff=7.0, alpha=0.45)
    pf = pdx.PairForce(pdx.LennardJones(0.01, 2.8, 7.0), coulomb=coul)
    pf.set_exceptions(N, exc)
    F1, E1, V1, t1 = _rows_eval(s, pf, 7.0, skin=0.5, q=s["q"])
    (F2, E2, V2, t2), _ = _cluster_eval(s, pf, skin=0.5, q=s["q"])
    assert _relrms(F2, F1) < 1e-6
    assert np.allclose(t2, t1, rtol=1e-6, atol=1e-6 * abs(t1[0]))
    assert np.allclose(E2, E1, rtol=1e-5, atol=1e-6)


def test_prune_and_rebuild_decisions(cuda):
    """Motion beyond prune_buffer/2 but within skin/2 prunes without a build; a requested build
    builds; an unchanged configuration does neither; beyond skin/2 builds. Forces stay exact
    (test/test_cluster_pruning_gpu.jl:29-55)."""
    import torch
    s = systems.fcc_lattice(4000, 0.8442, 0.05, seed=6)
    L = s["box"][0]
    pf = pdx.PairForce(pdx.LennardJones())
    eng = ClusterPairEngine(4000, s["box"], pf, 0.4, 0.1)
    P = np.stack([s["x"], s["y"], s["z"]], 1).astype(np.float64)

    def run(P):
        Pw = P - L * np.round(P / L)
        X, Y, Z = (dev(Pw[:, k].astype(np.float32)) for k in range(3))
        out = eng.compute(X, Y, Z, None, energy=True)
        F = _host(*out[:3])[0]
        ref = oracle.pair_direct(Pw[:, 0].astype(np.float32), Pw[:, 1].astype(np.float32),
                                 Pw[:, 2].astype(np.float32), s["box"], kind="lj")
        assert _relrms(F, ref["F"]) < 2e-5
        assert abs(out[5][0].item() - ref["energy"].sum()) < 1e-5 * abs(ref["energy"].sum())
        return eng.status()

    s0 = run(P)
    assert (s0.builds, s0.prunes) == (1, 1)
    P = P + np.array([0.06, 0.0, 0.0])
    s1 = run(P)
    assert (s1.builds, s1.prunes) == (1, 2)
    s2 = run(P)
    assert (s2.builds, s2.prunes) == (1, 2)
    eng.request_build()
    s3 = run(P)
    assert (s3.builds, s3.prunes) == (2, 3)
    rng = np.random.default_rng(1)
    P = P + rng.normal(0, 0.04, P.shape)
    P[17] += np.array([0.25, 0.0, 0.0])
    s4 = run(P)
    assert s4.builds == 3
    assert s4.working <= s4.outer
    torch.cuda.synchronize()


def test_capacity_overflow_and_recovery(cuda):
    s = systems.fcc_lattice(4000, 0.8442, 0.05, seed=8)
    pf = pdx.PairForce(pdx.LennardJones())
    eng = ClusterPairEngine(4000, s["box"], pf, 0.4, capacity=8)
    X, Y, Z = dev(s["x"]), dev(s["y"]), dev(s["z"])
    eng.compute(X, Y, Z, None, energy=False)
    with pytest.raises(pdx.NeighborCapacityError) as ei:
        eng.status()
    assert ei.value.required > 8
    assert eng.ensure_capacity()
    fx, fy, fz, E, V, tot = eng.compute(X, Y, Z, None, energy=True)
    st = eng.status()
    assert st.error == 0 and st.capacity >= ei.value.required
    ref = oracle.pair_direct(s["x"], s["y"], s["z"], s["box"], kind="lj")
    assert _relrms(_host(fx, fy, fz)[0], ref["F"]) < 2e-5


def test_graph_replay_equals_eager(cuda):
    """update + forces captured once and replayed: every replay equals the eager evaluation of
    the same coordinates, including a replay whose update prunes."""
    import torch
    s = systems.rock_salt(10, 2.82, 0.1, seed=4)
    coul = pdx.RealSpaceCoulomb(C=14.399645, cutoff=7.0, alpha=0.45)
    pf = pdx.PairForce(pdx.LennardJones(0.01, 2.8, 7.0), coulomb=coul)
    N = s["N"]
    X, Y, Z, Q = dev(s["x"]), dev(s["y"]), dev(s["z"]), dev(s["q"])
    eng = ClusterPairEngine(N, s["box"], pf, 0.6, 0.2)
    fx, fy, fz, E = (torch.zeros(N, device="cuda") for _ in range(4))
    V = torch.zeros(6, N, device="cuda")
    tot = torch.zeros(7, dtype=torch.float64, device="cuda")

    def step():
        eng.update(X, Y, Z, Q)
        eng.forces(fx, fy, fz, E, V, tot, energy=True)

    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(st):
        step()
    torch.cuda.current_stream().wait_stream(st)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    rng = np.random.default_rng(2)
    for k in range(4):
        d = torch.from_numpy(rng.normal(0, 0.03, (3, N)).astype(np.float32)).cuda()
        for t, dd in zip((X, Y, Z), d):
            t.add_(dd)
        g.replay()
        torch.cuda.synchronize()
        got = [t.clone() for t in (fx, fy, fz, E, tot)]
        step()
        torch.cuda.synchronize()
        for a, b in zip(got, (fx, fy, fz, E, tot)):
            assert torch.equal(a, b), k
    stt = eng.status()
    print(f"[clusters] graph replays {stt}", flush=True)
    assert stt.prunes >= 3


def test_restrictions(cuda):
    pf = pdx.PairForce(pdx.LennardJones())
    with pytest.raises(pdx.ArgumentError):
        ClusterPairEngine(100, (20.0, 20.0, 20.0), pf, 0.4, 0.5)   # prune buffer beyond the skin
    with pytest.raises(pdx.ArgumentError):
        ClusterPairEngine(100, (5.0, 20.0, 20.0), pf, 0.4)         # box side below 2 (cutoff + skin)
    with pytest.raises(pdx.ArgumentError):
        ClusterPairEngine(100, (20.0, 20.0, 20.0), pdx.PairForce(pdx.WCA(), accumulation="fp64"), 0.4)
    with pytest.raises(pdx.ArgumentError):
        ClusterPairEngine(100, (20.0, 20.0, 20.0), pf, 0.4, waves=3)


def test_engine_cluster_backend_with_mesh(cuda):
    """ForceEngine with the cluster backend (the reference's ClusterBackend: cluster tiles for the
    short range, the mesh in the cluster order) against the ELL backend on the same PME system,
    eager and as a captured graph."""
    import torch
    s = systems.rock_salt(12, 2.82, 0.1, seed=2)
    N, box = s["N"], s["box"]
    spec = pdx.PME(coulomb_prefactor=14.399645, real_cutoff=8.0, real_space_tolerance=1e-5, mesh_spacing=1.0,
                   order=5)
    lj = pdx.LennardJones(0.01, 2.8, 8.0)
    X, Y, Z, Q = dev(s["x"]), dev(s["y"]), dev(s["z"]), dev(s["q"])
    res = {}
    for backend in ("ell", "cluster"):
        eng = pdx.ForceEngine(N, box, lj, electrostatics=spec, charges=s["q"], skin=0.5, backend=backend)
        fx, fy, fz, E = (torch.empty(N, device="cuda") for _ in range(4))
        V = torch.empty(6, N, device="cuda")
        eng.evaluate(X, Y, Z, Q, fx, fy, fz, E, V, energy=True)
        tot = eng.totals()
        F = torch.stack([fx, fy, fz], 1).double().cpu().numpy()
        res[backend] = (F, E.double().cpu().numpy(), tot)
        if backend == "cluster":
            g = eng.capture(X, Y, Z, Q, fx, fy, fz, E, V, energy=True)
            eager = [t.clone() for t in (fx, fy, fz, E)]
            for _ in range(3):
                g.replay()
                torch.cuda.synchronize()
                for a, b in zip((fx, fy, fz, E), eager):
                    assert torch.equal(a, b)
    (F1, E1, t1), (F2, E2, t2) = res["ell"], res["cluster"]
    print(f"[clusters] engine rel rms vs ell {_relrms(F2, F1):.2e} dE {abs(t2['energy'] - t1['energy']):.2e}",
          flush=True)
    assert _relrms(F2, F1) < 1e-6
    assert abs(t2["energy"] - t1["energy"]) < 1e-6 * abs(t1["energy"])
    assert np.allclose(t2["virial"], t1["virial"], rtol=1e-6, atol=1e-6 * abs(t1["energy"]))
    assert np.allclose(E2, E1, rtol=1e-5, atol=1e-6)
