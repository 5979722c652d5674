"""Pair forces, energies and virials against the FP64 direct sums
(test/gpu_force_reference.jl:38-80, test/pair_reference_gpu.jl:25-64) and the closed forms
of test/test_force_precision_gpu.jl:7-20 and test/test_analytic_pair_forces.jl."""
import math

import numpy as np
import pytest

import oracle
import pdx
from conftest import dev
from pdx import systems

pytestmark = pytest.mark.gpu


def _eval(x, y, z, box, model, cap=64, lanes=0, q=None, coulomb=None, skin=0.4, cutoff=None, traversal="rows"):
    import torch
    N = len(x)
    X, Y, Z = dev(x), dev(y), dev(z)
    Q = dev(q) if q is not None else None
    rc = cutoff if cutoff is not None else model.cutoff
    ws = pdx.DeviceNeighborWorkspace(N, box, rc, skin, cap)
    ws.update(X, Y, Z, Q)
    pf = pdx.PairForce(model, coulomb=coulomb, lanes=lanes, traversal=traversal)
    fx, fy, fz, E = (torch.empty(N, device="cuda") for _ in range(4))
    V = torch.empty(6, N, device="cuda")
    tot = torch.zeros(7, dtype=torch.float64, device="cuda")
    pf.compute(ws, fx, fy, fz, E, V, tot, energy=True)
    ws.status()
    F = torch.stack([fx, fy, fz], 1).double().cpu().numpy()
    return F, E.double().cpu().numpy(), V.double().cpu().numpy().T, tot.cpu().numpy()


def test_two_particle_closed_forms(cuda):
    z2 = np.zeros(2, np.float32)
    F, E, V, tot = _eval(np.array([0, 1], np.float32), z2, z2, (10.0, 10.0, 10.0), pdx.LennardJones())
    assert np.array_equal(F[:, 0], [-24.0, 24.0]) and np.all(F[:, 1:] == 0)
    assert np.array_equal(E, [0.0, 0.0])
    assert V.sum(0)[0] == 24.0 and tot[1] == 24.0
    F, E, V, tot = _eval(np.array([0, 2], np.float32), z2, z2, (10.0, 10.0, 10.0), pdx.LennardJones())
    assert F[0, 0] == pytest.approx(93 / 512, rel=1e-6) and F[1, 0] == pytest.approx(-93 / 512, rel=1e-6)
    assert E == pytest.approx([-63 / 2048] * 2, rel=1e-6)
    assert tot[1] == pytest.approx(-93 / 256, rel=1e-6)


@pytest.mark.parametrize("rfac", [0.95, 1.5])
def test_analytic_lj(cuda, rfac):
    eps, sig = 2.5, 1.2
    r = rfac * sig
    z2 = np.zeros(2, np.float32)
    x = np.array([-r / 2, r / 2], np.float32)
    F, E, _, _ = _eval(x, z2, z2, (20.0, 20.0, 20.0), pdx.LennardJones(eps, sig, 3.0))
    rr = float(x[1] - x[0])
    fref = 24 * eps * (2 * (sig / rr) ** 12 - (sig / rr) ** 6) * (-rr / rr ** 2)
    assert F[0, 0] == pytest.approx(fref, rel=2e-6)
    assert F[0, 0] == -F[1, 0]
    assert E.sum() == pytest.approx(4 * eps * ((sig / rr) ** 12 - (sig / rr) ** 6), rel=2e-6)


def test_wca_and_soft_cutoffs(cuda):
    z2 = np.zeros(2, np.float32)
    eps = 1.7
    x = np.array([-0.45, 0.45], np.float32)
    F, E, _, _ = _eval(x, z2, z2, (20.0, 20.0, 20.0), pdx.WCA(eps, 1.0), cutoff=1.5)
    rr = 0.9
    assert E.sum() == pytest.approx(4 * eps * (rr ** -12 - rr ** -6) + eps, rel=2e-6)
    x = np.array([-0.5 * 1.01 * 2 ** (1 / 6), 0.5 * 1.01 * 2 ** (1 / 6)], np.float32)
    F, E, _, _ = _eval(x, z2, z2, (20.0, 20.0, 20.0), pdx.WCA(eps, 1.0), cutoff=1.5)
    assert np.all(F == 0) and np.all(E == 0)
    eps, sig = 8.0, 1.5
    x = np.array([-0.6, 0.6], np.float32)
    F, E, _, _ = _eval(x, z2, z2, (20.0, 20.0, 20.0), pdx.SoftRepulsive(eps, sig), cutoff=2.0)
    assert F[0, 0] == pytest.approx(-(eps / sig) * (1 - 1.2 / sig), rel=2e-6)
    assert E.sum() == pytest.approx(0.5 * eps * (1 - 1.2 / sig) ** 2, rel=2e-6)


def _relrms(a, b):
    return math.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())


@pytest.mark.parametrize("kind", ["lj", "wca"])
def test_fcc_against_direct_sum(cuda, kind):
    """Force rel. RMS < 2e-5, energy and virial rel. < 1e-5 (test/test_device_engine_pairs_gpu.jl:124-125)."""
    s = systems.fcc_lattice(4000, 0.8442, 0.05, seed=2)
    model = pdx.LennardJones(1.0, 1.0, 2.5) if kind == "lj" else pdx.WCA(1.0, 1.0)
    cap = pdx.default_capacity(4000, s["box"], model.cutoff + 0.4)
    F, E, V, tot = _eval(s["x"], s["y"], s["z"], s["box"], model, cap=cap)
    ref = oracle.pair_direct(s["x"], s["y"], s["z"], s["box"], kind=kind, cutoff=2.5)
    print(f"[pairs] {kind} relrms={_relrms(F, ref['F']):.3e}", flush=True)
    assert _relrms(F, ref["F"]) < 2e-5
    Eref = ref["energy"].sum()
    assert abs(tot[0] - Eref) / abs(Eref) < 1e-5
    assert abs(E.sum() - Eref) / abs(Eref) < 1e-5
    Wref = ref["virial"].sum(0)
    assert np.abs(tot[1:] - Wref).max() / abs(Wref[:3].sum()) < 1e-5
    assert np.abs(F.sum(0)).max() / np.abs(F).max() < 1e-4


def test_lanes_agree(cuda):
    s = systems.fcc_lattice(4000, 0.8442, 0.05, seed=4)
    cap = pdx.default_capacity(4000, s["box"], 2.9)
    base = None
    for G in (1, 4, 16, 64, -2, -4, -8, -16):  # wave lanes, and the split layout over waves
        F, E, V, tot = _eval(s["x"], s["y"], s["z"], s["box"], pdx.LennardJones(), cap=cap, lanes=G)
        if base is None:
            base = (F, tot)
        else:
            assert _relrms(F, base[0]) < 1e-6
            assert abs(tot[0] - base[1][0]) < 1e-6 * abs(base[1][0])


def test_split_layout_coulomb_matches_lanes(cuda):
    """The two traversal layouts give the same forces and FP64 totals for a charged system."""
    s = systems.rock_salt(12, 2.82, 0.1, seed=5)
    coul = pdx.RealSpaceCoulomb(C=14.399645, cutoff=8.0, alpha=0.39)
    lj = pdx.LennardJones(0.01, 2.8, 8.0)
    cap = pdx.default_capacity(s["N"], s["box"], 8.5)
    F1, E1, V1, t1 = _eval(s["x"], s["y"], s["z"], s["box"], lj, cap=cap, lanes=2, q=s["q"], coulomb=coul, skin=0.5)
    F2, E2, V2, t2 = _eval(s["x"], s["y"], s["z"], s["box"], lj, cap=cap, lanes=-8, q=s["q"], coulomb=coul, skin=0.5)
    assert _relrms(F2, F1) < 1e-6
    assert np.allclose(t1, t2, rtol=1e-6, atol=1e-6 * abs(t1[0]))


def test_real_space_coulomb(cuda):
    """ChargedPairs (PairModels.jl:90-110): LJ + C q q (erfc(a r)/r) through the fitted polynomial,
    against an FP64 evaluation of erfc over the same rows."""
    rng = np.random.default_rng(11)
    N, L, rc, a, C = 500, 16.0, 4.0, 0.8, 1.7
    P = (rng.random((N, 3)) * L - L / 2).astype(np.float32)
    q = rng.standard_normal(N).astype(np.float32)
    coul = pdx.RealSpaceCoulomb(C=C, cutoff=rc, alpha=a)
    F, E, V, tot = _eval(P[:, 0], P[:, 1], P[:, 2], (L, L, L), None, cap=200, q=q, coulomb=coul, skin=0.3,
                         cutoff=rc)
    Pd = P.astype(np.float64)
    d = Pd[:, None, :] - Pd[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    r = np.sqrt(r2)
    inside = r < rc
    qq = C * np.outer(q, q).astype(np.float64)
    from scipy.special import erfc
    fr = np.where(inside, qq * (erfc(a * r) / r + 2 * a / math.sqrt(math.pi) * np.exp(-a * a * r2)) / r2, 0.0)
    Fref = (fr[:, :, None] * d).sum(1)
    e = np.where(inside, qq * erfc(a * r) / r, 0.0).sum(1) * 0.5 - C * a / math.sqrt(math.pi) * q.astype(
        np.float64) ** 2
    print(f"[coulomb] relrms={_relrms(F, Fref):.3e}", flush=True)
    assert _relrms(F, Fref) < 1e-5
    assert abs(tot[0] - e.sum()) / abs(e.sum()) < 1e-5
    info = pdx.PairForce(None, coulomb=coul).info()
    assert info["force_fit_error"] < 1e-6


@pytest.mark.parametrize("kind", ["lj", "wca"])
def test_cell_traversal_against_direct_sum(cuda, kind):
    """The list-free cell traversal on the LJ fluid against the FP64 direct sums."""
    s = systems.fcc_lattice(4000, 0.8442, 0.05, seed=2)
    model = pdx.LennardJones() if kind == "lj" else pdx.WCA()
    cap = pdx.default_capacity(4000, s["box"], model.cutoff + 0.4)
    F, E, V, tot = _eval(s["x"], s["y"], s["z"], s["box"], model, cap=cap, traversal="cells")
    ref = oracle.pair_direct(s["x"], s["y"], s["z"], s["box"], kind=kind)
    assert _relrms(F, ref["F"]) < 2e-5
    assert abs(tot[0] - ref["energy"].sum()) < 1e-5 * abs(ref["energy"].sum())
    assert np.allclose(tot[1:], ref["virial"].sum(0), rtol=1e-5, atol=1e-5 * abs(ref["virial"][:, 0].sum()))
    assert np.allclose(E, ref["energy"], rtol=1e-4, atol=1e-5)


def test_cell_traversal_matches_rows_with_coulomb(cuda):
    """Cells and rows visit the same pairs: same forces and totals up to the FP32 sum order;
    the cell traversal is deterministic (two runs give the same bits)."""
    s = systems.rock_salt(12, 2.82, 0.1, seed=5)
    coul = pdx.RealSpaceCoulomb(C=14.399645, cutoff=8.0, alpha=0.39)
    lj = pdx.LennardJones(0.01, 2.8, 8.0)
    cap = pdx.default_capacity(s["N"], s["box"], 8.5)
    args = (s["x"], s["y"], s["z"], s["box"], lj)
    F1, E1, V1, t1 = _eval(*args, cap=cap, q=s["q"], coulomb=coul, skin=0.5)
    F2, E2, V2, t2 = _eval(*args, cap=cap, q=s["q"], coulomb=coul, skin=0.5, traversal="cells")
    F3, E3, V3, t3 = _eval(*args, cap=cap, q=s["q"], coulomb=coul, skin=0.5, traversal="cells")
    print(f"[cells] rel rms vs rows {_relrms(F2, F1):.2e}", flush=True)
    assert _relrms(F2, F1) < 1e-6
    assert np.allclose(t1, t2, rtol=1e-6, atol=1e-6 * abs(t1[0]))
    assert np.allclose(E2, E1, rtol=1e-5, atol=1e-6) and np.allclose(V2, V1, rtol=1e-4, atol=1e-4)
    assert np.array_equal(F2, F3) and np.array_equal(t2, t3)


def test_cell_traversal_restrictions(cuda):
    with pytest.raises(pdx.ArgumentError):
        pdx.PairForce(pdx.WCA(), traversal="tiles")
    import torch
    p = pdx.PairForce(pdx.WCA(), traversal="cells", accumulation="fixed64")
    nl = pdx.DeviceNeighborWorkspace(8, (12.0, 12.0, 12.0), 1.2, 0.3, 64)
    x = torch.zeros(8, device="cuda")
    with pytest.raises(pdx.ArgumentError):
        p.compute(nl, x, x, x, energy=False)
