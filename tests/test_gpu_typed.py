"""Typed pair models on the ELL rows (src/engine/PairModels.jl:16-19,61-81): mixed sigma and
pair matrices, LJ (plain and potential-shifted) and WCA, with and without real-space Coulomb,
against the FP64 direct sums of test/pair_reference_gpu.jl:8-64 (oracle.pair_typed) within the
reference's FP32 bounds (force 2e-5 relative RMS, energy 1e-5); the explicit position source of
pdx_pair_compute; refusals of the combinations that are not built."""
import math

import numpy as np
import pytest

import oracle
import pdx
from conftest import dev

pytestmark = pytest.mark.gpu


def _system(n=3000, rho=0.75, seed=0):
    rng = np.random.default_rng(seed)
    L = (n / rho) ** (1 / 3)
    # a jittered simple cubic lattice: no close contacts
    m = int(math.ceil(n ** (1 / 3)))
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)[:n]
    P = (g + 0.5) * (L / m) - L / 2 + rng.normal(0, 0.04, (n, 3))
    P = (P - L * np.floor(P / L + 0.5)).astype(np.float32)
    return P, L, rng


def _eval(model, P, L, q=None, coulomb=None, lanes=0):
    import torch
    n = len(P)
    pf = pdx.PairForce(model, coulomb=coulomb, lanes=lanes)
    rc = pf.cutoff()
    nl = pdx.DeviceNeighborWorkspace(n, (L, L, L), rc, 0.3, pdx.default_capacity(n, (L, L, L), rc + 0.3) + 32)
    X, Y, Z = (dev(P[:, d]) for d in range(3))
    Q = dev(q) if q is not None else None
    nl.update(X, Y, Z, Q)
    f = [torch.empty(n, device="cuda") for _ in range(3)]
    E = torch.empty(n, device="cuda")
    V = torch.empty(6, n, device="cuda")
    tot = torch.zeros(7, dtype=torch.float64, device="cuda")
    pf.compute_at(nl, X, Y, Z, Q, *f, E, V, tot)
    torch.cuda.synchronize()
    nl.status()
    F = torch.stack(f, 1).double().cpu().numpy()
    return F, E.double().cpu().numpy(), V.double().cpu().numpy().T, tot.cpu().numpy(), (pf, nl, X, Y, Z, Q)


def _pair_shift_energy(P, L, rc_of, u_of):
    """-1/2 u(rc_ij) per particle and neighbour within rc_ij (numpy brute force)."""
    d = P[:, None, :].astype(np.float64) - P[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    rc = rc_of()
    inside = r2 < rc ** 2
    return -0.5 * (inside * u_of()).sum(1)


def _check(F, E, V, tot, ref_F, ref_E, ref_V):
    err = np.linalg.norm(F - ref_F) / np.linalg.norm(ref_F)
    assert err < 2e-5, err
    assert abs(tot[0] - ref_E.sum()) < 1e-5 * max(1.0, abs(ref_E.sum()))
    assert abs(E.sum() - ref_E.sum()) < 1e-5 * max(1.0, np.abs(ref_E).sum())
    Wr = ref_V.sum(0)
    assert np.abs(tot[1:] - Wr).max() < 1e-5 * max(1.0, np.abs(Wr).max())
    assert np.abs(V.sum(0) - Wr).max() < 2e-5 * max(1.0, np.abs(Wr).max())
    return err


@pytest.mark.parametrize("kind,shift", [("lj", False), ("lj", True), ("wca", False)])
def test_mixed_sigma(cuda, kind, shift):
    P, L, rng = _system(seed=1)
    n = len(P)
    sig = np.where(rng.random(n) < 0.5, 0.9, 1.25).astype(np.float32)
    if kind == "lj":
        model, factor = pdx.MixedSigmaLJ(sig, epsilon=1.3, rcut_factor=2.5, shift=shift), 2.5
    else:
        model, factor = pdx.MixedSigmaWCA(sig, epsilon=1.3), 2 ** (1 / 6)
    F, E, V, tot, _ = _eval(model, P, L)
    Pd = P.astype(np.float64)
    ref = oracle.pair_typed((Pd[:, 0], Pd[:, 1], Pd[:, 2]), (L, L, L), kind=kind, style="mixed", epsilon=1.3,
                            sigma_particle=sig.astype(np.float64), factor=factor)
    Er = ref["energy"]
    if shift:
        s6 = (1 / factor) ** 6
        Er = Er + _pair_shift_energy(P, L, lambda: factor * 0.5 * (sig[:, None] + sig[None, :]),
                                     lambda: 4 * 1.3 * (s6 * s6 - s6))
    err = _check(F, E, V, tot, ref["F"], Er, ref["virial"])
    print(f"[typed] mixed {kind} shift={shift}: force {err:.2e}", flush=True)


@pytest.mark.parametrize("kind,shift", [("lj", False), ("lj", True), ("wca", False)])
def test_pair_matrix(cuda, kind, shift):
    P, L, rng = _system(seed=2)
    n = len(P)
    typ = rng.integers(1, 4, n).astype(np.int32)
    S = np.array([[1.0, 0.9, 1.2], [0.9, 0.8, 1.1], [1.2, 1.1, 1.3]])
    Ep = np.array([[1.0, 0.5, 2.0], [0.5, 1.5, 0.7], [2.0, 0.7, 1.2]])
    Rc = (2.2 if kind == "lj" else 2 ** (1 / 6)) * S
    Rc[0, 2] = Rc[2, 0] = 2.9 if kind == "lj" else Rc[0, 2]  # one pair with its own cut
    model = (pdx.PairMatrixLJ(typ, S, Ep, Rc, shift=shift) if kind == "lj" else pdx.PairMatrixWCA(typ, S, Ep, Rc))
    F, E, V, tot, _ = _eval(model, P, L)
    Pd = P.astype(np.float64)
    ref = oracle.pair_typed((Pd[:, 0], Pd[:, 1], Pd[:, 2]), (L, L, L), kind=kind, style="matrix", typeid=typ,
                            sigma_pair=S, epsilon_pair=Ep, rcut_pair=Rc)
    Er = ref["energy"]
    if shift:
        a, b = typ[:, None] - 1, typ[None, :] - 1
        x6 = (S[a, b] / Rc[a, b]) ** 6
        Er = Er + _pair_shift_energy(P, L, lambda: Rc[a, b], lambda: 4 * Ep[a, b] * (x6 * x6 - x6))
    err = _check(F, E, V, tot, ref["F"], Er, ref["virial"])
    print(f"[typed] matrix {kind} shift={shift}: force {err:.2e}", flush=True)


def test_typed_with_coulomb(cuda):
    """A charged two-species system: mixed-sigma LJ + real-space Ewald on the same rows, against
    the typed direct sums plus the FP64 cell-list Coulomb sums (ChargedPairs, PairModels.jl:84-110)."""
    P, L, rng = _system(n=2000, rho=0.6, seed=3)
    n = len(P)
    sig = np.where(rng.random(n) < 0.5, 0.9, 1.2).astype(np.float32)
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0).astype(np.float32)
    C, alpha, rc = 1.5, 1.1, 3.0
    F, E, V, tot, _ = _eval(pdx.MixedSigmaLJ(sig, 1.0, 2.5), P, L, q=q,
                            coulomb=pdx.RealSpaceCoulomb(C=C, cutoff=rc, alpha=alpha))
    Pd = P.astype(np.float64)
    lj = oracle.pair_typed((Pd[:, 0], Pd[:, 1], Pd[:, 2]), (L, L, L), kind="lj", style="mixed",
                           sigma_particle=sig.astype(np.float64), factor=2.5)
    co = oracle.cell_pairs(Pd[:, 0], Pd[:, 1], Pd[:, 2], q.astype(np.float64), (L, L, L), C_=C, alpha=alpha,
                           rc_coul=rc)
    selfe = -C * alpha / math.sqrt(math.pi) * q.astype(np.float64) ** 2
    err = _check(F, E, V, tot, lj["F"] + co["F"], lj["energy"] + co["E"] + selfe, lj["virial"] + co["V"])
    print(f"[typed] mixed LJ + Coulomb: force {err:.2e}", flush=True)


def test_explicit_positions_and_refusals(cuda):
    import torch
    P, L, rng = _system(n=500, seed=4)
    sig = np.full(500, 1.0, np.float32)
    F, E, V, tot, (pf, nl, X, Y, Z, Q) = _eval(pdx.MixedSigmaLJ(sig), P, L)
    f = [torch.empty(500, device="cuda") for _ in range(3)]
    other = X.clone()
    with pytest.raises(pdx.ArgumentError, match="packed"):
        pf.compute_at(nl, other, Y, Z, Q, *f)
    with pytest.raises(pdx.ArgumentError):
        pdx.PairForce(pdx.MixedSigmaLJ(sig), lanes=-4)
    with pytest.raises(pdx.ArgumentError):
        pdx.PairMatrixLJ(np.array([1, 2, 4], np.int32), np.ones((3, 3)), np.ones((3, 3)), np.ones((3, 3)))
        pdx.PairForce(pdx.PairMatrixLJ(np.array([1, 2, 4], np.int32), np.ones((3, 3)), np.ones((3, 3)),
                                       np.ones((3, 3))))
    with pytest.raises(pdx.DimensionMismatch):
        pf.compute(pdx.DeviceNeighborWorkspace(400, (L, L, L), 2.5, 0.3, 64), *f)


@pytest.mark.parametrize("style", ["mixed", "matrix"])
def test_typed_cluster_tiles(cuda, style):
    """The per-type table on the cluster tiles (lj_type_table, src/clusterpairs/PairFunctor.jl:
    166-179; the mixed-sigma kind LJ_MIXED): forces, energy and virial against the typed FP64 direct
    sums, as on the rows."""
    import torch
    P, L, rng = _system(seed=5)
    n = len(P)
    if style == "mixed":
        sig = np.where(rng.random(n) < 0.5, 0.9, 1.25).astype(np.float32)
        model = pdx.MixedSigmaLJ(sig, epsilon=1.1, rcut_factor=2.5)
        kw = dict(style="mixed", epsilon=1.1, sigma_particle=sig.astype(np.float64), factor=2.5)
    else:
        typ = rng.integers(1, 4, n).astype(np.int32)
        S = np.array([[1.0, 0.9, 1.2], [0.9, 0.8, 1.1], [1.2, 1.1, 1.3]])
        Ep = np.array([[1.0, 0.5, 2.0], [0.5, 1.5, 0.7], [2.0, 0.7, 1.2]])
        Rc = 2.2 * S
        model = pdx.PairMatrixLJ(typ, S, Ep, Rc)
        kw = dict(style="matrix", typeid=typ, sigma_pair=S, epsilon_pair=Ep, rcut_pair=Rc)
    eng = pdx.ForceEngine(n, (L, L, L), model, skin=0.3, backend="cluster", prune_buffer=0.1)
    X, Y, Z = (dev(P[:, d]) for d in range(3))
    f = [torch.empty(n, device="cuda") for _ in range(3)]
    E = torch.empty(n, device="cuda")
    V = torch.empty(6, n, device="cuda")
    eng.evaluate(X, Y, Z, None, *f, E, V, energy=True)
    tot = eng.totals()
    Pd = P.astype(np.float64)
    ref = oracle.pair_typed((Pd[:, 0], Pd[:, 1], Pd[:, 2]), (L, L, L), kind="lj", **kw)
    F = torch.stack(f, 1).double().cpu().numpy()
    t = np.array([tot["pair_energy"]] + list(tot["pair_virial"]))
    err = _check(F, E.double().cpu().numpy(), V.double().cpu().numpy().T, t, ref["F"], ref["energy"], ref["virial"])
    print(f"[typed] cluster {style}: force {err:.2e}", flush=True)
