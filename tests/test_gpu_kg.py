"""Bead-spring melt (WCA + FENE) with the three accumulation modes of the per-particle sums:
FENE closed forms, forces / energy / virial against the FP64 oracle, order independence of
the fixed-point sums, the FENE domain latch, and the modes' restrictions."""
import math

import numpy as np
import pytest

import oracle
import pdx
from conftest import dev
from pdx import melt, systems

pytestmark = pytest.mark.gpu

MODES = ("fp32", "fixed64", "fp64")


def _eval(sysd, mode, lanes=0, bonds=None, energy=True):
    import torch
    N, box = sysd["N"], sysd["box"]
    X, Y, Z = dev(sysd["x"]), dev(sysd["y"]), dev(sysd["z"])
    b = bonds if bonds is not None else pdx.FENEBonds(N, sysd["bonds"])
    eng = pdx.ForceEngine(N, box, pdx.WCA(epsilon=1.0, sigma=1.0), skin=0.4, lanes=lanes, accumulation=mode,
                          bonds=b)
    f = [torch.empty(N, device="cuda") for _ in range(3)]
    E = torch.empty(N, device="cuda")
    V = torch.empty(6, N, device="cuda")
    eng.evaluate(X, Y, Z, None, *f, E, V, energy=energy)
    torch.cuda.synchronize()
    tot = eng.totals() if energy else None
    return np.stack([t.cpu().numpy() for t in f], 1), E.cpu().numpy(), V.cpu().numpy(), tot, b


@pytest.mark.parametrize("mode", MODES)
def test_fene_two_beads(cuda, mode):
    """Two bonded beads beyond the WCA cutoff: the FENE closed forms."""
    k, R0, r = 30.0, 1.5, 1.25
    sysd = dict(x=np.array([-r / 2, r / 2], np.float32), y=np.zeros(2, np.float32), z=np.zeros(2, np.float32),
                bonds=np.array([[0, 1]]), box=(12.0, 12.0, 12.0), N=2)
    F, E, V, tot, _ = _eval(sysd, mode)
    f = -k * r / (1 - r * r / (R0 * R0))  # on bead 0 along +x: toward bead 1
    assert F[0, 0] == pytest.approx(-f, rel=2e-6) and F[1, 0] == pytest.approx(f, rel=2e-6)
    assert abs(F[:, 1:]).max() == 0.0
    u = -0.5 * k * R0 * R0 * math.log(1 - r * r / (R0 * R0))
    assert E.sum() == pytest.approx(u, rel=2e-6)
    assert tot["pair_energy"] == pytest.approx(u, rel=2e-6)
    assert tot["pair_virial"][0] == pytest.approx(f * r, rel=2e-6)


@pytest.fixture(scope="module")
def small_melt():
    sysd = systems.kg_chains(40, length=50, density=0.85, seed=3)
    melt.push_off(sysd, iterations=200)
    ref = oracle.kg_forces(sysd["x"], sysd["y"], sysd["z"], sysd["box"], sysd["bonds"])
    return sysd, ref


def test_melt_against_fp64(cuda, small_melt):
    sysd, ref = small_melt
    errs = {}
    for mode in MODES:
        F, E, V, tot, _ = _eval(sysd, mode)
        err = np.linalg.norm(F - ref["F"]) / np.linalg.norm(ref["F"])
        errs[mode] = err
        print(f"[kg] {mode}: force rel rms {err:.3e}, energy {tot['pair_energy']:.6f} vs {ref['energy']:.6f}",
              flush=True)
        assert err < 2e-6, mode
        assert tot["pair_energy"] == pytest.approx(ref["energy"], rel=1e-6)
        for c in range(6):
            assert tot["pair_virial"][c] == pytest.approx(ref["W"][c], rel=2e-5, abs=1e-3 * abs(ref["W"][0]))
        assert E.sum() == pytest.approx(ref["energy"], rel=1e-5)
    # accumulating in more precision does not make the sums worse
    assert errs["fp64"] <= errs["fp32"] * 1.05 and errs["fixed64"] <= errs["fp32"] * 1.05


def test_fixed_point_sums_independent_of_order(cuda, small_melt):
    """Fixed-point sums are exact: any split of the rows over lanes gives the same bits."""
    sysd, _ = small_melt
    base = _eval(sysd, "fixed64", lanes=1)
    for lanes in (2, 4, 8):
        other = _eval(sysd, "fixed64", lanes=lanes)
        assert np.array_equal(base[0], other[0]), lanes
        assert np.array_equal(base[1], other[1]), lanes
        assert base[3]["pair_energy"] == other[3]["pair_energy"]


def test_fene_domain_latch(cuda):
    sysd = systems.kg_chains(3, length=10, density=0.3, seed=1)
    x = sysd["x"].copy()
    # stretch the bond 4-5 of the first chain beyond R0 along x
    L = sysd["box"][0]
    x[5] = np.float32((x[4] + 1.6 + L / 2) % L - L / 2)
    sysd["x"] = x
    sysd["y"][5] = sysd["y"][4]
    sysd["z"][5] = sysd["z"][4]
    _, _, _, _, b = _eval(sysd, "fp32", energy=False)
    with pytest.raises(pdx.DomainError, match="particle 5 "):
        b.status()
    assert b.status() == -1  # reset by the first read


def test_modes_need_pair_model_without_coulomb(cuda):
    with pytest.raises(pdx.ArgumentError):
        pdx.PairForce(pdx.WCA(), accumulation="fp16")
    p = pdx.PairForce(pdx.LennardJones(epsilon=1.0, sigma=1.0, cutoff=2.5),
                      coulomb=pdx.RealSpaceCoulomb(C=1.0, cutoff=2.5, alpha=1.0), accumulation="fixed64")
    nl = pdx.DeviceNeighborWorkspace(8, (12.0, 12.0, 12.0), 2.5, 0.3, 64)
    import torch
    x = torch.zeros(8, device="cuda")
    with pytest.raises(pdx.ArgumentError):
        p.compute(nl, x, x, x, energy=False)


def test_bonded_exclusion_against_fp64(cuda, small_melt):
    """exclude_bonded (BondExclusionFeature, src/engine/Bonded.jl:60-92): bonded beads feel only
    their FENE bond; every accumulation mode against the FP64 oracle with the same exclusion."""
    import torch
    sysd, _ = small_melt
    ref = oracle.kg_forces(sysd["x"], sysd["y"], sysd["z"], sysd["box"], sysd["bonds"], exclude_bonded=True)
    full = oracle.kg_forces(sysd["x"], sysd["y"], sysd["z"], sysd["box"], sysd["bonds"])
    assert np.abs(full["F"] - ref["F"]).max() > 1e-3 * np.abs(ref["F"]).max()  # the exclusion matters here
    N, box = sysd["N"], sysd["box"]
    X, Y, Z = dev(sysd["x"]), dev(sysd["y"]), dev(sysd["z"])
    for mode in MODES:
        eng = pdx.ForceEngine(N, box, pdx.WCA(epsilon=1.0, sigma=1.0), skin=0.4, accumulation=mode,
                              bonds=pdx.FENEBonds(N, sysd["bonds"]), exclude_bonded=True)
        f = [torch.empty(N, device="cuda") for _ in range(3)]
        E = torch.empty(N, device="cuda")
        eng.evaluate(X, Y, Z, None, *f, E, None, energy=True)
        tot = eng.totals()
        F = torch.stack(f, 1).double().cpu().numpy()
        err = np.linalg.norm(F - ref["F"]) / np.linalg.norm(ref["F"])
        print(f"[kg exclusion] {mode}: force {err:.2e}", flush=True)
        assert err < 2e-6, mode
        assert tot["pair_energy"] == pytest.approx(ref["energy"], rel=1e-6)
        for c in range(6):
            assert tot["pair_virial"][c] == pytest.approx(ref["W"][c], rel=2e-5, abs=1e-3 * abs(ref["W"][0]))
    with pytest.raises(pdx.ArgumentError):
        pdx.ForceEngine(N, box, pdx.WCA(), exclude_bonded=True)
