"""The fused velocity-Verlet step of ForceEngine (kick-drift fused into the list check, force
evaluation, closing kick; src/engine/Run.jl:247-294, src/engine/Fusion.jl, NVEFusion.jl:3-24):
trajectories against the FP64 Verlet oracle, graph replays equal to eager steps bitwise, the
step counter driving the age-based rebuilds, per-particle masses, energy conservation with PME."""
import numpy as np
import pytest

import oracle
import pdx
from conftest import dev
from pdx import systems

pytestmark = pytest.mark.gpu


def _lj_state(seed=0):
    import torch
    s = systems.fcc_lattice(4000, 0.8442, 0.05, seed=seed)
    v = systems.maxwell_velocities(4000, 0.72, seed=seed)
    X, Y, Z = (dev(s[k]) for k in "xyz")
    VX, VY, VZ = (dev(v[:, d]) for d in range(3))
    f = [torch.empty(4000, device="cuda") for _ in range(3)]
    return s, v, [X, Y, Z], [VX, VY, VZ], f


def test_lj_trajectory_matches_fp64_verlet(cuda):
    s, v, P, Vv, f = _lj_state()
    eng = pdx.ForceEngine(4000, s["box"], pdx.LennardJones(1.0, 1.0, 2.5), skin=0.4)
    eng.evaluate(*P, None, *f, None, None, energy=False)
    dt = 0.005
    for _ in range(20):
        eng.step(*P, None, *Vv, *f, dt)
    st = eng.status()
    X, Vr, Fr, *_ = oracle.vv_steps(s["x"], s["y"], s["z"], v[:, 0], v[:, 1], v[:, 2], s["box"], kind="lj",
                                    cutoff=2.5, dt=dt, nsteps=20)
    L = s["box"][0]
    d = np.stack([t.cpu().numpy() for t in P], 1) - X
    d -= L * np.round(d / L)
    dv = np.stack([t.cpu().numpy() for t in Vv], 1) - Vr
    F = np.stack([t.cpu().numpy() for t in f], 1)
    print(f"[step] 20 LJ steps: max|dx| {np.abs(d).max():.2e} max|dv| {np.abs(dv).max():.2e} "
          f"force {np.linalg.norm(F - Fr) / np.linalg.norm(Fr):.2e} builds {st.builds} step {st.step}", flush=True)
    assert np.abs(d).max() < 1e-4 and np.abs(dv).max() < 1e-3
    assert np.linalg.norm(F - Fr) / np.linalg.norm(Fr) < 1e-4
    assert st.step == 20  # the closing kick advanced the step counter once per step
    assert all(np.abs(t.cpu().numpy()).max() <= L / 2 for t in P)


def test_step_graph_equals_eager_and_masses(cuda):
    import torch
    s, v, P, Vv, f = _lj_state(seed=1)
    eng = pdx.ForceEngine(4000, s["box"], pdx.LennardJones(1.0, 1.0, 2.5), skin=0.4, interval=5)
    eng.evaluate(*P, None, *f, None, None, energy=False)
    ref = [t.clone() for t in P + Vv + f]
    twin = pdx.ForceEngine(4000, s["box"], pdx.LennardJones(1.0, 1.0, 2.5), skin=0.4, interval=5)
    Pt, Vt, ft = [t.clone() for t in P], [t.clone() for t in Vv], [t.clone() for t in f]
    twin.evaluate(*Pt, None, *ft, None, None, energy=False)
    eng.capture_step(*P, None, *Vv, *f, 0.005)  # one warm step, then the captured one
    for _ in range(11):
        eng.replay()
    for _ in range(12):
        twin.step(*Pt, None, *Vt, *ft, 0.005)
    torch.cuda.synchronize()
    for a, b in zip(P + Vv + f, Pt + Vt + ft):
        assert torch.equal(a, b)
    st = eng.status()
    assert st.builds >= 3, st  # interval 5: the age policy rebuilt on the device-side step count
    # per-particle masses equal to 2 behave as mass = 2
    m = torch.full((4000,), 0.5, device="cuda")
    A = [t.clone() for t in ref]
    B = [t.clone() for t in ref]
    e1 = pdx.ForceEngine(4000, s["box"], pdx.LennardJones(1.0, 1.0, 2.5), skin=0.4)
    e2 = pdx.ForceEngine(4000, s["box"], pdx.LennardJones(1.0, 1.0, 2.5), skin=0.4)
    for _ in range(5):
        e1.step(*A[:3], None, *A[3:6], *A[6:], 0.005, mass=2.0)
        e2.step(*B[:3], None, *B[3:6], *B[6:], 0.005, inv_mass=m)
    torch.cuda.synchronize()
    for a, b in zip(A, B):
        assert torch.allclose(a, b, rtol=0, atol=1e-6)


def test_ionic_steps_conserve_energy(cuda):
    """Rock salt (1000 ions) with LJ + SPME, both potential-shifted (no energy jumps at the
    cutoffs): 200 steps of 1 fs keep the total energy within 1 % of kT per ion."""
    import torch
    s = systems.rock_salt(10, 2.82, 0.1, seed=3)
    N = s["N"]
    spec = pdx.PME(coulomb_prefactor=14.399645, real_cutoff=6.0, real_space_tolerance=1e-5, mesh_spacing=0.8,
                   order=6, real_space_modifier="potential_shift")
    eng = pdx.ForceEngine(N, s["box"], pdx.LennardJones(0.01, 2.8, 6.0, shift=True), electrostatics=spec,
                          charges=s["q"], skin=0.5)
    X, Y, Z, Q = (dev(s[k]) for k in ("x", "y", "z", "q"))
    mass = torch.from_numpy(s["mass"].astype(np.float32)).cuda()
    kT = 300 * 8.617333e-5
    rng = np.random.default_rng(0)
    v = rng.normal(0, 1, (N, 3)) * np.sqrt(kT / s["mass"])[:, None]
    v -= (v * s["mass"][:, None]).sum(0) / s["mass"].sum()
    VX, VY, VZ = (dev(v[:, d]) for d in range(3))
    f = [torch.empty(N, device="cuda") for _ in range(3)]
    E = torch.empty(N, device="cuda")
    dt = 1.0 / 10.1805  # 1 fs in the A / eV / amu time unit

    def total():
        eng.evaluate(X, Y, Z, Q, *f, E, None, energy=True)
        pe = eng.totals()["energy"]
        ke = 0.5 * float((mass.double() * (VX.double() ** 2 + VY.double() ** 2 + VZ.double() ** 2)).sum())
        return pe + ke

    e0 = total()
    inv = 1.0 / mass
    for _ in range(200):
        eng.step(X, Y, Z, Q, VX, VY, VZ, *f, dt, inv_mass=inv)
    e1 = total()
    drift = abs(e1 - e0) / N
    print(f"[step] ionic 200 fs: |dE|/N = {drift:.2e} eV (kT = {kT:.3e})", flush=True)
    assert drift < 1e-2 * kT
