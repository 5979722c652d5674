"""Learned correction: BF16 MFMA forward and backward against the FP64 checker (with BF16
rounding emulated), the coarse-grid features, device training, and the gain of the trained
network over coarse-mesh PME alone on a held-out configuration."""
import math

import numpy as np
import pytest

import mlp as ref_mlp
import pdx
from conftest import dev
from pdx import learned, systems

pytestmark = pytest.mark.gpu


def _rand_net(seed=0):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(learned.NPARAM) * 0.25).astype(np.float32)
    mu = (rng.standard_normal(learned.NFEAT) * 0.1).astype(np.float32)
    isd = (0.5 + rng.random(learned.NFEAT)).astype(np.float32)
    return p, mu, isd


def test_forward_matches_checker(cuda):
    import torch
    p, mu, isd = _rand_net(1)
    m = learned.MLP()
    m.set_params(p)
    m.set_norm(mu, isd, 3.0)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((1000, learned.NFEAT)).astype(np.float32)
    out = torch.empty(1000, 3, device="cuda")
    m.forward(dev(x), out)
    y, _ = ref_mlp.forward(p.astype(np.float64), x, mu, isd, emulate=True)
    got = out.double().cpu().numpy()
    err = np.linalg.norm(got - 3.0 * y) / np.linalg.norm(3.0 * y)
    print(f"[mlp] forward rel err {err:.2e}", flush=True)
    assert err < 5e-3
    # added into force arrays
    fx, fy, fz = (torch.ones(1000, device="cuda") for _ in range(3))
    m.forward(dev(x), None, fx, fy, fz)
    assert torch.allclose(fx - 1, out[:, 0], atol=1e-6) and torch.allclose(fz - 1, out[:, 2], atol=1e-6)


def test_gradient_matches_checker(cuda):
    p, mu, isd = _rand_net(3)
    m = learned.MLP()
    m.set_params(p)
    m.set_norm(mu, isd, 2.0)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((1000, learned.NFEAT)).astype(np.float32)
    t = rng.standard_normal((1000, 3)).astype(np.float32)
    g, loss = m.gradient(dev(x), dev(t))
    lref, gref = ref_mlp.loss_and_grad(p.astype(np.float64), x, t, mu, isd, 2.0, emulate=True)
    gg = g.double().cpu().numpy()
    assert abs(loss.item() - lref) < 1e-4 * lref
    o = 0
    for name, shape in ref_mlp.SIZES:
        k = int(np.prod(shape))
        a, b = gg[o:o + k], gref[o:o + k]
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"[mlp] grad {name} rel err {err:.2e}", flush=True)
        assert err < 2e-2, name
        o += k
    # batches through row ids equal the contiguous rows they name
    import torch
    rows = torch.arange(999, -1, -1, dtype=torch.int32, device="cuda")
    g2, loss2 = m.gradient(dev(x), dev(t), rows=rows)
    assert abs(loss2.item() - loss.item()) < 1e-6 * loss.item()
    assert torch.allclose(g2, g, rtol=1e-3, atol=1e-6 * g.abs().max().item())


def test_features_from_charge_grid(cuda):
    import torch
    R, q = systems.random_cloud(500, (14.0, 15.0, 16.0), seed=6)
    spec = pdx.PME(coulomb_prefactor=1.0, real_cutoff=4.0, alpha=0.9, mesh=(20, 22, 24), order=4)
    s = pdx.ElectrostaticsSolver(spec, (14.0, 15.0, 16.0), q)
    X, Y, Z, Q = dev(R[:, 0]), dev(R[:, 1]), dev(R[:, 2]), dev(q)
    s.spread(X, Y, Z, Q)
    grid = s.fixed_grid().cpu().numpy()  # (K3, K2, K1)
    f = learned.mesh_features(s, X, Y, Z, Q).cpu().numpy()
    Rf = R.astype(np.float32)
    qf = q.astype(np.float32)
    K = (20, 22, 24)
    Ls = (14.0, 15.0, 16.0)
    for i in (0, 17, 333, 499):
        u = []
        for d in range(3):
            sd = np.float32(Rf[i, d]) / np.float32(Ls[d])
            sd = sd - np.floor(sd)
            u.append(np.float32(sd * np.float32(K[d])))
        b = [int(np.floor(uu)) - 3 for uu in u]
        k = 0
        for d3 in range(4):
            for d2 in range(4):
                for d1 in range(4):
                    v = grid[(b[2] + d3) % K[2], (b[1] + d2) % K[1], (b[0] + d1) % K[0]]
                    assert f[i, k] == pytest.approx(qf[i] * v, rel=1e-6, abs=1e-9)
                    k += 1
        for d in range(3):
            assert f[i, 64 + d] == pytest.approx(qf[i] * (u[d] - np.floor(u[d])), rel=1e-6, abs=1e-7)
        assert f[i, 67] == qf[i] and f[i, 68] == pytest.approx(qf[i] * qf[i], rel=1e-7)
        assert not f[i, 69:].any()
    # the support patch holds the particle's own spread charge: summed over the patch of a
    # lone charge it is q^2
    R1 = np.array([[3.3, 7.1, 9.9]])
    s1 = pdx.ElectrostaticsSolver(pdx.PME(coulomb_prefactor=1.0, real_cutoff=4.0, alpha=0.9, mesh=(20, 22, 24),
                                          order=4, net_charge="background"), (14.0, 15.0, 16.0), np.array([0.7]))
    X1, Y1, Z1, Q1 = dev(R1[:, 0]), dev(R1[:, 1]), dev(R1[:, 2]), dev(np.array([0.7]))
    s1.spread(X1, Y1, Z1, Q1)
    f1 = learned.mesh_features(s1, X1, Y1, Z1, Q1).cpu().numpy()
    assert f1[0, :64].sum() == pytest.approx(0.49, rel=1e-6)


def test_forward_from_mesh_equals_feature_rows(cuda):
    """The network fed straight from the charge grid (features formed inside its kernel) gives
    bitwise the outputs of the feature pass + the network on the stored rows, including a ragged
    last tile; the separate add reproduces the in-kernel add."""
    import torch
    L = (14.0, 15.0, 16.0)
    for n in (1000, 37):
        R, q = systems.random_cloud(n, L, seed=7 + n)
        s = pdx.ElectrostaticsSolver(pdx.PME(coulomb_prefactor=1.0, real_cutoff=4.0, alpha=0.9, mesh=(20, 22, 24),
                                             order=4, net_charge="background"), L, q)
        X, Y, Z, Q = dev(R[:, 0]), dev(R[:, 1]), dev(R[:, 2]), dev(q)
        s.spread(X, Y, Z, Q)
        p, mu, isd = _rand_net(3)
        m = learned.MLP(seed=0)
        m.set_params(p)
        m.set_norm(mu, isd, 0.7)
        a = torch.empty(n, learned.NOUT, device="cuda")
        m.forward(learned.mesh_features(s, X, Y, Z, Q), a)
        b = torch.full((n, learned.NOUT), float("nan"), device="cuda")
        m.forward_mesh(s, X, Y, Z, Q, out=b)
        assert torch.equal(a, b), f"n {n}"
        f0 = [torch.rand(n, device="cuda") for _ in range(3)]
        f1 = [t.clone() for t in f0]
        m.forward_mesh(s, X, Y, Z, Q, fx=f0[0], fy=f0[1], fz=f0[2])
        learned.MLP.add(b, *f1)
        assert all(torch.equal(u, v) for u, v in zip(f0, f1))


def test_training_fits_a_known_map(cuda):
    """The device Adam loop on the MFMA kernels drives the loss down on a learnable target."""
    import torch
    rng = np.random.default_rng(8)
    x = rng.standard_normal((20000, learned.NFEAT)).astype(np.float32)
    A = rng.standard_normal((learned.NFEAT, 3)).astype(np.float32) / 9
    t = np.tanh(x @ A) * 0.5
    m = learned.MLP(seed=1)
    F, T = dev(x), dev(t.astype(np.float32))
    mu, isd, sc = learned.fit_standardisation(F, T)
    m.set_norm(mu, isd, sc)
    _, l0 = m.gradient(F, T)
    hist = learned.train(m, F, T, steps=600, batch=4096, lr=3e-3)
    _, l1 = m.gradient(F, T)
    print(f"[mlp] loss {l0.item():.3e} -> {l1.item():.3e}", flush=True)
    assert l1.item() < 0.05 * l0.item()


def test_correction_gains_over_coarse_mesh(cuda):
    """Trained on residuals the method generates (coarse 2.0 A / P=4 vs fine 0.5 A / P=6 PME),
    the network lowers the coarse-mesh force error on a held-out configuration."""
    import torch
    C = 14.399645
    coarse = pdx.PME(coulomb_prefactor=C, real_cutoff=9.0, real_space_tolerance=1e-5, mesh_spacing=2.0, order=4)
    fine = pdx.PME(coulomb_prefactor=C, real_cutoff=9.0, real_space_tolerance=1e-5, mesh_spacing=0.5, order=6)
    syss = [systems.rock_salt(40, 2.82, 0.1, seed=s, vacancy_fraction=0.05) for s in range(4)]
    F, T = learned.residual_dataset(syss[:3], coarse, fine)
    m = learned.MLP(seed=0)
    mu, isd, sc = learned.fit_standardisation(F, T)
    m.set_norm(mu, isd, sc)
    learned.train(m, F, T, steps=4000, batch=8192, lr=2e-3)
    Fc, ec, (X, Y, Z, Q) = learned.coulomb_forces(coarse, syss[3])
    Ff, _, _ = learned.coulomb_forces(fine, syss[3])
    feats = learned.mesh_features(ec.solver, X, Y, Z, Q)
    corr = torch.empty_like(Fc)
    m.forward(feats, corr)
    e0 = ((Fc - Ff).norm() / Ff.norm()).item()
    e1 = ((Fc + corr - Ff).norm() / Ff.norm()).item()
    print(f"[learned] coarse err {e0:.3e} -> corrected {e1:.3e} (x{e0 / e1:.2f})", flush=True)
    assert e1 < 0.5 * e0
