"""Pins the FP64 CPU oracle (oracle/oracle.c) against known answers, closed forms and
invariants, chosen so that a plausible mistake (a dropped term, a sign, an index, a wrong
image) fails one of them. CPU only."""
import math

import numpy as np
import pytest

import oracle

MADELUNG_NACL = 1.747564594633182


# ---------------------------------------------------------------- neighbor rows
def test_rows_known_answers():
    # test/test_neighbor_reference.jl:1-6
    rows, _, _ = oracle.neighbor_rows((np.array([0.0, 9.5, 5.0]), np.zeros(3)), (10.0, 10.0), np.ones(3))
    assert rows == [[1], [0], []]
    rows, _, _ = oracle.neighbor_rows((np.array([0.0, 1.0]), np.zeros(2), np.zeros(2)), (10.0, 10.0, 10.0),
                                      np.array([0.5, 1.5]))
    assert rows == [[], [0]]
    rows, _, _ = oracle.neighbor_rows((np.array([0.0]), np.array([0.0])), (10.0, 10.0), np.array([1.0]))
    assert rows == [[]]


def test_rows_symmetric_and_counted():
    rng = np.random.default_rng(3)
    n, L, r = 300, 12.0, 2.2
    c = [rng.random(n) * L - L / 2 for _ in range(3)]
    rows, counts, margin = oracle.neighbor_rows(tuple(c), (L, L, L), np.full(n, r))
    for i, row in enumerate(rows):
        assert len(row) == len(set(row)) == counts[i]
        for j in row:
            assert i in rows[j]
    # independent recount with numpy minimum image
    P = np.stack(c, 1)
    d = P[None, :, :] - P[:, None, :]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    assert (counts == (r2 <= r * r).sum(1)).all()
    assert margin > 0


# ---------------------------------------------------------------- pair sums
def _two(rx, kind="lj", eps=1.0, sig=1.0, rc=2.5):
    return oracle.pair_direct(np.array([0.0, rx]), np.zeros(2), np.zeros(2), (10.0, 10.0, 10.0), kind=kind,
                              sigma=sig, epsilon=eps, cutoff=rc, ids=[0, 1])


def test_pair_closed_forms():
    # test/test_force_precision_gpu.jl:7-20
    r = _two(1.0)
    assert np.array_equal(r["F"], [[-24.0, 0, 0], [24.0, 0, 0]])
    assert np.array_equal(r["energy"], [0.0, 0.0])
    assert np.array_equal(r["virial"].sum(0), [24.0, 0, 0, 0, 0, 0])
    r = _two(2.0)
    assert np.array_equal(r["F"][:, 0], [93 / 512, -93 / 512])
    assert np.array_equal(r["energy"], [-63 / 2048, -63 / 2048])
    assert r["virial"].sum(0)[0] == -93 / 256


@pytest.mark.parametrize("rfac", [0.95, 1.5])
def test_pair_analytic_lj(rfac):
    # test/test_analytic_pair_forces.jl: eps 2.5, sigma 1.2
    eps, sig = 2.5, 1.2
    r = rfac * sig
    out = _two(r, eps=eps, sig=sig, rc=3.0)
    fx = 24 * eps * (2 * (sig / r) ** 12 - (sig / r) ** 6) * (-r / (r * r))
    assert math.isclose(out["F"][0, 0], fx, rel_tol=1e-13)
    assert math.isclose(out["energy"].sum(), 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6), rel_tol=1e-13)
    assert out["F"][:, 0].sum() == pytest.approx(0.0, abs=1e-12)


def test_pair_analytic_wca_soft():
    eps, sig = 1.7, 1.0
    r = 0.9 * sig
    out = _two(r, kind="wca", eps=eps, sig=sig)
    lj = 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6)
    assert math.isclose(out["energy"].sum(), lj + eps, rel_tol=1e-13)
    out = _two(1.01 * 2 ** (1 / 6) * sig, kind="wca", eps=eps, sig=sig)
    assert np.all(out["F"] == 0) and np.all(out["energy"] == 0)
    eps, sig = 8.0, 1.5
    r = 0.8 * sig
    out = _two(r, kind="soft", eps=eps, sig=sig)
    assert math.isclose(out["F"][0, 0], (eps / sig) * (1 - r / sig) * (-1.0), rel_tol=1e-13)
    assert math.isclose(out["energy"].sum(), 0.5 * eps * (1 - r / sig) ** 2, rel_tol=1e-13)


def test_pair_gradient_and_newton():
    rng = np.random.default_rng(5)
    n, L = 60, 6.0
    x, y, z = (rng.random(n) * L - L / 2 for _ in range(3))
    out = oracle.pair_direct(x, y, z, (L, L, L), kind="wca")
    assert np.abs(out["F"].sum(0)).max() < 1e-9 * np.abs(out["F"]).max()
    # F = -dE/dx for a smooth potential (soft repulsion), central differences
    i, h = 7, 1e-6
    e = lambda xx: oracle.pair_direct(xx, y, z, (L, L, L), kind="soft", sigma=1.3)["energy"].sum()
    f = oracle.pair_direct(x, y, z, (L, L, L), kind="soft", sigma=1.3)["F"][i, 0]
    xp, xm = x.copy(), x.copy()
    xp[i] += h
    xm[i] -= h
    assert abs(-(e(xp) - e(xm)) / (2 * h) - f) < 1e-6 * max(1.0, abs(f))


# ---------------------------------------------------------------- erfc, B-spline
def test_erfc_matches_libm():
    for x in np.linspace(0.0, 6.0, 61):
        # the erf series loses a few ulp as 1 - erf(x) shrinks toward x = 2
        assert oracle.erfc(x) == pytest.approx(math.erfc(x), rel=1e-13, abs=1e-300)


@pytest.mark.parametrize("P", [3, 4, 5, 6, 8])
def test_bspline_identities(P):
    rng = np.random.default_rng(P)
    for u in 40 * rng.random(10):
        k0, w, dw = oracle.bspline(P, u)
        assert k0 == math.floor(u) - P + 1
        assert abs(w.sum() - 1) < 1e-14 and abs(dw.sum()) < 1e-13
        _, wp, _ = oracle.bspline(P, u + 1e-7)
        _, wm, _ = oracle.bspline(P, u - 1e-7)
        if math.floor(u + 1e-7) == math.floor(u - 1e-7):
            assert np.abs((wp - wm) / 2e-7 - dw).max() < 1e-6
    # cubic B-spline closed form M_4 at t = 0.5: node weights 1/48, 23/48, 23/48, 1/48
    _, w, _ = oracle.bspline(4, 10.5)
    assert np.allclose(w, [1 / 48, 23 / 48, 23 / 48, 1 / 48], atol=1e-15)


# ---------------------------------------------------------------- Ewald
def _nacl(n=4, a=1.0):
    g = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    q = np.where(g.sum(1) % 2 == 0, 1.0, -1.0)
    return (g + 0.25) * a, q, n * a


def test_ewald_madelung():
    R, q, L = _nacl()
    E, F, W = oracle.ewald(R, q, np.diag([L] * 3))
    assert E == pytest.approx(-len(q) / 2 * MADELUNG_NACL, rel=1e-12)
    assert np.abs(F).max() < 1e-12
    assert np.trace(W) == pytest.approx(E, rel=1e-12)
    assert np.allclose(np.diag(W), E / 3, rtol=1e-12)


def _cloud(n, L, seed):
    rng = np.random.default_rng(seed)
    R = rng.random((n, 3)) * np.asarray(L) - np.asarray(L) / 2
    q = rng.standard_normal(n)
    return R, q - q.mean()


def test_ewald_alpha_independent():
    R, q = _cloud(40, (7.0, 7.5, 8.0), 2)
    H = np.diag([7.0, 7.5, 8.0])
    E1, F1, W1 = oracle.ewald(R, q, H, C_=1.3, alpha=0.6)
    E2, F2, W2 = oracle.ewald(R, q, H, C_=1.3, alpha=1.1)
    assert E1 == pytest.approx(E2, rel=1e-12)
    assert np.abs(F1 - F2).max() < 1e-11 * np.abs(F1).max()
    assert np.abs(W1 - W2).max() < 1e-11 * np.abs(np.trace(W1))


def test_ewald_gradient_and_strain():
    Ls = (7.0, 7.5, 8.0)
    R, q = _cloud(30, Ls, 9)
    H = np.diag(Ls)
    E0, F0, W0 = oracle.ewald(R, q, H, alpha=0.8)
    h = 1e-5
    for i, d in ((3, 0), (17, 1), (25, 2)):
        Rp, Rm = R.copy(), R.copy()
        Rp[i, d] += h
        Rm[i, d] -= h
        dE = (oracle.ewald(Rp, q, H, alpha=0.8)[0] - oracle.ewald(Rm, q, H, alpha=0.8)[0]) / (2 * h)
        assert abs(-dE - F0[i, d]) < 1e-7 * np.abs(F0).max()
    for a in range(3):
        eps = 1e-6
        Es = []
        for e in (eps, -eps):
            s = np.ones(3)
            s[a] += e
            Es.append(oracle.ewald(R * s, q, np.diag(np.asarray(Ls) * s), alpha=0.8)[0])
        assert -(Es[0] - Es[1]) / (2 * eps) == pytest.approx(W0[a, a], rel=1e-6, abs=1e-6 * abs(np.trace(W0)))


def test_ewald_background_and_exceptions():
    Ls = (7.0, 7.5, 8.0)
    R, q = _cloud(30, Ls, 4)
    qn = q + 0.05
    H = np.diag(Ls)
    a = 0.9
    E0, _, W0 = oracle.ewald(R, qn, H, alpha=a)
    Eb, _, Wb = oracle.ewald(R, qn, H, alpha=a, background=True)
    Ebg = -math.pi * qn.sum() ** 2 / (2 * a * a * np.prod(Ls))
    assert Eb - E0 == pytest.approx(Ebg, rel=1e-12)
    assert np.allclose(np.diag(Wb - W0), Ebg, rtol=1e-10)
    # an excluded pair removes exactly C q_i q_j / r at the minimum image
    Ex, Fx, _ = oracle.ewald(R, q, H, alpha=a, exceptions=[(2, 5, 0.0)])
    E, F, _ = oracle.ewald(R, q, H, alpha=a)
    d = R[2] - R[5]
    d -= np.asarray(Ls) * np.round(d / np.asarray(Ls))
    assert Ex - E == pytest.approx(-q[2] * q[5] / np.linalg.norm(d), rel=1e-12)


def test_ewald_sampled_equals_full():
    Ls = (7.0, 7.5, 8.0)
    R, q = _cloud(40, Ls, 11)
    E, F, W = oracle.ewald(R, q, np.diag(Ls), C_=1.4, alpha=0.7, digits=14)
    ids = [0, 5, 39]
    Fs = oracle.ewald_sampled(R[:, 0], R[:, 1], R[:, 2], q, Ls, C_=1.4, alpha=0.7, digits=14, ids=ids)
    assert np.abs(Fs - F[ids]).max() < 1e-11 * np.abs(F).max()


def test_mlp_oracle_gradient_matches_finite_differences():
    """The FP64 MLP checker's backward equals central differences of its loss."""
    import mlp
    rng = np.random.default_rng(3)
    p = rng.standard_normal(mlp.NPARAM) * 0.3
    x = rng.standard_normal((20, mlp.NIN))
    t = rng.standard_normal((20, 3))
    mu, isd = rng.standard_normal(mlp.NIN) * 0.1, 1.0 + rng.random(mlp.NIN)
    loss, g = mlp.loss_and_grad(p, x, t, mu, isd, 2.0, emulate=False)
    h = 1e-6
    ob1 = mlp.NIN * mlp.NH  # first bias of each layer
    for k in list(range(0, mlp.NPARAM, 397)) + [mlp.NPARAM - 1, ob1, ob1 + mlp.NH + mlp.NH * mlp.NH]:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mlp.loss_and_grad(pp, x, t, mu, isd, 2.0, emulate=False)[0] -
              mlp.loss_and_grad(pm, x, t, mu, isd, 2.0, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 * max(1.0, abs(g[k])), k


def test_mlp_oracle_bf16_rounding():
    import mlp
    # BF16 keeps 7 fraction bits: 1 + 2^-8 is a tie between 1 and 1 + 2^-7 (to the even 1);
    # 1 + 3 2^-8 ties between 1 + 2^-7 and 1 + 2^-6 (to the even 1 + 2^-6)
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, -3.14159, 0.0, 1.0 + 3 * 2 ** -8])
    r = mlp.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 and r[5] == 1.0 + 2 ** -6
    assert abs(r[3] + 3.140625) < 1e-12


def test_vv_oracle_dimer_matches_converged_ode():
    """Velocity Verlet of an LJ dimer from rest against a converged ODE solution of the
    relative coordinate (r'' = 2 F(r) / m): the error is O(dt^2)."""
    from scipy.integrate import solve_ivp
    r0, dt, nsteps, m = 1.3, 0.002, 500, 1.0
    L = (10.0, 10.0, 10.0)
    X, V, F, pe, ke, w = oracle.vv_steps([-r0 / 2, r0 / 2], [0.0, 0.0], [0.0, 0.0], [0.0, 0.0], [0.0, 0.0],
                                         [0.0, 0.0], L, kind="lj", cutoff=2.5, dt=dt, mass=m, nsteps=nsteps)

    def rhs(t, s):
        r = s[0]
        f = 24.0 * (2.0 * r ** -12 - r ** -6) / r  # force on the atom at +r/2 along +x
        return [s[1], 2.0 * f / m]

    sol = solve_ivp(rhs, (0.0, dt * nsteps), [r0, 0.0], rtol=1e-12, atol=1e-14)
    r_vv = X[1, 0] - X[0, 0]
    assert abs(r_vv - sol.y[0, -1]) < 2e-5
    assert abs((V[1, 0] - V[0, 0]) - sol.y[1, -1]) < 2e-4
    # the error is second order: halving dt shrinks it about four times
    X2, *_ = oracle.vv_steps([-r0 / 2, r0 / 2], [0.0, 0.0], [0.0, 0.0], [0.0, 0.0], [0.0, 0.0], [0.0, 0.0], L,
                             kind="lj", cutoff=2.5, dt=dt / 2, mass=m, nsteps=2 * nsteps)
    e1 = abs(r_vv - sol.y[0, -1])
    e2 = abs((X2[1, 0] - X2[0, 0]) - sol.y[0, -1])
    assert 3.0 < e1 / e2 < 5.0


def test_vv_oracle_reversible_and_conserving():
    rng = np.random.default_rng(5)
    n, L = 64, 5.0
    g = np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), -1).reshape(-1, 3)
    X0 = (g + 0.5) * (L / 4) - L / 2 + rng.normal(0, 0.03, (n, 3))
    V0 = rng.normal(0, 0.8, (n, 3))
    V0 -= V0.mean(0)
    box = (L, L, L)
    X1, V1, F1, pe1, ke1, _ = oracle.vv_steps(*X0.T, *V0.T, box, kind="wca", dt=0.004, nsteps=60)
    X2, V2, *_ = oracle.vv_steps(*X1.T, *(-V1).T, box, kind="wca", dt=0.004, nsteps=60)
    d = X2 - X0
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-9
    assert np.abs(V2 + V0).max() < 1e-9
    _, _, _, pe0, ke0, _ = oracle.vv_steps(*X0.T, *V0.T, box, kind="wca", dt=0.004, nsteps=0)
    assert abs((pe1 + ke1) - (pe0 + ke0)) < 2e-3 * (ke0 + abs(pe0))


def _kg_small(seed=0, nchain=6, length=20, L=9.0):
    """Small bead-spring melt (random-walk chains, bond 0.97), no bead pair closer than 0.8."""
    rng = np.random.default_rng(seed)
    pts = []
    bonds = []
    for c in range(nchain):
        while True:
            p = rng.random(3) * L - L / 2
            if all(np.linalg.norm((p - q) - L * np.round((p - q) / L)) > 0.8 for q in pts):
                break
        pts.append(p)
        for s in range(1, length):
            for _ in range(1000):
                v = rng.normal(size=3)
                q = pts[-1] + 0.97 * v / np.linalg.norm(v)
                q = q - L * np.floor(q / L + 0.5)
                if all(np.linalg.norm((q - w) - L * np.round((q - w) / L)) > 0.8 for w in pts[:-1]):
                    break
            bonds.append((len(pts) - 1, len(pts)))
            pts.append(q)
    return np.array(pts), np.array(bonds), (L, L, L)


def test_kg_oracle_pair_part_equals_direct_sums():
    R, bonds, box = _kg_small()
    d = oracle.kg_forces(*R.T, box, bonds[:0])
    ref = oracle.pair_direct(*R.T, box, kind="wca")
    assert np.abs(d["F"] - ref["F"]).max() < 1e-9 * max(1.0, np.abs(ref["F"]).max())
    assert abs(d["energy"] - ref["energy"].sum()) < 1e-9 * max(1.0, abs(ref["energy"].sum()))
    assert np.allclose(d["W"], ref["virial"].sum(0), rtol=1e-9, atol=1e-9)


def test_kg_oracle_fene_closed_form():
    k, R0, r = 30.0, 1.5, 1.1
    d = oracle.kg_forces([0.0, r], [0.0, 0.0], [0.0, 0.0], (10.0, 10.0, 10.0), np.array([[0, 1]]), epsilon=0.0)
    f = -k * r / (1 - r * r / (R0 * R0))  # force on bead 0 along +x (toward bead 1)
    assert abs(d["F"][0, 0] + f) < 1e-12 * abs(f) and abs(d["F"][1, 0] - f) < 1e-12 * abs(f)
    u = -0.5 * k * R0 * R0 * math.log(1 - r * r / (R0 * R0))
    assert abs(d["energy"] - u) < 1e-12 * u
    assert abs(d["W"][0] - (-f) * (-r)) < 1e-12 * abs(f * r)
    with pytest.raises(ValueError):
        oracle.kg_forces([0.0, 1.6], [0.0, 0.0], [0.0, 0.0], (10.0, 10.0, 10.0), np.array([[0, 1]]))


def test_kg_oracle_gradient_and_strain():
    R, bonds, box = _kg_small(seed=2, nchain=4, length=15, L=7.0)
    d = oracle.kg_forces(*R.T, box, bonds)
    h = 1e-6
    for i in (0, 7, 31, 59):
        for a in range(3):
            Rp, Rm = R.copy(), R.copy()
            Rp[i, a] += h
            Rm[i, a] -= h
            fd = -(oracle.kg_forces(*Rp.T, box, bonds)["energy"] - oracle.kg_forces(*Rm.T, box, bonds)["energy"]) / (2 * h)
            assert abs(fd - d["F"][i, a]) < 1e-5 * max(1.0, abs(d["F"][i, a]))
    for a in range(3):
        s = np.ones(3)
        s[a] += h
        bp = tuple(np.array(box) * s)
        ep = oracle.kg_forces(*(R * s).T, bp, bonds)["energy"]
        s[a] -= 2 * h
        bm = tuple(np.array(box) * s)
        em = oracle.kg_forces(*(R * s).T, bm, bonds)["energy"]
        assert abs(-(ep - em) / (2 * h) - d["W"][a]) < 1e-5 * max(1.0, abs(d["W"][a]))


# ---------------------------------------------------------------- same-algorithm CPU evaluation
import mesh as cpu_mesh  # noqa: E402  (oracle/mesh.py)

L_ES = (14.0, 15.0, 16.0)


def _cloud(n, L, seed):
    rng = np.random.default_rng(seed)
    R = rng.random((n, 3)) * np.asarray(L) - np.asarray(L) / 2
    q = rng.standard_normal(n)
    q -= q.mean()
    return R, q


def test_cell_pairs_equal_direct_sums():
    """The cell-list pair sums equal the brute-force direct sums (LJ shifted and not, WCA)."""
    rng = np.random.default_rng(5)
    n, L = 500, 9.0
    P = rng.random((n, 3)) * L - L / 2
    for kind, shift in (("lj", False), ("lj", True), ("wca", False)):
        c = oracle.cell_pairs(P[:, 0], P[:, 1], P[:, 2], None, (L, L, L), kind=kind, cutoff=2.5, shift=shift)
        d = oracle.pair_direct(P[:, 0], P[:, 1], P[:, 2], (L, L, L), kind=kind, cutoff=2.5)
        scale = np.abs(d["F"]).max()
        assert np.abs(c["F"] - d["F"]).max() < 1e-12 * scale
        E = d["energy"]
        if kind == "lj" and shift:  # the direct sums are unshifted: 1/2 u(rc) per pair end
            s6 = (1 / 2.5) ** 6
            E = E - 0.5 * 4 * (s6 * s6 - s6) * d["counts"]
        assert np.abs(c["E"] - E).max() < 1e-12 * np.abs(E).max()
        assert np.abs(c["V"] - d["virial"]).max() < 1e-12 * np.abs(d["virial"]).max()


def test_cell_pairs_coulomb_equals_brute_force():
    """Real-space Ewald term against a numpy minimum-image erfc sum; F = -dE/dx."""
    from scipy.special import erfc
    R, q = _cloud(200, L_ES, 11)
    C_, alpha, rc = 1.7, 0.9, 4.0
    c = oracle.cell_pairs(R[:, 0], R[:, 1], R[:, 2], q, L_ES, C_=C_, alpha=alpha, rc_coul=rc)
    L = np.asarray(L_ES)
    d = R[:, None, :] - R[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    inside = r < rc
    rs = np.where(inside, r, 1.0)
    qq = C_ * q[:, None] * q[None, :]
    u = np.where(inside, qq * erfc(alpha * rs) / rs, 0.0)
    a = np.where(inside, qq * (erfc(alpha * rs) / rs + 2 / math.sqrt(math.pi) * alpha
                               * np.exp(-(alpha * rs) ** 2)) / rs ** 2, 0.0)
    F = (a[:, :, None] * d).sum(1)
    assert np.abs(c["F"] - F).max() < 1e-12 * np.abs(F).max()
    assert np.abs(c["E"] - 0.5 * u.sum(1)).max() < 1e-12 * np.abs(u).max()
    W = 0.5 * np.einsum("ij,ija,ijb->ab", a, d, d)
    Vt = c["V"].sum(0)
    assert np.allclose(Vt, [W[0, 0], W[1, 1], W[2, 2], W[0, 1], W[0, 2], W[1, 2]], rtol=1e-12, atol=1e-12 * abs(W).max())


def test_cpu_spme_against_ewald():
    """The CPU SPME (real space + mesh + self) on the reference's 120-charge case
    (test/test_electrostatics.jl:299-340): force rel. RMS < 3e-5, energy < 1e-4."""
    R, q = _cloud(120, L_ES, 3)
    Er, Fr, Wr = oracle.ewald(R, q, np.diag(L_ES), C_=1.7)
    out = cpu_mesh.evaluate(R[:, 0], R[:, 1], R[:, 2], q, L_ES, 1.7, 0.9, 4.0, (48, 48, 54), 6)
    err = math.sqrt(((out["F"] - Fr) ** 2).sum() / (Fr ** 2).sum())
    assert err < 3e-5
    assert abs(out["E"] - Er) / abs(Er) < 1e-4
    W = out["W"]
    Wm = np.array([[W[0], W[3], W[4]], [W[3], W[1], W[5]], [W[4], W[5], W[2]]])
    assert np.abs(Wm - Wr).max() / abs(np.trace(Wr)) < 3e-4
    assert abs(out["Epot"].sum() - out["E"]) < 1e-10 * abs(out["E"])
    # converged: order 8 on a finer mesh reaches the reference's "< 1e-6" (electrostatics.md:231)
    fine = cpu_mesh.evaluate(R[:, 0], R[:, 1], R[:, 2], q, L_ES, 1.7, 0.9, 4.0, (72, 72, 80), 8)
    assert math.sqrt(((fine["F"] - Fr) ** 2).sum() / (Fr ** 2).sum()) < 2e-6
    # the spread conserves charge: the k = 0 sum of the grid is sum q
    g = cpu_mesh.spread(R[:, 0], R[:, 1], R[:, 2], np.abs(q), L_ES, (48, 48, 54), 6)
    assert abs(g.sum() - np.abs(q).sum()) < 1e-12 * np.abs(q).sum()


@pytest.mark.parametrize("influence", ["spme", "p3m_ad"])
def test_cpu_mesh_gradient_and_strain(influence):
    """F = -dE/dx and W_aa = -dE/d eps_aa at fixed fractional coordinates for the discrete
    energy of the same algorithm (test/test_electrostatics.jl:341-375), FP64."""
    C_, rc, alpha, K, P = 1.0, 3.8, 0.9, (24, 25, 27), 5
    L = (13.0, 13.5, 14.0)
    R, q = _cloud(80, L, 9)

    def ev(Rr, Ll):
        return cpu_mesh.evaluate(Rr[:, 0], Rr[:, 1], Rr[:, 2], q, Ll, C_, alpha, rc, K, P, influence=influence)

    base = ev(R, L)
    Fmax = np.abs(base["F"]).max()
    h = 1e-5
    for i, d in ((2, 0), (39, 1), (76, 2)):
        Rp = R.copy(); Rp[i, d] += h
        Rm = R.copy(); Rm[i, d] -= h
        fd = -(ev(Rp, L)["E"] - ev(Rm, L)["E"]) / (2 * h)
        assert abs(fd - base["F"][i, d]) < 1e-6 * Fmax, (i, d, fd, base["F"][i, d])
    W = base["W"]
    tr = abs(W[0] + W[1] + W[2])
    for a in range(3):
        eps = 1e-5
        Es = []
        for e in (eps, -eps):
            s = np.ones(3); s[a] += e
            Es.append(ev(R * s, tuple(np.asarray(L) * s))["E"])
        fd = -(Es[0] - Es[1]) / (2 * eps)
        assert abs(fd - W[a]) < 1e-6 * max(abs(W[a]), tr), (a, fd, W[a])


def test_p3m_ad_alias_convergence():
    """The optimal AD influence function's alias sums converge: the force error against FP64
    Ewald changes by < 1 % from alias order 2 to 3, and the order-2 error is the converged
    method's error. On the reference's 120-charge case at a coarse mesh the converged AD error
    is within 25 % of SPME's: both use the same B-spline assignment and analytic derivative, and
    the optimal function minimises the error averaged over uncorrelated pair positions, not the
    error of one configuration (Ballenegger, Cerda & Holm 2012, the AD and SPME curves)."""
    R, q = _cloud(120, L_ES, 3)
    _, Fr, _ = oracle.ewald(R, q, np.diag(L_ES), C_=1.7)

    def err(influence, nmax=2):
        o = cpu_mesh.evaluate(R[:, 0], R[:, 1], R[:, 2], q, L_ES, 1.7, 0.9, 4.0, (20, 20, 24), 4,
                              influence=influence, alias_order=nmax)
        return math.sqrt(((o["F"] - Fr) ** 2).sum() / (Fr ** 2).sum())

    e1, e2, e3 = err("p3m_ad", 1), err("p3m_ad", 2), err("p3m_ad", 3)
    spme = err("spme")
    assert abs(e3 - e2) < 0.01 * e2
    assert abs(e2 - e1) < 0.1 * e2
    assert e2 < 1.25 * spme


# ---------------------------------------------------------------- typed pair models
def test_pair_typed_reduces_to_uniform():
    """Mixed sigma with equal diameters (factor = rc / sigma) and a one-type matrix equal the
    uniform direct sums; a two-type matrix of the Lorentz means equals the mixed-sigma model."""
    rng = np.random.default_rng(8)
    n, L = 400, 9.0
    P = rng.random((n, 3)) * L - L / 2
    for kind in ("lj", "wca"):
        sig = 1.1
        rc = 2.5 * sig if kind == "lj" else 1.2599211 ** 0.5 * sig
        d = oracle.pair_direct(P[:, 0], P[:, 1], P[:, 2], (L, L, L), kind=kind, sigma=sig, cutoff=rc)
        m = oracle.pair_typed((P[:, 0], P[:, 1], P[:, 2]), (L, L, L), kind=kind, style="mixed",
                              sigma_particle=np.full(n, sig), factor=rc / sig)
        t = oracle.pair_typed((P[:, 0], P[:, 1], P[:, 2]), (L, L, L), kind=kind, style="matrix",
                              typeid=np.ones(n, np.int32), sigma_pair=[[sig]], epsilon_pair=[[1.0]], rcut_pair=[[rc]])
        for r in (m, t):
            assert np.allclose(r["F"], d["F"], rtol=1e-12, atol=1e-12 * np.abs(d["F"]).max())
            assert np.allclose(r["energy"], d["energy"], rtol=1e-12, atol=1e-14)
            assert np.allclose(r["virial"], d["virial"][:, :6], rtol=1e-12, atol=1e-12 * np.abs(d["virial"]).max())
    typ = rng.integers(1, 3, n).astype(np.int32)
    sp = np.array([0.9, 1.3])
    S = 0.5 * (sp[:, None] + sp[None, :])
    mixed = oracle.pair_typed((P[:, 0], P[:, 1], P[:, 2]), (L, L, L), kind="lj", style="mixed",
                              sigma_particle=sp[typ - 1], factor=2.5)
    mat = oracle.pair_typed((P[:, 0], P[:, 1], P[:, 2]), (L, L, L), kind="lj", style="matrix", typeid=typ,
                            sigma_pair=S, epsilon_pair=np.ones((2, 2)), rcut_pair=2.5 * S)
    assert np.allclose(mat["F"], mixed["F"], rtol=1e-12, atol=1e-12 * np.abs(mixed["F"]).max())


def test_pair_typed_dimer_and_gradient():
    """A dimer at the LJ minimum 2^(1/6) sigma_ij has zero force; F = -dE/dx and Newton's third
    law on an asymmetric matrix of types; 2-D equals 3-D with all z equal."""
    sp = np.array([0.8, 1.4])
    rmin = 2 ** (1 / 6) * 0.5 * (sp[0] + sp[1])
    r = oracle.pair_typed((np.array([0.0, rmin]), np.zeros(2), np.zeros(2)), (10.0, 10.0, 10.0), kind="lj",
                          style="mixed", sigma_particle=sp, factor=2.5)
    assert np.abs(r["F"]).max() < 1e-12
    rng = np.random.default_rng(2)
    n, L = 60, 6.0
    P = rng.random((n, 3)) * L - L / 2
    typ = rng.integers(1, 4, n).astype(np.int32)
    Sg = np.array([[1.0, 0.9, 1.2], [0.9, 0.8, 1.1], [1.2, 1.1, 1.3]])
    Ep = np.array([[1.0, 0.5, 2.0], [0.5, 1.5, 0.7], [2.0, 0.7, 1.2]])
    Rc = 2.2 * Sg

    def total(Q):
        o = oracle.pair_typed((Q[:, 0], Q[:, 1], Q[:, 2]), (L, L, L), kind="lj", style="matrix", typeid=typ,
                              sigma_pair=Sg, epsilon_pair=Ep, rcut_pair=Rc)
        return o

    base = total(P)
    assert np.abs(base["F"].sum(0)).max() < 1e-9 * np.abs(base["F"]).max()
    h = 1e-6
    for i, d in ((3, 0), (17, 1), (44, 2)):
        Pp = P.copy(); Pp[i, d] += h
        Pm = P.copy(); Pm[i, d] -= h
        fd = -(total(Pp)["energy"].sum() - total(Pm)["energy"].sum()) / (2 * h)
        assert abs(fd - base["F"][i, d]) < 1e-6 * np.abs(base["F"]).max()
    P2 = P.copy()
    P2[:, 2] = 0.0
    a = oracle.pair_typed((P2[:, 0], P2[:, 1]), (L, L), kind="wca", style="mixed", sigma_particle=sp[typ % 2],
                          factor=1.2599211 ** 0.5)
    b = oracle.pair_typed((P2[:, 0], P2[:, 1], P2[:, 2]), (L, L, L), kind="wca", style="mixed",
                          sigma_particle=sp[typ % 2], factor=1.2599211 ** 0.5)
    assert np.allclose(a["F"], b["F"][:, :2], rtol=1e-13, atol=1e-13) and np.allclose(a["energy"], b["energy"])


def test_kg_oracle_bond_exclusion():
    """With exclude_bonded the WCA terms of bonded pairs vanish: the difference to the full sums is
    exactly the WCA of the bonded pairs (direct two-body evaluation)."""
    rng = np.random.default_rng(4)
    n, L = 200, 12.0
    m = 6  # a jittered lattice of 216 sites with bonded pairs along x
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)[:n]
    P = (g + 0.5) * (L / m) - L / 2 + rng.normal(0, 0.05, (n, 3))
    bonds = np.array([[i, i + 1] for i in range(0, n - 1, 2)], dtype=np.int32)
    # bonded partners within the WCA range so the exclusion matters
    for i, j in bonds:
        P[j] = P[i] + np.array([0.95, 0.0, 0.0])
    P = P - L * np.floor(P / L + 0.5)
    full = oracle.kg_forces(P[:, 0], P[:, 1], P[:, 2], (L, L, L), bonds)
    ex = oracle.kg_forces(P[:, 0], P[:, 1], P[:, 2], (L, L, L), bonds, exclude_bonded=True)
    dF = full["F"] - ex["F"]
    two = np.zeros_like(dF)
    e2 = 0.0
    for i, j in bonds:
        d = P[i] - P[j]
        d -= L * np.round(d / L)
        r2 = (d * d).sum()
        if r2 < 1.2599211:
            s6 = (1 / r2) ** 3
            a = 24 * s6 * (2 * s6 - 1) / r2
            two[i] += a * d
            two[j] -= a * d
            e2 += 4 * s6 * (s6 - 1) + 1
    assert np.abs(dF - two).max() < 1e-9 * np.abs(two).max()
    assert abs((full["energy"] - ex["energy"]) - e2) < 1e-9 * abs(e2)
    assert np.abs(two).max() > 1.0


# ---------------------------------------------------------------- ESP (prolate split and window)
import prolate  # noqa: E402


def test_prolate_shapes_of_table_2():
    """Golden shapes of Table 2 of Liang et al., as the reference pins them
    (test/test_electrostatics.jl:54-57)."""
    for eps, c in ((1e-3, 9.5392), (5e-4, 10.290), (1e-4, 12.024), (5e-5, 12.762), (1e-5, 14.471)):
        assert prolate.shape_for_tolerance(eps) == pytest.approx(c, rel=5e-4)


def test_prolate_function_identities():
    """psi_0^c is an eigenfunction of the finite Fourier transform on [-1, 1] (the ratio of its
    transform to itself is constant), its derivative is the central difference, psi(0) = 1, and
    its integral is the quadrature of its values (test/test_electrostatics.jl:58-63)."""
    pf = prolate.ProlateFunction(10.0)
    x, w = prolate.gauss_legendre(160)
    ratio = lambda t: float((w * pf(x) * np.cos(10.0 * t * x)).sum() / pf(t))
    assert ratio(0.2) == pytest.approx(ratio(0.9), rel=1e-10)
    assert float(pf.derivative(0.3)) == pytest.approx(float((pf(0.3 + 1e-6) - pf(0.3 - 1e-6)) / 2e-6), rel=1e-7)
    assert float(pf(0.0)) == pytest.approx(1.0, abs=1e-14)
    xq, wq = prolate.gauss_legendre(64, 0.0, 0.7)
    assert float(pf.integral(0.7)) == pytest.approx(float((wq * pf(xq)).sum()), rel=1e-13)
    # the tolerance rule: psi(1) / ||psi|| = eps at the rule's shape
    c = prolate.shape_for_tolerance(1e-4)
    p = prolate.ProlateFunction(c)
    assert float(p(1.0)) / p.l2_norm() == pytest.approx(1e-4, rel=1e-8)


def test_prolate_window_weights():
    """The centred stencil: u = 10.5, P = 6 starts at node 8 with weights symmetric about the
    particle, w_3 = psi(1/6) - psi(1) (test/test_electrostatics.jl:88-91); weights vanish at the
    support edge, so they are continuous where the stencil base changes (even P at integer u, odd
    P at half-integer u), and dw is the u-derivative of w."""
    pw = prolate.Window(6, 12.0)
    k0, w, _ = pw.stencil(np.array([10.5]))
    assert k0[0] == 8
    assert w[0, 2] == pytest.approx(w[0, 3], rel=1e-12)
    assert w[0, 2] == pytest.approx(float(pw.pf(1 / 6)) - pw.edge, rel=1e-12)
    for P, eps in ((4, 2.5e-4), (5, 5e-4), (6, 5e-5), (8, 5e-6)):
        win = prolate.Window(P, prolate.shape_for_tolerance(eps))
        u0 = 7.0 if P % 2 == 0 else 7.5
        kL, wL, _ = win.stencil(np.array([u0 - 1e-12]))
        kR, wR, _ = win.stencil(np.array([u0]))
        nodes = {}
        for k, ww, sgn in ((kL[0], wL[0], 1), (kR[0], wR[0], -1)):
            for j in range(P):
                nodes[k + j] = nodes.get(k + j, 0.0) + sgn * ww[j]
        assert kL[0] != kR[0]
        assert max(abs(v) for v in nodes.values()) < 1e-10
        u = np.array([13.37])
        _, wp, _ = win.stencil(u + 1e-7)
        _, wm, _ = win.stencil(u - 1e-7)
        _, _, dw = win.stencil(u)
        assert np.abs((wp - wm) / 2e-7 - dw).max() < 1e-6


def test_prolate_split_identities():
    """E_s(0) = 1, E_s vanishes at and beyond the cutoff, E_s(t) = 1 - chi(0) t + O(t^3), the force
    factor is -dE/dr / r (test/test_electrostatics.jl:281-297), and the long-range correction is
    1/r - E_s/r with fcorr = -(1/r) d corr/dr."""
    sp = prolate.Split(3.0, 11.0)
    assert float(sp.Es(0.0)) == 1.0
    assert float(sp.Es(1.0)) == pytest.approx(0.0, abs=1e-15)
    e, f = sp.pair(np.array([3.0, 3.5]))
    assert (e == 0).all() and (f == 0).all()
    assert float(sp.pair(np.array([1e-3]))[0][0]) * 1e-3 == pytest.approx(1 - sp.chi0 * 1e-3 / 3.0, abs=1e-9)
    r = np.linspace(0.3, 2.9, 40)
    h = 1e-6
    dE = (sp.pair(r + h)[0] - sp.pair(r - h)[0]) / (2 * h)
    assert np.abs(-dE - sp.pair(r)[1] * r).max() < 1e-7
    for rr in (0.5, 1.7, 2.9):
        corr, fcorr = sp.correction(rr * rr)
        assert corr == pytest.approx(1 / rr - float(sp.pair(np.array([rr]))[0][0]), rel=1e-10)
        cp = sp.correction((rr + h) ** 2)[0]
        cm = sp.correction((rr - h) ** 2)[0]
        assert fcorr == pytest.approx(-(cp - cm) / (2 * h) / rr, rel=1e-6)


def test_cpu_esp_against_ewald():
    """The FP64 ESP evaluation (real space + mesh + self) on the reference's 120-charge case
    (test/test_electrostatics.jl:299-320: ESP at tolerance 1e-6, mesh 32 x 32 x 36, width 8):
    forces < 1e-5 relative RMS, energy < 1e-6, virial < 1e-4 of its trace; the per-particle
    energies sum to the total; the plain deconvolution is less accurate than the optimal one."""
    R, q = _cloud(120, L_ES, 3)
    Er, Fr, Wr = oracle.ewald(R, q, np.diag(L_ES), C_=1.7)
    c, c1 = prolate.shape_for_tolerance(1e-6), prolate.shape_for_tolerance(5e-7)
    out = cpu_mesh.evaluate_esp(R[:, 0], R[:, 1], R[:, 2], q, L_ES, 1.7, 4.0, (32, 32, 36), 8, c, c1)
    err = math.sqrt(((out["F"] - Fr) ** 2).sum() / (Fr ** 2).sum())
    assert err < 1e-5
    assert abs(out["E"] - Er) / abs(Er) < 1e-6
    W = out["W"]
    Wm = np.array([[W[0], W[3], W[4]], [W[3], W[1], W[5]], [W[4], W[5], W[2]]])
    assert np.abs(Wm - Wr).max() / abs(np.trace(Wr)) < 1e-4
    assert abs(out["Epot"].sum() - out["E"]) < 1e-10 * abs(out["E"])
    dec = cpu_mesh.evaluate_esp(R[:, 0], R[:, 1], R[:, 2], q, L_ES, 1.7, 4.0, (32, 32, 36), 8, c, c1,
                                influence="esp_deconvolution")
    assert math.sqrt(((dec["F"] - Fr) ** 2).sum() / (Fr ** 2).sum()) > err


def test_cpu_esp_gradient_and_strain():
    """F = -dE/dx and W_aa = -dE/d eps_aa of the discrete ESP energy (test/test_electrostatics.jl:
    341-375), with the net-charge background: the energy of the same algorithm is differentiated."""
    C_, rc, K, P = 1.0, 3.8, (20, 21, 22), 6
    L = (13.0, 13.5, 14.0)
    R, q = _cloud(80, L, 9)
    q = q + 0.02
    c, c1 = prolate.shape_for_tolerance(1e-4), prolate.shape_for_tolerance(5e-5)
    mesh = lambda Ll: cpu_mesh.MeshESP(K, Ll, C_, rc, P, c, c1)

    def ev(Rr, Ll, m=None):
        return cpu_mesh.evaluate_esp(Rr[:, 0], Rr[:, 1], Rr[:, 2], q, Ll, C_, rc, K, P, c, c1, background=True,
                                     mesh=m)

    m0 = mesh(L)
    base = ev(R, L, m0)
    Fmax = np.abs(base["F"]).max()
    h = 1e-5
    for i, d in ((2, 0), (39, 1), (76, 2)):
        Rp = R.copy(); Rp[i, d] += h
        Rm = R.copy(); Rm[i, d] -= h
        fd = -(ev(Rp, L, m0)["E"] - ev(Rm, L, m0)["E"]) / (2 * h)
        assert abs(fd - base["F"][i, d]) < 1e-6 * Fmax, (i, d, fd, base["F"][i, d])
    W = base["W"]
    tr = abs(W[0] + W[1] + W[2])
    for a in range(3):
        eps = 1e-5
        Es = []
        for e in (eps, -eps):
            s = np.ones(3); s[a] += e
            Ls = tuple(np.asarray(L) * s)
            Es.append(ev(R * s, Ls, mesh(Ls))["E"])
        fd = -(Es[0] - Es[1]) / (2 * eps)
        assert abs(fd - W[a]) < 1e-6 * max(abs(W[a]), tr), (a, fd, W[a])
