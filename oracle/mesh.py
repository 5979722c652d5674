"""FP64 CPU restatement of the SPME / P3M-AD mesh evaluation — TEST INFRASTRUCTURE ONLY.

The same algorithm as the device mesh, in FP64, as a checker at equal parameters (mesh, order,
alpha) and as the same-algorithm CPU baseline: B-spline spread and analytic-derivative gather
(oracle.c: or_mesh_spread / or_mesh_gather, src/electrostatics/MeshKernels.jl:247-276,313-333),
the R2C / C2R transform (scipy.fft, pocketfft), and the influence functions restated from the
reference:
  * SPME: G = (C/V) g(k) D1 D2 D3 with g(k) = 4 pi exp(-k^2/4 alpha^2)/k^2 (spectral_g,
    src/electrostatics/Splits.jl:220-224) and the B-spline deconvolution D with the odd-order
    Nyquist neighbour average (window_deconvolution, src/electrostatics/Windows.jl:283-308);
    -dG/d eps_ab = G delta_ab + (C/V) D g'(k)/k k_a k_b (_influence_ewald_kernel! and
    _store_virial!, src/electrostatics/Influence.jl:18-66);
  * P3M-AD: G = sum U^2 g k_n^2 / (sum U^2 . sum U^2 k_n^2) over alias images |n_d| <= nmax,
    U = sinc^P (window_alias_table, Windows.jl:391-399; _influence_alias_kernel!,
    Influence.jl:77-181), with its -dG/d eps_ab.
Energy E = 1/2 sum_m |S(m)|^2 G(m), virial W_ab = 1/2 sum_m |S(m)|^2 Vc_ab(m) (the half
spectrum weighted 1/2 at m1 = 0 and the x Nyquist), self energy -C alpha/sqrt(pi) sum q^2,
background -C pi Q^2 / (2 alpha^2 V) (src/electrostatics/Solver.jl:382-429).

ESP (the prolate split and window of oracle/prolate.py) runs through the same transform and the
alias-sum influence with the window's transform (MeshESP, evaluate_esp at the end of this file).

Pinned in tests/test_oracle.py: the total (real space + mesh + self) against or_ewald on the
reference's 120-charge case (test/test_electrostatics.jl:299-340), for SPME and for ESP at the
reference's ESP case (tolerance 1e-6, 32 x 32 x 36, width 8: forces < 1e-5, energy < 1e-6);
F = -dE/dx and W_aa = -dE/d eps_aa by central differences (test/test_electrostatics.jl:341-375),
for ESP with the background; charge conservation of the spread; the P3M-AD error converging in
the alias order.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

import oracle as _o


def _workers():
    return max(1, _o.num_threads())


def deconvolution(P, K):
    """1/|sum_j M_P(j) e^{2 pi i m j/K}|^2, odd-order Nyquist zero -> neighbour average."""
    # M_P at the integers 0..P-1: or_bspline's node k0 + i carries M_P(u - k0 - i); at u = P - 1
    # (k0 = 0) that is w[i] = M_P(P - 1 - i)
    _, w, _ = _o.bspline(P, float(P - 1))
    M = w[::-1].copy()
    m = np.arange(K)[:, None]
    j = np.arange(P)[None, :]
    ph = 2 * np.pi * m * j / K
    re = (M[None, :] * np.cos(ph)).sum(1)
    im = (M[None, :] * np.sin(ph)).sum(1)
    den = re ** 2 + im ** 2
    D = np.where(den > 1e-12, 1.0 / np.where(den > 1e-12, den, 1.0), np.nan)
    for k in range(K):
        if np.isnan(D[k]):
            D[k] = 0.5 * (D[(k - 1) % K] + D[(k + 1) % K])
    return D


def _modes(K, box):
    K1, K2, K3 = K
    m1 = np.arange(K1 // 2 + 1)
    m2 = np.arange(K2)
    m3 = np.arange(K3)
    s2 = np.where(m2 <= K2 // 2, m2, m2 - K2)
    s3 = np.where(m3 <= K3 // 2, m3, m3 - K3)
    # arrays indexed [m3, m2, m1] (the rfftn layout of a (K3, K2, K1) array)
    kx = (2 * np.pi * m1 / box[0])[None, None, :]
    ky = (2 * np.pi * s2 / box[1])[None, :, None]
    kz = (2 * np.pi * s3 / box[2])[:, None, None]
    ny1 = ((K1 % 2 == 0) & (m1 == K1 // 2))[None, None, :]
    ny2 = ((K2 % 2 == 0) & (m2 == K2 // 2))[None, :, None]
    ny3 = ((K3 % 2 == 0) & (m3 == K3 // 2))[:, None, None]
    return (m1, m2, m3), (kx, ky, kz), (ny1, ny2, ny3)


def influence_spme(K, box, C_, alpha, P):
    """G and Vc (6 arrays: xx yy zz xy xz yz) on the half spectrum, [m3, m2, m1]."""
    (m1, m2, m3), (kx, ky, kz), (ny1, ny2, ny3) = _modes(K, box)
    V = box[0] * box[1] * box[2]
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    zero = k2 == 0
    k2s = np.where(zero, 1.0, k2)
    g = 4 * np.pi * np.exp(-k2s / (4 * alpha ** 2)) / k2s
    gpk = g * (-2.0 / k2s - 1.0 / (2 * alpha ** 2))
    D = (deconvolution(P, K[0])[m1][None, None, :] * deconvolution(P, K[1])[m2][None, :, None]
         * deconvolution(P, K[2])[m3][:, None, None])
    G = np.where(zero, 0.0, C_ / V * g * D)
    c = np.where(zero, 0.0, C_ / V * D * gpk)
    Vc = [G + c * kx * kx, G + c * ky * ky, G + c * kz * kz,
          np.where(ny1 | ny2, 0.0, c * kx * ky), np.where(ny1 | ny3, 0.0, c * kx * kz),
          np.where(ny2 | ny3, 0.0, c * ky * kz)]
    return G, [np.broadcast_to(v, G.shape) for v in Vc]


def _alias(P, K, nmax):
    m = np.arange(K)
    s = np.where(m <= K // 2, m, m - K)
    n = np.arange(-nmax, nmax + 1)
    x = np.pi * (s[:, None] + K * n[None, :]) / K
    sn = np.where(np.abs(x) < 1e-12, 1.0, np.sin(x) / np.where(np.abs(x) < 1e-12, 1.0, x))
    return sn ** P  # [m, n + nmax]


def _gaussian_spectral(alpha):
    i4a2 = 1.0 / (4 * alpha ** 2)

    def g_of(qs):
        g = 4 * np.pi * np.exp(-qs * i4a2) / qs
        return g, g * (-2.0 / qs - 2.0 * i4a2)
    return g_of


def influence_p3m_ad(K, box, C_, alpha, P, nmax=2):
    """Optimal influence function for analytic differentiation (and its -dG/d eps_ab)."""
    K1, K2, K3 = K
    return influence_alias_ad(K, box, C_, (_alias(P, K1, nmax), _alias(P, K2, nmax), _alias(P, K3, nmax)),
                              _gaussian_spectral(alpha), nmax)


def influence_alias_ad(K, box, C_, tables, spectral, nmax):
    """AD-optimal influence G = (C/V) sum U^2 g k_n^2 / (sum U^2 . sum U^2 k_n^2) over |n_d| <= nmax
    for window tables U_d[m, n + nmax] and a spectral function g(k^2) -> (g, g'(k)/k)
    (_influence_alias_kernel!, src/electrostatics/Influence.jl:77-181), with its -dG/d eps_ab."""
    K1, K2, K3 = K
    (m1, m2, m3), (kx, ky, kz), (ny1, ny2, ny3) = _modes(K, box)
    V = box[0] * box[1] * box[2]
    U1, U2, U3 = tables[0][m1], tables[1][m2], tables[2][m3]
    shape = (K3, K2, K1 // 2 + 1)
    A = np.zeros(shape); B = np.zeros(shape); Cc = np.zeros(shape)
    a = [np.zeros(shape) for _ in range(6)]
    c = [np.zeros(shape) for _ in range(6)]
    for i1, n1 in enumerate(range(-nmax, nmax + 1)):
        q1 = kx + 2 * np.pi * K1 * n1 / box[0]
        u1 = U1[:, i1][None, None, :]
        for i2, n2 in enumerate(range(-nmax, nmax + 1)):
            q2 = ky + 2 * np.pi * K2 * n2 / box[1]
            u2 = U2[:, i2][None, :, None]
            for i3, n3 in enumerate(range(-nmax, nmax + 1)):
                q3 = kz + 2 * np.pi * K3 * n3 / box[2]
                u3 = U3[:, i3][:, None, None]
                U2t = (u1 * u2 * u3) ** 2
                B += U2t
                qq = q1 ** 2 + q2 ** 2 + q3 ** 2
                nz = qq != 0
                qs = np.where(nz, qq, 1.0)
                g, gpk = spectral(qs)
                g = np.where(nz, g, 0.0)
                gpk = np.where(nz, gpk, 0.0)
                A += U2t * g * qq
                Cc += U2t * qq
                h = U2t * (gpk * qq + 2 * g)
                h2 = 2 * U2t * nz
                qa = (q1 * q1, q2 * q2, q3 * q3, q1 * q2, q1 * q3, q2 * q3)
                for k in range(6):
                    a[k] += h * qa[k]
                    c[k] += h2 * qa[k]
    zero = (kx == 0) & (ky == 0) & (kz == 0)
    den = np.where(zero, 1.0, B * Cc)
    G = np.where(zero, 0.0, C_ / V * A / den)
    f = np.where(zero, 0.0, C_ / V / den)
    r = np.where(zero, 0.0, C_ / V * A / (den * np.where(zero, 1.0, Cc)))
    Vc = [G + f * a[0] - r * c[0], G + f * a[1] - r * c[1], G + f * a[2] - r * c[2],
          np.where(ny1 | ny2, 0.0, f * a[3] - r * c[3]), np.where(ny1 | ny3, 0.0, f * a[4] - r * c[4]),
          np.where(ny2 | ny3, 0.0, f * a[5] - r * c[5])]
    return G, Vc


def spread(x, y, z, q, box, K, P):
    """FP64 charge grid, shape (K3, K2, K1) (x fastest)."""
    n = len(x)
    xa, ya, za, qa, bx = _o._d(x), _o._d(y), _o._d(z), _o._d(q), _o._d(box)
    Ka = _o._i(K)
    grid = np.zeros(K[0] * K[1] * K[2])
    rc = _o._lib.or_mesh_spread(n, xa[1], ya[1], za[1], qa[1], bx[1], Ka[1], int(P),
                                grid.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise ValueError("unsupported order")
    return grid.reshape(K[2], K[1], K[0])


def gather(x, y, z, q, box, K, P, pot):
    n = len(x)
    xa, ya, za, qa, bx = _o._d(x), _o._d(y), _o._d(z), _o._d(q), _o._d(box)
    Ka = _o._i(K)
    pa = np.ascontiguousarray(pot, dtype=np.float64)
    F = np.zeros(3 * n)
    phi = np.zeros(n)
    _o._lib.or_mesh_gather(n, xa[1], ya[1], za[1], qa[1], bx[1], Ka[1], int(P),
                           pa.ctypes.data_as(C.POINTER(C.c_double)), F.ctypes.data_as(C.POINTER(C.c_double)),
                           phi.ctypes.data_as(C.POINTER(C.c_double)))
    return F.reshape(n, 3), phi


class Mesh:
    """Precomputed influence function of one (K, box, C, alpha, P, influence) choice."""

    def __init__(self, K, box, C_, alpha, P, influence="spme", alias_order=2):
        self.K = tuple(int(k) for k in K)
        self.box = tuple(float(b) for b in box)
        self.C, self.alpha, self.P = float(C_), float(alpha), int(P)
        if influence == "spme":
            self.G, self.Vc = influence_spme(self.K, self.box, self.C, self.alpha, self.P)
        elif influence == "p3m_ad":
            self.G, self.Vc = influence_p3m_ad(self.K, self.box, self.C, self.alpha, self.P, alias_order)
        else:
            raise ValueError("influence must be 'spme' or 'p3m_ad'")
        K1 = self.K[0]
        w = np.ones(K1 // 2 + 1)
        w[0] = 0.5
        if K1 % 2 == 0:
            w[-1] = 0.5
        self.w = w[None, None, :]

    def evaluate(self, x, y, z, q):
        """Reciprocal-space energy, forces (n, 3), per-particle 1/2 q phi, virial (6)."""
        import scipy.fft as sf
        grid = spread(x, y, z, q, self.box, self.K, self.P)
        S = sf.rfftn(grid, workers=_workers())
        b2 = self.w * (S.real ** 2 + S.imag ** 2)
        E = float((b2 * self.G).sum())
        W = np.array([float((b2 * v).sum()) for v in self.Vc])
        pot = sf.irfftn(S * self.G, s=grid.shape, workers=_workers()) * grid.size
        F, phi = gather(x, y, z, q, self.box, self.K, self.P, pot)
        return E, F, 0.5 * np.asarray(q, dtype=np.float64) * phi, W


def self_energy(q, C_, alpha):
    return -C_ * alpha / math.sqrt(math.pi) * float(np.sum(np.asarray(q, dtype=np.float64) ** 2))


def background_energy(q, C_, alpha, box):
    V = box[0] * box[1] * box[2]
    Q = float(np.sum(np.asarray(q, dtype=np.float64)))
    return -C_ * math.pi * Q * Q / (2 * alpha ** 2 * V)


def evaluate(x, y, z, q, box, C_, alpha, rc, K, P, influence="spme", alias_order=2, lj=None, background=False,
             mesh=None):
    """Full FP64 evaluation, the device engine's algorithm: cell-list real space (Coulomb, and
    LJ / WCA when lj = dict(kind, sigma, epsilon, cutoff, shift)) + mesh + self (+ background).
    Returns dict E, F (n, 3), W (6: xx yy zz xy xz yz), Epot (n), and the parts."""
    m = mesh or Mesh(K, box, C_, alpha, P, influence, alias_order)
    lj = lj or {}
    pr = _o.cell_pairs(x, y, z, q, box, kind=lj.get("kind", "none"), sigma=lj.get("sigma", 1.0),
                       epsilon=lj.get("epsilon", 1.0), cutoff=lj.get("cutoff", 0.0), shift=lj.get("shift", False),
                       C_=C_, alpha=alpha, rc_coul=rc)
    Em, Fm, Epm, Wm = m.evaluate(x, y, z, q)
    Es = self_energy(q, C_, alpha)
    qd = np.asarray(q, dtype=np.float64)
    Ep = pr["E"] + Epm - C_ * alpha / math.sqrt(math.pi) * qd ** 2
    Eb = 0.0
    Wb = np.zeros(6)
    if background:
        Eb = background_energy(q, C_, alpha, box)
        V = box[0] * box[1] * box[2]
        Ep = Ep - C_ * math.pi * float(qd.sum()) / (2 * alpha ** 2 * V) * qd
        Wb[:3] = Eb
    Wp = pr["V"].sum(0)
    return dict(E=float(pr["E"].sum()) + Em + Es + Eb, F=pr["F"] + Fm, W=Wp + Wm + Wb, Epot=Ep,
                mesh_energy=Em, self_energy=Es, background_energy=Eb, pair_energy=float(pr["E"].sum()),
                pairs=pr["pairs"])


# ---------------------------------------------------------------------------------------------
# ESP (Ewald summation with prolates): the same evaluation with the prolate split and window of
# oracle/prolate.py. Spread and gather walk each particle's centred P^3 stencil in FP64
# (_spread_lines_kernel! / _gather_kernel! with window_stencil, src/electrostatics/MeshKernels.jl:
# 247-333, Windows.jl:239-250); the influence is the AD-optimal alias sum over the window's
# transform (:esp) or the plain deconvolution g / phi^2 (:esp_deconvolution, the alias sum at
# nmax = 0); the real space sums C q_i q_j E_s(r/rc)/r over the pairs within rc (Splits.jl:91-130);
# self energy -(C/2) chi(0)/rc sum q^2 and background -C pi rc^2 m_2 Q^2/V (Splits.jl:67-86).


def _grid_coordinate(x, L, K):
    s = np.asarray(x, dtype=np.float64) / L
    s = s - np.floor(s)
    u = s * K
    return np.where(u >= K, u - K, u)


def _stencils(x, y, z, box, K, window):
    out = []
    for c, L, Kd in zip((x, y, z), box, K):
        k0, w, dw = window.stencil(_grid_coordinate(c, L, Kd))
        nodes = (k0[:, None] + np.arange(window.P)[None, :]) % Kd
        out.append((nodes, w, dw))
    return out


def esp_spread(x, y, z, q, box, K, window):
    """FP64 charge grid of the prolate window, shape (K3, K2, K1)."""
    K1, K2, K3 = K
    (n1, w1, _), (n2, w2, _), (n3, w3, _) = _stencils(x, y, z, box, K, window)
    q = np.asarray(q, dtype=np.float64)
    grid = np.zeros(K1 * K2 * K3)
    for j3 in range(window.P):
        for j2 in range(window.P):
            idx = (n3[:, j3] * K2 + n2[:, j2])[:, None] * K1 + n1
            grid += np.bincount(idx.ravel(), weights=(q[:, None] * w3[:, j3, None] * w2[:, j2, None] * w1).ravel(),
                                minlength=grid.size)
    return grid.reshape(K3, K2, K1)


def esp_gather(x, y, z, q, box, K, window, pot):
    """Forces -q grad phi (n, 3) and phi (n) from the potential grid (K3, K2, K1)."""
    K1, K2, K3 = K
    (n1, w1, d1), (n2, w2, d2), (n3, w3, d3) = _stencils(x, y, z, box, K, window)
    p = np.asarray(pot, dtype=np.float64).ravel()
    phi = np.zeros(len(x))
    g = np.zeros((len(x), 3))
    for j3 in range(window.P):
        for j2 in range(window.P):
            v = p[(n3[:, j3] * K2 + n2[:, j2])[:, None] * K1 + n1]
            a = (w1 * v).sum(1)
            b = (d1 * v).sum(1)
            phi += w3[:, j3] * w2[:, j2] * a
            g[:, 0] += w3[:, j3] * w2[:, j2] * b
            g[:, 1] += w3[:, j3] * d2[:, j2] * a
            g[:, 2] += d3[:, j3] * w2[:, j2] * a
    q = np.asarray(q, dtype=np.float64)
    F = -q[:, None] * g * (np.asarray(K, dtype=np.float64) / np.asarray(box, dtype=np.float64))[None, :]
    return F, phi


class MeshESP:
    """Influence function and window of one ESP choice (K, box, C, rc, P, c, c1)."""

    def __init__(self, K, box, C_, rc, P, split_shape, window_shape, influence="esp", alias_order=2):
        import prolate
        self.K = tuple(int(k) for k in K)
        self.box = tuple(float(b) for b in box)
        self.C = float(C_)
        self.split = prolate.Split(rc, split_shape)
        self.window = prolate.Window(P, window_shape)
        nmax = 0 if influence == "esp_deconvolution" else int(alias_order)
        tables = tuple(self.window.alias_table(k, nmax) for k in self.K) if nmax > 0 else \
            tuple(self.window.transform(2 * np.pi * np.where(np.arange(k) <= k // 2, np.arange(k), np.arange(k) - k)
                                        / k)[:, None] for k in self.K)
        self.G, self.Vc = influence_alias_ad(self.K, self.box, self.C, tables,
                                             lambda qs: self.split.spectral(np.sqrt(qs)), nmax)
        K1 = self.K[0]
        w = np.ones(K1 // 2 + 1)
        w[0] = 0.5
        if K1 % 2 == 0:
            w[-1] = 0.5
        self.w = w[None, None, :]

    def evaluate(self, x, y, z, q):
        """Mesh energy, forces (n, 3), per-particle 1/2 q phi, virial (6)."""
        import scipy.fft as sf
        grid = esp_spread(x, y, z, q, self.box, self.K, self.window)
        S = sf.rfftn(grid, workers=_workers())
        b2 = self.w * (S.real ** 2 + S.imag ** 2)
        E = float((b2 * self.G).sum())
        W = np.array([float((b2 * v).sum()) for v in self.Vc])
        pot = sf.irfftn(S * self.G, s=grid.shape, workers=_workers()) * grid.size
        F, phi = esp_gather(x, y, z, q, self.box, self.K, self.window, pot)
        return E, F, 0.5 * np.asarray(q, dtype=np.float64) * phi, W


def esp_real_space(x, y, z, q, box, C_, split):
    """C q_i q_j E_s(r/rc)/r over the pairs within rc (minimum image, rc < L/2): per-particle
    half energies, forces (n, 3) and the virial (6: xx yy zz xy xz yz)."""
    from scipy.spatial import cKDTree
    L = np.asarray(box, dtype=np.float64)
    R = np.stack([np.asarray(c, dtype=np.float64) for c in (x, y, z)], 1)
    Rw = R - L * np.floor(R / L)
    Rw = np.where(Rw >= L, Rw - L, Rw)
    pairs = cKDTree(Rw, boxsize=L).query_pairs(split.rc, output_type="ndarray")
    n = len(R)
    E = np.zeros(n)
    F = np.zeros((n, 3))
    W = np.zeros(6)
    if len(pairs):
        i, j = pairs[:, 0], pairs[:, 1]
        d = R[i] - R[j]
        d -= L * np.round(d / L)
        r = np.sqrt((d * d).sum(1))
        e, f = split.pair(r)
        qq = C_ * np.asarray(q, dtype=np.float64)[i] * np.asarray(q, dtype=np.float64)[j]
        fv = (qq * f)[:, None] * d
        for c in range(3):
            F[:, c] += np.bincount(i, weights=fv[:, c], minlength=n) - np.bincount(j, weights=fv[:, c], minlength=n)
        eh = 0.5 * qq * e
        E += np.bincount(i, weights=eh, minlength=n) + np.bincount(j, weights=eh, minlength=n)
        for k, (a, b) in enumerate(((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))):
            W[k] = float((qq * f * d[:, a] * d[:, b]).sum())
    return E, F, W


def evaluate_esp(x, y, z, q, box, C_, rc, K, P, split_shape, window_shape, influence="esp", alias_order=2,
                 background=False, mesh=None):
    """Full FP64 ESP Coulomb evaluation: real space + mesh + self (+ background). Returns dict E,
    F (n, 3), W (6), Epot (n) and the parts."""
    m = mesh or MeshESP(K, box, C_, rc, P, split_shape, window_shape, influence, alias_order)
    qd = np.asarray(q, dtype=np.float64)
    Er, Fr, Wr = esp_real_space(x, y, z, q, box, C_, m.split)
    Em, Fm, Epm, Wm = m.evaluate(x, y, z, q)
    sc = m.split.self_coefficient()
    Es = -0.5 * C_ * sc * float((qd ** 2).sum())
    Ep = Er + Epm - 0.5 * C_ * sc * qd ** 2
    Eb = 0.0
    Wb = np.zeros(6)
    if background:
        V = box[0] * box[1] * box[2]
        Q = float(qd.sum())
        Eb = -C_ * m.split.background_coefficient() * Q * Q / V
        Ep = Ep - C_ * m.split.background_coefficient() * Q / V * qd
        Wb[:3] = Eb
    return dict(E=float(Er.sum()) + Em + Es + Eb, F=Fr + Fm, W=Wr + Wm + Wb, Epot=Ep, mesh_energy=Em,
                self_energy=Es, background_energy=Eb, pair_energy=float(Er.sum()))
