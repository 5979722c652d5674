"""FP64 CPU restatement of the ESP (Ewald summation with prolates) pieces — TEST INFRASTRUCTURE ONLY.

The checker of the device ESP path: the order-zero prolate spheroidal wave function, the prolate
split of 1/r and the prolate assignment window, restated from the reference's host code:
  * psi_0^c on [-1, 1] by Bouwkamp's method: the prolate operator is symmetric tridiagonal in the
    normalised even Legendre basis; its lowest eigenvector gives the Legendre coefficients,
    scaled so that psi(0) = 1 (ProlateFunction, src/electrostatics/Prolates.jl:26-50); values,
    derivatives and integrals from the Legendre series (Prolates.jl:52-104);
  * the tolerance rule psi(1)/||psi|| = eps (prolate_shape_for_tolerance, Prolates.jl:106-129);
  * the split chi(x) = psi(x)/int_0^1 psi on [0, 1]: short-range E_s(t) = 1 - a(t), self
    coefficient chi(0)/r_c, background pi r_c^2 m_2, the long-range correction a(t)/r and its
    force factor by Gauss-Legendre quadrature (ProlateSplit and ProlateCorrection,
    src/electrostatics/Splits.jl:30-87,163-202), the spectral function g(k) = 4 pi c^(k r_c)/k^2
    with c^(u) = int_0^1 chi(x) cos(u x) dx (spectral_g, Splits.jl:226-240);
  * the window W(y) = psi(y) - psi(1), y = (node - u) 2/P, |y| <= 1, centred stencil
    (ProlateWindow and window_stencil, src/electrostatics/Windows.jl:51-80,158-166,239-250), its
    continuum transform phi^(theta) = (P/2) int_{-1}^{1} W cos(theta P t / 2) dt, the deconvolution
    1/phi^^2 and the alias table (Windows.jl:272-283,340-411).
Everything is evaluated from the Legendre series directly (no per-node polynomial fits): the
device's FP32 node polynomials are checked against these values.

Pinned in tests/test_oracle.py: the shapes of Table 2 of Liang et al. (the reference's golden
values, test/test_electrostatics.jl:54-57), the finite-Fourier eigenfunction identity and the
derivative (:58-63), the window's symmetric weights at u = 10.5 (:88-91), E_s(0) = 1, E_s(1) = 0,
E_s(t) = 1 - chi(0) t + O(t^3) (:289-297), and the whole ESP evaluation against FP64 Ewald on the
reference's 120-charge case (:299-320), with F = -dE/dx and W_aa = -dE/d eps_aa by differences.
"""
from __future__ import annotations

import math

import numpy as np
from numpy.polynomial import legendre as npl


def gauss_legendre(n, a=-1.0, b=1.0):
    """n-point Gauss-Legendre rule on [a, b] (gauss_legendre, src/electrostatics/Quadrature.jl:13-26)."""
    x, w = npl.leggauss(int(n))
    half, mid = 0.5 * (b - a), 0.5 * (b + a)
    return mid + half * x, half * w


class ProlateFunction:
    """psi_0^c normalised to psi(0) = 1, as a series of even Legendre polynomials
    (src/electrostatics/Prolates.jl:26-50)."""

    def __init__(self, c, nterms=None):
        from scipy.linalg import eigh_tridiagonal
        if not c > 0:
            raise ValueError("Prolate bandlimit c must be positive.")
        nterms = int(math.ceil(c)) + 48 if nterms is None else int(nterms)
        cc = float(c) ** 2
        n = 2.0 * np.arange(nterms)
        diag = n * (n + 1) + cc * (2.0 * n * (n + 1) - 1.0) / ((2 * n + 3) * (2 * n - 1))
        nn = n[:-1]
        off = cc * (nn + 2) * (nn + 1) / ((2 * nn + 3) * np.sqrt((2.0 * nn + 1) * (2.0 * nn + 5)))
        _, v = eigh_tridiagonal(diag, off, select="i", select_range=(0, 0))
        beta = v[:, 0] * np.sqrt(n + 0.5)
        # P_{2m}(0) = (-1)^m (2m-1)!! / (2m)!!
        p0 = np.ones(nterms)
        for m in range(1, nterms):
            p0[m] = p0[m - 1] * -(2 * m - 1) / (2 * m)
        self.c = float(c)
        self.coeffs = beta / float((beta * p0).sum())
        leg = np.zeros(2 * nterms)
        leg[0::2] = self.coeffs
        self.leg = leg
        self.dleg = npl.legder(leg)
        self.ileg = npl.legint(leg, lbnd=0.0)

    def __call__(self, x):
        return npl.legval(np.asarray(x, dtype=np.float64), self.leg)

    def derivative(self, x):
        return npl.legval(np.asarray(x, dtype=np.float64), self.dleg)

    def integral(self, x):
        """int_0^x psi(t) dt."""
        return npl.legval(np.asarray(x, dtype=np.float64), self.ileg)

    def l2_norm(self):
        n = 2.0 * np.arange(len(self.coeffs))
        return math.sqrt(float((self.coeffs ** 2 * 2.0 / (2 * n + 1)).sum()))


def shape_for_tolerance(eps):
    """Bandlimit c with psi(1)/||psi|| = eps (bisection, src/electrostatics/Prolates.jl:106-129)."""
    if not 0 < eps < 1:
        raise ValueError("Prolate tolerance must lie in (0, 1).")
    target = math.log(eps)

    def f(c):
        pf = ProlateFunction(c)
        return math.log(abs(float(pf(1.0))) / pf.l2_norm()) - target

    lo, hi = 0.5, 80.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            lo = mid
        else:
            hi = mid
        if hi - lo < 1e-12 * hi:
            break
    return 0.5 * (lo + hi)


def max_order(shape):
    """Widest window for a window shape: P <= 2c/pi - 1/2 (max_prolate_order, Spec.jl:107-115)."""
    return int(math.floor(2 * shape / math.pi - 0.5 + 1e-9))


def order_for_tolerance(eps):
    """Window width of the ESP rule (esp_order_for_tolerance, src/workflow/Electrostatics.jl:178-189)."""
    d = -math.log10(eps)
    if d <= 3.35:
        return 5
    if d <= 4.05:
        return 6
    if d <= 4.35:
        return 7
    if d <= 5.05:
        return 8
    c1 = shape_for_tolerance(eps / 2)
    return min(max(int(math.floor(2 * c1 / math.pi - 1)), 8), 16)


class Split:
    """The prolate split of cutoff rc and shape c (src/electrostatics/Splits.jl:30-87)."""

    def __init__(self, rc, c, nquad=128):
        self.rc, self.c = float(rc), float(c)
        self.pf = ProlateFunction(c)
        self.norm = float(self.pf.integral(1.0))
        self.chi0 = 1.0 / self.norm
        xq, wq = gauss_legendre(nquad, 0.0, 1.0)
        self.xq = xq
        self.wchi = wq * self.pf(xq) / self.norm
        self.m2 = float((wq * xq ** 2 * self.pf(xq)).sum() / self.norm)
        uq, wu = gauss_legendre(96, 0.0, 1.0)
        self.uq, self.wu = uq, wu

    def Es(self, t):
        """Short-range factor E_s(t) = 1 - int_0^t chi, zero for t >= 1."""
        t = np.asarray(t, dtype=np.float64)
        return np.where(t < 1.0, 1.0 - self.pf.integral(np.minimum(t, 1.0)) / self.norm, 0.0)

    def dEs(self, t):
        t = np.asarray(t, dtype=np.float64)
        return np.where(t < 1.0, -self.pf(np.minimum(t, 1.0)) / self.norm, 0.0)

    def pair(self, r):
        """Short-range energy E_s/r and force factor f = (E_s/r - E_s'/rc)/r^2 per unit C q_i q_j."""
        r = np.asarray(r, dtype=np.float64)
        t = r / self.rc
        e = self.Es(t) / r
        f = (e - self.dEs(t) / self.rc) / (r * r)
        return np.where(t < 1.0, e, 0.0), np.where(t < 1.0, f, 0.0)

    def correction(self, s):
        """Long-range part at s = r^2: corr = a(t)/r and fcorr = -(1/r) d corr/dr by quadrature
        (long_range_correction, Splits.jl:180-200)."""
        r = math.sqrt(s)
        t = min(r / self.rc, 1.0)
        chit = float(self.pf(t)) / self.norm
        chiu = self.pf(t * self.uq) / self.norm
        corr = float((self.wu * chiu).sum()) / self.rc
        if t < 1e-4:
            h = 1e-3
            d2 = float(self.pf(h) - 2 * self.pf(0.0) + self.pf(-h)) / self.norm / h ** 2
            return corr, d2 * (1 / 3 - 1) / 2 / self.rc ** 3
        return corr, float((self.wu * (chiu - chit)).sum()) / (self.rc * r * r)

    def spectral(self, k):
        """g(k) = 4 pi c^(k rc)/k^2 and g'(k)/k (spectral_g, Splits.jl:226-240); k > 0 array."""
        k = np.asarray(k, dtype=np.float64)
        u = k * self.rc
        ux = u[..., None] * self.xq
        c = (self.wchi * np.cos(ux)).sum(-1)
        dc = -(self.wchi * self.xq * np.sin(ux)).sum(-1)
        k2 = k * k
        g = 4 * np.pi * c / k2
        gp = 4 * np.pi * (self.rc * dc / k2 - 2.0 * c / (k2 * k))
        return g, gp / k

    def self_coefficient(self):
        return self.chi0 / self.rc

    def background_coefficient(self):
        return math.pi * self.rc ** 2 * self.m2


class Window:
    """Prolate assignment window of width P grid points and shape c1 (Windows.jl:51-80)."""

    def __init__(self, P, c1):
        if not 2 <= P <= 16:
            raise ValueError("Prolate window width must be between 2 and 16 grid points.")
        self.P, self.c1 = int(P), float(c1)
        self.pf = ProlateFunction(c1)
        self.edge = float(self.pf(1.0))

    def stencil(self, u):
        """First node k0 (unwrapped) and the P weights and du-derivatives of the centred stencil
        of fractional grid coordinates u (array): k0 = floor(u) - (P/2 - 1) for even P,
        floor(u + 1/2) - (P - 1)/2 for odd P (window_stencil, Windows.jl:158-166,239-250)."""
        P = self.P
        u = np.asarray(u, dtype=np.float64)
        k0 = (np.floor(u) - (P // 2 - 1)) if P % 2 == 0 else (np.floor(u + 0.5) - (P - 1) // 2)
        nodes = k0[..., None] + np.arange(P)
        y = (nodes - u[..., None]) * (2.0 / P)
        inside = np.abs(y) <= 1.0
        yc = np.clip(y, -1.0, 1.0)
        w = np.where(inside, self.pf(yc) - self.edge, 0.0)
        dw = np.where(inside, -(2.0 / P) * self.pf.derivative(yc), 0.0)
        return k0.astype(np.int64), w, dw

    def transform(self, theta, nquad=192):
        """phi^(theta) = (P/2) int_{-1}^{1} (psi(t) - edge) cos(theta P t/2) dt (Windows.jl:272-283)."""
        x, wq = gauss_legendre(nquad)
        vals = self.pf(x) - self.edge
        half = 0.5 * self.P
        th = np.asarray(theta, dtype=np.float64)
        return half * (wq * vals * np.cos(th[..., None] * half * x)).sum(-1)

    def nyquist_attenuation(self):
        return float((self.transform(math.pi) / self.transform(0.0)) ** 2)

    def deconvolution(self, K):
        """1/phi^(2 pi m'/K)^2 for m = 0..K-1 (window_deconvolution, Windows.jl:340-352)."""
        m = np.arange(K)
        s = np.where(m <= K // 2, m, m - K)
        return 1.0 / self.transform(2 * np.pi * s / K) ** 2

    def alias_table(self, K, nmax):
        """U[m, n + nmax] = phi^(2 pi (m' + K n)/K) (window_alias_table, Windows.jl:401-411)."""
        m = np.arange(K)
        s = np.where(m <= K // 2, m, m - K)
        n = np.arange(-nmax, nmax + 1)
        return self.transform(2 * np.pi * (s[:, None] + K * n[None, :]) / K, nquad=256)
