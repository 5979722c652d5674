// Host-side set-up of ESP (Ewald summation with prolates, Liang, Lu, Barnett, Greengard &
// Jiang): the order-zero prolate spheroidal wave function psi_0^c, the prolate split and the
// prolate assignment window, in FP64 (src/electrostatics/Prolates.jl, Splits.jl:30-87,163-202,
// Windows.jl:51-166,272-411). Only set-up code: the device kernels take the results.
#include "pdx_prolate.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace pdx {

namespace {

// P_0..P_nmax at x (P_n(1) = 1), three-term recurrence.
void legendre_values(double x, int nmax, std::vector<double> &P) {
  P.assign(nmax + 1, 0.0);
  P[0] = 1.0;
  if (nmax >= 1) P[1] = x;
  for (int n = 1; n < nmax; ++n) P[n + 1] = ((2.0 * n + 1.0) * x * P[n] - n * P[n - 1]) / (n + 1.0);
}

// Number of eigenvalues of the symmetric tridiagonal (d, e) below x (Sturm sequence).
int sturm_below(const std::vector<double> &d, const std::vector<double> &e, double x) {
  int cnt = 0;
  double q = d[0] - x;
  if (q < 0) ++cnt;
  for (size_t i = 1; i < d.size(); ++i) {
    const double prev = q != 0.0 ? q : 1e-300;
    q = (d[i] - x) - e[i - 1] * e[i - 1] / prev;
    if (q < 0) ++cnt;
  }
  return cnt;
}

// Lowest eigenpair of the symmetric tridiagonal (d, e): the eigenvalue by Sturm bisection, the
// vector by inverse iteration with a shift just below it (T - sigma I positive definite, so
// the tridiagonal elimination needs no pivoting).
bool lowest_eigenpair(const std::vector<double> &d, const std::vector<double> &e, double &lambda,
                      std::vector<double> &v) {
  const int n = (int)d.size();
  double lo = d[0], hi = d[0];
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  for (int it = 0; it < 400 && hi - lo > 4 * std::numeric_limits<double>::epsilon() * std::max(std::fabs(lo), std::fabs(hi)); ++it) {
    const double mid = 0.5 * (lo + hi);
    if (sturm_below(d, e, mid) >= 1) hi = mid;
    else lo = mid;
  }
  lambda = 0.5 * (lo + hi);
  const double sigma = lambda - 1e-10 * std::max(1.0, std::fabs(lambda));
  v.assign(n, 1.0);
  std::vector<double> cp(n), dp(n);
  for (int iter = 0; iter < 4; ++iter) {
    // solve (T - sigma I) y = v (Thomas algorithm)
    double b = d[0] - sigma;
    if (!(b > 0)) return false;
    cp[0] = n > 1 ? e[0] / b : 0.0;
    dp[0] = v[0] / b;
    for (int i = 1; i < n; ++i) {
      b = (d[i] - sigma) - e[i - 1] * cp[i - 1];
      if (!(b > 0)) return false;
      cp[i] = i + 1 < n ? e[i] / b : 0.0;
      dp[i] = (v[i] - e[i - 1] * dp[i - 1]) / b;
    }
    v[n - 1] = dp[n - 1];
    for (int i = n - 2; i >= 0; --i) v[i] = dp[i] - cp[i] * v[i + 1];
    double nrm = 0.0;
    for (double x : v) nrm += x * x;
    nrm = std::sqrt(nrm);
    if (!(nrm > 0) || !std::isfinite(nrm)) return false;
    for (double &x : v) x /= nrm;
  }
  return true;
}

}  // namespace

bool Prolate::make(double bandlimit) {
  if (!(bandlimit > 0) || !std::isfinite(bandlimit)) return false;
  c = bandlimit;
  const int nt = (int)std::ceil(c) + 48;
  const double cc = c * c;
  // the prolate operator -((1 - x^2) psi')' + c^2 x^2 psi in the normalised even Legendre basis
  // sqrt(n + 1/2) P_n, n = 0, 2, 4, ...: diagonal and the n, n + 2 coupling
  std::vector<double> d(nt), e(nt - 1);
  for (int k = 0; k < nt; ++k) {
    const double n = 2.0 * k;
    d[k] = n * (n + 1) + cc * (2.0 * n * (n + 1) - 1.0) / ((2 * n + 3) * (2 * n - 1));
    if (k + 1 < nt) e[k] = cc * (n + 2) * (n + 1) / ((2 * n + 3) * std::sqrt((2 * n + 1) * (2 * n + 5)));
  }
  double lambda;
  std::vector<double> v;
  if (!lowest_eigenpair(d, e, lambda, v)) return false;
  a.assign(nt, 0.0);
  double p0 = 1.0, psi0 = 0.0;  // P_{2m}(0) = (-1)^m (2m-1)!! / (2m)!!
  for (int k = 0; k < nt; ++k) {
    a[k] = v[k] * std::sqrt(2.0 * k + 0.5);
    if (k > 0) p0 *= -(2.0 * k - 1.0) / (2.0 * k);
    psi0 += a[k] * p0;
  }
  if (!(std::fabs(psi0) > 0)) return false;
  for (double &x : a) x /= psi0;
  return true;
}

double Prolate::value(double x) const {
  std::vector<double> P;
  legendre_values(x, 2 * ((int)a.size() - 1), P);
  double s = 0.0;
  for (size_t k = 0; k < a.size(); ++k) s += a[k] * P[2 * k];
  return s;
}

double Prolate::derivative(double x) const {
  const int nmax = 2 * ((int)a.size() - 1);
  std::vector<double> P;
  legendre_values(x, nmax, P);
  std::vector<double> dP(nmax + 1, 0.0);  // P'_{n+1} = P'_{n-1} + (2n + 1) P_n
  if (nmax >= 1) dP[1] = 1.0;
  for (int n = 1; n < nmax; ++n) dP[n + 1] = dP[n - 1] + (2.0 * n + 1.0) * P[n];
  double s = 0.0;
  for (size_t k = 0; k < a.size(); ++k) s += a[k] * dP[2 * k];
  return s;
}

double Prolate::integral(double x) const {
  // int_0^x P_0 = x; int_0^x P_n = (P_{n+1}(x) - P_{n-1}(x)) / (2n + 1) (odd values vanish at 0)
  std::vector<double> P;
  legendre_values(x, 2 * ((int)a.size() - 1) + 1, P);
  double s = a[0] * x;
  for (size_t k = 1; k < a.size(); ++k) {
    const int n = 2 * (int)k;
    s += a[k] * (P[n + 1] - P[n - 1]) / (2.0 * n + 1.0);
  }
  return s;
}

double Prolate::l2_norm() const {
  double s = 0.0;  // int P_m P_n = 2/(2n + 1) delta_mn
  for (size_t k = 0; k < a.size(); ++k) s += a[k] * a[k] * 2.0 / (4.0 * k + 1.0);
  return std::sqrt(s);
}

void gauss_legendre(int n, double lo, double hi, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const double half = 0.5 * (hi - lo), mid = 0.5 * (hi + lo);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1.0, p0 = 0.0;
      for (int k = 1; k <= n; ++k) {
        const double p2 = p0;
        p0 = p1;
        p1 = ((2.0 * k - 1.0) * z * p0 - (k - 1.0) * p2) / k;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    const double wt = 2.0 / ((1.0 - z * z) * dp * dp);
    x[i] = mid - half * z;
    x[n - 1 - i] = mid + half * z;
    w[i] = w[n - 1 - i] = half * wt;
  }
}

double prolate_shape_for_tolerance(double eps) {
  if (!(eps > 0 && eps < 1)) return -1.0;
  const double target = std::log(eps);
  auto f = [&](double c) {
    Prolate pf;
    if (!pf.make(c)) return std::numeric_limits<double>::quiet_NaN();
    return std::log(std::fabs(pf.value(1.0)) / pf.l2_norm()) - target;
  };
  double lo = 0.5, hi = 80.0;
  if (!(f(lo) > 0) || !(f(hi) < 0)) return -1.0;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    (f(mid) > 0 ? lo : hi) = mid;
    if (hi - lo < 1e-12 * hi) break;
  }
  return 0.5 * (lo + hi);
}

int prolate_max_order(double shape) { return (int)std::floor(2.0 * shape / M_PI - 0.5 + 1e-9); }

double prolate_min_attenuation_fp32() {
  return 0.02 * std::pow((double)std::numeric_limits<float>::epsilon(), 2.0 / 3.0);
}

bool ProlateSplit::make(double cutoff, double shape) {
  if (!(cutoff > 0) || !std::isfinite(cutoff)) return false;
  rc = cutoff;
  c = shape;
  if (!pf.make(shape)) return false;
  norm = pf.integral(1.0);
  if (!(norm > 0)) return false;
  chi0 = 1.0 / norm;
  std::vector<double> wq;
  gauss_legendre(128, 0.0, 1.0, xq, wq);
  wchi.resize(xq.size());
  m2 = 0.0;
  for (size_t k = 0; k < xq.size(); ++k) {
    const double chi = pf.value(xq[k]) / norm;
    wchi[k] = wq[k] * chi;
    m2 += wq[k] * xq[k] * xq[k] * chi;
  }
  gauss_legendre(96, 0.0, 1.0, uq, wu);
  return true;
}

// With a(t) = int_0^t chi: corr = a/r = (1/rc) int_0^1 chi(t u) du and fcorr = (1/(rc r^2))
// int_0^1 (chi(t u) - chi(t)) du, by quadrature (no cancellation of 1 - E_s near r = 0).
void ProlateSplit::correction(double s, double &corr, double &fcorr) const {
  const double r = std::sqrt(s);
  const double t = std::min(r / rc, 1.0);
  const double chit = pf.value(t) / norm;
  double ic = 0.0, id = 0.0;
  for (size_t k = 0; k < uq.size(); ++k) {
    const double chiu = pf.value(t * uq[k]) / norm;
    ic += wu[k] * chiu;
    id += wu[k] * (chiu - chit);
  }
  corr = ic / rc;
  if (t < 1e-4) {  // limit: chi''(0) (1/3 - 1) / 2 / rc^3
    const double h = 1e-3;
    const double d2 = (pf.value(h) - 2.0 * pf.value(0.0) + pf.value(-h)) / norm / (h * h);
    fcorr = d2 * (1.0 / 3.0 - 1.0) / 2.0 / (rc * rc * rc);
    return;
  }
  fcorr = id / (rc * r * r);
}

double ProlateSplit::background_coefficient() const { return M_PI * rc * rc * m2; }

bool ProlateWindow::make(int width, double shape) {
  if (width < 2 || width > 16) return false;
  P = width;
  c1 = shape;
  if (!pf.make(shape)) return false;
  edge = pf.value(1.0);
  return true;
}

double ProlateWindow::transform(double theta, int nquad) const {
  std::vector<double> x, w;
  gauss_legendre(nquad, -1.0, 1.0, x, w);
  const double half = 0.5 * P;
  double s = 0.0;
  for (size_t k = 0; k < x.size(); ++k) s += w[k] * (pf.value(x[k]) - edge) * std::cos(theta * half * x[k]);
  return half * s;
}

double ProlateWindow::nyquist_attenuation() const {
  const double r = transform(M_PI, 192) / transform(0.0, 192);
  return r * r;
}

int ProlateWindow::node_polynomials(WinCoef &wc, double &err_w, double &err_d) const {
  const int H = (P + 1) / 2;
  if (H > WIN_NODES) return -1;
  auto yof = [&](int h, double tau) { return (h - 0.5 * (P - 1) - tau) * 2.0 / P; };
  const double eps32 = (double)std::numeric_limits<float>::epsilon();
  const int NCHK = 97;
  double dmax = 0.0;
  for (int h = 0; h < H; ++h)
    for (int i = 0; i < NCHK; ++i) dmax = std::max(dmax, std::fabs(2.0 / P * pf.derivative(yof(h, -0.5 + i / 96.0))));
  const double tolw = std::max(eps32 / 64.0, 4e-15), told = std::max(eps32 / 4.0, 1e-12) * dmax;
  for (int deg = 6; deg <= 2 * WIN_TERMS - 2; deg += 2) {
    // per node: the interpolant at the Chebyshev-Lobatto points s_i = cos(pi i / deg) of
    // f(s) = W_h(s/2), as monomials in s, then in tau = s / 2
    std::vector<std::vector<double>> C(H, std::vector<double>(deg + 1, 0.0));
    for (int h = 0; h < H; ++h) {
      std::vector<double> v(deg + 1), ak(deg + 1);
      for (int i = 0; i <= deg; ++i) v[i] = weight(yof(h, 0.5 * std::cos(M_PI * i / deg)));
      for (int k = 0; k <= deg; ++k) {
        double s = 0.0;
        for (int i = 0; i <= deg; ++i) s += (i == 0 || i == deg ? 0.5 : 1.0) * v[i] * std::cos(M_PI * i * k / deg);
        ak[k] = 2.0 / deg * s;
      }
      ak[0] *= 0.5;
      ak[deg] *= 0.5;
      // monomial coefficients of sum_k ak T_k(s): T_{k+1} = 2 s T_k - T_{k-1}
      std::vector<double> tp(deg + 1, 0.0), tc(deg + 1, 0.0), m(deg + 1, 0.0);
      tp[0] = 1.0;
      if (deg >= 1) tc[1] = 1.0;
      for (int j = 0; j <= deg; ++j) m[j] += ak[0] * tp[j];
      for (int k = 1; k <= deg; ++k) {
        for (int j = 0; j <= deg; ++j) m[j] += ak[k] * tc[j];
        std::vector<double> tn(deg + 1, 0.0);
        for (int j = 1; j <= deg; ++j) tn[j] = 2.0 * tc[j - 1];
        for (int j = 0; j <= deg; ++j) tn[j] -= tp[j];
        tp = tc;
        tc = tn;
      }
      double p2 = 1.0;
      for (int j = 0; j <= deg; ++j, p2 *= 2.0) C[h][j] = m[j] * p2;
    }
    if (P % 2 == 1)  // the centre node is its own mirror: an even function of tau
      for (int j = 1; j <= deg; j += 2) C[H - 1][j] = 0.0;
    double ew = 0.0, ed = 0.0;
    for (int h = 0; h < H; ++h)
      for (int i = 0; i < NCHK; ++i) {
        const double tau = -0.5 + i / 96.0;
        double p = C[h][deg], dp = 0.0;
        for (int k = deg - 1; k >= 0; --k) {
          dp = dp * tau + p;
          p = p * tau + C[h][k];
        }
        ew = std::max(ew, std::fabs(p - weight(yof(h, tau))));
        ed = std::max(ed, std::fabs(dp + 2.0 / P * pf.derivative(yof(h, tau))));
      }
    if (ew <= tolw && ed <= told) {
      wc = WinCoef{};
      for (int h = 0; h < H; ++h)
        for (int j = 0; j < WIN_TERMS; ++j) {
          const int ie = 2 * j, io = 2 * j + 1, ie1 = 2 * j + 2;
          wc.e[j][h] = ie <= deg ? (float)C[h][ie] : 0.f;
          wc.o[j][h] = io <= deg ? (float)C[h][io] : 0.f;
          wc.ed[j][h] = ie1 <= deg ? (float)(ie1 * C[h][ie1]) : 0.f;
          wc.od[j][h] = io <= deg ? (float)(io * C[h][io]) : 0.f;
        }
      err_w = ew;
      err_d = ed;
      return deg;
    }
  }
  return -1;
}

}  // namespace pdx
