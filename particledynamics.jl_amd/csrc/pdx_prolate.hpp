// Host-side (FP64) set-up of ESP, Ewald summation with prolates: the order-zero prolate
// spheroidal wave function, the prolate split of 1/r and the prolate assignment window
// (src/electrostatics/Prolates.jl, Splits.jl:30-87,163-202, Windows.jl:51-80,113-166,272-411).
// Nothing here runs on the device; the mesh kernels receive the window's FP32 node polynomials
// and the influence kernel the split's quadrature.
#pragma once

#include <vector>

namespace pdx {

// psi_0^c on [-1, 1], normalised so that psi(0) = 1, as a series in the even Legendre
// polynomials: value = sum_k a[k] P_{2k}(x).
struct Prolate {
  double c = 0;
  std::vector<double> a;
  // Bouwkamp's method: the lowest eigenvector of the prolate operator in the normalised even
  // Legendre basis (a symmetric tridiagonal matrix). False if c is not finite and positive.
  bool make(double bandlimit);
  double value(double x) const;
  double derivative(double x) const;
  double integral(double x) const;  // int_0^x psi
  double l2_norm() const;           // ||psi|| on [-1, 1]
};

// n-point Gauss-Legendre rule on [lo, hi] (Newton iteration on P_n).
void gauss_legendre(int n, double lo, double hi, std::vector<double> &x, std::vector<double> &w);

// Bandlimit c with psi(1) / ||psi|| = eps (Prolates.jl:106-129); negative if eps is out of range.
double prolate_shape_for_tolerance(double eps);
// Widest window (grid points) of a window shape: P <= 2c/pi - 1/2 (Spec.jl:107-115).
int prolate_max_order(double shape);

// The prolate split of cutoff rc: E_s(t) = 1 - int_0^t chi, chi = psi / int_0^1 psi on [0, 1].
struct ProlateSplit {
  double rc = 0, c = 0, norm = 0, chi0 = 0, m2 = 0;
  Prolate pf;
  std::vector<double> xq, wchi;  // 128-node quadrature of c^(u) = int_0^1 chi(x) cos(u x) dx
  std::vector<double> uq, wu;    // 96-node quadrature of the long-range correction
  bool make(double cutoff, double shape);
  // corr = (1 - E_s)/r and fcorr = -(1/r) d corr/dr at s = r^2 (long_range_correction)
  void correction(double s, double &corr, double &fcorr) const;
  double self_coefficient() const { return chi0 / rc; }
  double background_coefficient() const;
};

// Per-node polynomial form of the prolate window for the FP32 device kernels. Node h < H =
// ceil(P/2) of the centred stencil carries W_h(tau) = sum_k C[h][k] tau^k with tau the offset of
// the particle from the centre of its cell; node P-1-h carries W_h(-tau). Stored split into even
// and odd parts in t = tau^2: W_h(+-tau) = E_h(t) +- tau O_h(t), dW_h/dtau(+-tau) = +-tau Ed_h(t)
// + Od_h(t), so one pass over a stored node gives the node and its mirror.
constexpr int WIN_TERMS = 7;   // terms of each part: degree <= 13 in tau
constexpr int WIN_NODES = 4;   // stored nodes: widths up to 8
struct WinCoef {
  float e[WIN_TERMS][WIN_NODES];
  float o[WIN_TERMS][WIN_NODES];
  float ed[WIN_TERMS][WIN_NODES];
  float od[WIN_TERMS][WIN_NODES];
};

struct ProlateWindow {
  int P = 0;
  double c1 = 0, edge = 0;
  Prolate pf;
  bool make(int width, double shape);
  double weight(double y) const { return pf.value(y) - edge; }  // y = (node - u) 2/P, |y| <= 1
  // phi^(theta) = (P/2) int_{-1}^{1} (psi(t) - edge) cos(theta P t / 2) dt, nquad-node rule
  double transform(double theta, int nquad) const;
  double nyquist_attenuation() const;  // (phi^(pi) / phi^(0))^2
  // The lowest even degree (6..12) whose node polynomials meet FP32 round-off (weights within
  // eps32/64, derivatives within eps32/4 of the largest): fills wc and returns the degree, or
  // -1 if degree 12 does not reach it.
  int node_polynomials(WinCoef &wc, double &err_w, double &err_d) const;
};

// Smallest Nyquist attenuation an FP32 mesh resolves: 0.02 eps32^(2/3) (min_nyquist_attenuation).
double prolate_min_attenuation_fp32();

}  // namespace pdx
