/*
 * FP64 CPU oracle for the force hot path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's independent oracles, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's CPU baseline as the checker; nothing in the
 * product links or calls it. Every function cites the reference file it follows:
 *
 *   or_neighbor_rows  test/neighbor_reference.jl:6-16        (rows by brute force)
 *   or_pair_direct    test/gpu_force_reference.jl:38-80,      (direct FP64 pair sums)
 *                     test/pair_reference_gpu.jl:25-64
 *   or_pair_typed     test/pair_reference_gpu.jl:8-64        (mixed sigma, pair matrix, 2-D)
 *   or_ewald          test/ewald_reference.jl:50-126         (FP64 Ewald, general cell)
 *   or_erfc           test/ewald_reference.jl:20-40          (series / continued fraction)
 *   or_ewald_sampled  test/ewald_reference.jl:50-126 for an orthorhombic box, forces on a
 *                     sample only (minimum image + the images rcut reaches), for large N
 *   or_bspline        src/electrostatics/Windows.jl:188-223  (B-spline weights, FP64)
 *   or_cell_pairs     LJ / WCA (src/engine/PairModels.jl:13-54) + real-space Ewald
 *                     (src/electrostatics/RealSpace.jl:36-118) over a cell list, FP64
 *   or_mesh_spread,   B-spline spread and analytic-derivative gather
 *   or_mesh_gather    (src/electrostatics/MeshKernels.jl:247-276, 313-333), FP64; with the
 *                     influence function and FFT of oracle.py (mesh.py) the CPU SPME / P3M-AD
 *   or_kg_forces      Kremer-Grest melt: WCA over all pairs (bonded pairs included) by cells
 *                     plus FENE bonds (_fene3_EV!, src/bonded/FENEBondCUDA.jl), FP64
 *
 * Pinning (what each function is checked against, in tests/test_oracle.py):
 *   rows    — the three known answers of test/test_neighbor_reference.jl:1-6; symmetry
 *             j in row(i) <=> i in row(j) for a uniform radius; brute-force recount.
 *   pairs   — the closed forms of test/test_force_precision_gpu.jl:7-20 (r = 1, 2) and
 *             test/test_analytic_pair_forces.jl:1-113 (eps 2.5, sigma 1.2; WCA; soft);
 *             Newton's third law; F = -dE/dx by central differences.
 *   ewald_sampled — equals or_ewald's forces on the sampled particles (tests/test_oracle.py).
 *   ewald   — the rock-salt Madelung constant 1.747564594633 (textbook closed form),
 *             independence of the result from alpha (a wrong term breaks it), F = -dE/dx,
 *             W_aa = -dE/d eps_aa by strain differences, background closed form.
 *   erfc    — libm erfc to 1e-13 relative on [0, 6].
 *   bspline — partition of unity, sum dw = 0, derivative by differences, M_4 closed form.
 *   kg      — WCA part equals or_pair_direct; FENE two-bead closed form; F = -dE/dx and the
 *             virial W_aa = -dE/d eps_aa by central differences.
 *   typed   — style 2 with equal diameters and style 3 with one type equal or_pair_direct;
 *             a two-type matrix of the Lorentz means equals style 2; a dimer at
 *             r = 2^(1/6) sigma_ij has zero force; F = -dE/dx; Newton's third law; 2-D equals
 *             3-D with every z equal.
 *   cell_pairs — LJ / WCA part equals or_pair_direct; the Coulomb part equals a brute-force
 *             minimum-image erfc sum; F = -dE/dx by central differences.
 *   mesh    — the CPU SPME total (cell_pairs + mesh + self) against or_ewald on the
 *             reference's 120-charge case; F = -dE/dx and W_aa = -dE/d eps_aa by central
 *             differences for SPME and P3M-AD; spread conserves charge; the P3M-AD error
 *             converges in the alias order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static double jl_round(double x) { return nearbyint(x); } /* Julia round: ties to even */

/* ---------------------------------------------------------------- rows */
/* Row i = { j != i : sum_d (x_j - x_i - L round((x_j - x_i)/L))^2 <= r_i^2 } in ascending j
 * (0-based here). counts[i] keeps counting past cap. *margin receives min |d^2 - r_i^2|
 * over all pairs (the distance of the closest pair to the acceptance boundary). */
int or_neighbor_rows(int n, int D, const double *x, const double *y, const double *z, const double *box,
                     const double *radii, int cap, int *counts, int *rows, double *margin) {
  const double *c[3] = {x, y, z};
  double gmin = INFINITY;
#pragma omp parallel for schedule(dynamic, 16) reduction(min : gmin)
  for (int i = 0; i < n; ++i) {
    const double r2 = radii[i] * radii[i];
    int found = 0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      double s = 0.0;
      for (int d = 0; d < D; ++d) {
        const double dd = c[d][j] - c[d][i];
        const double m = dd - box[d] * jl_round(dd / box[d]);
        s += m * m;
      }
      const double g = fabs(s - r2);
      if (g < gmin) gmin = g;
      if (s <= r2) {
        if (found < cap) rows[(size_t)i * cap + found] = j;
        ++found;
      }
    }
    counts[i] = found;
  }
  if (margin) *margin = gmin;
  return 0;
}

/* ---------------------------------------------------------------- direct pair sums */
/* kind: 1 LJ (plain truncation, unshifted energy), 2 WCA (rc^2 = 1.2599211 sigma^2, u + eps),
 * 3 soft repulsion (rc = sigma). Per sample 12 values: F(3), half energy, 6 half virials
 * (xx yy zz xy xz yz), sum |f| r, count. */
int or_pair_direct(int n, const double *x, const double *y, const double *z, const double *box, int kind,
                   double sigma, double epsilon, double cutoff, int nsample, const int *ids, double *out) {
  double rc2;
  if (kind == 2) rc2 = 1.2599211 * sigma * sigma;
  else if (kind == 3) rc2 = sigma * sigma;
  else rc2 = cutoff * cutoff;
#pragma omp parallel for schedule(dynamic, 4)
  for (int s = 0; s < nsample; ++s) {
    const int i = ids[s];
    double v[12] = {0};
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      double dx = x[i] - x[j]; dx -= box[0] * jl_round(dx / box[0]);
      double dy = y[i] - y[j]; dy -= box[1] * jl_round(dy / box[1]);
      double dz = z[i] - z[j]; dz -= box[2] * jl_round(dz / box[2]);
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (!(0.0 < r2 && r2 < rc2)) continue;
      double a, u;
      if (kind == 3) {
        const double rr = sqrt(r2), xx = 1.0 - rr / sigma;
        a = (epsilon / sigma) * xx / rr;
        u = 0.5 * epsilon * xx * xx;
      } else {
        const double s2 = sigma * sigma / r2;
        const double s6 = s2 * s2 * s2;
        a = 24.0 * epsilon * s6 * (2.0 * s6 - 1.0) / r2;
        u = 4.0 * epsilon * s6 * (s6 - 1.0) + (kind == 2 ? epsilon : 0.0);
      }
      v[0] += a * dx; v[1] += a * dy; v[2] += a * dz;
      v[3] += 0.5 * u;
      v[4] += 0.5 * a * dx * dx; v[5] += 0.5 * a * dy * dy; v[6] += 0.5 * a * dz * dz;
      v[7] += 0.5 * a * dx * dy; v[8] += 0.5 * a * dx * dz; v[9] += 0.5 * a * dy * dz;
      v[10] += fabs(a) * sqrt(r2);
      v[11] += 1.0;
    }
    memcpy(out + (size_t)12 * s, v, sizeof(v));
  }
  return 0;
}

/* ---------------------------------------------------------------- erfc */
/* erfc(x): series of erf for x < 2, backward continued fraction beyond
 * (_erfc_ref, test/ewald_reference.jl:20-40). */
/* Typed pair models (test/pair_reference_gpu.jl:8-64): style 1 uniform (sigma, epsilon, cutoff),
 * style 2 mixed sigma (sigma_ij = (sigma_i + sigma_j)/2, rc = factor sigma_ij, one epsilon), style 3
 * pair matrix (sigma, epsilon, rc by 1-based type id; ntypes x ntypes, column-major). kind 1 LJ
 * (unshifted), 2 WCA (uniform rc^2 = 1.2599211 sigma^2; styles 2 and 3 take their rc; u + eps),
 * 3 soft. D = 2 ignores z. Per sample 10 values: F(3), half energy, 6 half virials. */
int or_pair_typed(int n, int D, const double *x, const double *y, const double *z, const double *box, int kind,
                  int style, double sigma, double epsilon, double cutoff, const double *sigma_p, double factor,
                  const int *typeid, int ntypes, const double *sig_m, const double *eps_m, const double *rc_m,
                  int nsample, const int *ids, double *out) {
  if (style == 3)
    for (int i = 0; i < n; ++i)
      if (typeid[i] < 1 || typeid[i] > ntypes) return -1;
#pragma omp parallel for schedule(dynamic, 4)
  for (int s = 0; s < nsample; ++s) {
    const int i = ids[s];
    double v[10] = {0};
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      double dx = x[i] - x[j]; dx -= box[0] * jl_round(dx / box[0]);
      double dy = y[i] - y[j]; dy -= box[1] * jl_round(dy / box[1]);
      double dz = 0.0;
      if (D == 3) { dz = z[i] - z[j]; dz -= box[2] * jl_round(dz / box[2]); }
      const double r2 = dx * dx + dy * dy + dz * dz;
      double sg, ep, rc2;
      if (style == 1) {
        sg = sigma; ep = epsilon;
        rc2 = kind == 2 ? 1.2599211 * sigma * sigma : (kind == 3 ? sigma * sigma : cutoff * cutoff);
      } else if (style == 2) {
        sg = 0.5 * (sigma_p[i] + sigma_p[j]); ep = epsilon;
        const double rc = factor * sg;
        rc2 = rc * rc;
      } else {
        const int a = typeid[i] - 1, b = typeid[j] - 1;
        sg = sig_m[a + (size_t)ntypes * b]; ep = eps_m[a + (size_t)ntypes * b];
        const double rc = rc_m[a + (size_t)ntypes * b];
        rc2 = rc * rc;
      }
      if (!(0.0 < r2 && r2 < rc2)) continue;
      double a, u;
      if (kind == 3) {
        const double rr = sqrt(r2), xx = 1.0 - rr / sg;
        a = (ep / sg) * xx / rr;
        u = 0.5 * ep * xx * xx;
      } else {
        const double s6 = pow(sg * sg / r2, 3.0);
        a = 24.0 * ep * s6 * (2.0 * s6 - 1.0) / r2;
        u = 4.0 * ep * s6 * (s6 - 1.0) + (kind == 2 ? ep : 0.0);
      }
      v[0] += a * dx; v[1] += a * dy; v[2] += a * dz;
      v[3] += 0.5 * u;
      v[4] += 0.5 * a * dx * dx; v[5] += 0.5 * a * dy * dy; v[6] += 0.5 * a * dz * dz;
      v[7] += 0.5 * a * dx * dy; v[8] += 0.5 * a * dx * dz; v[9] += 0.5 * a * dy * dz;
    }
    for (int k = 0; k < 10; ++k) out[10 * (size_t)s + k] = v[k];
  }
  return 0;
}

double or_erfc(double x) {
  if (x < 0) return 2.0 - or_erfc(-x);
  if (x < 2.0) {
    double term = x, acc = x;
    int n = 0;
    while (term > 1e-18 * acc) {
      n += 1;
      term *= 2.0 * x * x / (2 * n + 1);
      acc += term;
    }
    return 1.0 - 2.0 / sqrt(M_PI) * exp(-x * x) * acc;
  }
  double t = x;
  for (int n = 400; n >= 1; --n) t = x + (n / 2.0) / t;
  return exp(-x * x) / sqrt(M_PI) / t;
}

/* ---------------------------------------------------------------- Ewald */
static void mat3_inv(const double *H, double *Hi, double *det) {
  /* column-major 3x3: H[r + 3c] */
#define A(r, c) H[(r) + 3 * (c)]
  const double d = A(0, 0) * (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) -
                   A(0, 1) * (A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) +
                   A(0, 2) * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
  *det = d;
  Hi[0 + 3 * 0] = (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) / d;
  Hi[0 + 3 * 1] = (A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2)) / d;
  Hi[0 + 3 * 2] = (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1)) / d;
  Hi[1 + 3 * 0] = (A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2)) / d;
  Hi[1 + 3 * 1] = (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) / d;
  Hi[1 + 3 * 2] = (A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2)) / d;
  Hi[2 + 3 * 0] = (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0)) / d;
  Hi[2 + 3 * 1] = (A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1)) / d;
  Hi[2 + 3 * 2] = (A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0)) / d;
#undef A
}

static void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* Periodic Coulomb C sum' q_i q_j / |r_ij + H n| with tin-foil boundaries, optional uniform
 * background, minus (1 - s) C q_i q_j / r for each exception pair at its minimum image.
 * R: 3 x n column-major (R[3i + d]); H: cell matrix, column-major, columns = box vectors.
 * alpha <= 0 selects 3.6 / (smallest cell width). Outputs: E, F (3 x n), W (3 x 3 column-major)
 * with W_ab = sum r_a F_b = -dU/d eps_ab. */
int or_ewald(int n, const double *R, const double *q, const double *H, double C, double alpha, double digits,
             int background, int nexc, const int *exc_i, const int *exc_j, const double *exc_s, double *E_out,
             double *F, double *W_out) {
  double Hi[9], det;
  mat3_inv(H, Hi, &det);
  const double V = fabs(det);
  double widths[3];
  for (int d = 0; d < 3; ++d) {
    double c[3];
    cross3(&H[3 * ((d + 1) % 3)], &H[3 * ((d + 2) % 3)], c);
    widths[d] = V / sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  }
  double wmin = widths[0] < widths[1] ? widths[0] : widths[1];
  if (widths[2] < wmin) wmin = widths[2];
  const double a = alpha > 0 ? alpha : 3.6 / wmin;
  const double xx = sqrt(digits * log(10.0));
  const double rcut = xx / a, kmax = 2 * a * xx;
  int nimg[3];
  for (int d = 0; d < 3; ++d) nimg[d] = (int)ceil(rcut / widths[d]) + 1;
  memset(F, 0, sizeof(double) * 3 * n);
  double E = 0.0, W[9] = {0};
  const double tsp = 2 * a / sqrt(M_PI);
  /* real space */
#pragma omp parallel
  {
    double El = 0.0, Wl[9] = {0};
#pragma omp for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < n; ++j) {
        const double qq = C * q[i] * q[j];
        if (qq == 0.0) continue;
        for (int n1 = -nimg[0]; n1 <= nimg[0]; ++n1)
          for (int n2 = -nimg[1]; n2 <= nimg[1]; ++n2)
            for (int n3 = -nimg[2]; n3 <= nimg[2]; ++n3) {
              if (i == j && n1 == 0 && n2 == 0 && n3 == 0) continue;
              double d[3];
              for (int k = 0; k < 3; ++k)
                d[k] = R[3 * i + k] - R[3 * j + k] + H[k + 0] * n1 + H[k + 3] * n2 + H[k + 6] * n3;
              const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
              if (!(r < rcut)) continue;
              const double ec = or_erfc(a * r);
              El += 0.5 * qq * ec / r;
              const double fmag = qq * (ec / r + tsp * exp(-a * a * r * r)) / (r * r);
              for (int k = 0; k < 3; ++k) F[3 * i + k] += fmag * d[k];
              for (int c2 = 0; c2 < 3; ++c2)
                for (int r2 = 0; r2 < 3; ++r2) Wl[r2 + 3 * c2] += 0.5 * fmag * d[r2] * d[c2];
            }
      }
    }
#pragma omp critical
    {
      E += El;
      for (int k = 0; k < 9; ++k) W[k] += Wl[k];
    }
  }
  /* reciprocal space: B = 2 pi Hinv^T, columns are reciprocal vectors */
  double B[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) B[r + 3 * c] = 2 * M_PI * Hi[c + 3 * r];
  int mmax[3];
  for (int d = 0; d < 3; ++d) {
    const double nb = sqrt(B[0 + 3 * d] * B[0 + 3 * d] + B[1 + 3 * d] * B[1 + 3 * d] + B[2 + 3 * d] * B[2 + 3 * d]);
    mmax[d] = (int)ceil(kmax / nb * 1.0) + 1;
  }
  const int n1r = 2 * mmax[0] + 1, n2r = 2 * mmax[1] + 1, n3r = 2 * mmax[2] + 1;
  const long total = (long)n1r * n2r * n3r;
#pragma omp parallel
  {
    double El = 0.0, Wl[9] = {0};
    double *Fl = (double *)calloc((size_t)3 * n, sizeof(double));
#pragma omp for schedule(dynamic, 8)
    for (long t = 0; t < total; ++t) {
      const int m1 = (int)(t / ((long)n2r * n3r)) - mmax[0];
      const int m2 = (int)((t / n3r) % n2r) - mmax[1];
      const int m3 = (int)(t % n3r) - mmax[2];
      if (m1 == 0 && m2 == 0 && m3 == 0) continue;
      double k[3];
      for (int r = 0; r < 3; ++r) k[r] = B[r] * m1 + B[r + 3] * m2 + B[r + 6] * m3;
      const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
      if (!(k2 < kmax * kmax)) continue;
      const double g = 4 * M_PI * exp(-k2 / (4 * a * a)) / k2;
      double Sre = 0, Sim = 0;
      for (int i = 0; i < n; ++i) {
        const double ph = k[0] * R[3 * i] + k[1] * R[3 * i + 1] + k[2] * R[3 * i + 2];
        Sre += q[i] * cos(ph);
        Sim += q[i] * sin(ph);
      }
      const double Ek = C / (2 * V) * g * (Sre * Sre + Sim * Sim);
      El += Ek;
      for (int i = 0; i < n; ++i) {
        const double ph = k[0] * R[3 * i] + k[1] * R[3 * i + 1] + k[2] * R[3 * i + 2];
        /* imag(cis(ph) * conj(S)) */
        const double im = sin(ph) * Sre - cos(ph) * Sim;
        const double f = (C / V) * g * q[i] * im;
        for (int r = 0; r < 3; ++r) Fl[3 * i + r] += f * k[r];
      }
      const double fac = 2.0 / k2 * (1.0 + k2 / (4 * a * a));
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) Wl[r + 3 * c] += Ek * ((r == c ? 1.0 : 0.0) - fac * k[r] * k[c]);
    }
#pragma omp critical
    {
      E += El;
      for (int kk = 0; kk < 9; ++kk) W[kk] += Wl[kk];
      for (long i = 0; i < 3L * n; ++i) F[i] += Fl[i];
    }
    free(Fl);
  }
  /* self and background */
  double qsq = 0, Q = 0;
  for (int i = 0; i < n; ++i) { qsq += q[i] * q[i]; Q += q[i]; }
  E -= C * a / sqrt(M_PI) * qsq;
  if (background) {
    const double Ebg = -C * M_PI * Q * Q / (2 * a * a * V);
    E += Ebg;
    W[0] += Ebg; W[4] += Ebg; W[8] += Ebg;
  }
  /* exceptions at the minimum image */
  for (int e = 0; e < nexc; ++e) {
    const int i = exc_i[e], j = exc_j[e];
    const double s = exc_s ? exc_s[e] : 0.0;
    double d[3], f[3];
    for (int k = 0; k < 3; ++k) d[k] = R[3 * i + k] - R[3 * j + k];
    for (int k = 0; k < 3; ++k) f[k] = jl_round(Hi[k] * d[0] + Hi[k + 3] * d[1] + Hi[k + 6] * d[2]);
    for (int k = 0; k < 3; ++k) d[k] -= H[k] * f[0] + H[k + 3] * f[1] + H[k + 6] * f[2];
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double qq = (s - 1.0) * C * q[i] * q[j];
    E += qq / r;
    for (int k = 0; k < 3; ++k) {
      const double fk = qq / (r * r * r) * d[k];
      F[3 * i + k] += fk;
      F[3 * j + k] -= fk;
    }
    for (int c = 0; c < 3; ++c)
      for (int r2 = 0; r2 < 3; ++r2) W[r2 + 3 * c] += qq / (r * r * r) * d[r2] * d[c];
  }
  *E_out = E;
  memcpy(W_out, W, sizeof(W));
  return 0;
}

/* Forces on the `ns` particles `ids` of an orthorhombic box, FP64, same sums as or_ewald:
 * real space over every j and the periodic images within rcut (minimum image first), the
 * reciprocal sum over the half space of wave vectors (doubled). F: 3 x ns. */
int or_ewald_sampled(int n, const double *x, const double *y, const double *z, const double *q,
                     const double *box, double C, double alpha, double digits, int ns, const int *ids, double *F) {
  const double xx = sqrt(digits * log(10.0));
  const double rcut = xx / alpha, kmax = 2 * alpha * xx;
  int nimg[3], mmax[3];
  for (int d = 0; d < 3; ++d) {
    nimg[d] = (int)ceil(rcut / box[d] + 0.5) - 1;
    if (nimg[d] < 0) nimg[d] = 0;
    mmax[d] = (int)ceil(kmax * box[d] / (2 * M_PI)) + 1;
  }
  const double tsp = 2 * alpha / sqrt(M_PI), V = box[0] * box[1] * box[2];
  memset(F, 0, sizeof(double) * 3 * ns);
#pragma omp parallel for schedule(dynamic, 1)
  for (int s = 0; s < ns; ++s) {
    const int i = ids[s];
    double f[3] = {0, 0, 0};
    for (int j = 0; j < n; ++j) {
      const double qq = C * q[i] * q[j];
      if (qq == 0.0) continue;
      double d0[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
      for (int k = 0; k < 3; ++k) d0[k] -= box[k] * jl_round(d0[k] / box[k]);
      for (int n1 = -nimg[0]; n1 <= nimg[0]; ++n1)
        for (int n2 = -nimg[1]; n2 <= nimg[1]; ++n2)
          for (int n3 = -nimg[2]; n3 <= nimg[2]; ++n3) {
            if (i == j && n1 == 0 && n2 == 0 && n3 == 0) continue;
            const double d[3] = {d0[0] + n1 * box[0], d0[1] + n2 * box[1], d0[2] + n3 * box[2]};
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (!(r2 < rcut * rcut)) continue;
            const double r = sqrt(r2);
            const double fmag = qq * (or_erfc(alpha * r) / r + tsp * exp(-alpha * alpha * r2)) / r2;
            for (int k = 0; k < 3; ++k) f[k] += fmag * d[k];
          }
    }
    for (int k = 0; k < 3; ++k) F[3 * s + k] = f[k];
  }
  /* reciprocal: half space (m1 > 0, or m1 == 0 and (m2 > 0 or m2 == 0 and m3 > 0)), x2 */
  const int n2r = 2 * mmax[1] + 1, n3r = 2 * mmax[2] + 1;
  const long total = (long)(mmax[0] + 1) * n2r * n3r;
#pragma omp parallel
  {
    double *Fl = (double *)calloc((size_t)3 * ns, sizeof(double));
#pragma omp for schedule(dynamic, 16)
    for (long t = 0; t < total; ++t) {
      const int m1 = (int)(t / ((long)n2r * n3r));
      const int m2 = (int)((t / n3r) % n2r) - mmax[1];
      const int m3 = (int)(t % n3r) - mmax[2];
      if (m1 == 0 && (m2 < 0 || (m2 == 0 && m3 <= 0))) continue;
      const double k[3] = {2 * M_PI * m1 / box[0], 2 * M_PI * m2 / box[1], 2 * M_PI * m3 / box[2]};
      const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
      if (!(k2 < kmax * kmax)) continue;
      const double g = 4 * M_PI * exp(-k2 / (4 * alpha * alpha)) / k2;
      double Sre = 0, Sim = 0;
      for (int j = 0; j < n; ++j) {
        const double ph = k[0] * x[j] + k[1] * y[j] + k[2] * z[j];
        Sre += q[j] * cos(ph);
        Sim += q[j] * sin(ph);
      }
      for (int s = 0; s < ns; ++s) {
        const int i = ids[s];
        const double ph = k[0] * x[i] + k[1] * y[i] + k[2] * z[i];
        const double im = sin(ph) * Sre - cos(ph) * Sim;
        const double fm = 2.0 * (C / V) * g * q[i] * im;
        for (int r = 0; r < 3; ++r) Fl[3 * s + r] += fm * k[r];
      }
    }
#pragma omp critical
    for (long a = 0; a < 3L * ns; ++a) F[a] += Fl[a];
    free(Fl);
  }
  return 0;
}

/* ---------------------------------------------------------------- B-spline */
/* Cardinal B-spline weights of order P at fractional coordinate u: first node k0 =
 * floor(u) - P + 1 (returned), w[j] = M_P(u - (k0 + j)), dw[j] = dw/du, ascending nodes
 * (window_stencil / _bspline_step, src/electrostatics/Windows.jl:188-223). */
int or_bspline(int P, double u, double *w, double *dw) {
  const double fl = floor(u), t = u - fl;
  double a[16] = {0}, b[16];
  a[0] = t;
  a[1] = 1.0 - t;
  int n = 2;
  while (n < P - 1) {
    for (int j = 0; j <= n; ++j) {
      const double left = j < n ? a[j] : 0.0, right = j >= 1 ? a[j - 1] : 0.0;
      b[j] = ((t + j) * left + (n + 1 - t - j) * right) / n;
    }
    memcpy(a, b, sizeof(double) * (n + 1));
    ++n;
  }
  double full[16];
  for (int j = 0; j < P; ++j) {
    const double left = j < P - 1 ? a[j] : 0.0, right = j >= 1 ? a[j - 1] : 0.0;
    full[j] = ((t + j) * left + (P - t - j) * right) / (P - 1);
  }
  for (int i = 0; i < P; ++i) {
    const int j = P - 1 - i;
    w[i] = full[j];
    dw[i] = (j <= P - 2 ? a[j] : 0.0) - (j >= 1 ? a[j - 1] : 0.0);
  }
  return (int)fl - (P - 1);
}

/* ---------------------------------------------------------------- same-algorithm CPU evaluation */
/* Cell list of width >= the largest cutoff (27 cells, full rows, FP64) with the production pair
 * conventions: LJ (kind 1, energy shifted to zero at rc_lj when shift != 0) or WCA (kind 2,
 * rc^2 = 1.2599211 sigma^2, u + eps) (src/engine/PairModels.jl:13-54, src/engine/PairMath.jl:16-39),
 * plus the real-space Ewald term C q_i q_j erfc(alpha r)/r inside rc_coul when C != 0
 * (src/electrostatics/RealSpace.jl:36-118). Per particle: force F (3n, xyz), half energy E (n) and
 * half virial V (6n: xx yy zz xy xz yz per particle). With or_mesh_spread / or_mesh_gather and an
 * FFT this is the FP64 CPU restatement of the device SPME evaluation, used to compare the device
 * at equal parameters and to time a same-algorithm CPU baseline. Returns the number of pairs
 * within either cutoff, each counted once. */
long or_cell_pairs(int n, const double *x, const double *y, const double *z, const double *q, const double *box,
                   int kind, double sigma, double epsilon, double rc_lj, int shift, double C, double alpha,
                   double rc_coul, double *F, double *E, double *V) {
  const double rlj = kind == 2 ? sqrt(1.2599211) * sigma : (kind == 1 ? rc_lj : 0.0);
  const double rlj2 = kind == 2 ? 1.2599211 * sigma * sigma : rlj * rlj;
  const double rcc = C != 0.0 ? rc_coul : 0.0, rcc2 = rcc * rcc;
  const double rmax = rlj > rcc ? rlj : rcc;
  double ushift = 0.0;
  if (kind == 1 && shift) {
    const double s6 = pow(sigma / rc_lj, 6.0);
    ushift = 4.0 * epsilon * s6 * (s6 - 1.0);
  }
  int nc[3];
  for (int d = 0; d < 3; ++d) { nc[d] = (int)floor(box[d] / rmax); if (nc[d] < 3) nc[d] = 3; }
  const int ncell = nc[0] * nc[1] * nc[2];
  int *head = (int *)malloc(sizeof(int) * ncell), *next = (int *)malloc(sizeof(int) * n);
  int *cell = (int *)malloc(sizeof(int) * n);
  for (int c = 0; c < ncell; ++c) head[c] = -1;
  for (int i = 0; i < n; ++i) {
    const double p[3] = {x[i], y[i], z[i]};
    int ci[3];
    for (int d = 0; d < 3; ++d) {
      double s = p[d] / box[d] + 0.5;
      s -= floor(s);
      ci[d] = (int)(s * nc[d]);
      if (ci[d] >= nc[d]) ci[d] = nc[d] - 1;
    }
    cell[i] = (ci[2] * nc[1] + ci[1]) * nc[0] + ci[0];
  }
  for (int i = n - 1; i >= 0; --i) { next[i] = head[cell[i]]; head[cell[i]] = i; }
  const double tsp = 2.0 / sqrt(M_PI);
  long pairs = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : pairs)
  for (int i = 0; i < n; ++i) {
    const int c = cell[i];
    const int cx = c % nc[0], cy = (c / nc[0]) % nc[1], cz = c / (nc[0] * nc[1]);
    double ax = 0, ay = 0, az = 0, e = 0, v[6] = {0, 0, 0, 0, 0, 0};
    for (int oz = -1; oz <= 1; ++oz)
      for (int oy = -1; oy <= 1; ++oy)
        for (int ox = -1; ox <= 1; ++ox) {
          const int c2 = (((cz + oz + nc[2]) % nc[2]) * nc[1] + (cy + oy + nc[1]) % nc[1]) * nc[0] +
                         (cx + ox + nc[0]) % nc[0];
          for (int j = head[c2]; j >= 0; j = next[j]) {
            if (j == i) continue;
            double dx = x[i] - x[j]; dx -= box[0] * jl_round(dx / box[0]);
            double dy = y[i] - y[j]; dy -= box[1] * jl_round(dy / box[1]);
            double dz = z[i] - z[j]; dz -= box[2] * jl_round(dz / box[2]);
            const double r2 = dx * dx + dy * dy + dz * dz;
            double a = 0.0, u = 0.0;
            int hit = 0;
            if (kind != 0 && r2 < rlj2) {
              const double s2 = sigma * sigma / r2, s6 = s2 * s2 * s2;
              a += 24.0 * epsilon * s6 * (2.0 * s6 - 1.0) / r2;
              u += 4.0 * epsilon * s6 * (s6 - 1.0) + (kind == 2 ? epsilon : -ushift);
              hit = 1;
            }
            if (C != 0.0 && r2 < rcc2) {
              const double r = sqrt(r2), qq = C * q[i] * q[j];
              const double ec = erfc(alpha * r) / r;
              a += qq * (ec + tsp * alpha * exp(-alpha * alpha * r2)) / r2;
              u += qq * ec;
              hit = 1;
            }
            if (!hit) continue;
            ax += a * dx; ay += a * dy; az += a * dz;
            e += 0.5 * u;
            v[0] += 0.5 * a * dx * dx; v[1] += 0.5 * a * dy * dy; v[2] += 0.5 * a * dz * dz;
            v[3] += 0.5 * a * dx * dy; v[4] += 0.5 * a * dx * dz; v[5] += 0.5 * a * dy * dz;
            if (j > i) ++pairs;
          }
        }
    F[3 * i] = ax; F[3 * i + 1] = ay; F[3 * i + 2] = az;
    E[i] = e;
    for (int k = 0; k < 6; ++k) V[6 * (size_t)i + k] = v[k];
  }
  free(head); free(next); free(cell);
  return pairs;
}

/* Fractional mesh coordinate u = K (x/L - floor(x/L)) in [0, K) (_grid_coordinate,
 * src/electrostatics/MeshKernels.jl:9-15): node 0 sits at x = 0. */
static double mesh_u(double x, double L, int K) {
  double s = x / L;
  s -= floor(s);
  double u = s * K;
  if (!isfinite(u)) return 0.0;
  return u >= K ? u - K : u;
}
static int mesh_wrap(int n, int K) { return ((n % K) + K) % K; }

/* Charge assignment Q(n) = sum_i q_i M_P(u_i - n) on the grid (K1, K2, K3), x fastest
 * (_spread_lines_kernel!, src/electrostatics/MeshKernels.jl:247-276), FP64. Each OpenMP thread
 * owns a range of z planes and adds only into them, so the sums are deterministic. */
int or_mesh_spread(int n, const double *x, const double *y, const double *z, const double *q, const double *box,
                   const int *K, int P, double *grid) {
  if (P < 2 || P > 12) return -1;
  const int K1 = K[0], K2 = K[1], K3 = K[2];
  const size_t KK = (size_t)K1 * K2;
  memset(grid, 0, sizeof(double) * KK * K3);
#pragma omp parallel
  {
    int nt = 1, t = 0;
#ifdef _OPENMP
    nt = omp_get_num_threads();
    t = omp_get_thread_num();
#endif
    const int z0 = (int)((long)K3 * t / nt), z1 = (int)((long)K3 * (t + 1) / nt);
    double wx[16], wy[16], wz[16], d[16];
    for (int i = 0; i < n; ++i) {
      if (q[i] == 0.0) continue;
      const int k3 = or_bspline(P, mesh_u(z[i], box[2], K3), wz, d);
      int any = 0;
      for (int j3 = 0; j3 < P; ++j3) {
        const int n3 = mesh_wrap(k3 + j3, K3);
        if (n3 >= z0 && n3 < z1) any = 1;
      }
      if (!any) continue;
      const int k1 = or_bspline(P, mesh_u(x[i], box[0], K1), wx, d);
      const int k2 = or_bspline(P, mesh_u(y[i], box[1], K2), wy, d);
      for (int j3 = 0; j3 < P; ++j3) {
        const int n3 = mesh_wrap(k3 + j3, K3);
        if (n3 < z0 || n3 >= z1) continue;
        for (int j2 = 0; j2 < P; ++j2) {
          const int n2 = mesh_wrap(k2 + j2, K2);
          const double qw = q[i] * wz[j3] * wy[j2];
          double *line = grid + KK * n3 + (size_t)K1 * n2;
          for (int j1 = 0; j1 < P; ++j1) line[mesh_wrap(k1 + j1, K1)] += qw * wx[j1];
        }
      }
    }
  }
  return 0;
}

/* Gather with the analytic B-spline derivative (_gather_ad_lines_kernel!,
 * src/electrostatics/MeshKernels.jl:313-333): phi_i = sum_n M(u_i - n) pot(n) and
 * F_i = -q_i (K_d / L_d) sum_n dM/du_d pot(n). F (3n, xyz), phi (n). */
int or_mesh_gather(int n, const double *x, const double *y, const double *z, const double *q, const double *box,
                   const int *K, int P, const double *pot, double *F, double *phi) {
  if (P < 2 || P > 12) return -1;
  const int K1 = K[0], K2 = K[1], K3 = K[2];
  const size_t KK = (size_t)K1 * K2;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double wx[16], wy[16], wz[16], dx[16], dy[16], dz[16];
    const int k1 = or_bspline(P, mesh_u(x[i], box[0], K1), wx, dx);
    const int k2 = or_bspline(P, mesh_u(y[i], box[1], K2), wy, dy);
    const int k3 = or_bspline(P, mesh_u(z[i], box[2], K3), wz, dz);
    double p = 0, gx = 0, gy = 0, gz = 0;
    for (int j3 = 0; j3 < P; ++j3) {
      const int n3 = mesh_wrap(k3 + j3, K3);
      for (int j2 = 0; j2 < P; ++j2) {
        const double *line = pot + KK * n3 + (size_t)K1 * mesh_wrap(k2 + j2, K2);
        double lp = 0, lg = 0;
        for (int j1 = 0; j1 < P; ++j1) {
          const double v = line[mesh_wrap(k1 + j1, K1)];
          lp += wx[j1] * v;
          lg += dx[j1] * v;
        }
        p += wz[j3] * wy[j2] * lp;
        gx += wz[j3] * wy[j2] * lg;
        gy += wz[j3] * dy[j2] * lp;
        gz += dz[j3] * wy[j2] * lp;
      }
    }
    phi[i] = p;
    F[3 * i] = -q[i] * gx * K1 / box[0];
    F[3 * i + 1] = -q[i] * gy * K2 / box[1];
    F[3 * i + 2] = -q[i] * gz * K3 / box[2];
  }
  return 0;
}

/* Sets the OpenMP thread count of later calls; returns the previous one (1 without OpenMP). */
int or_set_threads(int n) {
#ifdef _OPENMP
  const int prev = omp_get_max_threads();
  if (n > 0) omp_set_num_threads(n);
  return prev;
#else
  (void)n;
  return 1;
#endif
}

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---------------------------------------------------------------- Kremer-Grest melt */
/* WCA (rc^2 = 1.2599211 sigma^2, u + eps) between every pair within the cutoff, bonded
 * pairs included (exclude_bonded = false, src/SimulationState.jl:152-155) or, with
 * exclude_bonded, skipped (BondExclusionFeature, src/engine/Bonded.jl:60-92), found with a cell
 * list of width >= rc, plus FENE bonds u = -k R0^2/2 ln(1 - r^2/R0^2), f/r = -k/(1 - r^2/R0^2)
 * (_fene3_EV!, src/bonded/FENEBondCUDA.jl). Bonds are pairs (bi[b], bj[b]), 0-based. Writes
 * F (3n, xyz interleaved), per-particle half energies E (n), total energy, and the virial
 * sum over pairs and bonds of r_a F_b (6: xx yy zz xy xz yz). Returns -1 - (first bond
 * index) when a bond has r >= R0, else the number of WCA pairs (each counted once). */
long or_kg_forces(int n, const double *x, const double *y, const double *z, const double *box, double sigma,
                  double epsilon, int nbond, const int *bi, const int *bj, double k, double R0, double *F,
                  double *E, double *energy, double *W, int exclude_bonded) {
  const double rc = pow(2.0, 1.0 / 6.0) * sigma, rc2 = 1.2599211 * sigma * sigma, R02 = R0 * R0;
  int nc[3];
  for (int d = 0; d < 3; ++d) { nc[d] = (int)floor(box[d] / rc); if (nc[d] < 3) nc[d] = 3; }
  const int ncell = nc[0] * nc[1] * nc[2];
  int *head = (int *)malloc(sizeof(int) * ncell), *next = (int *)malloc(sizeof(int) * n);
  int *cell = (int *)malloc(sizeof(int) * n);
  for (int c = 0; c < ncell; ++c) head[c] = -1;
  for (int i = 0; i < n; ++i) {
    const double p[3] = {x[i], y[i], z[i]};
    int ci[3];
    for (int d = 0; d < 3; ++d) {
      double s = p[d] / box[d] + 0.5;
      s -= floor(s);
      ci[d] = (int)(s * nc[d]);
      if (ci[d] >= nc[d]) ci[d] = nc[d] - 1;
    }
    cell[i] = (ci[2] * nc[1] + ci[1]) * nc[0] + ci[0];
  }
  for (int i = n - 1; i >= 0; --i) { next[i] = head[cell[i]]; head[cell[i]] = i; }
  /* bonded exclusions (exclude_bonded): partners of each particle, CSR */
  int *boff = (int *)calloc((size_t)n + 1, sizeof(int)), *bpart = (int *)malloc(sizeof(int) * (2 * (size_t)nbond + 1));
  for (int b = 0; b < nbond; ++b) { ++boff[bi[b] + 1]; ++boff[bj[b] + 1]; }
  for (int i = 0; i < n; ++i) boff[i + 1] += boff[i];
  {
    int *fill = (int *)calloc((size_t)n, sizeof(int));
    for (int b = 0; b < nbond; ++b) {
      bpart[boff[bi[b]] + fill[bi[b]]++] = bj[b];
      bpart[boff[bj[b]] + fill[bj[b]]++] = bi[b];
    }
    free(fill);
  }
  long pairs = 0;
  double etot = 0.0, w0 = 0, w1 = 0, w2 = 0, w3 = 0, w4 = 0, w5 = 0;
#pragma omp parallel for schedule(static) reduction(+ : pairs, etot, w0, w1, w2, w3, w4, w5)
  for (int i = 0; i < n; ++i) {
    const int c = cell[i];
    const int cx = c % nc[0], cy = (c / nc[0]) % nc[1], cz = c / (nc[0] * nc[1]);
    double ax = 0, ay = 0, az = 0, e = 0;
    for (int oz = -1; oz <= 1; ++oz)
      for (int oy = -1; oy <= 1; ++oy)
        for (int ox = -1; ox <= 1; ++ox) {
          const int c2 = (((cz + oz + nc[2]) % nc[2]) * nc[1] + (cy + oy + nc[1]) % nc[1]) * nc[0] +
                         (cx + ox + nc[0]) % nc[0];
          for (int j = head[c2]; j >= 0; j = next[j]) {
            if (j == i) continue;
            double dx = x[i] - x[j]; dx -= box[0] * jl_round(dx / box[0]);
            double dy = y[i] - y[j]; dy -= box[1] * jl_round(dy / box[1]);
            double dz = z[i] - z[j]; dz -= box[2] * jl_round(dz / box[2]);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 < rc2)) continue;
            if (exclude_bonded) {
              int bonded = 0;
              for (int t = boff[i]; t < boff[i + 1]; ++t) bonded |= bpart[t] == j;
              if (bonded) continue;
            }
            const double s2 = sigma * sigma / r2, s6 = s2 * s2 * s2;
            const double a = 24.0 * epsilon * s6 * (2.0 * s6 - 1.0) / r2;
            ax += a * dx; ay += a * dy; az += a * dz;
            e += 0.5 * (4.0 * epsilon * s6 * (s6 - 1.0) + epsilon);
            w0 += 0.5 * a * dx * dx; w1 += 0.5 * a * dy * dy; w2 += 0.5 * a * dz * dz;
            w3 += 0.5 * a * dx * dy; w4 += 0.5 * a * dx * dz; w5 += 0.5 * a * dy * dz;
            if (j > i) ++pairs;
          }
        }
    F[3 * i] = ax; F[3 * i + 1] = ay; F[3 * i + 2] = az;
    E[i] = e;
    etot += e;
  }
  long bad = -1;
  for (int b = 0; b < nbond; ++b) {
    const int i = bi[b], j = bj[b];
    double dx = x[i] - x[j]; dx -= box[0] * jl_round(dx / box[0]);
    double dy = y[i] - y[j]; dy -= box[1] * jl_round(dy / box[1]);
    double dz = z[i] - z[j]; dz -= box[2] * jl_round(dz / box[2]);
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < R02)) { if (bad < 0) bad = b; continue; }
    const double a = -k / ((R02 - r2) / R02);
    const double u = -0.5 * k * R02 * (r2 < 0.5 * R02 ? log1p(-r2 / R02) : log((R02 - r2) / R02));
    F[3 * i] += a * dx; F[3 * i + 1] += a * dy; F[3 * i + 2] += a * dz;
    F[3 * j] -= a * dx; F[3 * j + 1] -= a * dy; F[3 * j + 2] -= a * dz;
    E[i] += 0.5 * u; E[j] += 0.5 * u;
    etot += u;
    w0 += a * dx * dx; w1 += a * dy * dy; w2 += a * dz * dz;
    w3 += a * dx * dy; w4 += a * dx * dz; w5 += a * dy * dz;
  }
  *energy = etot;
  W[0] = w0; W[1] = w1; W[2] = w2; W[3] = w3; W[4] = w4; W[5] = w5;
  free(head); free(next); free(cell); free(boff); free(bpart);
  return bad >= 0 ? -1 - bad : pairs;
}
