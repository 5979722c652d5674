// Pair forces, per-particle half energies / half virials and FP64 totals over slot-major
// ELL rows: LJ, WCA, harmonic soft repulsion, each optionally with real-space Ewald Coulomb
// through an r^2 polynomial of the long-range correction.
//
// Thread layout: G lanes per particle (G in 1..64, a power of two) stride the particle's
// row (slots lane, lane+G, ...) and combine with xor-shuffles inside the wave64; the row
// slots of G=1 are read fully coalesced, larger G trade that for more parallelism on small
// systems (_pair_cooperative_kernel!, src/engine/PairTraversal.jl:66-125). Neighbor
// coordinates and charges come from one float4 gather (x, y, z, q).
#include "pdx_internal.hpp"

#include <algorithm>
#include <cmath>

using namespace pdx;

namespace {

constexpr int TPB = 256;
constexpr int MAXC = 24;

struct PairArgs {
  int N;
  const float4 *packed;
  const int32_t *nbr;
  const int32_t *counts;
  const int64_t *state;
  const float *geom;
  float *fx, *fy, *fz, *Epot, *V;
  double *totals;
  double *partial;
  unsigned *ticket;
  int add;
  float eps, sig2, rc2, eshift;   // LJ / WCA / soft
  float cC, crc2, csx, ceshift, halfself;
  float pe[MAXC];
  float pf[MAXC];
};

// Horner with a compile-time length: the coefficients are uniform kernel arguments held in
// scalar registers for the whole row loop (a runtime length would re-load them every step).
template <int NP>
__device__ __forceinline__ float horner(const float (&c)[NP], float x) {
  float acc = c[NP - 1];
#pragma unroll
  for (int k = NP - 2; k >= 0; --k) acc = fmaf(acc, x, c[k]);
  return acc;
}

// Per-visit physics shared by both traversal layouts: force factor f (force on i = f d) and
// pair energy u of one row entry (LJ/WCA/soft + real-space Coulomb polynomial).
template <int MODEL, bool COUL, bool ENERGY, int NP>
__device__ __forceinline__ void pair_terms(const PairArgs &a, const float (&pe)[NP], const float (&pf)[NP],
                                           float qi, float qj, float r2, float &f, float &u) {
  f = 0.f;
  u = 0.f;
  if (MODEL == PDX_PAIR_LJ || MODEL == PDX_PAIR_WCA) {
    const bool in = (r2 > 0.f) & (r2 < a.rc2);
    const float invr2 = __builtin_amdgcn_rcpf(r2);
    const float s2 = a.sig2 * invr2;
    const float s6 = s2 * s2 * s2;
    const float s12 = s6 * s6;
    const float fl = 24.0f * a.eps * (2.0f * s12 - s6) * invr2;
    f = in ? fl : 0.f;
    if (ENERGY) {
      const float ul = 4.0f * a.eps * (s12 - s6) + (MODEL == PDX_PAIR_WCA ? a.eps : -a.eshift);
      u = in ? ul : 0.f;
    }
  } else if (MODEL == PDX_PAIR_SOFT) {
    const bool in = (r2 > 0.f) & (r2 < a.rc2);
    const float r = sqrtf(r2);
    const float sg = sqrtf(a.sig2);
    const float xx = 1.0f - r / sg;
    f = in ? (a.eps / sg) * xx / r : 0.f;
    if (ENERGY) u = in ? 0.5f * a.eps * xx * xx : 0.f;
  }
  if (COUL) {
    const bool inr = (r2 > 0.f) & (r2 < a.crc2);
    const float x = fmaf(r2, a.csx, -1.0f);
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float rinv3 = rinv * rinv * rinv;
    const float cq = a.cC * qi * qj;
    const float fc = cq * (rinv3 - horner<NP>(pf, x));
    f += inr ? fc : 0.f;
    if (ENERGY) {
      const float uc = cq * (rinv + horner<NP>(pe, x) - a.ceshift);
      u += inr ? uc : 0.f;
    }
  }
}

// Writes of one particle's sums (lane that owns it) and the FP64 totals of the launch.
template <bool COUL, bool ENERGY>
__device__ __forceinline__ void pair_epilogue(const PairArgs &a, bool owner, int i, float qi, float ax, float ay,
                                              float az, float e, const float (&v)[6]) {
  float ei = 0.f, w[6] = {0, 0, 0, 0, 0, 0};
  if (owner) {
    if (a.add) {
      a.fx[i] += ax; a.fy[i] += ay; a.fz[i] += az;
    } else {
      a.fx[i] = ax; a.fy[i] = ay; a.fz[i] = az;
    }
    if (ENERGY) {
      // half of every pair (each is visited from both ends) plus the Coulomb self energy
      ei = 0.5f * e;
      if (COUL) ei -= a.halfself * qi * qi;
#pragma unroll
      for (int c = 0; c < 6; ++c) w[c] = 0.5f * v[c];
      if (a.Epot) a.Epot[i] = a.add ? a.Epot[i] + ei : ei;
      if (a.V) {
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          float *p = a.V + (size_t)c * a.N + i;
          *p = a.add ? *p + w[c] : w[c];
        }
      }
    }
  }
  if (ENERGY && a.totals) {
    const double d[7] = {(double)ei, (double)w[0], (double)w[1], (double)w[2],
                         (double)w[3], (double)w[4], (double)w[5]};
    grid_sum_finalize<7>(d, a.partial, a.ticket, a.totals, a.add != 0);
  }
}

// Layout A (small systems): G lanes of one wave per particle, xor-shuffle reduction.
template <int MODEL, bool COUL, bool ENERGY, int G, int NP>
__global__ __launch_bounds__(TPB) void k_pair(const PairArgs a) {
  const int64_t *s = a.state;
  const bool valid = s[PDX_DN_VALID] != 0 && s[PDX_DN_ERROR] == 0;  // uniform
  if (!valid) return;
  const int gid = blockIdx.x * TPB + threadIdx.x;
  const int i = gid / G;
  const int lane = gid % G;
  const int N = a.N;
  const bool active = i < N;
  const float Lx = a.geom[0], Ly = a.geom[1], Lz = a.geom[2];
  const float hx = 0.5f * Lx, hy = 0.5f * Ly, hz = 0.5f * Lz;
  float ax = 0.f, ay = 0.f, az = 0.f, e = 0.f;
  float v[6] = {0, 0, 0, 0, 0, 0};
  float qi = 0.f;
  float pe[NP], pf[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    pe[k] = a.pe[k];
    pf[k] = a.pf[k];
  }
  if (active) {
    const float4 pi = a.packed[i];
    qi = pi.w;
    const int n = a.counts[i];
    const int32_t *row = a.nbr + i;
    // U row entries per lane and iteration: U independent gathers in flight. Padding entries
    // point at i itself (r^2 = 0 is outside every model), so no branch is needed.
    constexpr int U = 4;
    for (int t0 = lane; t0 < n; t0 += U * G) {
      int jj[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int t = t0 + u * G;
        jj[u] = t < n ? row[(size_t)t * N] - 1 : i;
      }
      float4 pj[U];
#pragma unroll
      for (int u = 0; u < U; ++u) pj[u] = a.packed[jj[u]];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float dx = periodic_delta(pi.x, pj[u].x, hx, Lx);
        const float dy = periodic_delta(pi.y, pj[u].y, hy, Ly);
        const float dz = periodic_delta(pi.z, pj[u].z, hz, Lz);
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        float f, uu;
        pair_terms<MODEL, COUL, ENERGY, NP>(a, pe, pf, qi, pj[u].w, r2, f, uu);
        ax = fmaf(f, dx, ax);
        ay = fmaf(f, dy, ay);
        az = fmaf(f, dz, az);
        if (ENERGY) {
          e += uu;
          const float fxij = f * dx, fyij = f * dy, fzij = f * dz;
          v[0] = fmaf(dx, fxij, v[0]); v[1] = fmaf(dy, fyij, v[1]); v[2] = fmaf(dz, fzij, v[2]);
          v[3] = fmaf(dx, fyij, v[3]); v[4] = fmaf(dx, fzij, v[4]); v[5] = fmaf(dy, fzij, v[5]);
        }
      }
    }
  }
  if (G > 1) {
    ax = wave_sum_width(ax, G);
    ay = wave_sum_width(ay, G);
    az = wave_sum_width(az, G);
    if (ENERGY) {
      e = wave_sum_width(e, G);
#pragma unroll
      for (int c = 0; c < 6; ++c) v[c] = wave_sum_width(v[c], G);
    }
  }
  pair_epilogue<COUL, ENERGY>(a, active && lane == 0, i, qi, ax, ay, az, e, v);
}

// Layout B (large systems): a block owns 64 consecutive particles and S waves; wave w walks
// slots w, w+S, ... of all 64 rows, so every row load is one contiguous 256-byte access of
// the slot-major table, and the S partial sums meet in LDS.
template <int MODEL, bool COUL, bool ENERGY, int S, int NP>
__global__ __launch_bounds__(64 * S) void k_pair_split(const PairArgs a) {
  const int64_t *st = a.state;
  const bool valid = st[PDX_DN_VALID] != 0 && st[PDX_DN_ERROR] == 0;  // uniform
  if (!valid) return;
  constexpr int NV = ENERGY ? 10 : 3;
  __shared__ float red[S][NV][WAVE];
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = threadIdx.x / WAVE;
  const int i = blockIdx.x * WAVE + lane;
  const int N = a.N;
  const bool active = i < N;
  const float Lx = a.geom[0], Ly = a.geom[1], Lz = a.geom[2];
  const float hx = 0.5f * Lx, hy = 0.5f * Ly, hz = 0.5f * Lz;
  float acc[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) acc[c] = 0.f;
  float qi = 0.f;
  float pe[NP], pf[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    pe[k] = a.pe[k];
    pf[k] = a.pf[k];
  }
  if (active) {
    const float4 pi = a.packed[i];
    qi = pi.w;
    const int n = a.counts[i];
    const int32_t *row = a.nbr + i;
    int jn = w < n ? row[(size_t)w * N] - 1 : 0;
    for (int t = w; t < n; t += S) {
      const int j = jn;
      if (t + S < n) jn = row[(size_t)(t + S) * N] - 1;
      const float4 pj = a.packed[j];
      const float dx = periodic_delta(pi.x, pj.x, hx, Lx);
      const float dy = periodic_delta(pi.y, pj.y, hy, Ly);
      const float dz = periodic_delta(pi.z, pj.z, hz, Lz);
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      float f, u;
      pair_terms<MODEL, COUL, ENERGY, NP>(a, pe, pf, qi, pj.w, r2, f, u);
      acc[0] = fmaf(f, dx, acc[0]);
      acc[1] = fmaf(f, dy, acc[1]);
      acc[2] = fmaf(f, dz, acc[2]);
      if (ENERGY) {
        acc[3] += u;
        const float fxij = f * dx, fyij = f * dy, fzij = f * dz;
        acc[4] = fmaf(dx, fxij, acc[4]); acc[5] = fmaf(dy, fyij, acc[5]); acc[6] = fmaf(dz, fzij, acc[6]);
        acc[7] = fmaf(dx, fyij, acc[7]); acc[8] = fmaf(dx, fzij, acc[8]); acc[9] = fmaf(dy, fzij, acc[9]);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < NV; ++c) red[w][c][lane] = acc[c];
  __syncthreads();
  float sum[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < S; ++k) t += red[k][c][lane];  // fixed order: deterministic
    sum[c] = t;
  }
  float v[6] = {0, 0, 0, 0, 0, 0};
  if (ENERGY) {
#pragma unroll
    for (int c = 0; c < 6; ++c) v[c] = sum[4 + c];
  }
  pair_epilogue<COUL, ENERGY>(a, active && w == 0, i, qi, sum[0], sum[1], sum[2], ENERGY ? sum[3] : 0.f, v);
}

template <int MODEL, bool COUL, bool ENERGY, int NP>
void launch_g(int G, dim3 grid, hipStream_t st, const PairArgs &a) {
  if (G < 0) {  // split layout, S = -G waves per 64 particles
    const dim3 sg((a.N + WAVE - 1) / WAVE);
    switch (-G) {
      case 2: hipLaunchKernelGGL((k_pair_split<MODEL, COUL, ENERGY, 2, NP>), sg, dim3(128), 0, st, a); break;
      case 4: hipLaunchKernelGGL((k_pair_split<MODEL, COUL, ENERGY, 4, NP>), sg, dim3(256), 0, st, a); break;
      case 8: hipLaunchKernelGGL((k_pair_split<MODEL, COUL, ENERGY, 8, NP>), sg, dim3(512), 0, st, a); break;
      default: hipLaunchKernelGGL((k_pair_split<MODEL, COUL, ENERGY, 16, NP>), sg, dim3(1024), 0, st, a); break;
    }
    return;
  }
  switch (G) {
    case 1: hipLaunchKernelGGL((k_pair<MODEL, COUL, ENERGY, 1, NP>), grid, dim3(TPB), 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_pair<MODEL, COUL, ENERGY, 2, NP>), grid, dim3(TPB), 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_pair<MODEL, COUL, ENERGY, 4, NP>), grid, dim3(TPB), 0, st, a); break;
    case 8: hipLaunchKernelGGL((k_pair<MODEL, COUL, ENERGY, 8, NP>), grid, dim3(TPB), 0, st, a); break;
    case 16: hipLaunchKernelGGL((k_pair<MODEL, COUL, ENERGY, 16, NP>), grid, dim3(TPB), 0, st, a); break;
    case 32: hipLaunchKernelGGL((k_pair<MODEL, COUL, ENERGY, 32, NP>), grid, dim3(TPB), 0, st, a); break;
    default: hipLaunchKernelGGL((k_pair<MODEL, COUL, ENERGY, 64, NP>), grid, dim3(TPB), 0, st, a); break;
  }
}

// Coulomb polynomials are zero-padded to 16 or 24 monomials (Horner over exact zeros).
template <int MODEL>
void launch_model(bool coul, bool energy, int np, int G, dim3 grid, hipStream_t st, const PairArgs &a) {
  if (coul) {
    if (np <= 16) {
      if (energy) launch_g<MODEL, true, true, 16>(G, grid, st, a);
      else launch_g<MODEL, true, false, 16>(G, grid, st, a);
    } else {
      if (energy) launch_g<MODEL, true, true, 24>(G, grid, st, a);
      else launch_g<MODEL, true, false, 24>(G, grid, st, a);
    }
  } else {
    if (energy) launch_g<MODEL, false, true, 1>(G, grid, st, a);
    else launch_g<MODEL, false, false, 1>(G, grid, st, a);
  }
}

// ---------------------------------------------------------------- host: Ewald correction fit
// corr(s) = erf(alpha r)/r and fcorr(s) = -(1/r) d corr/dr at s = r^2 (long_range_correction,
// src/electrostatics/Splits.jl:142-161, restated in FP64): the alternating Taylor series in
// z = alpha^2 s below z = 1, the closed forms above (no cancellation there).
void ewald_correction(double alpha, double s, double &corr, double &fcorr) {
  const double pre = 2.0 * alpha / std::sqrt(M_PI);
  const double z = alpha * alpha * s;
  if (z < 1.0) {
    // corr = pre sum_n t_n/(2n+1) with t_n = (-z)^n/n!; fcorr = -2 d corr/ds
    //      = 2 pre alpha^2 sum_{n>=1} t_{n-1}/(2n+1)
    double ce = 1.0, cf = 0.0, term = 1.0;
    for (int n = 1; n < 60; ++n) {
      cf += term / (2 * n + 1);
      term = -term * z / n;
      ce += term / (2 * n + 1);
      if (std::fabs(term) < 1e-22) break;
    }
    corr = pre * ce;
    fcorr = 2.0 * pre * alpha * alpha * cf;
    return;
  }
  const double r = std::sqrt(s);
  const double ef = std::erf(alpha * r);
  corr = ef / r;
  fcorr = ef / (s * r) - pre * std::exp(-z) / s;
}

// Chebyshev interpolation on [-1, 1] at n first-kind nodes; truncate where the tail of
// |c_j| drops below tol * max|f|; return monomial coefficients in x.
std::vector<double> chebyshev_monomial(const std::vector<double> &fvals, double tol) {
  const int n = (int)fvals.size();
  std::vector<double> c(n, 0.0);
  double fmax = 0.0;
  for (double v : fvals) fmax = std::max(fmax, std::fabs(v));
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += fvals[k] * std::cos(M_PI * j * (k + 0.5) / n);
    c[j] = (j == 0 ? 1.0 : 2.0) * s / n;
  }
  int deg = n - 1;
  double tail = 0.0;
  for (int j = n - 1; j >= 1; --j) {
    tail += std::fabs(c[j]);
    if (tail > tol * std::max(fmax, 1e-300)) { deg = j; break; }
    deg = j - 1;
  }
  const int m = deg + 1;
  // T_k as monomials, accumulated: T_{k+1} = 2x T_k - T_{k-1}
  std::vector<double> mono(m, 0.0), t0(m, 0.0), t1(m, 0.0), t2(m, 0.0);
  t0[0] = 1.0;
  mono[0] += c[0];
  if (m > 1) {
    t1[1] = 1.0;
    mono[1] += c[1];
  }
  for (int k = 2; k < m; ++k) {
    std::fill(t2.begin(), t2.end(), 0.0);
    for (int q = 0; q + 1 < m; ++q) t2[q + 1] += 2.0 * t1[q];
    for (int q = 0; q < m; ++q) t2[q] -= t0[q];
    for (int q = 0; q < m; ++q) mono[q] += c[k] * t2[q];
    t0 = t1;
    t1 = t2;
  }
  return mono;
}

}  // namespace

// host helpers shared with the mesh translation unit
namespace pdx {
void host_ewald_correction(double alpha, double s, double &corr, double &fcorr) {
  ewald_correction(alpha, s, corr, fcorr);
}
}  // namespace pdx

extern "C" int pdx_pair_create(const pdx_pair_spec *spec, pdx_pair_t *out) {
  if (!spec || !out) return fail(PDX_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  const pdx_pair_spec &sp = *spec;
  if (sp.model < PDX_PAIR_NONE || sp.model > PDX_PAIR_SOFT) return fail(PDX_ERR_ARGUMENT, "unknown pair model");
  if (sp.model != PDX_PAIR_NONE) {
    if (!(std::isfinite(sp.epsilon) && std::isfinite(sp.sigma) && sp.sigma > 0))
      return fail(PDX_ERR_ARGUMENT, "pair epsilon and sigma must be finite, sigma positive");
    if (sp.model == PDX_PAIR_LJ && !(std::isfinite(sp.cutoff) && sp.cutoff > 0))
      return fail(PDX_ERR_ARGUMENT, "LJ cutoff must be finite and positive");
  }
  if (sp.coulomb) {
    if (!(sp.coulomb_C > 0 && std::isfinite(sp.coulomb_C)))
      return fail(PDX_ERR_ARGUMENT, "coulomb_prefactor must be finite and positive");
    if (!(sp.coulomb_cutoff > 0 && std::isfinite(sp.coulomb_cutoff)))
      return fail(PDX_ERR_ARGUMENT, "real_cutoff must be finite and positive");
    if (!(sp.alpha > 0 && std::isfinite(sp.alpha))) return fail(PDX_ERR_ARGUMENT, "Ewald alpha must be finite and positive.");
  }
  if (sp.lanes > 0 && (sp.lanes > 64 || (sp.lanes & (sp.lanes - 1))))
    return fail(PDX_ERR_ARGUMENT, "lanes per particle must be a power of two in 1..64");
  if (sp.lanes < 0 && !(sp.lanes == -2 || sp.lanes == -4 || sp.lanes == -8 || sp.lanes == -16))
    return fail(PDX_ERR_ARGUMENT, "split layouts use 2, 4, 8 or 16 waves (lanes = -2, -4, -8, -16)");
  auto *p = new pdx_pair_s();
  p->spec = sp;
  const float eps = (float)sp.epsilon, sig = (float)sp.sigma;
  p->eps = eps;
  p->sig2 = sig * sig;
  if (sp.model == PDX_PAIR_LJ) {
    const float rc = (float)sp.cutoff;
    p->rc2 = rc * rc;
    if (sp.shift && rc > 0) {  // lj_cutoff_energy (PairModels.jl:20-25), in the list precision
      const float s2 = p->sig2 / (rc * rc), s6 = s2 * s2 * s2;
      p->eshift = 4.0f * eps * (s6 * s6 - s6);
    }
  } else if (sp.model == PDX_PAIR_WCA) {
    p->rc2 = 1.2599211f * p->sig2;  // UniformWCA constant (PairModels.jl:16)
  } else if (sp.model == PDX_PAIR_SOFT) {
    p->rc2 = p->sig2;
  }
  if (sp.coulomb) {
    const double rc = sp.coulomb_cutoff, rc2 = rc * rc, a = sp.alpha;
    p->cC = (float)sp.coulomb_C;
    p->crc2 = (float)rc2;
    p->csx = (float)(2.0 / rc2);
    p->ceshift = sp.coulomb_shift ? (float)(std::erfc(a * rc) / rc) : 0.0f;
    p->halfself = (float)(0.5 * sp.coulomb_C * 2.0 * a / std::sqrt(M_PI));
    const int nodes = 64;
    std::vector<double> fe(nodes), ff(nodes);
    for (int k = 0; k < nodes; ++k) {
      const double x = std::cos(M_PI * (k + 0.5) / nodes);
      double corr, fcorr;
      ewald_correction(a, (x + 1.0) * rc2 / 2.0, corr, fcorr);
      fe[k] = -corr;
      ff[k] = fcorr;
    }
    // FP32 tolerances of coulomb_polynomial (PairFunctor.jl:58-61): 1e-8 energy, 1e-7 force
    std::vector<double> me = chebyshev_monomial(fe, 1e-8), mf = chebyshev_monomial(ff, 1e-7);
    if ((int)me.size() > MAXC || (int)mf.size() > MAXC) {
      delete p;
      return fail(PDX_ERR_ARGUMENT, "Ewald correction polynomial needs more than 24 terms (alpha*cutoff too large)");
    }
    p->npe = (int)me.size();
    p->npf = (int)mf.size();
    for (int k = 0; k < p->npe; ++k) p->pe[k] = (float)me[k];
    for (int k = 0; k < p->npf; ++k) p->pf[k] = (float)mf[k];
    // measured fit error of the FP32 Horner form against FP64, relative to max|f|
    double emax = 0, fmax = 0, eerr = 0, ferr = 0;
    for (int k = 0; k <= 2000; ++k) {
      const double x = -1.0 + 2.0 * k / 2000.0;
      double corr, fcorr;
      ewald_correction(a, (x + 1.0) * rc2 / 2.0, corr, fcorr);
      float he = p->pe[p->npe - 1], hf = p->pf[p->npf - 1];
      for (int q = p->npe - 2; q >= 0; --q) he = fmaf(he, (float)x, p->pe[q]);
      for (int q = p->npf - 2; q >= 0; --q) hf = fmaf(hf, (float)x, p->pf[q]);
      emax = std::max(emax, std::fabs(corr));
      fmax = std::max(fmax, std::fabs(fcorr));
      eerr = std::max(eerr, std::fabs((double)he + corr));
      ferr = std::max(ferr, std::fabs((double)hf - fcorr));
    }
    p->fit_err_e = eerr / emax;
    p->fit_err_f = ferr / fmax;
  }
  p->lanes = sp.lanes;
  int rc;
  if ((rc = p->ticket.alloc(1))) { delete p; return rc; }
  if (hipMemset(p->ticket.p, 0, sizeof(unsigned)) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    delete p;
    return fail(PDX_ERR_DEVICE, "pair model initialisation failed");
  }
  *out = p;
  return PDX_OK;
}

extern "C" int pdx_pair_destroy(pdx_pair_t p) {
  delete p;
  return PDX_OK;
}

extern "C" int pdx_pair_info(pdx_pair_t p, double out[4]) {
  if (!p || !out) return fail(PDX_ERR_ARGUMENT, "null argument");
  out[0] = p->npe; out[1] = p->npf; out[2] = p->fit_err_e; out[3] = p->fit_err_f;
  return PDX_OK;
}

extern "C" int pdx_pair_forces(pdx_pair_t p, pdx_nl_t nl, float *fx, float *fy, float *fz, float *Epot,
                               float *V, double *totals, int32_t flags, pdx_stream_t stream) {
  if (!p || !nl) return fail(PDX_ERR_ARGUMENT, "null pair model or neighbor list");
  if (!fx || !fy || !fz) return fail(PDX_ERR_ARGUMENT, "null force array");
  const bool energy = (flags & PDX_ENERGY) != 0;
  const int N = nl->N;
  // lanes > 0: G lanes of a wave per particle; lanes < 0: the split layout with -lanes waves
  // per block of 64 particles; 0: automatic
  int G = p->lanes;
  if (G == 0) {
    // about 2^17 threads: every CU busy while a wave's row slots stay in few cache lines
    // (measured on the 64 000-ion system: 2 lanes beat 1 and 4, and every split layout)
    const int64_t want = std::max<int64_t>(1, 131072 / std::max(N, 1));
    G = 1;
    while (G * 2 <= want && G < 64) G <<= 1;
  }
  const int blocks = G < 0 ? (N + WAVE - 1) / WAVE : (int)(((int64_t)N * G + TPB - 1) / TPB);
  if (energy && totals) {
    if ((size_t)blocks * 7 > p->partial.n) {
      PDX_HIP(hipStreamSynchronize(as_stream(stream)));
      int rc = p->partial.alloc((size_t)blocks * 7);
      if (rc) return rc;
    }
  }
  PairArgs a{};
  a.N = N;
  a.packed = nl->packed.p;
  a.nbr = nl->neighbors_flat.p;
  a.counts = nl->counts.p;
  a.state = nl->state.p;
  a.geom = nl->geom.p;
  a.fx = fx; a.fy = fy; a.fz = fz;
  a.Epot = Epot; a.V = V;
  a.totals = totals;
  a.partial = p->partial.p;
  a.ticket = p->ticket.p;
  a.add = (flags & PDX_ADD) ? 1 : 0;
  a.eps = p->eps; a.sig2 = p->sig2; a.rc2 = p->rc2; a.eshift = p->eshift;
  a.cC = p->cC; a.crc2 = p->crc2; a.csx = p->csx; a.ceshift = p->ceshift; a.halfself = p->halfself;
  for (int k = 0; k < MAXC; ++k) { a.pe[k] = p->pe[k]; a.pf[k] = p->pf[k]; }
  const bool coul = p->spec.coulomb != 0;
  const int np = std::max(p->npe, p->npf);
  hipStream_t st = as_stream(stream);
  switch (p->spec.model) {
    case PDX_PAIR_LJ: launch_model<PDX_PAIR_LJ>(coul, energy, np, G, dim3(blocks), st, a); break;
    case PDX_PAIR_WCA: launch_model<PDX_PAIR_WCA>(coul, energy, np, G, dim3(blocks), st, a); break;
    case PDX_PAIR_SOFT: launch_model<PDX_PAIR_SOFT>(coul, energy, np, G, dim3(blocks), st, a); break;
    default: launch_model<PDX_PAIR_NONE>(coul, energy, np, G, dim3(blocks), st, a); break;
  }
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}
