// Per-pair physics shared by the ELL-row and cluster-tile traversals: the model terms of one
// pair at squared distance r2, so both traversals evaluate every pair identically.
#pragma once

#include "pdx_internal.hpp"

namespace pdx {

// Horner with a compile-time length: the coefficients are uniform kernel arguments held in
// scalar registers for the whole row loop (a runtime length would re-load them every step).
template <int NP>
__device__ __forceinline__ float horner(const float (&c)[NP], float x) {
  float acc = c[NP - 1];
#pragma unroll
  for (int k = NP - 2; k >= 0; --k) acc = fmaf(acc, x, c[k]);
  return acc;
}

// LJ-form coefficients of one pair: uniform models take them from the arguments, the typed
// ones (mixed sigma, pair matrix) form them per pair.
struct LJCoef {
  float eps, sig2, rc2, eshift;
};

__host__ __device__ constexpr bool lj_form(int m) {
  return m == PDX_PAIR_LJ || m == PDX_PAIR_WCA || m == PDX_PAIR_MIXED_LJ || m == PDX_PAIR_MIXED_WCA ||
         m == PDX_PAIR_MATRIX_LJ || m == PDX_PAIR_MATRIX_WCA;
}
__host__ __device__ constexpr bool wca_energy(int m) {
  return m == PDX_PAIR_WCA || m == PDX_PAIR_MIXED_WCA || m == PDX_PAIR_MATRIX_WCA;
}
__host__ __device__ constexpr bool typed_model(int m) {
  return m == PDX_PAIR_MIXED_LJ || m == PDX_PAIR_MIXED_WCA || m == PDX_PAIR_MATRIX_LJ || m == PDX_PAIR_MATRIX_WCA;
}

// Per-visit physics shared by both traversal layouts: force factor f (force on i = f d) and
// pair energy u of one row entry (LJ/WCA/soft + real-space Coulomb polynomial), with the LJ
// coefficients c (PairModels.jl:39-81: LJ u - u_c, WCA u + eps).
// A: any argument block with the Coulomb fields cC, crc2, csx, ceshift (and, for the soft
// model, eps, sig2, rc2).
template <int MODEL, bool COUL, bool ENERGY, int NP, typename A>
__device__ __forceinline__ void pair_terms_c(const A &a, const LJCoef &c, const float (&pe)[NP],
                                             const float (&pf)[NP], float qi, float qj, float r2, float &f,
                                             float &u) {
  f = 0.f;
  u = 0.f;
  // with Coulomb, 1/r^2 is the square of the 1/r the Coulomb term needs (one quarter-rate
  // transcendental per pair instead of two)
  const float rinv = COUL ? __builtin_amdgcn_rsqf(r2) : 0.f;
  if (lj_form(MODEL)) {
    const bool in = (r2 > 0.f) & (r2 < c.rc2);
    const float invr2 = COUL ? rinv * rinv : __builtin_amdgcn_rcpf(r2);
    const float s2 = c.sig2 * invr2;
    const float s6 = s2 * s2 * s2;
    const float s12 = s6 * s6;
    const float fl = 24.0f * c.eps * (2.0f * s12 - s6) * invr2;
    f = in ? fl : 0.f;
    if (ENERGY) {
      const float ul = 4.0f * c.eps * (s12 - s6) + (wca_energy(MODEL) ? c.eps : -c.eshift);
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

// The LJ coefficients of the pair (i, j) of a typed model (PairModels.jl:61-81) from i's
// context (its diameter, or its 0-based type times ntypes) and j's datum (diameter or 0-based
// type). A: fields eps, eshift, rfac2 (mixed sigma) and tab, rows (sig2, eps, rc2, eshift).
template <int MODEL, typename A>
__device__ __forceinline__ LJCoef typed_coef(const A &a, float ci, float cj, int ti, int tj) {
  if (MODEL == PDX_PAIR_MIXED_LJ || MODEL == PDX_PAIR_MIXED_WCA) {
    const float s = 0.5f * (ci + cj);
    const float s2 = s * s;
    return LJCoef{a.eps, s2, a.rfac2 * s2, a.eshift};
  }
  const float4 t = a.tab[ti + tj];
  return LJCoef{t.y, t.x, t.z, t.w};
}

// The uniform models: coefficients from the argument block (fields eps, sig2, rc2, eshift).
template <int MODEL, bool COUL, bool ENERGY, int NP, typename A>
__device__ __forceinline__ void pair_terms(const A &a, const float (&pe)[NP], const float (&pf)[NP],
                                           float qi, float qj, float r2, float &f, float &u) {
  const LJCoef c{a.eps, a.sig2, a.rc2, a.eshift};
  pair_terms_c<MODEL, COUL, ENERGY, NP>(a, c, pe, pf, qi, qj, r2, f, u);
}

}  // namespace pdx
