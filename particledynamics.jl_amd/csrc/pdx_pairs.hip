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
#include "pdx_pairmath.hpp"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

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
  int add;
  float eps, sig2, rc2, eshift;   // LJ / WCA / soft
  float cC, crc2, csx, ceshift, halfself;
  float pe[MAXC];
  float pf[MAXC];
  // Coulomb exceptions (k_pair with COUL): CSR over particles, scale per entry
  const int32_t *exo, *exj;
  const float *exs;
  // FENE bonds (k_pair_acc only)
  const int32_t *boff, *bpart;
  int32_t *binvalid;
  float bk, bR02;
  // typed models (k_pair only): per-particle diameters, or 1-based types and the type-pair table
  const float *sigp;
  const int32_t *typ;
  const float4 *tab;  // (sig2, eps, rc2, eshift), row a*ntypes + b
  int ntypes;
  float rfac2;
};


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
    block_partials<7>(d, a.partial);
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
  constexpr bool TYPED = typed_model(MODEL);
  constexpr bool MIXED = MODEL == PDX_PAIR_MIXED_LJ || MODEL == PDX_PAIR_MIXED_WCA;
  if (active) {
    const float4 pi = a.packed[i];
    qi = pi.w;
    const int n = a.counts[i];
    const int32_t *row = a.nbr + i;
    // typed models: i's diameter, or its type's row of the table
    const float ci = MIXED ? a.sigp[i] : 0.f;
    const int ti = (TYPED && !MIXED) ? (a.typ[i] - 1) * a.ntypes : 0;
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
      float cj[U];
      int tj[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        cj[u] = MIXED ? a.sigp[jj[u]] : 0.f;
        tj[u] = (TYPED && !MIXED) ? a.typ[jj[u]] - 1 : 0;
      }
      float ddx[U], ddy[U], ddz[U];
      row_deltas<U>(pi, pj, hx, hy, hz, Lx, Ly, Lz, ddx, ddy, ddz);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float dx = ddx[u], dy = ddy[u], dz = ddz[u];
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        float f, uu;
        if (TYPED)
          pair_terms_c<MODEL, COUL, ENERGY, NP>(a, typed_coef<MODEL>(a, ci, cj[u], ti, tj[u]), pe, pf, qi, pj[u].w,
                                                r2, f, uu);
        else
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
    if (COUL && a.exo && qi != 0.0f) {
      // exceptions: + C q_i q_j (s - 1) / r at the minimum image, no cutoff, no potential shift
      for (int p = a.exo[i] + lane; p < a.exo[i + 1]; p += G) {
        const float4 pj = a.packed[a.exj[p]];
        const float sc = a.cC * qi * pj.w * (a.exs[p] - 1.0f);
        const float dx = periodic_delta(pi.x, pj.x, hx, Lx);
        const float dy = periodic_delta(pi.y, pj.y, hy, Ly);
        const float dz = periodic_delta(pi.z, pj.z, hz, Lz);
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const float invr = __builtin_amdgcn_rsqf(r2);
        const float f = sc * invr * invr * invr;
        ax = fmaf(f, dx, ax);
        ay = fmaf(f, dy, ay);
        az = fmaf(f, dz, az);
        if (ENERGY) {
          e += sc * invr + (r2 < a.crc2 ? a.cC * qi * pj.w * a.ceshift : 0.0f);
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

// ---- cell traversal: no rows. One workgroup per home cell of the list's grid stages the
// current (x, y, z, q) of its 27 neighbour cells into LDS (cell order, 2048 candidates per
// chunk), then each wave takes home particles in turn: 64 candidates at a time are tested
// against the cutoff, the hits are ballot-compacted into a wave-private LDS queue in candidate
// order and evaluated 64 at a time with every lane busy. Candidate order (cells z, y, x, then
// ids inside a cell) and all sums are fixed, so results are deterministic.
struct CellArgs {
  const float4 *sortedq;          // current (x, y, z, q) in cell order
  const int32_t *ids;             // 1-based storage id of cell slot t
  const int32_t *offsets;         // 1-based cell offsets (ncell + 1)
  int nx, ny, nz;
  float cut2;                     // the largest cutoff^2 of the pair terms
};

constexpr int CSTAGE = 2048;  // staged candidates per chunk (32 KiB)
constexpr int CHOME = 192;    // home particles per pass (their partial sums live in LDS)

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// cell-ordered copy of the current coordinates and charges (every step, after the list check)
__global__ __launch_bounds__(TPB) void k_refresh_sorted(int N, const int32_t *__restrict__ ids,
                                                         const float4 *__restrict__ packed, float4 *sortedq,
                                                         const int64_t *state) {
  if (!(state[PDX_DN_VALID] != 0 && state[PDX_DN_ERROR] == 0)) return;
  const int t = blockIdx.x * TPB + threadIdx.x;
  if (t < N) sortedq[t] = packed[ids[t] - 1];
}

template <int MODEL, bool COUL, bool ENERGY, int NP>
__global__ __launch_bounds__(TPB) void k_pair_cells(const PairArgs a, const CellArgs c) {
  const int64_t *st = a.state;
  const bool valid = st[PDX_DN_VALID] != 0 && st[PDX_DN_ERROR] == 0;  // uniform
  if (!valid) return;
  constexpr int WPB = TPB / WAVE;
  constexpr int NACC = ENERGY ? 10 : 3;
  __shared__ float4 stage[CSTAGE];
  __shared__ float4 queue[WPB][2 * WAVE];
  __shared__ float accs[CHOME][NACC];
  __shared__ int seg_start[27], seg_pre[28];
  const int w = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  const unsigned long long below = (1ull << lane) - 1ull;
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  const int home = blockIdx.x;
  const int cx = home % nx, cy = (home / nx) % ny, cz = home / (nx * ny);
  const float Lx = a.geom[0], Ly = a.geom[1], Lz = a.geom[2];
  const float hx = 0.5f * Lx, hy = 0.5f * Ly, hz = 0.5f * Lz;
  float pe[NP], pf[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    pe[k] = a.pe[k];
    pf[k] = a.pf[k];
  }
  if (threadIdx.x < 27) {
    const int k = threadIdx.x;
    int z2 = cz + k / 9 - 1, y2 = cy + (k / 3) % 3 - 1, x2 = cx + k % 3 - 1;
    z2 += z2 < 0 ? nz : 0; z2 -= z2 >= nz ? nz : 0;
    y2 += y2 < 0 ? ny : 0; y2 -= y2 >= ny ? ny : 0;
    x2 += x2 < 0 ? nx : 0; x2 -= x2 >= nx ? nx : 0;
    const int cc = (z2 * ny + y2) * nx + x2;
    seg_start[k] = c.offsets[cc] - 1;
    seg_pre[k + 1] = c.offsets[cc + 1] - c.offsets[cc];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    seg_pre[0] = 0;
    for (int k = 1; k <= 27; ++k) seg_pre[k] += seg_pre[k - 1];
  }
  __syncthreads();
  const int total = seg_pre[27];
  const int hs = c.offsets[home] - 1, he = c.offsets[home + 1] - 1;
  float4 *Q = queue[w];
  double tot[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int hg = hs; hg < he; hg += CHOME) {  // uniform
    const int hn = min(CHOME, he - hg);
    for (int k = threadIdx.x; k < hn * NACC; k += TPB) (&accs[0][0])[k] = 0.f;
    for (int c0 = 0; c0 < total; c0 += CSTAGE) {  // uniform
      const int nch = min(CSTAGE, total - c0);
      __syncthreads();  // the previous chunk is no longer read
      for (int f = threadIdx.x; f < nch; f += TPB) {
        const int g = c0 + f;
        int sg = 0;
        while (seg_pre[sg + 1] <= g) ++sg;
        stage[f] = c.sortedq[seg_start[sg] + (g - seg_pre[sg])];
      }
      __syncthreads();
      for (int p = w; p < hn; p += WPB) {
        const int t = hg + p;
        const float4 pi = c.sortedq[t];
        const float qi = pi.w;
        float ax = 0.f, ay = 0.f, az = 0.f, e = 0.f;
        float v[6] = {0, 0, 0, 0, 0, 0};
        auto evaluate = [&](int cnt) {
          if (lane < cnt) {
            const float4 d = Q[lane];
            const float r2 = fmaf(d.x, d.x, fmaf(d.y, d.y, d.z * d.z));
            float f, uu;
            pair_terms<MODEL, COUL, ENERGY, NP>(a, pe, pf, qi, d.w, r2, f, uu);
            ax = fmaf(f, d.x, ax);
            ay = fmaf(f, d.y, ay);
            az = fmaf(f, d.z, az);
            if (ENERGY) {
              e += uu;
              const float fxij = f * d.x, fyij = f * d.y, fzij = f * d.z;
              v[0] = fmaf(d.x, fxij, v[0]); v[1] = fmaf(d.y, fyij, v[1]); v[2] = fmaf(d.z, fzij, v[2]);
              v[3] = fmaf(d.x, fyij, v[3]); v[4] = fmaf(d.x, fzij, v[4]); v[5] = fmaf(d.y, fzij, v[5]);
            }
          }
        };
        int qn = 0;  // uniform
        for (int k0 = 0; k0 < nch; k0 += WAVE) {
          const int k = k0 + lane;
          bool hit = false;
          float4 d = make_float4(0.f, 0.f, 0.f, 0.f);
          if (k < nch) {  // the particle itself has r^2 = 0, which every pair term excludes
            const float4 pj = stage[k];
            d.x = periodic_delta(pi.x, pj.x, hx, Lx);
            d.y = periodic_delta(pi.y, pj.y, hy, Ly);
            d.z = periodic_delta(pi.z, pj.z, hz, Lz);
            d.w = pj.w;
            hit = fmaf(d.x, d.x, fmaf(d.y, d.y, d.z * d.z)) < c.cut2;
          }
          const unsigned long long m = __ballot(hit);
          if (hit) Q[qn + __popcll(m & below)] = d;
          qn += __popcll(m);
          if (qn >= WAVE) {
            wave_lds_sync();
            evaluate(WAVE);
            const float4 rest = lane < qn - WAVE ? Q[WAVE + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
            wave_lds_sync();
            if (lane < qn - WAVE) Q[lane] = rest;
            qn -= WAVE;
            wave_lds_sync();
          }
        }
        wave_lds_sync();
        evaluate(qn);
        wave_lds_sync();
        if (COUL && a.exo && c0 == 0 && qi != 0.0f) {
          const int i = c.ids[t] - 1;
          for (int q = a.exo[i] + lane; q < a.exo[i + 1]; q += WAVE) {
            const float4 pj = a.packed[a.exj[q]];
            const float sc = a.cC * qi * pj.w * (a.exs[q] - 1.0f);
            const float dx = periodic_delta(pi.x, pj.x, hx, Lx);
            const float dy = periodic_delta(pi.y, pj.y, hy, Ly);
            const float dz = periodic_delta(pi.z, pj.z, hz, Lz);
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            const float invr = __builtin_amdgcn_rsqf(r2);
            const float f = sc * invr * invr * invr;
            ax = fmaf(f, dx, ax);
            ay = fmaf(f, dy, ay);
            az = fmaf(f, dz, az);
            if (ENERGY) {
              e += sc * invr + (r2 < a.crc2 ? a.cC * qi * pj.w * a.ceshift : 0.0f);
              const float fxij = f * dx, fyij = f * dy, fzij = f * dz;
              v[0] = fmaf(dx, fxij, v[0]); v[1] = fmaf(dy, fyij, v[1]); v[2] = fmaf(dz, fzij, v[2]);
              v[3] = fmaf(dx, fyij, v[3]); v[4] = fmaf(dx, fzij, v[4]); v[5] = fmaf(dy, fzij, v[5]);
            }
          }
        }
        ax = wave_sum_width(ax, WAVE);
        ay = wave_sum_width(ay, WAVE);
        az = wave_sum_width(az, WAVE);
        if (ENERGY) {
          e = wave_sum_width(e, WAVE);
#pragma unroll
          for (int k = 0; k < 6; ++k) v[k] = wave_sum_width(v[k], WAVE);
        }
        if (lane == 0) {  // only this wave touches particle p: chunk sums in chunk order
          accs[p][0] += ax; accs[p][1] += ay; accs[p][2] += az;
          if (ENERGY) {
            accs[p][3] += e;
#pragma unroll
            for (int k = 0; k < 6; ++k) accs[p][4 + k] += v[k];
          }
        }
      }
    }
    __syncthreads();
    // outputs of the pass: thread per home particle (half energies and half virials)
    for (int p = threadIdx.x; p < hn; p += TPB) {
      const int t = hg + p;
      const int i = c.ids[t] - 1;
      const float qi = c.sortedq[t].w;
      const float fx = accs[p][0], fy = accs[p][1], fz = accs[p][2];
      if (a.add) {
        a.fx[i] += fx; a.fy[i] += fy; a.fz[i] += fz;
      } else {
        a.fx[i] = fx; a.fy[i] = fy; a.fz[i] = fz;
      }
      if (ENERGY) {
        float ei = 0.5f * accs[p][3];
        if (COUL) ei -= a.halfself * qi * qi;
        float wv[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) wv[k] = 0.5f * accs[p][4 + k];
        if (a.Epot) a.Epot[i] = a.add ? a.Epot[i] + ei : ei;
        if (a.V) {
#pragma unroll
          for (int k = 0; k < 6; ++k) {
            float *q = a.V + (size_t)k * a.N + i;
            *q = a.add ? *q + wv[k] : wv[k];
          }
        }
        tot[0] += (double)ei;
#pragma unroll
        for (int k = 0; k < 6; ++k) tot[1 + k] += (double)wv[k];
      }
    }
  }
  if (ENERGY && a.totals) block_partials<7>(tot, a.partial);
}

template <int MODEL, bool COUL, bool ENERGY, int NP>
void launch_cells_np(dim3 grid, hipStream_t st, const PairArgs &a, const CellArgs &c) {
  hipLaunchKernelGGL((k_pair_cells<MODEL, COUL, ENERGY, NP>), grid, dim3(TPB), 0, st, a, c);
}

template <int MODEL>
void launch_cells(bool coul, bool energy, int np, dim3 grid, hipStream_t st, const PairArgs &a, const CellArgs &c) {
  if (coul) {
    if (np <= 16) {
      if (energy) launch_cells_np<MODEL, true, true, 16>(grid, st, a, c);
      else launch_cells_np<MODEL, true, false, 16>(grid, st, a, c);
    } else {
      if (energy) launch_cells_np<MODEL, true, true, 24>(grid, st, a, c);
      else launch_cells_np<MODEL, true, false, 24>(grid, st, a, c);
    }
  } else {
    if (energy) launch_cells_np<MODEL, false, true, 1>(grid, st, a, c);
    else launch_cells_np<MODEL, false, false, 1>(grid, st, a, c);
  }
}

// ---- per-particle sums in a chosen precision (the bead-spring precision study). The pair
// arithmetic is the FP32 arithmetic of k_pair; only the accumulation of the row sums changes:
// FP32 (fma, as k_pair), 64-bit fixed point at 2^32 (exact, so independent of the order and of
// the lane split), or FP64 (exact products, rounded sums).
template <int ACC> struct Sum;
template <> struct Sum<PDX_ACC_FP32> {
  float v = 0.f;
  __device__ __forceinline__ void add_prod(float a, float b) { v = fmaf(a, b, v); }
  __device__ __forceinline__ void add(float a) { v += a; }
  __device__ __forceinline__ double value() const { return (double)v; }
  template <int G> __device__ __forceinline__ void combine() {
#pragma unroll
    for (int o = G >> 1; o >= 1; o >>= 1) v += __shfl_xor(v, o, G);
  }
};
template <> struct Sum<PDX_ACC_FIXED64> {
  long long v = 0;
  __device__ __forceinline__ void add_prod(float a, float b) { v += __float2ll_rn(a * b * 4294967296.0f); }
  __device__ __forceinline__ void add(float a) { v += __float2ll_rn(a * 4294967296.0f); }
  __device__ __forceinline__ double value() const { return (double)v * 2.3283064365386962890625e-10; }
  template <int G> __device__ __forceinline__ void combine() {
#pragma unroll
    for (int o = G >> 1; o >= 1; o >>= 1) v += __shfl_xor(v, o, G);
  }
};
template <> struct Sum<PDX_ACC_FP64> {
  double v = 0.0;
  __device__ __forceinline__ void add_prod(float a, float b) { v += (double)a * (double)b; }
  __device__ __forceinline__ void add(float a) { v += (double)a; }
  __device__ __forceinline__ double value() const { return v; }
  template <int G> __device__ __forceinline__ void combine() {
#pragma unroll
    for (int o = G >> 1; o >= 1; o >>= 1) v += __shfl_xor(v, o, G);
  }
};

template <int MODEL, bool ENERGY, int G, int ACC>
__global__ __launch_bounds__(TPB) void k_pair_acc(const PairArgs a, int bonded) {
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
  Sum<ACC> ax, ay, az, e, v[6];
  const float none[1] = {0.f};
  if (active) {
    const float4 pi = a.packed[i];
    const int n = a.counts[i];
    const int32_t *row = a.nbr + i;
    // bonded exclusions (BondExclusionFeature, src/engine/Bonded.jl:60-92): a row entry that is
    // one of i's bond partners gets no pair term; the first four partners live in registers
    const bool excl = (bonded & 2) != 0;
    int bp[4] = {-1, -1, -1, -1};
    int bx0 = 0, nbx = 0;
    if (excl) {
      bx0 = a.boff[i];
      nbx = a.boff[i + 1] - bx0;
#pragma unroll
      for (int k = 0; k < 4; ++k) bp[k] = k < nbx ? a.bpart[bx0 + k] : -1;
    }
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
      float ddx[U], ddy[U], ddz[U];
      row_deltas<U>(pi, pj, hx, hy, hz, Lx, Ly, Lz, ddx, ddy, ddz);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float dx = ddx[u], dy = ddy[u], dz = ddz[u];
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        float f, uu;
        pair_terms<MODEL, false, ENERGY, 1>(a, none, none, 0.f, 0.f, r2, f, uu);
        if (excl) {
          bool ex = (jj[u] == bp[0]) | (jj[u] == bp[1]) | (jj[u] == bp[2]) | (jj[u] == bp[3]);
          for (int k = 4; k < nbx; ++k) ex |= a.bpart[bx0 + k] == jj[u];
          f = ex ? 0.f : f;
          uu = ex ? 0.f : uu;
        }
        ax.add_prod(f, dx);
        ay.add_prod(f, dy);
        az.add_prod(f, dz);
        if (ENERGY) {
          e.add(uu);
          const float fxij = f * dx, fyij = f * dy, fzij = f * dz;
          v[0].add_prod(dx, fxij); v[1].add_prod(dy, fyij); v[2].add_prod(dz, fzij);
          v[3].add_prod(dx, fyij); v[4].add_prod(dx, fzij); v[5].add_prod(dy, fzij);
        }
      }
    }
    if (bonded & 1) {
      const int b0 = a.boff[i], nb = a.boff[i + 1] - b0;
      const float R02 = a.bR02;
      for (int t = lane; t < nb; t += G) {
        const float4 pj = a.packed[a.bpart[b0 + t]];
        const float dx = periodic_delta(pi.x, pj.x, hx, Lx);
        const float dy = periodic_delta(pi.y, pj.y, hy, Ly);
        const float dz = periodic_delta(pi.z, pj.z, hz, Lz);
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        if (!(r2 < R02)) {  // out of the FENE domain (or not finite): latch, add nothing
          atomicMin(a.binvalid, i);
          continue;
        }
        // subtraction before division keeps near-limit extensions; log1p for short bonds
        // (_fene_denom, _fene_log, src/bonded/FENEBondCUDA.jl)
        const float denom = (R02 - r2) / R02;
        const float f = -a.bk / denom;
        ax.add_prod(f, dx);
        ay.add_prod(f, dy);
        az.add_prod(f, dz);
        if (ENERGY) {
          const float lg = r2 < 0.5f * R02 ? log1pf(-r2 / R02) : logf(denom);
          e.add(-0.5f * a.bk * R02 * lg);
          const float fxij = f * dx, fyij = f * dy, fzij = f * dz;
          v[0].add_prod(dx, fxij); v[1].add_prod(dy, fyij); v[2].add_prod(dz, fzij);
          v[3].add_prod(dx, fyij); v[4].add_prod(dx, fzij); v[5].add_prod(dy, fzij);
        }
      }
    }
  }
  if (G > 1) {
    ax.template combine<G>();
    ay.template combine<G>();
    az.template combine<G>();
    if (ENERGY) {
      e.template combine<G>();
#pragma unroll
      for (int c = 0; c < 6; ++c) v[c].template combine<G>();
    }
  }
  // every pair and bond was visited from both ends: half energies and half virials
  const bool owner = active && lane == 0;
  double ei = 0.0, w[6] = {0, 0, 0, 0, 0, 0};
  if (owner) {
    const float fx = (float)ax.value(), fy = (float)ay.value(), fz = (float)az.value();
    if (a.add) {
      a.fx[i] += fx; a.fy[i] += fy; a.fz[i] += fz;
    } else {
      a.fx[i] = fx; a.fy[i] = fy; a.fz[i] = fz;
    }
    if (ENERGY) {
      ei = 0.5 * e.value();
#pragma unroll
      for (int c = 0; c < 6; ++c) w[c] = 0.5 * v[c].value();
      if (a.Epot) a.Epot[i] = a.add ? a.Epot[i] + (float)ei : (float)ei;
      if (a.V) {
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          float *p = a.V + (size_t)c * a.N + i;
          *p = a.add ? *p + (float)w[c] : (float)w[c];
        }
      }
    }
  }
  if (ENERGY && a.totals) {
    const double d[7] = {ei, w[0], w[1], w[2], w[3], w[4], w[5]};
    block_partials<7>(d, a.partial);
  }
}

template <int MODEL, bool ENERGY, int ACC>
void launch_acc_g(int G, dim3 grid, hipStream_t st, const PairArgs &a, int bonded) {
  switch (G) {
    case 1: hipLaunchKernelGGL((k_pair_acc<MODEL, ENERGY, 1, ACC>), grid, dim3(TPB), 0, st, a, bonded); break;
    case 2: hipLaunchKernelGGL((k_pair_acc<MODEL, ENERGY, 2, ACC>), grid, dim3(TPB), 0, st, a, bonded); break;
    case 4: hipLaunchKernelGGL((k_pair_acc<MODEL, ENERGY, 4, ACC>), grid, dim3(TPB), 0, st, a, bonded); break;
    default: hipLaunchKernelGGL((k_pair_acc<MODEL, ENERGY, 8, ACC>), grid, dim3(TPB), 0, st, a, bonded); break;
  }
}

template <int MODEL>
void launch_acc(int acc, bool energy, int G, dim3 grid, hipStream_t st, const PairArgs &a, int bonded) {
  if (acc == PDX_ACC_FIXED64) {
    if (energy) launch_acc_g<MODEL, true, PDX_ACC_FIXED64>(G, grid, st, a, bonded);
    else launch_acc_g<MODEL, false, PDX_ACC_FIXED64>(G, grid, st, a, bonded);
  } else if (acc == PDX_ACC_FP64) {
    if (energy) launch_acc_g<MODEL, true, PDX_ACC_FP64>(G, grid, st, a, bonded);
    else launch_acc_g<MODEL, false, PDX_ACC_FP64>(G, grid, st, a, bonded);
  } else {
    if (energy) launch_acc_g<MODEL, true, PDX_ACC_FP32>(G, grid, st, a, bonded);
    else launch_acc_g<MODEL, false, PDX_ACC_FP32>(G, grid, st, a, bonded);
  }
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

// Typed models: the lanes layout only.
template <int MODEL, bool COUL, bool ENERGY, int NP>
void launch_typed_g(int G, dim3 grid, hipStream_t st, const PairArgs &a) {
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

template <int MODEL>
void launch_typed(bool coul, bool energy, int np, int G, dim3 grid, hipStream_t st, const PairArgs &a) {
  if (coul) {
    if (np <= 16) {
      if (energy) launch_typed_g<MODEL, true, true, 16>(G, grid, st, a);
      else launch_typed_g<MODEL, true, false, 16>(G, grid, st, a);
    } else {
      if (energy) launch_typed_g<MODEL, true, true, 24>(G, grid, st, a);
      else launch_typed_g<MODEL, true, false, 24>(G, grid, st, a);
    }
  } else {
    if (energy) launch_typed_g<MODEL, false, true, 1>(G, grid, st, a);
    else launch_typed_g<MODEL, false, false, 1>(G, grid, st, a);
  }
}

// Coulomb polynomials are zero-padded to 14, 16 or 24 monomials (Horner over exact zeros); the
// 9-10 A real-space fits at 1e-5 take 13-14.
template <int MODEL>
void launch_model(bool coul, bool energy, int np, int G, dim3 grid, hipStream_t st, const PairArgs &a) {
  if (coul) {
    if (np <= 14) {
      if (energy) launch_g<MODEL, true, true, 14>(G, grid, st, a);
      else launch_g<MODEL, true, false, 14>(G, grid, st, a);
    } else if (np <= 16) {
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
  if (sp.model < PDX_PAIR_NONE || sp.model > PDX_PAIR_MATRIX_WCA) return fail(PDX_ERR_ARGUMENT, "unknown pair model");
  const bool typed = typed_model(sp.model);
  if (sp.model != PDX_PAIR_NONE && sp.model != PDX_PAIR_MATRIX_LJ && sp.model != PDX_PAIR_MATRIX_WCA) {
    if (!(std::isfinite(sp.epsilon) && (typed || (std::isfinite(sp.sigma) && sp.sigma > 0))))
      return fail(PDX_ERR_ARGUMENT, "pair epsilon and sigma must be finite, sigma positive");
    if ((sp.model == PDX_PAIR_LJ || sp.model == PDX_PAIR_MIXED_LJ || sp.model == PDX_PAIR_MIXED_WCA) &&
        !(std::isfinite(sp.cutoff) && sp.cutoff > 0))
      return fail(PDX_ERR_ARGUMENT, "LJ cutoff (the rcut factor of the mixed-sigma models) must be finite and positive");
  }
  if (typed && sp.lanes < 0) return fail(PDX_ERR_ARGUMENT, "typed pair models use the lanes layout (lanes >= 0)");
  if (sp.coulomb) {
    if (!(sp.coulomb_C > 0 && std::isfinite(sp.coulomb_C)))
      return fail(PDX_ERR_ARGUMENT, "coulomb_prefactor must be finite and positive");
    if (!(sp.coulomb_cutoff > 0 && std::isfinite(sp.coulomb_cutoff)))
      return fail(PDX_ERR_ARGUMENT, "real_cutoff must be finite and positive");
    if (sp.coulomb != 1 && sp.coulomb != 2) return fail(PDX_ERR_ARGUMENT, "coulomb must be 0, 1 (Gaussian) or 2 (prolate)");
    if (sp.coulomb == 1 && !(sp.alpha > 0 && std::isfinite(sp.alpha)))
      return fail(PDX_ERR_ARGUMENT, "Ewald alpha must be finite and positive.");
    if (sp.coulomb == 2 && !(sp.split_shape > 0 && std::isfinite(sp.split_shape)))
      return fail(PDX_ERR_ARGUMENT, "ESP split shape must be finite and positive.");
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
  } else if (sp.model == PDX_PAIR_MIXED_LJ || sp.model == PDX_PAIR_MIXED_WCA) {
    // rc = factor sigma_ij: (rc/sigma_ij)^2 = factor^2, and the LJ shift u(rc) is one constant
    const float f = (float)sp.cutoff;
    p->rfac2 = f * f;
    if (sp.model == PDX_PAIR_MIXED_LJ && sp.shift) {
      const float s2 = 1.0f / (f * f), s6 = s2 * s2 * s2;
      p->eshift = 4.0f * eps * (s6 * s6 - s6);
    }
  }
  if (sp.coulomb) {
    const double rc = sp.coulomb_cutoff, rc2 = rc * rc, a = sp.alpha;
    // the split's long-range part 1/r - E_s/r and its force factor as functions of s = r^2:
    // Gaussian (erf(alpha r)/r) or prolate (a(t)/r by quadrature, Splits.jl:163-202)
    ProlateSplit ps;
    if (sp.coulomb == 2 && !ps.make(rc, sp.split_shape)) {
      delete p;
      return fail(PDX_ERR_ARGUMENT, "prolate split set-up failed");
    }
    auto correction = [&](double s, double &corr, double &fcorr) {
      if (sp.coulomb == 2) ps.correction(s, corr, fcorr);
      else ewald_correction(a, s, corr, fcorr);
    };
    p->cC = (float)sp.coulomb_C;
    p->crc2 = (float)rc2;
    p->csx = (float)(2.0 / rc2);
    // cutoff energy E_s(rc)/rc: erfc(alpha rc)/rc, exactly zero for the prolate split (Splits.jl:77-78)
    p->ceshift = (sp.coulomb_shift && sp.coulomb == 1) ? (float)(std::erfc(a * rc) / rc) : 0.0f;
    p->halfself = (float)(0.5 * sp.coulomb_C * (sp.coulomb == 2 ? ps.self_coefficient() : 2.0 * a / std::sqrt(M_PI)));
    const int nodes = 64;
    std::vector<double> fe(nodes), ff(nodes);
    for (int k = 0; k < nodes; ++k) {
      const double x = std::cos(M_PI * (k + 0.5) / nodes);
      double corr, fcorr;
      correction((x + 1.0) * rc2 / 2.0, corr, fcorr);
      fe[k] = -corr;
      ff[k] = fcorr;
    }
    // FP32 tolerances of coulomb_polynomial (PairFunctor.jl:58-61): 1e-8 energy, 1e-7 force
    std::vector<double> me = chebyshev_monomial(fe, 1e-8), mf = chebyshev_monomial(ff, 1e-7);
    if ((int)me.size() > MAXC || (int)mf.size() > MAXC) {
      delete p;
      return fail(PDX_ERR_ARGUMENT, "the split's long-range polynomial needs more than 24 terms (alpha*cutoff or the split shape too large)");
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
      correction((x + 1.0) * rc2 / 2.0, corr, fcorr);
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

namespace {
// fixed_partial == nullptr: the model's own partial buffer, grown here when too small (a host
// boundary: it synchronises and allocates); otherwise that caller-provided workspace of
// fixed_blocks per-block slots, and a launch needing more is refused.
int pair_forces_impl(pdx_pair_t p, pdx_nl_t nl, float *fx, float *fy, float *fz, float *Epot, float *V,
                     double *totals, int32_t flags, pdx_stream_t stream, double *fixed_partial, int fixed_blocks) {
  if (!p || !nl) return fail(PDX_ERR_ARGUMENT, "null pair model or neighbor list");
  if (!fx || !fy || !fz) return fail(PDX_ERR_ARGUMENT, "null force array");
  const bool energy = (flags & PDX_ENERGY) != 0;
  const int N = nl->N;
  // lanes > 0: G lanes of a wave per particle; lanes < 0: the split layout with -lanes waves
  // per block of 64 particles; 0: automatic
  const bool acc_path = p->accumulate != PDX_ACC_FP32 || p->bonds != nullptr;
  const bool cells = p->traversal == PDX_TRAVERSAL_CELLS;
  const bool typed = typed_model(p->spec.model);
  if (typed) {
    if (acc_path || cells)
      return fail(PDX_ERR_ARGUMENT, "typed pair models use the rows traversal with FP32 sums and no bonds");
    if (p->typed_n != N)
      return fail(PDX_ERR_DIMENSION, p->typed_n == 0 ? "typed pair model without its per-particle data "
                                                       "(pdx_pair_set_sigmas / pdx_pair_set_types)"
                                                     : "typed pair data and neighbor list particle counts differ");
  }
  if (cells) {
    if (acc_path) return fail(PDX_ERR_ARGUMENT, "the cell traversal uses FP32 sums without bonds");
    if (nl->allpairs) return fail(PDX_ERR_ARGUMENT, "the cell traversal needs a cell grid (not all-pairs rows)");
  }
  if (acc_path) {
    if (p->spec.coulomb) return fail(PDX_ERR_ARGUMENT, "accumulation modes and bonds are not available with Coulomb");
    if (p->spec.model != PDX_PAIR_LJ && p->spec.model != PDX_PAIR_WCA && p->spec.model != PDX_PAIR_SOFT)
      return fail(PDX_ERR_ARGUMENT, "accumulation modes and bonds need the LJ, WCA or soft model");
    if (p->lanes < 0) return fail(PDX_ERR_ARGUMENT, "accumulation modes and bonds use the lanes layout (lanes >= 0)");
    if (p->bonds && p->bonds->N != nl->N) return fail(PDX_ERR_DIMENSION, "bonds and neighbor list particle counts differ");
  }
  int G = p->lanes;
  if (G == 0) {
    // about 2^18 threads: enough waves per SIMD to hide the neighbour gathers while a wave's
    // row slots stay in few cache lines (measured on the 64 000-ion system: 4 lanes beat 2,
    // 8 and 16, and every split layout)
    const int64_t want = std::max<int64_t>(1, 262144 / std::max(N, 1));
    G = 1;
    while (G * 2 <= want && G < 64) G <<= 1;
  }
  if (acc_path) G = std::min(G, 8);
  const int blocks = cells ? nl->ncell
                           : (G < 0 ? (N + WAVE - 1) / WAVE : (int)(((int64_t)N * G + TPB - 1) / TPB));
  double *partial = fixed_partial;
  if (energy && totals) {
    if (fixed_partial) {
      if (blocks > fixed_blocks)
        return fail(PDX_ERR_ARGUMENT, std::to_string(blocks) + " workgroups exceed the fixed workspace of " +
                                          std::to_string(fixed_blocks) + " (fewer particles or lanes, or pdx_pair_forces)");
    } else if ((size_t)blocks * 7 > p->partial.n) {
      PDX_HIP(hipStreamSynchronize(as_stream(stream)));
      int rc = p->partial.alloc((size_t)blocks * 7);
      if (rc) return rc;
    }
  }
  if (!fixed_partial) partial = p->partial.p;
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
  a.partial = partial;
  a.add = (flags & PDX_ADD) ? 1 : 0;
  a.eps = p->eps; a.sig2 = p->sig2; a.rc2 = p->rc2; a.eshift = p->eshift;
  a.cC = p->cC; a.crc2 = p->crc2; a.csx = p->csx; a.ceshift = p->ceshift; a.halfself = p->halfself;
  a.sigp = p->sigp.p; a.typ = p->typ.p; a.tab = p->tab.p; a.ntypes = p->ntypes; a.rfac2 = p->rfac2;
  for (int k = 0; k < MAXC; ++k) { a.pe[k] = p->pe[k]; a.pf[k] = p->pf[k]; }
  const bool coul = p->spec.coulomb != 0;
  const int np = std::max(p->npe, p->npf);
  hipStream_t st = as_stream(stream);
  if (p->nexc > 0) {
    if (p->exc_n != N) return fail(PDX_ERR_DIMENSION, "exceptions and neighbor list particle counts differ");
    if (G < 0 && !cells) return fail(PDX_ERR_ARGUMENT, "exceptions use the lanes layout (lanes >= 0)");
    a.exo = p->exc_off.p;
    a.exj = p->exc_j.p;
    a.exs = p->exc_s.p;
  }
  // FP64 totals: per-block partials, then one block adds them in block order
  auto finish = [&]() -> int {
    PDX_LAUNCH_CHECK();
    if (energy && totals) {
      hipLaunchKernelGGL(k_sum_partials<7>, dim3(1), dim3(SUM_THREADS), 0, st, (const double *)partial, blocks,
                         totals, a.add, (const int64_t *)nl->state.p);
      PDX_LAUNCH_CHECK();
    }
    return PDX_OK;
  };
  if (cells) {
    CellArgs c{};
    c.sortedq = nl->sortedq.p;
    c.ids = nl->particle_ids_sorted.p;
    c.offsets = nl->cell_offsets.p;
    c.nx = nl->nx; c.ny = nl->ny; c.nz = nl->nz;
    float cut2 = p->spec.model != PDX_PAIR_NONE ? p->rc2 : 0.0f;
    if (coul) cut2 = std::max(cut2, p->crc2);
    c.cut2 = cut2;
    hipLaunchKernelGGL(k_refresh_sorted, dim3((N + TPB - 1) / TPB), dim3(TPB), 0, st, N, nl->particle_ids_sorted.p,
                       nl->packed.p, nl->sortedq.p, nl->state.p);
    const dim3 grid(nl->ncell);
    switch (p->spec.model) {
      case PDX_PAIR_LJ: launch_cells<PDX_PAIR_LJ>(coul, energy, np, grid, st, a, c); break;
      case PDX_PAIR_WCA: launch_cells<PDX_PAIR_WCA>(coul, energy, np, grid, st, a, c); break;
      case PDX_PAIR_SOFT: launch_cells<PDX_PAIR_SOFT>(coul, energy, np, grid, st, a, c); break;
      default: launch_cells<PDX_PAIR_NONE>(coul, energy, np, grid, st, a, c); break;
    }
    return finish();
  }
  if (acc_path) {
    const int bonded = (p->bonds != nullptr ? 1 : 0) | (p->bonds != nullptr && p->exclude_bonded ? 2 : 0);
    if (bonded) {
      a.boff = p->bonds->offsets.p;
      a.bpart = p->bonds->partners.p;
      a.binvalid = p->bonds->invalid.p;
      a.bk = (float)p->bonds->k;
      a.bR02 = (float)(p->bonds->R0 * p->bonds->R0);
    }
    if (p->spec.model == PDX_PAIR_LJ) launch_acc<PDX_PAIR_LJ>(p->accumulate, energy, G, dim3(blocks), st, a, bonded);
    else if (p->spec.model == PDX_PAIR_WCA) launch_acc<PDX_PAIR_WCA>(p->accumulate, energy, G, dim3(blocks), st, a, bonded);
    else launch_acc<PDX_PAIR_SOFT>(p->accumulate, energy, G, dim3(blocks), st, a, bonded);
    return finish();
  }
  switch (p->spec.model) {
    case PDX_PAIR_MIXED_LJ: launch_typed<PDX_PAIR_MIXED_LJ>(coul, energy, np, G, dim3(blocks), st, a); break;
    case PDX_PAIR_MIXED_WCA: launch_typed<PDX_PAIR_MIXED_WCA>(coul, energy, np, G, dim3(blocks), st, a); break;
    case PDX_PAIR_MATRIX_LJ: launch_typed<PDX_PAIR_MATRIX_LJ>(coul, energy, np, G, dim3(blocks), st, a); break;
    case PDX_PAIR_MATRIX_WCA: launch_typed<PDX_PAIR_MATRIX_WCA>(coul, energy, np, G, dim3(blocks), st, a); break;
    case PDX_PAIR_LJ: launch_model<PDX_PAIR_LJ>(coul, energy, np, G, dim3(blocks), st, a); break;
    case PDX_PAIR_WCA: launch_model<PDX_PAIR_WCA>(coul, energy, np, G, dim3(blocks), st, a); break;
    case PDX_PAIR_SOFT: launch_model<PDX_PAIR_SOFT>(coul, energy, np, G, dim3(blocks), st, a); break;
    default: launch_model<PDX_PAIR_NONE>(coul, energy, np, G, dim3(blocks), st, a); break;
  }
  return finish();
}
}  // namespace

extern "C" int pdx_pair_forces(pdx_pair_t p, pdx_nl_t nl, float *fx, float *fy, float *fz, float *Epot,
                               float *V, double *totals, int32_t flags, pdx_stream_t stream) {
  return pair_forces_impl(p, nl, fx, fy, fz, Epot, V, totals, flags, stream, nullptr, 0);
}

extern "C" int pdx_pair_compute(pdx_pair_t p, pdx_nl_t nl, const float *x, const float *y, const float *z,
                                const float *q, float *fx, float *fy, float *fz, float *Epot, float *V, double *totals,
                                int32_t flags, pdx_stream_t stream) {
  if (!p || !nl) return fail(PDX_ERR_ARGUMENT, "null pair model or neighbor list");
  if (!x || !y || (!z && nl->dim == 3)) return fail(PDX_ERR_ARGUMENT, "null coordinate array");
  if (nl->packed_src[0] != x || nl->packed_src[1] != y || nl->packed_src[2] != z || nl->packed_src[3] != q)
    return fail(PDX_ERR_STATE, "the coordinates (and charges) differ from those the list's last check packed; "
                               "call pdx_nl_update / pdx_nl_check with them first");
  return pair_forces_impl(p, nl, fx, fy, fz, Epot, V, totals, flags, stream, nullptr, 0);
}

extern "C" int pdx_pair_set_sigmas(pdx_pair_t p, int32_t n, const float *sigma) {
  if (!p || (!sigma && n > 0)) return fail(PDX_ERR_ARGUMENT, "null argument");
  if (p->spec.model != PDX_PAIR_MIXED_LJ && p->spec.model != PDX_PAIR_MIXED_WCA)
    return fail(PDX_ERR_ARGUMENT, "per-particle diameters belong to the mixed-sigma models");
  if (n <= 0) return fail(PDX_ERR_ARGUMENT, "particle count must be positive");
  float smax = 0.f;
  for (int32_t i = 0; i < n; ++i) {
    if (!(std::isfinite(sigma[i]) && sigma[i] > 0)) return fail(PDX_ERR_ARGUMENT, "diameters must be finite and positive");
    smax = std::max(smax, sigma[i]);
  }
  PDX_HIP(hipDeviceSynchronize());
  int rc = p->sigp.alloc(n);
  if (rc) return rc;
  PDX_HIP(hipMemcpy(p->sigp.p, sigma, sizeof(float) * n, hipMemcpyHostToDevice));
  p->typed_n = n;
  p->typed_rmax = p->spec.cutoff * smax;
  return PDX_OK;
}

extern "C" int pdx_pair_set_types(pdx_pair_t p, int32_t n, int32_t ntypes, const int32_t *typ, const double *sigma,
                                  const double *epsilon, const double *rcut) {
  if (!p || !typ || !sigma || !epsilon || !rcut) return fail(PDX_ERR_ARGUMENT, "null argument");
  if (p->spec.model != PDX_PAIR_MATRIX_LJ && p->spec.model != PDX_PAIR_MATRIX_WCA)
    return fail(PDX_ERR_ARGUMENT, "type tables belong to the pair-matrix models");
  if (n <= 0) return fail(PDX_ERR_ARGUMENT, "particle count must be positive");
  if (ntypes < 1 || ntypes > 64) return fail(PDX_ERR_ARGUMENT, "ntypes must be in 1..64");
  for (int32_t i = 0; i < n; ++i)
    if (typ[i] < 1 || typ[i] > ntypes) return fail(PDX_ERR_ARGUMENT, "type ids must lie in 1..ntypes");
  // rows a*ntypes + b of (sigma^2, epsilon, rc^2, u(rc) with the shift) in FP32, from the
  // column-major matrices (element (a, b) at a + ntypes b)
  std::vector<float4> tab((size_t)ntypes * ntypes);
  double rmax = 0;
  for (int a = 0; a < ntypes; ++a)
    for (int b = 0; b < ntypes; ++b) {
      const size_t k = (size_t)a + (size_t)ntypes * b;
      const double sg = sigma[k], ep = epsilon[k], rc = rcut[k];
      if (!(std::isfinite(sg) && sg > 0 && std::isfinite(ep) && std::isfinite(rc) && rc >= 0))
        return fail(PDX_ERR_ARGUMENT, "pair sigma must be positive, epsilon and rcut finite, rcut >= 0");
      const float s2 = (float)sg * (float)sg, r2 = (float)rc * (float)rc;
      float esh = 0.f;
      if (p->spec.model == PDX_PAIR_MATRIX_LJ && p->spec.shift && rc > 0) {
        // lj_cutoff_energy (PairModels.jl:20-25) per pair
        const float x2 = s2 / r2, x6 = x2 * x2 * x2;
        esh = 4.0f * (float)ep * (x6 * x6 - x6);
      }
      tab[(size_t)a * ntypes + b] = make_float4(s2, (float)ep, r2, esh);
      rmax = std::max(rmax, rc);
    }
  PDX_HIP(hipDeviceSynchronize());
  int rc;
  if ((rc = p->typ.alloc(n)) || (rc = p->tab.alloc(tab.size()))) return rc;
  PDX_HIP(hipMemcpy(p->typ.p, typ, sizeof(int32_t) * n, hipMemcpyHostToDevice));
  PDX_HIP(hipMemcpy(p->tab.p, tab.data(), sizeof(float4) * tab.size(), hipMemcpyHostToDevice));
  p->typed_n = n;
  p->ntypes = ntypes;
  p->typed_rmax = rmax;
  return PDX_OK;
}

extern "C" int pdx_pair_cutoff(pdx_pair_t p, double *rc) {
  if (!p || !rc) return fail(PDX_ERR_ARGUMENT, "null argument");
  double r = 0;
  switch (p->spec.model) {
    case PDX_PAIR_LJ: r = p->spec.cutoff; break;
    case PDX_PAIR_WCA: r = std::sqrt(1.2599211) * p->spec.sigma; break;
    case PDX_PAIR_SOFT: r = p->spec.sigma; break;
    case PDX_PAIR_NONE: r = 0; break;
    default:
      if (p->typed_n == 0) return fail(PDX_ERR_STATE, "typed pair model without its per-particle data");
      r = p->typed_rmax;
  }
  if (p->spec.coulomb) r = std::max(r, p->spec.coulomb_cutoff);
  *rc = r;
  return PDX_OK;
}

extern "C" int pdx_pair_set_bond_exclusion(pdx_pair_t p, int32_t exclude) {
  if (!p) return fail(PDX_ERR_ARGUMENT, "null pair model");
  if (exclude && p->spec.coulomb)
    return fail(PDX_ERR_ARGUMENT, "bonded exclusions of charged pairs go through the Coulomb exceptions");
  p->exclude_bonded = exclude != 0;
  return PDX_OK;
}

extern "C" int pdx_pair_set_accumulation(pdx_pair_t p, int32_t mode) {
  if (!p) return fail(PDX_ERR_ARGUMENT, "null pair model");
  if (mode != PDX_ACC_FP32 && mode != PDX_ACC_FIXED64 && mode != PDX_ACC_FP64)
    return fail(PDX_ERR_ARGUMENT, "accumulation mode must be PDX_ACC_FP32, PDX_ACC_FIXED64 or PDX_ACC_FP64");
  p->accumulate = mode;
  return PDX_OK;
}

extern "C" int pdx_bonds_create(int32_t n, int32_t nbond, const int32_t *bi, const int32_t *bj, double k, double R0,
                                pdx_bonds_t *out) {
  if (!out) return fail(PDX_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  if (n < 1 || nbond < 0) return fail(PDX_ERR_ARGUMENT, "particle and bond counts must be positive");
  if (nbond > 0 && (!bi || !bj)) return fail(PDX_ERR_ARGUMENT, "null bond arrays");
  if (!(std::isfinite(k) && k >= 0)) return fail(PDX_ERR_ARGUMENT, "FENE k must be finite and nonnegative");
  if (!(std::isfinite(R0) && R0 > 0 && std::isfinite(R0 * R0)))
    return fail(PDX_ERR_ARGUMENT, "FENE R0 and its square must be finite and positive");
  std::vector<int32_t> cnt(n + 1, 0);
  for (int b = 0; b < nbond; ++b) {
    if (bi[b] < 0 || bi[b] >= n || bj[b] < 0 || bj[b] >= n || bi[b] == bj[b])
      return fail(PDX_ERR_ARGUMENT, "bond " + std::to_string(b) + " names an invalid particle pair");
    ++cnt[bi[b] + 1];
    ++cnt[bj[b] + 1];
  }
  for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  std::vector<int32_t> part(2 * (size_t)nbond), fill(cnt.begin(), cnt.end() - 1);
  for (int b = 0; b < nbond; ++b) {
    part[fill[bi[b]]++] = bj[b];
    part[fill[bj[b]]++] = bi[b];
  }
  for (int i = 0; i < n; ++i) std::sort(part.begin() + cnt[i], part.begin() + cnt[i + 1]);  // fixed order
  auto *b = new pdx_bonds_s();
  b->N = n;
  b->nbond = nbond;
  b->k = k;
  b->R0 = R0;
  int rc;
  if ((rc = b->offsets.alloc(n + 1)) || (rc = b->partners.alloc(std::max<size_t>(1, part.size()))) ||
      (rc = b->invalid.alloc(1))) {
    delete b;
    return rc;
  }
  const int32_t none = INT32_MAX;
  if (hipMemcpy(b->offsets.p, cnt.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice) != hipSuccess ||
      (nbond > 0 && hipMemcpy(b->partners.p, part.data(), sizeof(int32_t) * part.size(), hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(b->invalid.p, &none, sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
    delete b;
    return fail(PDX_ERR_DEVICE, "bond table upload failed");
  }
  *out = b;
  return PDX_OK;
}

extern "C" int pdx_bonds_destroy(pdx_bonds_t b) {
  delete b;
  return PDX_OK;
}

extern "C" int pdx_pair_set_bonds(pdx_pair_t p, pdx_bonds_t b) {
  if (!p) return fail(PDX_ERR_ARGUMENT, "null pair model");
  p->bonds = b;
  return PDX_OK;
}

extern "C" int pdx_bonds_status(pdx_bonds_t b, int32_t *first_invalid, int32_t reset, pdx_stream_t stream) {
  if (!b || !first_invalid) return fail(PDX_ERR_ARGUMENT, "null argument");
  hipStream_t st = as_stream(stream);
  int32_t v = INT32_MAX;
  PDX_HIP(hipMemcpyAsync(&v, b->invalid.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PDX_HIP(hipStreamSynchronize(st));
  if (reset && v != INT32_MAX) {
    const int32_t none = INT32_MAX;
    PDX_HIP(hipMemcpyAsync(b->invalid.p, &none, sizeof(int32_t), hipMemcpyHostToDevice, st));
    PDX_HIP(hipStreamSynchronize(st));
  }
  *first_invalid = v == INT32_MAX ? -1 : v;
  if (v != INT32_MAX)
    return fail(PDX_ERR_COORDINATE, "FENE bond incident on particle " + std::to_string(v + 1) +
                                        " has nonfinite length or r >= R0=" + std::to_string(b->R0) +
                                        ". Reduce the timestep or correct the configuration; forces are not clamped.");
  return PDX_OK;
}

extern "C" int pdx_pair_set_exceptions(pdx_pair_t p, int32_t n, int32_t nexc, const int32_t *ei, const int32_t *ej,
                                       const double *scale) {
  if (!p) return fail(PDX_ERR_ARGUMENT, "null pair model");
  if (nexc < 0 || n < 1) return fail(PDX_ERR_ARGUMENT, "particle and exception counts must be positive");
  if (nexc > 0 && (!ei || !ej || !scale)) return fail(PDX_ERR_ARGUMENT, "null exception arrays");
  if (nexc > 0 && !p->spec.coulomb) return fail(PDX_ERR_ARGUMENT, "exceptions need the real-space Coulomb term");
  // validate and deduplicate (normalize_exceptions, Spec.jl:200-224); key (min, max)
  std::vector<std::pair<std::pair<int32_t, int32_t>, double>> items;
  items.reserve(nexc);
  for (int k = 0; k < nexc; ++k) {
    const int32_t a = ei[k], b = ej[k];
    const double s = scale[k];
    if (a < 0 || a >= n || b < 0 || b >= n)
      return fail(PDX_ERR_ARGUMENT, "Exception pair (" + std::to_string(a + 1) + ", " + std::to_string(b + 1) +
                                        ") is out of range for N=" + std::to_string(n) + ".");
    if (a == b)
      return fail(PDX_ERR_ARGUMENT, "Exception pair (" + std::to_string(a + 1) + ", " + std::to_string(b + 1) +
                                        ") refers to a single particle.");
    if (!(std::isfinite(s) && s >= 0))
      return fail(PDX_ERR_ARGUMENT, "Exception scale must be finite and nonnegative.");
    items.push_back({{std::min(a, b), std::max(a, b)}, s});
  }
  std::sort(items.begin(), items.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
  std::vector<std::pair<std::pair<int32_t, int32_t>, double>> uniq;
  for (const auto &it : items) {
    if (!uniq.empty() && uniq.back().first == it.first) {
      if (uniq.back().second != it.second)
        return fail(PDX_ERR_ARGUMENT, "Exception pair (" + std::to_string(it.first.first + 1) + ", " +
                                          std::to_string(it.first.second + 1) + ") is listed with conflicting scales.");
      continue;
    }
    uniq.push_back(it);
  }
  std::vector<int32_t> off(n + 1, 0);
  size_t m = 0;
  for (const auto &it : uniq) {
    if (it.second == 1.0) continue;
    ++off[it.first.first + 1];
    ++off[it.first.second + 1];
    m += 2;
  }
  for (int i = 0; i < n; ++i) off[i + 1] += off[i];
  std::vector<int32_t> jj(std::max<size_t>(m, 1)), fill(off.begin(), off.end() - 1);
  std::vector<float> ss(std::max<size_t>(m, 1));
  for (const auto &it : uniq) {  // sorted by (i, j): each particle's list is in id order
    if (it.second == 1.0) continue;
    const int32_t a = it.first.first, b = it.first.second;
    jj[fill[a]] = b; ss[fill[a]++] = (float)it.second;
    jj[fill[b]] = a; ss[fill[b]++] = (float)it.second;
  }
  for (int i = 0; i < n; ++i) {  // partners of i in id order (the b-side entries arrive out of order)
    std::vector<std::pair<int32_t, float>> row;
    for (int t = off[i]; t < off[i + 1]; ++t) row.push_back({jj[t], ss[t]});
    std::sort(row.begin(), row.end());
    for (int t = off[i]; t < off[i + 1]; ++t) { jj[t] = row[t - off[i]].first; ss[t] = row[t - off[i]].second; }
  }
  p->nexc = 0;
  p->exc_n = n;
  if (m == 0) return PDX_OK;
  int rc;
  if ((rc = p->exc_off.alloc(n + 1)) || (rc = p->exc_j.alloc(m)) || (rc = p->exc_s.alloc(m))) return rc;
  PDX_HIP(hipMemcpy(p->exc_off.p, off.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice));
  PDX_HIP(hipMemcpy(p->exc_j.p, jj.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice));
  PDX_HIP(hipMemcpy(p->exc_s.p, ss.data(), sizeof(float) * m, hipMemcpyHostToDevice));
  p->nexc = (int32_t)(m / 2);
  return PDX_OK;
}

extern "C" int pdx_pair_set_lanes(pdx_pair_t p, int32_t lanes) {
  if (!p) return fail(PDX_ERR_ARGUMENT, "null pair model");
  if (lanes > 0 && (lanes > 64 || (lanes & (lanes - 1))))
    return fail(PDX_ERR_ARGUMENT, "lanes per particle must be a power of two in 1..64");
  if (lanes < 0 && !(lanes == -2 || lanes == -4 || lanes == -8 || lanes == -16))
    return fail(PDX_ERR_ARGUMENT, "split layouts use 2, 4, 8 or 16 waves (lanes = -2, -4, -8, -16)");
  if (lanes < 0 && (typed_model(p->spec.model) || p->nexc > 0 || p->bonds || p->accumulate != PDX_ACC_FP32))
    return fail(PDX_ERR_ARGUMENT, "typed models, exceptions, bonds and accumulation modes use the lanes layout (lanes >= 0)");
  p->lanes = lanes;
  return PDX_OK;
}

extern "C" int pdx_pair_set_traversal(pdx_pair_t p, int32_t mode) {
  if (!p) return fail(PDX_ERR_ARGUMENT, "null pair model");
  if (mode != PDX_TRAVERSAL_ROWS && mode != PDX_TRAVERSAL_CELLS)
    return fail(PDX_ERR_ARGUMENT, "traversal must be PDX_TRAVERSAL_ROWS or PDX_TRAVERSAL_CELLS");
  p->traversal = mode;
  return PDX_OK;
}
