// Internal declarations shared by the libpdx translation units: error plumbing,
// device helpers (periodic geometry, wave64 reductions) and the handle structs.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "pdx.h"
#include "pdx_prolate.hpp"

namespace pdx {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define PDX_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return ::pdx::fail(PDX_ERR_DEVICE, std::string(#call " failed: ") +          \
                                             hipGetErrorString(e_));               \
  } while (0)

#define PDX_LAUNCH_CHECK()                                                         \
  do {                                                                             \
    hipError_t e_ = hipGetLastError();                                             \
    if (e_ != hipSuccess)                                                          \
      return ::pdx::fail(PDX_ERR_DEVICE, std::string("kernel launch failed: ") +   \
                                             hipGetErrorString(e_));               \
  } while (0)

inline hipStream_t as_stream(pdx_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Owned device buffer (allocated at create time only).
template <typename T> struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  int alloc(size_t count) {
    release();
    n = count;
    if (count == 0) return PDX_OK;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      n = 0;
      return fail(PDX_ERR_DEVICE, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                      " bytes failed: " + hipGetErrorString(e));
    }
    return PDX_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

constexpr int WAVE = 64;

// ---------------------------------------------------------------- device helpers

// Fractional grid coordinate in [0, K) (_grid_coordinate, src/electrostatics/MeshKernels.jl:9-15).
__device__ __forceinline__ float grid_coordinate(float x, float L, int K) {
  float s = x / L;
  s -= floorf(s);
  float u = s * (float)K;
  if (!isfinite(u)) return 0.f;
  return u >= (float)K ? u - (float)K : u;
}
// A 64-bit fixed-point grid value as FP32 times inv (a power of two): the high and low words
// converted separately and joined by one fused step, the feature definition of the learned
// correction (the same arithmetic in every kernel that forms features).
__device__ __forceinline__ float fixed_to_float(unsigned long long v, float inv) {
  const int hi = (int)((long long)v >> 32);
  const unsigned lo = (unsigned)v;
  return fmaf((float)hi, 4294967296.0f, (float)lo) * inv;
}
// Mesh node index n wrapped into [0, K) for n in [-K, 2K).
__device__ __forceinline__ int wrap_node(int n, int K) { return n < 0 ? n + K : (n >= K ? n - K : n); }

// Minimum image for coordinates in one periodic image of [-L/2, L/2)
// (mic_fast, src/neighbors/NeighborTypes.jl:65-69): branch-free select.
__device__ __forceinline__ float mic_fast(float dx, float halfL, float L) {
  dx -= (dx > halfL) ? L : 0.0f;
  dx += (dx < -halfL) ? L : 0.0f;
  return dx;
}

// Compensated periodic difference (periodic_delta, src/geometry/PeriodicDisplacement.jl:4-14):
// the rounding error of xi - xj (TwoDiff) is carried past the exact (Sterbenz) subtraction
// of L, so short separations across the periodic seam keep their low bits.
__device__ __forceinline__ float periodic_delta(float xi, float xj, float halfL, float L) {
  float d = xi - xj;
  if (fabsf(d) > halfL) {
    float bv = xi - d;
    float res = (xi - (d + bv)) + (bv - xj);
    return (d - copysignf(L, d)) + res;
  }
  return d;
}

// The periodic differences pi - pj[u] of U row entries: plain differences (periodic_delta
// returns exactly these when |d| <= L/2), and periodic_delta itself only when some lane of the
// wave has an entry across the periodic seam; the compensated form then runs with selects.
template <int U>
__device__ __forceinline__ void row_deltas(const float4 &pi, const float4 (&pj)[U], float hx, float hy, float hz,
                                           float Lx, float Ly, float Lz, float (&dx)[U], float (&dy)[U],
                                           float (&dz)[U]) {
  bool wrap = false;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    dx[u] = pi.x - pj[u].x;
    dy[u] = pi.y - pj[u].y;
    dz[u] = pi.z - pj[u].z;
    wrap |= (fabsf(dx[u]) > hx) | (fabsf(dy[u]) > hy) | (fabsf(dz[u]) > hz);
  }
  if (__ballot(wrap) != 0) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      dx[u] = periodic_delta(pi.x, pj[u].x, hx, Lx);
      dy[u] = periodic_delta(pi.y, pj[u].y, hy, Ly);
      dz[u] = periodic_delta(pi.z, pj[u].z, hz, Lz);
    }
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}
__device__ __forceinline__ float wave_sum_width(float v, int width) {
  for (int off = width >> 1; off >= 1; off >>= 1) v += __shfl_xor(v, off, width);
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, WAVE));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, WAVE));
  return v;
}

// Deterministic grid-wide FP64 sums in two launches (no atomics): every block reduces its
// NV values in a fixed order and stores them in partial[block*NV + c]; then one block of
// k_sum_partials adds the partials in block order and writes (or adds into) out[c]. A
// device-scope ticket for a last-block reduction would serialise every block on one address
// in memory on a multi-XCD part. Must be called by every thread of the block (blockDim.x a
// multiple of 64, at most 1024).
template <int NV>
__device__ void block_partials(const double (&v)[NV], double *partial) {
  __shared__ double sh[NV][16];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wave = threadIdx.x / WAVE;
  const int nwave = blockDim.x / WAVE;
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    double s = wave_sum(v[c]);
    if (lane == 0) sh[c][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int w = 0; w < nwave; ++w) s += sh[threadIdx.x][w];
    partial[(size_t)blockIdx.x * NV + threadIdx.x] = s;
  }
}

// One block of 256 threads (the whole block calls it): out[c] (=, or += with add) = sum over
// blocks of partial[b*NV + c],
// thread t taking blocks t, t + 256, ... and a fixed tree over the threads.
template <int NV>
__device__ void sum_partials_block(const double *__restrict__ partial, int nblocks, double *out, int add) {
  __shared__ double sh[NV][4];
  double acc[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) acc[c] = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += 256) {
#pragma unroll
    for (int c = 0; c < NV; ++c) acc[c] += partial[(size_t)b * NV + c];
  }
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    double s = wave_sum(acc[c]);
    if (lane == 0) sh[c][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const double s = ((sh[threadIdx.x][0] + sh[threadIdx.x][1]) + sh[threadIdx.x][2]) + sh[threadIdx.x][3];
    out[threadIdx.x] = add ? out[threadIdx.x] + s : s;
  }
}

// The same sum as a kernel of one 1024-thread block (launch with 1024 threads): thread t adds the
// partials of blocks t, t + 1024, ... in that order, four blocks' loads in flight at a time, then
// a fixed tree over the threads.
constexpr int SUM_THREADS = 1024;
template <int NV>
__global__ __launch_bounds__(SUM_THREADS) void k_sum_partials(const double *__restrict__ partial, int nblocks,
                                                              double *out, int add, const int64_t *state) {
  if (state && !(state[PDX_DN_VALID] != 0 && state[PDX_DN_ERROR] == 0)) return;
  __shared__ double sh[NV][SUM_THREADS / WAVE];
  double acc[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) acc[c] = 0.0;
  int b = threadIdx.x;
  for (; b + 3 * SUM_THREADS < nblocks; b += 4 * SUM_THREADS) {
    double v[4][NV];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < NV; ++c) v[u][c] = partial[(size_t)(b + u * SUM_THREADS) * NV + c];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < NV; ++c) acc[c] += v[u][c];
  }
  for (; b < nblocks; b += SUM_THREADS) {
#pragma unroll
    for (int c = 0; c < NV; ++c) acc[c] += partial[(size_t)b * NV + c];
  }
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    double s = wave_sum(acc[c]);
    if (lane == 0) sh[c][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < SUM_THREADS / WAVE; ++w) s += sh[threadIdx.x][w];
    out[threadIdx.x] = add ? out[threadIdx.x] + s : s;
  }
}

}  // namespace pdx

// ---------------------------------------------------------------- handles

struct pdx_nl_s {
  int32_t N = 0, cap = 0, nx = 1, ny = 1, nz = 1, ncell = 1;
  int32_t interval = 20;
  int32_t row_lanes = 1;  // lanes per particle of the row build
  bool allpairs = false, diagnostics = false;
  int dim = 3;  // 2: no z coordinates (z = 0), box[2] is the list's own z extent of three radii
  double box[3] = {0, 0, 0};
  double cutoff = 0, skin = 0;
  float rl2 = 0;  // (cutoff + skin)^2 in FP32, as the reference's rl2
  pdx::DevBuf<int64_t> state;
  pdx::DevBuf<float> geom;  // device box lengths (3)
  pdx::DevBuf<int32_t> neighbors_flat, counts, cell_offsets, cell_counts, particle_ids_sorted,
      cell_of_particle, cell_rank;
  pdx::DevBuf<float> rref;      // 3*N: x | y | z
  pdx::DevBuf<float4> packed;   // N, storage order (x, y, z, q)
  pdx::DevBuf<float4> sorted;   // N, cell order (x, y, z, id bits)
  pdx::DevBuf<float4> sortedq;  // N, cell order, current (x, y, z, q): the cell traversal's copy
  pdx::DevBuf<float> disp_partial;  // ndisp per-block maxima of the last check
  int32_t ndisp = 0;
  pdx::DevBuf<int32_t> cap_partial;
  pdx::DevBuf<int64_t> stat_partial;  // 2 per block: candidates, accepted
  pdx::DevBuf<int32_t> scan_block;    // per-1024-cell block totals
  pdx::DevBuf<int64_t> candidates;    // N (diagnostics)
  // cell snapshot for a spread that runs beside the rebuild: offsets and ids of the last
  // completed build; snap_state[0] = build count copied, [2] = nonzero while none is valid,
  // [3] = nonzero when the last snapshot spread found a particle beyond its margin (written by
  // that spread, read by its fallback spread and by the tile gather of the same evaluation)
  pdx::DevBuf<int32_t> snap_offsets, snap_ids;
  pdx::DevBuf<int64_t> snap_state;
  int snap_host = 0;  // checks since create / box change (capped at 2): at 2 a snapshot was taken
  const void *packed_src[4] = {nullptr, nullptr, nullptr, nullptr};  // x, y, z, q of the last check
  int check_force = 0;  // the last check was a forced rebuild (pdx_nl_rebuild)
};

struct pdx_pair_s {
  pdx_pair_spec spec{};
  float eps = 0, sig2 = 0, rc2 = 0, eshift = 0, wca_shift = 0;
  float cC = 0, crc2 = 0, csx = 0, ceshift = 0, halfself = 0;
  int npe = 0, npf = 0;
  float pe[24] = {0}, pf[24] = {0};
  double fit_err_e = 0, fit_err_f = 0;
  int lanes = 4;
  int accumulate = PDX_ACC_FP32;      // force / energy / virial accumulator of the row sums
  int traversal = PDX_TRAVERSAL_ROWS;  // rows of the list, or the list's cells directly
  struct pdx_bonds_s *bonds = nullptr;  // FENE bonds folded into the same per-particle sums
  bool exclude_bonded = false;          // bonded pairs get no pair term (exclude_bonded)
  int32_t exc_n = 0, nexc = 0;          // Coulomb exceptions (CSR, both directions)
  pdx::DevBuf<int32_t> exc_off, exc_j;
  pdx::DevBuf<float> exc_s;
  pdx::DevBuf<double> partial;
  // typed models: per-particle diameters (mixed sigma), or type ids and the type-pair table
  // (sig2, eps, rc2, eshift) of the pair matrices; their particle count
  pdx::DevBuf<float> sigp;
  pdx::DevBuf<int32_t> typ;
  pdx::DevBuf<float4> tab;
  int32_t typed_n = 0, ntypes = 0;
  float rfac2 = 0;        // mixed sigma: (rc / sigma_ij)^2
  double typed_rmax = 0;  // the largest pair cutoff of the typed data
};

struct pdx_bonds_s {
  int32_t N = 0, nbond = 0;
  double k = 0, R0 = 0;
  pdx::DevBuf<int32_t> offsets;   // N + 1: bonds of particle i at partners[offsets[i] .. offsets[i+1])
  pdx::DevBuf<int32_t> partners;  // 2 nbond, 0-based (each bond listed from both ends)
  pdx::DevBuf<int32_t> invalid;   // lowest particle id with a bond at r >= R0 (INT32_MAX: none)
};

struct pdx_mesh_s {
  pdx_mesh_spec spec{};
  int32_t N = 0;
  int K1 = 0, K2 = 0, K3 = 0, K1h = 0, P = 0;
  double box[3] = {0, 0, 0};
  double C = 0, alpha = 0, selfc = 0, bgcoef = 0;
  double qtotal = 0, qsq = 0, qabs = 0, ubg = 0;
  float bg = 0;
  double fixed_scale = 0, fixed_inv = 0;  // 2^S and 2^-S of the fixed-point spread
  pdx::DevBuf<unsigned long long> grid64;  // K1*K2*K3 fixed-point charges
  pdx::DevBuf<float2> spec_c;              // K1h*K2*K3 half spectrum
  pdx::DevBuf<float> pot;                  // K1*K2*K3 potential
  pdx::DevBuf<float> G;                    // K1h*K2*K3 influence function (FP32)
  pdx::DevBuf<float2> tw;                  // twiddles exp(-2 pi i m / K) for K1 | K2 | K3
  pdx::DevBuf<double> acc;                 // 8: mesh energy + 6 virial + spare
  pdx::DevBuf<double> partial;
  // P3M: virial coefficients Vc (6 x half spectrum, SoA) of the optimal influence function;
  // ik: the z-field spectrum, a scratch spectrum and the three real field grids
  pdx::DevBuf<float> Vc;
  pdx::DevBuf<float2> spec_ez, spec_tmp;
  pdx::DevBuf<float> ex, ey, ez;
  pdx::DevBuf<double> Utab;                // alias tables (2 nmax + 1) x K per dimension
  bool pot_valid = false;                  // ik: the potential grid of the last solve exists
  int nmode_blocks = 0;
  // window of the spread and gather: wkey = P (B-spline order P) or 16 + P (prolate window of
  // width P, ESP) with its FP32 node polynomials; ESP's split quadrature (xq | wchi, 128 each)
  int wkey = 0;
  pdx::WinCoef wc{};
  double split_shape = 0, window_shape = 0;
  int window_degree = 0;
  pdx::DevBuf<double> espq;
  int T1 = 0, T2 = 0, T3 = 0;  // FFT tile widths, fixed at creation
  int plane_tp = 0;            // plane passes: lines per plane padded to this power of two (0: line passes)
  int spread_threads = 512;    // workgroup size of the pull-form spread (256, 512 or 1024)
  int transform_threads = 1024;  // workgroup size of the plane passes (512 or 1024)
  // per-tile particle lists of the last owner-computes spread (read by the tile gather):
  // tile_entries entries shared by the tiles, tile_n counts; valid for the positions of
  // tiles_nl's packed array while tiles_nl is set
  pdx::DevBuf<float4> tile_list;  // (u1, u2, u3, q)
  pdx::DevBuf<int32_t> tile_ids;  // particle index of each entry
  pdx::DevBuf<int32_t> tile_n;
  size_t tile_entries = 0;
  const void *tiles_nl = nullptr;
  bool tiles_snapshot = false;  // the lists came from a snapshot spread (its word 3 may send the gather per particle)
  int tiles_stile = 0, tiles_count = 0, tiles_cap = 0;
  // the owner-form spread's tile blocks ((STILE + P - 1)^3 fixed-point nodes per tile, folded into
  // the grid by a second pass); empty when they would exceed the size limit (the pull form runs)
  pdx::DevBuf<unsigned long long> tile_blocks;
  size_t tile_block_entries = 0;
};
