// Cell-list binning and slot-major ELL row construction on gfx950.
//
// Pipeline of one predicated rebuild (all kernels are enqueued every time; each returns at
// once unless the device state requests a rebuild, so the sequence can be captured in a
// hipGraph — the "predicated reference path" of src/neighbors/DeviceNeighborCUDA.jl:413-420):
//   k_decide       coverage check: per-block max |dr|^2, packing of (x, y, z, q), zeroing of
//                  the cell histogram
//   k_decide_final max over the blocks vs skin^2/4, age, adaptive interval (one block)
//   k_bin          cell id + histogram; the atomic's return value is the particle's slot
//   k_scan_single  (or k_scan_local / k_scan_blocks / k_scan_finish above 16384 cells)
//                  exclusive scan -> 1-based cell_offsets
//   k_scatter      particle_ids_sorted (1-based)
//   k_sort_cells   ascending id inside each cell (rows and traversal become reproducible),
//                  and the cell-ordered packed positions read by the row kernel
//   k_rows         27-cell stencil, r^2 <= rlist^2, slot-major rows t*N + i, 1-based ids
//   k_commit       per-block max row length and statistics
//   k_commit_final max row length vs cap, state words (one wave)
//   k_copy_refs    reference positions of the successful build
#include "pdx_internal.hpp"

#include <cmath>
#include <cstdlib>

using namespace pdx;

namespace {

constexpr int TPB = 256;
constexpr int SCAN_BLOCK = 1024;  // cells per block of the local scan (256 threads x 4)

__device__ __forceinline__ bool rebuild_requested(const int64_t *s) {
  return s[PDX_DN_REQUEST] != 0 && s[PDX_DN_ERROR] == 0;
}

// ------------------------------------------------------------------ decide
// Velocity-Verlet prefix of a step, fused into the check: v += c_i f, x += dt v wrapped into
// [-L/2, L/2) (_nve_kick_drift3_kernel!, src/simulation/NVEFusion.jl:3-24; the reference fuses
// it with the displacement pass the same way, src/engine/Fusion.jl:1-6). Applied only while the
// list is valid (_device_if_valid!, DeviceNeighborCUDA.jl:157-162).
struct KickDrift {
  float *vx, *vy, *vz;
  const float *fx, *fy, *fz;
  const float *inv_mass;  // per particle, or null: c for every particle
  float c, half_dt, dt;
};

// Coverage check (_device_displacement_kernel! + _device_decide_kernel!,
// DeviceNeighborCUDA.jl:169-236): per-block maximum squared displacement (plus the float4
// packing and the histogram clearing), then one block decides (k_decide_final). With KD the
// step's kick-drift moves x, y, z (in place) first.
template <bool KD>
__global__ __launch_bounds__(TPB) void k_decide(int N, float *__restrict__ x,
                                                 float *__restrict__ y,
                                                 float *__restrict__ z,
                                                 const float *__restrict__ q,
                                                 const float *__restrict__ rref, float4 *packed,
                                                 int32_t *cell_counts, int ncell, int allpairs,
                                                 const float *geom, const int64_t *state,
                                                 float *partial, const int32_t *__restrict__ offsets,
                                                 const int32_t *__restrict__ ids, int32_t *snap_offsets,
                                                 int32_t *snap_ids, int64_t *snap, KickDrift kd) {
  const int64_t err = state[PDX_DN_ERROR];
  if (err != 0) return;  // latched: every block leaves (uniform)
  const int i = blockIdx.x * TPB + threadIdx.x;
  // cell snapshot: the cells of a build completed since the last snapshot, copied before this
  // step's rebuild can overwrite them (k_decide_final records the build); a spread on another
  // stream may then read them beside the rebuild
  if (!allpairs && state[PDX_DN_BUILDS] > 0 && state[PDX_DN_BUILDS] != snap[0]) {
    for (int c = i; c <= ncell; c += gridDim.x * TPB) snap_offsets[c] = offsets[c];
    if (i < N) snap_ids[i] = ids[i];
    if (i == 0) snap[2] = 0;  // valid once this launch completes (snap[0] is written by k_decide_final)
  }
  const float L[3] = {geom[0], geom[1], geom[2]};
  const bool valid = state[PDX_DN_VALID] != 0;
  float value = 0.0f;
  if (i < N) {
    float c[3] = {x[i], y[i], z ? z[i] : 0.0f};  // 2-D lists: no z array, every z is 0
    if (KD && valid) {
      const float ci = kd.inv_mass ? kd.half_dt * kd.inv_mass[i] : kd.c;
      float *v[3] = {kd.vx, kd.vy, kd.vz};
      const float *f[3] = {kd.fx, kd.fy, kd.fz};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (d == 2 && !z) break;
        const float u = fmaf(ci, f[d][i], v[d][i]);
        v[d][i] = u;
        float p = fmaf(kd.dt, u, c[d]);
        p -= (p >= 0.5f * L[d]) ? L[d] : 0.0f;
        p += (p < -0.5f * L[d]) ? L[d] : 0.0f;
        c[d] = p;
      }
      x[i] = c[0];
      y[i] = c[1];
      if (z) z[i] = c[2];
    }
    // z, then y, then x: the nested muladd order of the reference (threshold ties included)
    for (int d = 2; d >= 0; --d) {
      const float h = 0.5f * L[d];
      if (!(isfinite(c[d]) && -h <= c[d] && c[d] <= h)) {
        value = INFINITY;
      } else if (valid && value != INFINITY) {
        float dx = mic_fast(c[d] - rref[(size_t)d * N + i], h, L[d]);
        value = fmaf(dx, dx, value);
      }
    }
    packed[i] = make_float4(c[0], c[1], c[2], q ? q[i] : 0.0f);
  }
  // histogram storage is scratch of the next rebuild: clear it here (one launch fewer)
  if (!allpairs)
    for (int c = i; c < ncell; c += gridDim.x * TPB) cell_counts[c] = 0;

  __shared__ float smax[TPB / WAVE];
  float m = wave_max(value);
  if ((threadIdx.x & (WAVE - 1)) == 0) smax[threadIdx.x / WAVE] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = smax[0];
    for (int w = 1; w < TPB / WAVE; ++w) b = fmaxf(b, smax[w]);
    // -1: not measured (no valid list to measure against); a snapshot spread reading these
    // maxima then cannot bound the motion since the snapshot's build
    partial[blockIdx.x] = (!valid && b != INFINITY) ? -1.0f : b;
  }
}

// The decision from the blocks' maxima (one block).
__global__ __launch_bounds__(TPB) void k_decide_final(int nblocks, const float *__restrict__ partial, int64_t *state,
                                                       float threshold, int force, int64_t *snap, int allpairs) {
  if (state[PDX_DN_ERROR] != 0) return;
  __shared__ float smax[TPB / WAVE];
  float v = 0.0f;
  for (int b = threadIdx.x; b < nblocks; b += TPB) v = fmaxf(v, partial[b]);
  v = wave_max(v);
  if ((threadIdx.x & (WAVE - 1)) == 0) smax[threadIdx.x / WAVE] = v;
  __syncthreads();
  if (threadIdx.x != 0) return;
  if (!allpairs && state[PDX_DN_BUILDS] > 0 && state[PDX_DN_BUILDS] != snap[0])  // k_decide copied it
    snap[0] = state[PDX_DN_BUILDS];
  float largest = smax[0];
  for (int w = 1; w < TPB / WAVE; ++w) largest = fmaxf(largest, smax[w]);
  const bool valid = state[PDX_DN_VALID] != 0;
  state[PDX_DN_CHECKS] += 1;
  state[PDX_DN_REQUEST] = 0;
  const bool diag = state[PDX_DN_DIAGNOSTICS] != 0;
  if (diag) state[PDX_DN_REASON] = 0;
  if (!isfinite(largest)) {
    state[PDX_DN_ERROR] = PDX_DN_COORDINATE_ERROR;
    state[PDX_DN_VALID] = 0;
    return;
  }
  if (force || !valid) {
    state[PDX_DN_REQUEST] = 1;
    state[PDX_DN_VALID] = 0;
    if (diag) state[PDX_DN_REASON] = PDX_DN_REASON_INVALID;
    return;
  }
  const int64_t age = state[PDX_DN_STEP] - state[PDX_DN_LAST_BUILD];
  const bool disp = largest > threshold;
  const bool old = age >= state[PDX_DN_INTERVAL];
  if (disp || old) {
    state[PDX_DN_REQUEST] = 1;
    state[PDX_DN_VALID] = 0;
    if (diag) state[PDX_DN_REASON] = (disp ? PDX_DN_REASON_DISPLACEMENT : 0) | (old ? PDX_DN_REASON_AGE : 0);
    // adaptive interval: x0.9 on displacement, x1.1 when the list aged out early, in [5, 100]
    if (disp) {
      float t = rintf(0.9f * (float)state[PDX_DN_INTERVAL]);
      state[PDX_DN_INTERVAL] = t < 5.0f ? 5 : (int64_t)t;
    } else if (largest < 0.1f * threshold) {
      float t = rintf(1.1f * (float)state[PDX_DN_INTERVAL]);
      state[PDX_DN_INTERVAL] = t > 100.0f ? 100 : (int64_t)t;
    }
  }
}

// ------------------------------------------------------------------ binning
// Cell id from the current box, fractional binning floor(x n/L) (_kernel_cell_ids3g!,
// DeviceNeighborCUDA.jl:464-481) + histogram (_kernel_cell_histogram!, DenseNeighborCUDA.jl:50).
__global__ __launch_bounds__(TPB) void k_bin(int N, const float4 *__restrict__ packed,
                                              const float *geom, int nx, int ny, int nz,
                                              int32_t *cell_of_particle, int32_t *cell_counts,
                                              int32_t *cell_rank, const int64_t *state) {
  if (!rebuild_requested(state)) return;
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= N) return;
  const float4 p = packed[i];
  const float Lx = geom[0], Ly = geom[1], Lz = geom[2];
  float xx = p.x + 0.5f * Lx; xx -= floorf(xx / Lx) * Lx;
  float yy = p.y + 0.5f * Ly; yy -= floorf(yy / Ly) * Ly;
  float zz = p.z + 0.5f * Lz; zz -= floorf(zz / Lz) * Lz;
  int cx = (int)floorf(xx * ((float)nx / Lx)); cx = min(max(cx, 0), nx - 1);
  int cy = (int)floorf(yy * ((float)ny / Ly)); cy = min(max(cy, 0), ny - 1);
  int cz = (int)floorf(zz * ((float)nz / Lz)); cz = min(max(cz, 0), nz - 1);
  const int c = (cz * ny + cy) * nx + cx;
  cell_of_particle[i] = c;
  cell_rank[i] = atomicAdd(&cell_counts[c], 1);
}

// Reference positions of a successful build (skipped when the commit latched an error).
__global__ __launch_bounds__(TPB) void k_copy_refs(int N, const float4 *__restrict__ packed,
                                                    float *rref, const int64_t *state) {
  if (!rebuild_requested(state)) return;
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= N) return;
  const float4 p = packed[i];
  rref[i] = p.x;
  rref[(size_t)N + i] = p.y;
  rref[(size_t)2 * N + i] = p.z;
}

// Cell offsets in one launch for grids of up to 16384 cells: 1024 threads, a contiguous run of
// at most 16 cells each, wave scans of the run totals; 1-based, last entry N+1
// (_device_offsets_kernel!, DeviceNeighborCUDA.jl:280-289).
constexpr int SCAN_SINGLE_MAX = 16384;
__global__ __launch_bounds__(1024) void k_scan_single(int ncell, int N, const int32_t *__restrict__ counts,
                                                       int32_t *offsets, const int64_t *state) {
  if (!rebuild_requested(state)) return;
  __shared__ int32_t wtot[16];
  const int per = (ncell + 1023) / 1024;
  const int c0 = min(threadIdx.x * per, ncell), c1 = min(c0 + per, ncell);
  int32_t run = 0;
  for (int c = c0; c < c1; ++c) run += counts[c];
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  int32_t v = run;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const int32_t u = __shfl_up(v, o, WAVE);
    if (lane >= o) v += u;
  }
  if (lane == WAVE - 1) wtot[w] = v;
  __syncthreads();
  int32_t base = v - run + 1;
  for (int k = 0; k < w; ++k) base += wtot[k];
  for (int c = c0; c < c1; ++c) {
    offsets[c] = base;
    base += counts[c];
  }
  if (threadIdx.x == 0) offsets[ncell] = N + 1;
}

// Exclusive scan of cell_counts, SCAN_BLOCK cells per block (block-local part).
__global__ __launch_bounds__(TPB) void k_scan_local(int ncell, const int32_t *counts,
                                                     int32_t *offsets, int32_t *block_totals,
                                                     const int64_t *state) {
  if (!rebuild_requested(state)) return;
  __shared__ int32_t sh[TPB];
  const int base = blockIdx.x * SCAN_BLOCK + threadIdx.x * 4;
  int32_t v[4];
  int32_t run = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = (base + k < ncell) ? counts[base + k] : 0;
    run += v[k];
  }
  sh[threadIdx.x] = run;
  __syncthreads();
  for (int off = 1; off < TPB; off <<= 1) {
    int32_t a = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += a;
    __syncthreads();
  }
  int32_t excl = sh[threadIdx.x] - run;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < ncell) offsets[base + k] = excl;
    excl += v[k];
  }
  if (threadIdx.x == TPB - 1) block_totals[blockIdx.x] = sh[TPB - 1];
}

// Exclusive scan of the block totals (one block, sequential over chunks of 1024).
__global__ __launch_bounds__(1024) void k_scan_blocks(int nblocks, int32_t *block_totals,
                                                       const int64_t *state) {
  if (!rebuild_requested(state)) return;
  __shared__ int32_t sh[1024];
  int32_t carry = 0;
  for (int base = 0; base < nblocks; base += 1024) {
    const int k = base + threadIdx.x;
    const int32_t v = k < nblocks ? block_totals[k] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      int32_t a = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += a;
      __syncthreads();
    }
    if (k < nblocks) block_totals[k] = carry + sh[threadIdx.x] - v;
    const int32_t tot = sh[1023];
    __syncthreads();
    carry += tot;
  }
}

// 1-based offsets with last entry N+1 (_device_offsets_kernel!, DeviceNeighborCUDA.jl:280-289).
__global__ __launch_bounds__(TPB) void k_scan_finish(int ncell, int N, int32_t *offsets,
                                                      const int32_t *block_totals,
                                                      const int64_t *state) {
  if (!rebuild_requested(state)) return;
  const int c = blockIdx.x * TPB + threadIdx.x;
  if (c < ncell) offsets[c] += block_totals[c / SCAN_BLOCK] + 1;
  if (c == ncell) offsets[ncell] = N + 1;
}

__global__ __launch_bounds__(TPB) void k_scatter(int N, const int32_t *cell_of_particle,
                                                  const int32_t *cell_rank,
                                                  const int32_t *offsets, int32_t *ids_sorted,
                                                  const int64_t *state) {
  if (!rebuild_requested(state)) return;
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= N) return;
  const int c = cell_of_particle[i];
  ids_sorted[offsets[c] - 1 + cell_rank[i]] = i + 1;
}

// Stable order inside each cell (ascending id) and the cell-ordered packed copy (x, y, z, id)
// the row kernel streams through. One wave per cell: the cell's ids are staged in LDS and each
// lands at its rank (ids are distinct); cells beyond the LDS stage are sorted by one lane.
constexpr int SORT_STAGE = 2048;
__global__ __launch_bounds__(WAVE) void k_sort_cells(int ncell, const int32_t *offsets, int32_t *ids_sorted,
                                                     const float4 *__restrict__ packed, float4 *sorted,
                                                     const int64_t *state) {
  if (!rebuild_requested(state)) return;
  __shared__ int32_t sh[SORT_STAGE];
  // grid-stride over the cells (a bounded grid: on a check without rebuild the launch is cheap
  // however many cells there are)
  for (int c = blockIdx.x; c < ncell; c += gridDim.x) {
    const int s = offsets[c] - 1, e = offsets[c + 1] - 1, n = e - s;
    __syncthreads();  // the previous cell's ranks are read
    if (n <= SORT_STAGE) {
      for (int k = threadIdx.x; k < n; k += WAVE) sh[k] = ids_sorted[s + k];
      __syncthreads();
      for (int k = threadIdx.x; k < n; k += WAVE) {
        const int32_t key = sh[k];
        int rank = 0;
        for (int m = 0; m < n; ++m) rank += sh[m] < key;
        ids_sorted[s + rank] = key;
        const float4 p = packed[key - 1];
        sorted[s + rank] = make_float4(p.x, p.y, p.z, __int_as_float(key - 1));
      }
      continue;
    }
    if (threadIdx.x == 0) {
      for (int a = s + 1; a < e; ++a) {
        const int32_t key = ids_sorted[a];
        int b = a - 1;
        while (b >= s && ids_sorted[b] > key) {
          ids_sorted[b + 1] = ids_sorted[b];
          --b;
        }
        ids_sorted[b + 1] = key;
      }
    }
    __syncthreads();
    for (int k = s + threadIdx.x; k < e; k += WAVE) {
      const int i = ids_sorted[k] - 1;
      const float4 p = packed[i];
      sorted[k] = make_float4(p.x, p.y, p.z, __int_as_float(i));
    }
  }
}

// ------------------------------------------------------------------ rows
// G lanes per particle, particles in cell order (the groups of a wave share their home cell,
// so candidate loads are broadcast). The group walks the 27-cell stencil G candidates at a
// time; accepted neighbours are placed with a wave64 ballot + popcount in candidate order, so
// every row is identical whatever G is. Distance test as the reference: mic_fast differences,
// r^2 = muladd(dx,dx,muladd(dy,dy,dz^2)) <= rl^2 (_kernel_neighbors3!, DenseNeighborCUDA.jl:153-206;
// cooperative placement as _cooperative_rows3_kernel!, src/engine/Builder.jl:39-129). Slots are
// written slot-major at t*N + i with 1-based ids; the count keeps growing past cap so the
// commit can report the required capacity.
// STAGE (opt-in): the workgroup's rows are kept in LDS as [slot][particle] (TPB/G particles, cap
// slots, at most 64 KiB) and written out slot by slot after the search, so a row slot of
// consecutive particles is one contiguous store; written straight from the search, the slots of
// a wave's particles advance at different rates and the 4-byte stores land in partially filled
// lines (measured: 8.6x the row bytes at 486 400 ions, 2.0x staged). Same rows either way.
constexpr size_t ROW_STAGE_BYTES = 65536;
template <int G, bool STAGE>
__global__ __launch_bounds__(TPB) void k_rows(int N, int cap, int nx, int ny, int nz,
                                               const float4 *__restrict__ sorted,
                                               const int32_t *__restrict__ offsets,
                                               const int32_t *__restrict__ cell_of_particle,
                                               const float *geom, float rl2, int32_t *nbr,
                                               int32_t *counts, int64_t *candidates,
                                               const int64_t *state) {
  constexpr int PB = TPB / G;  // particles per workgroup
  extern __shared__ int32_t srow[];  // STAGE: [cap][PB]
  __shared__ int sfound[PB], sid[PB];
  __shared__ int smax;
  if (!rebuild_requested(state)) return;  // uniform over the grid
  const int gid = blockIdx.x * TPB + threadIdx.x;
  const int t = gid / G;
  const int lg = gid % G;
  const int pb = threadIdx.x / G;
  const int wl = threadIdx.x & (WAVE - 1);
  const unsigned long long gmask = (G == WAVE) ? ~0ull : (((1ull << G) - 1ull) << (wl - lg));
  const unsigned long long below = (1ull << wl) - 1ull;
  const bool active = t < N;  // whole groups are active or not (G divides the block)
  const float Lx = geom[0], Ly = geom[1], Lz = geom[2];
  const float hx = 0.5f * Lx, hy = 0.5f * Ly, hz = 0.5f * Lz;
  int found = 0, i = -1;
  int64_t tested = 0;
  if (active) {
    const float4 pi = sorted[t];
    i = __float_as_int(pi.w);
    const int c0 = cell_of_particle[i];
    const int cx = c0 % nx, cy = (c0 / nx) % ny, cz = c0 / (nx * ny);
    for (int oz = -1; oz <= 1; ++oz) {
      int cz2 = cz + oz; cz2 -= (cz2 >= nz) ? nz : 0; cz2 += (cz2 < 0) ? nz : 0;
      for (int oy = -1; oy <= 1; ++oy) {
        int cy2 = cy + oy; cy2 -= (cy2 >= ny) ? ny : 0; cy2 += (cy2 < 0) ? ny : 0;
        for (int ox = -1; ox <= 1; ++ox) {
          int cx2 = cx + ox; cx2 -= (cx2 >= nx) ? nx : 0; cx2 += (cx2 < 0) ? nx : 0;
          const int c = (cz2 * ny + cy2) * nx + cx2;
          const int s = offsets[c] - 1, e = offsets[c + 1] - 1;
          for (int k0 = s; k0 < e; k0 += G) {
            const int k = k0 + lg;
            bool acc = false, cand = false;
            int j = 0;
            if (k < e) {
              const float4 pj = sorted[k];
              j = __float_as_int(pj.w);
              cand = j != i;
              const float dx = mic_fast(pj.x - pi.x, hx, Lx);
              const float dy = mic_fast(pj.y - pi.y, hy, Ly);
              const float dz = mic_fast(pj.z - pi.z, hz, Lz);
              const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
              acc = cand && (r2 <= rl2);
            }
            if (G == 1) {
              tested += cand;
              if (acc) {
                if (found < cap) {
                  if (STAGE) srow[found * PB + pb] = j + 1;
                  else nbr[(size_t)found * N + i] = j + 1;
                }
                ++found;
              }
            } else {
              const unsigned long long am = __ballot(acc) & gmask;
              tested += __popcll(__ballot(cand) & gmask);
              if (acc) {
                const int slot = found + __popcll(am & below);
                if (slot < cap) {
                  if (STAGE) srow[slot * PB + pb] = j + 1;
                  else nbr[(size_t)slot * N + i] = j + 1;
                }
              }
              found += __popcll(am);
            }
          }
        }
      }
    }
  }
  if (STAGE) {
    if (lg == 0) {
      sfound[pb] = active ? min(found, cap) : 0;
      sid[pb] = i;
    }
    if (threadIdx.x == 0) smax = 0;
    __syncthreads();
    if (lg == 0 && active) atomicMax(&smax, min(found, cap));
    __syncthreads();
    const int ns = smax;
    for (int e = threadIdx.x; e < ns * PB; e += TPB) {
      const int sl = e / PB, q = e - sl * PB;
      if (sl < sfound[q]) nbr[(size_t)sl * N + sid[q]] = srow[sl * PB + q];
    }
  }
  if (active && lg == 0) {
    counts[i] = found;
    if (candidates) candidates[i] = tested;
  }
}

// Rows from LDS-staged candidates: one workgroup per cell. Wave 0 lists the cell's 27 stencil cells
// in k_rows' order (z, y, x offsets -1..1, each cell's particles in sorted order) with the
// inclusive scan of their sizes; the flattened candidate list is staged into LDS RC_CHUNK float4s
// at a time, and the cell's particles, G lanes each (TPB/G per pass), test them from LDS. The
// ballot + popcount placement over the flattened order writes exactly k_rows' rows (the
// accepted candidates in stencil order), whatever G or the chunking.
constexpr int RC_CHUNK = 2048;
template <int G>
__global__ __launch_bounds__(TPB) void k_rows_cells(int N, int cap, int ncell, int nx, int ny, int nz,
                                                    const float4 *__restrict__ sorted,
                                                    const int32_t *__restrict__ offsets, const float *geom,
                                                    float rl2, int32_t *nbr, int32_t *counts,
                                                    int64_t *candidates, const int64_t *state) {
  constexpr int PB = TPB / G;  // particles per pass
  __shared__ float4 cand[RC_CHUNK];
  __shared__ int cstart[27], cpre[28];
  if (!rebuild_requested(state)) return;  // uniform over the grid
  for (int c = blockIdx.x; c < ncell; c += gridDim.x) {  // grid-stride over the cells
    const int p0 = offsets[c] - 1, np = offsets[c + 1] - offsets[c];
    if (np <= 0) continue;  // uniform over the workgroup
    const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
    __syncthreads();  // the previous cell's stencil and candidates are no longer read
    if (threadIdx.x < WAVE) {
      int cnt = 0;
      if (threadIdx.x < 27) {
        const int o = threadIdx.x;
        int cz2 = cz + o / 9 - 1; cz2 -= (cz2 >= nz) ? nz : 0; cz2 += (cz2 < 0) ? nz : 0;
        int cy2 = cy + (o / 3) % 3 - 1; cy2 -= (cy2 >= ny) ? ny : 0; cy2 += (cy2 < 0) ? ny : 0;
        int cx2 = cx + o % 3 - 1; cx2 -= (cx2 >= nx) ? nx : 0; cx2 += (cx2 < 0) ? nx : 0;
        const int cc = (cz2 * ny + cy2) * nx + cx2;
        cstart[o] = offsets[cc] - 1;
        cnt = offsets[cc + 1] - offsets[cc];
      }
#pragma unroll
      for (int d = 1; d < WAVE; d <<= 1) {
        const int u = __shfl_up(cnt, d, WAVE);
        if ((int)threadIdx.x >= d) cnt += u;
      }
      if (threadIdx.x < 27) cpre[threadIdx.x + 1] = cnt;
      if (threadIdx.x == 0) cpre[0] = 0;
    }
    __syncthreads();
    const int total = cpre[27];
    const float Lx = geom[0], Ly = geom[1], Lz = geom[2];
    const float hx = 0.5f * Lx, hy = 0.5f * Ly, hz = 0.5f * Lz;
    const int lg = threadIdx.x % G, pl = threadIdx.x / G;
    const int wl = threadIdx.x & (WAVE - 1);
    const unsigned long long gmask = (G == WAVE) ? ~0ull : (((1ull << G) - 1ull) << (wl - lg));
    const unsigned long long below = (1ull << wl) - 1ull;
    for (int pb = 0; pb < np; pb += PB) {
      const bool active = pb + pl < np;
      float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
      int i = -1;
      if (active) {
        pi = sorted[p0 + pb + pl];
        i = __float_as_int(pi.w);
      }
      int found = 0;
      int64_t tested = 0;
      for (int ch = 0; ch < total; ch += RC_CHUNK) {
        const int nch = min(RC_CHUNK, total - ch);
        __syncthreads();  // the previous chunk (or pass) is no longer read
        for (int f = threadIdx.x; f < nch; f += TPB) {
          const int g = ch + f;
          int cc = 0;
#pragma unroll
          for (int h = 16; h >= 1; h >>= 1)
            if (cc + h < 27 && cpre[cc + h] <= g) cc += h;
          cand[f] = sorted[cstart[cc] + (g - cpre[cc])];
        }
        __syncthreads();
        if (active) {
          for (int k0 = 0; k0 < nch; k0 += G) {
            const int k = k0 + lg;
            bool acc = false, cnd = false;
            int j = 0;
            if (k < nch) {
              const float4 pj = cand[k];
              j = __float_as_int(pj.w);
              cnd = j != i;
              const float dx = mic_fast(pj.x - pi.x, hx, Lx);
              const float dy = mic_fast(pj.y - pi.y, hy, Ly);
              const float dz = mic_fast(pj.z - pi.z, hz, Lz);
              const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
              acc = cnd && (r2 <= rl2);
            }
            if (G == 1) {
              tested += cnd;
              if (acc) {
                if (found < cap) nbr[(size_t)found * N + i] = j + 1;
                ++found;
              }
            } else {
              const unsigned long long am = __ballot(acc) & gmask;
              tested += __popcll(__ballot(cnd) & gmask);
              if (acc) {
                const int slot = found + __popcll(am & below);
                if (slot < cap) nbr[(size_t)slot * N + i] = j + 1;
              }
              found += __popcll(am);
            }
          }
        }
      }
      if (active && lg == 0) {
        counts[i] = found;
        if (candidates) candidates[i] = tested;
      }
    }
  }
}

// All-pairs rows for small boxes, one image (_kernel_neighbors_allpairs!, DenseNeighborCUDA.jl:128-151).
__global__ __launch_bounds__(TPB) void k_rows_allpairs(int N, int cap,
                                                        const float4 *__restrict__ packed,
                                                        const float *geom, float rl2,
                                                        int32_t *nbr, int32_t *counts,
                                                        int64_t *candidates,
                                                        const int64_t *state) {
  if (!rebuild_requested(state)) return;
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= N) return;
  const float Lx = geom[0], Ly = geom[1], Lz = geom[2];
  const float4 pi = packed[i];
  int found = 0;
  for (int j = 0; j < N; ++j) {
    if (j == i) continue;
    const float4 pj = packed[j];
    // dimension order of the reference's loop: r2 = muladd(dd, dd, r2) for d = 1..D
    float r2 = 0.0f;
    float dd = mic_fast(pj.x - pi.x, 0.5f * Lx, Lx); r2 = fmaf(dd, dd, r2);
    dd = mic_fast(pj.y - pi.y, 0.5f * Ly, Ly); r2 = fmaf(dd, dd, r2);
    dd = mic_fast(pj.z - pi.z, 0.5f * Lz, Lz); r2 = fmaf(dd, dd, r2);
    if (r2 <= rl2) {
      if (found < cap) nbr[(size_t)found * N + i] = j + 1;
      ++found;
    }
  }
  counts[i] = found;
  if (candidates) candidates[i] = N - 1;
}

// ------------------------------------------------------------------ commit
// Largest row vs cap, build counters (_device_capacity_partial_kernel! + _device_commit_kernel!,
// DeviceNeighborCUDA.jl:291-388), one launch with a last-block commit.
__global__ __launch_bounds__(TPB) void k_commit(int N, const int32_t *counts, const int64_t *candidates,
                                                 int32_t *partial, int64_t *stat_partial, const int64_t *state) {
  if (!rebuild_requested(state)) return;
  const int i = blockIdx.x * TPB + threadIdx.x;
  const int v = i < N ? counts[i] : 0;
  const int64_t cand = (candidates && i < N) ? candidates[i] : 0;
  __shared__ int smax[TPB / WAVE];
  __shared__ long long ssum[2][TPB / WAVE];
  int m = wave_max(v);
  long long sc = cand, sa = v;
  for (int off = 32; off >= 1; off >>= 1) {
    sc += __shfl_xor(sc, off, WAVE);
    sa += __shfl_xor(sa, off, WAVE);
  }
  const int w = threadIdx.x / WAVE;
  if ((threadIdx.x & (WAVE - 1)) == 0) { smax[w] = m; ssum[0][w] = sc; ssum[1][w] = sa; }
  __syncthreads();
  if (threadIdx.x == 0) {
    int b = smax[0]; long long c0 = ssum[0][0], a0 = ssum[1][0];
    for (int k = 1; k < TPB / WAVE; ++k) { b = max(b, smax[k]); c0 += ssum[0][k]; a0 += ssum[1][k]; }
    partial[blockIdx.x] = b;
    stat_partial[2 * blockIdx.x] = c0;
    stat_partial[2 * blockIdx.x + 1] = a0;
  }
}

// Capacity commit from the blocks' partials (one wave64; integer sums, so any order is exact).
__global__ __launch_bounds__(WAVE) void k_commit_final(int nblocks, int cap, const int32_t *__restrict__ partial,
                                                       const int64_t *__restrict__ stat_partial, int64_t *state) {
  if (!rebuild_requested(state)) return;
  int largest = 0;
  long long ctot = 0, atot = 0;
  for (int b = threadIdx.x; b < nblocks; b += WAVE) {
    largest = max(largest, partial[b]);
    ctot += stat_partial[2 * b];
    atot += stat_partial[2 * b + 1];
  }
  largest = wave_max(largest);
  for (int off = 32; off >= 1; off >>= 1) {
    ctot += __shfl_xor(ctot, off, WAVE);
    atot += __shfl_xor(atot, off, WAVE);
  }
  if (threadIdx.x != 0) return;
  state[PDX_DN_REQUIRED] = largest;
  const bool diag = state[PDX_DN_DIAGNOSTICS] != 0;
  if (diag) {
    state[PDX_DN_ATTEMPTS] += 1;
    state[PDX_DN_LAST_CANDIDATES] = ctot;
    state[PDX_DN_TOTAL_CANDIDATES] += ctot;
    state[PDX_DN_LAST_ACCEPTED] = atot;
    state[PDX_DN_TOTAL_ACCEPTED] += atot;
    if (largest > state[PDX_DN_MAX_ROW]) state[PDX_DN_MAX_ROW] = largest;
  }
  if (largest > cap) {
    state[PDX_DN_ERROR] = PDX_DN_CAPACITY_ERROR;
    state[PDX_DN_VALID] = 0;
    return;
  }
  state[PDX_DN_VALID] = 1;
  state[PDX_DN_LAST_BUILD] = state[PDX_DN_STEP];
  state[PDX_DN_BUILDS] += 1;
  if (diag) {
    const int64_t reason = state[PDX_DN_REASON];
    if (reason & PDX_DN_REASON_INVALID) state[PDX_DN_INVALID_BUILDS] += 1;
    else if (reason & PDX_DN_REASON_DISPLACEMENT) state[PDX_DN_DISPLACEMENT_BUILDS] += 1;
    else state[PDX_DN_AGE_BUILDS] += 1;
    if ((reason & (PDX_DN_REASON_DISPLACEMENT | PDX_DN_REASON_AGE)) ==
        (PDX_DN_REASON_DISPLACEMENT | PDX_DN_REASON_AGE))
      state[PDX_DN_OVERLAP_BUILDS] += 1;
  }
}


__global__ void k_advance(int64_t *state) {
  if (state[PDX_DN_VALID] != 0 && state[PDX_DN_ERROR] == 0) state[PDX_DN_STEP] += 1;
}

// Closing kick of a velocity-Verlet step, v += c_i f (_nve_kick3_kernel!, SimulationCore.jl:
// 581-591), while the list is valid; the first thread also advances the step counter (the
// reference folds advance_device_neighbor_step! into the kick, Run.jl:281-285).
__global__ __launch_bounds__(TPB) void k_kick(int N, KickDrift kd, int dim, int64_t *state) {
  if (!(state[PDX_DN_VALID] != 0 && state[PDX_DN_ERROR] == 0)) return;
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i < N) {
    const float ci = kd.inv_mass ? kd.half_dt * kd.inv_mass[i] : kd.c;
    kd.vx[i] = fmaf(ci, kd.fx[i], kd.vx[i]);
    kd.vy[i] = fmaf(ci, kd.fy[i], kd.vy[i]);
    if (dim == 3) kd.vz[i] = fmaf(ci, kd.fz[i], kd.vz[i]);
  }
  if (i == 0) state[PDX_DN_STEP] += 1;
}

__global__ void k_rearm(int64_t *state) {
  state[PDX_DN_ERROR] = 0;
  state[PDX_DN_REQUEST] = 1;
}

__global__ void k_invalidate(int64_t *state) { state[PDX_DN_VALID] = 0; }

// Spatial order of the last build (the permutation of OrderingWorkspace, src/engine/Ordering.jl:
// 19-79): slots in cell order (cells z, y, x), inside a cell by position (x, then y, then z, then
// the id for coincident particles), so the order depends on the configuration and not on the
// particles' labels. One wave per cell ranks its particles (cells hold a few dozen); perm[slot] =
// 0-based particle id. Cells beyond the LDS stage rank from global memory.
__global__ __launch_bounds__(WAVE) void k_spatial_order(int ncell, const int32_t *__restrict__ offsets,
                                                        const int32_t *__restrict__ ids,
                                                        const float4 *__restrict__ packed, int32_t *perm) {
  __shared__ float4 sh[SORT_STAGE / 2];
  const int c = blockIdx.x;
  const int s = offsets[c] - 1, e = offsets[c + 1] - 1, n = e - s;
  auto key = [&](int k) {
    const int i = ids[s + k] - 1;
    const float4 p = packed[i];
    return make_float4(p.x, p.y, p.z, __int_as_float(i));
  };
  const bool staged = n <= SORT_STAGE / 2;
  if (staged) {
    for (int k = threadIdx.x; k < n; k += WAVE) sh[k] = key(k);
    __syncthreads();
  }
  auto before = [](const float4 a, const float4 b) {
    if (a.x != b.x) return a.x < b.x;
    if (a.y != b.y) return a.y < b.y;
    if (a.z != b.z) return a.z < b.z;
    return __float_as_int(a.w) < __float_as_int(b.w);
  };
  for (int k = threadIdx.x; k < n; k += WAVE) {
    const float4 me = staged ? sh[k] : key(k);
    int rank = 0;
    for (int m = 0; m < n; ++m) rank += before(staged ? sh[m] : key(m), me);
    perm[s + rank] = __float_as_int(me.w);
  }
}

// Cell snapshot: offsets and ids of a build that completed since the last snapshot (the build
// counter moved), then the recorded count. Run after every step's evaluation, so that a spread
// running beside the next step's rebuild reads cells no rebuild is writing.
__global__ __launch_bounds__(TPB) void k_snap_copy(int N, int ncell, const int32_t *__restrict__ offsets,
                                                   const int32_t *__restrict__ ids, int32_t *so, int32_t *sid,
                                                   const int64_t *state, const int64_t *snap) {
  if (state[PDX_DN_ERROR] != 0 || state[PDX_DN_BUILDS] == snap[0]) return;
  const int k = blockIdx.x * TPB + threadIdx.x;
  if (k <= ncell) so[k] = offsets[k];
  if (k < N) sid[k] = ids[k];
}
__global__ void k_snap_mark(const int64_t *state, int64_t *snap) {
  if (state[PDX_DN_ERROR] != 0 || state[PDX_DN_BUILDS] == snap[0]) return;
  snap[0] = state[PDX_DN_BUILDS];
  snap[2] = 0;
}

inline int blocks_for(int64_t n, int tpb) { return (int)((n + tpb - 1) / tpb); }

int grid_for_box(const double box[3], double rl, int32_t out[3]) {
  for (int d = 0; d < 3; ++d) {
    if (!(std::isfinite(box[d]) && box[d] > 0))
      return fail(PDX_ERR_ARGUMENT, "Neighbor box lengths must be finite and positive.");
    // floor(L / cell) in the list precision, as _choose_grid (NeighborTypes.jl:52-63)
    const float n = std::floor((float)box[d] / (float)rl);
    out[d] = n < 1 ? 1 : (int32_t)n;
  }
  return PDX_OK;
}

int enqueue_rebuild(pdx_nl_t nl, hipStream_t st) {
  const int N = nl->N;
  const int nb = blocks_for(N, TPB);
  int64_t *s = nl->state.p;
  const float *geom = nl->geom.p;
  if (!nl->allpairs) {
    const int sblocks = blocks_for(nl->ncell, SCAN_BLOCK);
    hipLaunchKernelGGL(k_bin, dim3(nb), dim3(TPB), 0, st, N, nl->packed.p, geom, nl->nx, nl->ny, nl->nz,
                       nl->cell_of_particle.p, nl->cell_counts.p, nl->cell_rank.p, s);
    if (nl->ncell <= SCAN_SINGLE_MAX) {
      hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, st, nl->ncell, N, nl->cell_counts.p,
                         nl->cell_offsets.p, s);
    } else {
      hipLaunchKernelGGL(k_scan_local, dim3(sblocks), dim3(TPB), 0, st, nl->ncell, nl->cell_counts.p,
                         nl->cell_offsets.p, nl->scan_block.p, s);
      hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, sblocks, nl->scan_block.p, s);
      hipLaunchKernelGGL(k_scan_finish, dim3(blocks_for(nl->ncell + 1, TPB)), dim3(TPB), 0, st, nl->ncell, N,
                         nl->cell_offsets.p, nl->scan_block.p, s);
    }
    hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(TPB), 0, st, N, nl->cell_of_particle.p, nl->cell_rank.p,
                       nl->cell_offsets.p, nl->particle_ids_sorted.p, s);
    hipLaunchKernelGGL(k_sort_cells, dim3(std::min(nl->ncell, 8192)), dim3(WAVE), 0, st, nl->ncell, nl->cell_offsets.p,
                       nl->particle_ids_sorted.p, nl->packed.p, nl->sorted.p, s);
    // lanes per particle: the list's row lanes; with PDX_ROW_STAGE=1 the rows are staged in LDS and
    // the lanes raised until the workgroup's rows fit the 64 KiB stage. Staging cuts the build's HBM
    // writes at 486 400 ions from 3.97 GB to 0.93 GB but not its time (1.13 ms at 32 lanes against
    // 1.12 ms written straight at 8), so the default writes straight from the search.
    int G = nl->row_lanes;
    bool stage = false;
    if (const char *v = std::getenv("PDX_ROW_STAGE")) stage = std::atoi(v) != 0;
    // the cell-staged build (k_rows_cells, PDX_ROW_CELLS=1): at 486 400 ions (8 000 cells, one
    // workgroup each) 1.06-1.09 ms against 1.12 ms for k_rows in spatial storage order and 1.33
    // against 1.55 ms in lattice order, but on the checks without rebuild its LDS-holding
    // workgroups cost 10 us against 5 us, so spatially ordered storage (ForceEngine.reorder)
    // keeps k_rows; the LDS-staged row write (PDX_ROW_STAGE=1) is a k_rows option
    bool cells = false;
    if (const char *v = std::getenv("PDX_ROW_CELLS")) cells = std::atoi(v) != 0;
    if (stage) cells = false;
    if (cells) {
#define PDX_ROWC(GG)                                                                                        \
  hipLaunchKernelGGL((k_rows_cells<GG>), dim3(std::min(nl->ncell, 8192)), dim3(TPB), 0, st, N, nl->cap, nl->ncell, \
                     nl->nx, nl->ny, nl->nz, \
                     nl->sorted.p, nl->cell_offsets.p, geom, nl->rl2, nl->neighbors_flat.p, nl->counts.p,     \
                     nl->diagnostics ? nl->candidates.p : nullptr, s);
      switch (G) {
        case 1: PDX_ROWC(1) break;
        case 2: PDX_ROWC(2) break;
        case 4: PDX_ROWC(4) break;
        case 8: PDX_ROWC(8) break;
        case 16: PDX_ROWC(16) break;
        default: PDX_ROWC(32) break;
      }
#undef PDX_ROWC
    } else {
    if (stage) {
      int g = G;
      while (g < 32 && (size_t)(TPB / g) * nl->cap * sizeof(int32_t) > ROW_STAGE_BYTES) g <<= 1;
      if ((size_t)(TPB / g) * nl->cap * sizeof(int32_t) <= ROW_STAGE_BYTES) G = g;
      else stage = false;
    }
    const int rb = (int)(((int64_t)N * G + TPB - 1) / TPB);
    const size_t lds = stage ? (size_t)(TPB / G) * nl->cap * sizeof(int32_t) : 0;
#define PDX_ROWS(GG)                                                                                       \
  if (stage)                                                                                               \
    hipLaunchKernelGGL((k_rows<GG, true>), dim3(rb), dim3(TPB), lds, st, N, nl->cap, nl->nx, nl->ny, nl->nz, \
                       nl->sorted.p, nl->cell_offsets.p, nl->cell_of_particle.p, geom, nl->rl2,            \
                       nl->neighbors_flat.p, nl->counts.p, nl->diagnostics ? nl->candidates.p : nullptr, s); \
  else                                                                                                     \
    hipLaunchKernelGGL((k_rows<GG, false>), dim3(rb), dim3(TPB), 0, st, N, nl->cap, nl->nx, nl->ny, nl->nz, \
                       nl->sorted.p, nl->cell_offsets.p, nl->cell_of_particle.p, geom, nl->rl2,            \
                       nl->neighbors_flat.p, nl->counts.p, nl->diagnostics ? nl->candidates.p : nullptr, s);
    switch (G) {
      case 1: PDX_ROWS(1) break;
      case 2: PDX_ROWS(2) break;
      case 4: PDX_ROWS(4) break;
      case 8: PDX_ROWS(8) break;
      case 16: PDX_ROWS(16) break;
      default: PDX_ROWS(32) break;
    }
#undef PDX_ROWS
    }
  } else {
    hipLaunchKernelGGL(k_rows_allpairs, dim3(nb), dim3(TPB), 0, st, N, nl->cap, nl->packed.p, geom,
                       nl->rl2, nl->neighbors_flat.p, nl->counts.p,
                       nl->diagnostics ? nl->candidates.p : nullptr, s);
  }
  hipLaunchKernelGGL(k_commit, dim3(nb), dim3(TPB), 0, st, N, nl->counts.p,
                     nl->diagnostics ? nl->candidates.p : nullptr, nl->cap_partial.p, nl->stat_partial.p, s);
  hipLaunchKernelGGL(k_commit_final, dim3(1), dim3(64), 0, st, nb, nl->cap, nl->cap_partial.p, nl->stat_partial.p, s);
  hipLaunchKernelGGL(k_copy_refs, dim3(nb), dim3(TPB), 0, st, N, nl->packed.p, nl->rref.p, s);
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

int enqueue_check(pdx_nl_t nl, const float *x, const float *y, const float *z, const float *q, hipStream_t st,
                  int force, const KickDrift *kd = nullptr) {
  if (!nl) return fail(PDX_ERR_ARGUMENT, "null neighbor list");
  if (!x || !y || (!z && nl->dim == 3)) return fail(PDX_ERR_ARGUMENT, "null coordinate array");
  const int N = nl->N;
  const int nb = blocks_for(N, TPB);
  nl->packed_src[0] = x; nl->packed_src[1] = y; nl->packed_src[2] = z; nl->packed_src[3] = q;
  float *xw = const_cast<float *>(x), *yw = const_cast<float *>(y), *zw = const_cast<float *>(z);
  if (kd)
    hipLaunchKernelGGL(k_decide<true>, dim3(nb), dim3(TPB), 0, st, N, xw, yw, zw, q, nl->rref.p, nl->packed.p,
                       nl->cell_counts.p, nl->ncell, (int)nl->allpairs, nl->geom.p, nl->state.p, nl->disp_partial.p,
                       (const int32_t *)nl->cell_offsets.p, (const int32_t *)nl->particle_ids_sorted.p,
                       nl->snap_offsets.p, nl->snap_ids.p, nl->snap_state.p, *kd);
  else
    hipLaunchKernelGGL(k_decide<false>, dim3(nb), dim3(TPB), 0, st, N, xw, yw, zw, q, nl->rref.p, nl->packed.p,
                       nl->cell_counts.p, nl->ncell, (int)nl->allpairs, nl->geom.p, nl->state.p, nl->disp_partial.p,
                       (const int32_t *)nl->cell_offsets.p, (const int32_t *)nl->particle_ids_sorted.p,
                       nl->snap_offsets.p, nl->snap_ids.p, nl->snap_state.p, KickDrift{});
  PDX_LAUNCH_CHECK();
  // the first check after creation / a box change builds; the second one snapshots that build
  if (nl->snap_host < 2) nl->snap_host += 1;
  nl->check_force = force;
  return PDX_OK;
}

// The decision from the check's partial maxima (and the snapshot's record), then the
// predicated rebuild.
int enqueue_decide_build(pdx_nl_t nl, hipStream_t st) {
  const float s = (float)nl->skin;
  const float threshold = (0.25f * s) * s;
  const int nb = blocks_for(nl->N, TPB);
  hipLaunchKernelGGL(k_decide_final, dim3(1), dim3(TPB), 0, st, nb, nl->disp_partial.p, nl->state.p, threshold,
                     nl->check_force, nl->snap_state.p, (int)nl->allpairs);
  PDX_LAUNCH_CHECK();
  return enqueue_rebuild(nl, st);
}

int enqueue_update(pdx_nl_t nl, const float *x, const float *y, const float *z, const float *q,
                   hipStream_t st, int force) {
  int rc = enqueue_check(nl, x, y, z, q, st, force);
  if (rc) return rc;
  return enqueue_decide_build(nl, st);
}

}  // namespace

extern "C" int pdx_nl_create(int32_t n, const double box[3], double cutoff, double skin, int32_t cap,
                             int32_t interval, int32_t flags, pdx_nl_t *out) {
  if (!out || !box) return fail(PDX_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  if (!(n > 0 && n < INT32_MAX)) return fail(PDX_ERR_ARGUMENT, "Neighbor builders require 0 < N < typemax(Int32).");
  if (cap <= 0) return fail(PDX_ERR_ARGUMENT, "Neighbor capacity must be positive.");
  if (!(std::isfinite(cutoff) && cutoff >= 0)) return fail(PDX_ERR_ARGUMENT, "Neighbor cutoff must be finite and nonnegative.");
  if (!(std::isfinite(skin) && skin >= 0)) return fail(PDX_ERR_ARGUMENT, "Neighbor skin must be finite and nonnegative.");
  if (interval < 1 || interval > 100) return fail(PDX_ERR_ARGUMENT, "Rebuild interval must be between 1 and 100");
  if ((int64_t)n * cap > INT32_MAX)
    return fail(PDX_ERR_ARGUMENT, "Neighbor table of N particles x cap exceeds the Int32 index range; lower `cap`.");
  auto *nl = new pdx_nl_s();
  nl->N = n;
  nl->cap = cap;
  nl->interval = interval;
  nl->cutoff = cutoff;
  nl->skin = skin;
  nl->allpairs = (flags & PDX_NL_ALLPAIRS) != 0;
  nl->diagnostics = (flags & PDX_NL_DIAGNOSTICS) != 0;
  nl->dim = (flags & PDX_NL_2D) ? 2 : 3;
  for (int d = 0; d < 3; ++d) nl->box[d] = box[d];
  const float rlf = std::max(1e-6f, (float)cutoff + (float)skin);
  nl->rl2 = rlf * rlf;
  // 2-D (DenseNeighborCUDA.jl:77, D = 2): every z is 0 and the list's z extent is three list
  // radii, so the z layers of the stencil beside the particles' own are empty and no pair has a
  // z image: the 3-D kernels then build the 2-D rows, which equal the 9-cell 2-D stencil's
  double box3[3] = {box[0], box[1], box[2]};
  if (nl->dim == 2) {
    box3[2] = 3.0 * (double)rlf * (1.0 + 1e-6);
    nl->box[2] = box3[2];
  }
  int32_t g[3];
  int rc = grid_for_box(box3, rlf, g);
  if (rc) { delete nl; return rc; }
  if (nl->allpairs) {
    const double mn = std::min(box3[0], std::min(box3[1], box3[2]));
    if (!(2.0 * rlf <= mn)) {
      delete nl;
      return fail(PDX_ERR_ARGUMENT, "All-pairs neighbor rows need cutoff+skin <= L/2 (minimum image).");
    }
    g[0] = g[1] = g[2] = 1;
  } else if (g[0] < 3 || g[1] < 3 || g[2] < 3) {
    delete nl;
    return fail(PDX_ERR_ARGUMENT,
                "Dense cell-list neighbor builder requires at least 3 cells per periodic dimension to avoid "
                "duplicate neighbor-cell visits. Increase the box size, reduce cutoff+skin, or use all-pairs rows.");
  }
  nl->nx = g[0]; nl->ny = g[1]; nl->nz = g[2];
  nl->ncell = g[0] * g[1] * g[2];
  // lanes per particle of the row build: enough to fill ~2^19 resident threads (256 CUs x 2048),
  // and at least one lane per ~256 candidates (27 cells of the mean occupancy): measured at 486 400
  // ions (1 700 candidates per particle), 8 lanes build in 1.12 ms against 1.65 ms for 2; at most
  // 32; PDX_ROW_LANES overrides (tuning)
  {
    int G = 1;
    while ((int64_t)n * G < 524288 && G < 32) G <<= 1;
    const double cand = 27.0 * (double)n / (double)std::max(1, nl->ncell);
    while (G < 32 && cand > 256.0 * G) G <<= 1;
    if (const char *v = std::getenv("PDX_ROW_LANES")) {
      const int e = std::atoi(v);
      if (e == 1 || e == 2 || e == 4 || e == 8 || e == 16 || e == 32) G = e;
    }
    nl->row_lanes = G;
  }
  const int nb = blocks_for(n, TPB);
  nl->ndisp = nb;
  const int sblocks = blocks_for(nl->ncell, SCAN_BLOCK);
  if ((rc = nl->state.alloc(PDX_DN_NSTATE)) || (rc = nl->geom.alloc(3)) ||
      (rc = nl->neighbors_flat.alloc((size_t)n * cap)) || (rc = nl->counts.alloc(n)) ||
      (rc = nl->cell_offsets.alloc(nl->ncell + 1)) || (rc = nl->cell_counts.alloc(nl->ncell)) ||
      (rc = nl->particle_ids_sorted.alloc(n)) || (rc = nl->cell_of_particle.alloc(n)) ||
      (rc = nl->cell_rank.alloc(n)) || (rc = nl->rref.alloc((size_t)3 * n)) || (rc = nl->packed.alloc(n)) ||
      (rc = nl->sorted.alloc(n)) || (rc = nl->sortedq.alloc(n)) || (rc = nl->disp_partial.alloc(nb)) || (rc = nl->cap_partial.alloc(nb)) ||
      (rc = nl->stat_partial.alloc((size_t)2 * nb)) || (rc = nl->scan_block.alloc(sblocks)) ||
      (rc = nl->candidates.alloc(nl->diagnostics ? n : 1)) ||
      (rc = nl->snap_offsets.alloc((size_t)nl->ncell + 1)) || (rc = nl->snap_ids.alloc(n)) ||
      (rc = nl->snap_state.alloc(4))) {
    delete nl;
    return rc;
  }
  int64_t st[PDX_DN_NSTATE] = {0};
  st[PDX_DN_INTERVAL] = interval;
  st[PDX_DN_DIAGNOSTICS] = nl->diagnostics ? 1 : 0;
  const float geomf[3] = {(float)box3[0], (float)box3[1], (float)box3[2]};
  std::vector<int32_t> ones(nl->ncell + 1, 1);
  // identity traversal order until the first build (mesh kernels may read it before then)
  std::vector<int32_t> ident(n);
  for (int32_t k = 0; k < n; ++k) ident[k] = k + 1;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMemcpy(nl->state.p, st, sizeof(st), hipMemcpyHostToDevice);
  const int64_t snap0[4] = {-1, 0, 1, 0};
  if (e == hipSuccess) e = hipMemcpy(nl->snap_state.p, snap0, sizeof(snap0), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(nl->geom.p, geomf, sizeof(geomf), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(nl->neighbors_flat.p, 0xff, sizeof(int32_t) * (size_t)n * cap);  // -1 padding
  if (e == hipSuccess) e = hipMemset(nl->counts.p, 0, sizeof(int32_t) * n);
  if (e == hipSuccess) e = hipMemset(nl->cell_counts.p, 0, sizeof(int32_t) * nl->ncell);
  if (e == hipSuccess) e = hipMemcpy(nl->cell_offsets.p, ones.data(), sizeof(int32_t) * (nl->ncell + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(nl->particle_ids_sorted.p, ident.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(nl->cell_of_particle.p, 0, sizeof(int32_t) * n);
  if (e == hipSuccess) e = hipMemset(nl->packed.p, 0, sizeof(float4) * n);
  if (e == hipSuccess) e = hipMemset(nl->rref.p, 0, sizeof(float) * 3 * (size_t)n);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    delete nl;
    return fail(PDX_ERR_DEVICE, std::string("neighbor list initialisation failed: ") + hipGetErrorString(e));
  }
  *out = nl;
  return PDX_OK;
}

extern "C" int pdx_nl_destroy(pdx_nl_t nl) {
  delete nl;
  return PDX_OK;
}

extern "C" int pdx_nl_info(pdx_nl_t nl, int32_t out[8]) {
  if (!nl || !out) return fail(PDX_ERR_ARGUMENT, "null argument");
  out[0] = nl->N; out[1] = nl->cap; out[2] = nl->nx; out[3] = nl->ny; out[4] = nl->nz;
  out[5] = nl->ncell; out[6] = nl->allpairs; out[7] = nl->diagnostics;
  return PDX_OK;
}

extern "C" int pdx_nl_arrays_get(pdx_nl_t nl, pdx_nl_arrays *out) {
  if (!nl || !out) return fail(PDX_ERR_ARGUMENT, "null argument");
  out->neighbors_flat = nl->neighbors_flat.p;
  out->counts = nl->counts.p;
  out->cell_offsets = nl->cell_offsets.p;
  out->particle_ids_sorted = nl->particle_ids_sorted.p;
  out->cell_of_particle = nl->cell_of_particle.p;
  out->rref_x = nl->rref.p;
  out->rref_y = nl->rref.p + nl->N;
  out->rref_z = nl->rref.p + 2 * (size_t)nl->N;
  out->packed = reinterpret_cast<float *>(nl->packed.p);
  out->state = nl->state.p;
  return PDX_OK;
}

extern "C" int pdx_nl_update(pdx_nl_t nl, const float *x, const float *y, const float *z, const float *q,
                             pdx_stream_t stream) {
  return enqueue_update(nl, x, y, z, q, as_stream(stream), 0);
}

extern "C" int pdx_nl_rebuild(pdx_nl_t nl, const float *x, const float *y, const float *z, const float *q,
                              pdx_stream_t stream) {
  return enqueue_update(nl, x, y, z, q, as_stream(stream), 1);
}

extern "C" int pdx_nl_check(pdx_nl_t nl, const float *x, const float *y, const float *z, const float *q,
                            pdx_stream_t stream) {
  return enqueue_check(nl, x, y, z, q, as_stream(stream), 0);
}

namespace {
int make_kick_drift(pdx_nl_t nl, float *vx, float *vy, float *vz, const float *fx, const float *fy,
                    const float *fz, double dt, double mass, const float *inv_mass, KickDrift &kd) {
  if (!nl) return fail(PDX_ERR_ARGUMENT, "null neighbor list");
  if (!vx || !vy || !fx || !fy || (nl->dim == 3 && (!vz || !fz)))
    return fail(PDX_ERR_ARGUMENT, "null velocity or force array");
  if (!(std::isfinite(dt) && dt > 0)) return fail(PDX_ERR_ARGUMENT, "dt must be finite and positive");
  if (!inv_mass && !(std::isfinite(mass) && mass > 0))
    return fail(PDX_ERR_ARGUMENT, "Inertial integrators require mass > 0");
  kd.vx = vx; kd.vy = vy; kd.vz = vz;
  kd.fx = fx; kd.fy = fy; kd.fz = fz;
  kd.inv_mass = inv_mass;
  kd.dt = (float)dt;
  kd.half_dt = (float)(0.5 * dt);
  kd.c = inv_mass ? 0.0f : (float)dt / 2.0f / (float)mass;  // (dt / 2) / m in FP32, as Run.jl:249
  return PDX_OK;
}
}  // namespace

extern "C" int pdx_nl_check_kick_drift(pdx_nl_t nl, float *x, float *y, float *z, const float *q, float *vx,
                                       float *vy, float *vz, const float *fx, const float *fy, const float *fz,
                                       double dt, double mass, const float *inv_mass, pdx_stream_t stream) {
  KickDrift kd{};
  int rc = make_kick_drift(nl, vx, vy, vz, fx, fy, fz, dt, mass, inv_mass, kd);
  if (rc) return rc;
  return enqueue_check(nl, x, y, z, q, as_stream(stream), 0, &kd);
}

extern "C" int pdx_nve_kick(pdx_nl_t nl, float *vx, float *vy, float *vz, const float *fx, const float *fy,
                            const float *fz, double dt, double mass, const float *inv_mass, pdx_stream_t stream) {
  KickDrift kd{};
  int rc = make_kick_drift(nl, vx, vy, vz, fx, fy, fz, dt, mass, inv_mass, kd);
  if (rc) return rc;
  hipLaunchKernelGGL(k_kick, dim3(blocks_for(nl->N, TPB)), dim3(TPB), 0, as_stream(stream), nl->N, kd, nl->dim,
                     nl->state.p);
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

extern "C" int pdx_nl_spatial_order(pdx_nl_t nl, int32_t *perm, pdx_stream_t stream) {
  if (!nl || !perm) return fail(PDX_ERR_ARGUMENT, "null argument");
  if (nl->allpairs) return fail(PDX_ERR_ARGUMENT, "all-pairs lists have no cells to order by");
  int64_t st[PDX_DN_NSTATE];
  hipStream_t s = as_stream(stream);
  PDX_HIP(hipStreamSynchronize(s));
  PDX_HIP(hipMemcpy(st, nl->state.p, sizeof(st), hipMemcpyDeviceToHost));
  if (st[PDX_DN_ERROR] != 0 || st[PDX_DN_VALID] == 0)
    return fail(PDX_ERR_STATE, "spatial order needs a valid list (update it first)");
  hipLaunchKernelGGL(k_spatial_order, dim3(nl->ncell), dim3(WAVE), 0, s, nl->ncell,
                     (const int32_t *)nl->cell_offsets.p, (const int32_t *)nl->particle_ids_sorted.p,
                     (const float4 *)nl->packed.p, perm);
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

extern "C" int pdx_nl_invalidate(pdx_nl_t nl, pdx_stream_t stream) {
  if (!nl) return fail(PDX_ERR_ARGUMENT, "null neighbor list");
  hipStream_t s = as_stream(stream);
  PDX_HIP(hipStreamSynchronize(s));
  const int64_t snap0[4] = {-1, 0, 1, 0};
  PDX_HIP(hipMemcpy(nl->snap_state.p, snap0, sizeof(snap0), hipMemcpyHostToDevice));
  nl->snap_host = 0;
  hipLaunchKernelGGL(k_invalidate, dim3(1), dim3(1), 0, s, nl->state.p);
  PDX_LAUNCH_CHECK();
  PDX_HIP(hipStreamSynchronize(s));
  return PDX_OK;
}

extern "C" int pdx_nl_build(pdx_nl_t nl, pdx_stream_t stream) {
  if (!nl) return fail(PDX_ERR_ARGUMENT, "null neighbor list");
  return enqueue_decide_build(nl, as_stream(stream));
}

extern "C" int pdx_nl_snapshot(pdx_nl_t nl, pdx_stream_t stream) {
  if (!nl) return fail(PDX_ERR_ARGUMENT, "null neighbor list");
  if (nl->allpairs) return fail(PDX_ERR_ARGUMENT, "all-pairs lists have no cells to snapshot");
  hipStream_t st = as_stream(stream);
  const int n = std::max(nl->N, nl->ncell + 1);
  hipLaunchKernelGGL(k_snap_copy, dim3(blocks_for(n, TPB)), dim3(TPB), 0, st, nl->N, nl->ncell,
                     (const int32_t *)nl->cell_offsets.p, (const int32_t *)nl->particle_ids_sorted.p,
                     nl->snap_offsets.p, nl->snap_ids.p, (const int64_t *)nl->state.p,
                     (const int64_t *)nl->snap_state.p);
  hipLaunchKernelGGL(k_snap_mark, dim3(1), dim3(1), 0, st, (const int64_t *)nl->state.p, nl->snap_state.p);
  PDX_LAUNCH_CHECK();
  if (nl->snap_host >= 1) nl->snap_host = 2;
  return PDX_OK;
}

extern "C" int pdx_nl_advance(pdx_nl_t nl, pdx_stream_t stream) {
  if (!nl) return fail(PDX_ERR_ARGUMENT, "null neighbor list");
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, as_stream(stream), nl->state.p);
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

extern "C" int pdx_nl_rearm(pdx_nl_t nl, pdx_stream_t stream) {
  if (!nl) return fail(PDX_ERR_ARGUMENT, "null neighbor list");
  hipLaunchKernelGGL(k_rearm, dim3(1), dim3(1), 0, as_stream(stream), nl->state.p);
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

extern "C" int pdx_nl_resize_capacity(pdx_nl_t nl, int32_t cap) {
  if (!nl) return fail(PDX_ERR_ARGUMENT, "null neighbor list");
  if (cap <= 0) return fail(PDX_ERR_ARGUMENT, "Invalid ELL capacity");
  if ((int64_t)nl->N * cap > INT32_MAX) return fail(PDX_ERR_ARGUMENT, "Neighbor table exceeds the Int32 index range");
  PDX_HIP(hipDeviceSynchronize());
  int rc = nl->neighbors_flat.alloc((size_t)nl->N * cap);
  if (rc) return rc;
  PDX_HIP(hipMemset(nl->neighbors_flat.p, 0xff, sizeof(int32_t) * (size_t)nl->N * cap));
  nl->cap = cap;
  hipLaunchKernelGGL(k_invalidate, dim3(1), dim3(1), 0, 0, nl->state.p);
  PDX_HIP(hipDeviceSynchronize());
  return PDX_OK;
}

extern "C" int pdx_nl_set_box(pdx_nl_t nl, const double box[3], pdx_stream_t stream) {
  if (!nl || !box) return fail(PDX_ERR_ARGUMENT, "null argument");
  const float rlf = std::sqrt(nl->rl2);
  double box3[3] = {box[0], box[1], nl->dim == 2 ? nl->box[2] : box[2]};  // 2-D: z extent kept
  int32_t g[3];
  int rc = grid_for_box(box3, rlf, g);
  if (rc) return rc;
  if (nl->allpairs) {
    const double mn = std::min(box3[0], std::min(box3[1], box3[2]));
    if (!(2.0 * rlf <= mn)) return fail(PDX_ERR_ARGUMENT, "All-pairs neighbor rows need cutoff+skin <= L/2.");
  } else {
    if (g[0] < 3 || g[1] < 3 || g[2] < 3)
      return fail(PDX_ERR_ARGUMENT, "The box became too small for the device cell grid (3 cells per dimension).");
    if (g[0] != nl->nx || g[1] != nl->ny || g[2] != nl->nz) {
      PDX_HIP(hipStreamSynchronize(as_stream(stream)));
      const int ncell = g[0] * g[1] * g[2];
      if ((rc = nl->cell_offsets.alloc(ncell + 1)) || (rc = nl->cell_counts.alloc(ncell)) ||
          (rc = nl->scan_block.alloc(blocks_for(ncell, SCAN_BLOCK))) || (rc = nl->snap_offsets.alloc(ncell + 1)))
        return rc;
      std::vector<int32_t> ones(ncell + 1, 1);
      PDX_HIP(hipMemcpy(nl->cell_offsets.p, ones.data(), sizeof(int32_t) * (ncell + 1), hipMemcpyHostToDevice));
      PDX_HIP(hipMemset(nl->cell_counts.p, 0, sizeof(int32_t) * ncell));
      nl->nx = g[0]; nl->ny = g[1]; nl->nz = g[2]; nl->ncell = ncell;
    }
  }
  for (int d = 0; d < 3; ++d) nl->box[d] = box3[d];
  {  // cells of the old box are no cover for the new one: no snapshot until the next build
    PDX_HIP(hipStreamSynchronize(as_stream(stream)));
    const int64_t snap0[4] = {-1, 0, 1, 0};
    PDX_HIP(hipMemcpy(nl->snap_state.p, snap0, sizeof(snap0), hipMemcpyHostToDevice));
    nl->snap_host = 0;
  }
  const float geomf[3] = {(float)box3[0], (float)box3[1], (float)box3[2]};
  PDX_HIP(hipMemcpyAsync(nl->geom.p, geomf, sizeof(geomf), hipMemcpyHostToDevice, as_stream(stream)));
  hipLaunchKernelGGL(k_invalidate, dim3(1), dim3(1), 0, as_stream(stream), nl->state.p);
  PDX_HIP(hipStreamSynchronize(as_stream(stream)));
  return PDX_OK;
}

extern "C" int pdx_nl_status(pdx_nl_t nl, int64_t state[PDX_DN_NSTATE], int32_t check, pdx_stream_t stream) {
  if (!nl || !state) return fail(PDX_ERR_ARGUMENT, "null argument");
  PDX_HIP(hipStreamSynchronize(as_stream(stream)));
  PDX_HIP(hipMemcpy(state, nl->state.p, sizeof(int64_t) * PDX_DN_NSTATE, hipMemcpyDeviceToHost));
  if (!check) return PDX_OK;
  if (state[PDX_DN_ERROR] == PDX_DN_CAPACITY_ERROR)
    return fail(PDX_ERR_CAPACITY, "Neighbor capacity " + std::to_string(nl->cap) + " is insufficient; at least " +
                                      std::to_string(state[PDX_DN_REQUIRED]) +
                                      " entries per particle are required. Rebuild with a larger cap.");
  if (state[PDX_DN_ERROR] == PDX_DN_COORDINATE_ERROR)
    return fail(PDX_ERR_COORDINATE,
                "Device ELL requires finite, centered coordinates; execution stopped before rebuilding/forces");
  if (state[PDX_DN_ERROR] != 0) return fail(PDX_ERR_STATE, "Device list stopped with an unknown error code");
  if (state[PDX_DN_VALID] == 0) return fail(PDX_ERR_STATE, "Device ELL has no valid list");
  return PDX_OK;
}

extern "C" int pdx_nl_set_row_lanes(pdx_nl_t nl, int32_t lanes) {
  if (!nl) return fail(PDX_ERR_ARGUMENT, "null neighbor list");
  if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32)
    return fail(PDX_ERR_ARGUMENT, "row-build lanes must be 1, 2, 4, 8, 16 or 32");
  nl->row_lanes = lanes;
  return PDX_OK;
}
