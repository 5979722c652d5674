// Batched independent replica boxes (LJ / WCA), each advanced by velocity Verlet, all of them
// in one persistent launch: one workgroup owns one replica for every step, so replicas never
// synchronise with each other and a step costs no kernel launch. Per step, inside the
// workgroup:
//   A  kick-drift v += c f, x += dt v, wrap to [-L/2, L/2) (_nve_kick_drift3_kernel!,
//      src/simulation/NVEFusion.jl:3-24), cell of each atom, LDS histogram;
//   B  exclusive scan of the cell counts (LDS);
//   C  scatter (x, y, z, id) into cell order in LDS;
//   D  each cell's slice put in atom-id order (stable order: the traversal, and with it the
//      FP32 force sums, depend only on the replica's own coordinates);
//   E  forces from the LDS-resident sorted coordinates over the (2S+1)^3 stencil of cells of
//      side >= rc/S (or over all atoms for boxes with fewer than 2S+1 cells a side, as the
//      reference's all-pairs rows, DenseNeighborCUDA.jl:128-151), then the closing kick
//      v += c f (_nve_kick3_kernel!, src/SimulationCore.jl:581-591).
// Atom state (x, v, f) stays in HBM / L2 between phases; the sorted coordinates and the
// cell table live in LDS (up to 8192 atoms: 128 KiB of float4, within gfx950's 160 KiB).
#include "pdx_internal.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

using namespace pdx;

namespace {

constexpr int NMAX = 8192;   // atoms per replica (LDS-resident sorted coordinates)
constexpr int NCMAX = 8192;  // cells per replica
constexpr int RED_BYTES = 1024;

struct Rep {
  int off, n;
  int model;     // PDX_PAIR_LJ / PDX_PAIR_WCA
  int S;         // stencil half-width in cells; 0: all pairs
  int nc[3];     // cells per dimension
  float L[3], halfL[3], cinv[3];  // box, half box, cells per unit length
  float eps, sig2, rc2, eshift;
};

__device__ __forceinline__ int cell_coord(float x, float halfL, float cinv, int nc) {
  int c = (int)((x + halfL) * cinv);
  return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

__device__ __forceinline__ float wrap_centred(float x, float L) {
  x -= (x >= 0.5f * L) ? L : 0.0f;
  x += (x < -0.5f * L) ? L : 0.0f;
  return x;
}

// Force factor f (force on i = f d) and energy u of one pair (same arithmetic as the ELL
// pair kernel: 24e(2s^12 - s^6)/r^2; u = 4e(s^12 - s^6) - u_c, or + e for WCA).
__device__ __forceinline__ void lj_terms(const Rep &R, float r2, float &f, float &u) {
  const bool in = (r2 > 0.f) & (r2 < R.rc2);
  const float invr2 = __builtin_amdgcn_rcpf(r2);
  const float s2 = R.sig2 * invr2;
  const float s6 = s2 * s2 * s2;
  const float s12 = s6 * s6;
  f = in ? 24.0f * R.eps * (2.0f * s12 - s6) * invr2 : 0.f;
  u = in ? 4.0f * R.eps * (s12 - s6) + (R.model == PDX_PAIR_WCA ? R.eps : -R.eshift) : 0.f;
}

template <int BS>
__device__ __forceinline__ double block_sum(double v, double *red) {
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < BS / WAVE; ++w) s += red[w];
  return s;
}

template <int BS>
__global__ __launch_bounds__(BS) void k_replicas(const Rep *__restrict__ reps, const int32_t *__restrict__ order,
                                                 float *x, float *y, float *z, float *vx, float *vy, float *vz,
                                                 float *fx, float *fy, float *fz, int32_t *key, float c, float dt,
                                                 float mass, int nsteps, int first_force, int energy, int reg0,
                                                 int reg1, double *out) {
  extern __shared__ float4 smem[];
  float4 *spos = smem;
  int *cnt = reinterpret_cast<int *>(smem);  // aliases spos while binning
  unsigned short *cst = reinterpret_cast<unsigned short *>(reinterpret_cast<char *>(smem) + reg0);
  int *wsum = reinterpret_cast<int *>(reinterpret_cast<char *>(smem) + reg0 + reg1);
  double *red = reinterpret_cast<double *>(reinterpret_cast<char *>(smem) + reg0 + reg1 + 128);

  const int rid = order[blockIdx.x];
  const Rep R = reps[rid];
  const int t = threadIdx.x, lane = t & (WAVE - 1), wave = t / WAVE;
  const int n = R.n, off = R.off;
  const bool allpairs = R.S == 0;
  const int ncell = allpairs ? 0 : R.nc[0] * R.nc[1] * R.nc[2];
  const int chunk = (ncell + BS - 1) / BS;
  double pe = 0.0, ke = 0.0, vir = 0.0;

  for (int step = first_force ? 0 : 1; step <= nsteps; ++step) {
    const bool integrate = step > 0;
    const bool last = step == nsteps;
    // ---- A: kick-drift, cell counts
    for (int k = t; k < ncell; k += BS) cnt[k] = 0;
    __syncthreads();
    for (int i = t; i < n; i += BS) {
      const int g = off + i;
      float px = x[g], py = y[g], pz = z[g];
      if (integrate) {
        const float ux = fmaf(c, fx[g], vx[g]), uy = fmaf(c, fy[g], vy[g]), uz = fmaf(c, fz[g], vz[g]);
        vx[g] = ux; vy[g] = uy; vz[g] = uz;
        px = wrap_centred(fmaf(dt, ux, px), R.L[0]);
        py = wrap_centred(fmaf(dt, uy, py), R.L[1]);
        pz = wrap_centred(fmaf(dt, uz, pz), R.L[2]);
        x[g] = px; y[g] = py; z[g] = pz;
      }
      if (allpairs) {
        spos[i] = make_float4(px, py, pz, __int_as_float(i));
      } else {
        const int cell = (cell_coord(pz, R.halfL[2], R.cinv[2], R.nc[2]) * R.nc[1] +
                          cell_coord(py, R.halfL[1], R.cinv[1], R.nc[1])) * R.nc[0] +
                         cell_coord(px, R.halfL[0], R.cinv[0], R.nc[0]);
        const int r = atomicAdd(&cnt[cell], 1);
        key[g] = cell | (r << 16);
      }
    }
    __syncthreads();
    if (!allpairs) {
      // ---- B: exclusive scan of the counts (contiguous chunk per thread, wave scans)
      const int k0 = min(t * chunk, ncell), k1 = min(k0 + chunk, ncell);
      int sum = 0;
      for (int k = k0; k < k1; ++k) sum += cnt[k];
      int v = sum;
#pragma unroll
      for (int o = 1; o < WAVE; o <<= 1) {
        const int u = __shfl_up(v, o, WAVE);
        if (lane >= o) v += u;
      }
      if (lane == WAVE - 1) wsum[wave] = v;
      __syncthreads();
      int base = v - sum;
      for (int w = 0; w < wave; ++w) base += wsum[w];
      for (int k = k0; k < k1; ++k) {
        const int ck = cnt[k];
        cst[k] = (unsigned short)base;
        base += ck;
      }
      if (t == 0) cst[ncell] = (unsigned short)n;
      __syncthreads();
      // ---- C: scatter into cell order (overwrites the counts)
      for (int i = t; i < n; i += BS) {
        const int g = off + i;
        const int kk = key[g];
        const int slot = cst[kk & 0xffff] + (kk >> 16);
        spos[slot] = make_float4(x[g], y[g], z[g], __int_as_float(i));
      }
      __syncthreads();
      // ---- D: atom-id order inside every cell
      for (int cell = t; cell < ncell; cell += BS) {
        const int b = cst[cell], e = cst[cell + 1];
        for (int a = b + 1; a < e; ++a) {
          const float4 p = spos[a];
          const int id = __float_as_int(p.w);
          int j = a - 1;
          while (j >= b && __float_as_int(spos[j].w) > id) {
            spos[j + 1] = spos[j];
            --j;
          }
          spos[j + 1] = p;
        }
      }
      __syncthreads();
    }
    // ---- E: forces and the closing kick
    const bool acc = last && energy;
    for (int s = t; s < n; s += BS) {
      const float4 p = spos[s];
      float ax = 0.f, ay = 0.f, az = 0.f, e = 0.f, w = 0.f;
      if (allpairs) {
        for (int j = 0; j < n; ++j) {
          const float4 q = spos[j];
          const float dx = mic_fast(p.x - q.x, R.halfL[0], R.L[0]);
          const float dy = mic_fast(p.y - q.y, R.halfL[1], R.L[1]);
          const float dz = mic_fast(p.z - q.z, R.halfL[2], R.L[2]);
          const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          float f, u;
          lj_terms(R, j == s ? 0.f : r2, f, u);
          ax = fmaf(f, dx, ax); ay = fmaf(f, dy, ay); az = fmaf(f, dz, az);
          e += u;
          w = fmaf(f, r2, w);
        }
      } else {
        const int cx = cell_coord(p.x, R.halfL[0], R.cinv[0], R.nc[0]);
        const int cy = cell_coord(p.y, R.halfL[1], R.cinv[1], R.nc[1]);
        const int cz = cell_coord(p.z, R.halfL[2], R.cinv[2], R.nc[2]);
        for (int dz3 = -R.S; dz3 <= R.S; ++dz3) {
          int zz = cz + dz3;
          zz += zz < 0 ? R.nc[2] : 0;
          zz -= zz >= R.nc[2] ? R.nc[2] : 0;
          for (int dy3 = -R.S; dy3 <= R.S; ++dy3) {
            int yy = cy + dy3;
            yy += yy < 0 ? R.nc[1] : 0;
            yy -= yy >= R.nc[1] ? R.nc[1] : 0;
            const int row = (zz * R.nc[1] + yy) * R.nc[0];
            for (int dx3 = -R.S; dx3 <= R.S; ++dx3) {
              int xx = cx + dx3;
              xx += xx < 0 ? R.nc[0] : 0;
              xx -= xx >= R.nc[0] ? R.nc[0] : 0;
              const int cell = row + xx;
              const int b = cst[cell], en = cst[cell + 1];
              for (int j = b; j < en; ++j) {
                const float4 q = spos[j];
                const float dx = mic_fast(p.x - q.x, R.halfL[0], R.L[0]);
                const float dy = mic_fast(p.y - q.y, R.halfL[1], R.L[1]);
                const float dz = mic_fast(p.z - q.z, R.halfL[2], R.L[2]);
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                float f, u;
                lj_terms(R, j == s ? 0.f : r2, f, u);
                ax = fmaf(f, dx, ax); ay = fmaf(f, dy, ay); az = fmaf(f, dz, az);
                e += u;
                w = fmaf(f, r2, w);
              }
            }
          }
        }
      }
      const int g = off + __float_as_int(p.w);
      fx[g] = ax; fy[g] = ay; fz[g] = az;
      float ux = vx[g], uy = vy[g], uz = vz[g];
      if (integrate) {
        ux = fmaf(c, ax, ux); uy = fmaf(c, ay, uy); uz = fmaf(c, az, uz);
        vx[g] = ux; vy[g] = uy; vz[g] = uz;
      }
      if (acc) {
        pe += 0.5 * (double)e;
        vir += 0.5 * (double)w;
        ke += 0.5 * (double)mass * ((double)ux * ux + (double)uy * uy + (double)uz * uz);
      }
    }
    __syncthreads();  // spos / cnt are rewritten by the next step
  }
  if (energy) {
    const double a = block_sum<BS>(pe, red);
    const double b = block_sum<BS>(ke, red);
    const double d = block_sum<BS>(vir, red);
    if (t == 0) {
      out[4 * rid + 0] = a;
      out[4 * rid + 1] = b;
      out[4 * rid + 2] = d;
      out[4 * rid + 3] = (double)n;
    }
  }
}

size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }

}  // namespace

struct pdx_rep_s {
  std::vector<pdx_replica_spec> specs;
  std::vector<Rep> host;
  std::vector<int64_t> offsets;
  int nrep = 0, nmax = 0, ncmax = 0, block = 512, stencil = 1, nallpairs = 0;
  int64_t total = 0;
  size_t reg0 = 0, reg1 = 0, lds = 0;
  DevBuf<Rep> reps;
  DevBuf<int32_t> order, key;
};

namespace {

// Cell tables for stencil half-width S (1 or 2): cells of side >= rc / S, at least 2S + 1 a
// side (else S = 1, else all pairs), at most NCMAX cells (the widest dimension is coarsened).
int layout(pdx_rep_t r, int S) {
  r->host.resize(r->nrep);
  r->nmax = 0;
  r->ncmax = 0;
  r->nallpairs = 0;
  for (int k = 0; k < r->nrep; ++k) {
    const pdx_replica_spec &sp = r->specs[k];
    Rep R{};
    R.off = (int)r->offsets[k];
    R.n = sp.n;
    R.model = sp.model;
    const double rc = sp.model == PDX_PAIR_WCA ? std::pow(2.0, 1.0 / 6.0) * sp.sigma : sp.cutoff;
    const double rc2 = sp.model == PDX_PAIR_WCA ? 1.2599211 * sp.sigma * sp.sigma : sp.cutoff * sp.cutoff;
    R.eps = (float)sp.epsilon;
    R.sig2 = (float)(sp.sigma * sp.sigma);
    R.rc2 = (float)rc2;
    if (sp.model == PDX_PAIR_LJ && sp.shift) {
      const double s6 = std::pow(sp.sigma * sp.sigma / rc2, 3);
      R.eshift = (float)(4.0 * sp.epsilon * (s6 * s6 - s6));
    }
    for (int d = 0; d < 3; ++d) {
      R.L[d] = (float)sp.box[d];
      R.halfL[d] = (float)(0.5 * sp.box[d]);
    }
    int s = S;
    int nc[3];
    for (; s >= 1; --s) {
      bool ok = true;
      for (int d = 0; d < 3; ++d) {
        nc[d] = (int)std::floor(sp.box[d] * s / rc);
        ok &= nc[d] >= 2 * s + 1;
      }
      if (!ok) continue;
      while ((int64_t)nc[0] * nc[1] * nc[2] > NCMAX) {
        int dmax = 0;
        for (int d = 1; d < 3; ++d)
          if (nc[d] > nc[dmax]) dmax = d;
        --nc[dmax];
      }
      if (nc[0] >= 2 * s + 1 && nc[1] >= 2 * s + 1 && nc[2] >= 2 * s + 1) break;
    }
    if (s < 1) {
      R.S = 0;
      ++r->nallpairs;
      R.nc[0] = R.nc[1] = R.nc[2] = 1;
    } else {
      R.S = s;
      for (int d = 0; d < 3; ++d) {
        R.nc[d] = nc[d];
        R.cinv[d] = (float)(nc[d] / sp.box[d]);
      }
      r->ncmax = std::max(r->ncmax, nc[0] * nc[1] * nc[2]);
    }
    r->nmax = std::max(r->nmax, R.n);
    r->host[k] = R;
  }
  r->reg0 = align16(std::max<size_t>(16 * (size_t)r->nmax, 4 * (size_t)r->ncmax));
  r->reg1 = align16(2 * ((size_t)r->ncmax + 1));
  r->lds = r->reg0 + r->reg1 + 128 + RED_BYTES;
  if (r->lds > 160 * 1024) return fail(PDX_ERR_ARGUMENT, "replica tables exceed the 160 KiB of LDS");
  if (hipMemcpy(r->reps.p, r->host.data(), sizeof(Rep) * r->nrep, hipMemcpyHostToDevice) != hipSuccess)
    return fail(PDX_ERR_DEVICE, "replica table upload failed");
  // largest work first: the workgroups of the longest replicas start before the short ones
  std::vector<int32_t> ord(r->nrep);
  std::iota(ord.begin(), ord.end(), 0);
  auto work = [&](int k) {
    const Rep &R = r->host[k];
    const double cand = R.S == 0 ? (double)R.n : std::pow(2 * R.S + 1, 3) * R.n /
                                                     ((double)R.nc[0] * R.nc[1] * R.nc[2]);
    return (double)R.n * cand;
  };
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return work(a) > work(b); });
  if (hipMemcpy(r->order.p, ord.data(), sizeof(int32_t) * r->nrep, hipMemcpyHostToDevice) != hipSuccess)
    return fail(PDX_ERR_DEVICE, "replica order upload failed");
  r->stencil = S;
  return PDX_OK;
}

template <int BS>
int launch(pdx_rep_t r, float *x, float *y, float *z, float *vx, float *vy, float *vz, float *fx, float *fy,
           float *fz, float c, float dt, float mass, int nsteps, int first, int energy, double *out,
           hipStream_t st) {
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_replicas<BS>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipGetLastError();
    attr = true;
  }
  hipLaunchKernelGGL(k_replicas<BS>, dim3(r->nrep), dim3(BS), r->lds, st, r->reps.p, r->order.p, x, y, z, vx, vy,
                     vz, fx, fy, fz, r->key.p, c, dt, mass, nsteps, first, energy, (int)r->reg0, (int)r->reg1, out);
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

}  // namespace

extern "C" int pdx_replicas_create(int32_t nrep, const pdx_replica_spec *specs, int32_t block, int32_t stencil,
                                   pdx_rep_t *out) {
  if (!out || !specs) return fail(PDX_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  if (nrep < 1) return fail(PDX_ERR_ARGUMENT, "at least one replica is required");
  if (block == 0) block = 512;
  if (block != 256 && block != 512 && block != 1024) return fail(PDX_ERR_ARGUMENT, "block must be 256, 512 or 1024");
  if (stencil == 0) stencil = 1;
  if (stencil != 1 && stencil != 2) return fail(PDX_ERR_ARGUMENT, "stencil must be 1 or 2");
  int64_t total = 0;
  for (int k = 0; k < nrep; ++k) {
    const pdx_replica_spec &s = specs[k];
    if (s.n < 1 || s.n > NMAX)
      return fail(PDX_ERR_ARGUMENT, "replica " + std::to_string(k) + ": atom count must be in [1, 8192]");
    if (s.model != PDX_PAIR_LJ && s.model != PDX_PAIR_WCA)
      return fail(PDX_ERR_ARGUMENT, "replica " + std::to_string(k) + ": model must be LJ or WCA");
    if (!(s.sigma > 0) || !(s.epsilon >= 0))
      return fail(PDX_ERR_ARGUMENT, "replica " + std::to_string(k) + ": sigma must be > 0, epsilon >= 0");
    const double rc = s.model == PDX_PAIR_WCA ? std::pow(2.0, 1.0 / 6.0) * s.sigma : s.cutoff;
    if (!(rc > 0)) return fail(PDX_ERR_ARGUMENT, "replica " + std::to_string(k) + ": cutoff must be > 0");
    for (int d = 0; d < 3; ++d)
      if (!(s.box[d] >= 2.0 * rc))
        return fail(PDX_ERR_ARGUMENT, "replica " + std::to_string(k) +
                                          ": box lengths must be at least twice the cutoff (minimum image)");
    total += s.n;
  }
  if (total > INT32_MAX) return fail(PDX_ERR_ARGUMENT, "too many atoms in total");
  auto *r = new pdx_rep_s();
  r->nrep = nrep;
  r->specs.assign(specs, specs + nrep);
  r->offsets.resize(nrep + 1);
  r->offsets[0] = 0;
  for (int k = 0; k < nrep; ++k) r->offsets[k + 1] = r->offsets[k] + specs[k].n;
  r->total = total;
  r->block = block;
  int rc;
  if ((rc = r->reps.alloc(nrep)) || (rc = r->order.alloc(nrep)) || (rc = r->key.alloc(total)) ||
      (rc = layout(r, stencil))) {
    delete r;
    return rc;
  }
  *out = r;
  return PDX_OK;
}

extern "C" int pdx_replicas_destroy(pdx_rep_t r) {
  delete r;
  return PDX_OK;
}

extern "C" int pdx_replicas_set_launch(pdx_rep_t r, int32_t block, int32_t stencil) {
  if (!r) return fail(PDX_ERR_ARGUMENT, "null handle");
  if (block != 256 && block != 512 && block != 1024) return fail(PDX_ERR_ARGUMENT, "block must be 256, 512 or 1024");
  if (stencil != 1 && stencil != 2) return fail(PDX_ERR_ARGUMENT, "stencil must be 1 or 2");
  r->block = block;
  return stencil == r->stencil ? PDX_OK : layout(r, stencil);
}

extern "C" int pdx_replicas_info(pdx_rep_t r, int64_t out[8]) {
  if (!r || !out) return fail(PDX_ERR_ARGUMENT, "null argument");
  out[0] = r->total;
  out[1] = r->nrep;
  out[2] = r->nmax;
  out[3] = (int64_t)r->lds;
  out[4] = r->block;
  out[5] = r->stencil;
  out[6] = r->nallpairs;
  out[7] = r->ncmax;
  return PDX_OK;
}

extern "C" int pdx_replicas_offsets(pdx_rep_t r, int64_t *offsets) {
  if (!r || !offsets) return fail(PDX_ERR_ARGUMENT, "null argument");
  std::copy(r->offsets.begin(), r->offsets.end(), offsets);
  return PDX_OK;
}

extern "C" int pdx_replicas_run(pdx_rep_t r, float *x, float *y, float *z, float *vx, float *vy, float *vz, float *fx,
                                float *fy, float *fz, double mass, double dt, int32_t nsteps, int32_t flags,
                                double *out, pdx_stream_t stream) {
  if (!r) return fail(PDX_ERR_ARGUMENT, "null handle");
  if (!x || !y || !z || !vx || !vy || !vz || !fx || !fy || !fz) return fail(PDX_ERR_ARGUMENT, "null array");
  if (nsteps < 0) return fail(PDX_ERR_ARGUMENT, "nsteps must be >= 0");
  if (!(mass > 0) || !(dt >= 0)) return fail(PDX_ERR_ARGUMENT, "mass must be > 0 and dt >= 0");
  const bool energy = (flags & PDX_ENERGY) != 0;
  if (energy && !out) return fail(PDX_ERR_ARGUMENT, "PDX_ENERGY needs the output array");
  const int first = (flags & PDX_VV_FORCES_VALID) ? 0 : 1;
  if (nsteps == 0 && !first) return PDX_OK;
  const float c = (float)(0.5 * dt / mass);
  hipStream_t st = as_stream(stream);
  switch (r->block) {
    case 256: return launch<256>(r, x, y, z, vx, vy, vz, fx, fy, fz, c, (float)dt, (float)mass, nsteps, first, energy, out, st);
    case 512: return launch<512>(r, x, y, z, vx, vy, vz, fx, fy, fz, c, (float)dt, (float)mass, nsteps, first, energy, out, st);
    default: return launch<1024>(r, x, y, z, vx, vy, vz, fx, fy, fz, c, (float)dt, (float)mass, nsteps, first, energy, out, st);
  }
}
