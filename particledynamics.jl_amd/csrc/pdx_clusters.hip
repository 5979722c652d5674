// Cluster-pair lists and the Newton-3 tile kernel for the short-range pair forces
// (ClusterPairList / ClusterPairEngine: src/clusterpairs/List.jl, Engine.jl, Kernels.jl).
//
// Clustering (build): particles are counting-sorted by (x-y column, z bin, id); each column's
// sequence is cut into super-clusters of 64 and each super-cluster is split at its x, y and z
// medians into eight clusters of at most eight (List.jl:4-8, 210-330).
// Search (build): one wave64 per super-cluster S walks the columns within rlist of its own and
// keeps every j-cluster of a super-cluster S2 >= S whose bounding box comes within rlist of one
// of S's eight cluster boxes, with the 8-bit mask of those i-clusters (inside S only j-clusters
// cj >= ci); entries land in S's fixed segment in a fixed order (columns, S2, cj), so lists are
// identical from run to run.
// Force pass: a workgroup of W wave64s per super-cluster; lane (ti, tj) = (lane & 7, lane >> 3)
// holds i-atom ti of all eight i-clusters in registers and evaluates it against j-atom tj of the
// entry's j-cluster, so one 8 x 8 tile is one pass of the wave and the mask bits are
// wave-uniform branches. i-sums stay in registers over the super-cluster's entries and meet
// across the W waves in LDS in a fixed order. The j-sums of an entry are reduced over the eight
// ti lanes with DPP and stored in the entry's own row of a partial buffer instead of being added
// with float atomics (a device-scope atomic resolves behind eight separate L2s on this part);
// the scatter pass adds, per slot, its i-sum and the j-partials of the entries naming its
// cluster in entry order (a CSR built with the list), so the forces are bitwise reproducible.
// The pruning pass is the same kernel over the outer list: it also ballots, per tile, whether
// any pair lies inside rmax + prune_buffer, and compacts the kept entries into the working list
// (Kernels.jl:349-394, List.jl:719-759).
#include "pdx_internal.hpp"
#include "pdx_pairmath.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

using namespace pdx;

namespace {

constexpr int TPB = 256;
constexpr int CL = 8;    // atoms per cluster
constexpr int SUP = 64;  // atoms per super-cluster (eight clusters)
constexpr int KV = 10;   // partial values per atom: fx, fy, fz, energy, six virial components
constexpr int MAXC = 24;
constexpr int MAXW = 8;  // wave64s per workgroup of the force pass
constexpr int MAXSPLIT = 8;  // workgroups per super-cluster of the force pass
constexpr int MAXCHUNK = MAXW * MAXSPLIT;

// control words (device int32)
enum : int {
  W_NS = PDX_CL_NS,
  W_ERR = PDX_CL_ERROR,
  W_REQ = PDX_CL_REQUIRED,
  W_BUILD = 3,    // this update builds
  W_PRUNE = 4,    // the next force pass prunes (set with a build or on prune displacement)
  W_REQUEST = 5,  // a build is requested (creation, resize, caller)
  W_NBUILD = PDX_CL_BUILDS,
  W_NPRUNE = PDX_CL_PRUNES,
  W_LEN = 16
};

// Fraction of the box in [0, 1) of a coordinate (any image); NaN and the rounding edge map to 0.
__device__ __forceinline__ float wrap01(float x, float invL) {
  float u = x * invL + 0.5f;
  u -= floorf(u);
  return (u >= 0.0f && u < 1.0f) ? u : 0.0f;
}

// Squared gap between two axis-aligned boxes (centre c, half extents h) at the minimum image.
__device__ __forceinline__ float box_gap2(float4 ca, float4 ha, float4 cb, float4 hb, float Lx, float Ly,
                                          float Lz) {
  const float gx = fmaxf(fabsf(mic_fast(ca.x - cb.x, 0.5f * Lx, Lx)) - ha.x - hb.x, 0.0f);
  const float gy = fmaxf(fabsf(mic_fast(ca.y - cb.y, 0.5f * Ly, Ly)) - ha.y - hb.y, 0.0f);
  const float gz = fmaxf(fabsf(mic_fast(ca.z - cb.z, 0.5f * Lz, Lz)) - ha.z - hb.z, 0.0f);
  return gx * gx + gy * gy + gz * gz;
}

// Sum over the eight lanes of a half-row (lanes 8k .. 8k+7) with three DPP adds: xor 1 and xor 2
// as quad permutations, then the half-row mirror (lane l <-> 7 - l), which pairs the two quads.
__device__ __forceinline__ float sum8_dpp(float v) {
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false));
  return v;
}

// ---------------------------------------------------------------- decision
// Per particle the squared displacement since the build (ratio to (skin/2)^2) and since the
// last prune (ratio to (prune_buffer/2)^2) at the minimum image; per-block maxima.
// ref: 6 rows of N (build x, y, z, prune x, y, z).
__global__ __launch_bounds__(TPB) void k_cl_decide(int N, const float *__restrict__ x, const float *__restrict__ y,
                                                   const float *__restrict__ z, const float *__restrict__ ref,
                                                   const float *__restrict__ geom, float ib2, float ip2,
                                                   float *__restrict__ dpart, int32_t *ctl) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  float mb = 0.0f, mp = 0.0f;
  if (i < N) {
    const float xi = x[i], yi = y[i], zi = z[i];
    if (!(isfinite(xi) && isfinite(yi) && isfinite(zi))) ctl[W_ERR] = 2;
    const float Lx = geom[0], Ly = geom[1], Lz = geom[2];
    float dx = mic_fast(xi - ref[i], 0.5f * Lx, Lx);
    float dy = mic_fast(yi - ref[(size_t)N + i], 0.5f * Ly, Ly);
    float dz = mic_fast(zi - ref[(size_t)2 * N + i], 0.5f * Lz, Lz);
    mb = (dx * dx + dy * dy + dz * dz) * ib2;
    dx = mic_fast(xi - ref[(size_t)3 * N + i], 0.5f * Lx, Lx);
    dy = mic_fast(yi - ref[(size_t)4 * N + i], 0.5f * Ly, Ly);
    dz = mic_fast(zi - ref[(size_t)5 * N + i], 0.5f * Lz, Lz);
    mp = (dx * dx + dy * dy + dz * dz) * ip2;
  }
  mb = wave_max(mb);
  mp = wave_max(mp);
  __shared__ float sh[2][TPB / WAVE];
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  if (lane == 0) {
    sh[0][w] = mb;
    sh[1][w] = mp;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = sh[0][0], p = sh[1][0];
    for (int k = 1; k < TPB / WAVE; ++k) {
      b = fmaxf(b, sh[0][k]);
      p = fmaxf(p, sh[1][k]);
    }
    dpart[2 * blockIdx.x] = b;
    dpart[2 * blockIdx.x + 1] = p;
  }
}

// One block: this update's build / prune decision (_decide_kernel!, Engine.jl:301-324), and the
// bin histogram cleared for a build.
__global__ __launch_bounds__(TPB) void k_cl_decide_final(int nblocks, const float *__restrict__ dpart, int32_t *ctl,
                                                         int nbins, int32_t *bin_cnt) {
  __shared__ float sb[TPB / WAVE], sp[TPB / WAVE];
  __shared__ int build_s;
  float mb = 0.0f, mp = 0.0f;
  for (int b = threadIdx.x; b < nblocks; b += TPB) {
    mb = fmaxf(mb, dpart[2 * b]);
    mp = fmaxf(mp, dpart[2 * b + 1]);
  }
  mb = wave_max(mb);
  mp = wave_max(mp);
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  if (lane == 0) {
    sb[w] = mb;
    sp[w] = mp;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float B = sb[0], P = sp[0];
    for (int k = 1; k < TPB / WAVE; ++k) {
      B = fmaxf(B, sb[k]);
      P = fmaxf(P, sp[k]);
    }
    const bool err = ctl[W_ERR] != 0;
    const bool build = !err && (ctl[W_REQUEST] != 0 || B > 1.0f);
    const bool prune = !err && (build || P > 1.0f || ctl[W_PRUNE] != 0);
    if (build) {
      ctl[W_REQUEST] = 0;
      ctl[W_NBUILD] += 1;
      ctl[W_REQ] = 0;
    }
    if (prune && ctl[W_PRUNE] == 0) ctl[W_NPRUNE] += 1;
    ctl[W_BUILD] = build ? 1 : 0;
    ctl[W_PRUNE] = prune ? 1 : 0;
    build_s = build;
  }
  __syncthreads();
  if (build_s)
    for (int k = threadIdx.x; k < nbins; k += TPB) bin_cnt[k] = 0;
}

// ---------------------------------------------------------------- clustering
// Bin of every particle: column (x, y) and z bin (_column_key_kernel!, List.jl:214-232); the
// rank inside the bin comes from the histogram's atomic counter and is sorted away later.
__global__ __launch_bounds__(TPB) void k_cl_bin(int N, const float *__restrict__ x, const float *__restrict__ y,
                                                const float *__restrict__ z, const float *__restrict__ geom, int ncx,
                                                int ncy, int nzb, int32_t *bin_cnt, int32_t *bin_of,
                                                int32_t *bin_rank, const int32_t *ctl) {
  if (!ctl[W_BUILD]) return;
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= N) return;
  const float ux = wrap01(x[i], 1.0f / geom[0]);
  const float uy = wrap01(y[i], 1.0f / geom[1]);
  const float uz = wrap01(z[i], 1.0f / geom[2]);
  const int cx = min((int)(ux * ncx), ncx - 1);
  const int cy = min((int)(uy * ncy), ncy - 1);
  const int cz = min((int)(uz * nzb), nzb - 1);
  const int b = (cy * ncx + cx) * nzb + cz;
  bin_of[i] = b;
  bin_rank[i] = atomicAdd(&bin_cnt[b], 1);
}

// Exclusive prefix sums of cnt[0, n) into off[0, n] (off[n] = total) in one workgroup: thread t
// sums one contiguous run, the run totals are scanned with wave64 shuffles, then every run is
// written out. Predicated on control word `word`.
__global__ __launch_bounds__(1024) void k_cl_scan(int n, const int32_t *__restrict__ cnt, int32_t *off,
                                                  const int32_t *ctl, int word) {
  if (!ctl[word]) return;
  __shared__ int32_t wt[16];
  const int per = (n + 1023) / 1024;
  const int c0 = min((int)threadIdx.x * per, n), c1 = min(c0 + per, n);
  int32_t run = 0;
  for (int c = c0; c < c1; ++c) run += cnt[c];
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  int32_t v = run;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const int32_t u = __shfl_up(v, o, WAVE);
    if (lane >= o) v += u;
  }
  if (lane == WAVE - 1) wt[w] = v;
  __syncthreads();
  int32_t base = v - run;
  for (int k = 0; k < w; ++k) base += wt[k];
  for (int c = c0; c < c1; ++c) {
    off[c] = base;
    base += cnt[c];
  }
  if (threadIdx.x == 0) {
    int32_t tot = 0;
    for (int k = 0; k < 16; ++k) tot += wt[k];
    off[n] = tot;
  }
}

__global__ __launch_bounds__(TPB) void k_cl_order(int N, const int32_t *__restrict__ bin_of,
                                                  const int32_t *__restrict__ bin_rank,
                                                  const int32_t *__restrict__ bin_start, int32_t *order,
                                                  const int32_t *ctl) {
  if (!ctl[W_BUILD]) return;
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= N) return;
  order[bin_start[bin_of[i]] + bin_rank[i]] = i;
}

// The mesh traversal order (mesh_order, a copy of the cluster order): 1-based ids in slot order
// without the padding, i.e. the sorted order itself.
__global__ __launch_bounds__(TPB) void k_cl_order1(int N, const int32_t *__restrict__ order, int32_t *order1,
                                                   const int32_t *ctl) {
  if (!ctl[W_BUILD]) return;
  const int k = blockIdx.x * TPB + threadIdx.x;
  if (k < N) order1[k] = order[k] + 1;
}

// Ascending ids inside every bin (bins hold a few particles): the order, and with it the slots
// and the lists, do not depend on the histogram's atomic ranks.
__global__ __launch_bounds__(TPB) void k_cl_sort_bins(int nbins, const int32_t *__restrict__ bin_start, int32_t *order,
                                                      const int32_t *ctl) {
  if (!ctl[W_BUILD]) return;
  const int b = blockIdx.x * TPB + threadIdx.x;
  if (b >= nbins) return;
  const int s = bin_start[b], e = bin_start[b + 1];
  for (int a = s + 1; a < e; ++a) {
    const int32_t key = order[a];
    int q = a - 1;
    while (q >= s && order[q] > key) {
      order[q + 1] = order[q];
      --q;
    }
    order[q + 1] = key;
  }
}

// One workgroup: super-clusters per column (ceil(n/64)), their offsets, and per super-cluster
// its first position in the order, size and column (_super_count_kernel!, _super_offsets_kernel!,
// _super_meta_kernel!, List.jl:259-303).
__global__ __launch_bounds__(1024) void k_cl_layout(int ncol, int nzb, const int32_t *__restrict__ bin_start,
                                                    int32_t *col_sc_off, int32_t *sc_first, int32_t *sc_n,
                                                    int32_t *sc_col, int32_t *ctl) {
  if (!ctl[W_BUILD]) return;
  __shared__ int32_t wt[16];
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  int32_t carry = 0;
  for (int base = 0; base < ncol; base += 1024) {
    const int c = base + threadIdx.x;
    const int n = c < ncol ? bin_start[(c + 1) * nzb] - bin_start[c * nzb] : 0;
    const int32_t k = (n + SUP - 1) / SUP;
    int32_t v = k;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      const int32_t u = __shfl_up(v, o, WAVE);
      if (lane >= o) v += u;
    }
    if (lane == WAVE - 1) wt[w] = v;
    __syncthreads();
    int32_t excl = carry + v - k;
    for (int q = 0; q < w; ++q) excl += wt[q];
    if (c < ncol) col_sc_off[c] = excl;
    int32_t tot = 0;
    for (int q = 0; q < 16; ++q) tot += wt[q];
    __syncthreads();
    carry += tot;
  }
  if (threadIdx.x == 0) {
    col_sc_off[ncol] = carry;
    ctl[W_NS] = carry;
  }
  for (int c = threadIdx.x; c < ncol; c += 1024) {
    const int S0 = col_sc_off[c];
    const int st = bin_start[c * nzb];
    const int n = bin_start[(c + 1) * nzb] - st;
    for (int k = 0; k * SUP < n; ++k) {
      sc_first[S0 + k] = st + k * SUP;
      sc_n[S0 + k] = min(SUP, n - k * SUP);
      sc_col[S0 + k] = c;
    }
  }
}

// One wave64 per super-cluster slot range (every S < NSmax): split at the x median, then each
// half at its y median, then each quarter at its z median (ranks with ties broken by position;
// _cluster_assign_kernel!, List.jl:305-364), write perm / slot_of, the cluster and super-cluster
// bounding boxes of the build positions in [0, L), and clear the j-cluster counters. Unused
// slots get perm -1; boxes of empty clusters have h.w = 0.
__global__ __launch_bounds__(WAVE) void k_cl_assign(const int32_t *ctl, const int32_t *__restrict__ order,
                                                    const int32_t *__restrict__ sc_first,
                                                    const int32_t *__restrict__ sc_n, const float *__restrict__ x,
                                                    const float *__restrict__ y, const float *__restrict__ z,
                                                    const float *__restrict__ geom, int32_t *perm, int32_t *slot_of,
                                                    float4 *cbc, float4 *cbh, float4 *sbc, float4 *sbh,
                                                    int32_t *jcnt, int32_t *jcur) {
  if (!ctl[W_BUILD]) return;
  __shared__ float sx[SUP], sy[SUP], sz[SUP];
  __shared__ int32_t shx[SUP], shy[SUP];
  __shared__ int32_t sperm[SUP];
  __shared__ float px[SUP], py[SUP], pz[SUP];
  __shared__ float4 bc[CL], bh[CL];
  const int S = blockIdx.x, t = threadIdx.x;
  if (t < CL) {
    jcnt[S * CL + t] = 0;
    jcur[S * CL + t] = 0;
  }
  sperm[t] = -1;
  const int n = S < ctl[W_NS] ? sc_n[S] : 0;
  const float Lx = geom[0], Ly = geom[1], Lz = geom[2];
  int i = -1;
  float xw = 0.0f, yw = 0.0f, zw = 0.0f;
  if (t < n) {
    i = order[sc_first[S] + t];
    xw = wrap01(x[i], 1.0f / Lx) * Lx;
    yw = wrap01(y[i], 1.0f / Ly) * Ly;
    zw = wrap01(z[i], 1.0f / Lz) * Lz;
    sx[t] = xw;
    sy[t] = yw;
    sz[t] = zw;
  }
  __syncthreads();
  int hx = 0, hy = 0;
  if (t < n) {
    int r = 0;
    for (int u = 0; u < n; ++u) r += (sx[u] < xw) | ((sx[u] == xw) & (u < t));
    hx = r >= ((n + 1) >> 1);
    shx[t] = hx;
  }
  __syncthreads();
  if (t < n) {
    int m = 0, r = 0;
    for (int u = 0; u < n; ++u) {
      if (shx[u] == hx) {
        ++m;
        r += (sy[u] < yw) | ((sy[u] == yw) & (u < t));
      }
    }
    hy = r >= ((m + 1) >> 1);
    shy[t] = hy;
  }
  __syncthreads();
  if (t < n) {
    int m = 0, r = 0;
    for (int u = 0; u < n; ++u) {
      if (shx[u] == hx && shy[u] == hy) {
        ++m;
        r += (sz[u] < zw) | ((sz[u] == zw) & (u < t));
      }
    }
    const int m0 = (m + 1) >> 1;
    const int hz = r >= m0;
    const int pos = r - hz * m0;  // < 8: a quarter of at most 16 splits into halves of at most 8
    const int loc = (hx * 4 + hy * 2 + hz) * CL + pos;
    sperm[loc] = i;
    px[loc] = xw;
    py[loc] = yw;
    pz[loc] = zw;
    slot_of[i] = S * SUP + loc;
  }
  __syncthreads();
  perm[S * SUP + t] = sperm[t];
  if (t < CL) {
    float lx = INFINITY, ly = INFINITY, lz = INFINITY, ux = -INFINITY, uy = -INFINITY, uz = -INFINITY;
    int cnt = 0;
    for (int k = 0; k < CL; ++k) {
      const int l = t * CL + k;
      if (sperm[l] >= 0) {
        ++cnt;
        lx = fminf(lx, px[l]); ux = fmaxf(ux, px[l]);
        ly = fminf(ly, py[l]); uy = fmaxf(uy, py[l]);
        lz = fminf(lz, pz[l]); uz = fmaxf(uz, pz[l]);
      }
    }
    float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f), h = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (cnt > 0) {
      c = make_float4(0.5f * (lx + ux), 0.5f * (ly + uy), 0.5f * (lz + uz), 0.0f);
      h = make_float4(0.5f * (ux - lx), 0.5f * (uy - ly), 0.5f * (uz - lz), 1.0f);
    }
    bc[t] = c;
    bh[t] = h;
    cbc[S * CL + t] = c;
    cbh[S * CL + t] = h;
  }
  __syncthreads();
  if (t == 0) {
    float lx = INFINITY, ly = INFINITY, lz = INFINITY, ux = -INFINITY, uy = -INFINITY, uz = -INFINITY;
    bool any = false;
    for (int k = 0; k < CL; ++k) {
      if (bh[k].w > 0.0f) {
        any = true;
        lx = fminf(lx, bc[k].x - bh[k].x); ux = fmaxf(ux, bc[k].x + bh[k].x);
        ly = fminf(ly, bc[k].y - bh[k].y); uy = fmaxf(uy, bc[k].y + bh[k].y);
        lz = fminf(lz, bc[k].z - bh[k].z); uz = fmaxf(uz, bc[k].z + bh[k].z);
      }
    }
    sbc[S] = any ? make_float4(0.5f * (lx + ux), 0.5f * (ly + uy), 0.5f * (lz + uz), 0.0f)
                 : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    sbh[S] = any ? make_float4(0.5f * (ux - lx), 0.5f * (uy - ly), 0.5f * (uz - lz), 1.0f)
                 : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

// ---------------------------------------------------------------- search
struct SearchArgs {
  const float *geom;
  const int32_t *sc_col, *col_sc_off;
  const float4 *cbc, *cbh, *sbc, *sbh;
  int32_t *ctl, *ne, *oent, *jcnt;
  int ncx, ncy, mx, my, allx, ally, ecap;
  float rl2;
};

// One wave64 per super-cluster: lane (k, jc) = (lane >> 3, lane & 7) tests j-cluster jc of the
// k-th of eight candidate super-clusters of a neighbouring column against the eight i-cluster
// boxes; entries with a non-empty mask are placed in lane order by a ballot prefix
// (_search_warp_kernel!, List.jl:366-440). Also counts the entries per j-cluster for the CSR.
__global__ __launch_bounds__(WAVE) void k_cl_search(const SearchArgs a) {
  int32_t *ctl = a.ctl;
  if (!ctl[W_BUILD]) return;
  const int S = blockIdx.x, lane = threadIdx.x;
  if (S >= ctl[W_NS]) {
    if (lane == 0) a.ne[S] = 0;
    return;
  }
  const float Lx = a.geom[0], Ly = a.geom[1], Lz = a.geom[2];
  float4 ic[CL], ih[CL];
#pragma unroll
  for (int c = 0; c < CL; ++c) {
    ic[c] = a.cbc[S * CL + c];
    ih[c] = a.cbh[S * CL + c];
  }
  const float4 sC = a.sbc[S], sH = a.sbh[S];
  const int col = a.sc_col[S];
  const int cx = col % a.ncx, cy = col / a.ncx;
  const int nxr = a.allx ? a.ncx : 2 * a.mx + 1, nyr = a.ally ? a.ncy : 2 * a.my + 1;
  const int x0 = a.allx ? 0 : cx - a.mx, y0 = a.ally ? 0 : cy - a.my;
  const int k = lane >> 3, jc = lane & 7;
  const size_t seg = (size_t)S * a.ecap;
  const unsigned long long below = (1ull << lane) - 1ull;
  int count = 0;
  for (int dy = 0; dy < nyr; ++dy) {
    const int yy = ((y0 + dy) % a.ncy + a.ncy) % a.ncy;
    for (int dx = 0; dx < nxr; ++dx) {
      const int xx = ((x0 + dx) % a.ncx + a.ncx) % a.ncx;
      const int cc = yy * a.ncx + xx;
      const int lo = max(a.col_sc_off[cc], S), hi = a.col_sc_off[cc + 1];
      for (int base = lo; base < hi; base += 8) {
        const int S2 = base + k;
        unsigned m = 0;
        int cj = 0;
        if (S2 < hi && box_gap2(sC, sH, a.sbc[S2], a.sbh[S2], Lx, Ly, Lz) < a.rl2) {
          cj = S2 * CL + jc;
          const float4 jC = a.cbc[cj], jH = a.cbh[cj];
          if (jH.w > 0.0f) {
#pragma unroll
            for (int c = 0; c < CL; ++c)
              if ((S2 > S || jc >= c) && ih[c].w > 0.0f && box_gap2(ic[c], ih[c], jC, jH, Lx, Ly, Lz) < a.rl2)
                m |= 1u << c;
          }
        }
        const unsigned long long bal = __ballot(m != 0);
        if (m != 0) {
          const int pos = count + __popcll(bal & below);
          if (pos < a.ecap) {
            a.oent[seg + pos] = (cj << 8) | (int)m;
            atomicAdd(&a.jcnt[cj], 1);
          }
        }
        count += __popcll(bal);
      }
    }
  }
  if (lane == 0) {
    a.ne[S] = min(count, a.ecap);
    if (count > a.ecap) {
      ctl[W_ERR] = 1;
      atomicMax(&ctl[W_REQ], count);
    }
  }
}

// Entries grouped by j-cluster (CSR over the outer list): positions from the scanned counts,
// then every group sorted ascending so the scatter adds them in entry order.
__global__ __launch_bounds__(TPB) void k_cl_jfill(const int32_t *ctl, const int32_t *__restrict__ ne,
                                                  const int32_t *__restrict__ oent, int ecap,
                                                  const int32_t *__restrict__ joff, int32_t *jcur, int32_t *jlist) {
  if (!ctl[W_BUILD]) return;
  const int S = blockIdx.x;
  if (S >= ctl[W_NS]) return;
  const int n = ne[S];
  const size_t seg = (size_t)S * ecap;
  for (int k = threadIdx.x; k < n; k += TPB) {
    const int cj = oent[seg + k] >> 8;
    jlist[joff[cj] + atomicAdd(&jcur[cj], 1)] = (int32_t)(seg + k);
  }
}

__global__ __launch_bounds__(TPB) void k_cl_jsort(const int32_t *ctl, const int32_t *__restrict__ joff, int32_t *jlist) {
  if (!ctl[W_BUILD]) return;
  const int c = blockIdx.x * TPB + threadIdx.x;
  if (c >= ctl[W_NS] * CL) return;
  const int s = joff[c], e = joff[c + 1];
  for (int a = s + 1; a < e; ++a) {
    const int32_t key = jlist[a];
    int q = a - 1;
    while (q >= s && jlist[q] > key) {
      jlist[q + 1] = jlist[q];
      --q;
    }
    jlist[q + 1] = key;
  }
}

// Coordinates and charges in slot order, NaN on padding (_pack_kernel!, List.jl:366-404), and
// the reference positions of a build / prune.
// Typed models: the slot's diameter (mixed sigma) or 0-based type (pair matrix), from the pair
// model's per-particle data; padding slots take 1.0 / type 0 (their NaN coordinates drop them).
__global__ __launch_bounds__(TPB) void k_cl_pack(int nslot, int N, const int32_t *__restrict__ perm,
                                                 const float *__restrict__ x, const float *__restrict__ y,
                                                 const float *__restrict__ z, const float *__restrict__ q,
                                                 float4 *packed, float *ref, const int32_t *ctl,
                                                 const float *__restrict__ sigp, const int32_t *__restrict__ typ,
                                                 float *saux, int32_t *styp) {
  const int s = blockIdx.x * TPB + threadIdx.x;
  if (s >= nslot) return;
  const int i = perm[s];
  if (sigp) saux[s] = i < 0 ? 1.0f : sigp[i];
  if (typ) styp[s] = i < 0 ? 0 : typ[i] - 1;
  if (i < 0) {
    packed[s] = make_float4(NAN, NAN, NAN, 0.0f);
    return;
  }
  const float xi = x[i], yi = y[i], zi = z[i];
  packed[s] = make_float4(xi, yi, zi, q ? q[i] : 0.0f);
  if (ctl[W_BUILD]) {
    ref[i] = xi;
    ref[(size_t)N + i] = yi;
    ref[(size_t)2 * N + i] = zi;
  }
  if (ctl[W_PRUNE]) {
    ref[(size_t)3 * N + i] = xi;
    ref[(size_t)4 * N + i] = yi;
    ref[(size_t)5 * N + i] = zi;
  }
}

__global__ void k_cl_set(int32_t *ctl, int word, int32_t v) { ctl[word] = v; }

// ---------------------------------------------------------------- force pass
struct TileArgs {
  float eps, sig2, rc2, eshift;     // LJ / WCA / soft
  float rfac2;                      // mixed sigma: (rc / sigma_ij)^2
  const float *saux;                // typed: slot diameters (mixed sigma)
  const int32_t *styp;              // typed: slot 0-based types (pair matrix)
  const float4 *tab;                // typed: type-pair table (sig2, eps, rc2, eshift)
  int ntypes;
  float cC, crc2, csx, ceshift;     // real-space Coulomb
  float pe[MAXC], pf[MAXC];
  const float4 *packed;
  const int32_t *ctl;
  const float *geom;
  const int32_t *ne, *oent;
  int32_t *went, *wjb, *nw;
  uint8_t *keep;
  float *jbuf, *islot;
  int ecap, nslot, waves, splits;
  float rp2, rmax2;
};

// The tile kernel; PRUNE: the pass over the outer list that also writes the working list.
template <int MODEL, bool COUL, bool E, bool PRUNE, int NP>
__global__ __launch_bounds__(WAVE *MAXW) void k_cl_tiles(const TileArgs a) {
  const int32_t *ctl = a.ctl;
  if (ctl[W_ERR] != 0 || (ctl[W_PRUNE] != 0) != PRUNE) return;
  const int S = blockIdx.x / a.splits, part = blockIdx.x % a.splits;
  if (S >= ctl[W_NS]) return;
  constexpr int NV = E ? KV : 3;
  const int W = a.waves;
  const int nchunk = W * a.splits;
  const int tid = threadIdx.x, w = tid / WAVE, lane = tid & (WAVE - 1);
  const int ti = lane & 7, tj = lane >> 3;
  const float Lx = a.geom[0], Ly = a.geom[1], Lz = a.geom[2];
  const float hx = 0.5f * Lx, hy = 0.5f * Ly, hz = 0.5f * Lz;
  float pe[NP], pf[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    pe[k] = a.pe[k];
    pf[k] = a.pf[k];
  }
  float4 xi[CL];
#pragma unroll
  for (int c = 0; c < CL; ++c) xi[c] = a.packed[(size_t)S * SUP + c * CL + ti];
  // typed models: the i atoms' diameters or type rows (lane ti of each cluster c)
  constexpr bool TYPED = typed_model(MODEL);
  constexpr bool MIXED = MODEL == PDX_PAIR_MIXED_LJ || MODEL == PDX_PAIR_MIXED_WCA;
  float cai[TYPED ? CL : 1];
  int tai[TYPED ? CL : 1];
  if constexpr (TYPED) {
#pragma unroll
    for (int c = 0; c < CL; ++c) {
      const size_t sl = (size_t)S * SUP + c * CL + ti;
      cai[c] = MIXED ? a.saux[sl] : 0.0f;
      tai[c] = MIXED ? 0 : a.styp[sl] * a.ntypes;
    }
  }
  float si[CL][NV];
#pragma unroll
  for (int c = 0; c < CL; ++c)
#pragma unroll
    for (int k = 0; k < NV; ++k) si[c][k] = 0.0f;
  // the entries of S in W x splits contiguous chunks, one per wave
  const int n = a.ne[S];
  const int chunk = part * W + w;
  const int e0 = (n * chunk) / nchunk;
  const int e1 = PRUNE ? (n * (chunk + 1)) / nchunk : e0 + a.nw[S * nchunk + chunk];
  const size_t seg = (size_t)S * a.ecap;
  int wpos = e0;
  // the next entry and its j-atom are loaded one iteration ahead (the loads depend on each
  // other, and an iteration of up to eight tiles covers both latencies)
  int ent_n = 0, jb_n = 0;
  float4 xj_n = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float caj_n = 0.0f;
  int taj_n = 0;
  if (e0 < e1) {
    ent_n = PRUNE ? a.oent[seg + e0] : a.went[seg + e0];
    jb_n = PRUNE ? (int)(seg + e0) : a.wjb[seg + e0];
    xj_n = a.packed[(size_t)(ent_n >> 8) * CL + tj];
    if constexpr (TYPED) {
      if (MIXED) caj_n = a.saux[(size_t)(ent_n >> 8) * CL + tj];
      else taj_n = a.styp[(size_t)(ent_n >> 8) * CL + tj];
    }
  }
  for (int e = e0; e < e1; ++e) {
    const int ent = ent_n, jb = jb_n;
    const float4 xj = xj_n;
    const float caj = caj_n;
    const int taj = taj_n;
    if (e + 1 < e1) {
      ent_n = PRUNE ? a.oent[seg + e + 1] : a.went[seg + e + 1];
      jb_n = PRUNE ? (int)(seg + e + 1) : a.wjb[seg + e + 1];
      xj_n = a.packed[(size_t)(ent_n >> 8) * CL + tj];
      if constexpr (TYPED) {
        if (MIXED) caj_n = a.saux[(size_t)(ent_n >> 8) * CL + tj];
        else taj_n = a.styp[(size_t)(ent_n >> 8) * CL + tj];
      }
    }
    const int cj = ent >> 8;
    const unsigned mask = (unsigned)ent & 255u;
    const int cdiag = cj - S * CL;
    float sj[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) sj[k] = 0.0f;
    unsigned pm = 0;
#pragma unroll
    for (int c = 0; c < CL; ++c) {
      if (!((mask >> c) & 1u)) continue;
      float dx = xi[c].x - xj.x, dy = xi[c].y - xj.y, dz = xi[c].z - xj.z;
      if (__ballot(fabsf(dx) > hx || fabsf(dy) > hy || fabsf(dz) > hz) != 0) {
        dx = periodic_delta(xi[c].x, xj.x, hx, Lx);
        dy = periodic_delta(xi[c].y, xj.y, hy, Ly);
        dz = periodic_delta(xi[c].z, xj.z, hz, Lz);
      }
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      const bool ok = (r2 < a.rmax2) && (c != cdiag || tj > ti);
      if (PRUNE && __ballot(ok && r2 < a.rp2) != 0) pm |= 1u << c;
      float f, u;
      if constexpr (TYPED)
        pair_terms_c<MODEL, COUL, E, NP>(a, typed_coef<MODEL>(a, cai[c], caj, tai[c], taj), pe, pf, xi[c].w, xj.w,
                                         r2, f, u);
      else
        pair_terms<MODEL, COUL, E, NP>(a, pe, pf, xi[c].w, xj.w, r2, f, u);
      f = ok ? f : 0.0f;
      dx = ok ? dx : 0.0f;
      dy = ok ? dy : 0.0f;
      dz = ok ? dz : 0.0f;
      const float gx = f * dx, gy = f * dy, gz = f * dz;
      si[c][0] += gx; si[c][1] += gy; si[c][2] += gz;
      sj[0] -= gx; sj[1] -= gy; sj[2] -= gz;
      if (E) {
        u = ok ? u : 0.0f;
        const float v0 = dx * gx, v1 = dy * gy, v2 = dz * gz, v3 = dx * gy, v4 = dx * gz, v5 = dy * gz;
        si[c][3] += u;
        si[c][4] += v0; si[c][5] += v1; si[c][6] += v2;
        si[c][7] += v3; si[c][8] += v4; si[c][9] += v5;
        sj[3] += u;
        sj[4] += v0; sj[5] += v1; sj[6] += v2;
        sj[7] += v3; sj[8] += v4; sj[9] += v5;
      }
    }
    if (PRUNE) {
      if (lane == 0) a.keep[seg + e] = pm != 0;
      if (pm == 0) continue;  // every pair beyond rmax + prune_buffer: nothing to store
      if (lane == 0) {
        a.went[seg + wpos] = (cj << 8) | (int)pm;
        a.wjb[seg + wpos] = jb;
      }
      ++wpos;
    }
    // j-sums over the eight ti lanes; lane (ti, tj) stores value ti (and ti + 8) of atom tj
#pragma unroll
    for (int k = 0; k < NV; ++k) sj[k] = sum8_dpp(sj[k]);
    float v = sj[0];
#pragma unroll
    for (int k = 1; k < (NV < 8 ? NV : 8); ++k) v = ti == k ? sj[k] : v;
    float *row = a.jbuf + (size_t)jb * (KV * CL);
    if (ti < NV) row[ti * CL + tj] = v;
    if constexpr (NV > 8) {
      if (ti < NV - 8) row[(ti + 8) * CL + tj] = ti == 0 ? sj[8] : sj[NV - 1];
    }
  }
  if (PRUNE && lane == 0) a.nw[S * nchunk + chunk] = wpos - e0;
  // i-sums over the eight tj lanes, then over the waves in LDS in wave order
  __shared__ float sh[MAXW][KV][SUP];
#pragma unroll
  for (int c = 0; c < CL; ++c) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      float s = si[c][k];
      s += __shfl_xor(s, 8, WAVE);
      s += __shfl_xor(s, 16, WAVE);
      s += __shfl_xor(s, 32, WAVE);
      if (tj == c) sh[w][k][c * CL + ti] = s;
    }
  }
  __syncthreads();
  for (int idx = tid; idx < NV * SUP; idx += WAVE * W) {
    const int k = idx / SUP, l = idx % SUP;
    float s = sh[0][k][l];
    for (int q = 1; q < W; ++q) s += sh[q][k][l];
    a.islot[(size_t)(part * KV + k) * a.nslot + (size_t)S * SUP + l] = s;
  }
}

// Working-list CSR: the outer CSR filtered by the pruning pass's keep flags, in the same order.
__global__ __launch_bounds__(TPB) void k_cl_jfilter(const int32_t *ctl, const int32_t *__restrict__ joff,
                                                    const int32_t *__restrict__ jlist, const uint8_t *__restrict__ keep,
                                                    int32_t *wjlist, int32_t *wjcnt) {
  if (ctl[W_ERR] != 0 || !ctl[W_PRUNE]) return;
  const int c = blockIdx.x * TPB + threadIdx.x;
  if (c >= ctl[W_NS] * CL) return;
  const int s = joff[c], e = joff[c + 1];
  int n = 0;
  for (int p = s; p < e; ++p) {
    const int32_t jb = jlist[p];
    if (keep[jb]) wjlist[s + n++] = jb;
  }
  wjcnt[c] = n;
}

struct ScatterArgs {
  int32_t *ctl;
  const int32_t *perm, *slot_of, *joff, *wjcnt, *wjlist;
  const float *jbuf, *islot, *geom;
  const float4 *packed;
  const int32_t *exo, *exj;
  const float *exs;
  float *fx, *fy, *fz, *Epot, *V;
  double *partial;
  int N, nslot, add, totals, splits;
  float cC, crc2, ceshift, halfself;
};

// Slot sums -> storage order (_scatter_kernel!, Kernels.jl:207-240): i-sum plus the j-partials
// of the entries naming the slot's cluster, the Coulomb exceptions of the particle (+ C q_i q_j
// (s - 1) / r at the minimum image, no cutoff, no potential shift; _scaled_pairs_kernel!,
// Kernels.jl:307-339) and the Coulomb self energy; halves of energies and virials; per-block
// FP64 partials of the totals. Ends the pruning step (clears the prune flag).
template <bool COUL, bool E>
__global__ __launch_bounds__(TPB) void k_cl_scatter(const ScatterArgs a) {
  constexpr int NV = E ? KV : 3;
  const bool valid = a.ctl[W_ERR] == 0;
  const int s = blockIdx.x * TPB + threadIdx.x;
  float ei = 0.0f, wv[6] = {0, 0, 0, 0, 0, 0};
  if (valid && s < a.ctl[W_NS] * SUP) {
    const int i = a.perm[s];
    if (i >= 0) {
      float acc[NV];
#pragma unroll
      for (int k = 0; k < NV; ++k) acc[k] = a.islot[(size_t)k * a.nslot + s];
      for (int part = 1; part < a.splits; ++part) {
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] += a.islot[(size_t)(part * KV + k) * a.nslot + s];
      }
      const int c = s >> 3, t = s & 7;
      const int p0 = a.joff[c], p1 = p0 + a.wjcnt[c];
      for (int p = p0; p < p1; ++p) {
        const float *row = a.jbuf + (size_t)a.wjlist[p] * (KV * CL) + t;
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] += row[k * CL];
      }
      const float4 pi = a.packed[s];
      if (COUL && a.exo && pi.w != 0.0f) {
        const float Lx = a.geom[0], Ly = a.geom[1], Lz = a.geom[2];
        for (int p = a.exo[i]; p < a.exo[i + 1]; ++p) {
          const float4 pj = a.packed[a.slot_of[a.exj[p]]];
          const float sc = a.cC * pi.w * pj.w * (a.exs[p] - 1.0f);
          const float dx = periodic_delta(pi.x, pj.x, 0.5f * Lx, Lx);
          const float dy = periodic_delta(pi.y, pj.y, 0.5f * Ly, Ly);
          const float dz = periodic_delta(pi.z, pj.z, 0.5f * Lz, Lz);
          const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          const float invr = __builtin_amdgcn_rsqf(r2);
          const float f = sc * invr * invr * invr;
          const float gx = f * dx, gy = f * dy, gz = f * dz;
          acc[0] += gx; acc[1] += gy; acc[2] += gz;
          if (E) {
            acc[3] += sc * invr + (r2 < a.crc2 ? a.cC * pi.w * pj.w * a.ceshift : 0.0f);
            acc[4] += dx * gx; acc[5] += dy * gy; acc[6] += dz * gz;
            acc[7] += dx * gy; acc[8] += dx * gz; acc[9] += dy * gz;
          }
        }
      }
      if (a.add) {
        a.fx[i] += acc[0]; a.fy[i] += acc[1]; a.fz[i] += acc[2];
      } else {
        a.fx[i] = acc[0]; a.fy[i] = acc[1]; a.fz[i] = acc[2];
      }
      if (E) {
        ei = 0.5f * acc[3];
        if (COUL) ei -= a.halfself * pi.w * pi.w;
#pragma unroll
        for (int k = 0; k < 6; ++k) wv[k] = 0.5f * acc[4 + k];
        if (a.Epot) a.Epot[i] = a.add ? a.Epot[i] + ei : ei;
        if (a.V) {
#pragma unroll
          for (int k = 0; k < 6; ++k) {
            float *q = a.V + (size_t)k * a.N + i;
            *q = a.add ? *q + wv[k] : wv[k];
          }
        }
      }
    }
  }
  if (E && a.totals) {
    const double d[7] = {(double)ei, (double)wv[0], (double)wv[1], (double)wv[2],
                         (double)wv[3], (double)wv[4], (double)wv[5]};
    block_partials<7>(d, a.partial);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && a.ctl[W_PRUNE] != 0) a.ctl[W_PRUNE] = 0;
}

int blocks_of(int64_t n, int t) { return (int)std::max<int64_t>(1, (n + t - 1) / t); }

}  // namespace

// ---------------------------------------------------------------- handle
struct pdx_cluster_s {
  int32_t N = 0;
  double box[3] = {0, 0, 0};
  pdx_pair_t model = nullptr;
  double rmax = 0, skin = 0, pb = 0, rlist = 0;
  int ncx = 1, ncy = 1, nzb = 1, ncol = 1, nbins = 1, mx = 1, my = 1, allx = 0, ally = 0;
  int NSmax = 0, nslot = 0, ecap = 0, waves = 4, splits = 1;
  int dblocks = 0, sblocks = 0;
  bool updated = false;
  DevBuf<int32_t> ctl;
  DevBuf<float> geom, ref, dpart;
  DevBuf<int32_t> bin_of, bin_rank, bin_cnt, bin_start, order, order1;
  DevBuf<int32_t> col_sc_off, sc_first, sc_n, sc_col;
  DevBuf<int32_t> perm, slot_of;
  DevBuf<float4> cbc, cbh, sbc, sbh, packed;
  DevBuf<int32_t> ne, oent, went, wjb, nw;
  DevBuf<uint8_t> keep;
  DevBuf<int32_t> jcnt, jcur, joff, jlist, wjlist, wjcnt;
  DevBuf<float> jbuf, islot;
  DevBuf<double> partial;
  DevBuf<float> saux;    // typed models: diameter per slot (mixed sigma)
  DevBuf<int32_t> styp;  // typed models: 0-based type per slot (pair matrix)
};

namespace {

int alloc_entries(pdx_cluster_s *c, int ecap) {
  const size_t ne = (size_t)c->NSmax * ecap;
  if (ne > ((size_t)1 << 30))
    return fail(PDX_ERR_ARGUMENT, "cluster list: entry storage beyond 2^30 entries");
  int rc = 0;
  if ((rc = c->oent.alloc(ne)) || (rc = c->went.alloc(ne)) || (rc = c->wjb.alloc(ne)) || (rc = c->keep.alloc(ne)) ||
      (rc = c->jlist.alloc(ne)) || (rc = c->wjlist.alloc(ne)) || (rc = c->jbuf.alloc(ne * KV * CL)))
    return rc;
  c->ecap = ecap;
  return PDX_OK;
}

}  // namespace

extern "C" int pdx_cluster_create(int32_t n, const double box[3], pdx_pair_t model, double skin, double prune_buffer,
                                  int32_t entry_capacity, int32_t waves, pdx_cluster_t *out) {
  if (!out || !box || !model) return fail(PDX_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  if (n < 1) return fail(PDX_ERR_ARGUMENT, "cluster list: n must be >= 1");
  if (model->accumulate != PDX_ACC_FP32 || model->bonds)
    return fail(PDX_ERR_ARGUMENT, "cluster pairs use FP32 sums without bonds");
  if (model->spec.model == PDX_PAIR_NONE && !model->spec.coulomb)
    return fail(PDX_ERR_ARGUMENT, "cluster pairs need a pair model or Coulomb");
  if (!(skin >= 0.0) || !(prune_buffer >= 0.0) || prune_buffer > skin)
    return fail(PDX_ERR_ARGUMENT, "cluster pairs need 0 <= prune_buffer <= skin.");
  if (waves == 0) waves = 4;
  if (waves != 1 && waves != 2 && waves != 4 && waves != 8)
    return fail(PDX_ERR_ARGUMENT, "cluster pairs: waves must be 1, 2, 4 or 8");
  for (int d = 0; d < 3; ++d)
    if (!(box[d] > 0.0)) return fail(PDX_ERR_ARGUMENT, "box lengths must be positive");
  double rmax = 0.0;
  if (typed_model(model->spec.model)) {
    if (model->typed_n == 0) return fail(PDX_ERR_STATE, "typed pair model without its per-particle data");
    rmax = model->typed_rmax;
  } else if (model->spec.model != PDX_PAIR_NONE) {
    rmax = std::sqrt((double)model->rc2);
  }
  if (model->spec.coulomb) rmax = std::max(rmax, std::sqrt((double)model->crc2));
  const double rlist = rmax + skin;
  if (std::min(box[0], std::min(box[1], box[2])) <= 2.0 * rlist)
    return fail(PDX_ERR_ARGUMENT, "cluster pairs need every box side > 2 (cutoff + skin)");
  auto *c = new pdx_cluster_s();
  c->N = n;
  for (int d = 0; d < 3; ++d) c->box[d] = box[d];
  c->model = model;
  c->rmax = rmax;
  c->skin = skin;
  c->pb = prune_buffer;
  c->rlist = rlist;
  c->waves = waves;
  // about 2048 waves over the chip: small systems split each super-cluster's entries over
  // several workgroups
  {
    const int64_t ns = (n + SUP - 1) / SUP;
    int sp = 1;
    while (sp < MAXSPLIT && ns * waves * sp * 2 <= 2048) sp <<= 1;
    c->splits = sp;
  }
  // columns of about one super-cluster edge, z bins of about four particles
  const double V = box[0] * box[1] * box[2];
  const double rho = n / V;
  const double edge = std::cbrt(SUP / rho);
  c->ncx = std::max(1, (int)std::floor(box[0] / edge));
  c->ncy = std::max(1, (int)std::floor(box[1] / edge));
  c->ncol = c->ncx * c->ncy;
  c->nzb = std::max(1, (int)std::min<double>(n / (c->ncol * 4.0), 1 << 20));
  while ((int64_t)c->ncol * c->nzb > (int64_t)1 << 26) c->nzb = std::max(1, c->nzb / 2);
  c->nbins = c->ncol * c->nzb;
  const double wx = box[0] / c->ncx, wy = box[1] / c->ncy;
  c->mx = (int)std::ceil(rlist / wx);
  c->my = (int)std::ceil(rlist / wy);
  c->allx = 2 * c->mx + 1 >= c->ncx;
  c->ally = 2 * c->my + 1 >= c->ncy;
  c->NSmax = (n + SUP - 1) / SUP + c->ncol;
  c->nslot = SUP * c->NSmax;
  int ecap = entry_capacity;
  if (ecap <= 0) {
    // clusters within rlist of a super-cluster box (a column slab), with room: the half list
    // of Newton-3 needs about half of them
    const double ez = SUP / (rho * wx * wy);
    const double ce = edge / 2;
    const double vol = (wx + 2 * rlist + ce) * (wy + 2 * rlist + ce) * (ez + 2 * rlist + ce);
    ecap = (int)std::ceil(0.75 * vol * rho / CL) + 32;
  }
  ecap = std::max(8, std::min(ecap, CL * c->NSmax));
  const int N = n;
  c->dblocks = blocks_of(N, TPB);
  c->sblocks = blocks_of(c->nslot, TPB);
  int rc = 0;
  if ((rc = c->ctl.alloc(W_LEN)) || (rc = c->geom.alloc(3)) || (rc = c->ref.alloc((size_t)6 * N)) ||
      (rc = c->dpart.alloc((size_t)2 * c->dblocks)) || (rc = c->bin_of.alloc(N)) || (rc = c->bin_rank.alloc(N)) ||
      (rc = c->bin_cnt.alloc(c->nbins)) || (rc = c->bin_start.alloc((size_t)c->nbins + 1)) ||
      (rc = c->order.alloc(N)) || (rc = c->order1.alloc(N)) || (rc = c->col_sc_off.alloc((size_t)c->ncol + 1)) ||
      (rc = c->sc_first.alloc(c->NSmax)) || (rc = c->sc_n.alloc(c->NSmax)) || (rc = c->sc_col.alloc(c->NSmax)) ||
      (rc = c->perm.alloc(c->nslot)) || (rc = c->slot_of.alloc(N)) || (rc = c->cbc.alloc((size_t)CL * c->NSmax)) ||
      (rc = c->cbh.alloc((size_t)CL * c->NSmax)) || (rc = c->sbc.alloc(c->NSmax)) || (rc = c->sbh.alloc(c->NSmax)) ||
      (rc = c->packed.alloc(c->nslot)) || (rc = c->ne.alloc(c->NSmax)) ||
      (rc = c->nw.alloc((size_t)MAXCHUNK * c->NSmax)) || (rc = c->jcnt.alloc((size_t)CL * c->NSmax)) ||
      (rc = c->jcur.alloc((size_t)CL * c->NSmax)) || (rc = c->joff.alloc((size_t)CL * c->NSmax + 1)) ||
      (rc = c->wjcnt.alloc((size_t)CL * c->NSmax)) || (rc = c->islot.alloc((size_t)c->splits * KV * c->nslot)) ||
      (rc = c->partial.alloc((size_t)7 * c->sblocks)) || (rc = alloc_entries(c, ecap))) {
    delete c;
    return rc;
  }
  int32_t h[W_LEN] = {0};
  h[W_REQUEST] = 1;
  const float g[3] = {(float)box[0], (float)box[1], (float)box[2]};
  hipError_t e = hipMemcpy(c->ctl.p, h, sizeof(h), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->geom.p, g, sizeof(g), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(c->perm.p, 0xFF, sizeof(int32_t) * c->nslot);
  if (e == hipSuccess) e = hipMemset(c->ref.p, 0, sizeof(float) * 6 * (size_t)N);
  if (e == hipSuccess) e = hipMemset(c->ne.p, 0, sizeof(int32_t) * c->NSmax);
  if (e == hipSuccess) e = hipMemset(c->nw.p, 0, sizeof(int32_t) * MAXCHUNK * c->NSmax);
  if (e != hipSuccess) {
    delete c;
    return fail(PDX_ERR_DEVICE, std::string("cluster list set-up failed: ") + hipGetErrorString(e));
  }
  *out = c;
  return PDX_OK;
}

extern "C" int pdx_cluster_destroy(pdx_cluster_t c) {
  delete c;
  return PDX_OK;
}

extern "C" int pdx_cluster_info(pdx_cluster_t c, int64_t out[10]) {
  if (!c || !out) return fail(PDX_ERR_ARGUMENT, "null argument");
  out[0] = c->N; out[1] = c->ncx; out[2] = c->ncy; out[3] = c->nzb;
  out[4] = c->NSmax; out[5] = c->nslot; out[6] = c->ecap; out[7] = c->waves;
  out[8] = c->splits; out[9] = (int64_t)c->waves * c->splits;
  return PDX_OK;
}

extern "C" int pdx_cluster_set_launch(pdx_cluster_t c, int32_t waves, int32_t splits) {
  if (!c) return fail(PDX_ERR_ARGUMENT, "null cluster list");
  if (waves != 1 && waves != 2 && waves != 4 && waves != 8)
    return fail(PDX_ERR_ARGUMENT, "cluster pairs: waves must be 1, 2, 4 or 8");
  if (splits < 1 || splits > MAXSPLIT || (splits & (splits - 1)))
    return fail(PDX_ERR_ARGUMENT, "cluster pairs: splits must be 1, 2, 4 or 8");
  if (waves == c->waves && splits == c->splits) return PDX_OK;
  // the working list is laid out per chunk: the next update rebuilds and the pass prunes again
  PDX_HIP(hipDeviceSynchronize());
  if ((size_t)splits * KV * c->nslot > c->islot.n) {
    int rc = c->islot.alloc((size_t)splits * KV * c->nslot);
    if (rc) return rc;
  }
  c->waves = waves;
  c->splits = splits;
  int32_t h[W_LEN];
  PDX_HIP(hipMemcpy(h, c->ctl.p, sizeof(h), hipMemcpyDeviceToHost));
  h[W_REQUEST] = 1;
  PDX_HIP(hipMemcpy(c->ctl.p, h, sizeof(h), hipMemcpyHostToDevice));
  return PDX_OK;
}

extern "C" int pdx_cluster_arrays_get(pdx_cluster_t c, pdx_cluster_arrays *out) {
  if (!c || !out) return fail(PDX_ERR_ARGUMENT, "null argument");
  out->perm = c->perm.p;
  out->slot_of = c->slot_of.p;
  out->packed = reinterpret_cast<float *>(c->packed.p);
  out->n_outer = c->ne.p;
  out->outer = c->oent.p;
  out->control = c->ctl.p;
  out->order = c->order1.p;
  out->working = c->went.p;
  out->n_working = c->nw.p;
  return PDX_OK;
}

extern "C" int pdx_cluster_update(pdx_cluster_t c, const float *x, const float *y, const float *z, const float *q,
                                  pdx_stream_t stream) {
  if (!c || !x || !y || !z) return fail(PDX_ERR_ARGUMENT, "null argument");
  hipStream_t st = as_stream(stream);
  const int N = c->N;
  const float hb = (float)(0.5 * c->skin), hp = (float)(0.5 * c->pb);
  const float ib2 = hb > 0.0f ? 1.0f / (hb * hb) : INFINITY;
  const float ip2 = hp > 0.0f ? 1.0f / (hp * hp) : INFINITY;
  int32_t *ctl = c->ctl.p;
  hipLaunchKernelGGL(k_cl_decide, dim3(c->dblocks), dim3(TPB), 0, st, N, x, y, z, (const float *)c->ref.p,
                     (const float *)c->geom.p, ib2, ip2, c->dpart.p, ctl);
  hipLaunchKernelGGL(k_cl_decide_final, dim3(1), dim3(TPB), 0, st, c->dblocks, (const float *)c->dpart.p, ctl,
                     c->nbins, c->bin_cnt.p);
  hipLaunchKernelGGL(k_cl_bin, dim3(c->dblocks), dim3(TPB), 0, st, N, x, y, z, (const float *)c->geom.p, c->ncx,
                     c->ncy, c->nzb, c->bin_cnt.p, c->bin_of.p, c->bin_rank.p, (const int32_t *)ctl);
  hipLaunchKernelGGL(k_cl_scan, dim3(1), dim3(1024), 0, st, c->nbins, (const int32_t *)c->bin_cnt.p,
                     c->bin_start.p, (const int32_t *)ctl, (int)W_BUILD);
  hipLaunchKernelGGL(k_cl_order, dim3(c->dblocks), dim3(TPB), 0, st, N, (const int32_t *)c->bin_of.p,
                     (const int32_t *)c->bin_rank.p, (const int32_t *)c->bin_start.p, c->order.p,
                     (const int32_t *)ctl);
  hipLaunchKernelGGL(k_cl_sort_bins, dim3(blocks_of(c->nbins, TPB)), dim3(TPB), 0, st, c->nbins,
                     (const int32_t *)c->bin_start.p, c->order.p, (const int32_t *)ctl);
  hipLaunchKernelGGL(k_cl_order1, dim3(c->dblocks), dim3(TPB), 0, st, N, (const int32_t *)c->order.p, c->order1.p,
                     (const int32_t *)ctl);
  hipLaunchKernelGGL(k_cl_layout, dim3(1), dim3(1024), 0, st, c->ncol, c->nzb, (const int32_t *)c->bin_start.p,
                     c->col_sc_off.p, c->sc_first.p, c->sc_n.p, c->sc_col.p, ctl);
  hipLaunchKernelGGL(k_cl_assign, dim3(c->NSmax), dim3(WAVE), 0, st, (const int32_t *)ctl,
                     (const int32_t *)c->order.p, (const int32_t *)c->sc_first.p, (const int32_t *)c->sc_n.p, x, y,
                     z, (const float *)c->geom.p, c->perm.p, c->slot_of.p, c->cbc.p, c->cbh.p, c->sbc.p, c->sbh.p,
                     c->jcnt.p, c->jcur.p);
  SearchArgs sa{};
  sa.geom = c->geom.p;
  sa.sc_col = c->sc_col.p;
  sa.col_sc_off = c->col_sc_off.p;
  sa.cbc = c->cbc.p; sa.cbh = c->cbh.p; sa.sbc = c->sbc.p; sa.sbh = c->sbh.p;
  sa.ctl = ctl; sa.ne = c->ne.p; sa.oent = c->oent.p; sa.jcnt = c->jcnt.p;
  sa.ncx = c->ncx; sa.ncy = c->ncy; sa.mx = c->mx; sa.my = c->my; sa.allx = c->allx; sa.ally = c->ally;
  sa.ecap = c->ecap;
  sa.rl2 = (float)(c->rlist * c->rlist);
  hipLaunchKernelGGL(k_cl_search, dim3(c->NSmax), dim3(WAVE), 0, st, sa);
  hipLaunchKernelGGL(k_cl_scan, dim3(1), dim3(1024), 0, st, CL * c->NSmax, (const int32_t *)c->jcnt.p, c->joff.p,
                     (const int32_t *)ctl, (int)W_BUILD);
  hipLaunchKernelGGL(k_cl_jfill, dim3(c->NSmax), dim3(TPB), 0, st, (const int32_t *)ctl, (const int32_t *)c->ne.p,
                     (const int32_t *)c->oent.p, c->ecap, (const int32_t *)c->joff.p, c->jcur.p, c->jlist.p);
  hipLaunchKernelGGL(k_cl_jsort, dim3(blocks_of((int64_t)CL * c->NSmax, TPB)), dim3(TPB), 0, st,
                     (const int32_t *)ctl, (const int32_t *)c->joff.p, c->jlist.p);
  const pdx_pair_s *pm = c->model;
  const bool mixed = pm->spec.model == PDX_PAIR_MIXED_LJ || pm->spec.model == PDX_PAIR_MIXED_WCA;
  const bool matrix = pm->spec.model == PDX_PAIR_MATRIX_LJ || pm->spec.model == PDX_PAIR_MATRIX_WCA;
  if ((mixed || matrix) && pm->typed_n != N)
    return fail(PDX_ERR_DIMENSION, "typed pair data and cluster list particle counts differ");
  if (mixed && (int64_t)c->saux.n < c->nslot) {
    int rc = c->saux.alloc(c->nslot);
    if (rc) return rc;
  }
  if (matrix && (int64_t)c->styp.n < c->nslot) {
    int rc = c->styp.alloc(c->nslot);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_cl_pack, dim3(c->sblocks), dim3(TPB), 0, st, c->nslot, N, (const int32_t *)c->perm.p, x, y,
                     z, q, c->packed.p, c->ref.p, (const int32_t *)ctl, mixed ? (const float *)pm->sigp.p : nullptr,
                     matrix ? (const int32_t *)pm->typ.p : nullptr, c->saux.p, c->styp.p);
  PDX_LAUNCH_CHECK();
  c->updated = true;
  return PDX_OK;
}

namespace {

// The pruning pass, the working-list CSR filter, then the plain pass: each returns at once
// unless the device's prune flag selects it.
template <int MODEL>
void launch_tiles_model(bool coul, bool energy, int np, hipStream_t st, int grid, int threads, const TileArgs &a,
                        const pdx_cluster_s *c) {
  auto filter = [&]() {
    hipLaunchKernelGGL(k_cl_jfilter, dim3(blocks_of((int64_t)CL * c->NSmax, TPB)), dim3(TPB), 0, st,
                       (const int32_t *)c->ctl.p, (const int32_t *)c->joff.p, (const int32_t *)c->jlist.p,
                       (const uint8_t *)c->keep.p, c->wjlist.p, c->wjcnt.p);
  };
  if (coul) {
    if (np <= 16) {
      if (energy) {
        hipLaunchKernelGGL((k_cl_tiles<MODEL, true, true, true, 16>), dim3(grid), dim3(threads), 0, st, a);
        filter();
        hipLaunchKernelGGL((k_cl_tiles<MODEL, true, true, false, 16>), dim3(grid), dim3(threads), 0, st, a);
      } else {
        hipLaunchKernelGGL((k_cl_tiles<MODEL, true, false, true, 16>), dim3(grid), dim3(threads), 0, st, a);
        filter();
        hipLaunchKernelGGL((k_cl_tiles<MODEL, true, false, false, 16>), dim3(grid), dim3(threads), 0, st, a);
      }
    } else {
      if (energy) {
        hipLaunchKernelGGL((k_cl_tiles<MODEL, true, true, true, 24>), dim3(grid), dim3(threads), 0, st, a);
        filter();
        hipLaunchKernelGGL((k_cl_tiles<MODEL, true, true, false, 24>), dim3(grid), dim3(threads), 0, st, a);
      } else {
        hipLaunchKernelGGL((k_cl_tiles<MODEL, true, false, true, 24>), dim3(grid), dim3(threads), 0, st, a);
        filter();
        hipLaunchKernelGGL((k_cl_tiles<MODEL, true, false, false, 24>), dim3(grid), dim3(threads), 0, st, a);
      }
    }
  } else {
    if (energy) {
      hipLaunchKernelGGL((k_cl_tiles<MODEL, false, true, true, 1>), dim3(grid), dim3(threads), 0, st, a);
      filter();
      hipLaunchKernelGGL((k_cl_tiles<MODEL, false, true, false, 1>), dim3(grid), dim3(threads), 0, st, a);
    } else {
      hipLaunchKernelGGL((k_cl_tiles<MODEL, false, false, true, 1>), dim3(grid), dim3(threads), 0, st, a);
      filter();
      hipLaunchKernelGGL((k_cl_tiles<MODEL, false, false, false, 1>), dim3(grid), dim3(threads), 0, st, a);
    }
  }
}

}  // namespace

extern "C" int pdx_cluster_forces(pdx_cluster_t c, float *fx, float *fy, float *fz, float *Epot, float *V,
                                  double *totals, int32_t flags, pdx_stream_t stream) {
  if (!c) return fail(PDX_ERR_ARGUMENT, "null cluster list");
  if (!fx || !fy || !fz) return fail(PDX_ERR_ARGUMENT, "null force array");
  if (!c->updated) return fail(PDX_ERR_STATE, "cluster list: pdx_cluster_update must run before the forces");
  const pdx_pair_s *p = c->model;
  const bool energy = (flags & PDX_ENERGY) != 0;
  const bool coul = p->spec.coulomb != 0;
  if (p->nexc > 0 && p->exc_n != c->N)
    return fail(PDX_ERR_DIMENSION, "exceptions and cluster list particle counts differ");
  hipStream_t st = as_stream(stream);
  TileArgs a{};
  a.eps = p->eps; a.sig2 = p->sig2; a.rc2 = p->rc2; a.eshift = p->eshift;
  a.rfac2 = p->rfac2; a.saux = c->saux.p; a.styp = c->styp.p; a.tab = p->tab.p; a.ntypes = p->ntypes;
  if (typed_model(p->spec.model) && p->typed_n != c->N)
    return fail(PDX_ERR_DIMENSION, "typed pair data and cluster list particle counts differ");
  a.cC = p->cC; a.crc2 = p->crc2; a.csx = p->csx; a.ceshift = p->ceshift;
  for (int k = 0; k < MAXC; ++k) {
    a.pe[k] = p->pe[k];
    a.pf[k] = p->pf[k];
  }
  a.packed = c->packed.p;
  a.ctl = c->ctl.p;
  a.geom = c->geom.p;
  a.ne = c->ne.p;
  a.oent = c->oent.p;
  a.went = c->went.p;
  a.wjb = c->wjb.p;
  a.nw = c->nw.p;
  a.keep = c->keep.p;
  a.jbuf = c->jbuf.p;
  a.islot = c->islot.p;
  a.ecap = c->ecap;
  a.nslot = c->nslot;
  a.waves = c->waves;
  a.splits = c->splits;
  const double rp = c->rmax + c->pb;
  a.rp2 = (float)(rp * rp);
  a.rmax2 = (float)(c->rlist * c->rlist * 1.0001);
  const int np = std::max(p->npe, p->npf);
  const int grid = c->NSmax * c->splits, threads = WAVE * c->waves;
  switch (p->spec.model) {
    case PDX_PAIR_MIXED_LJ: launch_tiles_model<PDX_PAIR_MIXED_LJ>(coul, energy, np, st, grid, threads, a, c); break;
    case PDX_PAIR_MIXED_WCA: launch_tiles_model<PDX_PAIR_MIXED_WCA>(coul, energy, np, st, grid, threads, a, c); break;
    case PDX_PAIR_MATRIX_LJ: launch_tiles_model<PDX_PAIR_MATRIX_LJ>(coul, energy, np, st, grid, threads, a, c); break;
    case PDX_PAIR_MATRIX_WCA:
      launch_tiles_model<PDX_PAIR_MATRIX_WCA>(coul, energy, np, st, grid, threads, a, c);
      break;
    case PDX_PAIR_LJ: launch_tiles_model<PDX_PAIR_LJ>(coul, energy, np, st, grid, threads, a, c); break;
    case PDX_PAIR_WCA: launch_tiles_model<PDX_PAIR_WCA>(coul, energy, np, st, grid, threads, a, c); break;
    case PDX_PAIR_SOFT: launch_tiles_model<PDX_PAIR_SOFT>(coul, energy, np, st, grid, threads, a, c); break;
    default: launch_tiles_model<PDX_PAIR_NONE>(coul, energy, np, st, grid, threads, a, c); break;
  }
  ScatterArgs s{};
  s.ctl = c->ctl.p;
  s.perm = c->perm.p;
  s.slot_of = c->slot_of.p;
  s.joff = c->joff.p;
  s.wjcnt = c->wjcnt.p;
  s.wjlist = c->wjlist.p;
  s.jbuf = c->jbuf.p;
  s.islot = c->islot.p;
  s.geom = c->geom.p;
  s.packed = c->packed.p;
  if (coul && p->nexc > 0) {
    s.exo = p->exc_off.p;
    s.exj = p->exc_j.p;
    s.exs = p->exc_s.p;
  }
  s.fx = fx; s.fy = fy; s.fz = fz; s.Epot = Epot; s.V = V;
  s.partial = c->partial.p;
  s.N = c->N;
  s.nslot = c->nslot;
  s.add = (flags & PDX_ADD) ? 1 : 0;
  s.totals = totals != nullptr;
  s.splits = c->splits;
  s.cC = p->cC; s.crc2 = p->crc2; s.ceshift = p->ceshift; s.halfself = p->halfself;
  if (coul) {
    if (energy) hipLaunchKernelGGL((k_cl_scatter<true, true>), dim3(c->sblocks), dim3(TPB), 0, st, s);
    else hipLaunchKernelGGL((k_cl_scatter<true, false>), dim3(c->sblocks), dim3(TPB), 0, st, s);
  } else {
    if (energy) hipLaunchKernelGGL((k_cl_scatter<false, true>), dim3(c->sblocks), dim3(TPB), 0, st, s);
    else hipLaunchKernelGGL((k_cl_scatter<false, false>), dim3(c->sblocks), dim3(TPB), 0, st, s);
  }
  PDX_LAUNCH_CHECK();
  if (energy && totals) {
    hipLaunchKernelGGL(k_sum_partials<7>, dim3(1), dim3(SUM_THREADS), 0, st, (const double *)c->partial.p, c->sblocks, totals,
                       s.add, (const int64_t *)nullptr);
    PDX_LAUNCH_CHECK();
  }
  return PDX_OK;
}

extern "C" int pdx_cluster_request_build(pdx_cluster_t c, pdx_stream_t stream) {
  if (!c) return fail(PDX_ERR_ARGUMENT, "null cluster list");
  hipLaunchKernelGGL(k_cl_set, dim3(1), dim3(1), 0, as_stream(stream), c->ctl.p, (int)W_REQUEST, 1);
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

extern "C" int pdx_cluster_status(pdx_cluster_t c, int64_t out[8], int32_t check, pdx_stream_t stream) {
  if (!c || !out) return fail(PDX_ERR_ARGUMENT, "null argument");
  PDX_HIP(hipStreamSynchronize(as_stream(stream)));
  int32_t h[W_LEN];
  PDX_HIP(hipMemcpy(h, c->ctl.p, sizeof(h), hipMemcpyDeviceToHost));
  std::vector<int32_t> ne(c->NSmax), nw((size_t)MAXCHUNK * c->NSmax);
  PDX_HIP(hipMemcpy(ne.data(), c->ne.p, sizeof(int32_t) * ne.size(), hipMemcpyDeviceToHost));
  PDX_HIP(hipMemcpy(nw.data(), c->nw.p, sizeof(int32_t) * nw.size(), hipMemcpyDeviceToHost));
  int64_t outer = 0, work = 0;
  const int NS = std::min(std::max(h[W_NS], 0), c->NSmax);
  for (int S = 0; S < NS; ++S) {
    outer += ne[S];
    const int nchunk = c->waves * c->splits;
    for (int k = 0; k < nchunk; ++k) work += nw[(size_t)S * nchunk + k];
  }
  out[0] = h[W_NS]; out[1] = outer; out[2] = work; out[3] = h[W_NBUILD];
  out[4] = h[W_NPRUNE]; out[5] = h[W_ERR]; out[6] = h[W_REQ]; out[7] = c->ecap;
  if (check) {
    if (h[W_ERR] == 1)
      return fail(PDX_ERR_CAPACITY, "cluster list: a super-cluster needs " + std::to_string(h[W_REQ]) +
                                        " outer entries, capacity " + std::to_string(c->ecap));
    if (h[W_ERR] == 2) return fail(PDX_ERR_COORDINATE, "cluster list: non-finite coordinates");
  }
  return PDX_OK;
}

extern "C" int pdx_cluster_resize(pdx_cluster_t c, int32_t entry_capacity) {
  if (!c) return fail(PDX_ERR_ARGUMENT, "null cluster list");
  if (entry_capacity < 8) return fail(PDX_ERR_ARGUMENT, "cluster list: entry capacity must be >= 8");
  PDX_HIP(hipDeviceSynchronize());
  int rc = alloc_entries(c, std::min(entry_capacity, CL * c->NSmax));
  if (rc) return rc;
  int32_t h[W_LEN];
  PDX_HIP(hipMemcpy(h, c->ctl.p, sizeof(h), hipMemcpyDeviceToHost));
  if (h[W_ERR] == 1) h[W_ERR] = 0;
  h[W_REQ] = 0;
  h[W_REQUEST] = 1;
  PDX_HIP(hipMemcpy(c->ctl.p, h, sizeof(h), hipMemcpyHostToDevice));
  PDX_HIP(hipMemset(c->ne.p, 0, sizeof(int32_t) * c->NSmax));
  PDX_HIP(hipMemset(c->nw.p, 0, sizeof(int32_t) * MAXCHUNK * c->NSmax));
  return PDX_OK;
}
