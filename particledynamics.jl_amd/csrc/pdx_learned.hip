// Learned force correction on top of coarse-mesh PME: per-particle features from the coarse
// charge grid, and an 80-64-64-3 MLP (SiLU) run on BF16 MFMA tiles with FP32 accumulation,
// for inference (the correction is added to the forces) and for training (forward +
// backward in one kernel, per-wave weight-gradient partials reduced in a fixed order, Adam).
//
// MFMA tile: v_mfma_f32_16x16x16_bf16, one wave. Operand layouts (lane l, element e):
//   A[i][k]: i = l % 16, k = 4 (l / 16) + e;  B[k][j]: j = l % 16, k = 4 (l / 16) + e;
//   D[i][j]: j = l % 16, i = 4 (l / 16) + e.
// Activations move between layers through a wave-private LDS tile (16 x 64 BF16).
#include "pdx_internal.hpp"

#include <cmath>
#include <random>

using namespace pdx;

namespace {

constexpr int TPB = 256;
constexpr int NIN = 80, NH = 64, NOUT = 3, NOP = 16;  // NOP: output padded to one MFMA tile
constexpr int NKT = NIN / 16;                         // MFMA K steps of layer 1
static_assert(NIN % 16 == 0, "input width must be a multiple of the MFMA K");
// packed FP32 parameter vector: W1 (NIN x NH, row = input), b1, W2 (NH x NH), b2, W3 (NH x NOUT), b3
constexpr int OW1 = 0, OB1 = OW1 + NIN * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH * NOUT, NPARAM = OB3 + NOUT;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f2bf(float f) {  // round to nearest even (v_cvt_pk_bf16_f32)
  return __builtin_bit_cast(short, (__bf16)f);
}
__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((unsigned)(unsigned short)b) << 16); }

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// SiLU with the hardware reciprocal (1 ulp): the activations are rounded to BF16 next
__device__ __forceinline__ float silu(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

struct MLPDev {
  const short *w1, *w2, *w3;  // BF16: W1 [NIN][NH], W2 [NH][NH], W3 [NH][NOP] (zero-padded)
  const float *b1, *b2, *b3;  // FP32 biases (b3 padded to NOP)
  const float *mu, *isd;      // feature standardisation
  float out_scale;
};

// Weights into LDS (BF16) for the block.
struct __align__(16) WLds {
  short w1[NIN * NH], w2[NH * NH], w3[NH * NOP];
  float b1[NH], b2[NH], b3[NOP], mu[NIN], isd[NIN];
};

// 16-byte copies: the three weight blocks are multiples of 8 BF16 values and 16-byte aligned.
__device__ __forceinline__ void copy16(short *dst, const short *src, int count) {
  const uint4 *s = reinterpret_cast<const uint4 *>(src);
  uint4 *d = reinterpret_cast<uint4 *>(dst);
  for (int k = threadIdx.x; k < count / 8; k += blockDim.x) d[k] = s[k];
}

__device__ void load_weights(WLds &L, const MLPDev &m) {
  copy16(L.w1, m.w1, NIN * NH);
  copy16(L.w2, m.w2, NH * NH);
  copy16(L.w3, m.w3, NH * NOP);
  for (int k = threadIdx.x; k < NH; k += blockDim.x) { L.b1[k] = m.b1[k]; L.b2[k] = m.b2[k]; }
  for (int k = threadIdx.x; k < NOP; k += blockDim.x) L.b3[k] = m.b3[k];
  for (int k = threadIdx.x; k < NIN; k += blockDim.x) { L.mu[k] = m.mu[k]; L.isd[k] = m.isd[k]; }
}

// Inference weights in LDS in MFMA B-fragment order: fragment f of a [K][N] matrix is its 16 x 16
// block (k0, j0) = (16 (f / (N/16)), 16 (f % (N/16))), and lane l keeps that block's
// W[k0 + 4 (l/16) + e][j0 + l % 16], e = 0..3, as one 8-byte vector: one LDS read per MFMA operand.
struct __align__(16) WFrag {
  bf16x4 w1[(NIN / 16) * (NH / 16) * WAVE], w2[(NH / 16) * (NH / 16) * WAVE], w3[(NH / 16) * (NOP / 16) * WAVE];
  float b1[NH], b2[NH], b3[NOP], mu[NIN], isd[NIN];
};

__device__ void load_weights(WFrag &L, const MLPDev &m) {
  auto put = [&](bf16x4 *dst, const short *src, int K, int N) {
    const int nf = (K / 16) * (N / 16);
    for (int q = threadIdx.x; q < nf * WAVE; q += blockDim.x) {
      const int f = q / WAVE, l = q % WAVE;
      const int k0 = 16 * (f / (N / 16)) + 4 * (l / 16), j = 16 * (f % (N / 16)) + l % 16;
      bf16x4 v;
#pragma unroll
      for (int e = 0; e < 4; ++e) v[e] = src[(k0 + e) * N + j];
      dst[q] = v;
    }
  };
  put(L.w1, m.w1, NIN, NH);
  put(L.w2, m.w2, NH, NH);
  put(L.w3, m.w3, NH, NOP);
  for (int k = threadIdx.x; k < NH; k += blockDim.x) { L.b1[k] = m.b1[k]; L.b2[k] = m.b2[k]; }
  for (int k = threadIdx.x; k < NOP; k += blockDim.x) L.b3[k] = m.b3[k];
  for (int k = threadIdx.x; k < NIN; k += blockDim.x) { L.mu[k] = m.mu[k]; L.isd[k] = m.isd[k]; }
}

// Standardised input tile (16 particles x NIN) into the wave's LDS scratch as BF16 [i][k] (ld NIN).
// Each lane loads its float4 pieces of the tile first (all loads in flight at once), then
// standardises and converts them.
constexpr int NV4 = NIN / 4;              // float4 per feature row
constexpr int XPER = 16 * NV4 / WAVE;     // float4 per lane per tile
static_assert((16 * NV4) % WAVE == 0, "tile loads must divide evenly over the wave");
template <class W>
__device__ __forceinline__ void stage_inputs(short *xs, const float *__restrict__ feats, const int32_t *rows,
                                             int base, int n, const W &L, int lane) {
  float4 v[XPER];
#pragma unroll
  for (int u = 0; u < XPER; ++u) {
    const int e = lane + u * WAVE, i = e / NV4, k4 = e % NV4;
    const int p = base + i;
    v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (p < n) {
      const int r = rows ? rows[p] : p;
      v[u] = *reinterpret_cast<const float4 *>(feats + (size_t)r * NIN + 4 * k4);
    }
  }
#pragma unroll
  for (int u = 0; u < XPER; ++u) {
    const int e = lane + u * WAVE, i = e / NV4, k = 4 * (e % NV4);
    const bool in = base + i < n;
    short *d = xs + i * NIN + k;
    d[0] = f2bf(in ? (v[u].x - L.mu[k + 0]) * L.isd[k + 0] : 0.f);
    d[1] = f2bf(in ? (v[u].y - L.mu[k + 1]) * L.isd[k + 1] : 0.f);
    d[2] = f2bf(in ? (v[u].z - L.mu[k + 2]) * L.isd[k + 2] : 0.f);
    d[3] = f2bf(in ? (v[u].w - L.mu[k + 3]) * L.isd[k + 3] : 0.f);
  }
}

// The same standardised tile formed straight from the charge grid of the mesh's last spread, with
// the feature definition and arithmetic of the mesh feature pass (pdx_mesh_features); all loads
// are issued before the first conversion waits on them.
struct MeshFeat {
  const unsigned long long *grid;
  const float *x, *y, *z, *q;
  int K1, K2, K3;
  float Lx, Ly, Lz;
  float inv_scale;
};
// Lane l of a wave staging a 16-particle tile takes particle l / 4 and the grid x offset l % 4
// (features k = l % 4 + 4 s): four adjacent lanes read four adjacent nodes of one x row.
__device__ __forceinline__ int feat_particle(int lane) { return lane >> 2; }
__device__ __forceinline__ int feat_xoff(int lane) { return lane & 3; }
// One particle's position and charge as loaded (lane l holds particle base + l / 4).
struct PosQ {
  float x, y, z, q;
};
__device__ __forceinline__ PosQ load_posq(const MeshFeat &g, int base, int n, int lane) {
  const int p = base + feat_particle(lane);
  PosQ r{0.f, 0.f, 0.f, 0.f};
  if (p < n) r = PosQ{g.x[p], g.y[p], g.z[p], g.q[p]};
  return r;
}
// Issue the lane's 16 grid loads (feature k = s0 + 4 s: x offset s0, y offset s & 3, z offset s >> 2).
__device__ __forceinline__ void load_grid(unsigned long long (&gv)[16], const MeshFeat &g, const PosQ &pq, bool in,
                                          int s0) {
  const float f1 = floorf(grid_coordinate(pq.x, g.Lx, g.K1)), f2 = floorf(grid_coordinate(pq.y, g.Ly, g.K2)),
              f3 = floorf(grid_coordinate(pq.z, g.Lz, g.K3));
  const int n1 = wrap_node((int)f1 - 3 + s0, g.K1);
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int n2 = wrap_node((int)f2 - 3 + (s & 3), g.K2), n3 = wrap_node((int)f3 - 3 + (s >> 2), g.K3);
    gv[s] = in ? g.grid[(size_t)n3 * g.K1 * g.K2 + (size_t)n2 * g.K1 + n1] : 0ull;
  }
}
// The standardised BF16 features of the lane's particle into the tile, from its loads.
template <class W>
__device__ __forceinline__ void stage_features(short *xs, const MeshFeat &g, const PosQ &pq,
                                               const unsigned long long (&gv)[16], bool in, const W &L,
                                               int lane) {
  const int i = feat_particle(lane), s0 = feat_xoff(lane);
  const float qi = pq.q;
  const float u1 = grid_coordinate(pq.x, g.Lx, g.K1), u2 = grid_coordinate(pq.y, g.Ly, g.K2),
              u3 = grid_coordinate(pq.z, g.Lz, g.K3);
  const float f1 = floorf(u1), f2 = floorf(u2), f3 = floorf(u3);
  short *d = xs + i * NIN;
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int k = s0 + 4 * s;
    const float f = qi * fixed_to_float(gv[s], g.inv_scale);
    d[k] = f2bf(in ? (f - L.mu[k]) * L.isd[k] : 0.f);
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) {  // k = 64 + s0 + 4 s: q (u - floor u) per dimension, q, q^2, zeros
    const int e = s0 + 4 * s, k = 64 + e;
    float f = 0.f;
    f = e == 0 ? qi * (u1 - f1) : f;
    f = e == 1 ? qi * (u2 - f2) : f;
    f = e == 2 ? qi * (u3 - f3) : f;
    f = e == 3 ? qi : f;
    f = e == 4 ? qi * qi : f;
    d[k] = f2bf(in ? (f - L.mu[k]) * L.isd[k] : 0.f);
  }
}
static_assert(NIN == 80, "the mesh staging forms 64 grid features and 16 others");

// A operand (rows i, columns k0..k0+15) of a [16][ld] BF16 LDS tile.
__device__ __forceinline__ bf16x4 a_from(const short *t, int k0, int lane, int ld = NH) {
  const int i = lane % 16, k = k0 + 4 * (lane / 16);
  bf16x4 a;
#pragma unroll
  for (int e = 0; e < 4; ++e) a[e] = t[i * ld + k + e];
  return a;
}
// B operand (rows k0..k0+15, columns j0..j0+15) of a row-major [K][ld] BF16 matrix.
__device__ __forceinline__ bf16x4 b_from(const short *w, int ld, int k0, int j0, int lane) {
  const int j = j0 + lane % 16, k = k0 + 4 * (lane / 16);
  bf16x4 b;
#pragma unroll
  for (int e = 0; e < 4; ++e) b[e] = w[(k + e) * ld + j];
  return b;
}
// B operand of the transpose: B[k][j] = w[j][k] (w row-major [J][ld]).
__device__ __forceinline__ bf16x4 bt_from(const short *w, int ld, int k0, int j0, int lane) {
  const int j = j0 + lane % 16, k = k0 + 4 * (lane / 16);
  bf16x4 b;
#pragma unroll
  for (int e = 0; e < 4; ++e) b[e] = w[j * ld + k + e];
  return b;
}

// B operands of the three layers (rows k0..k0+15, columns 16 nt..): from the row-major weights of
// the training kernels, or from the fragment-ordered inference weights.
__device__ __forceinline__ bf16x4 w1_op(const WLds &L, int k0, int nt, int lane) { return b_from(L.w1, NH, k0, nt * 16, lane); }
__device__ __forceinline__ bf16x4 w2_op(const WLds &L, int k0, int nt, int lane) { return b_from(L.w2, NH, k0, nt * 16, lane); }
__device__ __forceinline__ bf16x4 w3_op(const WLds &L, int k0, int lane) { return b_from(L.w3, NOP, k0, 0, lane); }
__device__ __forceinline__ bf16x4 w1_op(const WFrag &L, int k0, int nt, int lane) {
  return L.w1[((k0 / 16) * (NH / 16) + nt) * WAVE + lane];
}
__device__ __forceinline__ bf16x4 w2_op(const WFrag &L, int k0, int nt, int lane) {
  return L.w2[((k0 / 16) * (NH / 16) + nt) * WAVE + lane];
}
__device__ __forceinline__ bf16x4 w3_op(const WFrag &L, int k0, int lane) { return L.w3[(k0 / 16) * WAVE + lane]; }

// Forward of one 16-particle tile: returns Z1, Z2 (pre-activations, C layout, 4 N-tiles) and
// the output (one tile, columns 0..2 valid); H1/H2 are left in the wave's LDS tiles.
template <class W>
__device__ __forceinline__ void forward_tile(const W &L, short *xs, short *h1s, short *h2s, int lane,
                                             f32x4 (&z1)[4], f32x4 (&z2)[4], f32x4 &y) {
  // layer 1: [16 x NIN] x [NIN x NH]; the four column tiles' accumulation chains advance together
  // (independent MFMAs back to back), each summed over k in ascending order
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) z1[nt] = f32x4{0, 0, 0, 0};
#pragma unroll
  for (int k0 = 0; k0 < NIN; k0 += 16) {
    const bf16x4 a = a_from(xs, k0, lane, NIN);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) z1[nt] = mfma(a, w1_op(L, k0, nt, lane), z1[nt]);
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const int j = nt * 16 + lane % 16;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      z1[nt][e] += L.b1[j];
      h1s[(4 * (lane / 16) + e) * NH + j] = f2bf(silu(z1[nt][e]));
    }
  }
  wave_sync();
  // layer 2: [16 x NH] x [NH x NH]
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) z2[nt] = f32x4{0, 0, 0, 0};
#pragma unroll
  for (int k0 = 0; k0 < NH; k0 += 16) {
    const bf16x4 a = a_from(h1s, k0, lane);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) z2[nt] = mfma(a, w2_op(L, k0, nt, lane), z2[nt]);
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const int j = nt * 16 + lane % 16;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      z2[nt][e] += L.b2[j];
      h2s[(4 * (lane / 16) + e) * NH + j] = f2bf(silu(z2[nt][e]));
    }
  }
  wave_sync();
  // layer 3: [16 x NH] x [NH x NOP]
  f32x4 c = {0, 0, 0, 0};
#pragma unroll
  for (int k0 = 0; k0 < NH; k0 += 16) c = mfma(a_from(h2s, k0, lane), w3_op(L, k0, lane), c);
#pragma unroll
  for (int e = 0; e < 4; ++e) c[e] += L.b3[lane % 16];
  y = c;
}

// Inference: out = out_scale * MLP(standardised features); added to fx, fy, fz when given.
__global__ __launch_bounds__(TPB) void k_mlp_forward(MLPDev m, const float *__restrict__ feats, int n,
                                                      float *out, float *fx, float *fy, float *fz) {
  __shared__ WFrag L;
  __shared__ short xtiles[TPB / WAVE][16 * NIN];
  __shared__ short tiles[TPB / WAVE][2][16 * NH];
  load_weights(L, m);
  __syncthreads();
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  short *xs = xtiles[w], *h1s = tiles[w][0], *h2s = tiles[w][1];
  const int ntile = (n + 15) / 16;
  for (int t = blockIdx.x * (TPB / WAVE) + w; t < ntile; t += gridDim.x * (TPB / WAVE)) {
    const int base = t * 16;
    stage_inputs(xs, feats, nullptr, base, n, L, lane);
    wave_sync();
    f32x4 z1[4], z2[4], y;
    forward_tile(L, xs, h1s, h2s, lane, z1, z2, y);
    const int j = lane % 16;
    if (j < NOUT) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int p = base + 4 * (lane / 16) + e;
        if (p >= n) continue;
        const float v = m.out_scale * y[e];
        if (out) out[(size_t)p * NOUT + j] = v;
        float *f = j == 0 ? fx : (j == 1 ? fy : fz);
        if (f) f[p] += v;
      }
    }
    wave_sync();
  }
}

// Inference from the mesh: the feature tile of each 16-particle tile is formed from the charge
// grid in the wave's LDS (no feature rows in memory); out (n x 3) and/or fx, fy, fz as in
// k_mlp_forward.
__global__ __launch_bounds__(TPB) void k_mlp_forward_mesh(MLPDev m, const MeshFeat g, int n, float *out, float *fx,
                                                           float *fy, float *fz) {
  __shared__ WFrag L;
  __shared__ short xtiles[TPB / WAVE][16 * NIN];
  __shared__ short tiles[TPB / WAVE][2][16 * NH];
  load_weights(L, m);
  __syncthreads();
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  short *xs = xtiles[w], *h1s = tiles[w][0], *h2s = tiles[w][1];
  const int ntile = (n + 15) / 16;
  const int s0 = feat_xoff(lane);
  for (int t = blockIdx.x * (TPB / WAVE) + w; t < ntile; t += gridDim.x * (TPB / WAVE)) {
    const int base = t * 16;
    const bool in = base + feat_particle(lane) < n;
    const PosQ pc = load_posq(g, base, n, lane);
    unsigned long long gc[16];
    load_grid(gc, g, pc, in, s0);
    stage_features(xs, g, pc, gc, in, L, lane);
    wave_sync();
    f32x4 z1[4], z2[4], y;
    forward_tile(L, xs, h1s, h2s, lane, z1, z2, y);
    const int j = lane % 16;
    if (j < NOUT) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int p = base + 4 * (lane / 16) + e;
        if (p >= n) continue;
        const float v = m.out_scale * y[e];
        if (out) out[(size_t)p * NOUT + j] = v;
        float *f = j == 0 ? fx : (j == 1 ? fy : fz);
        if (f) f[p] += v;
      }
    }
    wave_sync();
  }
}

__global__ __launch_bounds__(TPB) void k_add_correction(int n, const float *__restrict__ out, float *fx, float *fy,
                                                        float *fz) {
  const int p = blockIdx.x * TPB + threadIdx.x;
  if (p >= n) return;
  fx[p] += out[(size_t)p * NOUT + 0];
  fy[p] += out[(size_t)p * NOUT + 1];
  fz[p] += out[(size_t)p * NOUT + 2];
}

// Training gradients: per wave, forward + backward over its tiles of the batch, weight and
// bias gradients accumulated in MFMA registers; per-wave partials (NPARAM + 1 for the loss)
// written to `partial`. Loss = mean over batch and components of (y - t/out_scale)^2.
__global__ __launch_bounds__(TPB) void k_mlp_grad(MLPDev m, const float *__restrict__ feats,
                                                   const float *__restrict__ targets,
                                                   const int32_t *__restrict__ rows, int n, float *partial) {
  __shared__ WLds L;
  __shared__ short xtiles[TPB / WAVE][16 * NIN];
  __shared__ short tiles[TPB / WAVE][4][16 * NH];  // h1, h2, dz2 (as A), dz1 / dY staging
  load_weights(L, m);
  __syncthreads();
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  short *xs = xtiles[w], *h1s = tiles[w][0], *h2s = tiles[w][1], *gs = tiles[w][2], *ds = tiles[w][3];
  const float inv_scale = 1.0f / m.out_scale;
  const float gnorm = 2.0f / (float)(n * NOUT);
  f32x4 gw1[NKT][4], gw2[4][4], gw3[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    gw3[a] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int b = 0; b < 4; ++b) gw2[a][b] = f32x4{0, 0, 0, 0};
  }
#pragma unroll
  for (int a = 0; a < NKT; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) gw1[a][b] = f32x4{0, 0, 0, 0};
  float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gb3 = 0.f;
  double loss = 0.0;
  const int ntile = (n + 15) / 16;
  const int gw = blockIdx.x * (TPB / WAVE) + w;
  for (int t = gw; t < ntile; t += gridDim.x * (TPB / WAVE)) {
    const int base = t * 16;
    stage_inputs(xs, feats, rows, base, n, L, lane);
    wave_sync();
    f32x4 z1[4], z2[4], y;
    forward_tile(L, xs, h1s, h2s, lane, z1, z2, y);
    // dY (16 x NOP, columns >= NOUT zero), stored as BF16 [i][j] in ds (ld NH)
    const int j = lane % 16;
    f32x4 dy;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = 4 * (lane / 16) + e;
      const int p = base + i;
      float d = 0.f;
      if (j < NOUT && p < n) {
        const int r = rows ? rows[p] : p;
        const float diff = y[e] - targets[(size_t)r * NOUT + j] * inv_scale;
        loss += (double)diff * diff;
        d = gnorm * diff;
      }
      dy[e] = d;
      ds[i * NH + j] = f2bf(d);
      gb3 += (j < NOUT) ? d : 0.f;
    }
    wave_sync();
    // dW3 [NH x NOP] += H2^T (NH x 16) dY (16 x NOP): A[m][k] = H2[k][m], B[k][j] = dY[k][j]
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      bf16x4 a, b;
      const int mrow = mt * 16 + lane % 16, k = 4 * (lane / 16);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        a[e] = h2s[(k + e) * NH + mrow];
        b[e] = ds[(k + e) * NH + j];
      }
      gw3[mt] = mfma(a, b, gw3[mt]);
    }
    // dH2 (16 x NH) = dY (16 x NOP) W3^T: B[k][jj] = W3[jj][k]; then dZ2 = dH2 * silu'(Z2)
    f32x4 dz2[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      f32x4 c = {0, 0, 0, 0};
      c = mfma(a_from(ds, 0, lane), bt_from(L.w3, NOP, 0, nt * 16, lane), c);
#pragma unroll
      for (int e = 0; e < 4; ++e) c[e] *= dsilu(z2[nt][e]);
      dz2[nt] = c;
      const int jj = nt * 16 + j;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        gs[(4 * (lane / 16) + e) * NH + jj] = f2bf(c[e]);
        gb2[nt] += c[e];
      }
    }
    wave_sync();
    // dW2 [NH x NH] += H1^T dZ2
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      bf16x4 a;
      const int mrow = mt * 16 + lane % 16, k = 4 * (lane / 16);
#pragma unroll
      for (int e = 0; e < 4; ++e) a[e] = h1s[(k + e) * NH + mrow];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        bf16x4 b;
#pragma unroll
        for (int e = 0; e < 4; ++e) b[e] = gs[(k + e) * NH + nt * 16 + j];
        gw2[mt][nt] = mfma(a, b, gw2[mt][nt]);
      }
    }
    // dH1 (16 x NH) = dZ2 W2^T; dZ1 = dH1 * silu'(Z1) -> ds (BF16 [i][j])
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      f32x4 c = {0, 0, 0, 0};
#pragma unroll
      for (int k0 = 0; k0 < NH; k0 += 16) c = mfma(a_from(gs, k0, lane), bt_from(L.w2, NH, k0, nt * 16, lane), c);
#pragma unroll
      for (int e = 0; e < 4; ++e) c[e] *= dsilu(z1[nt][e]);
      const int jj = nt * 16 + j;
      wave_sync();
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        gb1[nt] += c[e];
        ds[(4 * (lane / 16) + e) * NH + jj] = f2bf(c[e]);
      }
    }
    wave_sync();
    // dW1 [NIN x NH] += X^T dZ1
#pragma unroll
    for (int mt = 0; mt < NKT; ++mt) {
      bf16x4 a;
      const int mrow = mt * 16 + lane % 16, k = 4 * (lane / 16);
#pragma unroll
      for (int e = 0; e < 4; ++e) a[e] = xs[(k + e) * NIN + mrow];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        bf16x4 b;
#pragma unroll
        for (int e = 0; e < 4; ++e) b[e] = ds[(k + e) * NH + nt * 16 + j];
        gw1[mt][nt] = mfma(a, b, gw1[mt][nt]);
      }
    }
    wave_sync();
  }
  // per-wave partials: weight gradients straight from the accumulator layouts; bias gradients
  // reduced over the four lane groups that share a column
  float *P = partial + (size_t)gw * (NPARAM + 1);
#pragma unroll
  for (int mt = 0; mt < NKT; ++mt)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = mt * 16 + 4 * (lane / 16) + e, c = nt * 16 + lane % 16;
        P[OW1 + r * NH + c] = gw1[mt][nt][e];
      }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = mt * 16 + 4 * (lane / 16) + e, c = nt * 16 + lane % 16;
        P[OW2 + r * NH + c] = gw2[mt][nt][e];
      }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int r = mt * 16 + 4 * (lane / 16) + e, c = lane % 16;
      if (c < NOUT) P[OW3 + r * NOUT + c] = gw3[mt][e];
    }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    float b1 = gb1[nt], b2 = gb2[nt];
    b1 += __shfl_xor(b1, 16, WAVE); b1 += __shfl_xor(b1, 32, WAVE);
    b2 += __shfl_xor(b2, 16, WAVE); b2 += __shfl_xor(b2, 32, WAVE);
    if (lane < 16) { P[OB1 + nt * 16 + lane] = b1; P[OB2 + nt * 16 + lane] = b2; }
  }
  float b3 = gb3;
  b3 += __shfl_xor(b3, 16, WAVE); b3 += __shfl_xor(b3, 32, WAVE);
  if (lane < NOUT) P[OB3 + lane] = b3;
  loss = wave_sum(loss);
  if (lane == 0) P[NPARAM] = (float)loss;
}

// Fixed-order reduction of the per-wave partials -> grad, loss. A block takes RED_P consecutive
// parameters (coalesced across lanes) and RED_S slices of the waves: slice s sums waves s, s +
// RED_S, ... in FP64, and the slices are added in slice order, so the result depends only on nwave.
constexpr int RED_P = 64, RED_S = 16;
__global__ __launch_bounds__(RED_P *RED_S) void k_mlp_reduce(const float *partial, int nwave, float *grad,
                                                               double *loss, int n) {
  __shared__ double acc[RED_S][RED_P];
  const int j = threadIdx.x % RED_P, sl = threadIdx.x / RED_P;
  const int k = blockIdx.x * RED_P + j;
  double s = 0.0;
  if (k <= NPARAM) {
    int w = sl;
    for (; w + 3 * RED_S < nwave; w += 4 * RED_S) {
      const float a = partial[(size_t)w * (NPARAM + 1) + k], b = partial[(size_t)(w + RED_S) * (NPARAM + 1) + k],
                  c = partial[(size_t)(w + 2 * RED_S) * (NPARAM + 1) + k],
                  d = partial[(size_t)(w + 3 * RED_S) * (NPARAM + 1) + k];
      s += a; s += b; s += c; s += d;
    }
    for (; w < nwave; w += RED_S) s += partial[(size_t)w * (NPARAM + 1) + k];
  }
  acc[sl][j] = s;
  __syncthreads();
  if (sl == 0 && k <= NPARAM) {
    double t = acc[0][j];
    for (int q = 1; q < RED_S; ++q) t += acc[q][j];
    if (k < NPARAM) grad[k] = (float)t;
    else if (loss) *loss = t / (double)(n * NOUT);
  }
}

// Adam on the FP32 master parameters, then the BF16 / padded device copies.
__global__ void k_mlp_adam(float *param, float *mom, float *vel, const float *grad, float lr, float b1, float b2,
                           float eps, float c1, float c2, short *w1, short *w2, short *w3, float *bb1, float *bb2,
                           float *bb3) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= NPARAM) return;
  const float g = grad[k];
  const float mm = b1 * mom[k] + (1.f - b1) * g;
  const float vv = b2 * vel[k] + (1.f - b2) * g * g;
  mom[k] = mm;
  vel[k] = vv;
  const float p = param[k] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  param[k] = p;
  if (k < OB1) w1[k - OW1] = f2bf(p);
  else if (k < OW2) bb1[k - OB1] = p;
  else if (k < OB2) w2[k - OW2] = f2bf(p);
  else if (k < OW3) bb2[k - OB2] = p;
  else if (k < OB3) { const int r = (k - OW3) / NOUT, c = (k - OW3) % NOUT; w3[r * NOP + c] = f2bf(p); }
  else bb3[k - OB3] = p;
}

__global__ void k_mlp_upload(const float *param, short *w1, short *w2, short *w3, float *bb1, float *bb2,
                             float *bb3) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= NPARAM) return;
  const float p = param[k];
  if (k < OB1) w1[k - OW1] = f2bf(p);
  else if (k < OW2) bb1[k - OB1] = p;
  else if (k < OB2) w2[k - OW2] = f2bf(p);
  else if (k < OW3) bb2[k - OB2] = p;
  else if (k < OB3) { const int r = (k - OW3) / NOUT, c = (k - OW3) % NOUT; w3[r * NOP + c] = f2bf(p); }
  else bb3[k - OB3] = p;
}

}  // namespace

struct pdx_mlp_s {
  DevBuf<float> param, mom, vel, grad, partial;
  DevBuf<short> w1, w2, w3;
  DevBuf<float> b1, b2, b3, mu, isd;
  DevBuf<double> loss;
  float out_scale = 1.0f;
  int64_t step = 0;
  int nwave_max = 0;
};

namespace {
MLPDev dev_of(pdx_mlp_t m) {
  MLPDev d{};
  d.w1 = m->w1.p; d.w2 = m->w2.p; d.w3 = m->w3.p;
  d.b1 = m->b1.p; d.b2 = m->b2.p; d.b3 = m->b3.p;
  d.mu = m->mu.p; d.isd = m->isd.p;
  d.out_scale = m->out_scale;
  return d;
}
}  // namespace

extern "C" int pdx_mlp_nparams(void) { return NPARAM; }

extern "C" int pdx_mlp_create(int32_t seed, pdx_mlp_t *out) {
  if (!out) return fail(PDX_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  auto *m = new pdx_mlp_s();
  int rc;
  if ((rc = m->param.alloc(NPARAM)) || (rc = m->mom.alloc(NPARAM)) || (rc = m->vel.alloc(NPARAM)) ||
      (rc = m->grad.alloc(NPARAM)) || (rc = m->w1.alloc(NIN * NH)) || (rc = m->w2.alloc(NH * NH)) ||
      (rc = m->w3.alloc(NH * NOP)) || (rc = m->b1.alloc(NH)) || (rc = m->b2.alloc(NH)) || (rc = m->b3.alloc(NOP)) ||
      (rc = m->mu.alloc(NIN)) || (rc = m->isd.alloc(NIN)) || (rc = m->loss.alloc(1))) {
    delete m;
    return rc;
  }
  // He-style initialisation from a seeded generator; zero biases
  std::vector<float> p(NPARAM, 0.0f);
  std::mt19937_64 g((uint64_t)seed * 0x9E3779B97F4A7C15ull + 1);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  for (int k = 0; k < NIN * NH; ++k) p[OW1 + k] = nd(g) * std::sqrt(2.0f / NIN);
  for (int k = 0; k < NH * NH; ++k) p[OW2 + k] = nd(g) * std::sqrt(2.0f / NH);
  for (int k = 0; k < NH * NOUT; ++k) p[OW3 + k] = nd(g) * std::sqrt(1.0f / NH);
  std::vector<float> zeros(NH * NH, 0.0f), ones(NIN, 1.0f);
  hipError_t e = hipMemcpy(m->param.p, p.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(m->mom.p, 0, sizeof(float) * NPARAM);
  if (e == hipSuccess) e = hipMemset(m->vel.p, 0, sizeof(float) * NPARAM);
  if (e == hipSuccess) e = hipMemset(m->w3.p, 0, sizeof(short) * NH * NOP);
  if (e == hipSuccess) e = hipMemset(m->b3.p, 0, sizeof(float) * NOP);
  if (e == hipSuccess) e = hipMemset(m->mu.p, 0, sizeof(float) * NIN);
  if (e == hipSuccess) e = hipMemcpy(m->isd.p, ones.data(), sizeof(float) * NIN, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mlp_upload, dim3((NPARAM + 255) / 256), dim3(256), 0, 0, m->param.p, m->w1.p, m->w2.p,
                       m->w3.p, m->b1.p, m->b2.p, m->b3.p);
    e = hipDeviceSynchronize();
  }
  if (e != hipSuccess) {
    delete m;
    return fail(PDX_ERR_DEVICE, std::string("mlp initialisation failed: ") + hipGetErrorString(e));
  }
  *out = m;
  return PDX_OK;
}

extern "C" int pdx_mlp_destroy(pdx_mlp_t m) {
  delete m;
  return PDX_OK;
}

extern "C" int pdx_mlp_set_params(pdx_mlp_t m, const float *params, int32_t n) {
  if (!m || !params) return fail(PDX_ERR_ARGUMENT, "null argument");
  if (n != NPARAM) return fail(PDX_ERR_DIMENSION, "parameter count differs");
  PDX_HIP(hipMemcpy(m->param.p, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  PDX_HIP(hipMemset(m->mom.p, 0, sizeof(float) * NPARAM));
  PDX_HIP(hipMemset(m->vel.p, 0, sizeof(float) * NPARAM));
  m->step = 0;
  hipLaunchKernelGGL(k_mlp_upload, dim3((NPARAM + 255) / 256), dim3(256), 0, 0, m->param.p, m->w1.p, m->w2.p,
                     m->w3.p, m->b1.p, m->b2.p, m->b3.p);
  PDX_HIP(hipDeviceSynchronize());
  return PDX_OK;
}

extern "C" int pdx_mlp_get_params(pdx_mlp_t m, float *params, int32_t n) {
  if (!m || !params) return fail(PDX_ERR_ARGUMENT, "null argument");
  if (n != NPARAM) return fail(PDX_ERR_DIMENSION, "parameter count differs");
  PDX_HIP(hipDeviceSynchronize());
  PDX_HIP(hipMemcpy(params, m->param.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  return PDX_OK;
}

extern "C" int pdx_mlp_set_norm(pdx_mlp_t m, const float *mu, const float *isd, float out_scale) {
  if (!m || !mu || !isd) return fail(PDX_ERR_ARGUMENT, "null argument");
  if (!(out_scale > 0 && std::isfinite(out_scale))) return fail(PDX_ERR_ARGUMENT, "out_scale must be positive");
  PDX_HIP(hipMemcpy(m->mu.p, mu, sizeof(float) * NIN, hipMemcpyHostToDevice));
  PDX_HIP(hipMemcpy(m->isd.p, isd, sizeof(float) * NIN, hipMemcpyHostToDevice));
  m->out_scale = out_scale;
  return PDX_OK;
}

extern "C" int pdx_mlp_forward(pdx_mlp_t m, const float *feats, int32_t n, float *out, float *fx, float *fy,
                               float *fz, pdx_stream_t stream) {
  if (!m || !feats) return fail(PDX_ERR_ARGUMENT, "null argument");
  if (n <= 0) return PDX_OK;
  if ((fx || fy || fz) && !(fx && fy && fz)) return fail(PDX_ERR_ARGUMENT, "give all three force arrays or none");
  // one tile per wave (the weights are staged into LDS once per block of four waves)
  const int ntile = (n + 15) / 16;
  const int blocks = std::max(1, std::min((ntile + 3) / 4, 2048));
  hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(TPB), 0, as_stream(stream), dev_of(m), feats, n, out, fx, fy,
                     fz);
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

extern "C" int pdx_mlp_forward_mesh(pdx_mlp_t m, pdx_mesh_t mesh, const float *x, const float *y, const float *z,
                                    const float *q, int32_t n, float *out, float *fx, float *fy, float *fz,
                                    pdx_stream_t stream) {
  if (!m || !mesh || !x || !y || !z || !q) return fail(PDX_ERR_ARGUMENT, "null argument");
  if (n <= 0) return PDX_OK;
  if (n > mesh->N) return fail(PDX_ERR_ARGUMENT, "more particles than the mesh was made for");
  if ((fx || fy || fz) && !(fx && fy && fz)) return fail(PDX_ERR_ARGUMENT, "give all three force arrays or none");
  if (!out && !fx) return fail(PDX_ERR_ARGUMENT, "give out, or fx, fy and fz");
  if (mesh->K1 < 4 || mesh->K2 < 4 || mesh->K3 < 4)
    return fail(PDX_ERR_ARGUMENT, "features need at least 4 mesh nodes a side");
  MeshFeat g{mesh->grid64.p, x, y, z, q, mesh->K1, mesh->K2, mesh->K3, (float)mesh->box[0], (float)mesh->box[1],
             (float)mesh->box[2], (float)mesh->fixed_inv};
  // one resident round: the kernel fits three 4-wave workgroups per CU (LDS), and each wave loops
  // over its tiles, so no second round of short-lived waves forms a tail
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;  // the part's CU count when the query fails
  const int ntile = (n + 15) / 16;
  const int blocks = std::max(1, std::min((ntile + 3) / 4, 3 * cus));
  hipLaunchKernelGGL(k_mlp_forward_mesh, dim3(blocks), dim3(TPB), 0, as_stream(stream), dev_of(m), g, n, out, fx, fy,
                     fz);
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

extern "C" int pdx_mlp_add(const float *out, int32_t n, float *fx, float *fy, float *fz, pdx_stream_t stream) {
  if (!out || !fx || !fy || !fz) return fail(PDX_ERR_ARGUMENT, "null argument");
  if (n <= 0) return PDX_OK;
  hipLaunchKernelGGL(k_add_correction, dim3((n + TPB - 1) / TPB), dim3(TPB), 0, as_stream(stream), n, out, fx, fy, fz);
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

extern "C" int pdx_mlp_gradient(pdx_mlp_t m, const float *feats, const float *targets, const int32_t *rows,
                                int32_t n, float *grad, double *loss, pdx_stream_t stream) {
  if (!m || !feats || !targets) return fail(PDX_ERR_ARGUMENT, "null argument");
  if (n <= 0) return fail(PDX_ERR_ARGUMENT, "empty batch");
  hipStream_t st = as_stream(stream);
  const int ntile = (n + 15) / 16;
  const int blocks = std::min((ntile + 3) / 4, 1024);
  const int nwave = blocks * (TPB / WAVE);
  if ((size_t)nwave * (NPARAM + 1) > m->partial.n) {
    PDX_HIP(hipStreamSynchronize(st));
    int rc = m->partial.alloc((size_t)nwave * (NPARAM + 1));
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(TPB), 0, st, dev_of(m), feats, targets, rows, n, m->partial.p);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((NPARAM + 1 + RED_P - 1) / RED_P), dim3(RED_P * RED_S), 0, st, m->partial.p, nwave,
                     grad ? grad : m->grad.p, loss ? loss : m->loss.p, n);
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

extern "C" int pdx_mlp_train_step(pdx_mlp_t m, const float *feats, const float *targets, const int32_t *rows,
                                  int32_t n, float lr, double *loss, pdx_stream_t stream) {
  int rc = pdx_mlp_gradient(m, feats, targets, rows, n, m->grad.p, loss, stream);
  if (rc) return rc;
  m->step += 1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, as_stream(stream), m->param.p, m->mom.p,
                     m->vel.p, m->grad.p, lr, b1, b2, eps, c1, c2, m->w1.p, m->w2.p, m->w3.p, m->b1.p, m->b2.p,
                     m->b3.p);
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}
