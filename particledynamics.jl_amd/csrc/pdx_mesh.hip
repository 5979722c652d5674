// Smooth PME mesh engine on gfx950: B-spline charge spreading into an int64 fixed-point
// grid, a mixed-radix (2,3,4,5,7,8,11,13,15) Stockham 3D FFT in LDS, the influence-function
// convolution with FP64 energy / virial reductions fused into the z transform, and the
// analytic-differentiation force gather.
//
// Pass structure of one solve (every pass reads and writes the mesh once):
//   k_fft_x_r2c   int64 fixed point -> float, R2C along x
//   k_fft_y       complex FFT along y (forward)
//   k_fft_z       forward z FFT, x G (+ FP64 energy/virial), inverse z FFT   [fused solve]
//   k_fft_y       inverse along y
//   k_fft_x_c2r   C2R along x into the potential grid (unnormalised, as cuFFT's brfft)
// Lines of a pass are staged in LDS as [element*T + line] for T adjacent lines, so global
// loads and stores of a tile are contiguous runs of T complex values.
#include "pdx_internal.hpp"
#include "pdx_prolate.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>

using namespace pdx;

namespace {

constexpr int TPB = 256;
constexpr int MAXSTAGE = 16;

// Mixed-radix Stockham plan of one axis: stage s has radix R_s, p_s = product of the
// radices before it, the magic multiplier ceil(2^32/p_s) (exact quotient of i/p_s for
// i p_s < 2^32) and the twiddle stride K/(p_s R_s).
struct FFTPlan {
  int K;
  int nstage;
  int radix[MAXSTAGE];
  int p[MAXSTAGE];
  unsigned magic[MAXSTAGE];
  int tstride[MAXSTAGE];
  int big;  // a radix of 7, 11 or 13 (compiled into separate, register-heavier kernels)
};

bool make_plan(int K, FFTPlan &pl) {
  pl.K = K;
  pl.nstage = 0;
  int n = K;
  const int rs[9] = {8, 15, 4, 2, 3, 5, 7, 11, 13};
  for (int r : rs) {
    while (n % r == 0) {
      if (pl.nstage >= MAXSTAGE) return false;
      pl.radix[pl.nstage++] = r;
      n /= r;
    }
  }
  int p = 1;
  pl.big = 0;
  for (int s = 0; s < pl.nstage; ++s) {
    if (pl.radix[s] == 7 || pl.radix[s] == 11 || pl.radix[s] == 13) pl.big = 1;
    pl.p[s] = p;
    pl.magic[s] = p > 1 ? (unsigned)((((unsigned long long)1 << 32) + p - 1) / p) : 0u;
    pl.tstride[s] = K / (p * pl.radix[s]);
    p *= pl.radix[s];
  }
  return n == 1;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by sign*i
__device__ __forceinline__ float2 cmuli(float2 a, float sign) { return make_float2(-sign * a.y, sign * a.x); }

// In-place radix-R DFT of u[0..R) with exponent sign `sg` (-1 forward, +1 inverse).
constexpr int MAXR = 16;
template <int R> __device__ __forceinline__ void dft(float2 (&u)[MAXR], float sg);

template <> __device__ __forceinline__ void dft<2>(float2 (&u)[MAXR], float) {
  const float2 a = u[0], b = u[1];
  u[0] = cadd(a, b);
  u[1] = csub(a, b);
}
template <> __device__ __forceinline__ void dft<3>(float2 (&u)[MAXR], float sg) {
  const float h = 0.86602540378443864676f;
  const float2 t = cadd(u[1], u[2]);
  const float2 d = cmuli(csub(u[1], u[2]), sg * h);
  const float2 m = make_float2(u[0].x - 0.5f * t.x, u[0].y - 0.5f * t.y);
  u[0] = cadd(u[0], t);
  u[1] = cadd(m, d);
  u[2] = csub(m, d);
}
template <> __device__ __forceinline__ void dft<4>(float2 (&u)[MAXR], float sg) {
  const float2 t0 = cadd(u[0], u[2]), t1 = csub(u[0], u[2]);
  const float2 t2 = cadd(u[1], u[3]), t3 = cmuli(csub(u[1], u[3]), sg);
  u[0] = cadd(t0, t2);
  u[2] = csub(t0, t2);
  u[1] = cadd(t1, t3);
  u[3] = csub(t1, t3);
}
// Radix 8 as two radix-4 DFTs: a_k = u_k + u_{k+4} gives the even outputs, b_k = (u_k - u_{k+4})
// W8^k the odd ones (W8 = exp(sg i pi / 4)).
template <> __device__ __forceinline__ void dft<8>(float2 (&u)[MAXR], float sg) {
  const float c = 0.70710678118654752440f;
  float2 a[MAXR], b[MAXR];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    a[k] = cadd(u[k], u[k + 4]);
    b[k] = csub(u[k], u[k + 4]);
  }
  b[1] = make_float2(c * (b[1].x - sg * b[1].y), c * (b[1].y + sg * b[1].x));
  b[2] = cmuli(b[2], sg);
  b[3] = make_float2(c * (-b[3].x - sg * b[3].y), c * (-b[3].y + sg * b[3].x));
  dft<4>(a, sg);
  dft<4>(b, sg);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    u[2 * m] = a[m];
    u[2 * m + 1] = b[m];
  }
}
template <> __device__ __forceinline__ void dft<5>(float2 (&u)[MAXR], float sg) {
  const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
  const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
  const float2 a1 = cadd(u[1], u[4]), b1 = csub(u[1], u[4]);
  const float2 a2 = cadd(u[2], u[3]), b2 = csub(u[2], u[3]);
  const float2 u0 = u[0];
  const float2 r1 = make_float2(u0.x + c1 * a1.x + c2 * a2.x, u0.y + c1 * a1.y + c2 * a2.y);
  const float2 r2 = make_float2(u0.x + c2 * a1.x + c1 * a2.x, u0.y + c2 * a1.y + c1 * a2.y);
  const float2 i1 = cmuli(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y), sg);
  const float2 i2 = cmuli(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y), sg);
  u[0] = cadd(u0, cadd(a1, a2));
  u[1] = cadd(r1, i1);
  u[4] = csub(r1, i1);
  u[2] = cadd(r2, i2);
  u[3] = csub(r2, i2);
}
// Radix 15 = 3 x 5 (Cooley-Tukey): n = 5 n1 + n2, k = k1 + 3 k2; radix 3 over n1, the twiddles
// W15^(n2 k1), radix 5 over n2 (W15 = exp(sg 2 pi i / 15), cos / sin of 2 pi m / 15 below).
template <> __device__ __forceinline__ void dft<15>(float2 (&u)[MAXR], float sg) {
  const float c[9] = {1.0f, 0.91354545764260086660f, 0.66913060635885823757f, 0.30901699437494745126f,
                      -0.10452846326765333207f, -0.49999999999999977796f, -0.80901699437494734024f,
                      -0.97814760073380568883f, -0.97814760073380568883f};
  const float s[9] = {0.0f, 0.40673664307580015276f, 0.74314482547739413310f, 0.95105651629515353118f,
                      0.99452189536827340088f, 0.86602540378443870761f, 0.58778525229247324813f,
                      0.20791169081775931482f, -0.20791169081775906502f};
  float2 y[3][5];
#pragma unroll
  for (int n2 = 0; n2 < 5; ++n2) {
    float2 a[MAXR];
    a[0] = u[n2];
    a[1] = u[5 + n2];
    a[2] = u[10 + n2];
    dft<3>(a, sg);
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) {
      float2 v = a[k1];
      if (k1 > 0 && n2 > 0) {
        const int m = n2 * k1;
        v = cmul(v, make_float2(c[m], sg * s[m]));
      }
      y[k1][n2] = v;
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < 3; ++k1) {
    float2 b[MAXR];
#pragma unroll
    for (int n2 = 0; n2 < 5; ++n2) b[n2] = y[k1][n2];
    dft<5>(b, sg);
#pragma unroll
    for (int k2 = 0; k2 < 5; ++k2) u[k1 + 3 * k2] = b[k2];
  }
}
// Odd primes 7, 11, 13: direct O(R^2) sum with the roots cos/sin(2 pi k/R) from a table.
template <int R>
__device__ __forceinline__ void dft_direct(float2 (&u)[MAXR], float sg, const float *c, const float *s) {
  float2 v[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int m = (q * r) % R;
      acc = cadd(acc, cmul(u[r], make_float2(c[m], sg * s[m])));
    }
    v[q] = acc;
  }
#pragma unroll
  for (int q = 0; q < R; ++q) u[q] = v[q];
}
template <> __device__ __forceinline__ void dft<7>(float2 (&u)[MAXR], float sg) {
  const float c[7] = {1.00000000000000000000f, 0.62348980185873359439f, -0.22252093395631433737f, -0.90096886790241903498f, -0.90096886790241914600f, -0.22252093395631458717f, 0.62348980185873337234f};
  const float s[7] = {0.00000000000000000000f, 0.78183148246802980363f, 0.97492791218182361934f, 0.43388373911755823142f, -0.43388373911755800938f, -0.97492791218182361934f, -0.78183148246802991466f};
  dft_direct<7>(u, sg, c, s);
}
template <> __device__ __forceinline__ void dft<11>(float2 (&u)[MAXR], float sg) {
  const float c[11] = {1.00000000000000000000f, 0.84125353283118120551f, 0.41541501300188643508f, -0.14231483827328500480f, -0.65486073394528498959f, -0.95949297361449736865f, -0.95949297361449747967f, -0.65486073394528521163f, -0.14231483827328522684f, 0.41541501300188604651f, 0.84125353283118120551f};
  const float s[11] = {0.00000000000000000000f, 0.54064081745559755543f, 0.90963199535451833011f, 0.98982144188093279524f, 0.75574957435425826890f, 0.28173255684142967104f, -0.28173255684142939348f, -0.75574957435425815788f, -0.98982144188093268422f, -0.90963199535451855215f, -0.54064081745559744441f};
  dft_direct<11>(u, sg, c, s);
}
template <> __device__ __forceinline__ void dft<13>(float2 (&u)[MAXR], float sg) {
  const float c[13] = {1.00000000000000000000f, 0.88545602565320991051f, 0.56806474673115592289f, 0.12053668025532300601f, -0.35460488704253545489f, -0.74851074817110119231f, -0.97094181742605201180f, -0.97094181742605212282f, -0.74851074817110130333f, -0.35460488704253589898f, 0.12053668025532320029f, 0.56806474673115481266f, 0.88545602565321002153f};
  const float s[13] = {0.00000000000000000000f, 0.46472317204376850652f, 0.82298386589365635224f, 0.99270887409805397272f, 0.93501624268541483342f, 0.66312265824079519305f, 0.23931566428755768339f, -0.23931566428755743359f, -0.66312265824079497101f, -0.93501624268541472240f, -0.99270887409805397272f, -0.82298386589365701838f, -0.46472317204376839550f};
  dft_direct<13>(u, sg, c, s);
}

template <int T> constexpr int ilog2() { return T <= 1 ? 0 : 1 + ilog2<T / 2>(); }

// One Stockham stage over T lines stored [element*S + line] (S >= T: a padded row stride);
// twiddles exp(-2 pi i m/K) from LDS (conjugated for the inverse, sg = +1).
template <int R, int T, int S = T>
__device__ __forceinline__ void stage(const float2 *__restrict__ in, float2 *__restrict__ out, int K, int p,
                                      unsigned magic, int tstride, const float2 *__restrict__ tw, float sg) {
  constexpr int TL = ilog2<T>();
  const int KR = K / R;
  const int nb = KR << TL;
  const int bs = blockDim.x;
  // two butterflies per iteration (one for the large radices, which need the registers): the LDS
  // loads of both are issued before either is computed
  constexpr int U = (R == 7 || R >= 11) ? 1 : 2;  // 7, 11, 13, 15: one butterfly per iteration
  for (int idx0 = threadIdx.x; idx0 < nb; idx0 += U * bs) {
    float2 u[U][MAXR];
    int jj[U], ll[U], pp = p;
    bool act[U];
#pragma unroll
    for (int h = 0; h < U; ++h) {
      const int idx = idx0 + h * bs;
      act[h] = idx < nb;
      const int l = idx & (T - 1);
      const int i = act[h] ? idx >> TL : 0;
      const int q = p == 1 ? i : (int)__umulhi((unsigned)i, magic);
      const int k = i - q * p;
      jj[h] = q * p * R + k;
      ll[h] = l;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 v = in[(i + r * KR) * S + l];
        if (r > 0 && k > 0) {
          float2 w = tw[r * k * tstride];
          w.y *= -sg;
          v = cmul(v, w);
        }
        u[h][r] = v;
      }
    }
#pragma unroll
    for (int h = 0; h < U; ++h) {
      if (!act[h]) continue;
      dft<R>(u[h], sg);
#pragma unroll
      for (int r = 0; r < R; ++r) out[(jj[h] + r * pp) * S + ll[h]] = u[h][r];
    }
  }
}

// Full transform of T lines held in buf0; returns the buffer with the result.
// sg = -1 forward, +1 inverse (unnormalised).
template <int T, bool BIG, int S = T>
__device__ float2 *lds_fft(float2 *buf0, float2 *buf1, const FFTPlan &pl, const float2 *__restrict__ tw, float sg) {
  float2 *a = buf0, *b = buf1;
  for (int s = 0; s < pl.nstage; ++s) {
    const int p = pl.p[s];
    const unsigned mg = pl.magic[s];
    const int ts = pl.tstride[s];
    switch (pl.radix[s]) {
      case 8: stage<8, T, S>(a, b, pl.K, p, mg, ts, tw, sg); break;
      case 15: stage<15, T, S>(a, b, pl.K, p, mg, ts, tw, sg); break;
      case 4: stage<4, T, S>(a, b, pl.K, p, mg, ts, tw, sg); break;
      case 2: stage<2, T, S>(a, b, pl.K, p, mg, ts, tw, sg); break;
      case 3: stage<3, T, S>(a, b, pl.K, p, mg, ts, tw, sg); break;
      case 5: stage<5, T, S>(a, b, pl.K, p, mg, ts, tw, sg); break;
      default:
        if (BIG) {
          switch (pl.radix[s]) {
            case 7: stage<7, T, S>(a, b, pl.K, p, mg, ts, tw, sg); break;
            case 11: stage<11, T, S>(a, b, pl.K, p, mg, ts, tw, sg); break;
            default: stage<13, T, S>(a, b, pl.K, p, mg, ts, tw, sg); break;
          }
        }
        break;
    }
    __syncthreads();
    float2 *t = a; a = b; b = t;
  }
  return a;
}

__device__ __forceinline__ void load_twiddles(float2 *twl, const float2 *__restrict__ tw, int K) {
  for (int k = threadIdx.x; k < K; k += blockDim.x) twl[k] = tw[k];
}

// Element e -> (global load, LDS store) for e in [0, n), CH loads in flight per thread: all
// loads of a chunk are issued before the first LDS store waits on them.
template <int CH, typename V, int BS = TPB, typename LOAD, typename STORE>
__device__ __forceinline__ void staged(int n, LOAD load, STORE store) {
  for (int base = 0; base < n; base += CH * BS) {
    V v[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int e = base + c * BS + (int)threadIdx.x;
      if (e < n) v[c] = load(e);
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int e = base + c * BS + (int)threadIdx.x;
      if (e < n) store(e, v[c]);
    }
  }
}

// ---- x pass, forward: 2T real lines per workgroup, packed two at a time as the real and
// imaginary parts of T complex lines (z = a + i b); after the transform
// A[m] = (Z[m] + conj Z[K-m])/2, B[m] = (Z[m] - conj Z[K-m])/(2i). The fixed-point charge
// grid is converted to float as it is read.
template <int T, bool FIXED, bool BIG>
__global__ __launch_bounds__(TPB) void k_fft_x_r2c(const void *src, unsigned long long *clear, float2 *spec,
                                                    int K1, int K1h, int nlines, FFTPlan pl,
                                                    const float2 *__restrict__ tw, double inv_scale) {
  extern __shared__ float2 lds[];
  float2 *twl = lds, *b0 = lds + K1, *b1 = b0 + (size_t)T * K1;
  load_twiddles(twl, tw, K1);
  const int line0 = blockIdx.x * 2 * T;
  const int nl = min(2 * T, nlines - line0);  // real lines of this tile, contiguous in memory
  const size_t g0 = (size_t)line0 * K1;
  const int n = nl * K1;
  if (nl < 2 * T)  // a partial last tile: zero the unused line slots
    for (int e = threadIdx.x; e < T * K1; e += TPB) b0[e] = make_float2(0.f, 0.f);
  __syncthreads();
  auto put = [&](int e, float v) {
    const int lr = e / K1, n1 = e - lr * K1;
    reinterpret_cast<float *>(&b0[n1 * T + (lr >> 1)])[lr & 1] = v;
  };
  if (FIXED) {
    const unsigned long long *g = (const unsigned long long *)src + g0;
    staged<8, float>(n, [&](int e) { return (float)((double)(long long)g[e] * inv_scale); }, put);
  } else {
    const float *g = (const float *)src + g0;
    staged<8, float>(n, [&](int e) { return g[e]; }, put);
  }
  __syncthreads();
  float2 *r = lds_fft<T, BIG>(b0, b1, pl, twl, -1.0f);
  const int nh = 2 * T * K1h;
  for (int e = threadIdx.x; e < nh; e += TPB) {
    const int lr = e / K1h, m = e - lr * K1h;
    const int line = line0 + lr;
    if (line >= nlines) continue;
    const int c = lr >> 1;
    const float2 z = r[m * T + c];
    const float2 w = r[(m == 0 ? 0 : K1 - m) * T + c];
    const float2 o = (lr & 1) ? make_float2(0.5f * (z.y + w.y), 0.5f * (w.x - z.x))
                              : make_float2(0.5f * (z.x + w.x), 0.5f * (z.y - w.y));
    spec[(size_t)line * K1h + m] = o;
  }
}

// ---- x pass, inverse: two half-spectrum lines per complex line (Z = A + i B after each is
// completed by Hermitian symmetry; the imaginary parts of the DC and Nyquist terms are
// dropped, as a C2R transform does); real part -> line a, imaginary part -> line b.
template <int T, bool BIG>
__global__ __launch_bounds__(TPB) void k_fft_x_c2r(const float2 *spec, float *out, int K1, int K1h, int nlines,
                                                   FFTPlan pl, const float2 *__restrict__ tw) {
  extern __shared__ float2 lds[];
  float2 *twl = lds, *b0 = lds + K1, *b1 = b0 + (size_t)T * K1;
  load_twiddles(twl, tw, K1);
  const int line0 = blockIdx.x * 2 * T;
  // stage the tile's half-spectrum lines (contiguous) in the ping-pong buffer b1 first
  const int nl = min(2 * T, nlines - line0);
  const float2 *g = spec + (size_t)line0 * K1h;
  staged<8, float2>(nl * K1h, [&](int e) { return g[e]; }, [&](int e, float2 v) { b1[e] = v; });
  __syncthreads();
  const int n = T * K1;
  for (int e = threadIdx.x; e < n; e += TPB) {
    const int c = e / K1, m = e - c * K1;
    const int la = 2 * c, lb = la + 1;
    const bool conj = m >= K1h;
    const int mm = conj ? K1 - m : m;
    float2 A = la < nl ? b1[la * K1h + mm] : make_float2(0.f, 0.f);
    float2 B = lb < nl ? b1[lb * K1h + mm] : make_float2(0.f, 0.f);
    if (mm == 0 || 2 * mm == K1) { A.y = 0.f; B.y = 0.f; }
    if (conj) { A.y = -A.y; B.y = -B.y; }
    b0[m * T + c] = make_float2(A.x - B.y, A.y + B.x);
  }
  __syncthreads();
  float2 *r = lds_fft<T, BIG>(b0, b1, pl, twl, +1.0f);
  const int nr = 2 * T * K1;
  for (int e = threadIdx.x; e < nr; e += TPB) {
    const int lr = e / K1, n1 = e - lr * K1;
    const int line = line0 + lr;
    if (line >= nlines) continue;
    const float2 z = r[n1 * T + (lr >> 1)];
    out[(size_t)line * K1 + n1] = (lr & 1) ? z.y : z.x;
  }
}

// ---- y pass: tiles of T adjacent m1 at fixed m3; elements m2 at stride K1h
template <int T, bool BIG>
__global__ __launch_bounds__(TPB) void k_fft_y(const float2 *src, float2 *spec, int K1h, int K2, FFTPlan pl,
                                               const float2 *__restrict__ tw, float sg, int ikdim, float kscale,
                                               int K1, const double *psum, int npsum, double *psum_out) {
  extern __shared__ float2 lds[];
  // block 0 first adds the z pass's energy / virial partials (one launch fewer on the path)
  if (psum && blockIdx.x == 0) {
    sum_partials_block<7>(psum, npsum, psum_out, 0);
    __syncthreads();
  }
  float2 *twl = lds, *b0 = lds + K2, *b1 = b0 + (size_t)T * K2;
  load_twiddles(twl, tw, K2);
  const int tiles = (K1h + T - 1) / T;
  const int m10 = (blockIdx.x % tiles) * T;
  const int m3 = blockIdx.x / tiles;
  const size_t base = (size_t)m3 * K2 * K1h;
  constexpr int TL = ilog2<T>();
  const int n = T * K2;
  staged<8, float2>(n, [&](int e) {
    const int l = e & (T - 1), m2 = e >> TL;
    return m10 + l < K1h ? src[base + (size_t)m2 * K1h + m10 + l] : make_float2(0.f, 0.f);
  }, [&](int e, float2 v) { b0[e] = v; });
  __syncthreads();
  if (ikdim) {  // field component: times -i k_d, the Nyquist component of k_d zeroed
    for (int e = threadIdx.x; e < n; e += TPB) {
      const int l = e & (T - 1), m2 = e >> TL;
      const int m1 = m10 + l;
      int sgn;
      if (ikdim == 1) sgn = ((K1 % 2 == 0) && m1 == K1 / 2) ? 0 : m1;
      else sgn = ((K2 % 2 == 0) && m2 == K2 / 2) ? 0 : (m2 <= K2 / 2 ? m2 : m2 - K2);
      const float kd = kscale * (float)sgn;
      const float2 z = b0[e];
      b0[e] = make_float2(z.y * kd, -z.x * kd);
    }
    __syncthreads();
  }
  float2 *r = lds_fft<T, BIG>(b0, b1, pl, twl, sg);
  for (int e = threadIdx.x; e < n; e += TPB) {
    const int l = e & (T - 1), m2 = e >> TL;
    const int m1 = m10 + l;
    if (m1 < K1h) spec[base + (size_t)m2 * K1h + m1] = r[m2 * T + l];
  }
}

struct SolveArgs {
  const float *G;
  const float *Vc;     // virial coefficients (6 x nmodes, SoA) or null: the Gaussian closed form
  float2 *ez;          // ik: receives G b (before the inverse z transform) or null
  size_t nmodes;
  int K1, K2, K3, K1h;
  double twopi_L[3];   // 2 pi / L_d
  double inv4a2;       // 1/(4 alpha^2): Vc_ab = G (delta_ab + c k_a k_b), c = -2/k^2 - 2 inv4a2
  float twopi_Lf[3];   // the same in FP32
  float inv2a2f;       // 2 inv4a2 = 1/(2 alpha^2)
  double *partial;
  double *acc;
};

// The influence multiply of one tile of T z-lines at fixed m2 (lines m1 = m10 + l), held in r
// as [m3*T + l], with the FP64 energy / virial terms (_solve_kernel!, MeshKernels.jl:57-81).
template <int T, bool ENERGY, int S = T>
__device__ __forceinline__ void solve_lines(float2 *r, int m10, int m2, const SolveArgs &s, double (&acc)[7]) {
  const int K1h = s.K1h, K2 = s.K2, K3 = s.K3;
  const size_t plane = (size_t)K1h * K2;
  const size_t base = (size_t)m2 * K1h;
  constexpr int TL = ilog2<T>();
  const int n = T * K3;
    const int s2 = m2 <= K2 / 2 ? m2 : m2 - K2;
    const bool ny2 = (K2 % 2 == 0) && (m2 == K2 / 2);
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
      const int l = e & (T - 1), m3 = e >> TL;
      const int m1 = m10 + l;
      if (m1 >= K1h) continue;
      const size_t idx = base + (size_t)m3 * plane + m1;
      const float Gm = s.G[idx];
      float2 z = r[m3 * S + l];
      if (ENERGY && Gm != 0.0f) {
        // per-mode terms in FP32 (the precision the coefficients are stored in), sums in FP64
        const float w = (m1 == 0 || ((s.K1 % 2 == 0) && m1 == s.K1 / 2)) ? 0.5f : 1.0f;
        const float b2 = w * fmaf(z.x, z.x, z.y * z.y);
        const float a = b2 * Gm;
        const int s3 = m3 <= K3 / 2 ? m3 : m3 - K3;
        const bool ny1 = (s.K1 % 2 == 0) && (m1 == s.K1 / 2);
        const bool ny3 = (K3 % 2 == 0) && (m3 == K3 / 2);
        acc[0] += (double)a;
        if (s.Vc) {  // tabulated -dG/d eps_ab (the optimal influence functions)
#pragma unroll
          for (int c = 0; c < 6; ++c) acc[1 + c] += (double)(b2 * s.Vc[(size_t)c * s.nmodes + idx]);
        } else {
          const float kx = s.twopi_Lf[0] * (float)m1, ky = s.twopi_Lf[1] * (float)s2, kz = s.twopi_Lf[2] * (float)s3;
          const float k2 = fmaf(kx, kx, fmaf(ky, ky, kz * kz));
          const float ac = a * (-2.0f / k2 - s.inv2a2f);
          acc[1] += (double)fmaf(ac * kx, kx, a);
          acc[2] += (double)fmaf(ac * ky, ky, a);
          acc[3] += (double)fmaf(ac * kz, kz, a);
          acc[4] += (ny1 || ny2) ? 0.0 : (double)(ac * kx * ky);
          acc[5] += (ny1 || ny3) ? 0.0 : (double)(ac * kx * kz);
          acc[6] += (ny2 || ny3) ? 0.0 : (double)(ac * ky * kz);
        }
      }
      const float2 gz = cscale(z, Gm);
      if (s.ez) s.ez[idx] = gz;
      r[m3 * S + l] = gz;
    }
}

// ---- z pass: MODE 0 forward only, 1 inverse only, 2 forward + solve + inverse
template <int T, int MODE, bool ENERGY, bool BIG>
__global__ __launch_bounds__(TPB) void k_fft_z(float2 *spec, FFTPlan pl, const float2 *__restrict__ tw,
                                               SolveArgs s) {
  extern __shared__ float2 lds[];
  const int K1h = s.K1h, K2 = s.K2, K3 = s.K3;
  float2 *twl = lds, *b0 = lds + K3, *b1 = b0 + (size_t)T * K3;
  load_twiddles(twl, tw, K3);
  const int tiles = (K1h + T - 1) / T;
  const int m10 = (blockIdx.x % tiles) * T;
  const int m2 = blockIdx.x / tiles;
  const size_t plane = (size_t)K1h * K2;
  const size_t base = (size_t)m2 * K1h;
  constexpr int TL = ilog2<T>();
  const int n = T * K3;
  staged<8, float2>(n, [&](int e) {
    const int l = e & (T - 1), m3 = e >> TL;
    return m10 + l < K1h ? spec[base + (size_t)m3 * plane + m10 + l] : make_float2(0.f, 0.f);
  }, [&](int e, float2 v) { b0[e] = v; });
  __syncthreads();
  float2 *r = b0;
  if (MODE == 0 || MODE == 2) r = lds_fft<T, BIG>(b0, b1, pl, twl, -1.0f);
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  if (MODE == 2) {
    solve_lines<T, ENERGY>(r, m10, m2, s, acc);
    __syncthreads();
    float2 *other = (r == b0) ? b1 : b0;
    r = lds_fft<T, BIG>(r, other, pl, twl, +1.0f);
  } else if (MODE == 1) {
    r = lds_fft<T, BIG>(b0, b1, pl, twl, +1.0f);
  } else if (MODE == 3) {  // ik z field: times -i k3 (Nyquist zeroed), then the inverse z transform
    for (int e = threadIdx.x; e < n; e += TPB) {
      const int m3 = e >> TL;
      const int s3 = ((K3 % 2 == 0) && m3 == K3 / 2) ? 0 : (m3 <= K3 / 2 ? m3 : m3 - K3);
      const float kd = (float)(s.twopi_L[2] * s3);
      const float2 z = b0[e];
      b0[e] = make_float2(z.y * kd, -z.x * kd);
    }
    __syncthreads();
    r = lds_fft<T, BIG>(b0, b1, pl, twl, +1.0f);
  }
  for (int e = threadIdx.x; e < n; e += TPB) {
    const int l = e & (T - 1), m3 = e >> TL;
    const int m1 = m10 + l;
    if (m1 < K1h) spec[base + (size_t)m3 * plane + m1] = r[m3 * T + l];
  }
  if (MODE == 2 && ENERGY) {
    // mesh energy = 1/2 sum_full |b|^2 G = sum_half w |b|^2 G (w = 1/2 on the m1 = 0 and
    // Nyquist planes); virial W_ab = sum_half w |b|^2 Vc_ab
    block_partials<7>(acc, s.partial);
  }
}

// ---- plane passes: one workgroup of PBS threads holds a whole mesh plane in LDS, so the x and
// y transforms of a z plane run in one launch with no round trip through memory between them,
// and the z solve takes a whole (m1, m3) plane. Lines are padded to TP (a power of two >= the
// lines of a plane); the LDS holds two TP x KM ping-pong buffers and the twiddles of the axes.
// The per-line arithmetic (stages, twiddles, R2C / C2R packing) is that of the line passes, so
// the potential grid is the same bit for bit.
constexpr int PBS = 1024;  // the default workgroup size of the plane passes (512 selectable)

// Sum of NV values per block over nblocks partials in a fixed order, for any block size (a
// multiple of 64, at most 1024); the whole block calls it.
template <int NV>
__device__ void sum_partials_any(const double *__restrict__ partial, int nblocks, double *out) {
  __shared__ double sh[NV][16];
  double acc[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) acc[c] = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
    for (int c = 0; c < NV; ++c) acc[c] += partial[(size_t)b * NV + c];
  }
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    const double v = wave_sum(acc[c]);
    if (lane == 0) sh[c][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double v = 0.0;
    for (int w = 0; w < nw; ++w) v += sh[threadIdx.x][w];
    out[threadIdx.x] = v;
  }
  __syncthreads();
}

// The plane passes are the critical path of an evaluation whose pair kernel runs beside them:
// their waves take the SIMD's issue slots first (wave priority 3 of 0..3; the pair kernel's
// waves keep 0 and use what is left), so sharing a CU slows the transform less.
__device__ __forceinline__ void mesh_chain_priority() { __builtin_amdgcn_s_setprio(3); }

// Forward x (R2C, two real lines per complex line) and y transforms of z plane n3 of the
// fixed-point charge grid into the half spectrum.
template <int TP, bool BIG, int BS>
__global__ __launch_bounds__(BS) void k_plane_fwd(const unsigned long long *__restrict__ grid, float2 *spec, int K1,
                                                   int K2, int K1h, int KM, FFTPlan p1, FFTPlan p2,
                                                   const float2 *__restrict__ tw, double inv_scale) {
  extern __shared__ float2 lds[];
  mesh_chain_priority();
  constexpr int TS = TP + 1;  // padded row stride: column walks of the transposes hit distinct banks
  float2 *tw1 = lds, *tw2 = lds + K1, *b0 = tw2 + K2, *b1 = b0 + (size_t)KM * TS;
  for (int k = threadIdx.x; k < K1 + K2; k += BS) lds[k] = tw[k];
  const int n3 = blockIdx.x;
  const unsigned long long *g = grid + (size_t)n3 * K1 * K2;
  const int n = K1 * K2;
  if (K2 & 1)  // the odd last real line has no partner: its imaginary slot stays zero
    for (int n1 = threadIdx.x; n1 < K1; n1 += BS) b0[n1 * TS + (K2 >> 1)] = make_float2(0.f, 0.f);
  __syncthreads();
  staged<16, float, BS>(n, [&](int e) { return (float)((double)(long long)g[e] * inv_scale); },
                         [&](int e, float v) {
                           const int n2 = e / K1, n1 = e - n2 * K1;
                           reinterpret_cast<float *>(&b0[n1 * TS + (n2 >> 1)])[n2 & 1] = v;
                         });
  __syncthreads();
  float2 *r = lds_fft<TP, BIG, TS>(b0, b1, p1, tw1, -1.0f);
  float2 *o = r == b0 ? b1 : b0;
  // A[m] = (Z[m] + conj Z[K-m])/2 (line 2c), B[m] = (Z[m] - conj Z[K-m])/(2i) (line 2c+1), stored
  // as the y lines: element n2, line m1
  for (int e = threadIdx.x; e < K2 * K1h; e += BS) {
    const int n2 = e / K1h, m = e - n2 * K1h;
    const int c = n2 >> 1;
    const float2 z = r[m * TS + c];
    const float2 w = r[(m == 0 ? 0 : K1 - m) * TS + c];
    o[n2 * TS + m] = (n2 & 1) ? make_float2(0.5f * (z.y + w.y), 0.5f * (w.x - z.x))
                              : make_float2(0.5f * (z.x + w.x), 0.5f * (z.y - w.y));
  }
  __syncthreads();
  r = lds_fft<TP, BIG, TS>(o, r, p2, tw2, -1.0f);
  float2 *out = spec + (size_t)n3 * K2 * K1h;
  for (int e = threadIdx.x; e < K2 * K1h; e += BS) {
    const int m2 = e / K1h, m1 = e - m2 * K1h;
    out[e] = r[m2 * TS + m1];
  }
}

// Forward z transform, influence multiply with the FP64 energy / virial partials, inverse z
// transform of TZ lines m1 = m10 .. m10 + TZ - 1 of the (m1, m3) plane m2: `tiles` workgroups per
// plane, so that the launch spans every CU.
template <int TZ, bool ENERGY, bool BIG, int BS>
__global__ __launch_bounds__(BS) void k_plane_z(float2 *spec, FFTPlan p3, const float2 *__restrict__ tw,
                                                 SolveArgs s, int tiles) {
  extern __shared__ float2 lds[];
  mesh_chain_priority();
  const int K1h = s.K1h, K2 = s.K2, K3 = s.K3;
  constexpr int TS = TZ + 1;
  float2 *twl = lds, *b0 = lds + K3, *b1 = b0 + (size_t)TS * K3;
  for (int k = threadIdx.x; k < K3; k += BS) twl[k] = tw[k];
  const int m2 = blockIdx.x / tiles, m10 = (blockIdx.x % tiles) * TZ;
  const int nl = min(TZ, K1h - m10);
  const size_t plane = (size_t)K1h * K2;
  const size_t base = (size_t)m2 * K1h + m10;
  const int n = K3 * nl;
  staged<8, float2, BS>(n, [&](int e) {
    const int m3 = e / nl, l = e - m3 * nl;
    return spec[base + (size_t)m3 * plane + l];
  }, [&](int e, float2 v) {
    const int m3 = e / nl, l = e - m3 * nl;
    b0[m3 * TS + l] = v;
  });
  __syncthreads();
  float2 *r = lds_fft<TZ, BIG, TS>(b0, b1, p3, twl, -1.0f);
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  solve_lines<TZ, ENERGY, TS>(r, m10, m2, s, acc);
  __syncthreads();
  r = lds_fft<TZ, BIG, TS>(r, r == b0 ? b1 : b0, p3, twl, +1.0f);
  for (int e = threadIdx.x; e < n; e += BS) {
    const int m3 = e / nl, l = e - m3 * nl;
    spec[base + (size_t)m3 * plane + l] = r[m3 * TS + l];
  }
  if (ENERGY) block_partials<7>(acc, s.partial);
}

// Inverse y and x (C2R) transforms of z plane m3 into the potential grid; block 0 first adds
// the z pass's energy / virial partials.
template <int TP, bool BIG, int BS>
__global__ __launch_bounds__(BS) void k_plane_inv(const float2 *__restrict__ spec, float *pot, int K1, int K2,
                                                   int K1h, int KM, FFTPlan p1, FFTPlan p2,
                                                   const float2 *__restrict__ tw, const double *psum, int npsum,
                                                   double *psum_out) {
  extern __shared__ float2 lds[];
  mesh_chain_priority();
  if (psum && blockIdx.x == 0) sum_partials_any<7>(psum, npsum, psum_out);
  constexpr int TS = TP + 1;  // padded row stride: column walks of the transposes hit distinct banks
  float2 *tw1 = lds, *tw2 = lds + K1, *b0 = tw2 + K2, *b1 = b0 + (size_t)KM * TS;
  for (int k = threadIdx.x; k < K1 + K2; k += BS) lds[k] = tw[k];
  const int m3 = blockIdx.x;
  const float2 *in = spec + (size_t)m3 * K2 * K1h;
  staged<8, float2, BS>(K2 * K1h, [&](int e) { return in[e]; }, [&](int e, float2 v) {
    const int m2 = e / K1h, m1 = e - m2 * K1h;
    b0[m2 * TS + m1] = v;
  });
  __syncthreads();
  float2 *r = lds_fft<TP, BIG, TS>(b0, b1, p2, tw2, +1.0f);
  float2 *o = r == b0 ? b1 : b0;
  // x lines: Z = A + i B with A, B the half-spectrum lines n2 = 2c, 2c+1 completed by Hermitian
  // symmetry (imaginary parts of the DC and Nyquist terms dropped)
  const int npair = (K2 + 1) >> 1;
  for (int e = threadIdx.x; e < npair * K1; e += BS) {
    const int c = e / K1, m = e - c * K1;
    const int la = 2 * c, lb = la + 1;
    const bool conj = m >= K1h;
    const int mm = conj ? K1 - m : m;
    float2 A = r[la * TS + mm];
    float2 B = lb < K2 ? r[lb * TS + mm] : make_float2(0.f, 0.f);
    if (mm == 0 || 2 * mm == K1) { A.y = 0.f; B.y = 0.f; }
    if (conj) { A.y = -A.y; B.y = -B.y; }
    o[m * TS + c] = make_float2(A.x - B.y, A.y + B.x);
  }
  __syncthreads();
  r = lds_fft<TP, BIG, TS>(o, r, p1, tw1, +1.0f);
  float *out = pot + (size_t)m3 * K1 * K2;
  for (int e = threadIdx.x; e < K1 * K2; e += BS) {
    const int n2 = e / K1, n1 = e - n2 * K1;
    const float2 z = r[n1 * TS + (n2 >> 1)];
    out[e] = (n2 & 1) ? z.y : z.x;
  }
}

// ---- B-spline weights (window_stencil, src/electrostatics/Windows.jl:209-223): order-P
// cardinal B-spline weights w[j] of nodes k0 + j (ascending) and derivatives dw[j].
// Every kernel that spreads or gathers must get bitwise the same weights (the tiled spread equals
// the atomic one bit for bit), so no product is left for the compiler to fuse or not as the
// surrounding code suggests: contraction is off and the fused steps are written as fmaf.
template <int P>
__device__ __forceinline__ int bspline(float u, float (&w)[P], float (&dw)[P]) {
#pragma clang fp contract(off)
  const float fl = floorf(u);
  const float t = u - fl;
  float a[P];  // a[j] = M_n(t + j), built up from M_2
  a[0] = t;
  a[1] = 1.0f - t;
#pragma unroll
  for (int j = 2; j < P; ++j) a[j] = 0.f;
#pragma unroll
  for (int n = 2; n < P - 1; ++n) {  // M_n -> M_{n+1}
    const float inv = 1.0f / (float)n;
#pragma unroll
    for (int j = n; j >= 0; --j) {
      const float left = j < n ? a[j] : 0.f;
      const float right = j >= 1 ? a[j - 1] : 0.f;
      a[j] = fmaf(t + j, left, ((float)(n + 1) - t - j) * right) * inv;
    }
  }
  // a = M_{P-1}(t + j), j = 0..P-2; final step to M_P and derivatives
  const float inv = 1.0f / (float)(P - 1);
  float full[P];
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const float left = j < P - 1 ? a[j] : 0.f;
    const float right = j >= 1 ? a[j - 1] : 0.f;
    full[j] = fmaf(t + j, left, ((float)P - t - j) * right) * inv;
  }
#pragma unroll
  for (int i = 0; i < P; ++i) {
    const int j = P - 1 - i;  // node k0 + i carries M_P(t + j)
    w[i] = full[j];
    const float left = j <= P - 2 ? a[j] : 0.f;
    const float right = j >= 1 ? a[j - 1] : 0.f;
    dw[i] = left - right;
  }
  return (int)fl - (P - 1);
}

// ---- window codes of the spread and gather kernels: WP = P, the order-P B-spline, or
// PROLATE + P, the prolate window of width P (ESP). Both windows cover P nodes ending at a base
// node: the B-spline's floor(u), the prolate window's centred stencil floor(u) + P/2 (even P) or
// floor(u + 1/2) + (P-1)/2 (odd P), i.e. the base of u + P/2. The tile spread and gather own a
// particle by its base node, so they serve both windows alike.
constexpr int PROLATE = 16;
__host__ __device__ constexpr int win_width(int WP) { return WP & 15; }
__host__ __device__ constexpr bool win_prolate(int WP) { return WP >= PROLATE; }
// the shift of the stencil against the B-spline's (the base node of u is that of u + shift)
template <int WP> __device__ __forceinline__ float win_shift() { return win_prolate(WP) ? 0.5f * win_width(WP) : 0.f; }

template <int WP>
__device__ __forceinline__ int window_base(float u) {
  constexpr int P = win_width(WP);
  if (!win_prolate(WP)) return (int)floorf(u);
  if (P % 2 == 0) return (int)floorf(u) + P / 2;
  return (int)floorf(u + 0.5f) + (P - 1) / 2;
}

// Prolate window weights (window_stencil of ProlatePolyStencil, src/electrostatics/Windows.jl:
// 158-166,231-250): tau = offset of u from the centre of its cell; stored node h carries
// E_h(t) + tau O_h(t) (t = tau^2), its mirror node P-1-h carries E_h(t) - tau O_h(t); the
// u-derivatives tau Ed_h(t) +- Od_h(t) are the exact derivatives of the same polynomials.
// Contraction is off and every fused step is an explicit fmaf, as in bspline above.
template <int P>
__device__ __forceinline__ int prolate_weights(float u, float (&w)[P], float (&dw)[P], const WinCoef &c) {
#pragma clang fp contract(off)
  constexpr int H = (P + 1) / 2;
  float tau;
  int k0;
  if (P % 2 == 0) {
    const float fl = floorf(u);
    tau = (u - fl) - 0.5f;
    k0 = (int)fl - (P / 2 - 1);
  } else {
    const float fl = floorf(u + 0.5f);
    tau = u - fl;
    k0 = (int)fl - (P - 1) / 2;
  }
  const float t = tau * tau;
#pragma unroll
  for (int h = 0; h < H; ++h) {
    float E = c.e[WIN_TERMS - 1][h], O = c.o[WIN_TERMS - 1][h];
    float Ed = c.ed[WIN_TERMS - 1][h], Od = c.od[WIN_TERMS - 1][h];
#pragma unroll
    for (int j = WIN_TERMS - 2; j >= 0; --j) {
      E = fmaf(E, t, c.e[j][h]);
      O = fmaf(O, t, c.o[j][h]);
      Ed = fmaf(Ed, t, c.ed[j][h]);
      Od = fmaf(Od, t, c.od[j][h]);
    }
    w[h] = fmaf(tau, O, E);
    dw[h] = fmaf(tau, Ed, Od);
    if (P - 1 - h != h) {
      w[P - 1 - h] = fmaf(-tau, O, E);
      dw[P - 1 - h] = fmaf(tau, Ed, -Od);
    }
  }
  return k0;
}

// Weights and u-derivatives of the P = win_width(WP) nodes k0 .. k0 + P - 1 (k0 returned,
// unwrapped: k0 + P - 1 = window_base<WP>(u)).
template <int WP>
__device__ __forceinline__ int window(float u, float (&w)[win_width(WP)], float (&dw)[win_width(WP)],
                                      const WinCoef &wc) {
  if constexpr (win_prolate(WP)) return prolate_weights<win_width(WP)>(u, w, dw, wc);
  else return bspline<win_width(WP)>(u, w, dw);
}

// v rounded to the nearest integer (ties to even) as a 64-bit integer, for |v| < 2^51: the FP64
// sum v + 1.5 * 2^52 lies in [2^52, 2^53), where the spacing of doubles is 1, so the addition
// itself rounds v and leaves it in the low mantissa bits. The same value as __float2ll_rn in
// four instructions instead of a dozen. The spread's fixed-point scale keeps every charge
// contribution q w_x w_y w_z 2^S below 2^50.
__device__ __forceinline__ long long fixed_round(float v) {
  constexpr double M = 6755399441055744.0;  // 1.5 * 2^52
  return __double_as_longlong((double)v + M) - __double_as_longlong(M);
}


template <int P>
__device__ __forceinline__ float sel(const float (&v)[P], int j) {
  float r = v[0];
#pragma unroll
  for (int k = 1; k < P; ++k) r = (j == k) ? v[k] : r;
  return r;
}

constexpr int lanes_for(int P) { return P * P <= 16 ? 16 : (P * P <= 32 ? 32 : 64); }
// gather: 8 lanes per particle, each walking several stencil lines, so the B-spline weights are
// computed by 8 lanes instead of P^2 and more loads are in flight per lane
constexpr int gather_lanes(int) { return 8; }

// Sum over the eight lanes 8k .. 8k+7 of a wave64, every lane receiving it: quad permutations
// (xor 1, xor 2) and the half-row mirror (lane l <-> 7 - l) as DPP operands of three adds.
__device__ __forceinline__ float sum8_lanes(float v) {
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false));
  return v;
}

// ---- spread: LPP lanes per particle, lane l owns the (x, y) line (l % P, l / P) of the
// particle's P x P x P stencil and walks its P z nodes with 64-bit integer atomics (the
// sum is exact and independent of the order of arrival).
template <int WP>
__device__ __forceinline__ void spread_particle(int lane, float px, float py, float pz, float qi,
                                                unsigned long long *grid, int K1, int K2, int K3, float Lx,
                                                float Ly, float Lz, float scale, const WinCoef &wc) {
  constexpr int P = win_width(WP);
  constexpr int LPP = lanes_for(P);
  float wx[P], wy[P], wz[P], d[P];
  const int k1 = window<WP>(grid_coordinate(px, Lx, K1), wx, d, wc);
  const int k2 = window<WP>(grid_coordinate(py, Ly, K2), wy, d, wc);
  const int k3 = window<WP>(grid_coordinate(pz, Lz, K3), wz, d, wc);
  const size_t KK = (size_t)K1 * K2;
  for (int l = lane; l < P * P; l += LPP) {
    const int j1 = l % P, j2 = l / P;
    const int n1 = wrap_node(k1 + j1, K1), n2 = wrap_node(k2 + j2, K2);
    const float q12 = qi * sel<P>(wx, j1) * sel<P>(wy, j2) * scale;
    const size_t b = (size_t)n1 + (size_t)K1 * n2;
#pragma unroll
    for (int j3 = 0; j3 < P; ++j3) {
      const int n3 = wrap_node(k3 + j3, K3);
      const long long v = fixed_round(q12 * wz[j3]);
      atomicAdd(&grid[b + KK * n3], (unsigned long long)v);
    }
  }
}

template <int WP>
__global__ __launch_bounds__(TPB) void k_spread(int N, const float *__restrict__ x, const float *__restrict__ y,
                                                const float *__restrict__ z, const float *__restrict__ q,
                                                const int32_t *__restrict__ order, unsigned long long *grid,
                                                int K1, int K2, int K3, float Lx, float Ly, float Lz,
                                                float scale, const WinCoef wc) {
  constexpr int LPP = lanes_for(win_width(WP));
  const int tid = blockIdx.x * TPB + threadIdx.x;
  const int t = tid / LPP;
  const int lane = tid % LPP;
  if (t >= N) return;
  const int i = order ? order[t] - 1 : t;
  const float qi = q[i];
  if (qi == 0.0f) return;
  spread_particle<WP>(lane, x[i], y[i], z[i], qi, grid, K1, K2, K3, Lx, Ly, Lz, scale, wc);
}

// ---- fallback of the snapshot spread: when the snapshot spread found a particle further from
// its snapshot cell than the snapshot's margin (stale[3] != 0, written by that spread, whose
// tiles then only cleared the grid), every particle of `packed` is spread with global 64-bit
// atomics, as k_spread; otherwise every workgroup leaves at once. Grid-stride, so a small grid
// suffices for this rare path.
template <int WP>
__global__ __launch_bounds__(TPB) void k_spread_stale(int N, const float4 *__restrict__ packed,
                                                      const int64_t *__restrict__ stale, unsigned long long *grid,
                                                      int K1, int K2, int K3, float Lx, float Ly, float Lz,
                                                      float scale, const WinCoef wc) {
  if (stale[3] == 0) return;
  constexpr int LPP = lanes_for(win_width(WP));
  const int64_t total = (int64_t)N * LPP;
  for (int64_t tid = (int64_t)blockIdx.x * TPB + threadIdx.x; tid < total; tid += (int64_t)gridDim.x * TPB) {
    const int i = (int)(tid / LPP), lane = (int)(tid % LPP);
    const float4 p = packed[i];
    if (p.w != 0.0f) spread_particle<WP>(lane, p.x, p.y, p.z, p.w, grid, K1, K2, K3, Lx, Ly, Lz, scale, wc);
  }
}

__global__ void k_zero64(unsigned long long *g, size_t n) {
  const size_t k = (size_t)blockIdx.x * TPB + threadIdx.x;
  if (k < n) g[k] = 0ull;
}

// ---- spread, owner computes: one workgroup per TILE^3 block of grid nodes, in two forms.
// Pull form (k_spread_tiles): the workgroup scans the cells of the neighbour list that can hold a
// particle whose P^3 stencil reaches the block (a particle at u covers nodes floor(u)-P+1 ..
// floor(u); cells are widened by the list's skin/2, the most any particle moved since it was
// binned), accumulates q w_x w_y w_z of the stencil nodes inside the block in 64-bit fixed point
// with LDS integer atomics (exact, order-independent) and stores the block with plain writes: no
// global atomics, and no clearing (every node is written). Up to SLIST reaching particles are
// staged in LDS; beyond that each is spread by its own thread (dense tiles, which take the owner
// form wherever its blocks fit).
// Owner form (k_spread_own, then k_spread_fold): the workgroup takes only the particles whose
// base node lies in the tile, spreads their whole stencils into an LDS block of the tile and its
// P - 1 lower halo nodes, and writes the block out; the fold then sums the blocks' overlapping
// entries per node. Each particle is handled by one workgroup (the pull form handles it in every
// tile its stencil reaches, about 2.3 at 16-node tiles), at the price of the blocks' writes and
// the fold, which only dense meshes repay.
// Tile edge: 16 nodes, or 8 on meshes too small to give 256 tiles of 16 (one per CU).
// Each block also lists, for the tile gather, the particles whose base node floor(u) lies in
// its tile (every particle in exactly one tile): entries (u1, u2, u3, q) and the particle index,
// up to tcap of them, and the count (which may exceed tcap) in tn[tile].
// One particle straight into the tile's accumulator (o = floor(u) - tile origin, wrapped), in the
// product order of k_spread; the path of particles beyond the staged list.
template <int WP, int STILE>
__device__ __forceinline__ void spread_direct(unsigned long long *acc, float u1, float u2, float u3,
                                                        float q, int o1, int o2, int o3, int e1, int e2, int e3,
                                                        int K1, int K2, int K3, float scale, const WinCoef &wc) {
  constexpr int P = win_width(WP);
  constexpr int SROW = STILE + 1;
  float wx[P], wy[P], wz[P], d[P];
  window<WP>(u1, wx, d, wc);
  window<WP>(u2, wy, d, wc);
  window<WP>(u3, wz, d, wc);
  for (int j2 = 0; j2 < P; ++j2) {
    int l2 = o2 - (P - 1) + j2; l2 += l2 < 0 ? K2 : 0; l2 -= l2 >= K2 ? K2 : 0;
    if (l2 >= e2) continue;
    for (int j1 = 0; j1 < P; ++j1) {
      int l1 = o1 - (P - 1) + j1; l1 += l1 < 0 ? K1 : 0; l1 -= l1 >= K1 ? K1 : 0;
      if (l1 >= e1) continue;
      const float q12 = q * wx[j1] * wy[j2] * scale;
      for (int j3 = 0; j3 < P; ++j3) {
        int l3 = o3 - (P - 1) + j3; l3 += l3 < 0 ? K3 : 0; l3 -= l3 >= K3 ? K3 : 0;
        if (l3 >= e3) continue;
        atomicAdd(&acc[(l3 * STILE + l2) * SROW + l1], (unsigned long long)fixed_round(q12 * wz[j3]));
      }
    }
  }
}

constexpr int SLIST = 768;   // particles reaching one block staged in LDS (a few hundred in practice)
constexpr int SCHUNK = 128;  // staged particles whose weights are computed together
template <int WP, int STILE, int SBS>
// (1024-thread workgroups of B-spline orders up to 5 are held to 64 VGPRs, 8 waves per SIMD: two
// of them, 32 waves, then fit on a CU; higher orders would spill)
__global__ __launch_bounds__(SBS, (SBS == 1024 && WP <= 5) ? 8 : 1) void k_spread_tiles(const float4 *__restrict__ packed,
                                                      const int32_t *__restrict__ offsets,
                                                      const int32_t *__restrict__ ids,
                                                      const int64_t *__restrict__ state, int nx, int ny,
                                                      int nz, float margin, unsigned long long *grid, int K1,
                                                      int K2, int K3, float Lx, float Ly, float Lz, float scale,
                                                      float4 *__restrict__ tlist, int32_t *__restrict__ tids,
                                                      int32_t *__restrict__ tn, int tcap,
                                                      const float *__restrict__ disp, int ndisp, float stale2,
                                                      int64_t *stale_word, const WinCoef wc) {
  if (state[PDX_DN_ERROR] != 0) return;  // no usable cells / positions: the grid keeps its last value
  constexpr int P = win_width(WP);
  constexpr int LPP = lanes_for(P);
  // x rows padded to STILE + 1 nodes: the P x P lanes of a particle then spread over the LDS banks
  constexpr int SROW = STILE + 1;
  __shared__ unsigned long long acc[STILE * STILE * SROW];
  __shared__ float4 list[SLIST];
  __shared__ int cstart[64], cpre[65];
  __shared__ int nlist;
  const int tx = (K1 + STILE - 1) / STILE, ty = (K2 + STILE - 1) / STILE;
  const int b = blockIdx.x;
  const int a1 = (b % tx) * STILE, a2 = ((b / tx) % ty) * STILE, a3 = (b / (tx * ty)) * STILE;
  const int e1 = min(STILE, K1 - a1), e2 = min(STILE, K2 - a2), e3 = min(STILE, K3 - a3);
  for (int k = threadIdx.x; k < STILE * STILE * SROW; k += SBS) acc[k] = 0ull;
  // snapshot cells (disp != null): the check's per-block maxima of |x - x_build|^2 bound the
  // motion since the snapshot's build. A maximum beyond the margin (or -1: not measured) makes
  // the cells unusable: every tile then only clears its nodes, and the fallback spread and the
  // per-particle gather take over for this evaluation (word 3 of the snapshot state).
  bool stale = false;
  if (disp) {
    int s = 0;
    for (int k = threadIdx.x; k < ndisp; k += SBS) {
      const float d = disp[k];
      s |= !(d >= 0.0f && d <= stale2);
    }
    stale = __syncthreads_or(s) != 0;
    if (b == 0 && threadIdx.x == 0) stale_word[3] = stale ? 1 : 0;
  }
  // candidate cells per dimension, in the list's frame (x + L/2) / L: a particle at u reaches the
  // block when the base node of u + shift lies in [A, A + E + P - 2]
  const int K[3] = {K1, K2, K3}, A[3] = {a1, a2, a3}, E[3] = {e1, e2, e3}, NC[3] = {nx, ny, nz};
  const float Ls[3] = {Lx, Ly, Lz};
  int c0[3], cn[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float f0 = ((float)A[d] - win_shift<WP>()) / (float)K[d] + 0.5f - margin / Ls[d] - 1e-5f;
    const float f1 = ((float)(A[d] + E[d] + P - 1) - win_shift<WP>()) / (float)K[d] + 0.5f + margin / Ls[d] + 1e-5f;
    const int lo = (int)floorf(f0 * NC[d]);
    c0[d] = lo;
    cn[d] = min((int)floorf(f1 * NC[d]) - lo + 1, NC[d]);
  }
  const int ncand = cn[0] * cn[1] * cn[2];
  // (the cell range is at most 64 cells: a block plus its halo spans < 4 cells per dimension
  // whenever the cells are wider than (STILE + P + skin) / 3 mesh spacings; otherwise the
  // remaining cells are taken in further rounds of 64)
  __shared__ int nown;
  if (threadIdx.x == 0) {
    nlist = 0;
    nown = 0;
  }
  float4 *own = tlist + (size_t)b * tcap;
  int32_t *oid = tids + (size_t)b * tcap;
  // one candidate particle: listed for the gather when its base node lies in the tile, staged
  // for the spread when its stencil reaches the tile (and its charge is nonzero)
  auto take = [&](int i, const float4 p) {
    const float u1 = grid_coordinate(p.x, Lx, K1), u2 = grid_coordinate(p.y, Ly, K2),
                u3 = grid_coordinate(p.z, Lz, K3);
    const int o1 = wrap_node(window_base<WP>(u1) - a1, K1);
    const int o2 = wrap_node(window_base<WP>(u2) - a2, K2);
    const int o3 = wrap_node(window_base<WP>(u3) - a3, K3);
    if (o1 < e1 && o2 < e2 && o3 < e3) {
      const int k = atomicAdd(&nown, 1);
      if (k < tcap) {
        own[k] = make_float4(u1, u2, u3, p.w);
        oid[k] = i;
      }
    }
    if (p.w == 0.0f) return;
    if (o1 > e1 + P - 2 || o2 > e2 + P - 2 || o3 > e3 + P - 2) return;
    const int slot = atomicAdd(&nlist, 1);
    if (slot < SLIST) {
      list[slot] = make_float4(u1, u2, u3, p.w);
      return;
    }
    // list full: this thread spreads the particle itself
    spread_direct<WP, STILE>(acc, u1, u2, u3, p.w, o1, o2, o3, e1, e2, e3, K1, K2, K3, scale, wc);
  };
  for (int round = 0; !stale && round * 64 < ncand; ++round) {
    const int nc = min(64, ncand - round * 64);
    if (threadIdx.x < 64) {  // wave 0: the round's cells and the inclusive scan of their sizes
      int cnt = 0;
      if (threadIdx.x < nc) {
        const int cc = round * 64 + threadIdx.x;
        int cx = c0[0] + cc % cn[0], cy = c0[1] + (cc / cn[0]) % cn[1], cz = c0[2] + cc / (cn[0] * cn[1]);
        cx = ((cx % nx) + nx) % nx;
        cy = ((cy % ny) + ny) % ny;
        cz = ((cz % nz) + nz) % nz;
        const int c = (cz * ny + cy) * nx + cx;
        cstart[threadIdx.x] = offsets[c] - 1;
        cnt = offsets[c + 1] - offsets[c];
      }
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(cnt, o, 64);
        if (threadIdx.x >= o) cnt += u;
      }
      cpre[threadIdx.x + 1] = cnt;
      if (threadIdx.x == 0) cpre[0] = 0;
    }
    __syncthreads();
    const int total = cpre[nc];
    // flat candidate index f -> its cell c (the last c < nc with cpre[c] <= f: six halvings)
    // -> particle; CB candidates per thread and pass, their loads issued together (two memory
    // round trips per pass: ids, then positions)
    auto cell_of = [&](int f) {
      int c = 0;
#pragma unroll
      for (int h = 32; h >= 1; h >>= 1)
        if (c + h < nc && cpre[c + h] <= f) c += h;
      return cstart[c] + (f - cpre[c]);
    };
    constexpr int CB = 2;
    for (int f0 = threadIdx.x; f0 < total; f0 += CB * SBS) {
      int ic[CB];
      float4 pc[CB];
#pragma unroll
      for (int u = 0; u < CB; ++u) {
        const int f = f0 + u * SBS;
        ic[u] = f < total ? ids[cell_of(f)] - 1 : -1;
      }
#pragma unroll
      for (int u = 0; u < CB; ++u) pc[u] = packed[ic[u] >= 0 ? ic[u] : ic[0]];
#pragma unroll
      for (int u = 0; u < CB; ++u)
        if (ic[u] >= 0) take(ic[u], pc[u]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) tn[b] = stale ? 0 : nown;
  // staged particles, SCHUNK at a time: per particle and dimension (three threads each) the
  // B-spline weights and the tile-local index of each stencil node (a byte; 0xFF: outside the
  // tile) go to LDS; then LPP lanes per particle, lane = one (x, y) line of the stencil walking
  // its P z nodes. The products and their order are those of k_spread: ((q w_x) w_y) 2^S, then
  // times w_z, rounded to the nearest integer.
  const int nl = stale ? 0 : min(nlist, SLIST);
  __shared__ float wsh[SCHUNK][3][8];
  __shared__ uint2 lsh[SCHUNK][3];
  const int grp = threadIdx.x / LPP, lane = threadIdx.x % LPP;
  auto node_byte = [](const uint2 x, int j) { return ((j < 4 ? x.x : x.y) >> (8 * (j & 3))) & 0xFFu; };
  for (int c0 = 0; c0 < nl; c0 += SCHUNK) {
    const int nc = min(SCHUNK, nl - c0);
    __syncthreads();
    for (int e = threadIdx.x; e < 3 * nc; e += SBS) {
      const int t = e / 3, d = e - 3 * t;
      const float4 u = list[c0 + t];
      const int K = d == 0 ? K1 : (d == 1 ? K2 : K3);
      const int a = d == 0 ? a1 : (d == 1 ? a2 : a3);
      const int ee = d == 0 ? e1 : (d == 1 ? e2 : e3);
      float w[P], dw[P];
      const int k = window<WP>(d == 0 ? u.x : (d == 1 ? u.y : u.z), w, dw, wc);
      unsigned lo = 0xFFFFFFFFu, hi = 0xFFFFFFFFu;
#pragma unroll
      for (int j = 0; j < P; ++j) {
        wsh[t][d][j] = w[j];
        int l = wrap_node(k + j, K) - a;
        l += l < 0 ? K : 0;
        const unsigned bj = l < ee ? (unsigned)l : 0xFFu;
        if (j < 4) lo = (lo & ~(0xFFu << (8 * j))) | (bj << (8 * j));
        else hi = (hi & ~(0xFFu << (8 * (j - 4)))) | (bj << (8 * (j - 4)));
      }
      lsh[t][d] = make_uint2(lo, hi);
    }
    __syncthreads();
    for (int t = grp; t < nc; t += SBS / LPP) {
      const float qv = list[c0 + t].w;
      for (int l = lane; l < P * P; l += LPP) {
        const int j1 = l % P, j2 = l / P;
        const unsigned l1 = node_byte(lsh[t][0], j1), l2 = node_byte(lsh[t][1], j2);
        if (l1 == 0xFFu || l2 == 0xFFu) continue;
        const float q12 = qv * wsh[t][0][j1] * wsh[t][1][j2] * scale;
        float wz[P];
#pragma unroll
        for (int j3 = 0; j3 < P; ++j3) wz[j3] = wsh[t][2][j3];
        const uint2 x3 = lsh[t][2];
        unsigned long long *line = acc + l2 * SROW + l1;
#pragma unroll
        for (int j3 = 0; j3 < P; ++j3) {
          const unsigned l3 = node_byte(x3, j3);
          if (l3 == 0xFFu) continue;
          atomicAdd(line + l3 * (STILE * SROW), (unsigned long long)fixed_round(q12 * wz[j3]));
        }
      }
    }
  }
  __syncthreads();
  const size_t KK = (size_t)K1 * K2;
  for (int k = threadIdx.x; k < STILE * STILE * STILE; k += SBS) {
    const int l1 = k % STILE, l2 = (k / STILE) % STILE, l3 = k / (STILE * STILE);
    if (l1 < e1 && l2 < e2 && l3 < e3)
      grid[(size_t)(a3 + l3) * KK + (size_t)(a2 + l2) * K1 + (a1 + l1)] = acc[(l3 * STILE + l2) * SROW + l1];
  }
}

// Owner form: the tile spreads only its own particles (base node in the tile), each into a block
// of (STILE + P - 1)^3 nodes, the tile and the P - 1 nodes below it in every dimension (unwrapped:
// block index = node - origin + P - 1), so every stencil node lands in the block. The block of
// own particle t starts at its base offset o (node k0 + j of window node j at block index o + j).
// 1. three threads per particle: window weights per dimension and the base offset (a byte);
// 2. one (particle, x node j1) per thread, walking its P x P (y, z) nodes: q12 = ((q w_x) w_y) 2^S,
//    then 64-bit LDS atomics of q12 w_z rounded to the nearest integer (the products and order of
//    k_spread), every address the item's base plus an immediate offset.
constexpr int OCHUNK = 160;  // own particles spread together (P items each: 960 at P = 6)
template <int WP, int STILE, int SBS, int RCAP>
__device__ __forceinline__ void own_chunk(unsigned long long *blk, const float4 *ring, int h, int nc,
                                          float (*wsh)[3][8], uchar4 *ob, const int (&K)[3], const int (&A)[3],
                                          float scale, const WinCoef &wc) {
  constexpr int P = win_width(WP);
  constexpr int BV = STILE + P - 1;
  __syncthreads();  // the previous chunk's lines are done with the chunk arrays
  for (int e = threadIdx.x; e < 3 * nc; e += SBS) {
    const int t = e / 3, d = e - 3 * t;
    const float4 u = ring[(h + t) % RCAP];
    float w[P], dw[P];
    const int k = window<WP>(d == 0 ? u.x : (d == 1 ? u.y : u.z), w, dw, wc);
#pragma unroll
    for (int j = 0; j < P; ++j) wsh[t][d][j] = w[j];
    // k + P - 1 is the base node, which lies in the tile: o in [0, STILE)
    ((unsigned char *)&ob[t])[d] = (unsigned char)wrap_node(k + P - 1 - A[d], K[d]);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nc * P; e += SBS) {
    const int t = e / P, j1 = e - t * P;
    const uchar4 o = ob[t];
    const float qx = ring[(h + t) % RCAP].w * wsh[t][0][j1];
    float wy[P], wz[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
      wy[j] = wsh[t][1][j];
      wz[j] = wsh[t][2][j];
    }
    unsigned long long *base = blk + ((int)o.z * BV + (int)o.y) * BV + (int)o.x + j1;
#pragma unroll
    for (int j2 = 0; j2 < P; ++j2) {
      const float q12 = qx * wy[j2] * scale;
#pragma unroll
      for (int j3 = 0; j3 < P; ++j3)
        atomicAdd(base + j2 * BV + j3 * (BV * BV), (unsigned long long)fixed_round(q12 * wz[j3]));
    }
  }
}

// Each node of the fixed-point grid = the sum of the tile blocks' entries that stand for it (the
// owner form above). Per dimension: its own tile's (block index r + P - 1, r = node - origin), the
// next tile's when the node lies within P - 1 below that tile's origin, and tile 0's when it lies
// within P - 1 below K (the periodic wrap), so one to three entries; the node's sum runs over their
// combinations. Integer sums: exact, in any order.
// The candidates of one dimension: slot 0 its own tile's (always), slot 1 the next tile's, slot 2
// tile 0's wrapped halo, each with its validity.
template <int STILE>
__device__ __forceinline__ void fold_slots(int x, int K, int nt, int P, int (&t)[3], int (&b)[3], bool (&v)[3]) {
  const int t0 = x / STILE, t1 = (x + P - 1) / STILE;
  t[0] = t0;
  b[0] = x - t0 * STILE + P - 1;
  v[0] = true;
  t[1] = t1;
  b[1] = x - t1 * STILE + P - 1;
  v[1] = t1 != t0 && t1 < nt;
  t[2] = 0;
  b[2] = x - K + P - 1;
  v[2] = x + P - 1 >= K;
}

// Each node of the fixed-point grid = the sum of the tile blocks' entries that stand for it (the
// owner form above). Per dimension: its own tile's (block index r + P - 1, r = node - origin), the
// next tile's when the node lies within P - 1 below that tile's origin, and tile 0's when it lies
// within P - 1 below K (the periodic wrap), so one to three entries; the node's sum runs over their
// combinations. Integer sums: exact, in any order. One wave per 64 nodes of an x row (block
// 64 x 4: four rows of y), so the y and z candidates are the same across a wave.
template <int STILE>
__global__ __launch_bounds__(256) void k_spread_fold(const unsigned long long *__restrict__ blocks,
                                                     unsigned long long *__restrict__ grid, int K1, int K2, int K3,
                                                     int P, const int64_t *__restrict__ state) {
  if (state[PDX_DN_ERROR] != 0) return;  // the tiles wrote nothing: the grid keeps its last value
  const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
  if (x >= K1 || y >= K2) return;
  const int BV = STILE + P - 1, BVOL = BV * BV * BV;
  const int nt1 = (K1 + STILE - 1) / STILE, nt2 = (K2 + STILE - 1) / STILE, nt3 = (K3 + STILE - 1) / STILE;
  int tx[3], bx[3], ty[3], by[3], tz[3], bz[3];
  bool vx[3], vy[3], vz[3];
  fold_slots<STILE>(x, K1, nt1, P, tx, bx, vx);
  fold_slots<STILE>(y, K2, nt2, P, ty, by, vy);
  fold_slots<STILE>(z, K3, nt3, P, tz, bz, vz);
  unsigned long long s = 0;
#pragma unroll
  for (int i3 = 0; i3 < 3; ++i3) {
    if (!vz[i3]) continue;
#pragma unroll
    for (int i2 = 0; i2 < 3; ++i2) {
      if (!vy[i2]) continue;
      const int row = tz[i3] * nt2 + ty[i2];
      const int off = (bz[i3] * BV + by[i2]) * BV;
#pragma unroll
      for (int i1 = 0; i1 < 3; ++i1)
        if (vx[i1]) s += blocks[(size_t)(row * nt1 + tx[i1]) * BVOL + off + bx[i1]];
    }
  }
  grid[((size_t)z * K2 + y) * K1 + x] = s;
}

// Tile lists from the particles' current positions (owner form): the counters cleared and the
// overflow word (word 3 of the list's snapshot state) reset, then k_tile_bin.
__global__ void k_tile_prep(int32_t *__restrict__ tn, int tiles, int64_t *__restrict__ over_word) {
  const int k = blockIdx.x * TPB + threadIdx.x;
  if (k < tiles) tn[k] = 0;
  if (k == 0) over_word[3] = 0;
}

// Each particle into the list of the tile that holds its base node, the lists the owner-form spread
// and the tile gather read: (u1, u2, u3, q) and the index, its slot from the tile's counter, one
// atomic per tile and wave (the wave's lanes grouped by tile). The particles are taken in the
// list's cell order (`order`: 1-based, a permutation of all particles), so that a wave meets a few
// tiles whatever the storage order. A slot beyond tcap sets the overflow word: the tiles then only
// clear their blocks, the predicated atomic spread adds every particle and the gather walks them all.
template <int WP, int STILE>
__global__ __launch_bounds__(TPB) void k_tile_bin(int N, const float4 *__restrict__ packed,
                                                  const int32_t *__restrict__ order,
                                                  const int64_t *__restrict__ state, int K1, int K2, int K3,
                                                  float Lx, float Ly, float Lz, float4 *__restrict__ tlist,
                                                  int32_t *__restrict__ tids, int32_t *__restrict__ tn, int tcap,
                                                  int64_t *__restrict__ over_word) {
  if (state[PDX_DN_ERROR] != 0) return;  // no usable positions: the grid keeps its last value
  const int t = blockIdx.x * TPB + threadIdx.x;
  const bool valid = t < N;
  const int i = valid ? order[t] - 1 : 0;
  const float4 p = packed[i];
  const float u1 = grid_coordinate(p.x, Lx, K1), u2 = grid_coordinate(p.y, Ly, K2),
              u3 = grid_coordinate(p.z, Lz, K3);
  const int tx = (K1 + STILE - 1) / STILE, ty = (K2 + STILE - 1) / STILE;
  const int b = ((wrap_node(window_base<WP>(u3), K3) / STILE) * ty + wrap_node(window_base<WP>(u2), K2) / STILE) * tx +
                wrap_node(window_base<WP>(u1), K1) / STILE;
  const int lane = threadIdx.x & (WAVE - 1);
  unsigned long long active = __ballot(valid);
  int slot = 0;
  while (active) {  // the same in every lane of the wave
    const int leader = __ffsll((unsigned long long)active) - 1;
    const int bl = __shfl(b, leader, WAVE);
    const unsigned long long m = __ballot(valid && b == bl) & active;
    int base = 0;
    if (lane == leader) base = atomicAdd(&tn[bl], (int)__popcll(m));
    base = __shfl(base, leader, WAVE);
    if ((m >> lane) & 1ull) slot = base + (int)__popcll(m & ((1ull << lane) - 1ull));
    active &= ~m;
  }
  if (!valid) return;
  if (slot < tcap) {
    tlist[(size_t)b * tcap + slot] = make_float4(u1, u2, u3, p.w);
    tids[(size_t)b * tcap + slot] = i;
  } else {
    over_word[3] = 1;
  }
}

// Own particles per chunk of the list-fed owner kernel: as many as leave room for two workgroups
// per CU (81 920 B of LDS each) beside the block, at most 160; 160 where the block alone is too
// large for two (a 1024-thread workgroup then runs alone on its CU).
template <int BV> constexpr int own_list_room() { return (81920 - BV * BV * BV * 8 - 512) / 116; }
template <int BV> constexpr bool own_list_pair() { return own_list_room<BV>() >= 32; }
template <int BV> constexpr int own_list_chunk() {
  return own_list_pair<BV>() ? (own_list_room<BV>() < 160 ? own_list_room<BV>() : 160) : 160;
}

// The owner form's tile kernel fed by the tile lists of k_tile_bin: the block, a chunk's ring and
// its weights in LDS, nothing else, so that two 512-thread workgroups share a CU where the block
// leaves room (orders up to 6 at 16-node tiles); one's list loads and block write then overlap the
// other's atomics. A list that overflowed (word 3) leaves every block clear: the predicated atomic
// spread then adds every particle.
template <int WP, int STILE, int SBS>
__global__ __launch_bounds__(SBS) void k_spread_own_listed(const int64_t *__restrict__ state, int K1, int K2, int K3,
                                                           float scale, const float4 *__restrict__ tlist,
                                                           const int32_t *__restrict__ tn, int tcap,
                                                           const int64_t *__restrict__ over_word, const WinCoef wc,
                                                           unsigned long long *__restrict__ blocks) {
  if (state[PDX_DN_ERROR] != 0) return;  // no usable positions: the grid keeps its last value
  constexpr int P = win_width(WP);
  constexpr int BV = STILE + P - 1;
  constexpr int ACC = BV * BV * BV;
  constexpr int CH = own_list_chunk<BV>();
  __shared__ unsigned long long acc[ACC];
  __shared__ float4 ring[CH];
  __shared__ float wsh[CH][3][8];
  __shared__ uchar4 ob[CH];
  const int tx = (K1 + STILE - 1) / STILE, ty = (K2 + STILE - 1) / STILE;
  const int b = blockIdx.x;
  const int a1 = (b % tx) * STILE, a2 = ((b / tx) % ty) * STILE, a3 = (b / (tx * ty)) * STILE;
  const int Kv[3] = {K1, K2, K3}, Av[3] = {a1, a2, a3};
  for (int k = threadIdx.x; k < ACC; k += SBS) acc[k] = 0ull;
  const int n = over_word[3] != 0 ? 0 : tn[b];  // <= tcap when no list overflowed
  const float4 *own = tlist + (size_t)b * tcap;
  for (int c0 = 0; c0 < n; c0 += CH) {
    const int nc = min(CH, n - c0);
    __syncthreads();  // the previous chunk's lines are done with the ring
    for (int t = threadIdx.x; t < nc; t += SBS) ring[t] = own[c0 + t];
    own_chunk<WP, STILE, SBS, CH>(acc, ring, 0, nc, wsh, ob, Kv, Av, scale, wc);
  }
  __syncthreads();
  unsigned long long *out = blocks + (size_t)b * ACC;  // the whole block, every entry written
  for (int k = threadIdx.x; k < ACC; k += SBS) out[k] = acc[k];
}

// The owner form's tile kernel over the list's cells (PDX_SPREAD_LISTS=0): the block, the ring and
// the cell scan's arrays take most of the LDS, so one workgroup per CU.
template <int WP, int STILE, int SBS>
__global__ __launch_bounds__(SBS) void k_spread_own(const float4 *__restrict__ packed,
                                                      const int32_t *__restrict__ offsets,
                                                      const int32_t *__restrict__ ids,
                                                      const int64_t *__restrict__ state, int nx, int ny,
                                                      int nz, float margin, unsigned long long *grid, int K1,
                                                      int K2, int K3, float Lx, float Ly, float Lz, float scale,
                                                      float4 *__restrict__ tlist, int32_t *__restrict__ tids,
                                                      int32_t *__restrict__ tn, int tcap,
                                                      const float *__restrict__ disp, int ndisp, float stale2,
                                                      int64_t *stale_word, const WinCoef wc,
                                                      unsigned long long *__restrict__ blocks) {
  if (state[PDX_DN_ERROR] != 0) return;  // no usable cells / positions: the grid keeps its last value
  constexpr int P = win_width(WP);
  // own particles wait in an LDS ring of (u1, u2, u3, q): one pass of the cell scan adds at most
  // CB * SBS of them, and OCHUNK of them are spread as soon as that many wait, so the ring never
  // holds more than OCHUNK - 1 + CB * SBS: every own particle is staged, at any density
  constexpr int CB = 2;
  constexpr int CH = OCHUNK;
  constexpr int RCAP = CH + CB * SBS;
  constexpr int BV = STILE + P - 1;
  constexpr int ACC = BV * BV * BV;  // the block
  __shared__ unsigned long long acc[ACC];
  __shared__ float4 ring[RCAP];
  __shared__ float wsh[CH][3][8];
  __shared__ uchar4 ob[OCHUNK];  // base offsets
  __shared__ int cstart[SBS], cpre[SBS + 1], wsum[SBS / 64];
  __shared__ int tail;  // particles staged so far; particle s waits in ring slot s % RCAP
  const int tx = (K1 + STILE - 1) / STILE, ty = (K2 + STILE - 1) / STILE;
  const int b = blockIdx.x;
  const int a1 = (b % tx) * STILE, a2 = ((b / tx) % ty) * STILE, a3 = (b / (tx * ty)) * STILE;
  const int e1 = min(STILE, K1 - a1), e2 = min(STILE, K2 - a2), e3 = min(STILE, K3 - a3);
  for (int k = threadIdx.x; k < ACC; k += SBS) acc[k] = 0ull;
  // snapshot cells (disp != null): the check's per-block maxima of |x - x_build|^2 bound the
  // motion since the snapshot's build. A maximum beyond the margin (or -1: not measured) makes
  // the cells unusable: every tile then only clears its nodes, and the fallback spread and the
  // per-particle gather take over for this evaluation (word 3 of the snapshot state).
  bool stale = false;
  if (disp) {
    int s = 0;
    for (int k = threadIdx.x; k < ndisp; k += SBS) {
      const float d = disp[k];
      s |= !(d >= 0.0f && d <= stale2);
    }
    stale = __syncthreads_or(s) != 0;
    if (b == 0 && threadIdx.x == 0) stale_word[3] = stale ? 1 : 0;
  }
  // candidate cells per dimension, in the list's frame (x + L/2) / L: a particle at u is own when
  // the base node of u + shift lies in [A, A + E - 1]
  const int K[3] = {K1, K2, K3}, A[3] = {a1, a2, a3}, E[3] = {e1, e2, e3}, NC[3] = {nx, ny, nz};
  const float Ls[3] = {Lx, Ly, Lz};
  int c0[3], cn[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float f0 = ((float)A[d] - win_shift<WP>()) / (float)K[d] + 0.5f - margin / Ls[d] - 1e-5f;
    const float f1 = ((float)(A[d] + E[d]) - win_shift<WP>()) / (float)K[d] + 0.5f + margin / Ls[d] + 1e-5f;
    const int lo = (int)floorf(f0 * NC[d]);
    c0[d] = lo;
    cn[d] = min((int)floorf(f1 * NC[d]) - lo + 1, NC[d]);
  }
  const int ncand = cn[0] * cn[1] * cn[2];
  // (the cell range is taken in rounds of SBS cells: one round whenever the tile's candidate cells
  // number at most the workgroup's threads)
  __shared__ int nown;
  if (threadIdx.x == 0) {
    tail = 0;
    nown = 0;
  }
  float4 *own = tlist + (size_t)b * tcap;
  int32_t *oid = tids + (size_t)b * tcap;
  // one candidate particle: own when its base node lies in the tile; then listed for the gather,
  // and staged for the spread when its charge is nonzero
  auto take = [&](int i, const float4 p) {
    const float u1 = grid_coordinate(p.x, Lx, K1), u2 = grid_coordinate(p.y, Ly, K2),
                u3 = grid_coordinate(p.z, Lz, K3);
    const int o1 = wrap_node(window_base<WP>(u1) - a1, K1);
    const int o2 = wrap_node(window_base<WP>(u2) - a2, K2);
    const int o3 = wrap_node(window_base<WP>(u3) - a3, K3);
    if (!(o1 < e1 && o2 < e2 && o3 < e3)) return;
    const int k = atomicAdd(&nown, 1);
    if (k < tcap) {
      own[k] = make_float4(u1, u2, u3, p.w);
      oid[k] = i;
    }
    if (p.w != 0.0f) ring[atomicAdd(&tail, 1) % RCAP] = make_float4(u1, u2, u3, p.w);
  };
  int head = 0;  // particles spread so far (the same in every thread)
  for (int round = 0; !stale && round * SBS < ncand; ++round) {
    // the round's cells, one per thread, and the block-wide inclusive scan of their sizes
    const int nc = min(SBS, ncand - round * SBS);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int cnt = 0;
    if (threadIdx.x < nc) {
      const int cc = round * SBS + threadIdx.x;
      int cx = c0[0] + cc % cn[0], cy = c0[1] + (cc / cn[0]) % cn[1], cz = c0[2] + cc / (cn[0] * cn[1]);
      cx = ((cx % nx) + nx) % nx;
      cy = ((cy % ny) + ny) % ny;
      cz = ((cz % nz) + nz) % nz;
      const int c = (cz * ny + cy) * nx + cx;
      cstart[threadIdx.x] = offsets[c] - 1;
      cnt = offsets[c + 1] - offsets[c];
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(cnt, o, 64);
      if (lane >= o) cnt += u;
    }
    if (lane == 63) wsum[w] = cnt;
    __syncthreads();
    if (w == 0) {  // the waves' totals, scanned
      int t = lane < SBS / 64 ? wsum[lane] : 0;
#pragma unroll
      for (int o = 1; o < SBS / 64; o <<= 1) {
        const int u = __shfl_up(t, o, 64);
        if (lane >= o) t += u;
      }
      if (lane < SBS / 64) wsum[lane] = t;
    }
    __syncthreads();
    cnt += w > 0 ? wsum[w - 1] : 0;
    cpre[threadIdx.x + 1] = cnt;
    if (threadIdx.x == 0) cpre[0] = 0;
    __syncthreads();
    const int total = cpre[nc];
    // flat candidate index f -> its cell c (the last c < nc with cpre[c] <= f, by halvings) ->
    // particle; CB candidates per thread and pass, their loads issued together (two memory round
    // trips per pass: ids, then positions)
    auto cell_of = [&](int f) {
      int c = 0;
#pragma unroll
      for (int hh = SBS / 2; hh >= 1; hh >>= 1)
        if (c + hh < nc && cpre[c + hh] <= f) c += hh;
      return cstart[c] + (f - cpre[c]);
    };
    for (int p0 = 0; p0 < total; p0 += CB * SBS) {
      int ic[CB];
      float4 pc[CB];
#pragma unroll
      for (int u = 0; u < CB; ++u) {
        const int f = p0 + threadIdx.x + u * SBS;
        ic[u] = f < total ? ids[cell_of(f)] - 1 : -1;
      }
#pragma unroll
      for (int u = 0; u < CB; ++u) pc[u] = packed[ic[u] >= 0 ? ic[u] : (ic[0] >= 0 ? ic[0] : 0)];
#pragma unroll
      for (int u = 0; u < CB; ++u)
        if (ic[u] >= 0) take(ic[u], pc[u]);
      __syncthreads();
      const int staged = tail;
      for (; staged - head >= CH; head += CH)
        own_chunk<WP, STILE, SBS, RCAP>(acc, ring, head, CH, wsh, ob, K, A, scale, wc);
      // every thread has read tail, and the chunks' lines are done with their ring slots, before
      // the next pass stages more
      __syncthreads();
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) tn[b] = stale ? 0 : nown;
  const int staged = stale ? 0 : tail;
  if (staged > head) own_chunk<WP, STILE, SBS, RCAP>(acc, ring, head, staged - head, wsh, ob, K, A, scale, wc);
  __syncthreads();
  unsigned long long *out = blocks + (size_t)b * ACC;  // the whole block, every entry written
  for (int k = threadIdx.x; k < ACC; k += SBS) out[k] = acc[k];
}

// Workgroups are dispatched to the 8 XCDs round robin (block b runs on XCD b mod 8), each XCD
// with its own L2. With `xcd` set, the blocks that share an XCD take a contiguous range of the
// logical block order, so one XCD's particles (consecutive in storage order, which for lattice-
// ordered or cell-sorted particles is a spatial slab) touch one slab of the grid.
__device__ __forceinline__ int xcd_block(int xcd) {
  const int b = blockIdx.x, nb = gridDim.x;
  if (!xcd) return b;
  const int x = b & 7, r = b >> 3, q = nb >> 3, rem = nb & 7;
  return (x < rem ? x * (q + 1) : rem * (q + 1) + (x - rem) * q) + r;
}

// ---- gather: 8 lanes per particle, lane = (y, z) lines of the stencil walked along x; line
// partial sums reduced over the particle's lanes.
template <int WP, bool ENERGY>
__global__ __launch_bounds__(TPB) void k_gather(int N, const float *__restrict__ x, const float *__restrict__ y,
                                                const float *__restrict__ z, const float *__restrict__ q,
                                                const int32_t *__restrict__ order, const float *__restrict__ pot,
                                                int K1, int K2, int K3, float Lx, float Ly, float Lz,
                                                float *fx, float *fy, float *fz, float *Epot, float bg, int set,
                                                float *Vglobal, const double *acc, double ubg, int xcd,
                                                const WinCoef wc) {
  constexpr int P = win_width(WP);
  constexpr int LPP = gather_lanes(P);
  const int tid = xcd_block(xcd) * TPB + threadIdx.x;
  if (Vglobal && tid < 6) Vglobal[tid] = (float)(acc[1 + tid] + (tid < 3 ? ubg : 0.0));
  const int t = tid / LPP;
  const int lane = tid % LPP;
  const bool active = t < N;
  const int i = active ? (order ? order[t] - 1 : t) : 0;
  const float qi = active ? q[i] : 0.0f;
  float phi = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
  if (active && qi != 0.0f) {
    float wx[P], wy[P], wz[P], dx[P], dy[P], dz[P];
    const int k1 = window<WP>(grid_coordinate(x[i], Lx, K1), wx, dx, wc);
    const int k2 = window<WP>(grid_coordinate(y[i], Ly, K2), wy, dy, wc);
    const int k3 = window<WP>(grid_coordinate(z[i], Lz, K3), wz, dz, wc);
    // lane j < P = x node k1 + j of every (y, z) line: one load instruction reads, per particle,
    // the P consecutive nodes of one line (one cache line) instead of P^2 scattered ones; the
    // P^2 lines are unrolled, so their loads are independent and the y / z weights are registers
    const size_t KK = (size_t)K1 * K2;
    const int jx = lane < P ? lane : 0;
    const int n1 = wrap_node(k1 + jx, K1);
    const float ax = lane < P ? sel<P>(wx, jx) : 0.0f;
    const float bx = lane < P ? sel<P>(dx, jx) : 0.0f;
    float c[P * P];
#pragma unroll
    for (int j3 = 0; j3 < P; ++j3) {
      const size_t b3 = KK * wrap_node(k3 + j3, K3) + n1;
#pragma unroll
      for (int j2 = 0; j2 < P; ++j2) c[j3 * P + j2] = pot[b3 + (size_t)K1 * wrap_node(k2 + j2, K2)];
    }
    // per z node the y sums A = sum wy v, B = sum dy v; phi = sum wz A, gy = sum wz B,
    // gz = sum dz A (two multiply-adds per node instead of six)
#pragma unroll
    for (int j3 = 0; j3 < P; ++j3) {
      float A = 0.f, B = 0.f;
#pragma unroll
      for (int j2 = 0; j2 < P; ++j2) {
        A = fmaf(wy[j2], c[j3 * P + j2], A);
        B = fmaf(dy[j2], c[j3 * P + j2], B);
      }
      phi = fmaf(wz[j3], A, phi);
      gy = fmaf(wz[j3], B, gy);
      gz = fmaf(dz[j3], A, gz);
    }
    // the lane's x weight (or its derivative, for gx) applies to its whole column of lines
    gx = bx * phi;
    phi *= ax;
    gy *= ax;
    gz *= ax;
  }
  // sums over the particle's 8 lanes: three DPP adds each (no LDS round trips)
  gx = sum8_lanes(gx);
  gy = sum8_lanes(gy);
  gz = sum8_lanes(gz);
  if (ENERGY) phi = sum8_lanes(phi);
  if (!active || lane != 0) return;
  const float ffx = -(qi * gx * ((float)K1 / Lx));
  const float ffy = -(qi * gy * ((float)K2 / Ly));
  const float ffz = -(qi * gz * ((float)K3 / Lz));
  if (set) {
    fx[i] = ffx; fy[i] = ffy; fz[i] = ffz;
    if (ENERGY && Epot) Epot[i] = 0.5f * qi * phi + bg * qi;
  } else if (qi != 0.0f) {
    fx[i] += ffx; fy[i] += ffy; fz[i] += ffz;
    if (ENERGY && Epot) Epot[i] += 0.5f * qi * phi + bg * qi;
  }
}

// ---- tile gather: the forces at the positions of the last owner-computes spread, from the
// particle lists it wrote (the tile kernels' own lists, or k_tile_bin's before the list-fed owner
// form). One workgroup per tile, in three phases with one
// memory round trip each: (1) the potential of the nodes the tile's particles reach (the tile
// and the P-1 nodes below it in each dimension, (STILE+P-1)^3 values) and the tile's list are
// loaded, every load issued before the first LDS store; (2) each listed particle is gathered
// from LDS with the lane layout of k_gather (lane = x node, (y, z) lines unrolled, DPP sums)
// into per-particle results in LDS; (3) the results are written (or added) to the particles'
// outputs. Tiles whose list overflowed (tn > tcap) are left to the workgroups past the tiles:
// they walk all particles and gather those of overflowed tiles from global memory, one
// particle per thread. Per particle the result is k_gather's up to FP32 rounding.
constexpr int GCHUNK = 128;  // listed particles per LDS chunk (a tile holds ~N/tiles)
constexpr int GTB = 512;     // threads per tile workgroup (measured faster than 256 and 128)
template <int WP, int STILE, bool ENERGY>
__global__ __launch_bounds__(GTB) void k_gather_tiles(int N, const float4 *__restrict__ packed,
                                                      const float4 *__restrict__ tlist,
                                                      const int32_t *__restrict__ tids,
                                                      const int32_t *__restrict__ tn, int tcap, int tiles,
                                                      const float *__restrict__ pot, int K1, int K2, int K3,
                                                      float Lx, float Ly, float Lz, float *fx, float *fy,
                                                      float *fz, float *Epot, float bg, int set, float *Vglobal,
                                                      const double *acc, double ubg,
                                                      const int64_t *__restrict__ stale_word, const WinCoef wc) {
  constexpr int P = win_width(WP);
  constexpr int R = STILE + P - 1;
  __shared__ float reg[R * R * R];
  __shared__ float4 pl[GCHUNK];
  __shared__ int pid[GCHUNK];
  __shared__ float res[4][GCHUNK];
  __shared__ float gw[GCHUNK][3][P], gd[GCHUNK][3][P];  // weights and derivatives per dimension
  __shared__ int go[GCHUNK][3];                         // tile-local base nodes
  if (Vglobal && blockIdx.x == 0 && threadIdx.x < 6)
    Vglobal[threadIdx.x] = (float)(acc[1 + threadIdx.x] + (threadIdx.x < 3 ? ubg : 0.0));
  const int tx = (K1 + STILE - 1) / STILE, ty = (K2 + STILE - 1) / STILE;
  const float s1 = (float)K1 / Lx, s2 = (float)K2 / Ly, s3 = (float)K3 / Lz;
  // a stale snapshot spread wrote no lists: the walk past the tiles then gathers every particle
  const bool stale = stale_word && stale_word[3] != 0;
  if ((int)blockIdx.x >= tiles) {
    const int stride = (gridDim.x - tiles) * GTB;
    const size_t KK = (size_t)K1 * K2;
    for (int i = (blockIdx.x - tiles) * GTB + threadIdx.x; i < N; i += stride) {
      const float4 p = packed[i];
      const float u1 = grid_coordinate(p.x, Lx, K1), u2 = grid_coordinate(p.y, Ly, K2),
                  u3 = grid_coordinate(p.z, Lz, K3);
      const int t = wrap_node(window_base<WP>(u1), K1) / STILE +
                    tx * (wrap_node(window_base<WP>(u2), K2) / STILE + ty * (wrap_node(window_base<WP>(u3), K3) / STILE));
      if (!stale && tn[t] <= tcap) continue;
      float wx[P], wy[P], wz[P], dx[P], dy[P], dz[P];
      const int k1 = window<WP>(u1, wx, dx, wc), k2 = window<WP>(u2, wy, dy, wc), k3 = window<WP>(u3, wz, dz, wc);
      float phi = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
      for (int j3 = 0; j3 < P; ++j3)
#pragma unroll
        for (int j2 = 0; j2 < P; ++j2) {
          const size_t b = KK * wrap_node(k3 + j3, K3) + (size_t)K1 * wrap_node(k2 + j2, K2);
          float lp = 0.f, lg = 0.f;
#pragma unroll
          for (int j1 = 0; j1 < P; ++j1) {
            const float v = pot[b + wrap_node(k1 + j1, K1)];
            lp = fmaf(wx[j1], v, lp);
            lg = fmaf(dx[j1], v, lg);
          }
          phi = fmaf(wy[j2] * wz[j3], lp, phi);
          gx = fmaf(wy[j2] * wz[j3], lg, gx);
          gy = fmaf(dy[j2] * wz[j3], lp, gy);
          gz = fmaf(wy[j2] * dz[j3], lp, gz);
        }
      const float qi = p.w;
      if (set) {
        fx[i] = -(qi * gx * s1); fy[i] = -(qi * gy * s2); fz[i] = -(qi * gz * s3);
        if (ENERGY && Epot) Epot[i] = 0.5f * qi * phi + bg * qi;
      } else if (qi != 0.0f) {
        fx[i] -= qi * gx * s1; fy[i] -= qi * gy * s2; fz[i] -= qi * gz * s3;
        if (ENERGY && Epot) Epot[i] += 0.5f * qi * phi + bg * qi;
      }
    }
    return;
  }
  if (stale) return;
  const int b = blockIdx.x;
  const int n = tn[b];
  if (n <= 0 || n > tcap) return;
  const int a1 = (b % tx) * STILE, a2 = ((b / tx) % ty) * STILE, a3 = (b / (tx * ty)) * STILE;
  const int r1 = min(STILE, K1 - a1) + P - 1, r2 = min(STILE, K2 - a2) + P - 1, r3 = min(STILE, K3 - a3) + P - 1;
  const float4 *own = tlist + (size_t)b * tcap;
  const int32_t *oid = tids + (size_t)b * tcap;
  // phase 1: region node l (per dimension) = grid node a - (P-1) + l, wrapped below zero; the
  // thread's region index k advances by GTB per load, its (l1, l2, l3) digits incrementally
  constexpr int NLD = (R * R * R + GTB - 1) / GTB;
  constexpr int D1 = GTB % R, D2 = (GTB / R) % R, D3 = GTB / (R * R);
  float v[NLD];
  {
    int l1 = threadIdx.x % R, l2 = (threadIdx.x / R) % R, l3 = threadIdx.x / (R * R);
#pragma unroll
    for (int j = 0; j < NLD; ++j) {
      v[j] = 0.f;
      if (l3 < r3 && l1 < r1 && l2 < r2) {
        int g1 = a1 - (P - 1) + l1; g1 += g1 < 0 ? K1 : 0;
        int g2 = a2 - (P - 1) + l2; g2 += g2 < 0 ? K2 : 0;
        int g3 = a3 - (P - 1) + l3; g3 += g3 < 0 ? K3 : 0;
        v[j] = pot[((size_t)g3 * K2 + g2) * K1 + g1];
      }
      l1 += D1;
      if (l1 >= R) { l1 -= R; ++l2; }
      l2 += D2;
      if (l2 >= R) { l2 -= R; ++l3; }
      l3 += D3;
    }
  }
  static_assert(GCHUNK <= GTB, "at most one list entry per thread and chunk");
  float4 e0 = make_float4(0.f, 0.f, 0.f, 0.f);
  int i0 = -1;
  if (threadIdx.x < min(n, GCHUNK)) {
    e0 = own[threadIdx.x];
    i0 = oid[threadIdx.x];
  }
#pragma unroll
  for (int j = 0; j < NLD; ++j) {
    const int k = threadIdx.x + j * GTB;
    if (k < R * R * R) reg[k] = v[j];
  }
  constexpr int LPP = gather_lanes(P);
  const int lane = threadIdx.x % LPP;
  for (int c0 = 0; c0 < n; c0 += GCHUNK) {
    const int nc = min(GCHUNK, n - c0);
    if (c0 > 0) {
      __syncthreads();  // the previous chunk's results are written out
      e0 = make_float4(0.f, 0.f, 0.f, 0.f);
      i0 = -1;
      if (threadIdx.x < nc) {
        e0 = own[c0 + threadIdx.x];
        i0 = oid[c0 + threadIdx.x];
      }
    }
    if (threadIdx.x < GCHUNK) {
      pl[threadIdx.x] = e0;
      pid[threadIdx.x] = i0;
    }
    __syncthreads();
    // weights: one thread per particle and dimension; the base node k + P - 1 (wrapped) lies in
    // the tile: local offsets 0 .. STILE-1
    for (int e = threadIdx.x; e < 3 * nc; e += GTB) {
      const int t = e / 3, d = e - 3 * t;
      const float4 u = pl[t];
      float w[P], dw[P];
      const int a = d == 0 ? a1 : (d == 1 ? a2 : a3);
      const int K = d == 0 ? K1 : (d == 1 ? K2 : K3);
      const int k = window<WP>(d == 0 ? u.x : (d == 1 ? u.y : u.z), w, dw, wc);
      go[t][d] = min(max(wrap_node(k + (P - 1) - a, K), 0), STILE - 1);
#pragma unroll
      for (int j = 0; j < P; ++j) {
        gw[t][d][j] = w[j];
        gd[t][d][j] = dw[j];
      }
    }
    __syncthreads();
    // phase 2 (t is uniform over a particle's LPP lanes, so the lane sums see only active lanes):
    // per z node, the y sums of the lane's x column, A = sum wy v and B = sum dy v; then
    // phi = sum wz A, gy = sum wz B, gz = sum dz A
    for (int t = threadIdx.x / LPP; t < nc; t += GTB / LPP) {
      const float4 e = pl[t];
      const int i = pid[t];
      const float qi = (i >= 0 && i < N) ? e.w : 0.0f;
      float phi = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
      if (qi != 0.0f) {
        float wy[P], wz[P], dy[P], dz[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
          wy[j] = gw[t][1][j]; dy[j] = gd[t][1][j];
          wz[j] = gw[t][2][j]; dz[j] = gd[t][2][j];
        }
        const int jx = lane < P ? lane : 0;
        const float ax = lane < P ? gw[t][0][jx] : 0.0f;
        const float bx = lane < P ? gd[t][0][jx] : 0.0f;
        const float *base = reg + (go[t][2] * R + go[t][1]) * R + go[t][0] + jx;
#pragma unroll
        for (int j3 = 0; j3 < P; ++j3) {
          float A = 0.f, B = 0.f;
#pragma unroll
          for (int j2 = 0; j2 < P; ++j2) {
            const float val = base[(j3 * R + j2) * R];
            A = fmaf(wy[j2], val, A);
            B = fmaf(dy[j2], val, B);
          }
          phi = fmaf(wz[j3], A, phi);
          gy = fmaf(wz[j3], B, gy);
          gz = fmaf(dz[j3], A, gz);
        }
        gx = bx * phi;
        phi *= ax;
        gy *= ax;
        gz *= ax;
      }
      gx = sum8_lanes(gx);
      gy = sum8_lanes(gy);
      gz = sum8_lanes(gz);
      if (ENERGY) phi = sum8_lanes(phi);
      if (lane == 0) {
        res[0][t] = -(qi * gx * s1);
        res[1][t] = -(qi * gy * s2);
        res[2][t] = -(qi * gz * s3);
        res[3][t] = 0.5f * qi * phi + bg * qi;
      }
    }
    __syncthreads();
    // phase 3: one listed particle per thread
    const int i = threadIdx.x < GCHUNK ? pid[threadIdx.x] : -1;
    if (threadIdx.x < nc && i >= 0 && i < N) {
      const float qi = pl[threadIdx.x].w;
      if (set) {
        fx[i] = res[0][threadIdx.x]; fy[i] = res[1][threadIdx.x]; fz[i] = res[2][threadIdx.x];
        if (ENERGY && Epot) Epot[i] = res[3][threadIdx.x];
      } else if (qi != 0.0f) {
        fx[i] += res[0][threadIdx.x]; fy[i] += res[1][threadIdx.x]; fz[i] += res[2][threadIdx.x];
        if (ENERGY && Epot) Epot[i] += res[3][threadIdx.x];
      }
    }
  }
}

// ---- ik gather (_gather_ik_kernel!, src/electrostatics/MeshKernels.jl:108-147): F_i = q_i
// sum_n W(u_i - n) E(n) from the three field grids; with ENERGY, E_i += 1/2 q_i phi_i + bg q_i.
// Same lane layout as k_gather.
template <int P, bool ENERGY>
__global__ __launch_bounds__(TPB) void k_gather_ik(int N, const float *__restrict__ x, const float *__restrict__ y,
                                                   const float *__restrict__ z, const float *__restrict__ q,
                                                   const int32_t *__restrict__ order, const float *__restrict__ ex,
                                                   const float *__restrict__ ey, const float *__restrict__ ez,
                                                   const float *__restrict__ pot, int K1, int K2, int K3, float Lx,
                                                   float Ly, float Lz, float *fx, float *fy, float *fz, float *Epot,
                                                   float bg, int set, float *Vglobal, const double *acc,
                                                   double ubg) {
  constexpr int LPP = gather_lanes(P);
  const int tid = blockIdx.x * TPB + threadIdx.x;
  if (Vglobal && tid < 6) Vglobal[tid] = (float)(acc[1 + tid] + (tid < 3 ? ubg : 0.0));
  const int t = tid / LPP;
  const int lane = tid % LPP;
  const bool active = t < N;
  const int i = active ? (order ? order[t] - 1 : t) : 0;
  const float qi = active ? q[i] : 0.0f;
  float phi = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
  if (active && qi != 0.0f) {
    float wx[P], wy[P], wz[P], d[P];
    const int k1 = bspline<P>(grid_coordinate(x[i], Lx, K1), wx, d);
    const int k2 = bspline<P>(grid_coordinate(y[i], Ly, K2), wy, d);
    const int k3 = bspline<P>(grid_coordinate(z[i], Lz, K3), wz, d);
    const size_t KK = (size_t)K1 * K2;
    for (int l = lane; l < P * P; l += LPP) {
      const int j1 = l % P, j2 = l / P;
      const int n1 = wrap_node(k1 + j1, K1), n2 = wrap_node(k2 + j2, K2);
      const size_t b = (size_t)n1 + (size_t)K1 * n2;
      float ax = 0.f, ay = 0.f, az = 0.f, a0 = 0.f;
#pragma unroll
      for (int j3 = 0; j3 < P; ++j3) {
        const size_t o = b + KK * wrap_node(k3 + j3, K3);
        ax = fmaf(wz[j3], ex[o], ax);
        ay = fmaf(wz[j3], ey[o], ay);
        az = fmaf(wz[j3], ez[o], az);
        if (ENERGY) a0 = fmaf(wz[j3], pot[o], a0);
      }
      const float w12 = sel<P>(wx, j1) * sel<P>(wy, j2);
      sx = fmaf(w12, ax, sx);
      sy = fmaf(w12, ay, sy);
      sz = fmaf(w12, az, sz);
      if (ENERGY) phi = fmaf(w12, a0, phi);
    }
  }
  sx = wave_sum_width(sx, LPP);
  sy = wave_sum_width(sy, LPP);
  sz = wave_sum_width(sz, LPP);
  if (ENERGY) phi = wave_sum_width(phi, LPP);
  if (!active || lane != 0) return;
  const float ffx = qi * sx, ffy = qi * sy, ffz = qi * sz;
  if (set) {
    fx[i] = ffx; fy[i] = ffy; fz[i] = ffz;
    if (ENERGY && Epot) Epot[i] = 0.5f * qi * phi + bg * qi;
  } else if (qi != 0.0f) {
    fx[i] += ffx; fy[i] += ffy; fz[i] += ffz;
    if (ENERGY && Epot) Epot[i] += 0.5f * qi * phi + bg * qi;
  }
}

// ---- optimal influence functions (Hockney-Eastwood; Ballenegger, Cerda & Holm 2012) for
// the B-spline window with alias sums |n_d| <= nmax, in FP64 on the device, one thread per
// half-spectrum mode (_influence_alias_kernel!, src/electrostatics/Influence.jl:77-181):
//   AD: G = (C/V) sum U^2 g k_n^2 / (sum U^2 . sum U^2 k_n^2)
//   ik: G = (C/V) sum U^2 g (D.k_n) / (|D|^2 (sum U^2)^2), D = k with Nyquist components zeroed
// with g(k) = 4 pi exp(-k^2/(4 alpha^2)) / k^2 (Gaussian split) or 4 pi c^(k rc)/k^2 with
// c^(u) = int_0^1 chi(x) cos(u x) dx by the nq-node quadrature (xq, wchi) (prolate split,
// spectral_g, src/electrostatics/Splits.jl:220-240), and the virial coefficients
// Vc_ab = -dG/d eps_ab at fixed fractional coordinates (Nyquist shears zeroed). U tables:
// U_d[m (2 nmax + 1) + n + nmax], the window transform at the alias frequencies (B-spline sinc^P
// or the prolate window's phi^). With nmax = 0 and U = phi^ this is ESP's plain deconvolution
// G = g / phi^^2 and its virial G delta_ab + (C/V) D g'(k)/k k_a k_b.
__device__ __forceinline__ void spectral_g(double qq, double i4a2, const double *__restrict__ xq,
                                           const double *__restrict__ wchi, int nq, double rc, double &g,
                                           double &gpk) {
  if (nq == 0) {
    g = 4.0 * M_PI * exp(-qq * i4a2) / qq;
    gpk = g * (-2.0 / qq - 2.0 * i4a2);  // g'(k)/k
    return;
  }
  const double k = sqrt(qq), u = k * rc;
  double c = 0.0, dc = 0.0;
  for (int j = 0; j < nq; ++j) {
    double sn, cs;
    sincos(u * xq[j], &sn, &cs);
    c += wchi[j] * cs;
    dc -= wchi[j] * xq[j] * sn;
  }
  g = 4.0 * M_PI * c / qq;
  gpk = 4.0 * M_PI * (rc * dc / qq - 2.0 * c / (qq * k)) / k;
}

__global__ void k_influence_alias(float *G, float *Vc, size_t nmodes, const double *__restrict__ U1t,
                                  const double *__restrict__ U2t, const double *__restrict__ U3t, int K1, int K2,
                                  int K3, double L1, double L2, double L3, double CoverV, double alpha, int nmax,
                                  int ik, const double *__restrict__ xq, const double *__restrict__ wchi, int nq,
                                  double rcs) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nmodes) return;
  const int K1h = K1 / 2 + 1;
  const int m1 = (int)(idx % K1h);
  const size_t rest = idx / K1h;
  const int m2 = (int)(rest % K2), m3 = (int)(rest / K2);
  const int s1 = m1, s2 = m2 <= K2 / 2 ? m2 : m2 - K2, s3 = m3 <= K3 / 2 ? m3 : m3 - K3;
  const bool ny1 = (K1 % 2 == 0) && m1 == K1 / 2, ny2 = (K2 % 2 == 0) && m2 == K2 / 2,
             ny3 = (K3 % 2 == 0) && m3 == K3 / 2;
  const double tp = 2.0 * M_PI;
  const double kx = tp * s1 / L1, ky = tp * s2 / L2, kz = tp * s3 / L3;
  const double dx = ny1 ? 0.0 : kx, dy = ny2 ? 0.0 : ky, dz = ny3 ? 0.0 : kz;
  const double d2 = dx * dx + dy * dy + dz * dz;
  const bool zero = (s1 == 0) && (s2 == 0) && (s3 == 0);
  if (zero || (ik && d2 == 0.0)) {
    G[idx] = 0.f;
    for (int c = 0; c < 6; ++c) Vc[(size_t)c * nmodes + idx] = 0.f;
    return;
  }
  const int W = 2 * nmax + 1;
  const double i4a2 = 1.0 / (4.0 * alpha * alpha);
  double A = 0, B = 0, Cc = 0, a[6] = {0, 0, 0, 0, 0, 0}, c[6] = {0, 0, 0, 0, 0, 0};
  for (int n1 = -nmax; n1 <= nmax; ++n1) {
    const double U1 = U1t[(size_t)m1 * W + n1 + nmax];
    const double q1 = tp * (s1 + (double)K1 * n1) / L1;
    for (int n2 = -nmax; n2 <= nmax; ++n2) {
      const double U2 = U2t[(size_t)m2 * W + n2 + nmax];
      const double q2 = tp * (s2 + (double)K2 * n2) / L2;
      for (int n3 = -nmax; n3 <= nmax; ++n3) {
        const double U3 = U3t[(size_t)m3 * W + n3 + nmax];
        const double q3 = tp * (s3 + (double)K3 * n3) / L3;
        const double u = U1 * U2 * U3, U2tot = u * u;
        B += U2tot;
        const double qq = q1 * q1 + q2 * q2 + q3 * q3;
        if (qq == 0.0) continue;
        double g, gpk;
        spectral_g(qq, i4a2, xq, wchi, nq, rcs, g, gpk);
        if (ik) {
          const double dq = dx * q1 + dy * q2 + dz * q3;
          A += U2tot * g * dq;
          const double h = U2tot * gpk * dq, e = U2tot * g;
          a[0] += h * q1 * q1 + e * 2.0 * dx * q1;
          a[1] += h * q2 * q2 + e * 2.0 * dy * q2;
          a[2] += h * q3 * q3 + e * 2.0 * dz * q3;
          a[3] += h * q1 * q2 + e * (dx * q2 + q1 * dy);
          a[4] += h * q1 * q3 + e * (dx * q3 + q1 * dz);
          a[5] += h * q2 * q3 + e * (dy * q3 + q2 * dz);
        } else {
          A += U2tot * g * qq;
          Cc += U2tot * qq;
          const double h = U2tot * (gpk * qq + 2.0 * g), h2 = 2.0 * U2tot;
          a[0] += h * q1 * q1; a[1] += h * q2 * q2; a[2] += h * q3 * q3;
          a[3] += h * q1 * q2; a[4] += h * q1 * q3; a[5] += h * q2 * q3;
          c[0] += h2 * q1 * q1; c[1] += h2 * q2 * q2; c[2] += h2 * q3 * q3;
          c[3] += h2 * q1 * q2; c[4] += h2 * q1 * q3; c[5] += h2 * q2 * q3;
        }
      }
    }
  }
  double Gm, v[6];
  if (ik) {
    const double den = d2 * B * B;
    Gm = CoverV * A / den;
    const double f = CoverV / den, r = CoverV * A / (B * B) * 2.0 / (d2 * d2);
    const double dd[6] = {dx * dx, dy * dy, dz * dz, dx * dy, dx * dz, dy * dz};
    for (int k = 0; k < 6; ++k) v[k] = (k < 3 ? Gm : 0.0) + f * a[k] - r * dd[k];
  } else {
    Gm = CoverV * A / (B * Cc);
    const double f = CoverV / (B * Cc), r = CoverV * A / (B * Cc * Cc);
    for (int k = 0; k < 6; ++k) v[k] = (k < 3 ? Gm : 0.0) + f * a[k] - r * c[k];
  }
  if (ny1 || ny2) v[3] = 0.0;
  if (ny1 || ny3) v[4] = 0.0;
  if (ny2 || ny3) v[5] = 0.0;
  G[idx] = (float)Gm;
  for (int k = 0; k < 6; ++k) Vc[(size_t)k * nmodes + idx] = (float)v[k];
}

// Learned-correction inputs from the charge grid of the last spread (80 per particle):
// q_i Q(n) at the 64 nodes n = floor(u) - 3 + {0..3}^3 that an order-4 B-spline at u covers
// (x fastest), q_i (u - floor(u)) per dimension, q_i, q_i^2, then zeros up to the MFMA width.
constexpr int NFEAT = 80;
// One wave per particle: lane l reads support node (l & 3, (l >> 2) & 3, l >> 4), so a row of
// 80 features is written by contiguous lanes.
__global__ __launch_bounds__(TPB) void k_features(int N, const float *__restrict__ x, const float *__restrict__ y,
                                                  const float *__restrict__ z, const float *__restrict__ q,
                                                  const unsigned long long *__restrict__ grid, int K1, int K2,
                                                  int K3, float Lx, float Ly, float Lz, float inv_scale,
                                                  float *feats) {
  const int i = blockIdx.x * (TPB / WAVE) + threadIdx.x / WAVE;
  const int lane = threadIdx.x % WAVE;
  if (i >= N) return;
  const float qi = q[i];
  const float u1 = grid_coordinate(x[i], Lx, K1), u2 = grid_coordinate(y[i], Ly, K2),
              u3 = grid_coordinate(z[i], Lz, K3);
  const float f1 = floorf(u1), f2 = floorf(u2), f3 = floorf(u3);
  const int n1 = wrap_node((int)f1 - 3 + (lane & 3), K1);
  const int n2 = wrap_node((int)f2 - 3 + ((lane >> 2) & 3), K2);
  const int n3 = wrap_node((int)f3 - 3 + (lane >> 4), K3);
  float *f = feats + (size_t)i * NFEAT;
  f[lane] = qi * fixed_to_float(grid[(size_t)n3 * K1 * K2 + (size_t)n2 * K1 + n1], inv_scale);
  if (lane < NFEAT - 64) {
    float e = 0.f;
    e = lane == 0 ? qi * (u1 - f1) : e;
    e = lane == 1 ? qi * (u2 - f2) : e;
    e = lane == 2 ? qi * (u3 - f3) : e;
    e = lane == 3 ? qi : e;
    e = lane == 4 ? qi * qi : e;
    f[64 + lane] = e;
  }
}

__global__ void k_fixed_to_double(const unsigned long long *g, double *out, size_t n, double inv) {
  const size_t k = (size_t)blockIdx.x * TPB + threadIdx.x;
  if (k < n) out[k] = (double)(long long)g[k] * inv;
}

// ---------------------------------------------------------------- host setup
double host_erfc(double x) { return std::erfc(x); }

int fft_friendly(int n) {
  int m = std::max(n, 1);
  while (true) {
    int r = m;
    for (int p : {2, 3, 5, 7})
      while (r % p == 0) r /= p;
    if (r == 1) return m;
    ++m;
  }
}

// M_P at the integers 0..P-1 (FP64), then the SPME deconvolution 1/|sum_j M_P(j) e^{2 pi i m j/K}|^2
// with the odd-order Nyquist zero replaced by the neighbour average (window_deconvolution,
// src/electrostatics/Windows.jl:283-308).
std::vector<double> deconvolution(int P, int K) {
  std::vector<double> a = {0.0, 1.0};
  while ((int)a.size() < P) {
    const int n = (int)a.size();
    std::vector<double> b(n + 1);
    for (int j = 0; j <= n; ++j) {
      const double left = j < n ? a[j] : 0.0, right = j >= 1 ? a[j - 1] : 0.0;
      b[j] = ((0.0 + j) * left + (n + 1 - 0.0 - j) * right) / n;
    }
    a = b;
  }
  std::vector<double> D(K);
  for (int m = 0; m < K; ++m) {
    double re = 0, im = 0;
    for (int j = 0; j < P; ++j) {
      re += a[j] * std::cos(2 * M_PI * m * j / K);
      im += a[j] * std::sin(2 * M_PI * m * j / K);
    }
    const double den = re * re + im * im;
    D[m] = den > 1e-12 ? 1.0 / den : NAN;
  }
  for (int m = 0; m < K; ++m)
    if (std::isnan(D[m])) D[m] = 0.5 * (D[(m - 1 + K) % K] + D[(m + 1) % K]);
  return D;
}

// Lines per workgroup tile (a power of two, 4..32): overridable per pass for tuning through
// PDX_FFT_TX / PDX_FFT_TY / PDX_FFT_TZ; kept within a 60 KiB LDS budget for the twiddles and
// the two ping-pong buffers (the reductions' static LDS comes on top).
int lines_per_tile(int K, int def, const char *env) {
  int T = def;
  if (const char *v = std::getenv(env)) T = std::atoi(v);
  if (T != 4 && T != 8 && T != 16 && T != 32) T = def;
  while (T > 4 && ((size_t)2 * T * K + K + 2 * T) * sizeof(float2) > 61440) T >>= 1;
  return T;
}

// B-spline window transform with alias images: U[m][n] = sinc(pi (s(m) + K n) / K)^P
// (window_alias_table, src/electrostatics/Windows.jl:391-399).
std::vector<double> alias_table(int P, int K, int nmax) {
  const int W = 2 * nmax + 1;
  std::vector<double> U((size_t)K * W);
  for (int m = 0; m < K; ++m) {
    const int sm = m <= K / 2 ? m : m - K;
    for (int n = -nmax; n <= nmax; ++n) {
      const double x = M_PI * (sm + (double)K * n) / K;
      const double sn = std::fabs(x) < 1e-12 ? 1.0 : std::sin(x) / x;
      U[(size_t)m * W + n + nmax] = std::pow(sn, P);
    }
  }
  return U;
}

bool is_esp(const pdx_mesh_s *m) {
  return m->spec.influence == PDX_INFLUENCE_ESP || m->spec.influence == PDX_INFLUENCE_ESP_DECONVOLUTION;
}
int alias_order_of(const pdx_mesh_s *m) {
  return m->spec.influence == PDX_INFLUENCE_ESP_DECONVOLUTION ? 0 : m->spec.alias_order;
}

// Prolate window transform at the alias frequencies, phi^(2 pi (s(m) + K n)/K) (window_alias_table,
// src/electrostatics/Windows.jl:401-411; with nmax = 0 the deconvolution's phi^, Windows.jl:340-352).
std::vector<double> prolate_alias_table(const ProlateWindow &w, int K, int nmax) {
  const int W = 2 * nmax + 1;
  std::vector<double> U((size_t)K * W);
  const int nq = nmax == 0 ? 192 : 256;
  for (int m = 0; m < K; ++m) {
    const int sm = m <= K / 2 ? m : m - K;
    for (int n = -nmax; n <= nmax; ++n) U[(size_t)m * W + n + nmax] = w.transform(2 * M_PI * (sm + (double)K * n) / K, nq);
  }
  return U;
}

int compute_influence(pdx_mesh_t m) {
  const int K1 = m->K1, K2 = m->K2, K3 = m->K3, K1h = m->K1h;
  const double L1 = m->box[0], L2 = m->box[1], L3 = m->box[2];
  const double CoverV = m->C / (L1 * L2 * L3);
  if (m->spec.influence != PDX_INFLUENCE_SPME) {
    const int nmax = alias_order_of(m), W = 2 * nmax + 1;
    std::vector<double> U;
    ProlateWindow pw;
    if (is_esp(m) && !pw.make(m->P, m->window_shape)) return fail(PDX_ERR_ARGUMENT, "prolate window set-up failed");
    for (int K : {K1, K2, K3}) {
      const std::vector<double> u = is_esp(m) ? prolate_alias_table(pw, K, nmax) : alias_table(m->P, K, nmax);
      U.insert(U.end(), u.begin(), u.end());
    }
    PDX_HIP(hipMemcpy(m->Utab.p, U.data(), sizeof(double) * U.size(), hipMemcpyHostToDevice));
    const size_t nmodes = (size_t)K1h * K2 * K3;
    hipLaunchKernelGGL(k_influence_alias, dim3((unsigned)((nmodes + 127) / 128)), dim3(128), 0, 0, m->G.p, m->Vc.p,
                       nmodes, m->Utab.p, m->Utab.p + (size_t)K1 * W, m->Utab.p + (size_t)(K1 + K2) * W, K1, K2,
                       K3, L1, L2, L3, CoverV, m->alpha, nmax, m->spec.influence == PDX_INFLUENCE_P3M_IK ? 1 : 0,
                       is_esp(m) ? (const double *)m->espq.p : nullptr,
                       is_esp(m) ? (const double *)m->espq.p + 128 : nullptr, is_esp(m) ? 128 : 0, m->spec.real_cutoff);
    PDX_LAUNCH_CHECK();
    PDX_HIP(hipDeviceSynchronize());
    return PDX_OK;
  }
  const std::vector<double> D1 = deconvolution(m->P, K1), D2 = deconvolution(m->P, K2), D3 = deconvolution(m->P, K3);
  std::vector<float> G((size_t)K1h * K2 * K3);
  const double a2 = m->alpha * m->alpha;
  for (int m3 = 0; m3 < K3; ++m3)
    for (int m2 = 0; m2 < K2; ++m2)
      for (int m1 = 0; m1 < K1h; ++m1) {
        const size_t idx = (size_t)m1 + (size_t)K1h * (m2 + (size_t)K2 * m3);
        const int s2 = m2 <= K2 / 2 ? m2 : m2 - K2, s3 = m3 <= K3 / 2 ? m3 : m3 - K3;
        if (m1 == 0 && s2 == 0 && s3 == 0) { G[idx] = 0.f; continue; }
        const double kx = 2 * M_PI * m1 / L1, ky = 2 * M_PI * s2 / L2, kz = 2 * M_PI * s3 / L3;
        const double k2 = kx * kx + ky * ky + kz * kz;
        const double g = 4.0 * M_PI * std::exp(-k2 / (4.0 * a2)) / k2;  // spectral_g (Splits.jl:220-224)
        G[idx] = (float)(CoverV * g * D1[m1] * D2[m2] * D3[m3]);
      }
  PDX_HIP(hipMemcpy(m->G.p, G.data(), sizeof(float) * G.size(), hipMemcpyHostToDevice));
  return PDX_OK;
}

int upload_twiddles(pdx_mesh_t m) {
  std::vector<float2> tw(m->K1 + m->K2 + m->K3);
  size_t o = 0;
  for (int K : {m->K1, m->K2, m->K3})
    for (int k = 0; k < K; ++k, ++o) tw[o] = make_float2((float)std::cos(2 * M_PI * k / K), (float)-std::sin(2 * M_PI * k / K));
  PDX_HIP(hipMemcpy(m->tw.p, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice));
  return PDX_OK;
}

void refresh_background(pdx_mesh_t m) {
  const double V = m->box[0] * m->box[1] * m->box[2];
  const bool bgon = m->spec.net_charge == PDX_NET_CHARGE_BACKGROUND;
  m->bg = bgon ? (float)(-m->C * m->bgcoef * m->qtotal / V) : 0.0f;
  m->ubg = bgon ? -m->C * m->bgcoef * m->qtotal * m->qtotal / V : 0.0;
}

struct Plans {
  FFTPlan p1, p2, p3;
  int T1, T2, T3;  // complex lines per tile (the x passes pack two real lines per complex line)
};

Plans plans_of(pdx_mesh_t m) {
  Plans pl;
  make_plan(m->K1, pl.p1);
  make_plan(m->K2, pl.p2);
  make_plan(m->K3, pl.p3);
  if (m->T1 == 0) {
    m->T1 = lines_per_tile(m->K1, 4, "PDX_FFT_TX");
    m->T2 = lines_per_tile(m->K2, 8, "PDX_FFT_TY");
    m->T3 = lines_per_tile(m->K3, 8, "PDX_FFT_TZ");
  }
  pl.T1 = m->T1;
  pl.T2 = m->T2;
  pl.T3 = m->T3;
  return pl;
}

SolveArgs solve_args(pdx_mesh_t m) {
  SolveArgs s{};
  s.G = m->G.p;
  s.Vc = m->Vc.p;
  s.ez = nullptr;
  s.nmodes = (size_t)m->K1h * m->K2 * m->K3;
  s.K1 = m->K1; s.K2 = m->K2; s.K3 = m->K3; s.K1h = m->K1h;
  for (int d = 0; d < 3; ++d) s.twopi_L[d] = 2 * M_PI / m->box[d];
  s.inv4a2 = m->alpha > 0 ? 1.0 / (4.0 * m->alpha * m->alpha) : 0.0;  // Gaussian closed form only (no Vc table)
  for (int d = 0; d < 3; ++d) s.twopi_Lf[d] = (float)s.twopi_L[d];
  s.inv2a2f = (float)(2.0 * s.inv4a2);
  s.partial = m->partial.p;
  s.acc = m->acc.p;
  return s;
}

// twiddles (K) + two ping-pong buffers (T K each) + 2T spare slots: the inverse x pass stages
// 2T half-spectrum lines (2T (K/2 + 1) values) in the second buffer
size_t fft_lds(int T, int K) { return ((size_t)2 * T * K + K + 2 * T) * sizeof(float2); }

#define PDX_T_SWITCH(T, CALL) \
  switch (T) {                \
    case 4: CALL(4); break;   \
    case 8: CALL(8); break;   \
    case 16: CALL(16); break; \
    default: CALL(32); break; \
  }

int enqueue_forward_xy(pdx_mesh_t m, const Plans &pl, const void *src, bool fixed, hipStream_t st) {
  const float2 *tw1 = m->tw.p, *tw2 = m->tw.p + m->K1;
  const int nlines = m->K2 * m->K3;
  const int gx = (nlines + 2 * pl.T1 - 1) / (2 * pl.T1);
  const size_t sh1 = fft_lds(pl.T1, m->K1);
#define PDX_XF2(TT, BB)                                                                                       \
  if (fixed)                                                                                                  \
    hipLaunchKernelGGL((k_fft_x_r2c<TT, true, BB>), dim3(gx), dim3(TPB), sh1, st, src, m->grid64.p,            \
                       m->spec_c.p, m->K1, m->K1h, nlines, pl.p1, tw1, m->fixed_inv);                          \
  else                                                                                                        \
    hipLaunchKernelGGL((k_fft_x_r2c<TT, false, BB>), dim3(gx), dim3(TPB), sh1, st, src, nullptr, m->spec_c.p,  \
                       m->K1, m->K1h, nlines, pl.p1, tw1, 1.0);
#define PDX_XF(TT) \
  if (pl.p1.big) { PDX_XF2(TT, true) } else { PDX_XF2(TT, false) }
  PDX_T_SWITCH(pl.T1, PDX_XF)
#undef PDX_XF
#undef PDX_XF2
  const int gy = ((m->K1h + pl.T2 - 1) / pl.T2) * m->K3;
#define PDX_YF(TT)                                                                                         \
  if (pl.p2.big)                                                                                             \
    hipLaunchKernelGGL((k_fft_y<TT, true>), dim3(gy), dim3(TPB), fft_lds(TT, m->K2), st, m->spec_c.p,          \
                       m->spec_c.p, m->K1h, m->K2, pl.p2, tw2, -1.0f, 0, 0.0f, m->K1, nullptr, 0,      \
                       nullptr);                        \
  else                                                                                                       \
    hipLaunchKernelGGL((k_fft_y<TT, false>), dim3(gy), dim3(TPB), fft_lds(TT, m->K2), st, m->spec_c.p,         \
                       m->spec_c.p, m->K1h, m->K2, pl.p2, tw2, -1.0f, 0, 0.0f, m->K1, nullptr, 0,      \
                       nullptr);
  PDX_T_SWITCH(pl.T2, PDX_YF)
#undef PDX_YF
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

// Inverse y pass from `src` into `spec` (in place when equal; times -i k_d first when ikdim is 1
// or 2), then the inverse x pass of `spec` into the real grid `dst`.
int enqueue_inverse_yx(pdx_mesh_t m, const Plans &pl, const float2 *src, float2 *spec, float *dst, int ikdim,
                       hipStream_t st, bool with_sum = false) {
  const double *ps = with_sum ? m->partial.p : nullptr;
  double *po = with_sum ? m->acc.p : nullptr;
  const int nps = m->nmode_blocks;
  const float2 *tw1 = m->tw.p, *tw2 = m->tw.p + m->K1;
  const int gy = ((m->K1h + pl.T2 - 1) / pl.T2) * m->K3;
  const float kscale = ikdim == 1 ? (float)(2 * M_PI / m->box[0]) : (float)(2 * M_PI / m->box[1]);
#define PDX_YI(TT)                                                                                         \
  if (pl.p2.big)                                                                                             \
    hipLaunchKernelGGL((k_fft_y<TT, true>), dim3(gy), dim3(TPB), fft_lds(TT, m->K2), st, src, spec, m->K1h,    \
                       m->K2, pl.p2, tw2, +1.0f, ikdim, kscale, m->K1, ps, nps, po);                                      \
  else                                                                                                       \
    hipLaunchKernelGGL((k_fft_y<TT, false>), dim3(gy), dim3(TPB), fft_lds(TT, m->K2), st, src, spec, m->K1h,   \
                       m->K2, pl.p2, tw2, +1.0f, ikdim, kscale, m->K1, ps, nps, po);
  PDX_T_SWITCH(pl.T2, PDX_YI)
#undef PDX_YI
  const int nlines = m->K2 * m->K3;
  const int gx = (nlines + 2 * pl.T1 - 1) / (2 * pl.T1);
#define PDX_XI(TT)                                                                                            \
  if (pl.p1.big)                                                                                                \
    hipLaunchKernelGGL((k_fft_x_c2r<TT, true>), dim3(gx), dim3(TPB), fft_lds(TT, m->K1), st, spec, dst,           \
                       m->K1, m->K1h, nlines, pl.p1, tw1);                                                      \
  else                                                                                                          \
    hipLaunchKernelGGL((k_fft_x_c2r<TT, false>), dim3(gx), dim3(TPB), fft_lds(TT, m->K1), st, spec, dst,          \
                       m->K1, m->K1h, nlines, pl.p1, tw1);
  PDX_T_SWITCH(pl.T1, PDX_XI)
#undef PDX_XI
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

template <int MODE, bool ENERGY>
void launch_z(pdx_mesh_t m, const Plans &pl, hipStream_t st, float2 *spec = nullptr, float2 *ez = nullptr) {
  const float2 *tw3 = m->tw.p + m->K1 + m->K2;
  const int gz = ((m->K1h + pl.T3 - 1) / pl.T3) * m->K2;
  if (!spec) spec = m->spec_c.p;
  SolveArgs sa = solve_args(m);
  sa.ez = ez;
#define PDX_ZZ(TT)                                                                                              \
  if (pl.p3.big)                                                                                                  \
    hipLaunchKernelGGL((k_fft_z<TT, MODE, ENERGY, true>), dim3(gz), dim3(TPB), fft_lds(TT, m->K3), st,              \
                       spec, pl.p3, tw3, sa);                                                                    \
  else                                                                                                            \
    hipLaunchKernelGGL((k_fft_z<TT, MODE, ENERGY, false>), dim3(gz), dim3(TPB), fft_lds(TT, m->K3), st,             \
                       spec, pl.p3, tw3, sa);
  PDX_T_SWITCH(pl.T3, PDX_ZZ)
#undef PDX_ZZ
}

// Plane passes (the default when a plane fits the LDS): lines of every plane at most 64
// (K1/2 + 1, ceil(K2/2)), sides at most 128 with radices 2, 3, 4, 5 only, not the ik path;
// PDX_FFT_PLANE=0 turns them off.
int plane_lines(pdx_mesh_t m) {
  if (const char *v = std::getenv("PDX_FFT_PLANE"))
    if (std::atoi(v) == 0) return 0;
  if (m->spec.influence == PDX_INFLUENCE_P3M_IK) return 0;
  if (m->K1 > 128 || m->K2 > 128 || m->K3 > 128) return 0;
  FFTPlan p;  // radices 7, 11, 13 need more registers than a 1024-thread workgroup leaves
  for (int K : {m->K1, m->K2, m->K3})
    if (!make_plan(K, p) || p.big) return 0;
  const int lines = std::max(m->K1h, (m->K2 + 1) / 2);
  if (lines > 64) return 0;
  return lines <= 32 ? 32 : 64;
}

size_t plane_lds_xy(pdx_mesh_t m) {
  return ((size_t)m->K1 + m->K2 + (size_t)2 * std::max(m->K1, m->K2) * (m->plane_tp + 1)) * sizeof(float2);
}
// z solve: lines of a plane in tiles of 32 (two or more workgroups per plane when K1/2 + 1 > 32)
constexpr int PLANE_TZ = 32;
int plane_z_tiles(pdx_mesh_t m) { return (m->K1h + PLANE_TZ - 1) / PLANE_TZ; }
size_t plane_lds_z(pdx_mesh_t m) { return ((size_t)m->K3 + (size_t)2 * m->K3 * (PLANE_TZ + 1)) * sizeof(float2); }

#define PDX_PLANE_SWITCH(CALL)                                    \
  (void)big;                                                       \
  if (m->plane_tp == 32) {                                         \
    CALL(32, false)                                                \
  } else {                                                         \
    CALL(64, false)                                                \
  }

int plane_attributes(pdx_mesh_t m, const Plans &pl) {
  const size_t lxy = plane_lds_xy(m), lz = plane_lds_z(m);
  bool big = pl.p1.big || pl.p2.big;
#define PDX_PA_XY_BS(TP, BB, BS)                                                                              \
  PDX_HIP(hipFuncSetAttribute((const void *)k_plane_fwd<TP, BB, BS>, hipFuncAttributeMaxDynamicSharedMemorySize,\
                              (int)lxy));                                                                     \
  PDX_HIP(hipFuncSetAttribute((const void *)k_plane_inv<TP, BB, BS>, hipFuncAttributeMaxDynamicSharedMemorySize,\
                              (int)lxy));
#define PDX_PA_XY(TP, BB) PDX_PA_XY_BS(TP, BB, 1024) PDX_PA_XY_BS(TP, BB, 512)
  PDX_PLANE_SWITCH(PDX_PA_XY)
#undef PDX_PA_XY
#undef PDX_PA_XY_BS
  big = pl.p3.big;
  PDX_HIP(hipFuncSetAttribute((const void *)k_plane_z<PLANE_TZ, true, false, 1024>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lz));
  PDX_HIP(hipFuncSetAttribute((const void *)k_plane_z<PLANE_TZ, false, false, 1024>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lz));
  PDX_HIP(hipFuncSetAttribute((const void *)k_plane_z<PLANE_TZ, true, false, 512>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lz));
  PDX_HIP(hipFuncSetAttribute((const void *)k_plane_z<PLANE_TZ, false, false, 512>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lz));
  return PDX_OK;
}

// The SPME / P3M-AD solve as three plane launches, in workgroups of BS threads.
template <int BS>
int enqueue_plane_solve_bs(pdx_mesh_t m, const Plans &pl, bool energy, hipStream_t st) {
  const int KM = std::max(m->K1, m->K2);
  const size_t lxy = plane_lds_xy(m), lz = plane_lds_z(m);
  const float2 *tw1 = m->tw.p, *tw3 = m->tw.p + m->K1 + m->K2;
  bool big = pl.p1.big || pl.p2.big;
#define PDX_PF(TP, BB)                                                                                        \
  hipLaunchKernelGGL((k_plane_fwd<TP, BB, BS>), dim3(m->K3), dim3(BS), lxy, st, (const unsigned long long *)m->grid64.p, \
                     m->spec_c.p, m->K1, m->K2, m->K1h, KM, pl.p1, pl.p2, tw1, m->fixed_inv);
  PDX_PLANE_SWITCH(PDX_PF)
#undef PDX_PF
  SolveArgs sa = solve_args(m);
  const int zt = plane_z_tiles(m);
  if (energy)
    hipLaunchKernelGGL((k_plane_z<PLANE_TZ, true, false, BS>), dim3(m->K2 * zt), dim3(BS), lz, st, m->spec_c.p, pl.p3,
                       tw3, sa, zt);
  else
    hipLaunchKernelGGL((k_plane_z<PLANE_TZ, false, false, BS>), dim3(m->K2 * zt), dim3(BS), lz, st, m->spec_c.p, pl.p3,
                       tw3, sa, zt);
  big = pl.p1.big || pl.p2.big;
  const double *ps = energy ? m->partial.p : nullptr;
  double *po = energy ? m->acc.p : nullptr;
#define PDX_PI(TP, BB)                                                                                        \
  hipLaunchKernelGGL((k_plane_inv<TP, BB, BS>), dim3(m->K3), dim3(BS), lxy, st, (const float2 *)m->spec_c.p, m->pot.p, \
                     m->K1, m->K2, m->K1h, KM, pl.p1, pl.p2, tw1, ps, m->nmode_blocks, po);
  PDX_PLANE_SWITCH(PDX_PI)
#undef PDX_PI
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

int enqueue_plane_solve(pdx_mesh_t m, const Plans &pl, bool energy, hipStream_t st) {
  if (m->transform_threads == 512) return enqueue_plane_solve_bs<512>(m, pl, energy, st);
  return enqueue_plane_solve_bs<PBS>(m, pl, energy, st);
}

// the window codes compiled: B-spline orders 3..8 and prolate widths 4..8
#define PDX_WINDOWS(CASE) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) \
  CASE(PROLATE + 4) CASE(PROLATE + 5) CASE(PROLATE + 6) CASE(PROLATE + 7) CASE(PROLATE + 8)

#define PDX_SPREAD_CASE(PP)                                                                                    \
  case PP:                                                                                                     \
    hipLaunchKernelGGL(k_spread<PP>, dim3(((int64_t)N * lanes_for(win_width(PP)) + TPB - 1) / TPB), dim3(TPB), \
                       0, st, N, x, y, z, q, order, m->grid64.p, m->K1, m->K2, m->K3, L[0], L[1], L[2], scale,  \
                       m->wc);                                                                                 \
    break;

#define PDX_GATHER_CASE(PP)                                                                                     \
  case PP:                                                                                                      \
    if (energy)                                                                                                 \
      hipLaunchKernelGGL((k_gather<PP, true>), dim3(blocks_of(PP)), dim3(TPB), 0, st, N, x, y, z, q, order,     \
                         m->pot.p, m->K1, m->K2, m->K3, L[0], L[1], L[2], fx, fy, fz, Epot, m->bg, set,         \
                         Vglobal, m->acc.p, m->ubg, xcd, m->wc);                                               \
    else                                                                                                        \
      hipLaunchKernelGGL((k_gather<PP, false>), dim3(blocks_of(PP)), dim3(TPB), 0, st, N, x, y, z, q, order,    \
                         m->pot.p, m->K1, m->K2, m->K3, L[0], L[1], L[2], fx, fy, fz, Epot, m->bg, set,         \
                         Vglobal, m->acc.p, m->ubg, xcd, m->wc);                                               \
    break;

#define PDX_GATHER_IK_CASE(PP)                                                                                  \
  case PP:                                                                                                      \
    if (energy)                                                                                                 \
      hipLaunchKernelGGL((k_gather_ik<PP, true>), dim3(blocks_of(PP)), dim3(TPB), 0, st, N, x, y, z, q, order,  \
                         m->ex.p, m->ey.p, m->ez.p, m->pot.p, m->K1, m->K2, m->K3, L[0], L[1], L[2], fx, fy,    \
                         fz, Epot, m->bg, set, Vglobal, m->acc.p, m->ubg);                                      \
    else                                                                                                        \
      hipLaunchKernelGGL((k_gather_ik<PP, false>), dim3(blocks_of(PP)), dim3(TPB), 0, st, N, x, y, z, q, order, \
                         m->ex.p, m->ey.p, m->ez.p, m->pot.p, m->K1, m->K2, m->K3, L[0], L[1], L[2], fx, fy,    \
                         fz, Epot, m->bg, set, Vglobal, m->acc.p, m->ubg);                                      \
    break;

}  // namespace

extern "C" int pdx_mesh_create(const pdx_mesh_spec *spec, int32_t n, const double box[3], pdx_mesh_t *out) {
  if (!spec || !box || !out) return fail(PDX_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  const pdx_mesh_spec &sp = *spec;
  if (!(std::isfinite(sp.coulomb_prefactor) && sp.coulomb_prefactor > 0))
    return fail(PDX_ERR_ARGUMENT, "coulomb_prefactor must be finite and positive (it carries the caller's E*L/Q^2 unit).");
  if (!(std::isfinite(sp.relative_permittivity) && sp.relative_permittivity > 0))
    return fail(PDX_ERR_ARGUMENT, "relative_permittivity must be finite and positive.");
  if (!(std::isfinite(sp.real_cutoff) && sp.real_cutoff > 0))
    return fail(PDX_ERR_ARGUMENT, "real_cutoff must be finite and positive.");
  const bool esp = sp.influence == PDX_INFLUENCE_ESP || sp.influence == PDX_INFLUENCE_ESP_DECONVOLUTION;
  if (sp.influence != PDX_INFLUENCE_SPME && sp.influence != PDX_INFLUENCE_P3M_AD && sp.influence != PDX_INFLUENCE_P3M_IK &&
      !esp)
    return fail(PDX_ERR_ARGUMENT, "Unknown influence kind.");
  if (sp.influence != PDX_INFLUENCE_SPME && (sp.alias_order < 0 || sp.alias_order > 6))
    return fail(PDX_ERR_ARGUMENT, "alias_order must be between 0 and 6.");
  if (esp) {
    if (sp.order < 2 || sp.order > 16)
      return fail(PDX_ERR_ARGUMENT, "Prolate window width must be between 2 and 16 grid points.");
    if (sp.order < 4 || sp.order > 8)
      return fail(PDX_ERR_ARGUMENT, "Prolate window widths other than 4..8 grid points are not compiled in this build.");
  } else {
    if (sp.order < 3 || sp.order > 12)
      return fail(PDX_ERR_ARGUMENT, "B-spline assignment order must be between 3 and 12.");
    if (sp.order > 8) return fail(PDX_ERR_ARGUMENT, "Assignment orders above 8 are not compiled in this build.");
  }
  if (sp.net_charge != PDX_NET_CHARGE_ERROR && sp.net_charge != PDX_NET_CHARGE_BACKGROUND)
    return fail(PDX_ERR_ARGUMENT, "net_charge must be :error or :background.");
  if (sp.real_space_modifier != 0 && sp.real_space_modifier != 1)
    return fail(PDX_ERR_ARGUMENT, "real_space_modifier must be :none or :potential_shift.");
  if (n <= 0) return fail(PDX_ERR_ARGUMENT, "particle count must be positive");
  for (int d = 0; d < 3; ++d)
    if (!(std::isfinite(box[d]) && box[d] > 0)) return fail(PDX_ERR_ARGUMENT, "box lengths must be finite and positive");
  const double mn = std::min(box[0], std::min(box[1], box[2]));
  if (!(2 * sp.real_cutoff < mn))
    return fail(PDX_ERR_ARGUMENT, "real_cutoff must be smaller than half the shortest box length (minimum-image real space).");
  // ESP: the prolate split and window (ProlateSplit / ProlateWindow, Splits.jl:48-60, Windows.jl:64-71;
  // the spec checks of Spec.jl:66-76, the FP32 resolution bound of Windows.jl:310-330)
  ProlateSplit psplit;
  ProlateWindow pwin;
  WinCoef wcoef{};
  int wdeg = 0;
  if (esp) {
    if (!(std::isfinite(sp.split_shape) && sp.split_shape > 0))
      return fail(PDX_ERR_ARGUMENT, "ESP split shape must be finite and positive.");
    if (!(std::isfinite(sp.window_shape) && sp.window_shape > 0))
      return fail(PDX_ERR_ARGUMENT, "Prolate window shape must be finite and positive.");
    if (!psplit.make(sp.real_cutoff, sp.split_shape) || !pwin.make(sp.order, sp.window_shape))
      return fail(PDX_ERR_ARGUMENT, "prolate function set-up failed (shape out of range)");
    if (sp.order > prolate_max_order(sp.window_shape))
      return fail(PDX_ERR_ARGUMENT, "A prolate window of shape " + std::to_string(sp.window_shape) + " supports at most " +
                                        std::to_string(prolate_max_order(sp.window_shape)) +
                                        " grid points (P <= 2c/pi - 1/2); got " + std::to_string(sp.order) +
                                        ". Use a larger window shape (tighter tolerance) or a narrower window.");
    if (!(pwin.nyquist_attenuation() >= prolate_min_attenuation_fp32()))
      return fail(PDX_ERR_ARGUMENT, "The prolate window is not resolved by an FP32 mesh: its Nyquist attenuation " +
                                        std::to_string(pwin.nyquist_attenuation()) + " is below 0.02 eps32^(2/3); "
                                        "use a larger window shape.");
    double ew, ed;
    wdeg = pwin.node_polynomials(wcoef, ew, ed);
    if (wdeg < 0)
      return fail(PDX_ERR_ARGUMENT, "the prolate window's node polynomials need more than degree 12 in FP32");
  }
  double alpha = sp.alpha;
  if (esp) {
    alpha = 0.0;
  } else if (!(alpha > 0)) {
    const double tol = sp.real_space_tolerance;
    if (!(tol > 0 && tol < 1)) return fail(PDX_ERR_ARGUMENT, "real_space_tolerance must lie in (0, 1).");
    double lo = 0.0, hi = 10.0;  // alpha_from_tolerance (Spec.jl:154-163): bisection of erfc(x) = tol
    for (int it = 0; it < 200; ++it) {
      const double mid = 0.5 * (lo + hi);
      (host_erfc(mid) - tol > 0) ? lo = mid : hi = mid;
    }
    alpha = 0.5 * (lo + hi) / sp.real_cutoff;
  }
  if (!esp && !(std::isfinite(alpha) && alpha > 0)) return fail(PDX_ERR_ARGUMENT, "Ewald alpha must be finite and positive.");
  int K[3];
  for (int d = 0; d < 3; ++d) {
    if (sp.mesh[0] > 0) {
      K[d] = sp.mesh[d];
    } else {
      if (!(std::isfinite(sp.mesh_spacing) && sp.mesh_spacing > 0))
        return fail(PDX_ERR_ARGUMENT, "Provide either `mesh=(K1,K2,K3)` or `mesh_spacing`.");
      K[d] = std::max(fft_friendly((int)std::ceil(box[d] / sp.mesh_spacing - 1e-9)), sp.order);
    }
    if (K[d] < sp.order) return fail(PDX_ERR_ARGUMENT, "Every mesh dimension must be at least the assignment order.");
    FFTPlan pl;
    if (!make_plan(K[d], pl))
      return fail(PDX_ERR_ARGUMENT, "Mesh dimensions must have prime factors in {2, 3, 5, 7, 11, 13}.");
    // one LDS tile holds 4 lines of K complex values twice plus the twiddles
    if (K[d] > 760) return fail(PDX_ERR_ARGUMENT, "Mesh dimensions above 760 are not supported.");
  }
  const int64_t nnodes = (int64_t)K[0] * K[1] * K[2];
  if (nnodes >= INT32_MAX) return fail(PDX_ERR_ARGUMENT, "Mesh is too large for 32-bit grid indexing.");
  auto *m = new pdx_mesh_s();
  m->spec = sp;
  m->N = n;
  m->K1 = K[0]; m->K2 = K[1]; m->K3 = K[2]; m->K1h = K[0] / 2 + 1;
  m->P = sp.order;
  for (int d = 0; d < 3; ++d) m->box[d] = box[d];
  m->alpha = alpha;
  m->C = sp.coulomb_prefactor / sp.relative_permittivity;
  m->selfc = esp ? psplit.self_coefficient() : 2.0 * alpha / std::sqrt(M_PI);  // self_coefficient (Splits.jl:67-68)
  m->bgcoef = esp ? psplit.background_coefficient() : M_PI / (2.0 * alpha * alpha);  // background_coefficient (Splits.jl:85-86)
  m->wkey = esp ? PROLATE + sp.order : sp.order;
  m->wc = wcoef;
  m->split_shape = esp ? sp.split_shape : 0.0;
  m->window_shape = esp ? sp.window_shape : 0.0;
  m->window_degree = wdeg;
  const size_t nmodes = (size_t)m->K1h * m->K2 * m->K3;
  const Plans pl = plans_of(m);
  m->plane_tp = plane_lines(m);
  m->nmode_blocks = m->plane_tp ? m->K2 * plane_z_tiles(m) : ((m->K1h + pl.T3 - 1) / pl.T3) * m->K2;
  int rc;
  if ((rc = m->grid64.alloc(nnodes)) || (rc = m->spec_c.alloc(nmodes)) || (rc = m->pot.alloc(nnodes)) ||
      (rc = m->G.alloc(nmodes)) || (rc = m->tw.alloc(K[0] + K[1] + K[2])) || (rc = m->acc.alloc(8)) ||
      (rc = m->partial.alloc((size_t)7 * m->nmode_blocks))) {
    delete m;
    return rc;
  }
  if (sp.influence != PDX_INFLUENCE_SPME &&
      ((rc = m->Vc.alloc(6 * nmodes)) ||
       (rc = m->Utab.alloc((size_t)(K[0] + K[1] + K[2]) * (2 * alias_order_of(m) + 1))))) {
    delete m;
    return rc;
  }
  if (esp) {
    std::vector<double> q(256);
    for (int j = 0; j < 128; ++j) {
      q[j] = psplit.xq[j];
      q[128 + j] = psplit.wchi[j];
    }
    if ((rc = m->espq.alloc(256))) {
      delete m;
      return rc;
    }
    if (hipMemcpy(m->espq.p, q.data(), sizeof(double) * 256, hipMemcpyHostToDevice) != hipSuccess) {
      delete m;
      return fail(PDX_ERR_DEVICE, "mesh initialisation failed");
    }
  }
  if (sp.influence == PDX_INFLUENCE_P3M_IK &&
      ((rc = m->spec_ez.alloc(nmodes)) || (rc = m->spec_tmp.alloc(nmodes)) || (rc = m->ex.alloc(nnodes)) ||
       (rc = m->ey.alloc(nnodes)) || (rc = m->ez.alloc(nnodes)))) {
    delete m;
    return rc;
  }
  {
    // tile lists: twice the mean occupancy plus 64 entries per tile, for the larger tile count
    const int nt8 = ((K[0] + 7) / 8) * ((K[1] + 7) / 8) * ((K[2] + 7) / 8);
    m->tile_entries = (size_t)2 * n + (size_t)64 * nt8;
    if ((rc = m->tile_list.alloc(m->tile_entries)) || (rc = m->tile_ids.alloc(m->tile_entries)) ||
        (rc = m->tile_n.alloc(nt8))) {
      delete m;
      return rc;
    }
    if (hipMemset(m->tile_n.p, 0, sizeof(int32_t) * nt8) != hipSuccess) {
      delete m;
      return fail(PDX_ERR_DEVICE, "mesh initialisation failed");
    }
    // owner-form tile blocks, for either tile edge, up to 1 GiB (a 256^3 mesh at 8-node tiles)
    const int P = sp.order;
    const int nt16 = ((K[0] + 15) / 16) * ((K[1] + 15) / 16) * ((K[2] + 15) / 16);
    const size_t b16 = (size_t)nt16 * (16 + P - 1) * (16 + P - 1) * (16 + P - 1);
    const size_t b8 = (size_t)nt8 * (8 + P - 1) * (8 + P - 1) * (8 + P - 1);
    const size_t nb = std::max(b16, b8);
    if (nb * sizeof(unsigned long long) <= ((size_t)1 << 30)) {
      if ((rc = m->tile_blocks.alloc(nb))) {
        delete m;
        return rc;
      }
      m->tile_block_entries = nb;
    }
  }
  if (hipMemset(m->grid64.p, 0, sizeof(unsigned long long) * nnodes) != hipSuccess ||
      hipMemset(m->acc.p, 0, sizeof(double) * 8) != hipSuccess ||
      hipMemset(m->pot.p, 0, sizeof(float) * nnodes) != hipSuccess) {
    delete m;
    return fail(PDX_ERR_DEVICE, "mesh initialisation failed");
  }
  if ((rc = compute_influence(m)) || (rc = upload_twiddles(m)) || (m->plane_tp && (rc = plane_attributes(m, pl)))) {
    delete m;
    return rc;
  }
  m->fixed_scale = std::ldexp(1.0, 40);
  m->fixed_inv = std::ldexp(1.0, -40);
  refresh_background(m);
  if (hipDeviceSynchronize() != hipSuccess) {
    delete m;
    return fail(PDX_ERR_DEVICE, "mesh initialisation failed");
  }
  *out = m;
  return PDX_OK;
}

extern "C" int pdx_esp_rule(double tolerance, int32_t order, double out[3]) {
  // the ESP defaults of a target relative force error (electrostatics_spec(::ESP, Float32) and
  // esp_order_for_tolerance / esp_window_shape, src/workflow/Electrostatics.jl:177-215,288-301)
  if (!out) return fail(PDX_ERR_ARGUMENT, "null argument");
  if (!(tolerance > 0 && tolerance < 1)) return fail(PDX_ERR_ARGUMENT, "ESP tolerance must lie in (0, 1).");
  const double c = prolate_shape_for_tolerance(tolerance);
  if (!(c > 0)) return fail(PDX_ERR_ARGUMENT, "Tolerance is out of the range of a prolate split.");
  int P = order;
  if (P <= 0) {
    const double d = -std::log10(tolerance);
    if (d <= 3.35) P = 5;
    else if (d <= 4.05) P = 6;
    else if (d <= 4.35) P = 7;
    else if (d <= 5.05) P = 8;
    else {
      const double c1 = prolate_shape_for_tolerance(tolerance / 2);
      P = std::min(std::max((int)std::floor(2 * c1 / M_PI - 1), 8), 16);
    }
  }
  // window shape: c(eps/2), or the smallest larger shape an FP32 mesh resolves
  auto resolved = [&](double shape) {
    ProlateWindow w;
    return w.make(P, shape) && w.nyquist_attenuation() >= prolate_min_attenuation_fp32();
  };
  double c1 = prolate_shape_for_tolerance(tolerance / 2);
  if (!(c1 > 0)) return fail(PDX_ERR_ARGUMENT, "Tolerance is out of the range of a prolate window.");
  if (!resolved(c1)) {
    double lo = c1, hi = c1 + 0.5;
    while (!resolved(hi)) {
      lo = hi;
      hi += 0.5;
      if (!(hi < 80)) return fail(PDX_ERR_ARGUMENT, "No prolate window of this width is resolved in FP32.");
    }
    while (hi - lo > 1e-3) {
      const double mid = 0.5 * (lo + hi);
      (resolved(mid) ? hi : lo) = mid;
    }
    c1 = hi;
  }
  out[0] = c;
  out[1] = P;
  out[2] = c1;
  return PDX_OK;
}

extern "C" int pdx_mesh_destroy(pdx_mesh_t m) {
  delete m;
  return PDX_OK;
}

extern "C" int pdx_mesh_info(pdx_mesh_t m, double out[9]) {
  if (!m || !out) return fail(PDX_ERR_ARGUMENT, "null argument");
  out[0] = m->K1; out[1] = m->K2; out[2] = m->K3; out[3] = m->P;
  out[4] = m->alpha; out[5] = m->C; out[6] = m->selfc; out[7] = m->bg;
  out[8] = m->plane_tp;
  return PDX_OK;
}

extern "C" int pdx_mesh_set_charges(pdx_mesh_t m, const float *q, int32_t n) {
  if (!m || (!q && n > 0)) return fail(PDX_ERR_ARGUMENT, "null argument");
  if (n != m->N) return fail(PDX_ERR_DIMENSION, "charge count differs from the mesh particle count");
  double Q = 0, qabs = 0, qsq = 0, qmax = 0;
  for (int i = 0; i < n; ++i) {
    qmax = std::max(qmax, (double)std::fabs(q[i]));
    if (!std::isfinite(q[i])) return fail(PDX_ERR_ARGUMENT, "Charges must be finite.");
    Q += q[i];
    qabs += std::fabs((double)q[i]);
    qsq += (double)q[i] * q[i];
  }
  if (m->spec.net_charge == PDX_NET_CHARGE_ERROR && std::fabs(Q) > m->spec.net_charge_tolerance * std::max(qabs, 1.0))
    return fail(PDX_ERR_ARGUMENT,
                "The system carries net charge. Neutralise it, or select net_charge=:background to add an explicit "
                "uniform neutralising background.");
  m->qtotal = Q;
  m->qabs = qabs;
  m->qsq = qsq;
  // fixed-point scale 2^S: every node sum is bounded by sum |q| (weights sum to 1 per
  // particle), so |node| * 2^S < 2^62 cannot overflow; S <= 40 keeps 2^-40 resolution, and
  // S <= 50 - log2 max|q| keeps every single contribution q w w w 2^S below 2^50 (the range
  // of the spreads' FP64 rounding; no change for charges up to 1024)
  int S = (int)std::floor(62.0 - std::log2(std::max(qabs, 1.0)));
  S = std::min(S, 40);
  S = std::min(S, (int)std::floor(50.0 - std::log2(std::max(qmax, 1.0))));
  m->fixed_scale = std::ldexp(1.0, S);
  m->fixed_inv = std::ldexp(1.0, -S);
  refresh_background(m);
  return PDX_OK;
}

extern "C" int pdx_mesh_set_box(pdx_mesh_t m, const double box[3]) {
  if (!m || !box) return fail(PDX_ERR_ARGUMENT, "null argument");
  for (int d = 0; d < 3; ++d)
    if (!(std::isfinite(box[d]) && box[d] > 0)) return fail(PDX_ERR_ARGUMENT, "box lengths must be finite and positive");
  const double mn = std::min(box[0], std::min(box[1], box[2]));
  if (!(2 * m->spec.real_cutoff < mn)) return fail(PDX_ERR_ARGUMENT, "real_cutoff must stay below half the shortest box length.");
  PDX_HIP(hipDeviceSynchronize());
  for (int d = 0; d < 3; ++d) m->box[d] = box[d];
  m->tiles_nl = nullptr;
  int rc = compute_influence(m);
  if (rc) return rc;
  refresh_background(m);
  return PDX_OK;
}

extern "C" int pdx_mesh_spread(pdx_mesh_t m, const float *x, const float *y, const float *z, const float *q,
                               const int32_t *order, pdx_stream_t stream) {
  if (!m || !x || !y || !z || !q) return fail(PDX_ERR_ARGUMENT, "null argument");
  const int N = m->N;
  m->tiles_nl = nullptr;  // this spread writes no tile lists
  m->tiles_snapshot = false;
  const float L[3] = {(float)m->box[0], (float)m->box[1], (float)m->box[2]};
  const float scale = (float)m->fixed_scale;
  hipStream_t st = as_stream(stream);
  const size_t nn = (size_t)m->K1 * m->K2 * m->K3;
  hipLaunchKernelGGL(k_zero64, dim3((nn + TPB - 1) / TPB), dim3(TPB), 0, st, m->grid64.p, nn);
  switch (m->wkey) {
    PDX_WINDOWS(PDX_SPREAD_CASE)
    default: return fail(PDX_ERR_ARGUMENT, "unsupported order");
  }
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

namespace {
// Owner-computes spread from the cells of nl: the current cells (margin skin/2, the most a
// particle moved since they were built) or the snapshot of the last completed build (margin skin:
// one step past a rebuild trigger moves no particle by another skin/2).
int spread_cells_impl(pdx_mesh_t m, pdx_nl_t nl, bool snapshot, pdx_stream_t stream) {
  if (!m || !nl) return fail(PDX_ERR_ARGUMENT, "null argument");
  if (nl->N != m->N) return fail(PDX_ERR_DIMENSION, "neighbor list and mesh particle counts differ");
  if (nl->dim != 3) return fail(PDX_ERR_ARGUMENT, "the mesh is three-dimensional; 2-D lists have no mesh");
  if (nl->allpairs) return fail(PDX_ERR_ARGUMENT, "all-pairs lists have no cells; use pdx_mesh_spread");
  for (int d = 0; d < 3; ++d)
    if (std::fabs(nl->box[d] - m->box[d]) > 1e-12 * m->box[d])
      return fail(PDX_ERR_ARGUMENT, "neighbor list and mesh boxes differ");
  const float L[3] = {(float)m->box[0], (float)m->box[1], (float)m->box[2]};
  const float scale = (float)m->fixed_scale;
  auto ntiles = [&](int T) { return ((m->K1 + T - 1) / T) * ((m->K2 + T - 1) / T) * ((m->K3 + T - 1) / T); };
  bool small = ntiles(16) < 256;
  if (const char *e = std::getenv("PDX_SPREAD_TILE")) small = std::atoi(e) == 8;
  const int tiles = ntiles(small ? 8 : 16);
  const float margin = (snapshot ? 1.0f : 0.5f) * (float)nl->skin;
  const int32_t *offsets = snapshot ? nl->snap_offsets.p : nl->cell_offsets.p;
  const int32_t *ids = snapshot ? nl->snap_ids.p : nl->particle_ids_sorted.p;
  const int64_t *guard = snapshot ? nl->snap_state.p : nl->state.p;  // word 2 nonzero: no usable cells
  hipStream_t st = as_stream(stream);
  // workgroup size (pdx_mesh_set_spread_threads): 512 (16 particles at a time per tile; half
  // the waves of 1024 to launch, which the evaluation measures faster), 1024, or 256 (more room
  // beside a concurrent pair kernel)
  int sbs = m->spread_threads;
  if (const char *e = std::getenv("PDX_SPREAD_BS")) {
    const int v = std::atoi(e);
    sbs = (v == 256 || v == 1024) ? v : 512;
  }
  int tcap = (int)std::min<size_t>(m->tile_entries / tiles, INT32_MAX);
  if (const char *e = std::getenv("PDX_TILE_CAP")) tcap = std::max(0, std::min(tcap, std::atoi(e)));  // tests: overflow
  m->tiles_nl = nl;
  m->tiles_stile = small ? 8 : 16;
  m->tiles_count = tiles;
  m->tiles_cap = tcap;
  // snapshot: the check's displacement maxima decide whether the snapshot still covers every
  // particle (|x - x_build| <= margin); the fallback spread runs only when it does not
  const float *disp = snapshot ? nl->disp_partial.p : nullptr;
  const int ndisp = nl->ndisp;
  const float stale2 = margin * margin;
  m->tiles_snapshot = snapshot;
  // the owner form (own particles into tile blocks, then the fold) where the blocks fit and the
  // mesh holds at least one particle per 12 nodes: its block writes and fold cost the same per
  // node at any density, while the pull form's extra work (each particle in every tile its stencil
  // reaches) grows with the particles. Measured on MI355X, graph replays: 0.119 particles per node
  // (486 400 ions, 160^3): owner 128 against pull 172 us; 0.26-0.51 (ESP, 50^3-100^3): owner
  // 20-22% faster evaluations; 0.019-0.037 (64 000 ions on 120^3-150^3, 512 000 on 300^3): pull
  // 4-14% faster. The pull form (every particle reaching a tile, straight into its nodes)
  // otherwise. PDX_SPREAD_OWNER=1 or 0 picks a form (the owner form still only where its blocks fit).
  const int stile = small ? 8 : 16, P = win_width(m->wkey);
  const size_t bvol = (size_t)(stile + P - 1) * (stile + P - 1) * (stile + P - 1);
  const bool fits = (size_t)tiles * bvol <= m->tile_block_entries;
  bool own = fits && (int64_t)m->N * 12 >= (int64_t)m->K1 * m->K2 * m->K3;
  if (const char *e = std::getenv("PDX_SPREAD_OWNER")) own = fits && std::atoi(e) != 0;
  // the owner form's particles: from tile lists that k_tile_bin builds from the current positions,
  // or from a scan of the list's cells. The lists' clearing, binning and fallback launches cost
  // ~15 us whatever the tiles, and repay it only where the tiles fill every CU twice (two 512-thread
  // tiles per CU): measured, 1 000 tiles (486 400 ions, 160^3) 25 us faster, 512 tiles 3 us faster,
  // 343 tiles 6-21 us slower. PDX_SPREAD_LISTS=1 or 0 picks. With the lists, word 3 of the snapshot
  // state is their overflow flag: the fallback spread and the per-particle gather read it as they
  // read a stale snapshot.
  static const int device_cus = [] {
    int dev = 0, c = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0)
      c = 256;
    return c;
  }();
  const bool pairs_fill = tiles >= 2 * device_cus;
  bool lists = own && pairs_fill;
  if (const char *e = std::getenv("PDX_SPREAD_LISTS")) lists = own && std::atoi(e) != 0;
  if (lists) m->tiles_snapshot = true;
#define PDX_TSPREAD_L(PP, TT, BB)                                                                              \
  hipLaunchKernelGGL((k_spread_tiles<PP, TT, BB>), dim3(tiles), dim3(BB), 0, st, nl->packed.p,                 \
                     offsets, ids, guard, nl->nx, nl->ny, nl->nz,                                              \
                     margin, m->grid64.p, m->K1, m->K2, m->K3, L[0], L[1], L[2], scale, m->tile_list.p,   \
                     m->tile_ids.p, m->tile_n.p, tcap, disp, ndisp, stale2, nl->snap_state.p, m->wc);
#define PDX_TSPREAD_O(PP, TT)                                                                                  \
  if (lists) {                                                                                                 \
    hipLaunchKernelGGL(k_tile_prep, dim3((tiles + TPB - 1) / TPB), dim3(TPB), 0, st, m->tile_n.p, tiles,       \
                       nl->snap_state.p);                                                                      \
    hipLaunchKernelGGL((k_tile_bin<PP, TT>), dim3((m->N + TPB - 1) / TPB), dim3(TPB), 0, st, m->N,            \
                       nl->packed.p, ids, guard, m->K1, m->K2, m->K3, L[0], L[1], L[2], m->tile_list.p,      \
                       m->tile_ids.p, m->tile_n.p, tcap, nl->snap_state.p);                                    \
  }                                                                                                            \
  if (lists && own_list_pair<TT + win_width(PP) - 1>() && pairs_fill)                                          \
    hipLaunchKernelGGL((k_spread_own_listed<PP, TT, 512>), dim3(tiles), dim3(512), 0, st, guard, m->K1, m->K2, \
                       m->K3, scale, m->tile_list.p, m->tile_n.p, tcap, nl->snap_state.p, m->wc,                \
                       m->tile_blocks.p);                                                                      \
  else if (lists)                                                                                              \
    hipLaunchKernelGGL((k_spread_own_listed<PP, TT, 1024>), dim3(tiles), dim3(1024), 0, st, guard, m->K1,     \
                       m->K2, m->K3, scale, m->tile_list.p, m->tile_n.p, tcap, nl->snap_state.p, m->wc,         \
                       m->tile_blocks.p);                                                                      \
  else                                                                                                         \
    hipLaunchKernelGGL((k_spread_own<PP, TT, 1024>), dim3(tiles), dim3(1024), 0, st, nl->packed.p,             \
                       offsets, ids, guard, nl->nx, nl->ny, nl->nz,                                            \
                       margin, m->grid64.p, m->K1, m->K2, m->K3, L[0], L[1], L[2], scale, m->tile_list.p, \
                       m->tile_ids.p, m->tile_n.p, tcap, disp, ndisp, stale2, nl->snap_state.p, m->wc,    \
                       m->tile_blocks.p);
#define PDX_TSPREAD(PP)                                                                                        \
  case PP:                                                                                                     \
    if (own) {                                                                                                 \
      if (small) { PDX_TSPREAD_O(PP, 8) }                                                                      \
      else { PDX_TSPREAD_O(PP, 16) }                                                                           \
    } else if (small) {                                                                                        \
      if (sbs == 1024) { PDX_TSPREAD_L(PP, 8, 1024) }                                                          \
      else if (sbs == 512) { PDX_TSPREAD_L(PP, 8, 512) }                                                       \
      else { PDX_TSPREAD_L(PP, 8, 256) }                                                                       \
    } else {                                                                                                   \
      if (sbs == 1024) { PDX_TSPREAD_L(PP, 16, 1024) }                                                         \
      else if (sbs == 512) { PDX_TSPREAD_L(PP, 16, 512) }                                                      \
      else { PDX_TSPREAD_L(PP, 16, 256) }                                                                      \
    }                                                                                                          \
    break;
  switch (m->wkey) {
    PDX_WINDOWS(PDX_TSPREAD)
    default: return fail(PDX_ERR_ARGUMENT, "unsupported order");
  }
#undef PDX_TSPREAD
#undef PDX_TSPREAD_L
#undef PDX_TSPREAD_O
  PDX_LAUNCH_CHECK();
  if (own) {
    const dim3 fg((m->K1 + 63) / 64, (m->K2 + 3) / 4, m->K3), fb(64, 4);
    if (small)
      hipLaunchKernelGGL(k_spread_fold<8>, fg, fb, 0, st, m->tile_blocks.p, m->grid64.p, m->K1, m->K2, m->K3, P, guard);
    else
      hipLaunchKernelGGL(k_spread_fold<16>, fg, fb, 0, st, m->tile_blocks.p, m->grid64.p, m->K1, m->K2, m->K3, P, guard);
    PDX_LAUNCH_CHECK();
  }
  if (snapshot || lists) {
    const int fb = (int)std::min<int64_t>(1024, ((int64_t)m->N * lanes_for(8) + TPB - 1) / TPB);
#define PDX_STALE(PP)                                                                                          \
  case PP:                                                                                                     \
    hipLaunchKernelGGL(k_spread_stale<PP>, dim3(fb), dim3(TPB), 0, st, m->N, nl->packed.p,                    \
                       (const int64_t *)nl->snap_state.p, m->grid64.p, m->K1, m->K2, m->K3, L[0], L[1], L[2], scale, \
                       m->wc);                                                                                 \
    break;
    switch (m->wkey) {
      PDX_WINDOWS(PDX_STALE)
      default: return fail(PDX_ERR_ARGUMENT, "unsupported order");
    }
#undef PDX_STALE
    PDX_LAUNCH_CHECK();
  }
  return PDX_OK;
}
}  // namespace

extern "C" int pdx_mesh_set_spread_threads(pdx_mesh_t m, int32_t threads) {
  if (!m) return fail(PDX_ERR_ARGUMENT, "null mesh");
  if (threads != 256 && threads != 512 && threads != 1024)
    return fail(PDX_ERR_ARGUMENT, "spread threads must be 256, 512 or 1024");
  m->spread_threads = threads;
  return PDX_OK;
}

extern "C" int pdx_mesh_set_transform_threads(pdx_mesh_t m, int32_t threads) {
  if (!m) return fail(PDX_ERR_ARGUMENT, "null mesh");
  if (threads != 512 && threads != 1024) return fail(PDX_ERR_ARGUMENT, "transform threads must be 512 or 1024");
  m->transform_threads = threads;
  return PDX_OK;
}

extern "C" int pdx_mesh_spread_cells(pdx_mesh_t m, pdx_nl_t nl, pdx_stream_t stream) {
  return spread_cells_impl(m, nl, false, stream);
}

extern "C" int pdx_mesh_spread_snapshot(pdx_mesh_t m, pdx_nl_t nl, pdx_stream_t stream) {
  if (nl && nl->snap_host != 2)
    return fail(PDX_ERR_STATE, "no cell snapshot yet: the second list check after creation or a box change takes it");
  return spread_cells_impl(m, nl, true, stream);
}

extern "C" int pdx_mesh_solve(pdx_mesh_t m, int32_t flags, pdx_stream_t stream) {
  if (!m) return fail(PDX_ERR_ARGUMENT, "null mesh");
  hipStream_t st = as_stream(stream);
  const Plans pl = plans_of(m);
  const bool energy = (flags & PDX_ENERGY) != 0;
  if (m->plane_tp) return enqueue_plane_solve(m, pl, energy, st);
  int rc = enqueue_forward_xy(m, pl, m->grid64.p, true, st);
  if (rc) return rc;
  if (m->spec.influence != PDX_INFLUENCE_P3M_IK) {
    if (energy) launch_z<2, true>(m, pl, st);
    else launch_z<2, false>(m, pl, st);
    PDX_LAUNCH_CHECK();
    return enqueue_inverse_yx(m, pl, m->spec_c.p, m->spec_c.p, m->pot.p, 0, st, energy);
  }
  // ik: G b is kept for the z field; E_d = IFFT(-i k_d G b) for d = z, x, y, and the potential
  // (three field transforms, plus the potential when energies are wanted)
  if (energy) launch_z<2, true>(m, pl, st, nullptr, m->spec_ez.p);
  else launch_z<2, false>(m, pl, st, nullptr, m->spec_ez.p);
  launch_z<3, false>(m, pl, st, m->spec_ez.p);
  PDX_LAUNCH_CHECK();
  if ((rc = enqueue_inverse_yx(m, pl, m->spec_ez.p, m->spec_ez.p, m->ez.p, 0, st, energy)) ||
      (rc = enqueue_inverse_yx(m, pl, m->spec_c.p, m->spec_tmp.p, m->ex.p, 1, st)) ||
      (rc = enqueue_inverse_yx(m, pl, m->spec_c.p, m->spec_tmp.p, m->ey.p, 2, st)))
    return rc;
  m->pot_valid = energy;
  if (energy) return enqueue_inverse_yx(m, pl, m->spec_c.p, m->spec_c.p, m->pot.p, 0, st);
  return PDX_OK;
}

extern "C" int pdx_mesh_gather(pdx_mesh_t m, const float *x, const float *y, const float *z, const float *q,
                               const int32_t *order, float *fx, float *fy, float *fz, float *Epot, float *Vglobal,
                               int32_t flags, int32_t set, pdx_stream_t stream) {
  if (!m || !x || !y || !z || !q || !fx || !fy || !fz) return fail(PDX_ERR_ARGUMENT, "null argument");
  const int N = m->N;
  const float L[3] = {(float)m->box[0], (float)m->box[1], (float)m->box[2]};
  const bool energy = (flags & PDX_ENERGY) != 0;
  if (!energy) Vglobal = nullptr;
  hipStream_t st = as_stream(stream);
  auto blocks_of = [&](int P) { return (int)(((int64_t)N * gather_lanes(P) + TPB - 1) / TPB); };
  int xcd = 1;
  if (const char *e = std::getenv("PDX_GATHER_XCD")) xcd = std::atoi(e);
  if (m->spec.influence == PDX_INFLUENCE_P3M_IK) {
    if (energy && !m->pot_valid)
      return fail(PDX_ERR_ARGUMENT, "ik gather with energies needs a solve with PDX_ENERGY first");
    switch (m->P) {
      PDX_GATHER_IK_CASE(3)
      PDX_GATHER_IK_CASE(4)
      PDX_GATHER_IK_CASE(5)
      PDX_GATHER_IK_CASE(6)
      PDX_GATHER_IK_CASE(7)
      PDX_GATHER_IK_CASE(8)
      default: return fail(PDX_ERR_ARGUMENT, "unsupported order");
    }
    PDX_LAUNCH_CHECK();
    return PDX_OK;
  }
  switch (m->wkey) {
    PDX_WINDOWS(PDX_GATHER_CASE)
    default: return fail(PDX_ERR_ARGUMENT, "unsupported order");
  }
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

extern "C" int pdx_mesh_gather_cells(pdx_mesh_t m, pdx_nl_t nl, float *fx, float *fy, float *fz, float *Epot,
                                     float *Vglobal, int32_t flags, int32_t set, pdx_stream_t stream) {
  if (!m || !nl || !fx || !fy || !fz) return fail(PDX_ERR_ARGUMENT, "null argument");
  if (m->spec.influence == PDX_INFLUENCE_P3M_IK)
    return fail(PDX_ERR_ARGUMENT, "the tile gather differentiates the potential (ad); ik meshes use pdx_mesh_gather");
  if (m->tiles_nl != (const void *)nl)
    return fail(PDX_ERR_STATE, "the last spread of this mesh was not an owner-computes spread from this list");
  const int N = m->N;
  const float L[3] = {(float)m->box[0], (float)m->box[1], (float)m->box[2]};
  const bool energy = (flags & PDX_ENERGY) != 0;
  if (!energy) Vglobal = nullptr;
  hipStream_t st = as_stream(stream);
  // workgroups past the tiles: the overflow walk, four particles per thread
  const int extra = (int)std::max<int64_t>(1, std::min<int64_t>(256, ((int64_t)N + 4 * GTB - 1) / (4 * GTB)));
  const dim3 grid(m->tiles_count + extra);
#define PDX_GT_L(PP, TT, EE)                                                                                   \
  hipLaunchKernelGGL((k_gather_tiles<PP, TT, EE>), grid, dim3(GTB), 0, st, N, nl->packed.p, m->tile_list.p,    \
                     m->tile_ids.p, m->tile_n.p, m->tiles_cap, m->tiles_count, m->pot.p, m->K1, m->K2, m->K3, L[0], L[1],     \
                     L[2], fx, fy, fz, Epot, m->bg, set, Vglobal, m->acc.p, m->ubg,                        \
                     m->tiles_snapshot ? (const int64_t *)nl->snap_state.p : nullptr, m->wc);
#define PDX_GT(PP)                                                                                             \
  case PP:                                                                                                     \
    if (m->tiles_stile == 8) {                                                                                 \
      if (energy) { PDX_GT_L(PP, 8, true) } else { PDX_GT_L(PP, 8, false) }                                    \
    } else {                                                                                                   \
      if (energy) { PDX_GT_L(PP, 16, true) } else { PDX_GT_L(PP, 16, false) }                                  \
    }                                                                                                          \
    break;
  switch (m->wkey) {
    PDX_WINDOWS(PDX_GT)
    default: return fail(PDX_ERR_ARGUMENT, "unsupported order");
  }
#undef PDX_GT
#undef PDX_GT_L
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

extern "C" int pdx_mesh_totals(pdx_mesh_t m, double out[9], pdx_stream_t stream) {
  if (!m || !out) return fail(PDX_ERR_ARGUMENT, "null argument");
  PDX_HIP(hipStreamSynchronize(as_stream(stream)));
  double acc[8];
  PDX_HIP(hipMemcpy(acc, m->acc.p, sizeof(acc), hipMemcpyDeviceToHost));
  out[0] = acc[0];
  for (int c = 0; c < 6; ++c) out[1 + c] = acc[1 + c] + (c < 3 ? m->ubg : 0.0);
  out[7] = -0.5 * m->C * m->selfc * m->qsq;  // self_energy (Solver.jl:421-423)
  out[8] = m->ubg;                           // background_energy (Solver.jl:427-429)
  return PDX_OK;
}

extern "C" int pdx_mesh_features(pdx_mesh_t m, const float *x, const float *y, const float *z, const float *q,
                                 float *feats, pdx_stream_t stream) {
  if (!m || !x || !y || !z || !q || !feats) return fail(PDX_ERR_ARGUMENT, "null argument");
  if (m->K1 < 4 || m->K2 < 4 || m->K3 < 4) return fail(PDX_ERR_ARGUMENT, "features need at least 4 mesh nodes a side");
  const float L[3] = {(float)m->box[0], (float)m->box[1], (float)m->box[2]};
  hipLaunchKernelGGL(k_features, dim3((m->N + TPB / WAVE - 1) / (TPB / WAVE)), dim3(TPB), 0, as_stream(stream), m->N, x, y, z, q,
                     m->grid64.p, m->K1, m->K2, m->K3, L[0], L[1], L[2], (float)m->fixed_inv, feats);
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

extern "C" int pdx_mesh_copy_grid(pdx_mesh_t m, int32_t which, void *dst, pdx_stream_t stream) {
  if (!m || !dst) return fail(PDX_ERR_ARGUMENT, "null argument");
  const size_t nn = (size_t)m->K1 * m->K2 * m->K3;
  hipStream_t st = as_stream(stream);
  if (which == 0) {
    hipLaunchKernelGGL(k_fixed_to_double, dim3((nn + TPB - 1) / TPB), dim3(TPB), 0, st, m->grid64.p, (double *)dst, nn,
                       m->fixed_inv);
    PDX_LAUNCH_CHECK();
  } else if (which == 1) {
    PDX_HIP(hipMemcpyAsync(dst, m->pot.p, sizeof(float) * nn, hipMemcpyDeviceToDevice, st));
  } else {
    return fail(PDX_ERR_ARGUMENT, "which must be 0 or 1");
  }
  return PDX_OK;
}

extern "C" int pdx_mesh_fft(pdx_mesh_t m, int32_t dir, const void *in, void *out, pdx_stream_t stream) {
  if (!m || !in || !out) return fail(PDX_ERR_ARGUMENT, "null argument");
  hipStream_t st = as_stream(stream);
  const Plans pl = plans_of(m);
  const size_t nmodes = (size_t)m->K1h * m->K2 * m->K3;
  if (dir == -1) {
    int rc = enqueue_forward_xy(m, pl, in, false, st);
    if (rc) return rc;
    launch_z<0, false>(m, pl, st);
    PDX_LAUNCH_CHECK();
    PDX_HIP(hipMemcpyAsync(out, m->spec_c.p, sizeof(float2) * nmodes, hipMemcpyDeviceToDevice, st));
    return PDX_OK;
  }
  if (dir == 1) {
    PDX_HIP(hipMemcpyAsync(m->spec_c.p, in, sizeof(float2) * nmodes, hipMemcpyDeviceToDevice, st));
    launch_z<1, false>(m, pl, st);
    PDX_LAUNCH_CHECK();
    return enqueue_inverse_yx(m, pl, m->spec_c.p, m->spec_c.p, (float *)out, 0, st);
  }
  return fail(PDX_ERR_ARGUMENT, "dir must be -1 (R2C) or +1 (C2R)");
}
