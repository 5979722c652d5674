// FP64 Ewald summation on the device: the converged reference force and energy that PME
// results (and the learned correction's training targets) are measured against.
// Mirrors the independent oracle test/ewald_reference.jl:50-126 for an orthorhombic box:
// real space over explicit periodic images to rcut = sqrt(digits ln 10)/alpha, reciprocal
// vectors to kmax = 2 alpha sqrt(digits ln 10) (half space, doubled), self energy
// -C alpha/sqrt(pi) sum q^2, optional uniform background -C pi Q^2/(2 alpha^2 V).
#include "pdx_internal.hpp"

#include <cmath>

using namespace pdx;

namespace {

constexpr int TPB = 256;

struct EwaldArgs {
  int n;
  const double *x, *y, *z, *q;
  double L[3];
  double C, alpha, rcut2;
  int nimg[3];
};

// Real-space sum over j (and images) for particle i: force, and energy/virial halves.
__device__ void real_terms(const EwaldArgs &a, int i, int j0, int jstride, double f[3], double ev[7]) {
  const double xi = a.x[i], yi = a.y[i], zi = a.z[i], qi = a.q[i];
  const double two_a_sqrtpi = 2.0 * a.alpha / sqrt(M_PI);
  for (int j = j0; j < a.n; j += jstride) {
    const double qq = a.C * qi * a.q[j];
    if (qq == 0.0) continue;
    double dx0 = xi - a.x[j], dy0 = yi - a.y[j], dz0 = zi - a.z[j];
    dx0 -= a.L[0] * rint(dx0 / a.L[0]);
    dy0 -= a.L[1] * rint(dy0 / a.L[1]);
    dz0 -= a.L[2] * rint(dz0 / a.L[2]);
    for (int n1 = -a.nimg[0]; n1 <= a.nimg[0]; ++n1)
      for (int n2 = -a.nimg[1]; n2 <= a.nimg[1]; ++n2)
        for (int n3 = -a.nimg[2]; n3 <= a.nimg[2]; ++n3) {
          if (i == j && n1 == 0 && n2 == 0 && n3 == 0) continue;
          const double dx = dx0 + n1 * a.L[0], dy = dy0 + n2 * a.L[1], dz = dz0 + n3 * a.L[2];
          const double r2 = dx * dx + dy * dy + dz * dz;
          if (!(r2 < a.rcut2)) continue;
          const double r = sqrt(r2);
          const double ec = erfc(a.alpha * r);
          const double fmag = qq * (ec / r + two_a_sqrtpi * exp(-a.alpha * a.alpha * r2)) / r2;
          f[0] += fmag * dx; f[1] += fmag * dy; f[2] += fmag * dz;
          ev[0] += 0.5 * qq * ec / r;
          ev[1] += 0.5 * fmag * dx * dx; ev[2] += 0.5 * fmag * dy * dy; ev[3] += 0.5 * fmag * dz * dz;
          ev[4] += 0.5 * fmag * dx * dy; ev[5] += 0.5 * fmag * dx * dz; ev[6] += 0.5 * fmag * dy * dz;
        }
  }
}

__device__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int w = threadIdx.x / WAVE;
  __syncthreads();
  if ((threadIdx.x & (WAVE - 1)) == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < TPB / WAVE; ++k) s += sh[k];
  return s;
}

// One block per sampled particle: real-space force.
__global__ __launch_bounds__(TPB) void k_real_force(EwaldArgs a, const int32_t *sample, double *F) {
  __shared__ double sh[TPB / WAVE];
  const int i = sample[blockIdx.x];
  double f[3] = {0, 0, 0}, ev[7] = {0, 0, 0, 0, 0, 0, 0};
  real_terms(a, i, threadIdx.x, TPB, f, ev);
  for (int d = 0; d < 3; ++d) {
    const double s = block_sum(f[d], sh);
    if (threadIdx.x == 0) F[3 * blockIdx.x + d] = s;
  }
}

// One block per particle: real-space energy and virial (halves of full rows).
__global__ __launch_bounds__(TPB) void k_real_energy(EwaldArgs a, double *partial) {
  double f[3] = {0, 0, 0}, ev[7] = {0, 0, 0, 0, 0, 0, 0};
  real_terms(a, blockIdx.x, threadIdx.x, TPB, f, ev);
  block_partials<7>(ev, partial);
}

// Structure factor S(k) = sum_j q_j e^{i k.r_j}, one block per wave vector.
__global__ __launch_bounds__(TPB) void k_structure(EwaldArgs a, const double *kv, double2 *S) {
  __shared__ double sh[TPB / WAVE];
  const double kx = kv[3 * blockIdx.x], ky = kv[3 * blockIdx.x + 1], kz = kv[3 * blockIdx.x + 2];
  double re = 0.0, im = 0.0;
  for (int j = threadIdx.x; j < a.n; j += TPB) {
    double s, c;
    sincos(kx * a.x[j] + ky * a.y[j] + kz * a.z[j], &s, &c);
    re += a.q[j] * c;
    im += a.q[j] * s;
  }
  re = block_sum(re, sh);
  im = block_sum(im, sh);
  if (threadIdx.x == 0) S[blockIdx.x] = make_double2(re, im);
}

// Reciprocal energy and virial over the half space (x2).
__global__ __launch_bounds__(TPB) void k_recip_energy(int nk, const double *kv, const double2 *S, double CoverV,
                                                      double inv4a2, double *partial) {
  const int k = blockIdx.x * TPB + threadIdx.x;
  double ev[7] = {0, 0, 0, 0, 0, 0, 0};
  if (k < nk) {
    const double kx = kv[3 * k], ky = kv[3 * k + 1], kz = kv[3 * k + 2];
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double g = 4.0 * M_PI * exp(-k2 * inv4a2) / k2;
    const double Ek = 2.0 * 0.5 * CoverV * g * (S[k].x * S[k].x + S[k].y * S[k].y);
    const double b = 2.0 / k2 * (1.0 + k2 * inv4a2);
    ev[0] = Ek;
    ev[1] = Ek * (1.0 - b * kx * kx); ev[2] = Ek * (1.0 - b * ky * ky); ev[3] = Ek * (1.0 - b * kz * kz);
    ev[4] = -Ek * b * kx * ky; ev[5] = -Ek * b * kx * kz; ev[6] = -Ek * b * ky * kz;
  }
  block_partials<7>(ev, partial);
}

// Reciprocal force on sampled particles, one block per sample, threads over wave vectors.
__global__ __launch_bounds__(TPB) void k_recip_force(EwaldArgs a, int nk, const double *kv, const double2 *S,
                                                     double CoverV, double inv4a2, const int32_t *sample,
                                                     double *F) {
  __shared__ double sh[TPB / WAVE];
  const int i = sample[blockIdx.x];
  const double xi = a.x[i], yi = a.y[i], zi = a.z[i], qi = a.q[i];
  double f[3] = {0, 0, 0};
  for (int k = threadIdx.x; k < nk; k += TPB) {
    const double kx = kv[3 * k], ky = kv[3 * k + 1], kz = kv[3 * k + 2];
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double g = 4.0 * M_PI * exp(-k2 * inv4a2) / k2;
    double s, c;
    sincos(kx * xi + ky * yi + kz * zi, &s, &c);
    // Im(e^{ik.r} conj(S)) = s S.re - c S.im
    const double im = s * S[k].x - c * S[k].y;
    const double fm = 2.0 * CoverV * g * qi * im;
    f[0] += fm * kx; f[1] += fm * ky; f[2] += fm * kz;
  }
  for (int d = 0; d < 3; ++d) {
    const double s = block_sum(f[d], sh);
    if (threadIdx.x == 0) F[3 * blockIdx.x + d] += s;
  }
}

}  // namespace

extern "C" int pdx_ewald_fp64(int32_t n, const double *x, const double *y, const double *z, const double *q,
                              const double box[3], double C, double alpha, double digits, int32_t background,
                              int32_t nsample, const int32_t *sample, double *F, int32_t want_energy, double out[7],
                              pdx_stream_t stream) {
  if (n <= 0 || !x || !y || !z || !q || !box) return fail(PDX_ERR_ARGUMENT, "invalid argument");
  if (nsample > 0 && (!sample || !F)) return fail(PDX_ERR_ARGUMENT, "sample and F are required when nsample > 0");
  if (!(alpha > 0) || !(digits > 0) || !(C > 0)) return fail(PDX_ERR_ARGUMENT, "alpha, digits and C must be positive");
  hipStream_t st = as_stream(stream);
  EwaldArgs a{};
  a.n = n; a.x = x; a.y = y; a.z = z; a.q = q;
  for (int d = 0; d < 3; ++d) a.L[d] = box[d];
  a.C = C; a.alpha = alpha;
  const double xr = std::sqrt(digits * std::log(10.0));
  const double rcut = xr / alpha, kmax = 2.0 * alpha * xr;
  a.rcut2 = rcut * rcut;
  for (int d = 0; d < 3; ++d) a.nimg[d] = std::max(0, (int)std::ceil(rcut / box[d] + 0.5) - 1);
  // half-space wave vectors within kmax
  std::vector<double> kv;
  int mmax[3];
  for (int d = 0; d < 3; ++d) mmax[d] = (int)std::ceil(kmax * box[d] / (2 * M_PI)) + 1;
  for (int m1 = 0; m1 <= mmax[0]; ++m1)
    for (int m2 = -mmax[1]; m2 <= mmax[1]; ++m2)
      for (int m3 = -mmax[2]; m3 <= mmax[2]; ++m3) {
        if (m1 == 0 && (m2 < 0 || (m2 == 0 && m3 <= 0))) continue;
        const double kx = 2 * M_PI * m1 / box[0], ky = 2 * M_PI * m2 / box[1], kz = 2 * M_PI * m3 / box[2];
        if (kx * kx + ky * ky + kz * kz < kmax * kmax) { kv.push_back(kx); kv.push_back(ky); kv.push_back(kz); }
      }
  const int nk = (int)(kv.size() / 3);
  const double V = box[0] * box[1] * box[2];
  const double CoverV = C / V, inv4a2 = 1.0 / (4 * alpha * alpha);
  // a synchronous boundary call (the error judge, not an enqueue of the force path): its
  // workspace lives for the call
  DevBuf<double> dkv, part, acc;
  DevBuf<double2> S;
  int rc;
  const int ekblocks = std::max(1, (nk + TPB - 1) / TPB);
  if ((rc = dkv.alloc(std::max<size_t>(kv.size(), 3))) || (rc = S.alloc(std::max(nk, 1))) ||
      (rc = part.alloc((size_t)7 * std::max(n, ekblocks))) || (rc = acc.alloc(16)))
    return rc;
  PDX_HIP(hipMemsetAsync(acc.p, 0, sizeof(double) * 16, st));
  if (nk > 0) PDX_HIP(hipMemcpyAsync(dkv.p, kv.data(), sizeof(double) * kv.size(), hipMemcpyHostToDevice, st));
  if (nk > 0) hipLaunchKernelGGL(k_structure, dim3(nk), dim3(TPB), 0, st, a, dkv.p, S.p);
  if (nsample > 0) {
    hipLaunchKernelGGL(k_real_force, dim3(nsample), dim3(TPB), 0, st, a, sample, F);
    if (nk > 0)
      hipLaunchKernelGGL(k_recip_force, dim3(nsample), dim3(TPB), 0, st, a, nk, dkv.p, S.p, CoverV, inv4a2, sample, F);
  }
  if (want_energy) {
    // per-block FP64 partials, then one block sums them in block order (no device-scope tickets)
    hipLaunchKernelGGL(k_real_energy, dim3(n), dim3(TPB), 0, st, a, part.p);
    hipLaunchKernelGGL(k_sum_partials<7>, dim3(1), dim3(256), 0, st, (const double *)part.p, n, acc.p, 0,
                       (const int64_t *)nullptr);
    if (nk > 0) {
      hipLaunchKernelGGL(k_recip_energy, dim3(ekblocks), dim3(TPB), 0, st, nk, dkv.p, S.p, CoverV, inv4a2, part.p);
      hipLaunchKernelGGL(k_sum_partials<7>, dim3(1), dim3(256), 0, st, (const double *)part.p, ekblocks, acc.p + 8, 0,
                         (const int64_t *)nullptr);
    }
  }
  PDX_LAUNCH_CHECK();
  PDX_HIP(hipStreamSynchronize(st));
  if (want_energy && out) {
    double h[16];
    PDX_HIP(hipMemcpy(h, acc.p, sizeof(h), hipMemcpyDeviceToHost));
    double Q = 0, qsq = 0;
    std::vector<double> qh(n);
    PDX_HIP(hipMemcpy(qh.data(), q, sizeof(double) * n, hipMemcpyDeviceToHost));
    for (double v : qh) { Q += v; qsq += v * v; }
    for (int c = 0; c < 7; ++c) out[c] = h[c] + h[8 + c];
    out[0] -= C * alpha / std::sqrt(M_PI) * qsq;
    if (background) {
      const double Ebg = -C * M_PI * Q * Q / (2 * alpha * alpha * V);
      out[0] += Ebg;
      out[1] += Ebg; out[2] += Ebg; out[3] += Ebg;
    }
  }
  return PDX_OK;
}
