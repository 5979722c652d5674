// FP64 Ewald summation on the device: the converged reference force and energy that PME
// results (and the learned correction's training targets) are measured against.
// Mirrors the independent oracle test/ewald_reference.jl:50-126 for an orthorhombic box:
// real space over explicit periodic images to rcut = sqrt(digits ln 10)/alpha, reciprocal
// vectors to kmax = 2 alpha sqrt(digits ln 10) (half space, doubled), self energy
// -C alpha/sqrt(pi) sum q^2, optional uniform background -C pi Q^2/(2 alpha^2 V).
//
// Wave vectors are not listed: every kernel walks the box of half-space indices
// m1 in [0, M1], m2 in [-M2, M2], m3 in [-M3, M3] and skips the vectors outside the sphere,
// so the only device memory beyond the caller's arrays is the structure factors (one per box
// index), the energy passes' block partials (at most EBLK blocks each) and 32 result words.
// pdx_ewald_fp64 allocates them per call; pdx_ewald_fp64_noalloc takes them from a static
// device array of fixed size. Both run the same kernels with the same launch shapes, so they
// return the same bits.
#include "pdx_internal.hpp"

#include <cmath>
#include <mutex>

using namespace pdx;

namespace {

constexpr int TPB = 256;
constexpr int EBLK = 1024;            // blocks of each grid-stride energy pass
constexpr int NBOX_MAX = 1 << 18;     // largest wave-vector box of the static-scratch variant
constexpr int NACC = 32;              // result words: real [0,7), reciprocal [8,15), Q and sum q^2 [16,18)
constexpr size_t SCRATCH_DOUBLES = 2 * (size_t)NBOX_MAX + 7 * (size_t)EBLK + NACC;

// The static-scratch variant's workspace (4.1 MiB per device): structure factors | partials | acc.
// Every call on a device shares it; calls that overlap on one device race.
__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

struct EwaldArgs {
  int n;
  const double *x, *y, *z, *q;
  double L[3];
  double C, alpha, rcut2, kmax2;
  int nimg[3];
  int M[3];  // half-space index box
  int nbox;  // (M1 + 1)(2 M2 + 1)(2 M3 + 1)
};

// Wave vector of box index b; false outside the half space or the sphere |k| < kmax.
__device__ __forceinline__ bool wave_vector(const EwaldArgs &a, int b, double &kx, double &ky, double &kz) {
  const int n3 = 2 * a.M[2] + 1, n2 = 2 * a.M[1] + 1;
  const int m3 = b % n3 - a.M[2];
  const int t = b / n3;
  const int m2 = t % n2 - a.M[1];
  const int m1 = t / n2;
  if (m1 == 0 && (m2 < 0 || (m2 == 0 && m3 <= 0))) return false;
  kx = 2 * M_PI * m1 / a.L[0];
  ky = 2 * M_PI * m2 / a.L[1];
  kz = 2 * M_PI * m3 / a.L[2];
  return kx * kx + ky * ky + kz * kz < a.kmax2;
}

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

// Real-space energy and virial (halves of full rows): block b takes particles b, b + grid, ...
__global__ __launch_bounds__(TPB) void k_real_energy(EwaldArgs a, double *partial) {
  double f[3] = {0, 0, 0}, ev[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int i = blockIdx.x; i < a.n; i += gridDim.x) real_terms(a, i, threadIdx.x, TPB, f, ev);
  block_partials<7>(ev, partial);
}

// Total charge and sum of squares (self and background energies), grid-stride partials.
__global__ __launch_bounds__(TPB) void k_charge_sums(EwaldArgs a, double *partial) {
  double v[2] = {0, 0};
  for (int i = blockIdx.x * TPB + threadIdx.x; i < a.n; i += gridDim.x * TPB) {
    v[0] += a.q[i];
    v[1] += a.q[i] * a.q[i];
  }
  block_partials<2>(v, partial);
}

// Structure factor S(k) = sum_j q_j e^{i k.r_j}, one block per box index (zero outside the sphere).
__global__ __launch_bounds__(TPB) void k_structure(EwaldArgs a, double2 *S) {
  __shared__ double sh[TPB / WAVE];
  double kx, ky, kz;
  if (!wave_vector(a, blockIdx.x, kx, ky, kz)) {
    if (threadIdx.x == 0) S[blockIdx.x] = make_double2(0.0, 0.0);
    return;
  }
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

// Reciprocal energy and virial over the half space (x2), grid-stride over the box.
__global__ __launch_bounds__(TPB) void k_recip_energy(EwaldArgs a, const double2 *S, double CoverV, double inv4a2,
                                                      double *partial) {
  double ev[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int b = blockIdx.x * TPB + threadIdx.x; b < a.nbox; b += gridDim.x * TPB) {
    double kx, ky, kz;
    if (!wave_vector(a, b, kx, ky, kz)) continue;
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double g = 4.0 * M_PI * exp(-k2 * inv4a2) / k2;
    const double Ek = 2.0 * 0.5 * CoverV * g * (S[b].x * S[b].x + S[b].y * S[b].y);
    const double c = 2.0 / k2 * (1.0 + k2 * inv4a2);
    ev[0] += Ek;
    ev[1] += Ek * (1.0 - c * kx * kx); ev[2] += Ek * (1.0 - c * ky * ky); ev[3] += Ek * (1.0 - c * kz * kz);
    ev[4] -= Ek * c * kx * ky; ev[5] -= Ek * c * kx * kz; ev[6] -= Ek * c * ky * kz;
  }
  block_partials<7>(ev, partial);
}

// Reciprocal force on sampled particles, one block per sample, threads over the box.
__global__ __launch_bounds__(TPB) void k_recip_force(EwaldArgs a, const double2 *S, double CoverV, double inv4a2,
                                                     const int32_t *sample, double *F) {
  __shared__ double sh[TPB / WAVE];
  const int i = sample[blockIdx.x];
  const double xi = a.x[i], yi = a.y[i], zi = a.z[i], qi = a.q[i];
  double f[3] = {0, 0, 0};
  for (int b = threadIdx.x; b < a.nbox; b += TPB) {
    double kx, ky, kz;
    if (!wave_vector(a, b, kx, ky, kz)) continue;
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double g = 4.0 * M_PI * exp(-k2 * inv4a2) / k2;
    double s, c;
    sincos(kx * xi + ky * yi + kz * zi, &s, &c);
    // Im(e^{ik.r} conj(S)) = s S.re - c S.im
    const double im = s * S[b].x - c * S[b].y;
    const double fm = 2.0 * CoverV * g * qi * im;
    f[0] += fm * kx; f[1] += fm * ky; f[2] += fm * kz;
  }
  for (int d = 0; d < 3; ++d) {
    const double s = block_sum(f[d], sh);
    if (threadIdx.x == 0) F[3 * blockIdx.x + d] += s;
  }
}

// Arguments, cutoffs and the wave-vector box of one call (host).
int ewald_setup(int32_t n, const double *x, const double *y, const double *z, const double *q, const double box[3],
                double C, double alpha, double digits, int32_t nsample, const int32_t *sample, double *F,
                EwaldArgs &a) {
  if (n <= 0 || !x || !y || !z || !q || !box) return fail(PDX_ERR_ARGUMENT, "invalid argument");
  if (nsample < 0 || (nsample > 0 && (!sample || !F)))
    return fail(PDX_ERR_ARGUMENT, "sample and F are required when nsample > 0");
  if (!(alpha > 0) || !(digits > 0) || !(C > 0)) return fail(PDX_ERR_ARGUMENT, "alpha, digits and C must be positive");
  for (int d = 0; d < 3; ++d)
    if (!(box[d] > 0) || !std::isfinite(box[d])) return fail(PDX_ERR_ARGUMENT, "box lengths must be finite and positive");
  a = EwaldArgs{};
  a.n = n; a.x = x; a.y = y; a.z = z; a.q = q;
  for (int d = 0; d < 3; ++d) a.L[d] = box[d];
  a.C = C; a.alpha = alpha;
  const double xr = std::sqrt(digits * std::log(10.0));
  const double rcut = xr / alpha, kmax = 2.0 * alpha * xr;
  a.rcut2 = rcut * rcut;
  a.kmax2 = kmax * kmax;
  long long nbox = 1;
  for (int d = 0; d < 3; ++d) {
    a.nimg[d] = std::max(0, (int)std::ceil(rcut / box[d] + 0.5) - 1);
    const double m = std::ceil(kmax * box[d] / (2 * M_PI)) + 1;
    if (!(m < 1e6)) return fail(PDX_ERR_ARGUMENT, "too many wave vectors (alpha * box too large)");
    a.M[d] = (int)m;
    nbox *= d == 0 ? a.M[0] + 1 : 2 * a.M[d] + 1;
  }
  if (nbox > INT32_MAX / 2) return fail(PDX_ERR_ARGUMENT, "too many wave vectors (alpha * box too large)");
  a.nbox = (int)nbox;
  return PDX_OK;
}

// Enqueue the passes on st. S: nbox structure factors; part: 7 EBLK partials; acc: NACC words.
int ewald_enqueue(const EwaldArgs &a, int32_t nsample, const int32_t *sample, double *F, int32_t want_energy,
                  double2 *S, double *part, double *acc, hipStream_t st) {
  const double V = a.L[0] * a.L[1] * a.L[2];
  const double CoverV = a.C / V, inv4a2 = 1.0 / (4 * a.alpha * a.alpha);
  hipLaunchKernelGGL(k_structure, dim3(a.nbox), dim3(TPB), 0, st, a, S);
  if (nsample > 0) {
    hipLaunchKernelGGL(k_real_force, dim3(nsample), dim3(TPB), 0, st, a, sample, F);
    hipLaunchKernelGGL(k_recip_force, dim3(nsample), dim3(TPB), 0, st, a, (const double2 *)S, CoverV, inv4a2, sample,
                       F);
  }
  if (want_energy) {
    // per-block FP64 partials, then one block sums them in block order (no device-scope tickets)
    const int rb = std::min(a.n, EBLK);
    hipLaunchKernelGGL(k_real_energy, dim3(rb), dim3(TPB), 0, st, a, part);
    hipLaunchKernelGGL(k_sum_partials<7>, dim3(1), dim3(SUM_THREADS), 0, st, (const double *)part, rb, acc, 0,
                       (const int64_t *)nullptr);
    const int kb = std::min((a.nbox + TPB - 1) / TPB, EBLK);
    hipLaunchKernelGGL(k_recip_energy, dim3(kb), dim3(TPB), 0, st, a, (const double2 *)S, CoverV, inv4a2, part);
    hipLaunchKernelGGL(k_sum_partials<7>, dim3(1), dim3(SUM_THREADS), 0, st, (const double *)part, kb, acc + 8, 0,
                       (const int64_t *)nullptr);
    const int qb = std::min((a.n + TPB - 1) / TPB, EBLK);
    hipLaunchKernelGGL(k_charge_sums, dim3(qb), dim3(TPB), 0, st, a, part);
    hipLaunchKernelGGL(k_sum_partials<2>, dim3(1), dim3(SUM_THREADS), 0, st, (const double *)part, qb, acc + 16, 0,
                       (const int64_t *)nullptr);
  }
  PDX_LAUNCH_CHECK();
  return PDX_OK;
}

// Wait for the passes and form E and W on the host: self energy and background from the
// device charge sums.
int ewald_finish(const EwaldArgs &a, int32_t background, int32_t want_energy, const double *acc, double out[7],
                 hipStream_t st) {
  PDX_HIP(hipStreamSynchronize(st));
  if (!(want_energy && out)) return PDX_OK;
  double h[18];
  PDX_HIP(hipMemcpy(h, acc, sizeof(h), hipMemcpyDeviceToHost));
  const double Q = h[16], qsq = h[17];
  for (int c = 0; c < 7; ++c) out[c] = h[c] + h[8 + c];
  out[0] -= a.C * a.alpha / std::sqrt(M_PI) * qsq;
  if (background) {
    const double V = a.L[0] * a.L[1] * a.L[2];
    const double Ebg = -a.C * M_PI * Q * Q / (2 * a.alpha * a.alpha * V);
    out[0] += Ebg;
    out[1] += Ebg; out[2] += Ebg; out[3] += Ebg;
  }
  return PDX_OK;
}

// Address of `scratch` on each device, taken once per device with that device current.
constexpr int MAXDEV = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[MAXDEV] = {nullptr};

int scratch_on_current_device(double **out) {
  int dev = 0;
  PDX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAXDEV) return fail(PDX_ERR_DEVICE, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    PDX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(PDX_ERR_DEVICE, "hipGetSymbolAddress returned no address for the Ewald scratch");
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return PDX_OK;
}

}  // namespace

extern "C" int pdx_ewald_fp64(int32_t n, const double *x, const double *y, const double *z, const double *q,
                              const double box[3], double C, double alpha, double digits, int32_t background,
                              int32_t nsample, const int32_t *sample, double *F, int32_t want_energy, double out[7],
                              pdx_stream_t stream) {
  EwaldArgs a;
  int rc;
  if ((rc = ewald_setup(n, x, y, z, q, box, C, alpha, digits, nsample, sample, F, a))) return rc;
  hipStream_t st = as_stream(stream);
  // a synchronous boundary call (the error judge, not an enqueue of the force path): its
  // workspace lives for the call
  DevBuf<double2> S;
  DevBuf<double> part, acc;
  if ((rc = S.alloc(a.nbox)) || (rc = part.alloc((size_t)7 * EBLK)) || (rc = acc.alloc(NACC))) return rc;
  if ((rc = ewald_enqueue(a, nsample, sample, F, want_energy, S.p, part.p, acc.p, st))) return rc;
  return ewald_finish(a, background, want_energy, acc.p, out, st);
}

extern "C" int pdx_ewald_fp64_noalloc(int32_t n, const double *x, const double *y, const double *z, const double *q,
                                      const double box[3], double C, double alpha, double digits,
                                      int32_t background, int32_t nsample, const int32_t *sample, double *F,
                                      int32_t want_energy, double out[7], pdx_stream_t stream) {
  EwaldArgs a;
  int rc;
  if ((rc = ewald_setup(n, x, y, z, q, box, C, alpha, digits, nsample, sample, F, a))) return rc;
  if (a.nbox > NBOX_MAX)
    return fail(PDX_ERR_ARGUMENT, "pdx_ewald_fp64_noalloc: the wave-vector box holds " + std::to_string(a.nbox) +
                                      " indices, more than the static scratch's " + std::to_string(NBOX_MAX) +
                                      " (lower alpha * box or digits, or call pdx_ewald_fp64)");
  double *base = nullptr;
  if ((rc = scratch_on_current_device(&base))) return rc;
  hipStream_t st = as_stream(stream);
  double2 *S = reinterpret_cast<double2 *>(base);
  double *part = base + 2 * (size_t)NBOX_MAX;
  double *acc = part + 7 * (size_t)EBLK;
  if ((rc = ewald_enqueue(a, nsample, sample, F, want_energy, S, part, acc, st))) return rc;
  return ewald_finish(a, background, want_energy, acc, out, st);
}
