// Transform-kit generation on the GPU in FP64.
//
// Replaces tfm%init (src/submodules/mlegs_spectfm_init.f90:6-154): Gauss-
// Legendre nodes/weights (:181-214), log-norms (:218-250), the fully
// normalised associated Legendre table (:254-300) and at0/at1 (:147-150);
// plus the band-operator rows of src/submodules/mlegs_spectfm_diff.f90 and
// the V/D/T projection tables that vector_projection rebuilds on every call
// (src/submodules/mlegs_scalar_ops.f90:1387-1411) -- here built once.
//
// The table uses the normalised three-term recurrence in FP64 rather than the
// reference's 50-digit recurrence on unnormalised values followed by an FP64
// log-space renormalisation; both give the orthonormal P_n^m with the
// Condon-Shortley phase, and they agree to ~1e-13 relative (the precision of
// the reference's own log-space normalisation).
#include <cmath>
#include <cstring>

#include "mlx_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr double kPi = 3.14159265358979323846;

// one thread per root i < ceil((n+1)/2); Newton on P_n, tolerance 1e-15.
__global__ void k_gauleg(int n, double *z_out, double *pp_out) {
  int m = (n + 2) / 2;  // ceiling((n+1)/2) for integer n
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double z = cos(kPi * ((i + 1) - 0.25) / (n + 0.5));
  double z1 = z + 1.0;
  double pp = 0.0;
  while (fabs(z - z1) > 1e-15) {
    double p1 = 1.0, p2 = 0.0;
    for (int j = 1; j <= n; ++j) {
      double p3 = p2;
      p2 = p1;
      p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
    }
    pp = n * (z * p1 - p2) / (z * z - 1);
    z1 = z;
    z = z1 - p1 / pp;
  }
  z_out[i] = z;
  pp_out[i] = pp;
}

// sequential placement of the roots as the reference's loop writes them
// (the last root overwrites the middle pair for even n).
__global__ void k_gauleg_place(int n, const double *z, const double *pp, double *x, double *w) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int m = (n + 2) / 2;
  for (int i = 0; i < m; ++i) {
    double zi = z[i];
    double wi = 2.0 / ((1 - zi * zi) * pp[i] * pp[i]);
    x[i] = -zi;
    x[n - 1 - i] = zi;
    w[i] = wi;
    w[n - 1 - i] = wi;
  }
}

__global__ void k_lnx_r(int nr, double ell, const double *x, double *lnx, double *r) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  lnx[i] = -log(1.0 - x[i]);
  r[i] = ell * sqrt((1.0 + x[i]) / (1.0 - x[i]));
}

// lognorm[m][nn] = 0.5*log((2n+1)/2 (n-m)!/(n+m)!), n = m + nn
// (mlegs_spectfm_init.f90:218-250), one thread per m.
__global__ void k_lognorm(int ne, int npchop, double *out) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= npchop) return;
  double wk = log(0.5);
  for (int k = 1; k <= m; ++k)
    wk = wk + log(2.0 * k + 1.0) - log(2.0 * k * ((2.0 * k - 1.0) * (2.0 * k - 1.0)));
  double v = wk;
  out[(size_t)m * ne + 0] = 0.5 * v;
  for (int nn = 1; nn < ne; ++nn) {
    int n = m + nn;
    v = v + log(2.0 * n + 1.0) - log(2.0 * n - 1.0) + log(1.0 * (n - m)) - log(1.0 * (n + m));
    out[(size_t)m * ne + nn] = 0.5 * v;
  }
}

// Normalised P_n^m at x_i, n = m .. m+ne-1 (one thread per (i, m)).
__device__ void legendre_column(double x, int m, int ne, double *dst, size_t stride) {
  double s = sqrt((1.0 - x) * (1.0 + x));
  double p0 = 1.0 / sqrt(2.0);
  for (int k = 1; k <= m; ++k) p0 = -sqrt((2.0 * k + 1.0) / (2.0 * k)) * s * p0;
  dst[0] = p0;
  if (ne == 1) return;
  double p1 = sqrt(2.0 * m + 3.0) * x * p0;
  dst[stride] = p1;
  for (int nn = 2; nn < ne; ++nn) {
    int n = m + nn;
    double a = sqrt((4.0 * n * n - 1.0) / ((double)n * n - (double)m * m));
    double b = sqrt(((double)(n - 1) * (n - 1) - (double)m * m) / (4.0 * (n - 1) * (n - 1) - 1.0));
    double p2 = a * (x * p1 - b * p0);
    dst[(size_t)nn * stride] = p2;
    p0 = p1;
    p1 = p2;
  }
}

__global__ void k_legtbl(int nrh, int ne, int npchop, const double *x, double *pf) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int m = blockIdx.y;
  if (i >= nrh || m >= npchop) return;
  legendre_column(x[i], m, ne, pf + (size_t)m * ne * nrh + i, (size_t)nrh);
}

__global__ void k_at01(int nrchop, double *at0, double *at1) {
  if (threadIdx.x == 0) legendre_column(-1.0 + 1e-15, 0, nrchop, at0, 1);
  if (threadIdx.x == 1) legendre_column(1.0 - 1e-15, 0, nrchop, at1, 1);
}

// Band rows of leg_del2 (without the -ak^2 on the diagonal) and leg_xxdx in
// the normalised basis: entry (i, j) * exp(lognorm(j) - lognorm(i))
// (mlegs_spectfm_diff.f90:6-43, 98-152).
__global__ void k_bands(int ne, int npchop, double ell, const double *lognorm, double *d2c, double *xxc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int am = blockIdx.y;
  if (i >= ne || am >= npchop) return;
  const double *ln = lognorm + (size_t)am * ne;
  double n = (double)(am + i);
  double ell2 = ell * ell;
  double g[5];
  g[0] = -(n - am - 1.0) * (n - am) * (n - 2.0) * (n - 1.0) / (2.0 * n - 3.0) / (2.0 * n - 1.0);
  g[1] = 2.0 * n * (n - am) * (n - 1.0) / (2.0 * n - 1.0);
  g[2] = (-2.0 * n * (n + 1.0) * (3.0 * n * n + 3.0 * n - am * am - 2.0)) / (2.0 * n - 1.0) / (2.0 * n + 3.0);
  g[3] = 2.0 * (n + 1.0) * (n + am + 1.0) * (n + 2.0) / (2.0 * n + 3.0);
  g[4] = -(n + am + 1.0) * (n + am + 2.0) * (n + 3.0) * (n + 2.0) / (2.0 * n + 3.0) / (2.0 * n + 5.0);
  double *d = d2c + ((size_t)am * ne + i) * 5;
  for (int q = 0; q < 5; ++q) {
    int j = i + q - 2;
    double v = g[q] / ell2;
    if (j < 0 || j >= ne) v = 0.0;
    else if (q != 2) v = v * exp(ln[j] - ln[i]);
    d[q] = v;
  }
  double xs = -(n - 1.0) * (n - am) / (2.0 * n - 1.0);
  double xp = (n + 2.0) * (n + am + 1.0) / (2.0 * n + 3.0);
  double *xr = xxc + ((size_t)am * ne + i) * 2;
  xr[0] = (i >= 1) ? xs * exp(ln[i - 1] - ln[i]) : 0.0;
  xr[1] = (i + 1 < ne) ? xp * exp(ln[i + 1] - ln[i]) : 0.0;
}

// V, D, T tables of vector_projection (mlegs_scalar_ops.f90:1387-1411):
//   pfd(:, i) = sum_k pf(:, k) X(k, i)  with X = leg_xxdx (tridiagonal),
//   V(i,x) = pf(x,i) / (n(n+1)) / fff(x),  D(i,x) = pfd(x,i) / (n(n+1)) / fff(x),
//   T(i,x) = pf(x,i) w(x),  fff = (1-x^2)/w,  n = max(1, m+i).
__global__ void k_v2t_tables(int nrh, int ne, int npchop, int nt, const double *x, const double *w,
                             const double *pf, const double *xxc, double *tV, double *tD, double *tT) {
  int xi = blockIdx.x * blockDim.x + threadIdx.x;
  int i = blockIdx.y;
  int m = blockIdx.z;
  if (xi >= nrh || i >= nt || m >= npchop) return;
  const double *P = pf + (size_t)m * ne * nrh;
  const double *X = xxc + (size_t)m * ne * 2;
  double pfi = P[(size_t)i * nrh + xi];
  // X(i-1, i) = superdiagonal of row i-1; X(i+1, i) = subdiagonal of row i+1
  double pfd = 0.0;
  if (i >= 1) pfd = pfd + P[(size_t)(i - 1) * nrh + xi] * X[(size_t)(i - 1) * 2 + 1];
  if (i + 1 < ne) pfd = pfd + P[(size_t)(i + 1) * nrh + xi] * X[(size_t)(i + 1) * 2 + 0];
  double xv = x[xi], wv = w[xi];
  double fff = (1.0 - xv * xv) / wv;
  int nv = m + i;
  if (nv < 1) nv = 1;
  size_t o = ((size_t)m * nt + i) * nrh + xi;
  tV[o] = pfi / nv / (nv + 1) / fff;
  tD[o] = pfd / nv / (nv + 1) / fff;
  tT[o] = pfi * wv;
}

__global__ void k_to_f32(size_t n, const double *a, float *b) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

FftPlan make_plan(int n) {
  FftPlan p;
  std::memset(&p, 0, sizeof(p));
  p.n = n;
  int r = n;
  while (r > 1 && r % 4 == 0) { p.radix[p.npass++] = 4; r /= 4; }
  while (r > 1 && r % 2 == 0) { p.radix[p.npass++] = 2; r /= 2; }
  while (r > 1 && r % 3 == 0) { p.radix[p.npass++] = 3; r /= 3; }
  while (r > 1 && r % 5 == 0) { p.radix[p.npass++] = 5; r /= 5; }
  if (r != 1) MLX_THROW(MLX_E_ARG, "fft length must only have factors of 2, 3 and 5");
  return p;
}

double2 *upload_twiddles(int n, int count, double denom) {
  // exp(-2 pi i t / denom), t < count, evaluated in extended precision
  std::vector<double2> h(count > 0 ? count : 1);
  for (int t = 0; t < count; ++t) {
    long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)denom;
    h[t].x = (double)cosl(a);
    h[t].y = (double)sinl(a);
  }
  double2 *d = nullptr;
  HIPCHK(hipMalloc(&d, sizeof(double2) * h.size()));
  HIPCHK(hipMemcpy(d, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice));
  (void)n;
  return d;
}

template <class T>
T *dalloc(size_t n) {
  T *p = nullptr;
  if (n == 0) n = 1;
  if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) MLX_THROW(MLX_E_ALLOC, "kit: device allocation failed");
  return p;
}

}  // namespace

namespace mlx {

void kit_free(KitDev &k) {
  void *ptrs[] = {k.x, k.w, k.lnx, k.r, k.ak, k.lognorm, k.pf, k.pf32, k.at0, k.at1, k.d2c,
                  k.xxc, k.tV, k.tD, k.tT, k.tV32, k.tD32, k.tT32, k.tw_nph, k.tw_nz, k.tw_half,
                  k.pfg32, k.tVg32, k.tDg32, k.tTg32, k.pfg64, k.tVg64, k.tDg64, k.tTg64};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  k.x = k.w = k.lnx = k.r = k.ak = k.lognorm = k.pf = k.at0 = k.at1 = k.d2c = k.xxc = nullptr;
  k.tV = k.tD = k.tT = nullptr;
  k.pf32 = k.tV32 = k.tD32 = k.tT32 = nullptr;
  k.pfg32 = k.tVg32 = k.tDg32 = k.tTg32 = nullptr;
  k.pfg64 = k.tVg64 = k.tDg64 = k.tTg64 = nullptr;
  k.mfma_ap = 0;
  k.tw_nph = k.tw_nz = k.tw_half = nullptr;
  k.ready = false;
}

void kit_build(mlx_ctx *c) {
  KitDev &k = c->kit;
  const mlx_params &p = c->p;
  hipStream_t s = c->stream;
  kit_free(k);
  k.nr = p.nr; k.np = p.np; k.nz = p.nz;
  k.nrchop = p.nrchop; k.npchop = p.npchop; k.nzchop = p.nzchop;
  k.hyperpow = p.hyperpow;
  k.ell = p.ell; k.zlen = p.zlen;
  k.nrh = p.nr / 2;
  k.nrdim_ref = p.nr + (p.hyperpow > 3 ? p.hyperpow : 3);
  k.nrdim = (k.nrdim_ref + 7) / 8 * 8;
  k.npdim = p.np / 2 + 1;
  k.nzdim = p.nz;
  k.nph = p.np / 2;
  k.ne = p.nrchop + 14;
  k.chopzl = p.nzchop;
  k.chopzu = p.nz - p.nzchop + 2;
  k.nt = p.nrchop + 3;

  k.x = dalloc<double>(k.nr);
  k.w = dalloc<double>(k.nr);
  k.lnx = dalloc<double>(k.nr);
  k.r = dalloc<double>(k.nr);
  k.ak = dalloc<double>(k.nz);
  k.lognorm = dalloc<double>((size_t)k.ne * k.npchop);
  k.pf = dalloc<double>((size_t)k.nrh * k.ne * k.npchop);
  k.pf32 = dalloc<float>((size_t)k.nrh * k.ne * k.npchop);
  k.at0 = dalloc<double>(k.nrchop);
  k.at1 = dalloc<double>(k.nrchop);
  k.d2c = dalloc<double>((size_t)k.ne * k.npchop * 5);
  k.xxc = dalloc<double>((size_t)k.ne * k.npchop * 2);
  size_t ntab = (size_t)k.nt * k.nrh * k.npchop;
  k.tV = dalloc<double>(ntab);
  k.tD = dalloc<double>(ntab);
  k.tT = dalloc<double>(ntab);
  k.tV32 = dalloc<float>(ntab);
  k.tD32 = dalloc<float>(ntab);
  k.tT32 = dalloc<float>(ntab);

  // Gauss-Legendre nodes and weights
  int m = (k.nr + 2) / 2;
  double *zt = dalloc<double>(m), *ppt = dalloc<double>(m);
  hipLaunchKernelGGL(k_gauleg, dim3((m + 63) / 64), dim3(64), 0, s, k.nr, zt, ppt);
  hipLaunchKernelGGL(k_gauleg_place, dim3(1), dim3(1), 0, s, k.nr, zt, ppt, k.x, k.w);
  hipLaunchKernelGGL(k_lnx_r, dim3((k.nr + 255) / 256), dim3(256), 0, s, k.nr, k.ell, k.x, k.lnx, k.r);
  // log-norms, table, at0/at1, band rows, projection tables
  hipLaunchKernelGGL(k_lognorm, dim3((k.npchop + 63) / 64), dim3(64), 0, s, k.ne, k.npchop, k.lognorm);
  hipLaunchKernelGGL(k_legtbl, dim3((k.nrh + 63) / 64, k.npchop), dim3(64), 0, s, k.nrh, k.ne, k.npchop,
                     k.x, k.pf);
  hipLaunchKernelGGL(k_at01, dim3(1), dim3(64), 0, s, k.nrchop, k.at0, k.at1);
  hipLaunchKernelGGL(k_bands, dim3((k.ne + 63) / 64, k.npchop), dim3(64), 0, s, k.ne, k.npchop, k.ell,
                     k.lognorm, k.d2c, k.xxc);
  hipLaunchKernelGGL(k_v2t_tables, dim3((k.nrh + 63) / 64, k.nt, k.npchop), dim3(64), 0, s, k.nrh, k.ne,
                     k.npchop, k.nt, k.x, k.w, k.pf, k.xxc, k.tV, k.tD, k.tT);
  size_t npf = (size_t)k.nrh * k.ne * k.npchop;
  hipLaunchKernelGGL(k_to_f32, dim3((npf + 255) / 256), dim3(256), 0, s, npf, k.pf, k.pf32);
  hipLaunchKernelGGL(k_to_f32, dim3((ntab + 255) / 256), dim3(256), 0, s, ntab, k.tV, k.tV32);
  hipLaunchKernelGGL(k_to_f32, dim3((ntab + 255) / 256), dim3(256), 0, s, ntab, k.tD, k.tD32);
  hipLaunchKernelGGL(k_to_f32, dim3((ntab + 255) / 256), dim3(256), 0, s, ntab, k.tT, k.tT32);
  HIPCHK(hipGetLastError());

  // axial wavenumbers (mlegs_spectfm_init.f90:132-134)
  k.ak_h.assign(k.nz, 0.0);
  for (int i = 0; i < k.nz; ++i) k.ak_h[i] = 2.0 * kPi / k.zlen * (i - k.nz);
  for (int i = 0; i <= k.nz / 2 && i < k.nz; ++i) k.ak_h[i] = 2.0 * kPi / k.zlen * i;
  HIPCHK(hipMemcpyAsync(k.ak, k.ak_h.data(), sizeof(double) * k.nz, hipMemcpyHostToDevice, s));

  // FFT plans and twiddles
  if (k.np > 1) {
    k.plan_nph = make_plan(k.nph);
    k.tw_nph = upload_twiddles(k.nph, k.nph, k.nph);
    k.tw_half = upload_twiddles(k.np, k.nph + 1, k.np);
  }
  if (k.nz > 1) {
    k.plan_nz = make_plan(k.nz);
    k.tw_nz = upload_twiddles(k.nz, k.nz, k.nz);
  }

  mfma_tables(c);
  HIPCHK(hipStreamSynchronize(s));
  (void)hipFree(zt);
  (void)hipFree(ppt);

  k.lognorm_h.resize((size_t)k.ne * k.npchop);
  HIPCHK(hipMemcpy(k.lognorm_h.data(), k.lognorm, sizeof(double) * k.lognorm_h.size(), hipMemcpyDeviceToHost));
  k.at1_h.resize(k.nrchop);
  k.at0_h.resize(k.nrchop);
  HIPCHK(hipMemcpy(k.at1_h.data(), k.at1, sizeof(double) * k.nrchop, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(k.at0_h.data(), k.at0, sizeof(double) * k.nrchop, hipMemcpyDeviceToHost));
  k.x_h.resize(k.nr);
  k.w_h.resize(k.nr);
  HIPCHK(hipMemcpy(k.x_h.data(), k.x, sizeof(double) * k.nr, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(k.w_h.data(), k.w, sizeof(double) * k.nr, hipMemcpyDeviceToHost));
  k.ready = true;
}

}  // namespace mlx
