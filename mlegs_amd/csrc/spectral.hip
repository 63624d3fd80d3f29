// Spectral-space and pointwise passes of the time step (FP64, HBM-bound).
//
//   chop, dealias, svv_filter, calcat/zeroat1, delsqp/idelsqp
//       (src/submodules/mlegs_scalar_ops.f90:6-155, 237-416)
//   banded matvec: xxdx, del2h, del2  (:418-573) -- one wavefront per
//       (m, k) column staged in LDS, band rows precomputed at kit build
//   banded solve: ihelm, idel2        (:575-854) -- LU with partial pivoting
//       (LAPACK gbtf2/gbtrs order of operations) factorised once per alpha
//       and cached; one thread per (m, k) system
//   steppers' linear combinations (febe/abcn, :1157-1262), vecprod
//       (:1264-1306), the tp2vec/vec2tp pointwise pieces (:1337-1373,
//       :1488-1527), check_stability and the FP64 energy reduction.
//
// z-slice lists follow the reference's two kk-loops: slices k < nzchop and
// k >= nz - nzchop + 1 (0-based); when nzchop = nz/2 + 1 the slice
// k = nzchop - 1 belongs to both loops and the reference applies the
// operator to it twice -- reproduced here by an apply count of 2.
#include <cmath>

#include "mlx_internal.h"

namespace {

constexpr int kT = 256;

__device__ inline bool z_active(int k, int nzc, int nzcu) { return (k + 1 <= nzc) || (k + 1 >= nzcu); }
__device__ inline int z_count(int k, int nzc, int nzcu) {
  return ((k + 1 <= nzc) ? 1 : 0) + ((k + 1 >= nzcu) ? 1 : 0);
}
__device__ inline int nrcs_dev(int nrc, int npc, int m) {
  if (m >= npc) return 0;
  int v = nrc - m;
  return v > 0 ? v : 0;
}

__device__ inline double block_sum(double v, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) / 64;
    for (int i = 0; i < nw; ++i) r += sh[i];
  }
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
__global__ void k_ln_column_(double2 *e, const double *ln, const double *lnx, double sign, int nr, int nrdim,
                             int npdim, int nz) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  if (i >= nr || k >= nz) return;
  double v = sign * ln[0] * lnx[i];
  double2 *p = e + (size_t)k * nrdim * npdim + i;
  p->x += v;
}

__global__ void k_chop_(double2 *e, int nrdim, int npdim, int nz, int fr, int fp, int fz, int nrc, int npc,
                        int nzc, int nzcu) {
  size_t n = (size_t)nrdim * npdim * nz;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(t % nrdim);
    int j = (int)((t / nrdim) % npdim);
    int k = (int)(t / ((size_t)nrdim * npdim));
    bool z = false;
    if (fr && i >= nrcs_dev(nrc, npc, j)) z = true;
    if (fp && j >= npc) z = true;
    if (fz && nzc < nzcu && k >= nzc && k <= nzcu - 2) z = true;
    if (z) e[t] = make_double2(0.0, 0.0);
  }
}

__global__ void k_dealias_(double2 *e, int nrdim, int npdim, int nz, int pcut, int zcut, int zupper) {
  size_t n = (size_t)nrdim * npdim * nz;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    int j = (int)((t / nrdim) % npdim);
    int k = (int)(t / ((size_t)nrdim * npdim));
    bool z = (pcut >= 0 && j >= pcut) || (zcut >= 0 && k >= zcut && k <= zupper - 2);
    if (z) e[t] = make_double2(0.0, 0.0);
  }
}

// ---- SVV ------------------------------------------------------------------
__device__ inline double svv_q(int i, int j, int k, double inv_r, double inv_p, const double *ak, double kmax) {
  double qr = fmin(1.0, (double)(i > 0 ? i : 0) * inv_r);
  double qp = fmin(1.0, (double)j * inv_p);
  double qz = fmin(1.0, fabs(ak[k]) / kmax);
  return fmin(1.0, fmax(fmax(qr, qp), qz));
}

// SVV on one or two fields (blockIdx.z / blockIdx.y selects the field).
// Reduction: workgroup (bx, k) sums |c|^2 and the tail (q >= cutoff) over a
// range of columns j < pcut of slice k; slices zlo..zhi are skipped (the
// caller has just de-aliased the field: they are zero and add nothing).
// flag (optional): set to 1 when an entry is not finite, the stability check
// of the time loop taken on the same pass.  k_svv_gain then adds the partial
// sums in a fixed order and updates each field's gain and strength
// (mlegs_scalar_ops.f90:72-126): out[3 f] = strength, out[3 f + 1] = total,
// out[3 f + 2] = tail.
struct SvvFields {
  double2 *e[2];
  double *gain[2];
};

__global__ void k_svv_reduce(SvvFields F, int nrdim, int npdim, double inv_r, double inv_p, const double *ak,
                             double kmax, double cutoff, double *part, int *flag, int pcut, int zlo, int zhi) {
  __shared__ double sh[16];
  const int f = blockIdx.z, k = blockIdx.y, nbx = gridDim.x;
  double tot = 0.0, tail = 0.0;
  int bad = 0;
  if (k < zlo || k > zhi) {
    // q >= cutoff (cutoff < 1) when any of the row, column and slice ratios
    // reaches the cutoff
    const bool tz = fmin(1.0, fabs(ak[k]) / kmax) >= cutoff;
    const int ncol = pcut < npdim ? pcut : npdim;
    const int per = (ncol + nbx - 1) / nbx, j0 = blockIdx.x * per;
    const int j1 = j0 + per < ncol ? j0 + per : ncol;
    // columns j0..j1-1 are contiguous: item t is row t % nrdim of column
    // j0 + t / nrdim
    const double2 *ep = F.e[f] + (size_t)k * nrdim * npdim + (size_t)j0 * nrdim;
    const int cnt = j1 > j0 ? (j1 - j0) * nrdim : 0;
    constexpr int U = 4;
    for (int t0 = threadIdx.x; t0 < cnt; t0 += U * blockDim.x) {
      double2 v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int t = t0 + u * blockDim.x;
        v[u] = t < cnt ? ld_nt(&ep[t]) : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int t = t0 + u * blockDim.x;
        const int jj = t / nrdim, i = t - jj * nrdim;
        if (!isfinite(v[u].x) || !isfinite(v[u].y)) bad = 1;
        const double a2 = v[u].x * v[u].x + v[u].y * v[u].y;
        tot += a2;
        const bool tq = tz || fmin(1.0, (double)i * inv_r) >= cutoff || fmin(1.0, (double)(j0 + jj) * inv_p) >= cutoff;
        if (t < cnt && tq) tail += a2;
      }
    }
  }
  if (flag && bad) atomicOr(flag, 1);
  const double a = block_sum(tot, sh);
  const double b = block_sum(tail, sh);
  if (threadIdx.x == 0) {
    const int nb = nbx * gridDim.y;
    double *pf = part + (size_t)f * 2 * nb;
    const int slot = k * nbx + blockIdx.x;
    pf[2 * slot] = a;
    pf[2 * slot + 1] = b;
  }
}

// one workgroup: the partial sums of field blockIdx.x, its gain and strength
__global__ void k_svv_gain(SvvFields F, const double *part, int nb, double *out, double target, double relax,
                           double strength0) {
  __shared__ double sh[16];
  const int g = blockIdx.x;
  const double *pg = part + (size_t)g * 2 * nb;
  double x = 0.0, y = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    x += pg[2 * i];
    y += pg[2 * i + 1];
  }
  const double T = block_sum(x, sh), L = block_sum(y, sh);
  if (threadIdx.x != 0) return;
  double *o = out + 3 * g;
  o[1] = T;
  o[2] = L;
  if (T <= 2.2250738585072014e-308) {
    o[0] = 0.0;
    return;
  }
  const double feedback = fmin(fmax((L / T) / target - 1.0, 0.0), 1.0);
  const double gn = (1.0 - relax) * F.gain[g][0] + relax * feedback;
  F.gain[g][0] = gn;
  o[0] = fmin(strength0 * gn, 1.0);
}

__global__ void k_svv_apply(SvvFields F, int nrdim, int npdim, int nz, double inv_r, double inv_p,
                            const double *ak, double kmax, double cutoff, const double *out) {
  const int f = blockIdx.y;
  const double st = out[3 * f];
  if (!(st > 0.0)) return;
  double2 *e = F.e[f];
  size_t n = (size_t)nrdim * npdim * nz;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(t % nrdim);
    int j = (int)((t / nrdim) % npdim);
    int k = (int)(t / ((size_t)nrdim * npdim));
    double q = svv_q(i, j, k, inv_r, inv_p, ak, kmax);
    if (q > cutoff) {
      double s = (q - cutoff) / (1.0 - cutoff);
      double s2 = s * s, s4 = s2 * s2;
      double f2 = exp(-st * (s4 * s4));
      double2 v = e[t];
      e[t] = make_double2(f2 * v.x, f2 * v.y);
    }
  }
}

// ---- calcat / zeroat1 -----------------------------------------------------
__global__ void k_calcat_(const double2 *e, const double *tbl, int n, int nrdim, int npdim, double2 *out) {
  __shared__ double sh[16];
  int k = blockIdx.x;
  const double2 *c = e + (size_t)k * nrdim * npdim;
  double sr = 0.0, si = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    sr += c[i].x * tbl[i];
    si += c[i].y * tbl[i];
  }
  double a = block_sum(sr, sh);
  double b = block_sum(si, sh);
  if (threadIdx.x == 0) out[k] = make_double2(a, b);
}

// zeroat1 on one or two fields (blockIdx.y selects the field); lnz[f]
// (optional): the field's log coefficient, set to zero by the same launch
struct Zat1Fields {
  double2 *e[2];
  int n[2];
  double *lnz[2];
};
__global__ void k_zeroat1_(Zat1Fields F, const double *at1, int nrdim, int npdim) {
  __shared__ double sh[16];
  const int k = blockIdx.x, f = blockIdx.y, n = F.n[f];
  double2 *c = F.e[f] + (size_t)k * nrdim * npdim;
  double sr = 0.0, si = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    sr += c[i].x * at1[i];
    si += c[i].y * at1[i];
  }
  double a = block_sum(sr, sh);
  double b = block_sum(si, sh);
  if (threadIdx.x == 0) {
    double d = at1[0];
    c[0] = make_double2(c[0].x - a / d, c[0].y - b / d);
    if (k == 0 && F.lnz[f]) F.lnz[f][0] = 0.0;
  }
}

// ---- delsqp / idelsqp -------------------------------------------------------
__global__ void k_delsqp_(double2 *e, double *ln, const double *lognorm, int nrdim, int npdim, int nz, int nrc,
                          int npc, double ell, int inverse) {
  size_t n = (size_t)nrdim * npdim * nz;
  double ell2 = ell * ell;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(t % nrdim);
    int j = (int)((t / nrdim) % npdim);
    int k = (int)(t / ((size_t)nrdim * npdim));
    if (j >= npc || i >= nrcs_dev(nrc, npc, j)) continue;
    double nv = (double)(j + i);
    double2 v = e[t];
    if (!inverse) {
      double2 o = make_double2(-v.x * nv * (nv + 1.0) / ell2, -v.y * nv * (nv + 1.0) / ell2);
      if (t == 0) {
        o = make_double2(ln[0] / ell2 / exp(lognorm[0]), 0.0);
        ln[0] = 0.0;
      }
      e[t] = o;
    } else {
      if (t == 0) ln[0] = v.x * ell2 * exp(lognorm[0]);
      double2 o;
      if (j + i == 0 || (i == 0 && j == 0)) o = make_double2(0.0, 0.0);
      else o = make_double2(-v.x / nv / (nv + 1.0) * ell2, -v.y / nv / (nv + 1.0) * ell2);
      e[t] = o;
    }
  }
}

__global__ void k_zero_00k(double2 *e, int nrdim, int npdim, int nz) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nz) e[(size_t)k * nrdim * npdim] = make_double2(0.0, 0.0);
}

// ---- banded matvec ----------------------------------------------------------
// op 0: del2 (5 bands, diag -= ak^2), 1: del2h (5 bands), 2: xxdx (2 bands)
__global__ void __launch_bounds__(64) k_bandmul_(double2 *e, double *ln, const double *d2c, const double *xxc,
                                                 const double *lognorm, const double *ak, int op, int nrdim,
                                                 int npdim, int nz, int ne, int nrc, int npc, int nzc, int nzcu,
                                                 double ell) {
  extern __shared__ double2 col[];
  const int m = blockIdx.x, k = blockIdx.y;
  const int cnt = z_count(k, nzc, nzcu);
  const int nn = nrcs_dev(nrc, npc, m);
  double2 *g = e + (size_t)k * nrdim * npdim + (size_t)nrdim * m;
  if (cnt > 0 && nn > 0) {
    for (int i = threadIdx.x; i < nn; i += 64) col[i] = g[i];
    __syncthreads();
    const double akk = (op == 0) ? ak[k] * ak[k] : 0.0;
    for (int rep = 0; rep < cnt; ++rep) {
      double2 res[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int i = threadIdx.x + 64 * q;
        if (i >= nn) continue;
        double2 s = make_double2(0.0, 0.0);
        if (op < 2) {
          const double *b = d2c + ((size_t)m * ne + i) * 5;
#pragma unroll
          for (int d = 0; d < 5; ++d) {
            int j = i + d - 2;
            if (j < 0 || j >= nn) continue;
            double c = b[d];
            if (d == 2) c = c - akk;
            s.x += c * col[j].x;
            s.y += c * col[j].y;
          }
        } else {
          const double *b = xxc + ((size_t)m * ne + i) * 2;
          if (i >= 1) { s.x += b[0] * col[i - 1].x; s.y += b[0] * col[i - 1].y; }
          if (i + 1 < nn) { s.x += b[1] * col[i + 1].x; s.y += b[1] * col[i + 1].y; }
        }
        res[q] = s;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int i = threadIdx.x + 64 * q;
        if (i < nn) col[i] = res[q];
      }
      __syncthreads();
    }
    for (int i = threadIdx.x; i < nn; i += 64) g[i] = col[i];
  }
  __syncthreads();
  if (m == 0 && k == 0 && threadIdx.x == 0) {
    double l = ln[0];
    if (op < 2) {
      double ell2 = ell * ell;
      g[0].x += 4.0 / 3.0 / ell2 / exp(lognorm[0]) * l;
      g[1].x -= 2.0 / 1.0 / ell2 / exp(lognorm[1]) * l;
      g[2].x += 2.0 / 3.0 / ell2 / exp(lognorm[2]) * l;
    } else {
      g[0].x += 1.0 / exp(lognorm[0]) * l;
      g[1].x += 1.0 / exp(lognorm[1]) * l;
    }
    ln[0] = 0.0;
  }
}

// ---- helmholtz / poisson solves -----------------------------------------------
__global__ void k_helm_pre_(double2 *e, double *ln_out, const double *ln_in, const double *lognorm, double alpha,
                            double ell) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double l = ln_in[0] / alpha;
  ln_out[0] = l;
  double ell2 = ell * ell;
  e[0].x -= 4.0 / 3.0 / ell2 / exp(lognorm[0]) * l;
  e[1].x += 2.0 / 1.0 / ell2 / exp(lognorm[1]) * l;
  e[2].x -= 2.0 / 3.0 / ell2 / exp(lognorm[2]) * l;
}

__global__ void k_idel2_post_(const double2 *e, double *ln, const double *lognorm) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ln[0] = e[0].x * exp(lognorm[0]);
}

// idel2 with a prescribed log coefficient, (m,k)=(0,0) system only
// (mlegs_scalar_ops.f90:603-625): rows shifted down, first row pins the
// log coefficient; banded LU with partial pivoting on one thread.
__global__ void k_idel2_preln00_(double2 *e, double *ln, const double *d2c, const double *lognorm, int ne, int n,
                                 double preln, double ell, double *work) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  // The system: row 0 prescribes the log coefficient, rows 1.. are the
  // Laplacian's rows 0..n-2 (bandwidths 2, 2) shifted down one, and column 0
  // carries the log term's Laplacian: lower bandwidth 3, upper 1.  Banded LU
  // with partial pivoting (kl = 3, ku = 2 as the reference's gbsv call), band
  // storage with fill-in rows: element (i, j) at ab[KV + i - j + LD j].
  constexpr int KLP = 3, KUP = 2, KVP = KLP + KUP, LD = 2 * KLP + KUP + 1;
  double *ab = work;
  double *br = work + (size_t)LD * n;
  double *bi = br + n;
  double ell2 = ell * ell;
  for (size_t t = 0; t < (size_t)LD * n; ++t) ab[t] = 0.0;
#define AB(i, j) ab[(KVP + (i) - (j)) + (size_t)LD * (j)]
  for (int i = 0; i < n - 1; ++i)
    for (int d = 0; d < 5; ++d) {
      int j = i + d - 2;
      if (j < 0 || j >= n) continue;
      AB(i + 1, j) = d2c[((size_t)0 * ne + i) * 5 + d];
    }
  AB(0, 0) = 1.0;
  if (n > 1) AB(1, 0) += 4.0 / 3.0 / ell2;
  if (n > 2) AB(2, 0) -= 2.0 / 1.0 / ell2 * exp(lognorm[0] - lognorm[1]);
  if (n > 3) AB(3, 0) += 2.0 / 3.0 / ell2 * exp(lognorm[0] - lognorm[2]);
  for (int i = n - 1; i >= 1; --i) { br[i] = e[i - 1].x; bi[i] = e[i - 1].y; }
  br[0] = preln / exp(lognorm[0]);
  bi[0] = 0.0;
  // factor and forward-eliminate the right-hand side together
  int ju = 0;
  for (int j = 0; j < n; ++j) {
    const int km = (KLP < n - 1 - j) ? KLP : n - 1 - j;
    int p = 0;
    double best = fabs(AB(j, j));
    for (int r = 1; r <= km; ++r)
      if (fabs(AB(j + r, j)) > best) { best = fabs(AB(j + r, j)); p = r; }
    int jn = j + KUP + p;
    if (jn > n - 1) jn = n - 1;
    if (jn > ju) ju = jn;
    if (p != 0) {
      for (int c = j; c <= ju; ++c) { double t = AB(j, c); AB(j, c) = AB(j + p, c); AB(j + p, c) = t; }
      double t = br[j]; br[j] = br[j + p]; br[j + p] = t;
      t = bi[j]; bi[j] = bi[j + p]; bi[j + p] = t;
    }
    const double rp = 1.0 / AB(j, j);
    for (int r = 1; r <= km; ++r) {
      const double l = AB(j + r, j) * rp;
      if (l == 0.0) continue;
      for (int c = j + 1; c <= ju; ++c) AB(j + r, c) -= l * AB(j, c);
      br[j + r] -= l * br[j];
      bi[j + r] -= l * bi[j];
    }
  }
  // back substitution, U has KL + KU superdiagonals
  for (int j = n - 1; j >= 0; --j) {
    double d = AB(j, j);
    br[j] /= d;
    bi[j] /= d;
    for (int i = (j - KVP > 0 ? j - KVP : 0); i < j; ++i) {
      br[i] -= br[j] * AB(i, j);
      bi[i] -= bi[j] * AB(i, j);
    }
  }
#undef AB
  for (int i = 0; i < n; ++i) e[i] = make_double2(br[i], bi[i]);
  ln[0] = br[0] * exp(lognorm[0]);
}

// ---- linear combinations -------------------------------------------------------
__global__ void k_axpby_(double2 *d, double *dln, double a, const double2 *s, const double *sln, double b, size_t n) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    double2 x = d[t], y = s[t];
    d[t] = make_double2(a * x.x + b * y.x, a * x.y + b * y.y);
  }
  if (dln && sln && blockIdx.x == 0 && threadIdx.x == 0) dln[0] = a * dln[0] + b * sln[0];
}

// mode 1 (abcn, mlegs_scalar_ops.f90:1214-1240):
//   sh = a * ((s + dt*(1.5 nl - 0.5 nlp)) + dt/2 * (visc*d2s)),  sh.ln = a*(s.ln + dt*(1.5 nl.ln - 0.5 nlp.ln))
// mode 0 (febe, :1169-1180): sh = a * (s + dt*nl), sh.ln = a*(s.ln + dt*nl.ln)
__global__ void k_abcn_rhs_(double2 *sh, double *shln, const double2 *s, const double *sln, const double2 *nl,
                            const double *nlln, const double2 *nlp, const double *nlpln, const double2 *d2s,
                            double dt, double a, double visc, int mode, size_t n) {
  double hdt = dt / 2.0;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    double2 x = s[t], y = nl[t];
    double2 o;
    if (mode == 1) {
      double2 z = nlp[t], v = d2s[t];
      double tx = x.x + dt * (1.5 * y.x - 0.5 * z.x);
      double ty = x.y + dt * (1.5 * y.y - 0.5 * z.y);
      double vx = v.x * visc, vy = v.y * visc;
      o = make_double2(a * (tx + hdt * vx), a * (ty + hdt * vy));
    } else {
      o = make_double2(a * (x.x + dt * y.x), a * (x.y + dt * y.y));
    }
    sh[t] = o;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double l;
    if (mode == 1) l = sln[0] + dt * (1.5 * nlln[0] - 0.5 * nlpln[0]);
    else l = sln[0] + dt * nlln[0];
    shln[0] = a * l;
  }
}

__device__ inline void band_pow(double2 *y, const double2 *x, double2 *tmp, int n, int cnt, int op, int m, int ne,
                                const double *d2c, const double *xxc, double akk);

// AB2/CN right-hand side with the viscous term formed in place, one (m, kappa)
// column per workgroup (abcn, mlegs_scalar_ops.f90:1218-1240):
//   sh = a (s + dt (1.5 nl - 0.5 nl_p) + dt/2 visc svis),  svis = del2(s)
// where del2 acts on the rows of s's radial offset (applied as often as the
// kk-loops visit the slice, log-term correction of (0,0) after) and leaves
// every other entry equal to s, as the reference's copy-then-del2 does.
// pre (nlln, nlpln given): k_abcn_ln_ followed by ihelm's log-term step
// (k_helm_pre_ with alpha = a), in the (0,0) workgroup: ln <- a (ln + dt
// (1.5 nl.ln - 0.5 nl_p.ln)) / a, and rows 0..2 of the (0,0) column lose the
// log term's Laplacian.  No other workgroup reads ln.
__global__ void __launch_bounds__(256) k_abcn_rhs_d2_(double2 *e, double *sln, const double2 *nl,
                                                      const double2 *nlp, const double *d2c, const double *lognorm,
                                                      const double *ak, int nrdim, int npdim, int ne, int nrc,
                                                      int npc, int nzc, int nzcu, double ell, double dt, double a,
                                                      double visc, int pcut, int zlo, int zhi, const double *nlln,
                                                      const double *nlpln, int act) {
  extern __shared__ double2 lds[];
  double2 *S = lds, *V = lds + nrdim, *T = lds + 2 * nrdim;
  const int m = blockIdx.x, k = blockIdx.y;
  // columns m >= pcut and slices zlo..zhi: left as they are (the caller
  // de-aliases the result)
  if (m >= pcut || (k >= zlo && k <= zhi)) return;
  const size_t off = (size_t)k * nrdim * npdim + (size_t)nrdim * m;
  const int cnt = z_count(k, nzc, nzcu);
  const int nn = (cnt > 0) ? nrcs_dev(nrc, npc, m) : 0;
  // act: s, N and N_p are chopped (rows >= nn zero), so the right-hand side
  // is zero there too and only rows < nn are read and written (the band
  // product itself reads rows < nn only)
  const int nrow = act ? nn : nrdim;
  for (int i = threadIdx.x; i < nrow; i += blockDim.x) S[i] = e[off + i];
  __syncthreads();
  if (nn > 0) {
    band_pow(V, S, T, nn, cnt, 0, m, ne, d2c, nullptr, ak[k] * ak[k]);
    if (m == 0 && k == 0 && threadIdx.x == 0) {
      const double l = sln[0], ell2 = ell * ell;
      V[0].x += 4.0 / 3.0 / ell2 / exp(lognorm[0]) * l;
      if (nn > 1) V[1].x -= 2.0 / 1.0 / ell2 / exp(lognorm[1]) * l;
      if (nn > 2) V[2].x += 2.0 / 3.0 / ell2 / exp(lognorm[2]) * l;
    }
    __syncthreads();
  }
  const double hdt = dt / 2.0;
  const bool pre = nlln && m == 0 && k == 0;
  double l = 0.0;
  if (pre) {
    const double sh = a * (sln[0] + dt * (1.5 * nlln[0] - 0.5 * nlpln[0]));
    l = sh / a;
  }
  const double ell2 = ell * ell;
  for (int i = threadIdx.x; i < nrow; i += blockDim.x) {
    const double2 x = S[i], y = ld_nt(&nl[off + i]), z = ld_nt(&nlp[off + i]);
    const double2 v = (i < nn) ? V[i] : x;
    const double tx = x.x + dt * (1.5 * y.x - 0.5 * z.x);
    const double ty = x.y + dt * (1.5 * y.y - 0.5 * z.y);
    const double vx = v.x * visc, vy = v.y * visc;
    double2 o = make_double2(a * (tx + hdt * vx), a * (ty + hdt * vy));
    if (pre) {
      if (i == 0) o.x -= 4.0 / 3.0 / ell2 / exp(lognorm[0]) * l;
      if (i == 1) o.x += 2.0 / 1.0 / ell2 / exp(lognorm[1]) * l;
      if (i == 2) o.x -= 2.0 / 3.0 / ell2 / exp(lognorm[2]) * l;
    }
    e[off + i] = o;
  }
  if (pre) {
    __syncthreads();  // every read of sln above is done
    if (threadIdx.x == 0) sln[0] = l;
  }
}

// The hyperviscous AB2/CN right-hand side (abcn's HYPERPOW branch,
// mlegs_scalar_ops.f90:1241-1255), one (m, kappa) column per workgroup:
//   sh = a (s + dt (1.5 nl - 0.5 nl_p) + dt/2 g svis),
//   svis = helmp(s, p, 0, b) = b del2(s) + del2^(p/2)(s)
// with every del2 applied as often as the kk-loops visit the slice and the
// first one taking the (0,0) log-term correction (helmp's bandmul steps);
// rows at or past the system's rows keep s in both products, as bandmul
// leaves them.  act: s, N and N_p are chopped, so only rows < nn are read
// and written (the rest are zero on both sides).  The log coefficient is
// k_abcn_ln_'s, launched after.
__global__ void __launch_bounds__(256) k_abcn_rhs_hyp_(double2 *e, const double *sln, const double2 *nl,
                                                       const double2 *nlp, const double *d2c, const double *lognorm,
                                                       const double *ak, int nrdim, int npdim, int ne, int nrc,
                                                       int npc, int nzc, int nzcu, double ell, double dt, double a,
                                                       double g, double b, int half, int pcut, int zlo, int zhi,
                                                       int act) {
  extern __shared__ double2 lds[];
  double2 *S = lds, *V = lds + nrdim, *W = lds + 2 * nrdim, *T = lds + 3 * nrdim;
  const int m = blockIdx.x, k = blockIdx.y;
  // columns m >= pcut and slices zlo..zhi: left as they are (the caller
  // de-aliases the result)
  if (m >= pcut || (k >= zlo && k <= zhi)) return;
  const size_t off = (size_t)k * nrdim * npdim + (size_t)nrdim * m;
  const int cnt = z_count(k, nzc, nzcu);
  const int nn = (cnt > 0) ? nrcs_dev(nrc, npc, m) : 0;
  const int nrow = act ? nn : nrdim;
  for (int i = threadIdx.x; i < nrow; i += blockDim.x) S[i] = e[off + i];
  __syncthreads();
  if (nn > 0) {
    const double akk = ak[k] * ak[k];
    // V = del2(s) (the (0,0) log term after), W = del2^(half - 1)(V)
    band_pow(V, S, T, nn, cnt, 0, m, ne, d2c, nullptr, akk);
    if (m == 0 && k == 0 && threadIdx.x == 0) {
      const double l = sln[0], ell2 = ell * ell;
      V[0].x += 4.0 / 3.0 / ell2 / exp(lognorm[0]) * l;
      if (nn > 1) V[1].x -= 2.0 / 1.0 / ell2 / exp(lognorm[1]) * l;
      if (nn > 2) V[2].x += 2.0 / 3.0 / ell2 / exp(lognorm[2]) * l;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nn; i += blockDim.x) W[i] = V[i];
    __syncthreads();
    for (int st = 1; st < half; ++st) band_pow(W, W, T, nn, cnt, 0, m, ne, d2c, nullptr, akk);
  }
  const double hdt = dt / 2.0;
  for (int i = threadIdx.x; i < nrow; i += blockDim.x) {
    const double2 x = S[i], y = ld_nt(&nl[off + i]), z = ld_nt(&nlp[off + i]);
    const double2 v = (i < nn) ? V[i] : x, w = (i < nn) ? W[i] : x;
    // helmp's two axpby steps: 0 s + b del2(s), then + del2^(p/2)(s)
    const double2 sv = make_double2((0.0 * x.x + b * v.x) + w.x, (0.0 * x.y + b * v.y) + w.y);
    const double tx = x.x + dt * (1.5 * y.x - 0.5 * z.x);
    const double ty = x.y + dt * (1.5 * y.y - 0.5 * z.y);
    const double vx = sv.x * g, vy = sv.y * g;
    e[off + i] = make_double2(a * (tx + hdt * vx), a * (ty + hdt * vy));
  }
}

// sh.ln = a (s.ln + dt (1.5 nl.ln - 0.5 nl_p.ln))  (svis has no log term)
__global__ void k_abcn_ln_(double *shln, const double *sln, const double *nlln, const double *nlpln, double dt,
                           double a) {
  shln[0] = a * (sln[0] + dt * (1.5 * nlln[0] - 0.5 * nlpln[0]));
}

__global__ void k_scale_rows_(double2 *e, const double *r, int nr, int nrdim, int npdim, int nz, int inverse) {
  size_t n = (size_t)nrdim * npdim * nz;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(t % nrdim);
    if (i >= nr) continue;
    double2 v = e[t];
    double rr = r[i];
    e[t] = inverse ? make_double2(v.x / rr, v.y / rr) : make_double2(v.x * rr, v.y * rr);
  }
}

// ur <- i m psi + i kv ur ; up <- -up - m kv uz  (mlegs_scalar_ops.f90:1488-1502)
__global__ void k_tp2vec_combine_(double2 *ur, double2 *up, const double2 *psi, const double2 *uz, const double *ak,
                                  int nrdim, int npdim, int nz, int nrc, int npc, int nzc, int nzcu) {
  size_t n = (size_t)nrdim * npdim * nz;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(t % nrdim);
    int j = (int)((t / nrdim) % npdim);
    int k = (int)(t / ((size_t)nrdim * npdim));
    if (j >= npc || i >= nrcs_dev(nrc, npc, j) || !z_active(k, nzc, nzcu)) continue;
    double mv = (double)j, kv = ak[k];
    double2 p = psi[t], r = ur[t], u = up[t], z = uz[t];
    ur[t] = make_double2(-mv * p.y - kv * r.y, mv * p.x + kv * r.x);
    up[t] = make_double2(-u.x - mv * kv * z.x, -u.y - mv * kv * z.y);
  }
}

// ---------------------------------------------------------------------------
// Front half of tp2vec / tp2curlvec in one pass per (m, kappa) column
// (vector_reconstruction, mlegs_scalar_ops.f90:1455-1545; curl_vector_
// reconstruction, :1547-1560).  With P, C the two input scalars (mode 0:
// P = psi, C = chi; mode 1: P = -del2(chi) as tp2curlvec forms it, C = psi)
// and all band products taken over the rows of the +3 radial offset:
//   ur = i m P + i kappa xxdx(C),  up = -xxdx(P) - m kappa C,  uz = -del2h(C),
// then chopped to the offset-0 rows, columns and slices (rows, columns and
// slices outside are written as 0).  Every band product is applied as many
// times as the reference's two kk-loops visit the slice, and the log-term
// corrections of the (m, kappa) = (0, 0) column are added after the band
// products, as in xxdx / del2h / del2.  The results have no log term.
// ---------------------------------------------------------------------------
__device__ inline double2 band_row(const double2 *x, int i, int n, int op, int m, int ne, const double *d2c,
                                   const double *xxc, double akk) {
  double2 s = make_double2(0.0, 0.0);
  if (op < 2) {
    const double *b = d2c + ((size_t)m * ne + i) * 5;
#pragma unroll
    for (int d = 0; d < 5; ++d) {
      const int j = i + d - 2;
      if (j < 0 || j >= n) continue;
      double c = b[d];
      if (d == 2) c = c - akk;
      s.x += c * x[j].x;
      s.y += c * x[j].y;
    }
  } else {
    const double *b = xxc + ((size_t)m * ne + i) * 2;
    if (i >= 1) { s.x += b[0] * x[i - 1].x; s.y += b[0] * x[i - 1].y; }
    if (i + 1 < n) { s.x += b[1] * x[i + 1].x; s.y += b[1] * x[i + 1].y; }
  }
  return s;
}

// y = B^cnt x over rows < n with tmp as scratch (all in LDS); y may alias x
__device__ inline void band_pow(double2 *y, const double2 *x, double2 *tmp, int n, int cnt, int op, int m, int ne,
                                const double *d2c, const double *xxc, double akk) {
  const double2 *src = x;
  for (int rep = 0; rep < cnt; ++rep) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) tmp[i] = band_row(src, i, n, op, m, ne, d2c, xxc, akk);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) y[i] = tmp[i];
    __syncthreads();
    src = y;
  }
}

__device__ __forceinline__ void tp2vec_front_body(double2 *ur, double2 *up, double2 *uz, const double2 *p_in,
                                                  const double2 *c_in, const double *p_ln, const double *c_ln,
                                                  int mode, int md_off, const double *d2c, const double *xxc,
                                                  const double *lognorm, const double *ak, int nrdim, int npdim,
                                                  int nz, int ne, int nrc, int npc, int nzc, int nzcu, double ell,
                                                  double *ln_r, double *ln_p, double *ln_z, int zlo, int zhi,
                                                  bool staged = false) {
  extern __shared__ double2 lds[];
  // the three outputs carry no log term (their ln slots are distinct from
  // the inputs' p_ln and c_ln)
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
    if (ln_r) ln_r[0] = 0.0;
    if (ln_p) ln_p[0] = 0.0;
    if (ln_z) ln_z[0] = 0.0;
  }

  double2 *P = lds, *C = lds + nrdim, *A = lds + 2 * nrdim, *B = lds + 3 * nrdim;
  // m fastest: consecutive workgroups take adjacent columns of one slice
  const int m = blockIdx.x, k = blockIdx.y;
  // slices zlo..zhi: zero in the inputs, and the z transform that follows
  // takes them as zero without reading them
  if (k >= zlo && k <= zhi) return;
  const size_t off = (size_t)k * nrdim * npdim + (size_t)nrdim * m;
  const int cnt = z_count(k, nzc, nzcu);
  const int nn0 = nrcs_dev(nrc, npc, m);
  const double2 zero = make_double2(0.0, 0.0);
  if (cnt == 0 || nn0 == 0) {
    // a chopped kappa slice of an active column is input to the z transform
    for (int i = threadIdx.x; i < nn0; i += blockDim.x) ur[off + i] = up[off + i] = uz[off + i] = zero;
    return;
  }
  int nn3 = nrcs_dev(nrc + 3, npc, m);
  if (nn3 > nrdim) nn3 = nrdim;
  const bool c00 = (m == 0 && k == 0);
  const double ell2 = ell * ell;
  const double akk = ak[k] * ak[k];
  double lnP = 0.0;
  const double lnC = c_ln[0];
  if (cnt == 1) {
    // staged (mode 1 right after the mode 0 pass of the same column, which
    // left p_in in LDS slot 1 and c_in in slot 0): the columns are taken
    // from there and -del2(p_in) goes to slot 2
    if (staged) {
      P = lds + 2 * nrdim;
      C = lds;
      A = lds + nrdim;
    }
    // one application of each band product: every output row comes straight
    // from the staged columns.  A thread's rows i = tid, tid + 256, ... (at
    // most three): their band coefficients are loaded before the barrier so
    // they arrive with the staged columns.
    constexpr int RT = 3;
    double cx[RT][2], cd[RT][5];
#pragma unroll
    for (int q = 0; q < RT; ++q) {
      const int i = threadIdx.x + q * blockDim.x;
      const bool in = i < nn0;
      const double *bx = xxc + ((size_t)m * ne + (in ? i : 0)) * 2;
      const double *bd = d2c + ((size_t)m * ne + (in ? i : 0)) * 5;
      cx[q][0] = in ? bx[0] : 0.0;
      cx[q][1] = in ? bx[1] : 0.0;
#pragma unroll
      for (int d = 0; d < 5; ++d) cd[q][d] = in ? bd[d] : 0.0;
    }
    const double mv1 = (double)m, kv1 = ak[k];
    if (mode == 0) {
      lnP = p_ln[0];
      for (int i = threadIdx.x; i < nn3; i += blockDim.x) {
        P[i] = p_in[off + i];
        C[i] = c_in[off + i];
      }
      __syncthreads();
    } else {
      int nmd = nrcs_dev(nrc + md_off, npc, m);
      if (nmd > nn3) nmd = nn3;
      const double lch = p_ln[0];
      if (!staged) {
        for (int i = threadIdx.x; i < nn3; i += blockDim.x) {
          A[i] = p_in[off + i];
          C[i] = c_in[off + i];
        }
        __syncthreads();
      }
      for (int i = threadIdx.x; i < nn3; i += blockDim.x) {
        double2 v = A[i];
        if (i < nmd) {
          v = band_row(A, i, nmd, 0, m, ne, d2c, xxc, akk);
          if (c00) {
            if (i == 0) v.x += 4.0 / 3.0 / ell2 / exp(lognorm[0]) * lch;
            if (i == 1) v.x -= 2.0 / 1.0 / ell2 / exp(lognorm[1]) * lch;
            if (i == 2) v.x += 2.0 / 3.0 / ell2 / exp(lognorm[2]) * lch;
          }
        }
        P[i] = make_double2(-v.x, -v.y);
      }
      __syncthreads();
    }
    // Rows i >= nn0 are not written: the transforms that consume these
    // outputs (z transform restricted to the active lines, then the radial
    // transform) read rows i < nrchop - m of columns m < npchop only.
#pragma unroll
    for (int q = 0; q < RT; ++q) {
      const int i = threadIdx.x + q * blockDim.x;
      if (i >= nn0) break;
      // xxdx: coefficients of rows i - 1 and i + 1; del2h: rows i - 2 .. i + 2
      double2 xc = make_double2(0.0, 0.0), xp = xc, hc = xc;
      if (i >= 1) {
        xc.x += cx[q][0] * C[i - 1].x; xc.y += cx[q][0] * C[i - 1].y;
        xp.x += cx[q][0] * P[i - 1].x; xp.y += cx[q][0] * P[i - 1].y;
      }
      if (i + 1 < nn3) {
        xc.x += cx[q][1] * C[i + 1].x; xc.y += cx[q][1] * C[i + 1].y;
        xp.x += cx[q][1] * P[i + 1].x; xp.y += cx[q][1] * P[i + 1].y;
      }
#pragma unroll
      for (int d = 0; d < 5; ++d) {
        const int j = i + d - 2;
        if (j < 0 || j >= nn3) continue;
        hc.x += cd[q][d] * C[j].x;
        hc.y += cd[q][d] * C[j].y;
      }
      if (c00) {
        if (i < 2) {
          xc.x += 1.0 / exp(lognorm[i]) * lnC;
          xp.x += 1.0 / exp(lognorm[i]) * lnP;
        }
        if (i == 0) hc.x += 4.0 / 3.0 / ell2 / exp(lognorm[0]) * lnC;
        if (i == 1) hc.x -= 2.0 / 1.0 / ell2 / exp(lognorm[1]) * lnC;
        if (i == 2) hc.x += 2.0 / 3.0 / ell2 / exp(lognorm[2]) * lnC;
      }
      ur[off + i] = make_double2(-mv1 * P[i].y - kv1 * xc.y, mv1 * P[i].x + kv1 * xc.x);
      up[off + i] = make_double2(-xp.x - mv1 * kv1 * C[i].x, -xp.y - mv1 * kv1 * C[i].y);
      uz[off + i] = make_double2(-hc.x, -hc.y);
    }
    return;
  }
  if (mode == 0) {
    lnP = p_ln[0];
    for (int i = threadIdx.x; i < nn3; i += blockDim.x) P[i] = p_in[off + i];
  } else {
    // md = -del2(chi) with the copy semantics of del2 (rows past the offset-0
    // system keep chi's values), log terms of chi added, then negated
    for (int i = threadIdx.x; i < nn3; i += blockDim.x) P[i] = p_in[off + i];
    __syncthreads();
    int nmd = nrcs_dev(nrc + md_off, npc, m);
    if (nmd > nn3) nmd = nn3;
    band_pow(P, P, A, nmd, cnt, 0, m, ne, d2c, xxc, akk);
    if (c00 && threadIdx.x == 0) {
      const double l = p_ln[0];
      P[0].x += 4.0 / 3.0 / ell2 / exp(lognorm[0]) * l;
      if (nn3 > 1) P[1].x -= 2.0 / 1.0 / ell2 / exp(lognorm[1]) * l;
      if (nn3 > 2) P[2].x += 2.0 / 3.0 / ell2 / exp(lognorm[2]) * l;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nn3; i += blockDim.x) P[i] = make_double2(-P[i].x, -P[i].y);
  }
  for (int i = threadIdx.x; i < nn3; i += blockDim.x) C[i] = c_in[off + i];
  __syncthreads();
  const double mv = (double)m, kv = ak[k];
  // ur = i m P + i kappa xxdx(C)
  band_pow(A, C, B, nn3, cnt, 2, m, ne, d2c, xxc, 0.0);
  if (c00 && threadIdx.x == 0) {
    A[0].x += 1.0 / exp(lognorm[0]) * lnC;
    if (nn3 > 1) A[1].x += 1.0 / exp(lognorm[1]) * lnC;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nrdim; i += blockDim.x) {
    double2 v = zero;
    if (i < nn0) v = make_double2(-mv * P[i].y - kv * A[i].y, mv * P[i].x + kv * A[i].x);
    ur[off + i] = v;
  }
  __syncthreads();
  // up = -xxdx(P) - m kappa C
  band_pow(A, P, B, nn3, cnt, 2, m, ne, d2c, xxc, 0.0);
  if (c00 && threadIdx.x == 0) {
    A[0].x += 1.0 / exp(lognorm[0]) * lnP;
    if (nn3 > 1) A[1].x += 1.0 / exp(lognorm[1]) * lnP;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nrdim; i += blockDim.x) {
    double2 v = zero;
    if (i < nn0) v = make_double2(-A[i].x - mv * kv * C[i].x, -A[i].y - mv * kv * C[i].y);
    up[off + i] = v;
  }
  __syncthreads();
  // uz = -del2h(C)
  band_pow(A, C, B, nn3, cnt, 1, m, ne, d2c, xxc, 0.0);
  if (c00 && threadIdx.x == 0) {
    A[0].x += 4.0 / 3.0 / ell2 / exp(lognorm[0]) * lnC;
    if (nn3 > 1) A[1].x -= 2.0 / 1.0 / ell2 / exp(lognorm[1]) * lnC;
    if (nn3 > 2) A[2].x += 2.0 / 3.0 / ell2 / exp(lognorm[2]) * lnC;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nrdim; i += blockDim.x) {
    double2 v = zero;
    if (i < nn0) v = make_double2(-A[i].x, -A[i].y);
    uz[off + i] = v;
  }
}

__global__ void __launch_bounds__(256) k_tp2vec_front_(double2 *ur, double2 *up, double2 *uz, const double2 *p_in,
                                                      const double2 *c_in, const double *p_ln, const double *c_ln,
                                                      int mode, int md_off, const double *d2c, const double *xxc,
                                                      const double *lognorm, const double *ak, int nrdim, int npdim,
                                                      int nz, int ne, int nrc, int npc, int nzc, int nzcu,
                                                      double ell, double *ln_r, double *ln_p, double *ln_z, int zlo,
                                                      int zhi) {
  tp2vec_front_body(ur, up, uz, p_in, c_in, p_ln, c_ln, mode, md_off, d2c, xxc, lognorm, ak, nrdim, npdim, nz, ne,
                    nrc, npc, nzc, nzcu, ell, ln_r, ln_p, ln_z, zlo, zhi);
}

// tp2vec's and tp2curlvec's front passes over the same column in one
// workgroup: the second pass finds psi and chi in the cache
struct FrontOut {
  double2 *r, *p, *z;
  double *lr, *lp, *lz;
};
__global__ void __launch_bounds__(256) k_tp2vec_front2_(FrontOut v, FrontOut w, const double2 *psi,
                                                       const double2 *chi, const double *psi_ln,
                                                       const double *chi_ln, int md_off, const double *d2c,
                                                       const double *xxc, const double *lognorm, const double *ak,
                                                       int nrdim, int npdim, int nz, int ne, int nrc, int npc,
                                                       int nzc, int nzcu, double ell, int zlo, int zhi) {
  tp2vec_front_body(v.r, v.p, v.z, psi, chi, psi_ln, chi_ln, 0, 0, d2c, xxc, lognorm, ak, nrdim, npdim, nz, ne, nrc,
                    npc, nzc, nzcu, ell, v.lr, v.lp, v.lz, zlo, zhi);
  __syncthreads();  // the staging buffers are reused
  // where the first pass staged psi and chi (one band application), the
  // second takes them from LDS
  tp2vec_front_body(w.r, w.p, w.z, chi, psi, chi_ln, psi_ln, 1, md_off, d2c, xxc, lognorm, ak, nrdim, npdim, nz, ne,
                    nrc, npc, nzc, nzcu, ell, w.lr, w.lp, w.lz, zlo, zhi, true);
}

// v <- v x u on the two packed phi points (mlegs_scalar_ops.f90:1284-1296)
__global__ void k_vecprod_(double2 *vr, double2 *vp, double2 *vz, const double2 *ur, const double2 *up,
                           const double2 *uz, int nr, int nph, int nrdim, int npdim, int nz) {
  size_t n = (size_t)nrdim * npdim * nz;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(t % nrdim);
    int j = (int)((t / nrdim) % npdim);
    if (i >= nr || j >= nph) {
      vr[t] = vp[t] = vz[t] = make_double2(0.0, 0.0);
      continue;
    }
    double2 a1 = vr[t], a2 = vp[t], a3 = vz[t], b1 = ur[t], b2 = up[t], b3 = uz[t];
    vr[t] = make_double2(a2.x * b3.x - a3.x * b2.x, a2.y * b3.y - a3.y * b2.y);
    vp[t] = make_double2(a3.x * b1.x - a1.x * b3.x, a3.y * b1.y - a1.y * b3.y);
    vz[t] = make_double2(a1.x * b2.x - a2.x * b1.x, a1.y * b2.y - a2.y * b1.y);
  }
}

// psi.ln = -Re(f_inf(kappa=0))/2 ; up(:nr, m=0, k=0) += psi.ln (1 + x)  (:1366-1373)

// ---------------------------------------------------------------------------
// ihelmp (mlegs_scalar_ops.f90:900-1000): per (m, kappa) column the system
// (leg_del2^(P/2) + beta leg_del2 + alpha) x = b of the nn x nn truncations,
// banded with kl = ku = P, factored with partial pivoting as LAPACK gbtf2
// (first maximum as pivot, multipliers scaled by the pivot's reciprocal) and
// solved as gbtrs, once per visit of the slice.  The matrix depends on kappa
// only through kappa^2, so one factorization per (m, |kappa|) serves both
// slices, and it is kept (HypLu) for every later call with the same
// (P, alpha, beta, offset).  Three kernels: the factorization (once), the
// (0,0) column's log term, and the solve.
// ---------------------------------------------------------------------------

// One workgroup per (m, |kappa|): the band storage AB (2P + P + 1 rows,
// A(i, j) at AB[2P + i - j][j]) formed and factored in LDS, then written out
// in HypLu's layout.  A zero pivot sets *sing and leaves the system.
template <int P>
__global__ void __launch_bounds__(256) k_ihelmp_fac_(double *__restrict__ ab, int8_t *__restrict__ pv,
                                                     const double *d2c, const double *ak, double alpha, double beta,
                                                     int ne, int nrc, int npc, int nkk, int *sing) {
  constexpr int Q = P / 2, KV = 2 * P, LDAB = 3 * P + 1, W = 2 * P + 1;
  extern __shared__ double hsm[];
  const int m = blockIdx.x, kk = blockIdx.y, tid = threadIdx.x;
  const int nn = nrcs_dev(nrc, npc, m);
  if (nn <= 0) return;
  // HypLu's block of 32 systems and the system's place in it
  const size_t nblk = (size_t)npc * ((nkk + 31) / 32), blk = (size_t)m * ((nkk + 31) / 32) + kk / 32;
  const int ib = kk % 32;
  double *AB = hsm;                                                             // LDAB * nn
  int *piv = reinterpret_cast<int *>(AB + (((size_t)LDAB * nn + 1) & ~(size_t)1));  // nn
  __shared__ int s_jp, s_ju, s_sing;
  const double akk = ak[kk] * ak[kk];
  const double *dm = d2c + (size_t)m * ne * 5;
  auto dcoef = [&](int r, int d) { return dm[(size_t)r * 5 + d] - (d == 2 ? akk : 0.0); };
  for (int q = tid; q < LDAB * nn; q += blockDim.x) AB[q] = 0.0;
  __syncthreads();
  // row i of W: e_i^T D^Q (support i - P .. i + P) + beta D(i, :) + alpha
  for (int i = tid; i < nn; i += blockDim.x) {
    double r[W];
#pragma unroll
    for (int o = 0; o < W; ++o) r[o] = (o == P) ? 1.0 : 0.0;
#pragma unroll
    for (int st = 0; st < Q; ++st) {
      double x[W];
#pragma unroll
      for (int o = 0; o < W; ++o) x[o] = 0.0;
#pragma unroll
      for (int o = 0; o < W; ++o) {
        const int kr = i + o - P;
        if (kr < 0 || kr >= nn || r[o] == 0.0) continue;
#pragma unroll
        for (int d = 0; d < 5; ++d) {
          const int oo = o + d - 2, j = kr + d - 2;
          if (oo < 0 || oo >= W || j < 0 || j >= nn) continue;
          x[oo] += r[o] * dcoef(kr, d);
        }
      }
#pragma unroll
      for (int o = 0; o < W; ++o) r[o] = x[o];
    }
#pragma unroll
    for (int d = 0; d < 5; ++d) {
      const int j = i + d - 2;
      if (j >= 0 && j < nn) r[P + d - 2] += beta * dcoef(i, d);
    }
    r[P] += alpha;
#pragma unroll
    for (int o = 0; o < W; ++o) {
      const int j = i + o - P;
      if (j >= 0 && j < nn) AB[(size_t)j * LDAB + (KV + i - j)] = r[o];
    }
  }
  if (tid == 0) s_ju = 0;
  __syncthreads();
  // gbtf2
  for (int j = 0; j < nn; ++j) {
    const int km = (P < nn - 1 - j) ? P : nn - 1 - j;
    double *cj = AB + (size_t)j * LDAB;
    if (tid == 0) {
      int jp = 0;
      double best = fabs(cj[KV]);
      for (int q = 1; q <= km; ++q)
        if (fabs(cj[KV + q]) > best) { best = fabs(cj[KV + q]); jp = q; }
      piv[j] = jp;
      s_jp = jp;
      // the pivot test is made once, here, and read by every thread from
      // LDS: after the barrier tid 0 may already be swapping cj[KV + jp]
      s_sing = (cj[KV + jp] == 0.0);
      if (!s_sing) {
        const int cand = (j + P + jp < nn - 1) ? j + P + jp : nn - 1;
        if (cand > s_ju) s_ju = cand;
      }
    }
    __syncthreads();
    const int jp = s_jp, ju = s_ju;
    if (s_sing) {  // uniform over the workgroup: every thread leaves here
      if (tid == 0) atomicOr(sing, 1);
      return;  // singular: reported (the reference stops)
    }
    if (jp != 0)
      for (int cc = j + tid; cc <= ju; cc += blockDim.x) {
        double *col = AB + (size_t)cc * LDAB;
        const double x = col[KV + j - cc];
        col[KV + j - cc] = col[KV + j + jp - cc];
        col[KV + j + jp - cc] = x;
      }
    __syncthreads();
    const double rp = 1.0 / cj[KV];
    for (int q = 1 + tid; q <= km; q += blockDim.x) cj[KV + q] *= rp;
    __syncthreads();
    const int ncol = ju - j;
    for (int w = tid; w < km * ncol; w += blockDim.x) {
      const int q = 1 + w % km, cc = j + 1 + w / km;
      double *col = AB + (size_t)cc * LDAB;
      col[KV + j + q - cc] -= cj[KV + q] * col[KV + j - cc];
    }
    __syncthreads();
  }
  // HypLu's layout: entry r of row j (U's r = 0 .. 2P at r, L's at r + 1)
  constexpr int ROW = 3 * P + 2;
  for (int q = tid; q < LDAB * nn; q += blockDim.x) {
    const int j = q / LDAB, r = q % LDAB;
    ab[(((size_t)j * nblk + blk) * ROW + (r <= KV ? r : r + 1)) * 32 + ib] = AB[q];
  }
  for (int j = tid; j < nn; j += blockDim.x) pv[((size_t)j * nblk + blk) * 32 + ib] = (int8_t)piv[j];
}

// The (0,0) column's log term before the solve (the reference's bl, bl2):
// lnp = ln / alpha; b(:, 0, 0) -= leg_del2(0, ak_0)^(Q-1) bl2 + beta bl2,
// with bl2 the log term's Laplacian rows; ln <- lnp.  One workgroup.
template <int P>
__global__ void __launch_bounds__(256) k_ihelmp_ln_(double2 *e, double *sln, const double *d2c, const double *ak,
                                                    const double *lognorm, double alpha, double beta, int ne, int nn,
                                                    double ell) {
  constexpr int Q = P / 2;
  extern __shared__ double hsm[];
  double *v = hsm, *t = hsm + nn;
  const int tid = threadIdx.x;
  const double akk = ak[0] * ak[0];
  auto dcoef = [&](int r, int d) { return d2c[(size_t)r * 5 + d] - (d == 2 ? akk : 0.0); };
  const double lnp = sln[0] / alpha;
  const double ell2 = ell * ell;
  const double b0 = 4.0 / 3.0 / ell2 / exp(lognorm[0]) * lnp, b1 = -2.0 / ell2 / exp(lognorm[1]) * lnp,
               b2 = 2.0 / 3.0 / ell2 / exp(lognorm[2]) * lnp;
  for (int i = tid; i < nn; i += blockDim.x) v[i] = (i == 0) ? b0 : (i == 1) ? b1 : (i == 2) ? b2 : 0.0;
  __syncthreads();
  for (int st = 0; st < Q - 1; ++st) {
    for (int i = tid; i < nn; i += blockDim.x) {
      double a = 0.0;
      for (int d = 0; d < 5; ++d) {
        const int j = i + d - 2;
        if (j >= 0 && j < nn) a += dcoef(i, d) * v[j];
      }
      t[i] = a;
    }
    __syncthreads();
    for (int i = tid; i < nn; i += blockDim.x) v[i] = t[i];
    __syncthreads();
  }
  for (int i = tid; i < nn; i += blockDim.x) {
    const double bl2 = (i == 0) ? b0 : (i == 1) ? b1 : (i == 2) ? b2 : 0.0;
    double2 x = e[i];
    x.x = x.x - v[i] - beta * bl2;
    e[i] = x;
  }
  __syncthreads();
  if (tid == 0) sln[0] = lnp;
}

// gbtrs on the kept factors.  A 64-lane workgroup owns column m and 32
// consecutive |kappa| values; lanes 2i and 2i + 1 solve slices kk and
// nz - kk of system kk (they read the same factor entries).  Each lane runs
// the two sweeps over its own column in registers, in the reference's order
// of operations: the forward sweep keeps rows j .. j + P (interchange, then
// the P eliminations of row j), the back substitution the 2P pending rows
// above row j (x_j, then its 2P updates).  The factor entries of the rows
// ahead are loaded RF (forward) and RB (back) rows early.  The loops have no
// branches around loads or stores, so the loads stay in flight across rows:
// rows past the system's end run with zero multipliers and no interchange
// and store into `sink`, band entries outside the matrix are stored zeros
// (their updates leave values unchanged), and a lane with nothing to solve
// (its slice chopped, the twin of a self-conjugate slice, past the last
// |kappa|) works in `sink` (garbage in, garbage out, never read as data).
// A slice visited twice by the chop (z_count 2) is solved twice.
template <int P>
__global__ void __launch_bounds__(64) k_ihelmp_sol_(double2 *e, const double *__restrict__ ab,
                                                    const int8_t *__restrict__ pv, double2 *sink, int nrdim,
                                                    int npdim, int nz, int nrc, int npc, int nkk, int nzc, int nzcu,
                                                    int zlo, int zhi) {
  constexpr int KV = 2 * P, ROW = 3 * P + 2, RF = 4, RB = 2;
  const int m = blockIdx.y, lane = threadIdx.x;
  const int nn = nrcs_dev(nrc, npc, m);  // uniform over the workgroup
  if (nn <= 0) return;
  const int kk0 = blockIdx.x * 32 + lane / 2, twin = lane & 1;
  const bool have = kk0 < nkk;
  const int kk = have ? kk0 : 0;
  const int k = twin ? (nz - kk) % nz : kk;
  const bool valid = have && (!twin || (kk != 0 && 2 * kk != nz));
  // slices zlo..zhi are left as they are
  const int cnt = (valid && !(k >= zlo && k <= zhi)) ? z_count(k, nzc, nzcu) : 0;
  // HypLu's block of the workgroup's 32 systems, the lane's place in it
  const size_t nblk = (size_t)npc * ((nkk + 31) / 32), blk = (size_t)m * ((nkk + 31) / 32) + blockIdx.x;
  const int ib = lane / 2;
  double2 *const colv = e + (size_t)k * nrdim * npdim + (size_t)nrdim * m;
  double2 *const snk = sink + (size_t)lane * (nrc + KV + RF);  // rows j < nrc + RF written, -KV-1.. read
  const double2 z2 = make_double2(0.0, 0.0);
  // row j's factors (clamped to the system's rows): entry r at rw[32 r]
#define MLX_JC(j) ((j) < 0 ? 0 : ((j) > nn - 1 ? nn - 1 : (j)))
#define MLX_ROW(j) (ab + ((size_t)MLX_JC(j) * nblk + blk) * ROW * 32 + ib)
  // a right-hand-side entry (zero outside 0 .. nn - 1), read without a branch
  // (component selects: a select of whole double2 values becomes a select of
  // addresses)
#define MLX_IDX(i) (((i) >= 0 && (i) < nn) ? (i) : 0)
#define MLX_IN(i) ((i) >= 0 && (i) < nn)
#define MLX_ZSEL(dst, in, v)      \
  do {                            \
    const bool in_ = (in);        \
    dst.x = in_ ? (v).x : 0.0;    \
    dst.y = in_ ? (v).y : 0.0;    \
  } while (0)
  // ring slot u: the raw entries of row j (loads only: the row-validity
  // selects are applied where the values are used, so no instruction waits
  // on a load before its row comes up)
#define MLX_LOAD_F(j, u)                                                            \
  do {                                                                              \
    const double *rw = MLX_ROW(j) + (KV + 2) * 32;                                  \
    _Pragma("unroll") for (int q = 0; q < P; ++q) L[u][q] = rw[32 * q];             \
    jp[u] = pv[((size_t)MLX_JC(j) * nblk + blk) * 32 + ib];                         \
    nb[u] = col[MLX_IDX((j) + P + 1)];                                              \
  } while (0)
#define MLX_LOAD_B(j, u)                                                            \
  do {                                                                              \
    const double *rw = MLX_ROW(j);                                                  \
    _Pragma("unroll") for (int q = 0; q <= KV; ++q) U[u][q] = rw[32 * (KV - q)];    \
    yb[u] = col[MLX_IDX((j) - KV - 1)];                                             \
  } while (0)
#pragma unroll 1
  for (int rep = 0; rep < 2; ++rep) {
    if (!__any(rep < cnt)) break;
    // the lane's working column: its slice, or the sink
    double2 *const col = (rep < cnt) ? colv : snk + KV + 1;
    // forward sweep over rows 0 .. nn - 1 (and padding to a multiple of RF):
    // F[q] = b(j + q); row j stores y_j
    {
      double L[RF][P];
      int jp[RF];
      double2 nb[RF];  // b(j + P + 1) for the rows ahead
      double2 F[P + 1];
#pragma unroll
      for (int q = 0; q <= P; ++q) {
        const double2 v = col[MLX_IDX(q)];
        MLX_ZSEL(F[q], MLX_IN(q), v);
      }
#pragma unroll
      for (int u = 0; u < RF; ++u) MLX_LOAD_F(u, u);
#pragma unroll 1
      for (int j0 = 0; j0 < nn; j0 += RF) {
#pragma unroll
        for (int u = 0; u < RF; ++u) {
          const int j = j0 + u;
          // interchange rows j and j + jp, as selects over the window: the
          // empty asm hides that the mask has one bit, so the compiler keeps
          // F in registers instead of indexing it through memory; component
          // selects, since a select of whole double2 values becomes a select
          // of addresses
          const bool el = j < nn - 1;  // row j has eliminations
          unsigned pm = 1u << (el ? jp[u] : 0);
          asm volatile("" : "+v"(pm));
          double f0x = F[0].x, f0y = F[0].y;
#pragma unroll
          for (int q = 1; q <= P; ++q) {
            const bool sw = (pm >> q) & 1u;
            const double ax = F[q].x, ay = F[q].y;
            F[q].x = sw ? F[0].x : ax;
            F[q].y = sw ? F[0].y : ay;
            f0x = sw ? ax : f0x;
            f0y = sw ? ay : f0y;
          }
          F[0].x = f0x;
          F[0].y = f0y;
          // multipliers of rows below nn - 1 are stored zeros
#pragma unroll
          for (int q = 1; q <= P; ++q) {
            const double l = el ? L[u][q - 1] : 0.0;
            F[q].x -= l * F[0].x;
            F[q].y -= l * F[0].y;
          }
          double2 *const o = (j < nn) ? col + j : snk + j;
          *o = F[0];
#pragma unroll
          for (int q = 0; q < P; ++q) F[q] = F[q + 1];
          MLX_ZSEL(F[P], MLX_IN(j + P + 1), nb[u]);
          MLX_LOAD_F(j + RF, u);
        }
      }
    }
    // back substitution over rows nn - 1 .. 0 (and padding): B[q] = b(j - q),
    // pending; row j stores x_j
    {
      double U[RB][KV + 1];
      double2 yb[RB];  // b(j - 2P - 1) for the rows ahead
      double2 B[KV + 1];
#pragma unroll
      for (int q = 0; q <= KV; ++q) {
        const double2 v = col[MLX_IDX(nn - 1 - q)];
        MLX_ZSEL(B[q], MLX_IN(nn - 1 - q), v);
      }
#pragma unroll
      for (int u = 0; u < RB; ++u) MLX_LOAD_B(nn - 1 - u, u);
#pragma unroll 1
      for (int j0 = nn - 1; j0 >= 0; j0 -= RB) {
#pragma unroll
        for (int u = 0; u < RB; ++u) {
          const int j = j0 - u;
          const bool el = j >= 0;  // a row of the system
          const double d = el ? U[u][0] : 1.0;
          const double2 x = make_double2(B[0].x / d, B[0].y / d);
          double2 *const o = (j >= 0) ? col + j : snk + (j + KV + 1);
          *o = x;
          // U(j - q, j) of rows above the first is a stored zero
#pragma unroll
          for (int q = 1; q <= KV; ++q) {
            const double a = el ? U[u][q] : 0.0;
            B[q].x -= a * x.x;
            B[q].y -= a * x.y;
          }
#pragma unroll
          for (int q = 0; q < KV; ++q) B[q] = B[q + 1];
          MLX_ZSEL(B[KV], MLX_IN(j - KV - 1), yb[u]);
          MLX_LOAD_B(j - RB, u);
        }
      }
    }
  }
#undef MLX_LOAD_B
#undef MLX_LOAD_F
#undef MLX_ZSEL
#undef MLX_IN
#undef MLX_IDX
#undef MLX_ROW
#undef MLX_JC
}

// ---------------------------------------------------------------------------
// fftreat (mlegs_scalar_ops.f90:1002-1063), the physical-space part: on the
// radial lines of the field after delsqp and rtrans_backward, per column
// m < npc and kappa slice, (1 - x)^2 weighting of rows < nr, rows
// >= ns0 - 1 zeroed in columns m >= 1 (every slice), five passes of smooth
// (:2149-2175) over rows ns - 1 .. nrow - 1 on the chopped slices, and the
// weighting removed again.  nrow is the reference's column length nr +
// max(3, hyperpow): the smoothing reaches into the rows past nr.
// smooth(a), a of length ni: out(1) = a(1); a(ni) goes to out(ni-2:ni) with
// weights 0.2, 0.3, 0.5; each a(i), 1 < i < ni, to out(i-1:i+1) with
// weights (1-f)/2, f, (1-f)/2, f = (1 + (ni-i)/(ni-1))/2 (1-based).
// One workgroup per (m, kappa) line, the line in LDS.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fftreat_mid_(double2 *e, const double *x, int nrdim, int npdim, int nr,
                                                      int nrow, int ns, int ns0, int npc, int nzc, int nzcu) {
  extern __shared__ double2 line[];  // 2 nrow
  double2 *a = line, *b = line + nrow;
  const int m = blockIdx.x, k = blockIdx.y, tid = threadIdx.x;
  double2 *g = e + (size_t)k * nrdim * npdim + (size_t)nrdim * m;
  const bool act = (k + 1 <= nzc) || (k + 1 >= nzcu);
  if (!act) {
    if (m >= 1)
      for (int i = ns0 - 1 + tid; i < nrow; i += blockDim.x) g[i] = make_double2(0.0, 0.0);
    return;
  }
  for (int i = tid; i < nrow; i += blockDim.x) {
    double2 v = g[i];
    if (i < nr) {
      const double w = (1.0 - x[i]) * (1.0 - x[i]);
      v = make_double2(v.x * w, v.y * w);
    }
    if (m >= 1 && i >= ns0 - 1) v = make_double2(0.0, 0.0);
    a[i] = v;
  }
  __syncthreads();
  const int o = ns - 1, ni = nrow - o;
  for (int pass = 0; pass < 5; ++pass) {
    for (int j = tid; j < ni; j += blockDim.x) {
      double2 s = make_double2(0.0, 0.0);
      if (j == 0) s = a[o];
      // interior sources i = j-1, j, j+1 (1 <= i <= ni-2)
      for (int i = j - 1; i <= j + 1; ++i) {
        if (i < 1 || i > ni - 2) continue;
        const double f = (1.0 + (double)(ni - (i + 1)) / (ni - 1.0)) * 0.5;
        const double wgt = (i == j) ? f : (1.0 - f) / 2.0;
        s.x += a[o + i].x * wgt;
        s.y += a[o + i].y * wgt;
      }
      if (j >= ni - 3) {
        const double fae = (j == ni - 3) ? 0.2 : (j == ni - 2) ? 0.3 : 0.5;
        s.x += a[o + ni - 1].x * fae;
        s.y += a[o + ni - 1].y * fae;
      }
      b[o + j] = s;
    }
    __syncthreads();
    for (int j = tid; j < ni; j += blockDim.x) a[o + j] = b[o + j];
    __syncthreads();
  }
  for (int i = tid; i < nrow; i += blockDim.x) {
    double2 v = a[i];
    if (i < nr) {
      const double w = (1.0 - x[i]) * (1.0 - x[i]);
      v = make_double2(v.x / w, v.y / w);
    }
    g[i] = v;
  }
}

__global__ void k_vec2tp_up_ln_(double2 *up, const double2 *inf, double *psiln, const double *x, int nr) {
  double l = -1.0 / 2.0 * inf[0].x;
  if (blockIdx.x == 0 && threadIdx.x == 0) psiln[0] = l;
  for (int i = threadIdx.x; i < nr; i += blockDim.x) up[i].x += l * (1.0 + x[i]);
}

__global__ void k_nonfinite_(const double2 *a, const double2 *b, size_t n, int *flag) {
  int bad = 0;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    double2 u = a[t], v = b[t];
    if (!isfinite(u.x) || !isfinite(u.y) || !isfinite(v.x) || !isfinite(v.y)) bad = 1;
  }
  if (bad) atomicOr(flag, 1);
}

__global__ void k_energy_(const double2 *vr, const double2 *vp, const double2 *vz, const double *x, const double *w,
                          int nr, int nph, int nrdim, int npdim, int nz, double ell2, double cell, double *part) {
  __shared__ double sh[16];
  size_t n = (size_t)nrdim * npdim * nz;
  double s = 0.0;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(t % nrdim);
    int j = (int)((t / nrdim) % npdim);
    if (i >= nr || j >= nph) continue;
    double2 a = vr[t], b = vp[t], c = vz[t];
    double q = a.x * a.x + b.x * b.x + c.x * c.x + a.y * a.y + b.y * b.y + c.y * c.y;
    double om = 1.0 - x[i];
    s += w[i] * ell2 / (om * om) * q;
  }
  double tot = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = 0.5 * cell * tot;
}

__global__ void k_sum_parts(const double *part, int nb, double *out) {
  __shared__ double sh[16];
  double a = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) a += part[i];
  double t = block_sum(a, sh);
  if (threadIdx.x == 0) out[0] = t;
}

// plain copy of a whole field (also the known-size reference for HBM counters)
__global__ void k_copy_field_(double2 *__restrict__ d, const double2 *__restrict__ s, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    d[i] = s[i];
}

// read-only and write-only passes over a whole field (the HBM ceilings of a
// kernel that only reads or only writes): the read pass leaves one partial
// sum per workgroup so the loads are not dropped
__global__ void k_read_field_(const double2 *__restrict__ s, size_t n, double *__restrict__ part) {
  __shared__ double sh[16];
  double a = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double2 v = s[i];
    a += v.x + v.y;
  }
  const double t = block_sum(a, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void k_write_field_(double2 *__restrict__ d, size_t n, double v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    d[i] = make_double2(v, v);
}

// |w| at each point, on the real and the imaginary parts separately (the two
// azimuthal points a packed PPP entry holds): save_vort_mag
// (vortical_flow_3d.f90:433-443)
__global__ void k_vormag_(double2 *__restrict__ o, const double2 *__restrict__ a, const double2 *__restrict__ b,
                          const double2 *__restrict__ d, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double2 x = a[i], y = b[i], z = d[i];
    o[i] = make_double2(sqrt(x.x * x.x + y.x * y.x + z.x * z.x), sqrt(x.y * x.y + y.y * y.y + z.y * z.y));
  }
}

__global__ void k_set_scalar_(double *d, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) d[0] = v;
}
__global__ void k_set_flag_(int *d, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) d[0] = v;
}
__global__ void k_copy_scalar_(double *d, const double *s, double scale) {
  if (threadIdx.x == 0 && blockIdx.x == 0) d[0] = scale * s[0];
}

inline int grid_for(size_t n) {
  size_t b = (n + kT - 1) / kT;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace

namespace mlx {

void k_ln_column(mlx_ctx *c, double2 *e, const double *ln, double sign) {
  const KitDev &k = c->kit;
  dim3 grid((k.nr + 127) / 128, k.nzdim);
  hipLaunchKernelGGL(k_ln_column_, grid, dim3(128), 0, c->stream, e, ln, k.lnx, sign, k.nr, k.nrdim, k.npdim,
                     k.nzdim);
}

void k_chop(mlx_ctx *c, double2 *e, const char *space, int nroff) {
  const KitDev &k = c->kit;
  hipLaunchKernelGGL(k_chop_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, e, k.nrdim, k.npdim, k.nzdim,
                     space[0] == 'F', space[1] == 'F', space[2] == 'F', k.nrchop + nroff, k.npchop, k.chopzl,
                     k.chopzu);
}

void k_dealias(mlx_ctx *c, double2 *e, const char *space) {
  const KitDev &k = c->kit;
  int pcut = -1, zcut = -1, zupper = 0;
  if (k.np > 1 && space[1] == 'F') pcut = (k.np / 3 + 1 > 1) ? k.np / 3 + 1 : 1;
  if (k.nz > 1 && space[2] == 'F') {
    zcut = (k.nz / 3 + 1 > 1) ? k.nz / 3 + 1 : 1;
    zupper = k.nz - zcut + 2;
  }
  if (pcut < 0 && zcut < 0) return;
  hipLaunchKernelGGL(k_dealias_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, e, k.nrdim, k.npdim, k.nzdim,
                     pcut, zcut, zupper);
}

void k_svv(mlx_ctx *c, double2 *e0, double *g0, double2 *e1, double *g1, int *flag, int pcut, int zlo, int zhi,
           double *part_ws, double *out_ws) {
  const KitDev &k = c->kit;
  const mlx_params &p = c->p;
  double inv_r = 1.0 / fmax((double)(k.nrchop - 1), 1.0);
  double inv_p = 1.0 / fmax((double)(k.np / 2), 1.0);
  double kmax = 1.0;
  for (double a : k.ak_h) kmax = fmax(kmax, fabs(a));
  double cutoff = fmin(fmax(p.svv_cutoff, 0.0), 0.99);
  double target = fmax(p.svv_target, 1e-12);
  double relax = fmin(fmax(p.svv_relax, 0.0), 1.0);
  double st0 = fmax(p.svv_strength, 0.0);
  const int nf = e1 ? 2 : 1;
  SvvFields F{{e0, e1 ? e1 : e0}, {g0, g1 ? g1 : g0}};
  // workgroups per slice, each over a range of about eight columns (a few
  // loads per thread: the pass is latency-bound otherwise); partial sums of
  // at most kSvvParts workgroups per field (c->dpart or part_ws)
  const int ncol = (pcut > 0 && pcut < k.npdim) ? pcut : k.npdim;
  int nbx = (ncol + 7) / 8;
  if (nbx * k.nzdim > kSvvParts) nbx = kSvvParts / k.nzdim;
  if (nbx < 1) nbx = 1;
  if ((size_t)nbx * k.nzdim > kSvvParts) MLX_THROW(MLX_E_ARG, "svv_filter: nz above 4096 is not supported");
  double *part = part_ws ? part_ws : c->dpart;
  double *out = out_ws ? out_ws : c->dsmall + 2048;
  hipLaunchKernelGGL(k_svv_reduce, dim3(nbx, k.nzdim, nf), dim3(kT), 0, c->stream, F, k.nrdim, k.npdim, inv_r,
                     inv_p, k.ak, kmax, cutoff, part, flag, ncol, zlo, zhi);
  hipLaunchKernelGGL(k_svv_gain, dim3(nf), dim3(1024), 0, c->stream, F, part, nbx * k.nzdim, out, target, relax,
                     st0);
  hipLaunchKernelGGL(k_svv_apply, dim3(grid_for(c->nelem), nf), dim3(kT), 0, c->stream, F, k.nrdim, k.npdim,
                     k.nzdim, inv_r, inv_p, k.ak, kmax, cutoff, out);
}

void k_calcat(mlx_ctx *c, const double2 *e, const double *tbl, int nrc, double2 *out) {
  const KitDev &k = c->kit;
  int n = nrc < k.nrchop ? nrc : k.nrchop;
  if (n > k.nrdim) n = k.nrdim;
  hipLaunchKernelGGL(k_calcat_, dim3(k.nzdim), dim3(kT), 0, c->stream, e, tbl, n, k.nrdim, k.npdim, out);
}

void k_zeroat1(mlx_ctx *c, double2 *e, int nrc, double2 *e2, int nrc2, double *lnz, double *lnz2) {
  const KitDev &k = c->kit;
  auto rows = [&](int nrc_) {
    int n = nrc_ < k.nrchop ? nrc_ : k.nrchop;
    return n > k.nrdim ? k.nrdim : n;
  };
  const int nf = e2 ? 2 : 1;
  Zat1Fields F{{e, e2 ? e2 : e}, {rows(nrc), rows(e2 ? nrc2 : nrc)}, {lnz, e2 ? lnz2 : nullptr}};
  hipLaunchKernelGGL(k_zeroat1_, dim3(k.nzdim, nf), dim3(kT), 0, c->stream, F, k.at1, k.nrdim, k.npdim);
}

void k_delsqp(mlx_ctx *c, double2 *e, double *ln, int nroff, int inverse) {
  const KitDev &k = c->kit;
  hipLaunchKernelGGL(k_delsqp_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, e, ln, k.lognorm, k.nrdim,
                     k.npdim, k.nzdim, k.nrchop + nroff, k.npchop, k.ell, inverse);
  if (inverse)
    hipLaunchKernelGGL(k_zero_00k, dim3((k.nzdim + 255) / 256), dim3(256), 0, c->stream, e, k.nrdim, k.npdim,
                       k.nzdim);
}

void k_bandmul(mlx_ctx *c, double2 *e, double *ln, int op, int nroff) {
  const KitDev &k = c->kit;
  int nrc = k.nrchop + nroff;
  if (nrc > k.nrdim_ref) MLX_THROW(MLX_E_ARG, "chop_index: chopping in r too large");
  if (nrc > 8 * 64) MLX_THROW(MLX_E_ARG, "banded operators support at most 512 radial modes");
  if (op == 2 && nrc > k.ne) MLX_THROW(MLX_E_ARG, "xxdx: table too small");
  dim3 grid(k.npdim, k.nzdim);
  size_t lds = sizeof(double2) * (size_t)k.nrdim;
  hipLaunchKernelGGL(k_bandmul_, grid, dim3(64), lds, c->stream, e, ln, k.d2c, k.xxc, k.lognorm, k.ak, op, k.nrdim,
                     k.npdim, k.nzdim, k.ne, nrc, k.npchop, k.chopzl, k.chopzu, k.ell);
  HIPCHK(hipGetLastError());
}

void k_helm_pre(mlx_ctx *c, double2 *e, double *ln_out, const double *ln_in, double alpha) {
  hipLaunchKernelGGL(k_helm_pre_, dim3(1), dim3(64), 0, c->stream, e, ln_out, ln_in, c->kit.lognorm, alpha,
                     c->kit.ell);
}

void k_idel2_post(mlx_ctx *c, const double2 *e, double *ln) {
  hipLaunchKernelGGL(k_idel2_post_, dim3(1), dim3(64), 0, c->stream, e, ln, c->kit.lognorm);
}

void k_idel2_preln00(mlx_ctx *c, double2 *e, double *ln, double preln, int nroff) {
  const KitDev &k = c->kit;
  int n = nrcs_of(k, 0, nroff);
  if (n <= 0) return;
  // band storage (9 n) and the right-hand side (2 n) in the second
  // field-sized scratch buffer, which no caller of this function holds
  double *work = reinterpret_cast<double *>(c->scr[1]);
  if (11 * (size_t)n > 2 * c->nelem) MLX_THROW(MLX_E_ARG, "idel2: field too small for the (0,0) solve");
  hipLaunchKernelGGL(k_idel2_preln00_, dim3(1), dim3(64), 0, c->stream, e, ln, k.d2c, k.lognorm, k.ne, n, preln,
                     k.ell, work);
}

namespace {
constexpr int kHypEntries = 4;

void hyp_free(HypLu &h) {
  if (h.ab) (void)hipFree(h.ab);
  if (h.pv) (void)hipFree(h.pv);
  if (h.sink) (void)hipFree(h.sink);
  h.ab = nullptr;
  h.pv = nullptr;
  h.sink = nullptr;
}

// rows of the solve's sink per lane (k_ihelmp_sol_: nrc + 2P + 4)
size_t hyp_sink_rows(int nrc, int power) { return (size_t)nrc + 2 * power + 4; }

// The factors for (power, alpha, beta, nroff): kept ones, else factored now
// (a zero pivot throws, nothing is kept)
HypLu &hyp_get(mlx_ctx *c, int power, double alpha, double beta, int nroff, int nrc, int *sing) {
  for (HypLu &h : c->hyp)
    if (h.power == power && h.alpha == alpha && h.beta == beta && h.nroff == nroff) {
      h.stamp = ++c->lu_clock;
      return h;
    }
  const KitDev &k = c->kit;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  HIPCHK(hipStreamIsCapturing(c->stream, &st));
  if (st != hipStreamCaptureStatusNone)
    MLX_THROW(MLX_E_ARG, "ihelmp: first use of (power, alpha, beta) under stream capture (run one step before capturing)");
  if ((int)c->hyp.size() >= kHypEntries) {  // evict the least recently used entry
    size_t victim = 0;
    for (size_t i = 1; i < c->hyp.size(); ++i)
      if (c->hyp[i].stamp < c->hyp[victim].stamp) victim = i;
    HIPCHK(hipStreamSynchronize(c->stream));
    hyp_free(c->hyp[victim]);
    c->hyp.erase(c->hyp.begin() + victim);
  }
  HypLu h;
  h.power = power;
  h.alpha = alpha;
  h.beta = beta;
  h.nroff = nroff;
  h.nkk = (k.nzdim > 1) ? k.nzdim / 2 + 1 : 1;
  h.nsys = k.npchop * ((h.nkk + 31) / 32) * 32;  // blocks of 32 systems
  const int LDAB = 3 * power + 1;
  // factor rows of 3P + 2 doubles (HypLu's layout)
  if (hipMalloc(&h.ab, sizeof(double) * (size_t)nrc * (LDAB + 1) * h.nsys) != hipSuccess ||
      hipMalloc(&h.pv, (size_t)nrc * h.nsys) != hipSuccess ||
      hipMalloc(&h.sink, sizeof(double2) * 64 * hyp_sink_rows(nrc, power)) != hipSuccess) {
    hyp_free(h);
    MLX_THROW(MLX_E_ALLOC, "ihelmp: cannot allocate the banded factors");
  }
  const size_t lds = (((size_t)LDAB * nrc + 1) & ~(size_t)1) * sizeof(double) + (size_t)nrc * sizeof(int);
  dim3 grid(k.npchop, h.nkk);
#define MLX_IHF(P_)                                                                                           \
  hipLaunchKernelGGL(k_ihelmp_fac_<P_>, grid, dim3(256), lds, c->stream, h.ab, h.pv, k.d2c, k.ak, alpha, beta, \
                     k.ne, nrc, k.npchop, h.nkk, sing)
  switch (power) {
    case 4: MLX_IHF(4); break;
    case 6: MLX_IHF(6); break;
    default: MLX_IHF(8); break;
  }
#undef MLX_IHF
  HIPCHK(hipGetLastError());
  int hs = 0;
  HIPCHK(hipMemcpyAsync(&hs, sing, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (hs) {
    hyp_free(h);
    MLX_THROW(MLX_E_ARG, "ihelmp: zero pivot in a banded system (singular); the field is left unsolved");
  }
  h.stamp = ++c->lu_clock;
  c->hyp.push_back(h);
  return c->hyp.back();
}
}  // namespace

void hyp_cache_free(mlx_ctx *c) {
  for (HypLu &h : c->hyp) hyp_free(h);
  c->hyp.clear();
}

void k_ihelmp(mlx_ctx *c, double2 *e, double *sln, int power, double alpha, double beta, int nroff, int *sing,
              int pcut, int zlo, int zhi) {
  const KitDev &k = c->kit;
  const int nrc = k.nrchop + nroff;
  // chop_index's check (mlegs_scalar_ops.f90:2053): the rows of a system lie
  // inside the field's column
  if (nrc > k.nrdim_ref || nrc > k.ne) MLX_THROW(MLX_E_ARG, "chop_index: chopping in r too large");
  if (power != 4 && power != 6 && power != 8) MLX_THROW(MLX_E_ARG, "ihelmp: power must be 4, 6 or 8");
  const size_t lds = (size_t)nrc * ((3 * power + 1) * sizeof(double) + sizeof(int)) + sizeof(double);
  if (lds > 65536) MLX_THROW(MLX_E_ARG, "ihelmp: the banded system of one column exceeds the workgroup's LDS");
  if (nrc <= 0) return;
  HypLu &h = hyp_get(c, power, alpha, beta, nroff, nrc, sing);
  const int npc_grid = (pcut > 0 && pcut < k.npchop) ? pcut : k.npchop;
  dim3 grid((h.nkk + 31) / 32, npc_grid);
#define MLX_IHS(P_)                                                                                             \
  do {                                                                                                          \
    hipLaunchKernelGGL(k_ihelmp_ln_<P_>, dim3(1), dim3(256), 2 * sizeof(double) * nrc, c->stream, e, sln, k.d2c, \
                       k.ak, k.lognorm, alpha, beta, k.ne, nrc, k.ell);                                        \
    hipLaunchKernelGGL(k_ihelmp_sol_<P_>, grid, dim3(64), 0, c->stream, e, h.ab, h.pv, h.sink, k.nrdim, k.npdim, \
                       k.nzdim, \
                       nrc, k.npchop, h.nkk, k.chopzl, k.chopzu, zlo, zhi);                                    \
  } while (0)
  switch (power) {
    case 4: MLX_IHS(4); break;
    case 6: MLX_IHS(6); break;
    default: MLX_IHS(8); break;
  }
#undef MLX_IHS
  HIPCHK(hipGetLastError());
}

void k_fftreat_mid(mlx_ctx *c, double2 *e, int npc) {
  const KitDev &k = c->kit;
  const int ns = k.nr * 3 / 4, ns0 = (ns + 4 < k.nr) ? ns + 4 : k.nr;
  const int nrow = k.nrdim_ref;
  if (nrow - ns + 1 < 3) MLX_THROW(MLX_E_ARG, "smooth: the input length must be longer than or equal to 3.");
  if (npc > k.npdim) MLX_THROW(MLX_E_ARG, "chop_index: chopping in p too large");
  if (npc <= 0) return;
  const size_t lds = 2 * sizeof(double2) * (size_t)nrow;
  if (lds > 65536) MLX_THROW(MLX_E_ARG, "fftreat: the radial line exceeds the workgroup's LDS");
  hipLaunchKernelGGL(k_fftreat_mid_, dim3(npc, k.nzdim), dim3(256), lds, c->stream, e, k.x, k.nrdim, k.npdim, k.nr,
                     nrow, ns, ns0, npc, k.chopzl, k.chopzu);
  HIPCHK(hipGetLastError());
}

void k_axpby(mlx_ctx *c, double2 *dst, double *dln, double a, const double2 *src, const double *sln, double b) {
  hipLaunchKernelGGL(k_axpby_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, dst, dln, a, src, sln, b,
                     c->nelem);
}

void k_read_field(mlx_ctx *c, const double2 *src, double *part) {
  hipLaunchKernelGGL(k_read_field_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, src, c->nelem, part);
}

void k_write_field(mlx_ctx *c, double2 *dst, double v) {
  hipLaunchKernelGGL(k_write_field_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, dst, c->nelem, v);
}

void k_copy_field(mlx_ctx *c, double2 *dst, const double2 *src) {
  hipLaunchKernelGGL(k_copy_field_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, dst, src, c->nelem);
}

void k_abcn_rhs(mlx_ctx *c, double2 *sh, double *shln, const double2 *s, const double *sln, const double2 *nl,
                const double *nlln, const double2 *nlp, const double *nlpln, const double2 *svis, double dt,
                double a, int ab2) {
  hipLaunchKernelGGL(k_abcn_rhs_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, sh, shln, s, sln, nl, nlln, nlp,
                     nlpln, svis, dt, a, c->p.visc, ab2, c->nelem);
}

// the same with the diffusive term's factor given (hyperviscous abcn: the
// powered operator's factor hypervisc (-1)^(power/2+1) in place of visc)
void k_abcn_rhs_g(mlx_ctx *c, double2 *sh, double *shln, const double2 *s, const double *sln, const double2 *nl,
                  const double *nlln, const double2 *nlp, const double *nlpln, const double2 *svis, double dt,
                  double a, int ab2, double g) {
  hipLaunchKernelGGL(k_abcn_rhs_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, sh, shln, s, sln, nl, nlln, nlp,
                     nlpln, svis, dt, a, g, ab2, c->nelem);
}

void k_abcn_rhs_d2(mlx_ctx *c, double2 *e, double *ln, const double2 *nl, const double *nlln, const double2 *nlp,
                   const double *nlpln, double dt, double a, int nroff, int pcut, int zlo, int zhi, bool helm_pre) {
  const KitDev &k = c->kit;
  const int nrc = k.nrchop + nroff;
  if (nrc > k.nrdim_ref) MLX_THROW(MLX_E_ARG, "chop_index: chopping in r too large");
  if (pcut <= 0 || pcut > k.npdim) pcut = k.npdim;
  if (helm_pre && zlo <= 0 && zhi >= 0) MLX_THROW(MLX_E_ARG, "abcn: the (0,0) column must be in the solved range");
  dim3 grid(pcut, k.nzdim);
  size_t lds = 3 * sizeof(double2) * (size_t)k.nrdim;
  // the log coefficient first: the data kernel reads s's ln for the (0,0) correction
  // with helm_pre the (0,0) workgroup also takes ihelm's log-term step
  // algorithmic bytes: s, N and N_p's active rows in, the right-hand side out
  const double nzk = (double)k.nzdim - ((zhi >= zlo) ? (double)(zhi - zlo + 1) : 0.0);
  ProbeScope probe(c, MLX_PROBE_ABCN_RHS, 64.0 * nzk * active_pairs(k, nroff, pcut), 0.0);
  probed_launch(probe, k_abcn_rhs_d2_, grid, dim3(256), lds, c->stream, e, ln, nl, nlp, k.d2c, k.lognorm, k.ak,
                k.nrdim, k.npdim, k.ne, nrc, k.npchop, k.chopzl, k.chopzu, k.ell, dt, a, c->p.visc, pcut, zlo, zhi,
                helm_pre ? nlln : nullptr, helm_pre ? nlpln : nullptr, helm_pre ? 1 : 0);
  if (!helm_pre) hipLaunchKernelGGL(k_abcn_ln_, dim3(1), dim3(1), 0, c->stream, ln, ln, nlln, nlpln, dt, a);
  HIPCHK(hipGetLastError());
}

void k_abcn_rhs_hyp(mlx_ctx *c, double2 *e, double *ln, const double2 *nl, const double *nlln, const double2 *nlp,
                    const double *nlpln, double dt, double a, double g, double b, int power, int nroff, int pcut,
                    int zlo, int zhi, bool act) {
  const KitDev &k = c->kit;
  const int nrc = k.nrchop + nroff;
  if (nrc > k.nrdim_ref) MLX_THROW(MLX_E_ARG, "chop_index: chopping in r too large");
  if (pcut <= 0 || pcut > k.npdim) pcut = k.npdim;
  dim3 grid(pcut, k.nzdim);
  const size_t lds = 4 * sizeof(double2) * (size_t)k.nrdim;
  hipLaunchKernelGGL(k_abcn_rhs_hyp_, grid, dim3(256), lds, c->stream, e, ln, nl, nlp, k.d2c, k.lognorm, k.ak,
                     k.nrdim, k.npdim, k.ne, nrc, k.npchop, k.chopzl, k.chopzu, k.ell, dt, a, g, b, power / 2, pcut,
                     zlo, zhi, act ? 1 : 0);
  // the log coefficient after every read of s.ln above
  hipLaunchKernelGGL(k_abcn_ln_, dim3(1), dim3(1), 0, c->stream, ln, ln, nlln, nlpln, dt, a);
  HIPCHK(hipGetLastError());
}

void k_scale_rows_r(mlx_ctx *c, double2 *e, int inverse) {
  const KitDev &k = c->kit;
  hipLaunchKernelGGL(k_scale_rows_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, e, k.r, k.nr, k.nrdim, k.npdim,
                     k.nzdim, inverse);
}

void k_tp2vec_combine(mlx_ctx *c, double2 *ur, double2 *up, const double2 *psi, const double2 *uz, int nroff) {
  const KitDev &k = c->kit;
  hipLaunchKernelGGL(k_tp2vec_combine_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, ur, up, psi, uz, k.ak,
                     k.nrdim, k.npdim, k.nzdim, k.nrchop + nroff, k.npchop, k.chopzl, k.chopzu);
}

void k_tp2vec_front(mlx_ctx *c, double2 *ur, double2 *up, double2 *uz, const double2 *p_in, const double2 *c_in,
                    const double *p_ln, const double *c_ln, int mode, int md_off, double *ln_r, double *ln_p,
                    double *ln_z, int ncols, int zlo, int zhi) {
  const KitDev &k = c->kit;
  if (k.nrchop + 3 > k.nrdim_ref || k.nrchop + 3 > k.ne) MLX_THROW(MLX_E_ARG, "chop_index: chopping in r too large");
  if (k.nrchop > 768) MLX_THROW(MLX_E_ARG, "vector_reconstruction: nrchop above 768 is not supported");
  if (md_off < 0 || md_off > 3 || k.nrchop + md_off > k.nrdim_ref)
    MLX_THROW(MLX_E_ARG, "curl_vector_reconstruction: chopping offset of chi out of range");
  // ncols > 0: columns m >= ncols are neither read nor written (their
  // inputs are zero and the caller's transforms skip them)
  dim3 grid((ncols > 0 && ncols < k.npdim) ? ncols : k.npdim, k.nzdim);
  size_t lds = 4 * sizeof(double2) * (size_t)k.nrdim;
  hipLaunchKernelGGL(k_tp2vec_front_, grid, dim3(256), lds, c->stream, ur, up, uz, p_in, c_in, p_ln, c_ln, mode,
                     md_off, k.d2c, k.xxc, k.lognorm, k.ak, k.nrdim, k.npdim, k.nzdim, k.ne, k.nrchop, k.npchop, k.chopzl,
                     k.chopzu, k.ell, ln_r, ln_p, ln_z, zlo, zhi);
  HIPCHK(hipGetLastError());
}

void k_tp2vec_front2(mlx_ctx *c, double2 *vr, double2 *vp, double2 *vz, double *lvr, double *lvp, double *lvz,
                     double2 *wr, double2 *wp, double2 *wz, double *lwr, double *lwp, double *lwz,
                     const double2 *psi, const double2 *chi, const double *psi_ln, const double *chi_ln,
                     int md_off, int ncols, int zlo, int zhi) {
  const KitDev &k = c->kit;
  if (k.nrchop + 3 > k.nrdim_ref || k.nrchop + 3 > k.ne) MLX_THROW(MLX_E_ARG, "chop_index: chopping in r too large");
  if (k.nrchop > 768) MLX_THROW(MLX_E_ARG, "vector_reconstruction: nrchop above 768 is not supported");
  if (md_off < 0 || md_off > 3 || k.nrchop + md_off > k.nrdim_ref)
    MLX_THROW(MLX_E_ARG, "curl_vector_reconstruction: chopping offset of chi out of range");
  dim3 grid((ncols > 0 && ncols < k.npdim) ? ncols : k.npdim, k.nzdim);
  size_t lds = 4 * sizeof(double2) * (size_t)k.nrdim;
  const FrontOut v{vr, vp, vz, lvr, lvp, lvz}, w{wr, wp, wz, lwr, lwp, lwz};
  // algorithmic bytes: psi and chi's active rows in, the six fields' rows of
  // the +3 offset out, over the computed columns and z slices
  const int nc = (ncols > 0 && ncols < k.npchop) ? ncols : k.npchop;
  const double nzk = (double)k.nzdim - ((zhi >= zlo) ? (double)(zhi - zlo + 1) : 0.0);
  ProbeScope probe(c, MLX_PROBE_TP_FRONT, 16.0 * nzk * (2.0 * active_pairs(k, 0, nc) + 6.0 * active_pairs(k, 3, nc)),
                   0.0);
  probed_launch(probe, k_tp2vec_front2_, grid, dim3(256), lds, c->stream, v, w, psi, chi, psi_ln, chi_ln, md_off,
                k.d2c, k.xxc, k.lognorm, k.ak, k.nrdim, k.npdim, k.nzdim, k.ne, k.nrchop, k.npchop, k.chopzl, k.chopzu,
                k.ell, zlo, zhi);
  HIPCHK(hipGetLastError());
}

void k_vecprod(mlx_ctx *c, double2 *vr, double2 *vp, double2 *vz, const double2 *ur, const double2 *up,
               const double2 *uz) {
  const KitDev &k = c->kit;
  hipLaunchKernelGGL(k_vecprod_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, vr, vp, vz, ur, up, uz, k.nr,
                     k.nph, k.nrdim, k.npdim, k.nzdim);
}

void k_vec2tp_up_ln(mlx_ctx *c, double2 *up, const double2 *inf, double *psiln) {
  hipLaunchKernelGGL(k_vec2tp_up_ln_, dim3(1), dim3(256), 0, c->stream, up, inf, psiln, c->kit.x, c->kit.nr);
}

void k_nonfinite(mlx_ctx *c, const double2 *a, const double2 *b, int *flag) {
  hipLaunchKernelGGL(k_nonfinite_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, a, b, c->nelem, flag);
}

void k_energy(mlx_ctx *c, const double2 *vr, const double2 *vp, const double2 *vz, double *out) {
  const KitDev &k = c->kit;
  const int nb = 1024;
  double *parts = c->dsmall;  // 1024 partial sums
  double cell = (2.0 * 3.14159265358979323846 / k.np) * (k.zlen / k.nz);
  hipLaunchKernelGGL(k_energy_, dim3(nb), dim3(kT), 0, c->stream, vr, vp, vz, k.x, k.w, k.nr, k.nph, k.nrdim,
                     k.npdim, k.nzdim, k.ell * k.ell, cell, parts);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(1024), 0, c->stream, parts, nb, out);
}

void k_vormag(mlx_ctx *c, double2 *out, const double2 *wr, const double2 *wp, const double2 *wz) {
  hipLaunchKernelGGL(k_vormag_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, out, wr, wp, wz, c->nelem);
}

void k_set_scalar(mlx_ctx *c, double *dst, double v) {
  hipLaunchKernelGGL(k_set_scalar_, dim3(1), dim3(64), 0, c->stream, dst, v);
}
void k_set_flag(mlx_ctx *c, int *dst, int v) {
  hipLaunchKernelGGL(k_set_flag_, dim3(1), dim3(64), 0, c->stream, dst, v);
}
void k_copy_scalar(mlx_ctx *c, double *dst, const double *src, double scale) {
  hipLaunchKernelGGL(k_copy_scalar_, dim3(1), dim3(64), 0, c->stream, dst, src, scale);
}

}  // namespace mlx
