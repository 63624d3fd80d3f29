// Kernels of the adjoint (transposed) operators: the transforms, the
// vector product, the Laplacian and the Helmholtz solve.  The reference has
// no adjoint; these are the transposes of this library's forward operators
// with respect to the Euclidean inner products of the stored arrays (the real
// and imaginary parts of every entry, plus the log coefficient ln).
//
// Transforms.  With P the normalised mapped-Legendre table and W the Gauss
// weights, the radial transforms are L_f = P^T W (rtrans_forward,
// mlegs_scalar_ops.f90:1852-1927) and L_b = P (rtrans_backward, :1931-2008),
// so L_b^T = L_f W^-1 and L_f^T = W L_b: the forward kernels serve the
// adjoint after a diagonal row scaling.  The complex z transforms satisfy
// Z_b^T = nz Z_f and Z_f^T = Z_b / nz.  The real phi transform of np samples
// to m = 0..np/2: Phi_b(C)_t = Re C_0 + 2 Re sum_{0<m<np/2} C_m e^{i m phi_t}
// + Re C_{np/2} (-1)^t, so Phi_b^T g = np mu Phi_f(g) and Phi_f^T G =
// Phi_b(G / (np mu)), mu_m = 1 at m = 0 and np/2 and 2 between, with the
// imaginary parts at m = 0 and np/2 (which Phi_b does not read on real
// fields) set to zero.  The log term enters the backward transform as
// e(i, 0, k) += ln lnx(i) (physical rows i < nr, every z plane), whose
// transpose adds sum lnx(i) Re g(i, 0, k) to the cotangent of ln.
#include "mlx_internal.h"

namespace {

constexpr int kT = 256;

inline int grid_for(size_t n) {
  size_t g = (n + kT - 1) / kT;
  return (int)(g < 8192 ? (g > 0 ? g : 1) : 8192);
}

// mode +1: after a forward phi transform (Phi_b^T): x np mu_m;
// mode -1: before a backward phi transform (Phi_f^T): / (np mu_m);
// the imaginary parts at m = 0 and m = np/2 are zeroed in both
__global__ void k_adj_phi_scale_(double2 *e, size_t n, int nrdim, int npdim, int nph, int np, int mode) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int m = (int)((t / nrdim) % npdim);
    const double mu = (m == 0 || m == nph) ? 1.0 : 2.0;
    const double f = mode > 0 ? np * mu : 1.0 / (np * mu);
    double2 v = e[t];
    v.x *= f;
    v.y = (m == 0 || m == nph) ? 0.0 : v.y * f;
    e[t] = v;
  }
}

// physical rows i < nr times w(i) (mode +1) or / w(i) (mode -1); rows >= nr zeroed
__global__ void k_adj_wscale_(double2 *e, size_t n, const double *w, int nr, int nrdim, int mode) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % nrdim);
    double2 v = e[t];
    if (i >= nr) {
      v = make_double2(0.0, 0.0);
    } else {
      const double f = mode > 0 ? w[i] : 1.0 / w[i];
      v.x *= f;
      v.y *= f;
    }
    e[t] = v;
  }
}

// ln_bar += sign sum_{k, i < nr} lnx(i) Re e(i, 0, k), one workgroup, fixed order
__global__ void __launch_bounds__(256) k_ln_adjoint_(const double2 *e, const double *lnx, double *ln_bar, double sign,
                                                     int nr, int nrdim, int npdim, int nz) {
  __shared__ double part[256];
  double s = 0.0;
  const int n = nr * nz;
  for (int t = threadIdx.x; t < n; t += 256) {
    const int i = t % nr, k = t / nr;
    s += lnx[i] * e[(size_t)k * nrdim * npdim + i].x;
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) ln_bar[0] += sign * part[0];
}

__global__ void k_adj_scale_(double2 *e, size_t n, double s) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    double2 v = e[t];
    v.x *= s;
    v.y *= s;
    e[t] = v;
  }
}

// zero outside the spectral support of a real, chopped field: rows
// n >= nrcs(m), columns m >= npc, the z slices between the chops, and the
// Nyquist column m = np/2 and slice kappa = nz/2
__global__ void k_adj_project_(double2 *e, size_t n, int nrdim, int npdim, int nz, int nrc, int npc, int nzc,
                               int nzcu, int nph, int nzh) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % nrdim);
    const int m = (int)((t / nrdim) % npdim);
    const int k = (int)(t / ((size_t)nrdim * npdim));
    const int nn = m < npc ? nrc - m : 0;
    bool keep = i < nn;
    if (nz > 1) {
      keep = keep && ((k + 1 <= nzc) || (k + 1 >= nzcu)) && k != nzh;
    }
    if (nph > 0 && m == nph) keep = false;
    if (!keep) e[t] = make_double2(0.0, 0.0);
  }
}

// out(j) = sum_i A(i, j) in(i), A = del2h band (plus -ak(k)^2 on the
// diagonal when akk): the transpose of the Laplacian's band product on rows
// j < nrcs(m) of every z slice inside the chops
__global__ void __launch_bounds__(64) k_del2_t_(double2 *out, const double2 *in, const double *d2c, const double *ak,
                                                int nrdim, int npdim, int nz, int ne, int nrc, int npc, int nzc,
                                                int nzcu, int with_ak) {
  const int m = blockIdx.x, k = blockIdx.y;
  const size_t base = (size_t)k * nrdim * npdim + (size_t)nrdim * m;
  const int nn = m < npc ? nrc - m : 0;
  const bool slice = nz == 1 || (k + 1 <= nzc) || (k + 1 >= nzcu);
  const double akk = with_ak ? ak[k] * ak[k] : 0.0;
  for (int j = threadIdx.x; j < nrdim; j += 64) {
    double2 s = make_double2(0.0, 0.0);
    if (slice && j < nn) {
      for (int i = j - 2; i <= j + 2; ++i) {
        if (i < 0 || i >= nn) continue;
        double a = d2c[((size_t)m * ne + i) * 5 + (j - i + 2)];
        if (i == j) a -= akk;
        const double2 x = in[base + i];
        s.x += a * x.x;
        s.y += a * x.y;
      }
    }
    out[base + j] = s;
  }
}

// ln_bar += scale sum_r c_r Re e(r, 0, 0), c = the log term's Laplacian
// coefficients of rows 0..2 (4/3, -2, 2/3) / ell^2 / exp(lognorm(r, 0))
__global__ void k_lap_ln_t_(const double2 *e, double *ln_bar, const double *lognorm, double ell, double scale) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double ell2 = ell * ell;
  const double s = 4.0 / 3.0 / ell2 / exp(lognorm[0]) * e[0].x - 2.0 / 1.0 / ell2 / exp(lognorm[1]) * e[1].x +
                   2.0 / 3.0 / ell2 / exp(lognorm[2]) * e[2].x;
  ln_bar[0] += scale * s;
}

constexpr int KL = 2, KU = 2, KV = 4, LDAB = 7;

// (del2 + alpha)^T of system (m, kk): entry (i, j) of the transpose is the
// band entry (j, i) of the forward matrix
__device__ inline double d2t_entry(const double *d2c, int m, int ne, int i, int j, double akk, double alpha) {
  const int d = i - j + 2;  // forward row j, column i
  if (d < 0 || d > 4) return 0.0;
  double v = d2c[((size_t)m * ne + j) * 5 + d];
  if (d == 2) v = v - akk + alpha;
  return v;
}

// one thread per (m, |kappa|) system: gbtf2 of the transposed matrix in
// LAPACK band storage (element (i, j) at ab[(KV + i - j) + LDAB j])
__global__ void k_lu_factor_t_(double *fac, int8_t *piv, const double *d2c, const double *ak, int nrdim, int nkk,
                               int ne, int nrc, int npc, double alpha, int *info) {
  const int kk = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (kk >= nkk || m >= npc) return;
  const int n = nrc - m;
  if (n <= 0) return;
  const size_t sys = (size_t)m * nkk + kk;
  double *ab = fac + sys * (size_t)nrdim * LDAB;
  int8_t *ip = piv + sys * (size_t)nrdim;
  const double akk = ak[kk] * ak[kk];
  for (int j = 0; j < n; ++j) {
    for (int r = 0; r < LDAB; ++r) ab[r + LDAB * j] = 0.0;
    for (int i = (j - KU > 0 ? j - KU : 0); i <= j + KL && i < n; ++i)
      ab[(KV + i - j) + LDAB * j] = d2t_entry(d2c, m, ne, i, j, akk, alpha);
  }
  int ju = 0;
  for (int j = 0; j < n; ++j) {
    const int km = (KL < n - 1 - j) ? KL : n - 1 - j;
    int jp = 0;
    double best = fabs(ab[KV + LDAB * j]);
    for (int r = 1; r <= km; ++r) {
      const double v = fabs(ab[KV + r + LDAB * j]);
      if (v > best) {
        best = v;
        jp = r;
      }
    }
    ip[j] = (int8_t)jp;
    if (ab[KV + jp + LDAB * j] != 0.0) {
      int jnew = j + KU + jp;
      if (jnew > n - 1) jnew = n - 1;
      if (jnew > ju) ju = jnew;
      if (jp != 0) {
        for (int c = j; c <= ju; ++c) {
          double *a = &ab[(KV + j - c) + LDAB * c];
          double *b = &ab[(KV + j + jp - c) + LDAB * c];
          const double t = *a;
          *a = *b;
          *b = t;
        }
      }
      if (km > 0) {
        const double rp = 1.0 / ab[KV + LDAB * j];
        for (int r = 1; r <= km; ++r) ab[KV + r + LDAB * j] *= rp;
        for (int c = j + 1; c <= ju; ++c) {
          const double u = ab[(KV + j - c) + LDAB * c];
          if (u != 0.0)
            for (int r = 1; r <= km; ++r) ab[(KV + j + r - c) + LDAB * c] -= ab[KV + r + LDAB * j] * u;
        }
      }
    } else {
      atomicExch(info, 1);
    }
  }
}

// one thread per (m, z slice): gbtrs with the factors of |kappa| on rows
// n < nrcs(m) of every slice inside the chops, each slice once
__global__ void k_lu_solve_t_(const double *fac, const int8_t *piv, double2 *e, int nrdim, int npdim, int nz,
                              int nkk, int nrc, int npc, int nzc, int nzcu) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (k >= nz || m >= npc) return;
  if (nz > 1 && !((k + 1 <= nzc) || (k + 1 >= nzcu))) return;
  const int n = nrc - m;
  if (n <= 0) return;
  const int kk = k <= nz - k ? k : nz - k;
  if (kk >= nkk) return;
  const size_t sys = (size_t)m * nkk + kk;
  const double *ab = fac + sys * (size_t)nrdim * LDAB;
  const int8_t *ip = piv + sys * (size_t)nrdim;
  double2 *b = e + (size_t)k * nrdim * npdim + (size_t)nrdim * m;
  for (int j = 0; j < n - 1; ++j) {
    const int lm = (KL < n - 1 - j) ? KL : n - 1 - j;
    const int jp = ip[j];
    if (jp != 0) {
      const double2 t = b[j];
      b[j] = b[j + jp];
      b[j + jp] = t;
    }
    const double2 bj = b[j];
    for (int r = 1; r <= lm; ++r) {
      const double l = ab[KV + r + LDAB * j];
      b[j + r].x -= l * bj.x;
      b[j + r].y -= l * bj.y;
    }
  }
  for (int j = n - 1; j >= 0; --j) {
    const double d = ab[KV + LDAB * j];
    double2 bj = b[j];
    bj.x /= d;
    bj.y /= d;
    b[j] = bj;
    const int i0 = j - KV > 0 ? j - KV : 0;
    for (int i = i0; i < j; ++i) {
      const double u = ab[(KV + i - j) + LDAB * j];
      b[i].x -= u * bj.x;
      b[i].y -= u * bj.y;
    }
  }
}

// the vector product's adjoint at (v, u): with a the cotangent of v x u,
// v_bar = u x a and u_bar = a x v, on each packed real sample
__global__ void k_vecprod_adj_(double2 *a1, double2 *a2, double2 *a3, const double2 *v1, const double2 *v2,
                               const double2 *v3, const double2 *u1, const double2 *u2, const double2 *u3,
                               double2 *ub1, double2 *ub2, double2 *ub3, size_t n) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const double2 A1 = a1[t], A2 = a2[t], A3 = a3[t];
    const double2 V1 = v1[t], V2 = v2[t], V3 = v3[t];
    const double2 U1 = u1[t], U2 = u2[t], U3 = u3[t];
    // v_bar = u x a
    a1[t] = make_double2(U2.x * A3.x - U3.x * A2.x, U2.y * A3.y - U3.y * A2.y);
    a2[t] = make_double2(U3.x * A1.x - U1.x * A3.x, U3.y * A1.y - U1.y * A3.y);
    a3[t] = make_double2(U1.x * A2.x - U2.x * A1.x, U1.y * A2.y - U2.y * A1.y);
    // u_bar = a x v
    ub1[t] = make_double2(A2.x * V3.x - A3.x * V2.x, A2.y * V3.y - A3.y * V2.y);
    ub2[t] = make_double2(A3.x * V1.x - A1.x * V3.x, A3.y * V1.y - A1.y * V3.y);
    ub3[t] = make_double2(A1.x * V2.x - A2.x * V1.x, A1.y * V2.y - A2.y * V1.y);
  }
}

}  // namespace

namespace mlx {

void k_adj_phi_scale(mlx_ctx *c, double2 *e, int mode) {
  const KitDev &k = c->kit;
  hipLaunchKernelGGL(k_adj_phi_scale_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, e, c->nelem, k.nrdim,
                     k.npdim, k.nph, k.np, mode);
}

void k_adj_wscale(mlx_ctx *c, double2 *e, int mode) {
  const KitDev &k = c->kit;
  hipLaunchKernelGGL(k_adj_wscale_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, e, c->nelem, k.w, k.nr,
                     k.nrdim, mode);
}

void k_ln_adjoint(mlx_ctx *c, const double2 *e, double *ln_bar, double sign) {
  const KitDev &k = c->kit;
  hipLaunchKernelGGL(k_ln_adjoint_, dim3(1), dim3(256), 0, c->stream, e, k.lnx, ln_bar, sign, k.nr, k.nrdim, k.npdim,
                     k.nzdim);
}

void k_adj_scale(mlx_ctx *c, double2 *e, double s) {
  hipLaunchKernelGGL(k_adj_scale_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, e, c->nelem, s);
}

void k_adj_project(mlx_ctx *c, double2 *e) {
  const KitDev &k = c->kit;
  hipLaunchKernelGGL(k_adj_project_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, e, c->nelem, k.nrdim,
                     k.npdim, k.nzdim, k.nrchop, k.npchop, k.chopzl, k.chopzu, k.np > 1 ? k.nph : -1,
                     k.nz > 1 ? k.nz / 2 : -1);
}

void k_del2_t(mlx_ctx *c, double2 *out, const double2 *in, bool with_ak) {
  const KitDev &k = c->kit;
  if (k.nrchop > k.ne) MLX_THROW(MLX_E_ARG, "del2 adjoint: table too small");
  hipLaunchKernelGGL(k_del2_t_, dim3(k.npdim, k.nzdim), dim3(64), 0, c->stream, out, in, k.d2c, k.ak, k.nrdim,
                     k.npdim, k.nzdim, k.ne, k.nrchop, k.npchop, k.chopzl, k.chopzu, with_ak ? 1 : 0);
}

void k_lap_ln_t(mlx_ctx *c, const double2 *e, double *ln_bar, double scale) {
  hipLaunchKernelGGL(k_lap_ln_t_, dim3(1), dim3(64), 0, c->stream, e, ln_bar, c->kit.lognorm, c->kit.ell, scale);
}

// (del2 + alpha)^-T per (m, kappa) system on rows n < nrcs(m), factors
// cached per alpha (two sets kept)
void k_helm_solve_t(mlx_ctx *c, double2 *e, double alpha) {
  const KitDev &k = c->kit;
  const int nkk = (k.nzdim > 1) ? k.nzdim / 2 + 1 : 1;
  AdjLu *hit = nullptr;
  for (AdjLu &a : c->adj_lu)
    if (a.alpha == alpha) hit = &a;
  if (!hit) {
    if (c->adj_lu.size() >= 2) {
      HIPCHK(hipStreamSynchronize(c->stream));
      (void)hipFree(c->adj_lu.front().fac);
      (void)hipFree(c->adj_lu.front().piv);
      c->adj_lu.erase(c->adj_lu.begin());
    }
    AdjLu a;
    a.alpha = alpha;
    const size_t nsys = (size_t)k.npchop * nkk;
    if (hipMalloc(&a.fac, sizeof(double) * nsys * k.nrdim * LDAB) != hipSuccess ||
        hipMalloc(&a.piv, nsys * k.nrdim) != hipSuccess) {
      if (a.fac) (void)hipFree(a.fac);
      MLX_THROW(MLX_E_ALLOC, "ihelm adjoint: cannot allocate LU factors");
    }
    int *dinfo = reinterpret_cast<int *>(c->dsmall + 4000);
    k_set_flag(c, dinfo, 0);
    hipLaunchKernelGGL(k_lu_factor_t_, dim3((nkk + 63) / 64, k.npchop), dim3(64), 0, c->stream, a.fac, a.piv, k.d2c,
                       k.ak, k.nrdim, nkk, k.ne, k.nrchop, k.npchop, alpha, dinfo);
    HIPCHK(hipGetLastError());
    int info = 0;
    HIPCHK(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (info != 0) {
      (void)hipFree(a.fac);
      (void)hipFree(a.piv);
      MLX_THROW(MLX_E_ARG, "lurc: lu factorization resulted in failure");
    }
    c->adj_lu.push_back(a);
    hit = &c->adj_lu.back();
  }
  hipLaunchKernelGGL(k_lu_solve_t_, dim3((k.nzdim + 63) / 64, k.npchop), dim3(64), 0, c->stream, hit->fac, hit->piv,
                     e, k.nrdim, k.npdim, k.nzdim, nkk, k.nrchop, k.npchop, k.chopzl, k.chopzu);
  HIPCHK(hipGetLastError());
}

void k_vecprod_adj(mlx_ctx *c, double2 *a1, double2 *a2, double2 *a3, const double2 *v1, const double2 *v2,
                   const double2 *v3, const double2 *u1, const double2 *u2, const double2 *u3, double2 *ub1,
                   double2 *ub2, double2 *ub3) {
  hipLaunchKernelGGL(k_vecprod_adj_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, a1, a2, a3, v1, v2, v3, u1,
                     u2, u3, ub1, ub2, ub3, c->nelem);
}

void adj_cache_free(mlx_ctx *c) {
  for (AdjLu &a : c->adj_lu) {
    (void)hipFree(a.fac);
    (void)hipFree(a.piv);
  }
  c->adj_lu.clear();
}

}  // namespace mlx
