// Kernels of the adjoint (transposed) operators: the transforms, the
// vector product, the Laplacian, the toroidal-poloidal reconstruction and
// projection, zeroat1 and the energy's gradient (the transposed Helmholtz
// and idel2 solves run on the banded LU of solve.hip).  The reference has
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

// one pass of the diagonal factors of a transform's transpose:
// e(i, m, k) *= s W(i) R(i) P(m), W = w(i) (wmode +1), 1/w(i) (-1) or 1 (0);
// R = r(i) (rmode +1), 1/r(i) (-1) or 1 (0); P = np mu_m (phimode +1),
// 1 / (np mu_m) (-1) or 1 (0), the imaginary parts at m = 0 and np/2 zeroed
// when phimode != 0; rows i >= nr zeroed when wmode or rmode is set
__global__ void k_adj_diag_(double2 *e, size_t n, const double *w, const double *r, int nr, int nrdim, int npdim,
                            int nph, int np, int wmode, int rmode, int phimode, double s) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % nrdim);
    const int m = (int)((t / nrdim) % npdim);
    double2 v = e[t];
    double f = s;
    bool zim = false;
    if (wmode != 0 || rmode != 0) {
      if (i >= nr) {
        f = 0.0;
      } else {
        if (wmode > 0) f *= w[i];
        if (wmode < 0) f /= w[i];
        if (rmode > 0) f *= r[i];
        if (rmode < 0) f /= r[i];
      }
    }
    if (phimode != 0) {
      const double mu = (m == 0 || m == nph) ? 1.0 : 2.0;
      f = phimode > 0 ? f * (np * mu) : f / (np * mu);
      zim = (m == 0 || m == nph);
    }
    v.x *= f;
    v.y = zim ? 0.0 : v.y * f;
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

// sb <- sb + g (data and log coefficient), kept on the chopped support
// (as k_adj_project_) and outside the de-aliased modes (columns m >= pcut,
// slices zcut .. zupper - 2), zero elsewhere: the sum, the projection and
// dealias in one pass
__global__ void k_adj_sum_mask_(double2 *sb, double *sbl, const double2 *g, const double *gl, size_t n, int nrdim,
                                int npdim, int nz, int nrc, int npc, int nzc, int nzcu, int nph, int nzh, int pcut,
                                int zcut, int zupper) {
  if (blockIdx.x == 0 && threadIdx.x == 0) sbl[0] += gl[0];
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % nrdim);
    const int m = (int)((t / nrdim) % npdim);
    const int k = (int)(t / ((size_t)nrdim * npdim));
    const int nn = m < npc ? nrc - m : 0;
    bool keep = i < nn;
    if (nz > 1) keep = keep && ((k + 1 <= nzc) || (k + 1 >= nzcu)) && k != nzh;
    if (nph > 0 && m == nph) keep = false;
    if ((pcut >= 0 && m >= pcut) || (zcut >= 0 && k >= zcut && k <= zupper - 2)) keep = false;
    if (keep) {
      const double2 a = sb[t], b = g[t];
      sb[t] = make_double2(a.x + b.x, a.y + b.y);
    } else {
      sb[t] = make_double2(0.0, 0.0);
    }
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

// the vector product's adjoint at (v, u): with a the cotangent of v x u,
// v_bar = u x a and u_bar = a x v, on each packed real sample
// (the product's samples: rows i < nr, packed columns < np/2; zero elsewhere).
// F32: v and u stored in FP32 at f32_index (index t32), as the forward's
// fused product reads them; ub (optional): added to v3 (the background
// axial flow)
template <bool F32>
__device__ inline double2 ld_vw(const double2 *p, size_t t, size_t t32) {
  if (F32) {
    const float2 f = reinterpret_cast<const float2 *>(p)[t32];
    return make_double2((double)f.x, (double)f.y);
  }
  return p[t];
}
template <bool F32>
__global__ void k_vecprod_adj_(double2 *a1, double2 *a2, double2 *a3, const double2 *v1, const double2 *v2,
                               const double2 *v3, const double2 *u1, const double2 *u2, const double2 *u3,
                               double2 *ub1, double2 *ub2, double2 *ub3, const double2 *bg, size_t n, int nr,
                               int nrdim, int npdim, int nph) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % nrdim), j = (int)((t / nrdim) % npdim);
    if (i >= nr || j >= nph) {
      const double2 z2 = make_double2(0.0, 0.0);
      a1[t] = a2[t] = a3[t] = z2;
      ub1[t] = ub2[t] = ub3[t] = z2;
      continue;
    }
    const size_t t32 = F32 ? f32_index(t / ((size_t)nrdim * npdim), j, i, nrdim, npdim) : t;
    const double2 A1 = a1[t], A2 = a2[t], A3 = a3[t];
    const double2 V1 = ld_vw<F32>(v1, t, t32), V2 = ld_vw<F32>(v2, t, t32);
    double2 V3 = ld_vw<F32>(v3, t, t32);
    if (bg) {
      const double2 g = bg[t];
      V3 = make_double2(V3.x + g.x, V3.y + g.y);
    }
    const double2 U1 = ld_vw<F32>(u1, t, t32), U2 = ld_vw<F32>(u2, t, t32), U3 = ld_vw<F32>(u3, t, t32);
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

// ---- tp2vec transpose ----------------------------------------------------------
// Forward (vector_reconstruction, mlegs_scalar_ops.f90:1455-1545), per (m,
// kappa) column of a kept slice on rows j < nn = nrcs(m), before the radial
// synthesis and the 1/r weight:
//   ur_s = i m psi + i kappa X chi,  up_s = -X psi - m kappa chi,  uz_s = -H chi
// with X the xxdx band (offsets -1, +1) and H the del2h band; psi.ln enters
// up_s and chi.ln enters uz_s through the bands' log rows of column (0, 0).
// The transpose:
//   psi_b = -i m ur_b - X^T up_b,  chi_b = X^T (-i kappa ur_b) - m kappa up_b - H^T uz_b
// on rows j < nn of the kept slices, zero elsewhere.
// zlo..zhi (zlo > zhi: none) and columns m >= npc: taken as zero (written as
// zero, their inputs not read), for cotangents that are de-aliased after
__global__ void __launch_bounds__(64) k_tp2vec_t_(double2 *psib, double2 *chib, const double2 *urb,
                                                  const double2 *upb, const double2 *uzb, const double *d2c,
                                                  const double *xxc, const double *ak, int nrdim, int npdim, int nz,
                                                  int ne, int nrc, int npc, int nzc, int nzcu, int zlo, int zhi) {
  const int m = blockIdx.x, k = blockIdx.y;
  const size_t base = (size_t)k * nrdim * npdim + (size_t)nrdim * m;
  const int nn = (m < npc && nrc - m > 0) ? nrc - m : 0;
  const bool slice = (nz == 1 || (k + 1 <= nzc) || (k + 1 >= nzcu)) && !(k >= zlo && k <= zhi);
  const double kv = ak[k], mv = (double)m;
  for (int j = threadIdx.x; j < nrdim; j += 64) {
    double2 pb = make_double2(0.0, 0.0), cb = pb;
    if (slice && j < nn) {
      const double2 r = urb[base + j], p = upb[base + j];
      pb = make_double2(mv * r.y, -mv * r.x);                 // -i m ur_b
      cb = make_double2(-mv * kv * p.x, -mv * kv * p.y);      // -m kappa up_b
      // X(i, j) for i = j - 1 (its +1 entry) and i = j + 1 (its -1 entry)
      for (int s = -1; s <= 1; s += 2) {
        const int i = j + s;
        if (i < 0 || i >= nn) continue;
        const double a = xxc[((size_t)m * ne + i) * 2 + (s < 0 ? 1 : 0)];
        const double2 pi = upb[base + i], ri = urb[base + i];
        pb.x -= a * pi.x;
        pb.y -= a * pi.y;
        cb.x += a * kv * ri.y;  // a (-i kappa) ur_b(i)
        cb.y -= a * kv * ri.x;
      }
      for (int i = j - 2; i <= j + 2; ++i) {
        if (i < 0 || i >= nn) continue;
        const double a = d2c[((size_t)m * ne + i) * 5 + (j - i + 2)];
        const double2 zi = uzb[base + i];
        cb.x -= a * zi.x;
        cb.y -= a * zi.y;
      }
    }
    psib[base + j] = pb;
    chib[base + j] = cb;
  }
}

// the log rows: xxdx adds psi.ln / exp(lognorm(n, 0)) to rows n = 0, 1 of
// column (0, 0) and up_s = -X psi; del2h adds (4/3, -2, 2/3) chi.ln / ell^2 /
// exp(lognorm(n, 0)) to rows 0..2 and uz_s = -H chi
__global__ void k_tp2vec_t_ln_(const double2 *upb, const double2 *uzb, double *psib_ln, double *chib_ln,
                               const double *lognorm, double ell) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double ell2 = ell * ell;
  psib_ln[0] = -(upb[0].x / exp(lognorm[0]) + upb[1].x / exp(lognorm[1]));
  chib_ln[0] = -(4.0 / 3.0 / ell2 / exp(lognorm[0]) * uzb[0].x - 2.0 / 1.0 / ell2 / exp(lognorm[1]) * uzb[1].x +
                 2.0 / 3.0 / ell2 / exp(lognorm[2]) * uzb[2].x);
}

// ---- vec2tp projection transpose ------------------------------------------------
// Forward (vector_projection's five products, mlegs_scalar_ops.f90:1387-1436;
// the plain FP64 kernel k_v2t_f64), per (m < npc, kept slice kappa) and table
// row i < M = nrcs(m) + 3, with s_i = +1 (i even) or -1 (i odd), x < nr/2,
// x~ = nr - 1 - x:
//   A = V (ur(x) + s ur(x~)),  B = D (up(x) - s up(x~)),  C = D (ur(x) - s ur(x~)),
//   Dd = V (up(x) + s up(x~)),  E = T (uz(x) + s uz(x~))   (A = Dd = 0 at m = 0)
//   psi_i = -i m A - B,  chi_i = i kappa C + m kappa Dd - E
// The transpose, from the cotangents of psi and chi:
//   Ab = i m psib, Bb = -psib, Cb = -i kappa chib, Db = m kappa chib, Eb = -chib
//   ur(x) += V Ab + D Cb, ur(x~) += s (V Ab - D Cb), up(x) += D Bb + V Db,
//   up(x~) += s (V Db - D Bb), uz(x) += T Eb, uz(x~) += s T Eb
// One workgroup per (m, kappa) column, one thread per x; rows staged in LDS
// in chunks.  Every entry of the outputs is written (zero outside).
// One workgroup of 256 threads per (m, kappa) column: RG = 256 / NX row
// groups of NX threads (NX = nr/2 rounded up to 64, 128 or 256), one thread
// per x; group g takes the table rows i = g, g + RG, ... of each staged
// chunk, and the groups' partial sums meet in LDS at the end.
constexpr int V2T_CH = 64;
template <int NX>
__global__ void __launch_bounds__(256) k_v2t_t_(double2 *urb, double2 *upb, double2 *uzb, const double2 *psib,
                                                const double2 *chib, const double *tV, const double *tD,
                                                const double *tT, const double *ak, int nrdim, int npdim, int nz,
                                                int nr, int nrh, int nt, int nrc3, int npc, int nzc, int nzcu,
                                                int zlo, int zhi) {
  constexpr int RG = 256 / NX;
  __shared__ double2 sQ[5][V2T_CH];
  __shared__ double2 sR[RG > 1 ? RG : 1][6][RG > 1 ? NX : 1];
  const int m = blockIdx.x, k = blockIdx.y, tid = threadIdx.x, x = tid % NX, grp = tid / NX;
  const size_t base = (size_t)k * nrdim * npdim + (size_t)nrdim * m;
  int M = (m < npc && nrc3 - m > 0) ? nrc3 - m : 0;
  if (M > nt) M = nt;
  const bool slice = (nz == 1 || (k + 1 <= nzc) || (k + 1 >= nzcu)) && !(k >= zlo && k <= zhi);
  if (!slice) M = 0;
  const double kv = ak[k], mv = (double)m;
  const bool mz = (m == 0);
  const size_t tb = (size_t)m * nt * nrh;
  const double2 z2 = make_double2(0.0, 0.0);
  double2 acc[6] = {z2, z2, z2, z2, z2, z2};  // ur(x), ur(x~), up(x), up(x~), uz(x), uz(x~)
  for (int i0 = 0; i0 < M; i0 += V2T_CH) {
    __syncthreads();
    if (tid < V2T_CH) {
      const int i = i0 + tid;
      double2 ps = z2, cs = z2;
      if (i < M) {
        ps = psib[base + i];
        cs = chib[base + i];
      }
      sQ[0][tid] = mz ? z2 : make_double2(-mv * ps.y, mv * ps.x);  // i m psib
      sQ[1][tid] = make_double2(-ps.x, -ps.y);
      sQ[2][tid] = make_double2(kv * cs.y, -kv * cs.x);            // -i kappa chib
      sQ[3][tid] = mz ? z2 : make_double2(mv * kv * cs.x, mv * kv * cs.y);
      sQ[4][tid] = make_double2(-cs.x, -cs.y);
    }
    __syncthreads();
    if (x < nrh) {
      const int iend = (M - i0 < V2T_CH) ? M - i0 : V2T_CH;
#pragma unroll 4
      for (int t = grp; t < iend; t += RG) {
        const int i = i0 + t;
        const size_t o = tb + (size_t)i * nrh + x;
        const double v = tV[o], d = tD[o], tt = tT[o];
        const double sg = (i & 1) ? -1.0 : 1.0;
        const double2 A = sQ[0][t], B = sQ[1][t], C = sQ[2][t], D = sQ[3][t], E = sQ[4][t];
        const double vax = v * A.x, vay = v * A.y, dcx = d * C.x, dcy = d * C.y;
        const double dbx = d * B.x, dby = d * B.y, vdx = v * D.x, vdy = v * D.y;
        const double tex = tt * E.x, tey = tt * E.y;
        acc[0].x += vax + dcx;
        acc[0].y += vay + dcy;
        acc[1].x += sg * (vax - dcx);
        acc[1].y += sg * (vay - dcy);
        acc[2].x += dbx + vdx;
        acc[2].y += dby + vdy;
        acc[3].x += sg * (vdx - dbx);
        acc[3].y += sg * (vdy - dby);
        acc[4].x += tex;
        acc[4].y += tey;
        acc[5].x += sg * tex;
        acc[5].y += sg * tey;
      }
    }
  }
  if (RG > 1) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 6; ++q) sR[grp][q][x] = acc[q];
    __syncthreads();
    if (grp == 0) {
      for (int g = 1; g < RG; ++g)
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          acc[q].x += sR[g][q][x].x;
          acc[q].y += sR[g][q][x].y;
        }
    }
  }
  for (int i = nr + tid; i < nrdim; i += 256) {
    urb[base + i] = z2;
    upb[base + i] = z2;
    uzb[base + i] = z2;
  }
  if (grp == 0 && x < nrh) {
    urb[base + x] = acc[0];
    urb[base + nr - 1 - x] = acc[1];
    upb[base + x] = acc[2];
    upb[base + nr - 1 - x] = acc[3];
    uzb[base + x] = acc[4];
    uzb[base + nr - 1 - x] = acc[5];
  }
}

// vec2tp's psi.ln = -1/2 Re calcat1(up)(kappa = 0) (:1371-1374) and the term
// up(i, 0, 0) += psi.ln (1 + x_i), i < nr (:1377-1379): given the cotangent of
// the PFF up (upP) and the FFF up (upF), upF(n, 0, 0) gains
// -1/2 at1(n) sum_i (1 + x_i) Re upP(i, 0, 0) on rows n < nrchop.  One workgroup.
__global__ void __launch_bounds__(256) k_v2t_up_ln_t_(double2 *upF, const double2 *upP, const double *x,
                                                      const double *at1, int nr, int n) {
  __shared__ double part[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < nr; i += 256) s += (1.0 + x[i]) * upP[i].x;
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  const double lb = part[0];
  for (int i = threadIdx.x; i < n; i += 256) upF[i].x += -0.5 * at1[i] * lb;
}

// zeroat1 (mlegs_scalar_ops.f90:311-325) maps column m = 0 of each slice by
// e(0) <- e(0) - sum_{n < N} at1(n) e(n) / at1(0); the transpose:
// e(n) <- e(n) - at1(n) e(0) / at1(0) for n < N (e(0) becomes zero)
__global__ void k_zeroat1_t_(double2 *e, const double *at1, int n, int nrdim, int npdim) {
  double2 *c = e + (size_t)blockIdx.x * nrdim * npdim;
  const double2 t = c[0];
  __syncthreads();
  const double d = at1[0];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double f = at1[i] / d;
    double2 v = c[i];
    v.x -= f * t.x;
    v.y -= f * t.y;
    c[i] = v;
  }
}

// the gradient of the kinetic energy 0.5 sum w_i ell^2/(1-x_i)^2 (2pi/np)
// (zlen/nz) |v|^2 of a PPP velocity (k_energy_'s quadrature) with respect to
// the stored samples: v <- cell w_i ell^2/(1-x_i)^2 v on rows i < nr and
// packed columns < np/2, zero elsewhere
__global__ void k_energy_grad_(double2 *vr, double2 *vp, double2 *vz, const double *x, const double *w, double ell,
                               double cell, int nr, int nrdim, int npdim, int nph, size_t n) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % nrdim);
    const int j = (int)((t / nrdim) % npdim);
    double f = 0.0;
    if (i < nr && j < nph) {
      const double om = 1.0 - x[i];
      f = cell * w[i] * ell * ell / (om * om);
    }
    double2 a = vr[t], b = vp[t], c = vz[t];
    vr[t] = make_double2(f * a.x, f * a.y);
    vp[t] = make_double2(f * b.x, f * b.y);
    vz[t] = make_double2(f * c.x, f * c.y);
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

void k_adj_diag(mlx_ctx *c, double2 *e, int wmode, int rmode, int phimode, double s) {
  const KitDev &k = c->kit;
  hipLaunchKernelGGL(k_adj_diag_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, e, c->nelem, k.w, k.r, k.nr,
                     k.nrdim, k.npdim, k.nph, k.np, wmode, rmode, phimode, s);
}

void k_adj_project(mlx_ctx *c, double2 *e) {
  const KitDev &k = c->kit;
  hipLaunchKernelGGL(k_adj_project_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, e, c->nelem, k.nrdim,
                     k.npdim, k.nzdim, k.nrchop, k.npchop, k.chopzl, k.chopzu, k.np > 1 ? k.nph : -1,
                     k.nz > 1 ? k.nz / 2 : -1);
}

void k_adj_sum_mask(mlx_ctx *c, double2 *sb, double *sbl, const double2 *g, const double *gl) {
  const KitDev &k = c->kit;
  int pcut = -1, zcut = -1, zupper = 0;
  if (k.np > 1) pcut = (k.np / 3 + 1 > 1) ? k.np / 3 + 1 : 1;
  if (k.nz > 1) {
    zcut = (k.nz / 3 + 1 > 1) ? k.nz / 3 + 1 : 1;
    zupper = k.nz - zcut + 2;
  }
  hipLaunchKernelGGL(k_adj_sum_mask_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, sb, sbl, g, gl, c->nelem,
                     k.nrdim, k.npdim, k.nzdim, k.nrchop, k.npchop, k.chopzl, k.chopzu, k.np > 1 ? k.nph : -1,
                     k.nz > 1 ? k.nz / 2 : -1, pcut, zcut, zupper);
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

// (del2 + alpha)^-T per (m, kappa) system on rows n < nrcs(m) + nroff;
// kind 1: idel2_proln's systems (alpha = 0, the (0,0) system bordered).  The
// transposed matrices are factored and solved by the banded LU of solve.hip
// (kind + 4), over every column and kept slice.
void k_helm_solve_t(mlx_ctx *c, double2 *e, double alpha, int kind, int nroff, int pcut, int zlo, int zhi,
                    double2 *e2) {
  k_helm_solve(c, e, alpha, kind | 4, nroff, pcut, zlo, zhi, e2);
}

// abcn's transpose after the solve: nlbar = c1 y, nlpbar = c2 y, then
// y = a y + c3 t (t = del2^T y), one pass
__global__ void k_abcn_adj_out_(double2 *y, const double2 *t, double2 *nlbar, double2 *nlpbar, size_t n, double c1,
                                double c2, double a, double c3) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double2 v = y[i], w = t[i];
    nlbar[i] = make_double2(c1 * v.x, c1 * v.y);
    nlpbar[i] = make_double2(c2 * v.x, c2 * v.y);
    y[i] = make_double2(a * v.x + c3 * w.x, a * v.y + c3 * w.y);
  }
}

// abcn's transpose, the log terms (one thread, before y's update): with
// q(y) = (4/3, -2, 2/3) . Re y(0..2, 0, 0) / ell^2 / exp(lognorm(0..2)),
// L = ysc - q(y), nl = 1.5 dt L, nlp = -0.5 dt L, ysc = L + c2 q(y)
__global__ void k_abcn_adj_ln_(const double2 *y, double *ysc, double *nlln, double *nlpln, const double *lognorm,
                               double ell, double dt, double c2) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double ell2 = ell * ell;
  const double q = 4.0 / 3.0 / ell2 / exp(lognorm[0]) * y[0].x - 2.0 / 1.0 / ell2 / exp(lognorm[1]) * y[1].x +
                   2.0 / 3.0 / ell2 / exp(lognorm[2]) * y[2].x;
  const double L = ysc[0] - q;
  nlln[0] = 1.5 * dt * L;
  nlpln[0] = -0.5 * dt * L;
  ysc[0] = L + c2 * q;
}

void k_abcn_adj_out(mlx_ctx *c, double2 *y, const double2 *t, double2 *nlbar, double2 *nlpbar, double c1, double c2,
                    double a, double c3) {
  hipLaunchKernelGGL(k_abcn_adj_out_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, y, t, nlbar, nlpbar, c->nelem,
                     c1, c2, a, c3);
}

void k_abcn_adj_ln(mlx_ctx *c, const double2 *y, double *ysc, double *nlln, double *nlpln, double dt, double c2) {
  hipLaunchKernelGGL(k_abcn_adj_ln_, dim3(1), dim3(64), 0, c->stream, y, ysc, nlln, nlpln, c->kit.lognorm, c->kit.ell,
                     dt, c2);
}

void k_tp2vec_t(mlx_ctx *c, double2 *psib, double2 *chib, const double2 *urb, const double2 *upb, const double2 *uzb,
                double *psib_ln, double *chib_ln, int ncols, int zlo, int zhi) {
  const KitDev &k = c->kit;
  if (k.nrchop > k.ne || k.nrchop < 3) MLX_THROW(MLX_E_ARG, "tp2vec adjoint: table too small");
  const int npc = (ncols > 0 && ncols < k.npchop) ? ncols : k.npchop;
  hipLaunchKernelGGL(k_tp2vec_t_, dim3(k.npdim, k.nzdim), dim3(64), 0, c->stream, psib, chib, urb, upb, uzb, k.d2c,
                     k.xxc, k.ak, k.nrdim, k.npdim, k.nzdim, k.ne, k.nrchop, npc, k.chopzl, k.chopzu, zlo, zhi);
  hipLaunchKernelGGL(k_tp2vec_t_ln_, dim3(1), dim3(64), 0, c->stream, upb, uzb, psib_ln, chib_ln, k.lognorm, k.ell);
  HIPCHK(hipGetLastError());
}

void k_v2t_t(mlx_ctx *c, double2 *urb, double2 *upb, double2 *uzb, const double2 *psib, const double2 *chib,
             int ncols, int zlo, int zhi) {
  const KitDev &k = c->kit;
  if (k.nrh > 256 || (k.nr & 1)) MLX_THROW(MLX_E_ARG, "vec2tp adjoint: nr must be even and at most 512");
  if (k.nt < k.nrchop + 3) MLX_THROW(MLX_E_ARG, "vec2tp adjoint: projection tables too small");
  const int npc = (ncols > 0 && ncols < k.npchop) ? ncols : k.npchop;
#define MLX_V2TT(NX_)                                                                                            \
  hipLaunchKernelGGL(k_v2t_t_<NX_>, dim3(k.npdim, k.nzdim), dim3(256), 0, c->stream,                             \
                     urb, upb, uzb, psib, chib, k.tV, k.tD, k.tT, k.ak, k.nrdim, k.npdim, k.nzdim, k.nr, k.nrh, k.nt,  \
                     k.nrchop + 3, npc, k.chopzl, k.chopzu, zlo, zhi)
  if (k.nrh <= 64) MLX_V2TT(64);
  else if (k.nrh <= 128) MLX_V2TT(128);
  else MLX_V2TT(256);
#undef MLX_V2TT
  HIPCHK(hipGetLastError());
}

void k_v2t_up_ln_t(mlx_ctx *c, double2 *upF, const double2 *upP) {
  const KitDev &k = c->kit;
  hipLaunchKernelGGL(k_v2t_up_ln_t_, dim3(1), dim3(256), 0, c->stream, upF, upP, k.x, k.at1, k.nr,
                     k.nrchop < k.nrdim ? k.nrchop : k.nrdim);
  HIPCHK(hipGetLastError());
}

void k_zeroat1_t(mlx_ctx *c, double2 *e) {
  const KitDev &k = c->kit;
  const int n = k.nrchop < k.nrdim ? k.nrchop : k.nrdim;
  hipLaunchKernelGGL(k_zeroat1_t_, dim3(k.nzdim), dim3(kT), 0, c->stream, e, k.at1, n, k.nrdim, k.npdim);
  HIPCHK(hipGetLastError());
}

void k_energy_grad(mlx_ctx *c, double2 *vr, double2 *vp, double2 *vz) {
  const KitDev &k = c->kit;
  const double cell = (2.0 * 3.14159265358979323846 / k.np) * (k.zlen / k.nz);
  hipLaunchKernelGGL(k_energy_grad_, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, vr, vp, vz, k.x, k.w, k.ell,
                     cell, k.nr, k.nrdim, k.npdim, k.nph, c->nelem);
  HIPCHK(hipGetLastError());
}

void k_vecprod_adj(mlx_ctx *c, double2 *a1, double2 *a2, double2 *a3, const double2 *v1, const double2 *v2,
                   const double2 *v3, const double2 *u1, const double2 *u2, const double2 *u3, double2 *ub1,
                   double2 *ub2, double2 *ub3, bool f32, const double2 *bg) {
  const KitDev &k = c->kit;
  if (f32)
    hipLaunchKernelGGL(k_vecprod_adj_<true>, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, a1, a2, a3, v1, v2, v3,
                       u1, u2, u3, ub1, ub2, ub3, bg, c->nelem, k.nr, k.nrdim, k.npdim, k.nph);
  else
    hipLaunchKernelGGL(k_vecprod_adj_<false>, dim3(grid_for(c->nelem)), dim3(kT), 0, c->stream, a1, a2, a3, v1, v2,
                       v3, u1, u2, u3, ub1, ub2, ub3, bg, c->nelem, k.nr, k.nrdim, k.npdim, k.nph);
}



}  // namespace mlx
