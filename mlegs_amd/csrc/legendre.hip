// Mapped-Legendre radial transforms and the vec2tp projection contractions.
//
// Forward (rtrans_forward, src/submodules/mlegs_scalar_ops.f90:1852-1927):
//   for every m < npc and parity p, rows n = m + 2a + p < m + nrcs(m):
//     e(2a+p, m, k) = sum_{x<nr/2} pf(x, 2a+p, m) * fold_p(f)(x, m, k),
//     fold_0 = (f(x) + f(-x)) w(x),  fold_1 = (f(x) - f(-x)) w(x),
//   with f(-x_i) = e(nr-1-i); every other entry of the output is zero.
// Backward (rtrans_backward, :1931-2008):
//   be/bo(x, k) = sum_a pf(x, 2a+p, m) e(2a+p, m, k);
//   f(x) = be + bo,  f(-x) = be - bo, rows >= nr and columns >= npc zero.
// The reference does both through complex zgemm after promoting the real
// table to complex (mlegs_genmat_ops.f90:41-57); here the real table
// multiplies interleaved complex data directly (half the flops), all m and
// both parities in one launch over a table of (m, parity, row-tile) items,
// with operand tiles staged in LDS.
//
// vec2tp (vector_projection, :1387-1436): five parity-folded contractions
// per (m, k) against the precomputed V, D, T tables, combined in the
// epilogue into psi and chi.
#include "mlx_internal.h"

namespace {

constexpr int BM = 32, BN = 32, BK = 32;

// ---------------------------------------------------------------------------
// FP64 forward: work item = (m, p, a0, M)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_leg_fwd_f64(const double2 *__restrict__ in, double2 *__restrict__ out,
                                                     const double *__restrict__ pf, const double *__restrict__ w,
                                                     const int4 *__restrict__ work, int nrdim, int npdim,
                                                     int nz, int nr, int nrh, int ne) {
  const int4 it = work[blockIdx.x];
  const int m = it.x, p = it.y, a0 = it.z, M = it.w;
  const int k0 = blockIdx.y * BN;
  __shared__ double As[BK][BM + 1];
  __shared__ double2 Bs[BK][BN + 1];
  const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
  const size_t zs = (size_t)nrdim * npdim;
  const double *P = pf + (size_t)m * ne * nrh;
  const double2 *col = in + (size_t)nrdim * m;
  double2 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = make_double2(0.0, 0.0);

  for (int x0 = 0; x0 < nrh; x0 += BK) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int idx = tid + 256 * q;
      int x = idx % BK, a = idx / BK;
      int arow = a0 + a, xg = x0 + x;
      double v = 0.0;
      if (arow < M && xg < nrh) v = P[(size_t)(2 * arow + p) * nrh + xg];
      As[x][a] = v;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int idx = tid + 256 * q;
      int x = idx % BK, kk = idx / BK;
      int kg = k0 + kk, xg = x0 + x;
      double2 v = make_double2(0.0, 0.0);
      if (kg < nz && xg < nrh) {
        double2 a = col[(size_t)kg * zs + xg];
        double2 b = col[(size_t)kg * zs + (nr - 1 - xg)];
        double wx = w[xg];
        v = (p == 0) ? make_double2((a.x + b.x) * wx, (a.y + b.y) * wx)
                     : make_double2((a.x - b.x) * wx, (a.y - b.y) * wx);
      }
      Bs[x][kk] = v;
    }
    __syncthreads();
#pragma unroll 8
    for (int xx = 0; xx < BK; ++xx) {
      double a0v = As[xx][tx], a1v = As[xx][tx + 16];
      double2 b0 = Bs[xx][ty], b1 = Bs[xx][ty + 16];
      acc[0][0].x += a0v * b0.x; acc[0][0].y += a0v * b0.y;
      acc[0][1].x += a0v * b1.x; acc[0][1].y += a0v * b1.y;
      acc[1][0].x += a1v * b0.x; acc[1][0].y += a1v * b0.y;
      acc[1][1].x += a1v * b1.x; acc[1][1].y += a1v * b1.y;
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      int a = a0 + tx + 16 * i, kg = k0 + ty + 16 * j;
      if (a < M && kg < nz) out[(size_t)kg * zs + (size_t)nrdim * m + (2 * a + p)] = acc[i][j];
    }
}

// ---------------------------------------------------------------------------
// FP64 backward: work item = (m, nn, x0, 0)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_leg_bwd_f64(const double2 *__restrict__ in, double2 *__restrict__ out,
                                                     const double *__restrict__ pf, const int4 *__restrict__ work,
                                                     int nrdim, int npdim, int nz, int nr, int nrh, int ne) {
  const int4 it = work[blockIdx.x];
  const int m = it.x, nn = it.y, x0 = it.z;
  const int k0 = blockIdx.y * BN;
  const int Me = (nn + 1) / 2, Mo = nn / 2;
  __shared__ double Ae[BK][BM + 1], Ao[BK][BM + 1];
  __shared__ double2 Be[BK][BN + 1], Bo[BK][BN + 1];
  const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
  const size_t zs = (size_t)nrdim * npdim;
  const double *P = pf + (size_t)m * ne * nrh;
  const double2 *col = in + (size_t)nrdim * m;
  double2 ce[2][2], co[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      ce[i][j] = make_double2(0.0, 0.0);
      co[i][j] = make_double2(0.0, 0.0);
    }
  for (int a0 = 0; a0 < Me; a0 += BK) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int idx = tid + 256 * q;
      int x = idx % BM, a = idx / BM;
      int ag = a0 + a, xg = x0 + x;
      double ve = 0.0, vo = 0.0;
      if (xg < nrh) {
        if (ag < Me) ve = P[(size_t)(2 * ag) * nrh + xg];
        if (ag < Mo) vo = P[(size_t)(2 * ag + 1) * nrh + xg];
      }
      Ae[a][x] = ve;
      Ao[a][x] = vo;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int idx = tid + 256 * q;
      int a = idx % BK, kk = idx / BK;
      int ag = a0 + a, kg = k0 + kk;
      double2 ve = make_double2(0.0, 0.0), vo = make_double2(0.0, 0.0);
      if (kg < nz) {
        if (ag < Me) ve = col[(size_t)kg * zs + 2 * ag];
        if (ag < Mo) vo = col[(size_t)kg * zs + 2 * ag + 1];
      }
      Be[a][kk] = ve;
      Bo[a][kk] = vo;
    }
    __syncthreads();
#pragma unroll 4
    for (int aa = 0; aa < BK; ++aa) {
      double e0 = Ae[aa][tx], e1 = Ae[aa][tx + 16];
      double o0 = Ao[aa][tx], o1 = Ao[aa][tx + 16];
      double2 be0 = Be[aa][ty], be1 = Be[aa][ty + 16];
      double2 bo0 = Bo[aa][ty], bo1 = Bo[aa][ty + 16];
      ce[0][0].x += e0 * be0.x; ce[0][0].y += e0 * be0.y;
      ce[0][1].x += e0 * be1.x; ce[0][1].y += e0 * be1.y;
      ce[1][0].x += e1 * be0.x; ce[1][0].y += e1 * be0.y;
      ce[1][1].x += e1 * be1.x; ce[1][1].y += e1 * be1.y;
      co[0][0].x += o0 * bo0.x; co[0][0].y += o0 * bo0.y;
      co[0][1].x += o0 * bo1.x; co[0][1].y += o0 * bo1.y;
      co[1][0].x += o1 * bo0.x; co[1][0].y += o1 * bo0.y;
      co[1][1].x += o1 * bo1.x; co[1][1].y += o1 * bo1.y;
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      int xg = x0 + tx + 16 * i, kg = k0 + ty + 16 * j;
      if (xg < nrh && kg < nz) {
        double2 *o = out + (size_t)kg * zs + (size_t)nrdim * m;
        o[xg] = make_double2(ce[i][j].x + co[i][j].x, ce[i][j].y + co[i][j].y);
        o[nr - 1 - xg] = make_double2(ce[i][j].x - co[i][j].x, ce[i][j].y - co[i][j].y);
      }
    }
}

// ---------------------------------------------------------------------------
// FP64 vec2tp projection: work item = (m, p, a0, M) with M rows of parity p
// among the nn3 = nrchop + 3 - m projected rows.
// ---------------------------------------------------------------------------
constexpr int VBM = 32, VBN = 16, VBK = 16;

__global__ void __launch_bounds__(256) k_v2t_f64(const double2 *__restrict__ ur, const double2 *__restrict__ up,
                                                 const double2 *__restrict__ uz, double2 *__restrict__ psi,
                                                 double2 *__restrict__ chi, const double *__restrict__ tV,
                                                 const double *__restrict__ tD, const double *__restrict__ tT,
                                                 const double *__restrict__ ak, const int4 *__restrict__ work,
                                                 int nrdim, int npdim, int nz, int nr, int nrh, int nt, int nzc,
                                                 int nzcu) {
  const int4 it = work[blockIdx.x];
  const int m = it.x, p = it.y, a0 = it.z, M = it.w;
  const int k0 = blockIdx.y * VBN;
  __shared__ double Vs[VBK][VBM + 1], Ds[VBK][VBM + 1], Ts[VBK][VBM + 1];
  __shared__ double2 U[5][VBK][VBN + 1];  // eo(ur), oe(ur), eo(up), oe(up), eo(uz)
  const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
  const size_t zs = (size_t)nrdim * npdim;
  const size_t tb = (size_t)m * nt * nrh;
  const double2 *cr = ur + (size_t)nrdim * m;
  const double2 *cp = up + (size_t)nrdim * m;
  const double2 *cz = uz + (size_t)nrdim * m;
  double2 acc[2][5];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int q = 0; q < 5; ++q) acc[i][q] = make_double2(0.0, 0.0);
  const double seo = (p == 0) ? 1.0 : -1.0;  // eo(p): even fold for p = 0
  for (int x0 = 0; x0 < nrh; x0 += VBK) {
    for (int idx = tid; idx < VBK * VBM; idx += 256) {
      int x = idx % VBK, a = idx / VBK;
      int ar = a0 + a, xg = x0 + x;
      double v = 0.0, d = 0.0, t = 0.0;
      if (ar < M && xg < nrh) {
        size_t o = tb + (size_t)(2 * ar + p) * nrh + xg;
        v = tV[o];
        d = tD[o];
        t = tT[o];
      }
      Vs[x][a] = v;
      Ds[x][a] = d;
      Ts[x][a] = t;
    }
    {
      int x = tid % VBK, kk = tid / VBK;  // 16 x 16 = 256
      int xg = x0 + x, kg = k0 + kk;
      double2 r1 = make_double2(0, 0), r2 = r1, p1 = r1, p2 = r1, z1 = r1;
      if (xg < nrh && kg < nz) {
        size_t oa = (size_t)kg * zs + xg, ob = (size_t)kg * zs + (nr - 1 - xg);
        double2 a = cr[oa], b = cr[ob];
        r1 = make_double2(a.x + seo * b.x, a.y + seo * b.y);
        r2 = make_double2(a.x - seo * b.x, a.y - seo * b.y);
        a = cp[oa];
        b = cp[ob];
        p1 = make_double2(a.x + seo * b.x, a.y + seo * b.y);
        p2 = make_double2(a.x - seo * b.x, a.y - seo * b.y);
        a = cz[oa];
        b = cz[ob];
        z1 = make_double2(a.x + seo * b.x, a.y + seo * b.y);
      }
      U[0][x][kk] = r1;
      U[1][x][kk] = r2;
      U[2][x][kk] = p1;
      U[3][x][kk] = p2;
      U[4][x][kk] = z1;
    }
    __syncthreads();
#pragma unroll 4
    for (int xx = 0; xx < VBK; ++xx) {
      double2 u0 = U[0][xx][ty], u1 = U[1][xx][ty], u2 = U[2][xx][ty], u3 = U[3][xx][ty], u4 = U[4][xx][ty];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        double v = Vs[xx][tx + 16 * i], d = Ds[xx][tx + 16 * i], t = Ts[xx][tx + 16 * i];
        acc[i][0].x += v * u0.x; acc[i][0].y += v * u0.y;  // V . eo(ur)
        acc[i][1].x += d * u3.x; acc[i][1].y += d * u3.y;  // D . oe(up)
        acc[i][2].x += d * u1.x; acc[i][2].y += d * u1.y;  // D . oe(ur)
        acc[i][3].x += v * u2.x; acc[i][3].y += v * u2.y;  // V . eo(up)
        acc[i][4].x += t * u4.x; acc[i][4].y += t * u4.y;  // T . eo(uz)
      }
    }
    __syncthreads();
  }
  const int kg = k0 + ty;
  if (kg >= nz) return;
  if (!((kg + 1 <= nzc) || (kg + 1 >= nzcu))) return;
  const double kv = ak[kg];
  const double mv = (double)m;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    int a = a0 + tx + 16 * i;
    if (a >= M) continue;
    size_t o = (size_t)kg * zs + (size_t)nrdim * m + (2 * a + p);
    double2 A = acc[i][0], B = acc[i][1], C = acc[i][2], Dd = acc[i][3], E = acc[i][4];
    if (m == 0) A = make_double2(0.0, 0.0), Dd = make_double2(0.0, 0.0);
    // psi = -i m A - B
    psi[o] = make_double2(mv * A.y - B.x, -mv * A.x - B.y);
    // chi = i kv C + m kv Dd - E
    chi[o] = make_double2(-kv * C.y + mv * kv * Dd.x - E.x, kv * C.x + mv * kv * Dd.y - E.y);
  }
}

// ---------------------------------------------------------------------------
// work tables
// ---------------------------------------------------------------------------
const WorkTable &get_work(mlx_ctx *c, int kind, int nroff) {
  long key = (long)kind * 1000 + nroff;
  auto itw = c->work.find(key);
  if (itw != c->work.end()) return itw->second;
  const KitDev &k = c->kit;
  int npc = k.npchop;
  std::vector<int4> h;
  if (kind == 0 || kind == 2) {
    int tile = (kind == 0) ? BM : VBM;
    for (int m = 0; m < npc; ++m) {
      int nn = nrcs_of(k, m, nroff);
      for (int p = 0; p < 2; ++p) {
        int M = (nn - p + 1) / 2;
        if (M <= 0) continue;
        for (int a0 = 0; a0 < M; a0 += tile) h.push_back(make_int4(m, p, a0, M));
      }
    }
  } else {
    for (int m = 0; m < npc; ++m) {
      int nn = nrcs_of(k, m, nroff);
      for (int x0 = 0; x0 < k.nrh; x0 += BM) h.push_back(make_int4(m, nn, x0, 0));
    }
  }
  WorkTable wt;
  wt.n = (int)h.size();
  if (wt.n > 0) {
    HIPCHK(hipMalloc(&wt.d, sizeof(int4) * h.size()));
    HIPCHK(hipMemcpy(wt.d, h.data(), sizeof(int4) * h.size(), hipMemcpyHostToDevice));
  }
  c->work[key] = wt;
  return c->work[key];
}

}  // namespace

namespace mlx {

void leg_forward(mlx_ctx *c, const double2 *in, double2 *out, int nroff, bool zero, int ncols, bool ztail) {
  const KitDev &k = c->kit;
  if (k.nrchop + nroff > k.nrdim_ref) MLX_THROW(MLX_E_ARG, "rtrans_forward: chopping in r too large");
  if (c->prec != MLX_PREC_FP64 && mfma_ok(k, nroff)) {
    leg_forward_mfma(c, in, out, ncols, ztail);
    return;
  }
  if (zero) HIPCHK(hipMemsetAsync(out, 0, sizeof(double2) * c->nelem, c->stream));
  const WorkTable &wt = get_work(c, 0, nroff);
  if (wt.n == 0) return;
  dim3 grid(wt.n, (k.nzdim + BN - 1) / BN);
  ProbeScope probe(c, MLX_PROBE_LEG_FWD, leg_bytes(k, nroff, 8), leg_flops(k, nroff));
  probed_launch(probe, k_leg_fwd_f64, grid, dim3(256), 0, c->stream, in, out, k.pf, k.w, wt.d, k.nrdim, k.npdim,
                     k.nzdim, k.nr, k.nrh, k.ne);
  HIPCHK(hipGetLastError());
}

void leg_backward(mlx_ctx *c, const double2 *in, double2 *out, int nroff, bool zero, int ncols, bool ztail) {
  const KitDev &k = c->kit;
  if (k.nrchop + nroff > k.nrdim_ref) MLX_THROW(MLX_E_ARG, "rtrans_backward: chopping in r too large");
  if (c->prec != MLX_PREC_FP64 && mfma_ok(k, nroff)) {
    leg_backward_mfma(c, in, out, ncols, ztail);
    return;
  }
  if (zero) HIPCHK(hipMemsetAsync(out, 0, sizeof(double2) * c->nelem, c->stream));
  const WorkTable &wt = get_work(c, 1, nroff);
  if (wt.n == 0) return;
  dim3 grid(wt.n, (k.nzdim + BN - 1) / BN);
  ProbeScope probe(c, MLX_PROBE_LEG_BWD, leg_bytes(k, nroff, 8), leg_flops(k, nroff));
  probed_launch(probe, k_leg_bwd_f64, grid, dim3(256), 0, c->stream, in, out, k.pf, wt.d, k.nrdim, k.npdim, k.nzdim,
                     k.nr, k.nrh, k.ne);
  HIPCHK(hipGetLastError());
}

// ncols, zlo, zhi: as vec2tp_project_f32 on the MFMA path; the FP64 path
// computes every column and slice
void vec2tp_project(mlx_ctx *c, const double2 *ur, const double2 *up, const double2 *uz, double2 *psi,
                    double2 *chi, int nroff, int ncols, int zlo, int zhi, int prec_m0) {
  const KitDev &k = c->kit;
  if (c->prec != MLX_PREC_FP64 && mfma_ok(k, 0)) {
    vec2tp_project_mfma(c, ur, up, uz, psi, chi, ncols, zlo, zhi, prec_m0);
    return;
  }
  HIPCHK(hipMemsetAsync(psi, 0, sizeof(double2) * c->nelem, c->stream));
  HIPCHK(hipMemsetAsync(chi, 0, sizeof(double2) * c->nelem, c->stream));
  const WorkTable &wt = get_work(c, 2, nroff);
  if (wt.n == 0) return;
  dim3 grid(wt.n, (k.nzdim + VBN - 1) / VBN);
  ProbeScope probe(c, MLX_PROBE_V2T, v2t_bytes(k, 0, 8), 5.0 * leg_flops(k, 0));
  probed_launch(probe, k_v2t_f64, grid, dim3(256), 0, c->stream, ur, up, uz, psi, chi, k.tV, k.tD, k.tT, k.ak, wt.d,
                     k.nrdim, k.npdim, k.nzdim, k.nr, k.nrh, k.nt, k.chopzl, k.chopzu);
  HIPCHK(hipGetLastError());
}

}  // namespace mlx
