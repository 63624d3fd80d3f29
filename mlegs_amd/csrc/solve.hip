// Banded LU of (del2 + alpha) per (m, k) system and its solve, FP64.
//
// Replaces lsolve -> zgbtrf/zgbtrs behind ihelm / idel2
// (src/submodules/mlegs_bndmat_ops.f90:369-448, called from
// src/submodules/mlegs_scalar_ops.f90:575-854).  The reference rebuilds and
// refactorises every (m, k) matrix on every call; here each (alpha, kind)
// set of factors is computed once (gbtf2: partial pivoting, multipliers by
// the reciprocal pivot, rank-1 updates skipped on zero entries) and cached.
// The matrix depends on k only through ak(k)^2, which is bit-identical for
// k and nz - k, so factors are stored once per |kappa|.
//
// Solve: the recurrences of the two triangular sweeps are split over the
// lanes.  A system's rows are cut into 32 segments of S rows, one per lane,
// and a wavefront holds the systems of k and nz - k (same factors, lanes
// 0-31 and 32-63).  Each sweep is an affine map of a small carried state
// (forward: the three rows w0..w2 being eliminated; back: the four solved
// rows x(i+1..i+4)); every lane builds its segment's map (state matrix from
// unit states, offset from its right-hand side), a five-step shuffle scan
// composes the maps of the preceding segments, and each lane then reruns
// its segment from the true incoming state with the sequential arithmetic.
// All of a lane's rows and factors are loaded at once, so a wavefront makes
// one round trip to memory instead of one per chunk of rows.
//
// Factor layout, per system sys = kk * npc + m and row j = S s + u (lane s):
//   FL (double2) L(j+1,j), L(j+2,j)        at (sys S + u) 32 + s
//   FU (double2) U(j,j+1), U(j,j+2)        same index
//   FD (double)  U(j,j)                    same index
//   FX (double2) U(j,j+3), U(j,j+4)        fill-in, read by pivoting systems
//   PW (uint64)  2 bits per row interchange, one word per (sys, s)
// Consecutive lanes read consecutive factor entries.
#include <cmath>
#include <vector>

#include "mlx_internal.h"

namespace {

constexpr int KL = 2, KU = 2, KV = 4, LDAB = 7;
constexpr int NSEG = 32;  // row segments per system (lanes per system)

struct FacPtrs {
  double2 *fl;
  double2 *fu;
  double *fd;
  double2 *fx;
};
// nrows = systems x S x 32 padded rows
__host__ __device__ inline FacPtrs fac_ptrs(double *base, size_t nrows) {
  FacPtrs p;
  p.fl = reinterpret_cast<double2 *>(base);
  p.fu = reinterpret_cast<double2 *>(base + 2 * nrows);
  p.fd = base + 4 * nrows;
  p.fx = reinterpret_cast<double2 *>(base + 5 * nrows);
  return p;
}

__device__ inline int nrcs_dev(int nrc, int npc, int m) {
  if (m >= npc) return 0;
  int v = nrc - m;
  return v > 0 ? v : 0;
}
__device__ inline int z_count(int k, int nzc, int nzcu) {
  return ((k + 1 <= nzc) ? 1 : 0) + ((k + 1 >= nzcu) ? 1 : 0);
}
// (del2 + alpha)(i, j) in the normalised basis; bordered (m=0,k=0 of
// idel2_proln): column 0 gains the 4/3, -2, 2/3 log-term entries
// (mlegs_scalar_ops.f90:705-711)
__device__ inline double d2_entry(const double *d2c, const double *lognorm, int m, int ne, int i, int j, double akk,
                                  double alpha, int bordered, double ell) {
  int d = j - i + 2;
  double v = 0.0;
  if (d >= 0 && d <= 4) {
    v = d2c[((size_t)m * ne + i) * 5 + d];
    if (d == 2) v = v - akk + alpha;
  }
  if (bordered && j == 0) {
    double ell2 = ell * ell;
    if (i == 0) v += 4.0 / 3.0 / ell2;
    if (i == 1) v -= 2.0 / 1.0 / ell2 * exp(lognorm[0] - lognorm[1]);
    if (i == 2) v += 2.0 / 3.0 / ell2 * exp(lognorm[0] - lognorm[2]);
  }
  return v;
}

// one thread per (m, kk) system: gbtf2 in LAPACK band storage
// (element (i, j) at ab[(KV + i - j) + LDAB j], rows 0..KL-1 fill-in)
__global__ void k_lu_factor(double *fac, int8_t *piv, const double *d2c, const double *lognorm, const double *ak,
                            int nrdim, int nkk, int ne, int nrc, int npc, double alpha, int kind, double ell,
                            int *info) {
  int kk = blockIdx.x * blockDim.x + threadIdx.x;
  int m = blockIdx.y;
  if (kk >= nkk || m >= npc) return;
  int n = nrcs_dev(nrc, npc, m);
  if (n <= 0) return;
  const bool tr = (kind & 4) != 0;  // the transposed matrix (the adjoint's solves)
  kind &= 3;
  if (kind == 2 && m == 0 && kk == 0) return;  // idel2_preln: (0,0) solved densely
  size_t sys = (size_t)m * nkk + kk;
  double *ab = fac + sys * (size_t)nrdim * LDAB;
  int8_t *ip = piv + sys * (size_t)nrdim;
  double akk = ak[kk] * ak[kk];
  int bordered = (kind == 1 && m == 0 && kk == 0) ? 1 : 0;
  for (int j = 0; j < n; ++j) {
    for (int r = 0; r < LDAB; ++r) ab[r + LDAB * j] = 0.0;
    for (int i = (j - KU > 0 ? j - KU : 0); i <= j + KL && i < n; ++i)
      ab[(KV + i - j) + LDAB * j] = tr ? d2_entry(d2c, lognorm, m, ne, j, i, akk, alpha, bordered, ell)
                                       : d2_entry(d2c, lognorm, m, ne, i, j, akk, alpha, bordered, ell);
  }
  int ju = 0;
  for (int j = 0; j < n; ++j) {
    int km = (KL < n - 1 - j) ? KL : n - 1 - j;
    int jp = 0;
    double best = fabs(ab[KV + LDAB * j]);
    for (int r = 1; r <= km; ++r) {
      double v = fabs(ab[KV + r + LDAB * j]);
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
          double t = *a;
          *a = *b;
          *b = t;
        }
      }
      if (km > 0) {
        double rp = 1.0 / ab[KV + LDAB * j];
        for (int r = 1; r <= km; ++r) ab[KV + r + LDAB * j] *= rp;
        for (int c = j + 1; c <= ju; ++c) {
          double u = ab[(KV + j - c) + LDAB * c];
          if (u != 0.0)
            for (int r = 1; r <= km; ++r) ab[(KV + j + r - c) + LDAB * c] -= ab[KV + r + LDAB * j] * u;
        }
      }
    } else {
      atomicExch(info, 1);
    }
  }
}

// factors of one system into the segment layout; rows n.. of the padded
// system are identity rows (U(j,j) = 1, no multipliers)
__global__ void k_lu_repack(const double *fac, const int8_t *piv, FacPtrs F, uint64_t *pw, int *anyp, int nrdim,
                            int nkk, int nrc, int npc, int kind, int S) {
  int kk = blockIdx.x * blockDim.x + threadIdx.x;
  int m = blockIdx.y;
  if (kk >= nkk || m >= npc) return;
  kind &= 3;
  if (kind == 2 && m == 0 && kk == 0) return;
  int n = nrcs_dev(nrc, npc, m);
  const size_t sysk = (size_t)kk * npc + m;  // segment-layout system index
  const size_t sysf = (size_t)m * nkk + kk;  // factor-kernel system index
  const double *ab = fac + sysf * (size_t)nrdim * LDAB;
  const int8_t *ip = piv + sysf * (size_t)nrdim;
  int any = 0;
  for (int sg = 0; sg < NSEG; ++sg) {
    uint64_t word = 0;
    for (int u = 0; u < S; ++u) {
      const int j = sg * S + u;
      const size_t o = (sysk * S + u) * NSEG + sg;
      if (j < n) {
        double uu[5];
        for (int d = 0; d <= 4; ++d) uu[d] = (j + d < n) ? ab[(KV - d) + LDAB * (j + d)] : 0.0;
        F.fl[o] = make_double2((j + 1 < n) ? ab[KV + 1 + LDAB * j] : 0.0, (j + 2 < n) ? ab[KV + 2 + LDAB * j] : 0.0);
        F.fu[o] = make_double2(uu[1], uu[2]);
        F.fd[o] = uu[0];
        F.fx[o] = make_double2(uu[3], uu[4]);
        const int p = ip[j];
        any |= p;
        word |= (uint64_t)p << (2 * u);
      } else {
        F.fl[o] = make_double2(0.0, 0.0);
        F.fu[o] = make_double2(0.0, 0.0);
        F.fd[o] = 1.0;
        F.fx[o] = make_double2(0.0, 0.0);
      }
    }
    pw[sysk * NSEG + sg] = word;
  }
  anyp[sysk] = any;
}

__device__ __forceinline__ double2 shup2(double2 v, int d) {
  return make_double2(__shfl_up(v.x, d, NSEG), __shfl_up(v.y, d, NSEG));
}
__device__ __forceinline__ double2 shdn2(double2 v, int d) {
  return make_double2(__shfl_down(v.x, d, NSEG), __shfl_down(v.y, d, NSEG));
}
__device__ __forceinline__ void axpy2(double2 &y, double a, double2 x) {
  y.x -= a * x.x;
  y.y -= a * x.y;
}
__device__ __forceinline__ void fma2(double2 &y, double a, double2 x) {
  y.x += a * x.x;
  y.y += a * x.y;
}
// affine maps of a two-row state: (t, p) <- (t, p) o (o, q) = (t o, t q + p)
__device__ __forceinline__ void compose2(double (&t)[2][2], double2 (&p)[2], const double (&o)[2][2],
                                         const double2 (&q)[2]) {
  double2 n0 = p[0], n1 = p[1];
  fma2(n0, t[0][0], q[0]);
  fma2(n0, t[0][1], q[1]);
  fma2(n1, t[1][0], q[0]);
  fma2(n1, t[1][1], q[1]);
  p[0] = n0;
  p[1] = n1;
  const double a00 = t[0][0] * o[0][0] + t[0][1] * o[1][0], a01 = t[0][0] * o[0][1] + t[0][1] * o[1][1];
  const double a10 = t[1][0] * o[0][0] + t[1][1] * o[1][0], a11 = t[1][0] * o[0][1] + t[1][1] * o[1][1];
  t[0][0] = a00;
  t[0][1] = a01;
  t[1][0] = a10;
  t[1][1] = a11;
}
// inclusive scan of two-row maps over the 32 lanes of a half: prefix
// (UP: lane s composes lanes 0..s, later maps applied last) or suffix
template <bool UP>
__device__ __forceinline__ void scan2(double (&t)[2][2], double2 (&p)[2], int s) {
#pragma unroll
  for (int d = 1; d < NSEG; d *= 2) {
    double o[2][2];
    double2 q[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
#pragma unroll
      for (int c = 0; c < 2; ++c) o[r][c] = UP ? __shfl_up(t[r][c], d, NSEG) : __shfl_down(t[r][c], d, NSEG);
      q[r] = UP ? shup2(p[r], d) : shdn2(p[r], d);
    }
    if (UP ? (s >= d) : (s + d < NSEG)) compose2(t, p, o, q);
  }
}
// LDS of one wavefront for the four-row chain of pivoting systems: segment
// maps (state matrix t, offset p) and the incoming states handed back
struct SegLds {
  double t[64][16];
  double2 p[64][4];
  double2 in[64][4];
};

// Forward sweep L y = P b over the segments.  b[u] holds row s S + u on
// entry and y on exit.  Row j (j < n - 1): interchange w0 with w1 / w2,
// w1 -= L(j+1,j) w0, w2 -= L(j+2,j) w0, emit w0, shift in b(j+3).
// Entering a segment, w2 is always the untouched b(j0 + 2) (shifted in
// after the previous row's interchange), so the carried state is (w0, w1).
template <int S, bool PIV>
__device__ __forceinline__ void fwd_sweep(double2 (&b)[S], const double2 (&fl)[S], uint64_t pw, int s, int n) {
  const double2 z2 = make_double2(0.0, 0.0);
  const int j0 = s * S;
  // rows j0 + S .. j0 + S + 2 belong to the next lane
  double2 nb[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    nb[q] = shdn2(b[q], 1);
    if (s == NSEG - 1) nb[q] = z2;
  }
  auto inj = [&](int u) -> double2 { return (u + 3 < S) ? b[u + 3] : nb[u + 3 - S]; };
  // segment map (w0, w1)' = T (w0, w1) + p; lane 0 starts from the true
  // state (b0, b1, b2) and has T = 0
  double2 w0 = (s == 0) ? b[0] : z2, w1 = (s == 0) ? b[1] : z2, w2 = b[2];
  double v[3][2];  // state rows for unit inputs in w0 / w1
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) v[r][c] = (s != 0 && r == c) ? 1.0 : 0.0;
#pragma unroll
  for (int u = 0; u < S; ++u) {
    const int j = j0 + u;
    if (j < n - 1) {
      if (PIV) {
        const unsigned pv = (unsigned)(pw >> (2 * u)) & 3u;
        if (pv == 1) {
          double2 x = w0; w0 = w1; w1 = x;
#pragma unroll
          for (int c = 0; c < 2; ++c) { double y = v[0][c]; v[0][c] = v[1][c]; v[1][c] = y; }
        } else if (pv == 2) {
          double2 x = w0; w0 = w2; w2 = x;
#pragma unroll
          for (int c = 0; c < 2; ++c) { double y = v[0][c]; v[0][c] = v[2][c]; v[2][c] = y; }
        }
      }
      axpy2(w1, fl[u].x, w0);
      axpy2(w2, fl[u].y, w0);
      w0 = w1;
      w1 = w2;
      w2 = inj(u);
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const double a0 = v[0][c];
        v[0][c] = v[1][c] - fl[u].x * a0;
        v[1][c] = v[2][c] - fl[u].y * a0;
        v[2][c] = 0.0;
      }
    }
  }
  double t[2][2] = {{v[0][0], v[0][1]}, {v[1][0], v[1][1]}};
  double2 p[2] = {w0, w1};
  scan2<true>(t, p, s);
  // incoming state: the previous lane's outgoing state
  {
    const double2 q0 = shup2(p[0], 1), q1 = shup2(p[1], 1);
    w0 = (s == 0) ? b[0] : q0;
    w1 = (s == 0) ? b[1] : q1;
    w2 = b[2];
  }
#pragma unroll
  for (int u = 0; u < S; ++u) {
    const int j = j0 + u;
    if (j < n - 1) {
      if (PIV) {
        const unsigned pv = (unsigned)(pw >> (2 * u)) & 3u;
        if (pv == 1) { double2 x = w0; w0 = w1; w1 = x; }
        else if (pv == 2) { double2 x = w0; w0 = w2; w2 = x; }
      }
      axpy2(w1, fl[u].x, w0);
      axpy2(w2, fl[u].y, w0);
      const double2 nx = inj(u);  // read before b[u] is overwritten (u + 3 > u)
      b[u] = w0;
      w0 = w1;
      w1 = w2;
      w2 = nx;
    } else if (j == n - 1) {
      b[u] = w0;
    }
  }
}

// Back substitution U x = y over the segments, rows in descending order:
// x(i) = (y(i) - U(i,i+4) x(i+4) - U(i,i+3) x(i+3) - U(i,i+2) x(i+2)
//         - U(i,i+1) x(i+1)) / U(i,i)
// The carried state is x(i+1..i+D): D = 2 without interchanges (U has two
// superdiagonals), D = 4 with them (fill-in).  Two-row maps are scanned
// across the lanes; four-row maps are chained in order by one lane per half.
template <int S, bool PIV>
__device__ __forceinline__ void back_sweep(double2 (&b)[S], const double (&fd)[S], const double2 (&fu)[S],
                                           const double2 (&fx)[S], int s, int n, SegLds &sc) {
  constexpr int D = PIV ? 4 : 2;
  const double2 z2 = make_double2(0.0, 0.0);
  const int j0 = s * S;
  double2 x[4] = {z2, z2, z2, z2};
  double t[D][D];
#pragma unroll
  for (int r = 0; r < D; ++r)
#pragma unroll
    for (int c = 0; c < D; ++c) t[r][c] = (r == c) ? 1.0 : 0.0;
#pragma unroll
  for (int u = S - 1; u >= 0; --u) {
    const int i = j0 + u;
    if (i < n) {
      double2 acc = b[u];
      if (PIV) {
        axpy2(acc, fx[u].y, x[3]);
        axpy2(acc, fx[u].x, x[2]);
      }
      axpy2(acc, fu[u].y, x[1]);
      axpy2(acc, fu[u].x, x[0]);
      const double d = fd[u];
      const double2 xi = make_double2(acc.x / d, acc.y / d);
      x[3] = x[2];
      x[2] = x[1];
      x[1] = x[0];
      x[0] = xi;
      const double rd = 1.0 / d;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        double a = -fu[u].x * t[0][c] - fu[u].y * t[1][c];
        if constexpr (D == 4) a -= fx[u].x * t[2][c] + fx[u].y * t[3][c];
#pragma unroll
        for (int r = D - 1; r > 0; --r) t[r][c] = t[r - 1][c];
        t[0][c] = a * rd;
      }
    }
  }
  if constexpr (D == 2) {
    double2 p[2] = {x[0], x[1]};
    scan2<false>(t, p, s);
    const double2 q0 = shdn2(p[0], 1), q1 = shdn2(p[1], 1);
    const bool last = (s == NSEG - 1);
    x[0] = last ? z2 : q0;
    x[1] = last ? z2 : q1;
  } else {
    // chain over the segments in descending order
    double *ts = sc.t[threadIdx.x];
    double2 *ps = sc.p[threadIdx.x];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) ts[4 * r + c] = t[r][c];
#pragma unroll
    for (int r = 0; r < 4; ++r) ps[r] = x[r];
    __syncthreads();
    if (s == 0) {
      const int h0 = threadIdx.x;
      double2 a[4] = {z2, z2, z2, z2};
#pragma unroll 4
      for (int q = NSEG - 1; q >= 0; --q) {
        const double *tq = sc.t[h0 + q];
        const double2 *pq = sc.p[h0 + q];
        double2 c[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          sc.in[h0 + q][r] = a[r];
          c[r] = pq[r];
#pragma unroll
          for (int k = 0; k < 4; ++k) fma2(c[r], tq[4 * r + k], a[k]);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) a[r] = c[r];
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) x[r] = sc.in[threadIdx.x][r];
    __syncthreads();
  }
#pragma unroll
  for (int u = S - 1; u >= 0; --u) {
    const int i = j0 + u;
    if (i < n) {
      double2 acc = b[u];
      if (PIV) {
        axpy2(acc, fx[u].y, x[3]);
        axpy2(acc, fx[u].x, x[2]);
      }
      axpy2(acc, fu[u].y, x[1]);
      axpy2(acc, fu[u].x, x[0]);
      const double d = fd[u];
      const double2 xi = make_double2(acc.x / d, acc.y / d);
      b[u] = xi;
      x[3] = x[2];
      x[2] = x[1];
      x[1] = x[0];
      x[0] = xi;
    }
  }
}

template <int S, bool PIV>
__device__ __forceinline__ void solve_rows(double2 (&b)[S], const FacPtrs &F, const uint64_t *__restrict__ pw,
                                           size_t sys, int s, int n, SegLds &sc) {
  {
    double2 fl[S];
#pragma unroll
    for (int u = 0; u < S; ++u) fl[u] = F.fl[(sys * S + u) * NSEG + s];
    const uint64_t w = PIV ? pw[sys * NSEG + s] : 0;
    fwd_sweep<S, PIV>(b, fl, w, s, n);
  }
  double2 fu[S], fx[S];
  double fd[S];
#pragma unroll
  for (int u = 0; u < S; ++u) {
    const size_t o = (sys * S + u) * NSEG + s;
    fu[u] = F.fu[o];
    fd[u] = F.fd[o];
    fx[u] = PIV ? F.fx[o] : make_double2(0.0, 0.0);
  }
  back_sweep<S, PIV>(b, fd, fu, fx, s, n, sc);
}

// one wavefront per (m, |kappa|): lanes 0-31 solve slice k = kk, lanes 32-63
// slice nz - kk (same factors); lane s of a half owns rows s S .. s S + S - 1.
// Systems with and without row interchanges are solved by separate
// instantiations (PIV), each returning at once for the other kind.
// blockIdx.z selects the right-hand sides: e (0) or e2 (1), solved with the
// same factors.  mixed = 0: every system is of this launch's kind (anyp is
// not read).
template <int S, bool PIV>
__global__ void __launch_bounds__(64) k_lu_solve(double2 *e0, double2 *e2, FacPtrs F, const uint64_t *__restrict__ pw,
                                                 const int *__restrict__ anyp, int mixed, int nrdim, int npdim, int nz,
                                                 int nrc, int npc, int nzc, int nzcu, int skip00, int zlo, int zhi) {
  const int m = blockIdx.x, kk = blockIdx.y;
  double2 *e = blockIdx.z ? e2 : e0;
  // slices zlo..zhi left unsolved (both kk and nz - kk in the range)
  if (kk >= zlo && kk <= zhi && (kk == 0 || 2 * kk == nz || (nz - kk >= zlo && nz - kk <= zhi))) return;
  const int lane = threadIdx.x, s = lane % NSEG;
  const bool twin = lane >= NSEG;
  if (m >= npc) return;
  const int n = nrc - m;
  if (n <= 0) return;
  const size_t sys = (size_t)kk * npc + m;
  if (mixed && (anyp[sys] != 0) != PIV) return;
  const int k = twin ? nz - kk : kk;
  const bool valid = !twin || (kk != 0 && 2 * kk != nz);
  int cnt = valid ? z_count(k, nzc, nzcu) : 0;
  if (skip00 && m == 0 && k == 0) cnt = 0;
  double2 *col = e + (size_t)(valid ? k : kk) * nrdim * npdim + (size_t)nrdim * m;
  const int j0 = s * S;
  double2 b[S];
  bool nonzero = false;
#pragma unroll
  for (int u = 0; u < S; ++u) {
    const int j = j0 + u;
    b[u] = (cnt > 0 && j < n) ? col[j] : make_double2(0.0, 0.0);
    nonzero |= (b[u].x != 0.0) || (b[u].y != 0.0);
  }
  // an exactly zero right-hand side has the zero solution: de-aliased fields
  // are zero on a third of the kappa slices and above m = np/3
  if (!__any(nonzero)) return;
  __shared__ SegLds sc;
  // the duplicated Nyquist slice (z_count 2) is solved twice, as by the
  // reference's loop over both chopped ranges
  for (int rep = 0; rep < 2; ++rep) {
    if (!__any(rep < cnt)) break;
    solve_rows<S, PIV>(b, F, pw, sys, s, n, sc);
    if (rep < cnt) {
#pragma unroll
      for (int u = 0; u < S; ++u)
        if (j0 + u < n) col[j0 + u] = b[u];
    }
  }
}

// rows per segment: the smallest instantiated S with 32 S >= nrc
int seg_rows(int nrc) {
  static const int opts[] = {4, 8, 9, 12, 16, 17};
  for (int S : opts)
    if (NSEG * S >= nrc) return S;
  return 0;
}

// factor rows of all systems: (|kappa| values) x npchop systems x S x 32
size_t fac_rows(const KitDev &k, int S) {
  const int nkk = (k.nzdim > 1) ? k.nzdim / 2 + 1 : 1;
  return (size_t)nkk * k.npchop * S * NSEG;
}

LuCache &lu_get(mlx_ctx *c, double alpha, int kind, int nroff) {
  for (auto &e : c->lu)
    if (e.alpha == alpha && e.kind == kind && e.nroff == nroff) {
      e.stamp = ++c->lu_clock;
      return e;
    }
  const KitDev &k = c->kit;
  const int nrc = k.nrchop + nroff;
  if (nrc > k.nrdim_ref) MLX_THROW(MLX_E_ARG, "chop_index: chopping in r too large");
  const int S = seg_rows(nrc);
  if (S == 0) MLX_THROW(MLX_E_ARG, "ihelm: more than 544 radial rows per system are not supported");
  if (c->lu.size() >= 6) {  // evict the least recently used entry
    size_t victim = 0;
    for (size_t i = 1; i < c->lu.size(); ++i)
      if (c->lu[i].stamp < c->lu[victim].stamp) victim = i;
    HIPCHK(hipStreamSynchronize(c->stream));
    (void)hipFree(c->lu[victim].sol);
    (void)hipFree(c->lu[victim].pvw);
    (void)hipFree(c->lu[victim].anyp);
    c->lu.erase(c->lu.begin() + victim);
  }
  LuCache e;
  e.alpha = alpha;
  e.kind = kind;
  e.nroff = nroff;
  e.seg = S;
  const int nkk = (k.nzdim > 1) ? k.nzdim / 2 + 1 : 1;
  const size_t nsys = (size_t)k.npchop * nkk;
  const size_t nrows = fac_rows(k, S);
  double *fac = nullptr;
  int8_t *piv = nullptr;
  auto cleanup = [&]() {
    if (fac) (void)hipFree(fac);
    if (piv) (void)hipFree(piv);
    if (e.sol) (void)hipFree(e.sol);
    if (e.pvw) (void)hipFree(e.pvw);
    if (e.anyp) (void)hipFree(e.anyp);
  };
  if (hipMalloc(&fac, sizeof(double) * nsys * k.nrdim * LDAB) != hipSuccess ||
      hipMalloc(&piv, nsys * k.nrdim) != hipSuccess ||
      hipMalloc(&e.sol, sizeof(double) * 7 * nrows) != hipSuccess ||
      hipMalloc(&e.pvw, sizeof(uint64_t) * nsys * NSEG) != hipSuccess ||
      hipMalloc(&e.anyp, sizeof(int) * nsys) != hipSuccess) {
    cleanup();
    MLX_THROW(MLX_E_ALLOC, "ihelm: cannot allocate LU factors");
  }
  // systems the factor kernel skips ((0,0) of kind 2) keep zero words
  HIPCHK(hipMemsetAsync(e.anyp, 0, sizeof(int) * nsys, c->stream));
  HIPCHK(hipMemsetAsync(e.pvw, 0, sizeof(uint64_t) * nsys * NSEG, c->stream));
  int *dinfo = reinterpret_cast<int *>(c->dsmall + 4000);
  HIPCHK(hipMemsetAsync(dinfo, 0, sizeof(int), c->stream));
  dim3 grid((nkk + 63) / 64, k.npchop);
  hipLaunchKernelGGL(k_lu_factor, grid, dim3(64), 0, c->stream, fac, piv, k.d2c, k.lognorm, k.ak, k.nrdim, nkk, k.ne,
                     nrc, k.npchop, alpha, kind, k.ell, dinfo);
  hipLaunchKernelGGL(k_lu_repack, grid, dim3(64), 0, c->stream, fac, piv, fac_ptrs(e.sol, nrows), e.pvw, e.anyp,
                     k.nrdim, nkk, nrc, k.npchop, kind, S);
  HIPCHK(hipGetLastError());
  int info = 0;
  std::vector<int> anyp(nsys);
  HIPCHK(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(anyp.data(), e.anyp, sizeof(int) * nsys, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  (void)hipFree(fac);
  (void)hipFree(piv);
  fac = nullptr;
  piv = nullptr;
  if (info != 0) {
    cleanup();
    MLX_THROW(MLX_E_ARG, "lurc: lu factorization resulted in failure");
  }
  // factor bytes one solve must read: 40 per row, 16 more per row and the
  // interchange words for systems with interchanges
  e.fac_bytes = 0.0;
  e.npiv = 0;
  e.pivh.assign(nsys, 0);
  for (int kk = 0; kk < nkk; ++kk)
    for (int m = 0; m < k.npchop; ++m) {
      const int n = nrc - m > 0 ? nrc - m : 0;
      const bool p = anyp[(size_t)kk * k.npchop + m] != 0;
      e.fac_bytes += (double)n * (p ? 56.0 : 40.0) + (p ? 8.0 * NSEG : 0.0);
      e.npiv += p ? 1 : 0;
      e.pivh[(size_t)kk * k.npchop + m] = p ? 1 : 0;
    }
  e.nsys = (int64_t)nsys;
  e.stamp = ++c->lu_clock;
  c->lu.push_back(e);
  return c->lu.back();
}

}  // namespace

namespace mlx {

void k_helm_solve(mlx_ctx *c, double2 *e, double alpha, int kind, int nroff, int pcut, int zlo, int zhi,
                  double2 *e2) {
  const KitDev &k = c->kit;
  LuCache &lu = lu_get(c, alpha, kind, nroff);
  const int nkk = (k.nzdim > 1) ? k.nzdim / 2 + 1 : 1;
  const int npc_grid = (pcut > 0 && pcut < k.npchop) ? pcut : k.npchop;
  const int nf = e2 ? 2 : 1;
  dim3 grid(npc_grid, nkk, nf);
  // right-hand sides of the solved systems read and written once, their
  // factors (one set per |kappa|) read once
  const int nrc0 = k.nrchop + nroff;
  auto in_band = [&](int kz) { return kz >= zlo && kz <= zhi; };
  double rows = 0.0, fbytes = 0.0;
  for (int kk = 0; kk < nkk; ++kk) {
    int nk = 0;  // slices kk and nz - kk that are solved
    if (!in_band(kk)) ++nk;
    if (kk != 0 && 2 * kk != k.nzdim && !in_band(k.nzdim - kk)) ++nk;
    if (nk == 0) continue;
    for (int m = 0; m < npc_grid; ++m) {
      const int n = nrc0 - m > 0 ? nrc0 - m : 0;
      const bool pv = lu.pivh[(size_t)kk * k.npchop + m] != 0;
      rows += (double)n * nk;
      fbytes += (double)n * (pv ? 56.0 : 40.0) + (pv ? 8.0 * NSEG : 0.0);
    }
  }
  ProbeScope probe(c, MLX_PROBE_LU_SOLVE, 32.0 * rows * nf + fbytes, 2.0 * 4.0 * 5.0 * rows * nf);
  const FacPtrs F = fac_ptrs(lu.sol, fac_rows(k, lu.seg));
  const int nrc = k.nrchop + nroff, sk = (kind & 3) == 2 ? 1 : 0;
  const int mixed = lu.npiv > 0 && lu.npiv < lu.nsys;
#define MLX_LU_LAUNCH(SV, PV)                                                                              \
  probed_launch(probe, k_lu_solve<SV, PV>, grid, dim3(64), 0, c->stream, e, e2 ? e2 : e, F, lu.pvw, lu.anyp, mixed, \
                k.nrdim, k.npdim, k.nzdim, nrc, k.npchop, k.chopzl, k.chopzu, sk, zlo, zhi)
#define MLX_LU_BOTH(SV)                               \
  do {                                                \
    if (lu.npiv < lu.nsys) MLX_LU_LAUNCH(SV, false); \
    if (lu.npiv > 0) MLX_LU_LAUNCH(SV, true);         \
  } while (0)
  switch (lu.seg) {
    case 4: MLX_LU_BOTH(4); break;
    case 8: MLX_LU_BOTH(8); break;
    case 9: MLX_LU_BOTH(9); break;
    case 12: MLX_LU_BOTH(12); break;
    case 16: MLX_LU_BOTH(16); break;
    case 17: MLX_LU_BOTH(17); break;
    default: MLX_THROW(MLX_E_ARG, "ihelm: bad segment size");
  }
#undef MLX_LU_BOTH
#undef MLX_LU_LAUNCH
  HIPCHK(hipGetLastError());
}

void lu_cache_free(mlx_ctx *c) {
  for (auto &e : c->lu) {
    (void)hipFree(e.sol);
    (void)hipFree(e.pvw);
    (void)hipFree(e.anyp);
  }
  c->lu.clear();
}

}  // namespace mlx
