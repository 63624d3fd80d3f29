// phi and z Fourier transforms of the field, FP64, in LDS.
//
// Replaces horizontal_fft_forward/backward and vertical_fft_forward/backward
// (src/submodules/mlegs_scalar_ops.f90:1567-1848), which call FFTE's
// dzfft2d/zdfft2d/zfft1d per row.  Conventions kept from the reference:
//   phi forward : real row f_0..f_{np-1} packed as e(i, j) = f_{2j} + i f_{2j+1}
//                 -> C_m = (1/np) sum_t f_t e^{-2 pi i m t/np}, m = 0..np/2
//   phi backward: f_t = sum_m C_m e^{+2 pi i m t/np} (Hermitian), packed back;
//                 column np/2+1 (scratch in the reference) is written as 0
//   z forward   : (1/nz) sum e^{-2 pi i jk/nz};  z backward: sum e^{+...}
//   every row (including radial padding) and every column is transformed.
//
// Design: one workgroup owns TI consecutive radial rows i of one (k) plane
// (phi) or one (j) column (z); it stages the TI lines in LDS with the line
// index fastest (coalesced global access, conflict-free LDS), runs a
// mixed-radix (4,2,3,5) Stockham FFT over the lines, and writes back.  The
// real-to-complex phi transform is a complex FFT of length np/2 plus the
// even/odd untangling step.  HBM-bound: one read and one write per element.
#include <cmath>

#include "mlx_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kLdsElems = 4096;  // complex doubles per workgroup (64 KiB)

// The fields of a batched in-place z transform (one per blockIdx.z)
struct ZFields {
  double2 *p[mlx::kZBatch];
};

struct Cx {
  double x, y;
};
__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 cmuli(double2 a, double s) { return make_double2(-s * a.y, s * a.x); }  // i*s*a

// DFT of size R on v with sign sg (-1 forward, +1 inverse)
template <int R>
__device__ inline void dft(double2 *v, double sg);

template <>
__device__ inline void dft<2>(double2 *v, double) {
  double2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
template <>
__device__ inline void dft<4>(double2 *v, double sg) {
  double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
  double2 b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
  double2 ib1 = cmuli(b1, sg);  // (sg*i) * (v1 - v3)
  v[0] = cadd(a0, b0);
  v[2] = csub(a0, b0);
  v[1] = cadd(a1, ib1);
  v[3] = csub(a1, ib1);
}
template <>
__device__ inline void dft<3>(double2 *v, double sg) {
  const double c1 = -0.5, s1 = 0.86602540378443864676;
  double2 t1 = cadd(v[1], v[2]);
  double2 t2 = make_double2(v[0].x + c1 * t1.x, v[0].y + c1 * t1.y);
  double2 t3 = cmuli(csub(v[1], v[2]), sg * s1);
  v[0] = cadd(v[0], t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}
template <>
__device__ inline void dft<5>(double2 *v, double sg) {
  const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
  const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
  double2 a14 = cadd(v[1], v[4]), b14 = csub(v[1], v[4]);
  double2 a23 = cadd(v[2], v[3]), b23 = csub(v[2], v[3]);
  double2 r1 = make_double2(v[0].x + c1 * a14.x + c2 * a23.x, v[0].y + c1 * a14.y + c2 * a23.y);
  double2 r2 = make_double2(v[0].x + c2 * a14.x + c1 * a23.x, v[0].y + c2 * a14.y + c1 * a23.y);
  double2 i1 = make_double2(s1 * b14.x + s2 * b23.x, s1 * b14.y + s2 * b23.y);
  double2 i2 = make_double2(s2 * b14.x - s1 * b23.x, s2 * b14.y - s1 * b23.y);
  double2 j1 = cmuli(i1, sg), j2 = cmuli(i2, sg);
  v[0] = make_double2(v[0].x + a14.x + a23.x, v[0].y + a14.y + a23.y);
  v[1] = cadd(r1, j1);
  v[4] = csub(r1, j1);
  v[2] = cadd(r2, j2);
  v[3] = csub(r2, j2);
}

// One Stockham radix-R pass over TI lines of length N held in buf[t*TI + l].
template <int R, int MAXB>
__device__ inline void stockham_pass(double2 *buf, const double2 *tw, int N, int TI, int ns, double sg) {
  const int NR = N / R;
  const int total = TI * NR;
  double2 v[MAXB][R];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    int b = threadIdx.x + q * kThreads;
    if (b < total) {
      int l = b % TI, j = b / TI;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = buf[(j + r * NR) * TI + l];
    }
  }
  __syncthreads();
  const int step0 = N / (ns * R);
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    int b = threadIdx.x + q * kThreads;
    if (b < total) {
      int l = b % TI, j = b / TI;
      int k = j % ns;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        double2 t = tw[r * k * step0];
        if (sg > 0) t.y = -t.y;
        v[q][r] = cmul(v[q][r], t);
      }
      dft<R>(v[q], sg);
      int idxD = (j / ns) * ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(idxD + r * ns) * TI + l] = v[q][r];
    }
  }
  __syncthreads();
}

__device__ void lds_fft(double2 *buf, const double2 *tw, const FftPlan &pl, int TI, double sg) {
  int ns = 1;
  for (int p = 0; p < pl.npass; ++p) {
    int R = pl.radix[p];
    if (R == 4) stockham_pass<4, kLdsElems / 4 / kThreads>(buf, tw, pl.n, TI, ns, sg);
    else if (R == 2) stockham_pass<2, kLdsElems / 2 / kThreads>(buf, tw, pl.n, TI, ns, sg);
    else if (R == 3) stockham_pass<3, (kLdsElems / 3 + kThreads - 1) / kThreads>(buf, tw, pl.n, TI, ns, sg);
    else stockham_pass<5, (kLdsElems / 5 + kThreads - 1) / kThreads>(buf, tw, pl.n, TI, ns, sg);
    ns *= R;
  }
}

// phi transform; grid.x = row tiles, grid.y = z planes.  Reads `in`, writes
// `out` (they may be the same field).  rscale > 0 multiplies the input rows
// i < nr by r_i before the forward transform; rscale < 0 divides the output
// rows i < nr by r_i after the backward one (the r-weighting of vec2tp and
// tp2vec, mlegs_scalar_ops.f90:1330-1331, 1511-1512).
__global__ void __launch_bounds__(kThreads) k_fft_phi(double2 *out, const double2 *in, int nrdim, int npdim,
                                                      int nzdim, int TI, FftPlan pl, const double2 *twg,
                                                      const double2 *twh, int dir, const double *r, int nr,
                                                      int rscale) {
  __shared__ double2 buf[kLdsElems];
  __shared__ double2 tw[512];
  const int N = pl.n;  // np/2
  const int i0 = blockIdx.x * TI;
  const int k = blockIdx.y;
  const size_t plane = (size_t)nrdim * npdim * k;
  for (int t = threadIdx.x; t < N; t += kThreads) tw[t] = twg[t];
  if (dir < 0) {
    // load packed reals z_j = e(i, j, k), j < N
    for (int idx = threadIdx.x; idx < TI * N; idx += kThreads) {
      int l = idx % TI, j = idx / TI;
      int i = i0 + l;
      double2 v = make_double2(0.0, 0.0);
      if (i < nrdim) v = in[plane + (size_t)j * nrdim + i];
      if (rscale > 0 && i < nr) {
        const double rr = r[i];
        v = make_double2(v.x * rr, v.y * rr);
      }
      buf[j * TI + l] = v;
    }
    __syncthreads();
    lds_fft(buf, tw, pl, TI, -1.0);
    // untangle: C_m = E + W^m O, W = e^{-2 pi i/np}; scale 1/np
    const double sc = 1.0 / (2.0 * N);
    for (int idx = threadIdx.x; idx < TI * (N + 1); idx += kThreads) {
      int l = idx % TI, mm = idx / TI;
      int i = i0 + l;
      double2 zm = buf[(mm % N) * TI + l];
      double2 zc = buf[((N - mm) % N) * TI + l];
      zc.y = -zc.y;
      double2 E = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
      double2 D = make_double2(0.5 * (zm.x - zc.x), 0.5 * (zm.y - zc.y));
      double2 O = make_double2(D.y, -D.x);  // -i * D
      double2 C = cadd(E, cmul(twh[mm], O));
      if (i < nrdim) out[plane + (size_t)mm * nrdim + i] = make_double2(C.x * sc, C.y * sc);
    }
  } else {
    // Z_m = (C_m + conj C_{N-m}) + i (C_m - conj C_{N-m}) e^{+2 pi i m/np}
    for (int idx = threadIdx.x; idx < TI * N; idx += kThreads) {
      int l = idx % TI, mm = idx / TI;
      int i = i0 + l;
      double2 Z = make_double2(0.0, 0.0);
      if (i < nrdim) {
        double2 cm = in[plane + (size_t)mm * nrdim + i];
        double2 cc = in[plane + (size_t)(N - mm) * nrdim + i];
        if (mm == 0) { cm.y = 0.0; cc.y = 0.0; }  // imaginary parts of DC and Nyquist are ignored
        cc.y = -cc.y;
        double2 E = cadd(cm, cc);
        double2 w = twh[mm];
        w.y = -w.y;
        double2 O = cmul(csub(cm, cc), w);
        Z = make_double2(E.x - O.y, E.y + O.x);
      }
      buf[mm * TI + l] = Z;
    }
    __syncthreads();
    lds_fft(buf, tw, pl, TI, +1.0);
    for (int idx = threadIdx.x; idx < TI * (N + 1); idx += kThreads) {
      int l = idx % TI, j = idx / TI;
      int i = i0 + l;
      if (i >= nrdim) continue;
      double2 v = (j < N) ? buf[j * TI + l] : make_double2(0.0, 0.0);
      if (rscale < 0 && i < nr) {
        const double rr = r[i];
        v = make_double2(v.x / rr, v.y / rr);
      }
      out[plane + (size_t)j * nrdim + i] = v;
    }
  }
}

// z transform; grid.x = row tiles, grid.y = columns j
__global__ void __launch_bounds__(kThreads) k_fft_z(double2 *e, int nrdim, int npdim, int nzdim, int TI,
                                                    FftPlan pl, const double2 *twg, int dir) {
  __shared__ double2 buf[kLdsElems];
  __shared__ double2 tw[512];
  const int N = pl.n;
  const int i0 = blockIdx.x * TI;
  const int j = blockIdx.y;
  const size_t col = (size_t)nrdim * j;
  const size_t zs = (size_t)nrdim * npdim;
  for (int t = threadIdx.x; t < N; t += kThreads) tw[t] = twg[t];
  for (int idx = threadIdx.x; idx < TI * N; idx += kThreads) {
    int l = idx % TI, t = idx / TI;
    int i = i0 + l;
    double2 v = make_double2(0.0, 0.0);
    if (i < nrdim) v = e[col + (size_t)t * zs + i];
    buf[t * TI + l] = v;
  }
  __syncthreads();
  lds_fft(buf, tw, pl, TI, dir < 0 ? -1.0 : 1.0);
  const double sc = dir < 0 ? 1.0 / N : 1.0;
  for (int idx = threadIdx.x; idx < TI * N; idx += kThreads) {
    int l = idx % TI, t = idx / TI;
    int i = i0 + l;
    if (i >= nrdim) continue;
    double2 v = buf[t * TI + l];
    e[col + (size_t)t * zs + i] = make_double2(v.x * sc, v.y * sc);
  }
}

// ---------------------------------------------------------------------------
// Fast path for power-of-two lengths N = N1 * N2 (N1, N2 <= 16): four-step
// FFT.  Stage 1 loads x[n2 + N2 n1] of a line straight into registers, runs
// the N1-point DFT over n1 there and applies W_N^{n2 k1}; one LDS exchange;
// stage 2 runs the N2-point DFT over n2 and writes X[k1 + N1 k2].  A
// workgroup owns TI consecutive rows (lines) of one column (z) or plane (phi),
// so every global access is a run of TI consecutive complex numbers.
// The real phi transform adds its even/odd untangling (forward) or tangling
// (backward) through a second LDS exchange.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr double c16(int q) {
  // cos(2 pi q / 16), q = 0..7
  return q == 0 ? 1.0 : q == 1 ? 0.92387953251128675613 : q == 2 ? 0.70710678118654752440
       : q == 3 ? 0.38268343236508977173 : q == 4 ? 0.0 : q == 5 ? -0.38268343236508977173
       : q == 6 ? -0.70710678118654752440 : -0.92387953251128675613;
}
__host__ __device__ constexpr double s16(int q) {
  // sin(2 pi q / 16), q = 0..7
  return q == 0 ? 0.0 : q == 1 ? 0.38268343236508977173 : q == 2 ? 0.70710678118654752440
       : q == 3 ? 0.92387953251128675613 : q == 4 ? 1.0 : q == 5 ? 0.92387953251128675613
       : q == 6 ? 0.70710678118654752440 : 0.38268343236508977173;
}
__host__ __device__ constexpr int bitrev(int i, int n) {
  int r = 0;
  for (int b = 1; b < n; b <<= 1) {
    r = (r << 1) | (i & 1);
    i >>= 1;
  }
  return r;
}

// in-register radix-2 DIT DFT of size N (N <= 16), sign sg of the exponent
// FP32 storage of physical fields (two floats per packed entry)
typedef float mlx_fv2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_nt_f32(double2 *base, size_t idx, double2 x) {
  const mlx_fv2 v = {(float)x.x, (float)x.y};
  __builtin_nontemporal_store(v, reinterpret_cast<mlx_fv2 *>(base) + idx);
}
__device__ __forceinline__ double2 ld_nt_f32(const double2 *base, size_t idx) {
  const mlx_fv2 v = __builtin_nontemporal_load(reinterpret_cast<const mlx_fv2 *>(base) + idx);
  return make_double2((double)v.x, (double)v.y);
}

template <int N>
__device__ __forceinline__ void dft_reg(double2 *v, double sg) {
  double2 t[N];
#pragma unroll
  for (int i = 0; i < N; ++i) t[i] = v[bitrev(i, N)];
#pragma unroll
  for (int len = 2; len <= N; len <<= 1) {
    const int h = len / 2;
#pragma unroll
    for (int i0 = 0; i0 < N; i0 += len) {
#pragma unroll
      for (int j = 0; j < h; ++j) {
        const double2 a = t[i0 + j];
        double2 b = t[i0 + j + h];
        if (j != 0) {
          const int q = j * (16 / len);
          const double cw = c16(q), sw = sg * s16(q);
          b = make_double2(b.x * cw - b.y * sw, b.x * sw + b.y * cw);
        }
        t[i0 + j] = make_double2(a.x + b.x, a.y + b.y);
        t[i0 + j + h] = make_double2(a.x - b.x, a.y - b.y);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = t[i];
}

__device__ __forceinline__ double2 twid(const double2 *tw, int idx, double sg) {
  double2 w = tw[idx];
  if (sg > 0) w.y = -w.y;
  return w;
}

// MODE 0: complex transform (z); MODE 1: real forward (phi); MODE 2: real backward (phi)
// MODE 2: modes m >= mlim are taken as zero without being read.  MODE 0:
// line entries zin_lo..zin_hi are taken as zero without being read and
// outputs zout_lo..zout_hi are not written (their consumer does not read them)
// PAIR (MODE 2): the input is the backward radial contraction's float4
// (be, bo) per radial pair x < nrow at ((z npdim + m) pstride + x), npdim =
// gs / es (each line of pairs starts on a 256-byte boundary); a tile holds
// the pairs x0 .. x0 + TI/2 - 1 and synthesizes rows x = be + bo and
// nr - 1 - x = be - bo (formed in FP64 on load), rows < nr only.
// F32OUT (MODE 2): the output is stored in FP32 at f32_index (columns
// padded to 128-byte lines, half the bytes).
// MODE 0 with gridDim.z > 1: the in-place transform of the fields zf.p[z]
// (same shape and limits), one launch for all.
template <int N1, int N2, int TI, int MODE, bool PAIR = false, bool F32OUT = false>
__global__ void __launch_bounds__(kThreads) k_fft4(double2 *out, const double2 *in, size_t es, size_t gs, int nrow,
                                                   int act_nrc, const double2 *tw, const double2 *twh, double sg,
                                                   double scale, const double *r, int nr, int rscale, int mlim,
                                                   int zin_lo, int zin_hi, int zout_lo, int zout_hi,
                                                   int pstride, ZFields zf) {
  constexpr int N = N1 * N2;
  if (MODE == 0 && gridDim.z > 1) {
    out = zf.p[blockIdx.z];
    in = out;
  }
  constexpr int T1 = (TI * N2 + kThreads - 1) / kThreads;
  constexpr int T2 = (TI * N1 + kThreads - 1) / kThreads;
  constexpr bool HALF = (MODE == 0 && N == 256);
  __shared__ double2 buf[HALF ? TI * N / 2 : TI * N];
  __shared__ double2 nyq[MODE == 2 ? TI : 1];
  constexpr int TH = PAIR ? TI / 2 : TI;  // input rows (or pairs) per tile
  const int i0 = blockIdx.x * TH;
  const size_t base = (size_t)blockIdx.y * gs;
  int rlim = nrow;
  if (act_nrc > 0) {
    const int a = act_nrc - (int)blockIdx.y;
    rlim = a < 0 ? 0 : (a < nrow ? a : nrow);
  }
  if (i0 >= rlim) return;  // the whole tile lies outside the active rows
  const int tid = threadIdx.x;

  if (MODE == 2 && PAIR) {
    // stage 0 from radial pairs: tile rows l (x) and TH + l (nr - 1 - x)
    const float4 *pin = reinterpret_cast<const float4 *>(in);
    for (int b = tid; b < TH * (N + 1); b += kThreads) {
      const int l = b % TH, m = b / TH, x = i0 + l;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      // an ordinary load: the pairs were just written and are read from the cache
      if (x < rlim && m < mlim) v = pin[((size_t)blockIdx.y * (gs / es) + m) * pstride + x];
      const double ex = v.x, ey = v.y, ox = v.z, oy = v.w;
      const double2 hi = make_double2(ex + ox, ey + oy), lo = make_double2(ex - ox, ey - oy);
      if (m < N) {
        buf[m * TI + l] = hi;
        buf[m * TI + TH + l] = lo;
      } else {
        nyq[l] = hi;
        nyq[TH + l] = lo;
      }
    }
    __syncthreads();
  } else if (MODE == 2) {
    // stage 0: C_m, m = 0..N, into LDS (C_N apart)
    for (int b = tid; b < TI * (N + 1); b += kThreads) {
      const int l = b % TI, m = b / TI, i = i0 + l;
      const double2 v = (i < rlim && m < mlim) ? ld_nt(&in[base + (size_t)m * es + i]) : make_double2(0.0, 0.0);
      if (m < N) buf[m * TI + l] = v;
      else nyq[l] = v;
    }
    __syncthreads();
  }

  // stage 1
  double2 y1[T1][N1];
#pragma unroll
  for (int t = 0; t < T1; ++t) {
    const int b = tid + t * kThreads;
    if (b < TI * N2) {
      const int l = b % TI, n2 = b / TI, i = i0 + l;
      double2 v[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) {
        const int j = n2 + N2 * n1;
        if (MODE == 2) {
          double2 cm = buf[j * TI + l];
          double2 cc = (j == 0) ? nyq[l] : buf[(N - j) * TI + l];
          if (j == 0) { cm.y = 0.0; cc.y = 0.0; }  // imaginary parts of DC and Nyquist are ignored
          cc.y = -cc.y;
          const double2 E = cadd(cm, cc);
          double2 w = twh[j];
          w.y = -w.y;
          const double2 O = cmul(csub(cm, cc), w);
          v[n1] = make_double2(E.x - O.y, E.y + O.x);
        } else {
          const bool zin = MODE == 0 && j >= zin_lo && j <= zin_hi;
          double2 x = (i < rlim && !zin) ? ld_nt(&in[base + (size_t)j * es + i]) : make_double2(0.0, 0.0);
          if (MODE == 1 && rscale > 0 && i < nr) {
            const double rr = r[i];
            x = make_double2(x.x * rr, x.y * rr);
          }
          v[n1] = x;
        }
      }
      dft_reg<N1>(v, sg);
#pragma unroll
      for (int k1 = 1; k1 < N1; ++k1) {
        const int q = (n2 * k1) & (N - 1);
        if (q) v[k1] = cmul(v[k1], twid(tw, q, sg));
      }
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) y1[t][k1] = v[k1];
    }
  }
  if (MODE == 2) __syncthreads();  // stage 0 data fully consumed
  double2 y2[T2][N2];
  if (HALF) {
    // the exchange in two halves through TI N doubles: real parts, then
    // imaginary parts
    double *bd = reinterpret_cast<double *>(buf);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h) __syncthreads();
#pragma unroll
      for (int t = 0; t < T1; ++t) {
        const int b = tid + t * kThreads;
        if (b < TI * N2) {
          const int l = b % TI, n2 = b / TI;
#pragma unroll
          for (int k1 = 0; k1 < N1; ++k1) bd[(k1 * N2 + n2) * TI + l] = h ? y1[t][k1].y : y1[t][k1].x;
        }
      }
      __syncthreads();
#pragma unroll
      for (int t = 0; t < T2; ++t) {
        const int b = tid + t * kThreads;
        if (b < TI * N1) {
          const int l = b % TI, k1 = b / TI;
#pragma unroll
          for (int n2 = 0; n2 < N2; ++n2) {
            if (h) y2[t][n2].y = bd[(k1 * N2 + n2) * TI + l];
            else y2[t][n2].x = bd[(k1 * N2 + n2) * TI + l];
          }
        }
      }
    }
#pragma unroll
    for (int t = 0; t < T2; ++t)
      if (tid + t * kThreads < TI * N1) dft_reg<N2>(y2[t], sg);
  } else {
#pragma unroll
    for (int t = 0; t < T1; ++t) {
      const int b = tid + t * kThreads;
      if (b < TI * N2) {
        const int l = b % TI, n2 = b / TI;
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) buf[(k1 * N2 + n2) * TI + l] = y1[t][k1];
      }
    }
    __syncthreads();

    // stage 2
#pragma unroll
    for (int t = 0; t < T2; ++t) {
      const int b = tid + t * kThreads;
      if (b < TI * N1) {
        const int l = b % TI, k1 = b / TI;
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) y2[t][n2] = buf[(k1 * N2 + n2) * TI + l];
        dft_reg<N2>(y2[t], sg);
      }
    }
  }
  if (MODE == 0 || MODE == 2) {
#pragma unroll
    for (int t = 0; t < T2; ++t) {
      const int b = tid + t * kThreads;
      if (b < TI * N1) {
        const int l = b % TI, k1 = b / TI;
        // the output row of tile row l (PAIR: pair x = i0 + l % TH)
        const int lx = PAIR ? l % TH : l;
        const int i = (PAIR && l >= TH) ? nr - 1 - (i0 + lx) : i0 + lx;
        if (i0 + lx < rlim) {
          double dv = 1.0;
          if (MODE == 2 && rscale < 0 && i < nr) dv = r[i];
#pragma unroll
          for (int k2 = 0; k2 < N2; ++k2) {
            const int ko = k1 + N1 * k2;
            if (MODE == 0 && ko >= zout_lo && ko <= zout_hi) continue;
            double2 x = y2[t][k2];
            if (MODE == 0) x = make_double2(x.x * scale, x.y * scale);
            else if (dv != 1.0) x = make_double2(x.x / dv, x.y / dv);
            if (F32OUT) st_nt_f32(out, f32_index(blockIdx.y, ko, i, es, gs / es), x);
            else st_nt(&out[base + (size_t)ko * es + i], x);
          }
          if (MODE == 2 && k1 == 0) {
            if (F32OUT) st_nt_f32(out, f32_index(blockIdx.y, N, i, es, gs / es), make_double2(0.0, 0.0));
            else st_nt(&out[base + (size_t)N * es + i], make_double2(0.0, 0.0));
          }
        }
      }
    }
    return;
  }
  // MODE 1: natural order into LDS, then C_m = E + W^m O, m = 0..N, scaled 1/np
  __syncthreads();
#pragma unroll
  for (int t = 0; t < T2; ++t) {
    const int b = tid + t * kThreads;
    if (b < TI * N1) {
      const int l = b % TI, k1 = b / TI;
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) buf[(k1 + N1 * k2) * TI + l] = y2[t][k2];
    }
  }
  __syncthreads();
  for (int b = tid; b < TI * (N + 1); b += kThreads) {
    const int l = b % TI, m = b / TI, i = i0 + l;
    const double2 zm = buf[(m % N) * TI + l];
    double2 zc = buf[((N - m) % N) * TI + l];
    zc.y = -zc.y;
    const double2 E = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
    const double2 D = make_double2(0.5 * (zm.x - zc.x), 0.5 * (zm.y - zc.y));
    const double2 O = make_double2(D.y, -D.x);  // -i * D
    const double2 C = cadd(E, cmul(twh[m], O));
    if (i < rlim) st_nt(&out[base + (size_t)m * es + i], make_double2(C.x * scale, C.y * scale));
  }
}

// Forward phi transform of the three components of v x w, straight from the
// six physical (PPP, packed real) fields: vector_product
// (mlegs_scalar_ops.f90:1561-1566, here with the background axial flow ub
// added to v_z first, vortical_flow_3d.f90:381) followed by the first step
// of vector_projection (:1320-1336: r v_r, r v_phi and v_z to PFP).  Stage 1
// of the four-step transform forms the products while loading, so each
// input is read once and no physical product field is written.  Rows
// i >= nr give zero products, as vector_product leaves them.
// F32IN: the six velocity and vorticity fields are stored in FP32 at
// f32_index (ub, the background flow, stays FP64).
template <int N1, int N2, int TI, bool F32IN = false>
__global__ void __launch_bounds__(kThreads) k_fft4_cross(double2 *o0, double2 *o1, double2 *o2, const double2 *v0,
                                                         const double2 *v1, const double2 *v2, const double2 *w0,
                                                         const double2 *w1, const double2 *w2, const double2 *ub,
                                                         size_t es, size_t gs, int nrow, const double2 *tw,
                                                         const double2 *twh, double scale, const double *r, int nr,
                                                         double *ln0, double *ln1, double *ln2, int mmax) {
  constexpr int N = N1 * N2;
  constexpr int T1 = (TI * N2 + kThreads - 1) / kThreads;
  constexpr int T2 = (TI * N1 + kThreads - 1) / kThreads;
  const double sg = -1.0;
  __shared__ double2 buf[TI * N];
  const int i0 = blockIdx.x * TI;
  const size_t base = (size_t)blockIdx.y * gs;
  // the products carry no log term
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
    if (ln0) ln0[0] = 0.0;
    if (ln1) ln1[0] = 0.0;
    if (ln2) ln2[0] = 0.0;
  }
  if (i0 >= nrow) return;
  const int tid = threadIdx.x;

  // stage 1, all three components
  double2 y1[3][T1][N1];
#pragma unroll
  for (int t = 0; t < T1; ++t) {
    const int b = tid + t * kThreads;
    if (b < TI * N2) {
      const int l = b % TI, n2 = b / TI, i = i0 + l;
      double2 c0[N1], c1[N1], c2[N1];
      const double rr = (i < nr) ? r[i] : 0.0;
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) {
        const int j = n2 + N2 * n1;
        const double2 z2 = make_double2(0.0, 0.0);
        double2 a1 = z2, a2 = z2, a3 = z2, b1 = z2, b2 = z2, b3 = z2;
        if (i < nr) {
          const size_t o = base + (size_t)j * es + i;
          if (F32IN) {
            const size_t o32 = f32_index(blockIdx.y, j, i, es, gs / es);
            a1 = ld_nt_f32(v0, o32);
            a2 = ld_nt_f32(v1, o32);
            a3 = ld_nt_f32(v2, o32);
            b1 = ld_nt_f32(w0, o32);
            b2 = ld_nt_f32(w1, o32);
            b3 = ld_nt_f32(w2, o32);
          } else {
            a1 = ld_nt(&v0[o]);
            a2 = ld_nt(&v1[o]);
            a3 = ld_nt(&v2[o]);
            b1 = ld_nt(&w0[o]);
            b2 = ld_nt(&w1[o]);
            b3 = ld_nt(&w2[o]);
          }
          if (ub) {
            const double2 u = ld_nt(&ub[o]);
            a3 = make_double2(a3.x + u.x, a3.y + u.y);
          }
        }
        const double2 pr = make_double2(a2.x * b3.x - a3.x * b2.x, a2.y * b3.y - a3.y * b2.y);
        const double2 pp = make_double2(a3.x * b1.x - a1.x * b3.x, a3.y * b1.y - a1.y * b3.y);
        c2[n1] = make_double2(a1.x * b2.x - a2.x * b1.x, a1.y * b2.y - a2.y * b1.y);
        c0[n1] = make_double2(pr.x * rr, pr.y * rr);
        c1[n1] = make_double2(pp.x * rr, pp.y * rr);
      }
      dft_reg<N1>(c0, sg);
      dft_reg<N1>(c1, sg);
      dft_reg<N1>(c2, sg);
#pragma unroll
      for (int k1 = 1; k1 < N1; ++k1) {
        const int q = (n2 * k1) & (N - 1);
        if (q) {
          const double2 w = twid(tw, q, sg);
          c0[k1] = cmul(c0[k1], w);
          c1[k1] = cmul(c1[k1], w);
          c2[k1] = cmul(c2[k1], w);
        }
      }
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) {
        y1[0][t][k1] = c0[k1];
        y1[1][t][k1] = c1[k1];
        y1[2][t][k1] = c2[k1];
      }
    }
  }
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    double2 *out = comp == 0 ? o0 : comp == 1 ? o1 : o2;
    if (comp > 0) __syncthreads();  // the previous component's untangling has read buf
#pragma unroll
    for (int t = 0; t < T1; ++t) {
      const int b = tid + t * kThreads;
      if (b < TI * N2) {
        const int l = b % TI, n2 = b / TI;
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) buf[(k1 * N2 + n2) * TI + l] = y1[comp][t][k1];
      }
    }
    __syncthreads();
    double2 y2[T2][N2];
#pragma unroll
    for (int t = 0; t < T2; ++t) {
      const int b = tid + t * kThreads;
      if (b < TI * N1) {
        const int l = b % TI, k1 = b / TI;
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) y2[t][n2] = buf[(k1 * N2 + n2) * TI + l];
        dft_reg<N2>(y2[t], sg);
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < T2; ++t) {
      const int b = tid + t * kThreads;
      if (b < TI * N1) {
        const int l = b % TI, k1 = b / TI;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) buf[(k1 + N1 * k2) * TI + l] = y2[t][k2];
      }
    }
    __syncthreads();
    // C_m = E + W^m O, m = 0..N, scaled 1/np
    for (int b = tid; b < TI * (N + 1); b += kThreads) {
      const int l = b % TI, m = b / TI, i = i0 + l;
      const double2 zm = buf[(m % N) * TI + l];
      double2 zc = buf[((N - m) % N) * TI + l];
      zc.y = -zc.y;
      const double2 E = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
      const double2 D = make_double2(0.5 * (zm.x - zc.x), 0.5 * (zm.y - zc.y));
      const double2 O = make_double2(D.y, -D.x);  // -i * D
      const double2 C = cadd(E, cmul(twh[m], O));
      if (i < nrow && m < mmax) st_nt(&out[base + (size_t)m * es + i], make_double2(C.x * scale, C.y * scale));
    }
  }
}

// grid_y.z > 1 (MODE 0): the batched in-place transform of the fields zf.p[0 .. grid_y.z)
template <int MODE, bool PAIR = false, bool F32OUT = false>
bool launch_fft4(mlx_ctx *c, ProbeScope &ps, int N, dim3 grid_y, double2 *out, const double2 *in, size_t es,
                 size_t gs, int nrow, int act_nrc, const double2 *tw, const double2 *twh, double sg, double scale,
                 int rscale, int mlim = 1 << 30, int zin_lo = 1, int zin_hi = 0, int zout_lo = 1, int zout_hi = 0,
                 int pstride = 0, ZFields zf = ZFields()) {
  const KitDev &k = c->kit;
  if (grid_y.z > 1 && (MODE != 0 || grid_y.z > (unsigned)mlx::kZBatch))
    MLX_THROW(MLX_E_ARG, "fft: batched launch outside the z transform's limits");
#define MLX_FFT4(N1_, N2_, TI_)                                                                                    \
  do {                                                                                                             \
    constexpr int TH_ = PAIR ? TI_ / 2 : TI_;                                                                      \
    dim3 grid((nrow + TH_ - 1) / TH_, grid_y.y, grid_y.z);                                                         \
    probed_launch(ps, k_fft4<N1_, N2_, TI_, MODE, PAIR, F32OUT>, grid, dim3(kThreads), 0, c->stream, out, in, es, \
                  gs, nrow, act_nrc, tw, twh, sg, scale, k.r, k.nr, rscale, mlim, zin_lo, zin_hi, zout_lo,         \
                  zout_hi, pstride, zf);                                                                           \
    return true;                                                                                                   \
  } while (0)
  switch (N) {
    case 16: MLX_FFT4(4, 4, 64);
    case 32: MLX_FFT4(4, 8, 64);
    case 64: MLX_FFT4(8, 8, 32);
    case 128: MLX_FFT4(8, 16, 16);
    case 256:
      if (MODE == 0) MLX_FFT4(16, 16, 16);  // 64 KiB of LDS
      MLX_FFT4(16, 16, 8);                  // the real modes also hold the Nyquist column
    default: return false;
  }
#undef MLX_FFT4
}

int tile_rows(int n_line, int nrdim) {
  int ti = kLdsElems / (n_line + 1);
  if (ti > 64) ti = 64;
  if (ti > nrdim) ti = nrdim;
  if (ti < 1) ti = 1;
  return ti;
}

}  // namespace

namespace mlx {

bool fft_z_fast(const KitDev &k) {
  const int N = k.nz;
  return k.nz > 1 && (N == 16 || N == 32 || N == 64 || N == 128 || N == 256);
}

bool fft_phi_fast(const KitDev &k) {
  const int N = k.nph;
  return k.np > 1 && (N == 16 || N == 32 || N == 64 || N == 128 || N == 256);
}

void fft_phi(mlx_ctx *c, double2 *out, const double2 *in, int dir, int rscale, int mlim, bool phys_rows) {
  const KitDev &k = c->kit;
  if (k.np <= 1) return;
  if (mlim > 0 && dir > 0 && !fft_phi_fast(k)) MLX_THROW(MLX_E_ARG, "fft_phi: mode limit without the fast path");
  if (k.nph > 512) MLX_THROW(MLX_E_ARG, "fft_phi: np/2 above 512 is not supported");
  int TI = tile_rows(k.nph, k.nrdim);
  dim3 grid((k.nrdim + TI - 1) / TI, k.nzdim);
  // packed reals in, nph+1 modes out (or back): nr rows x nz slices
  const double nread = (mlim > 0 && mlim < k.nph + 1) ? mlim : k.nph + 1;
  ProbeScope probe(c, MLX_PROBE_FFT_PHI, 16.0 * k.nr * k.nz * (dir > 0 ? nread + k.nph : 2.0 * k.nph + 1),
                   2.5 * k.np * std::log2((double)k.np) * k.nr * k.nz);
  const size_t plane = (size_t)k.nrdim * k.npdim;
  bool fast = (dir < 0) ? launch_fft4<1>(c, probe, k.nph, dim3(1, k.nzdim), out, in, k.nrdim, plane, k.nrdim, 0,
                                         k.tw_nph, k.tw_half, -1.0, 1.0 / k.np, rscale)
                        : launch_fft4<2>(c, probe, k.nph, dim3(1, k.nzdim), out, in, k.nrdim, plane,
                                         phys_rows ? k.nr : k.nrdim, 0, k.tw_nph, k.tw_half, +1.0, 1.0, rscale,
                                         mlim > 0 ? mlim : 1 << 30);
  if (fast) {
    HIPCHK(hipGetLastError());
    return;
  }
  probed_launch(probe, k_fft_phi, grid, dim3(kThreads), 0, c->stream, out, in, k.nrdim, k.npdim, k.nzdim, TI,
                     k.plan_nph, k.tw_nph, k.tw_half, dir, k.r, k.nr, rscale);
  HIPCHK(hipGetLastError());
}

// The backward phi transform of the radial pairs that leg_backward_pairs_f32
// writes: rows x and nr - 1 - x (x < nr / 2) of the PPP output, rows >= nr
// left as they are.  mlim: modes m >= mlim are taken as zero.
void fft_phi_pairs(mlx_ctx *c, double2 *out, const float4 *pairs, int rscale, int mlim, bool f32out) {
  const KitDev &k = c->kit;
  if (!fft_phi_fast(k) || (k.nr & 1)) MLX_THROW(MLX_E_ARG, "fft_phi: radial pairs need the four-step kernels");
  const double nread = (mlim > 0 && mlim < k.nph + 1) ? mlim : k.nph + 1;
  ProbeScope probe(c, MLX_PROBE_FFT_PHI, 8.0 * k.nr * k.nz * nread + (f32out ? 8.0 : 16.0) * k.nr * k.nz * k.nph,
                   2.5 * k.np * std::log2((double)k.np) * k.nr * k.nz);
  const size_t plane = (size_t)k.nrdim * k.npdim;
  const double2 *pin = reinterpret_cast<const double2 *>(pairs);
  const int ml = mlim > 0 ? mlim : 1 << 30;
  if (f32out)
    launch_fft4<2, true, true>(c, probe, k.nph, dim3(1, k.nzdim), out, pin, k.nrdim, plane, k.nrh, 0, k.tw_nph,
                               k.tw_half, +1.0, 1.0, rscale, ml, 1, 0, 1, 0, pair_stride(k));
  else
    launch_fft4<2, true>(c, probe, k.nph, dim3(1, k.nzdim), out, pin, k.nrdim, plane, k.nrh, 0, k.tw_nph, k.tw_half,
                         +1.0, 1.0, rscale, ml, 1, 0, 1, 0, pair_stride(k));
  HIPCHK(hipGetLastError());
}

bool fft_phi_cross_ok(const KitDev &k) {
  const int N = k.nph;
  return k.np > 1 && (N == 16 || N == 32 || N == 64 || N == 128 || N == 256);
}

bool fft_phi_cross(mlx_ctx *c, double2 *o0, double2 *o1, double2 *o2, const double2 *v0, const double2 *v1,
                   const double2 *v2, const double2 *w0, const double2 *w1, const double2 *w2, const double2 *ub,
                   double *ln0, double *ln1, double *ln2, int mmax, bool f32in) {
  const KitDev &k = c->kit;
  const int N = k.nph;
  if (mmax <= 0 || mmax > N + 1) mmax = N + 1;
  if (!fft_phi_cross_ok(k)) return false;
  const size_t plane = (size_t)k.nrdim * k.npdim;
  // six physical fields (FP32 or FP64) and ub read (rows < nr), three PFP
  // fields written
  ProbeScope probe(c, MLX_PROBE_PHI_CROSS, k.nr * k.nz * ((f32in ? 48.0 : 96.0) * N + 16.0 * N + 48.0 * mmax),
                   3.0 * 2.5 * k.np * std::log2((double)k.np) * k.nr * k.nz + 18.0 * k.nr * k.np * k.nz);
#define MLX_FFTX(N1_, N2_, TI_)                                                                                 \
  do {                                                                                                          \
    dim3 grid((k.nrdim + TI_ - 1) / TI_, k.nzdim);                                                              \
    if (f32in)                                                                                                  \
      probed_launch(probe, k_fft4_cross<N1_, N2_, TI_, true>, grid, dim3(kThreads), 0, c->stream, o0, o1, o2, v0,  \
                    v1, v2, w0, w1, w2, ub, (size_t)k.nrdim, plane, k.nrdim, k.tw_nph, k.tw_half, 1.0 / k.np, k.r, \
                    k.nr, ln0, ln1, ln2, mmax);                                                                 \
    else                                                                                                        \
      probed_launch(probe, k_fft4_cross<N1_, N2_, TI_, false>, grid, dim3(kThreads), 0, c->stream, o0, o1, o2, v0, \
                    v1, v2, w0, w1, w2, ub, (size_t)k.nrdim, plane, k.nrdim, k.tw_nph, k.tw_half, 1.0 / k.np, k.r, \
                    k.nr, ln0, ln1, ln2, mmax);                                                                 \
  } while (0)
  switch (N) {
    case 16: MLX_FFTX(4, 4, 64); break;
    case 32: MLX_FFTX(4, 8, 64); break;
    case 64: MLX_FFTX(8, 8, 32); break;
    case 128: MLX_FFTX(8, 16, 16); break;
    default: MLX_FFTX(16, 16, 8); break;
  }
#undef MLX_FFTX
  HIPCHK(hipGetLastError());
  return true;
}

void fft_z_batch(mlx_ctx *c, double2 *const *e, int nf, int dir, double ncol, int act_nrc, int ncols, int zin_lo,
                 int zin_hi, int zout_lo, int zout_hi) {
  const KitDev &k = c->kit;
  if (k.nz <= 1 || nf <= 0) return;
  if (k.nz > 512) MLX_THROW(MLX_E_ARG, "fft_z: nz above 512 is not supported");
  if (nf > kZBatch) MLX_THROW(MLX_E_ARG, "fft_z: at most kZBatch fields per launch");
  if (ncols <= 0 || ncols > k.npdim) ncols = k.npdim;
  auto span = [&](int lo, int hi) {
    lo = lo < 0 ? 0 : lo;
    hi = hi > k.nz - 1 ? k.nz - 1 : hi;
    return hi >= lo ? hi - lo + 1 : 0;
  };
  // entries read and written per line (the skipped ranges neither)
  ProbeScope probe(c, MLX_PROBE_FFT_Z,
                   16.0 * nf * ncol * (2.0 * k.nz - span(zin_lo, zin_hi) - span(zout_lo, zout_hi)),
                   5.0 * nf * k.nz * std::log2((double)k.nz) * ncol);
  ZFields zf;
  for (int f = 0; f < kZBatch; ++f) zf.p[f] = e[f < nf ? f : 0];
  if (launch_fft4<0>(c, probe, k.nz, dim3(1, ncols, nf), e[0], e[0], (size_t)k.nrdim * k.npdim, k.nrdim, k.nrdim,
                     act_nrc, k.tw_nz, nullptr, dir < 0 ? -1.0 : 1.0, dir < 0 ? 1.0 / k.nz : 1.0, 0, 1 << 30, zin_lo,
                     zin_hi, zout_lo, zout_hi, 0, zf)) {
    HIPCHK(hipGetLastError());
    return;
  }
  // other lengths: the Stockham kernel, one launch per field
  int TI = tile_rows(k.nz, k.nrdim);
  dim3 grid((k.nrdim + TI - 1) / TI, k.npdim);
  for (int f = 0; f < nf; ++f)
    probed_launch(probe, k_fft_z, grid, dim3(kThreads), 0, c->stream, e[f], k.nrdim, k.npdim, k.nzdim, TI,
                  k.plan_nz, k.tw_nz, dir);
  HIPCHK(hipGetLastError());
}

void fft_z(mlx_ctx *c, double2 *e, int dir, double ncol, int act_nrc, int ncols, int zin_lo, int zin_hi,
           int zout_lo, int zout_hi) {
  fft_z_batch(c, &e, 1, dir, ncol, act_nrc, ncols, zin_lo, zin_hi, zout_lo, zout_hi);
}

}  // namespace mlx
