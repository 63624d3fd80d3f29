// Legendre contractions on MFMA, in FP32 (v_mfma_f32_16x16x4_f32) or FP64
// (v_mfma_f64_16x16x4_f64).
//
// Same contractions as legendre.hip (rtrans_forward / rtrans_backward /
// vector_projection, src/submodules/mlegs_scalar_ops.f90:1852-2008,
// 1387-1436).  Fields stay FP64 in HBM; the FP32 kernels round them to FP32
// when staging into LDS and accumulate in FP32, the FP64 kernels stage and
// accumulate in FP64 (the same kernels, templated on the operand type).
// These kernels also write every entry of their output (zeros outside the
// active set), so no separate zero-fill pass is needed.
//
// GEMM orientation.  Both directions are computed transposed: the MFMA "A"
// operand is the field (rows = real columns j = 2k + re/im, contraction
// dimension in LDS), the "B" operand is the table (columns = radial index).
// The FP32 instruction's 16x16 accumulator holds rows 4(l/16)+r, column l%16
// in element r of lane l; the FP64 instruction's holds rows l/16 + 4r
// (measured on gfx950).  The FP64 kernels feed A row i the field column
// 4(i%4) + i/4, so in both types element r of lane l holds field column
// 4(l/16)+r: each lane has the (re, im) pairs of two axial slices for one
// radial index, and rows 2a and 2a+1 (both parities) or x and nr-1-x are
// written as contiguous FP64 pairs.
//
// Work decomposition: one workgroup (4 wavefronts) per (m, 32 axial slices);
// wave w owns 16-row tiles w, w+4, ... of the radial index.  Contraction in
// chunks of KC = 8 staged in LDS.
#include "mlx_internal.h"

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int KC = 8;         // contraction chunk
constexpr int KT = 32;        // axial slices per workgroup
constexpr int BJ = 2 * KT;    // real columns per workgroup
// LDS row pitches are 16 (mod 64) elements: the four rows an MFMA operand
// read touches (lane / 16) then fall in disjoint banks
constexpr int BJP = BJ + 16;  // padded LDS row (elements)

// the MFMA of one operand type: 16x16 accumulator (4 per lane), 16-byte
// table loads of VE elements
// Bf3: 3-term split-BF16.  Operands are staged in FP32 (from FP64, and the
// FP32 tables); at the MFMA each FP32 value x is split exactly into three
// BF16 terms h + m + l (h = bf16(x), m = bf16(x - h), l = bf16(x - h - m):
// 24 significant bits), and the six products h h, h m, m h, h l, m m, l h
// accumulate in FP32 on v_mfma_f32_16x16x16_bf16 (products of BF16 pairs
// are exact in FP32; the dropped m l, l m, l l are below 2^-24).
struct Bf3 {};
typedef short shortx4 __attribute__((ext_vector_type(4)));

// Per operand type: the MFMA accumulator, the 16-byte table load, the LDS
// staging type, the contraction chunk (KCH) staged per barrier, the k step of
// one instruction (KSTEP) and the values a lane feeds it per operand (NV, at
// rows kk + NV (l/16) + j of the chunk), the packed operand (pk) and its MFMA.
template <typename T>
struct Mma;
template <>
struct Mma<float> {
  typedef floatx4 acc;
  typedef float4 vec;
  typedef float stage;
  typedef float pk;
  static constexpr int VE = 4, KCH = 8, KSTEP = 4, NV = 1;
  __device__ static inline acc op(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline pk prep(const stage *v) { return v[0]; }
  __device__ static inline acc mma(pk a, pk b, acc c) { return op(a, b, c); }
  __device__ static inline acc zero() { return acc{0.f, 0.f, 0.f, 0.f}; }
  __device__ static inline vec vzero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
  // field column that A row i carries (accumulator element r of lane l is row 4(l/16)+r)
  __device__ static inline int arow(int i) { return i; }
};
template <>
struct Mma<double> {
  typedef doublex4 acc;
  typedef double2 vec;
  typedef double stage;
  typedef double pk;
  static constexpr int VE = 2, KCH = 8, KSTEP = 4, NV = 1;
  __device__ static inline acc op(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline pk prep(const stage *v) { return v[0]; }
  __device__ static inline acc mma(pk a, pk b, acc c) { return op(a, b, c); }
  __device__ static inline acc zero() { return acc{0.0, 0.0, 0.0, 0.0}; }
  __device__ static inline vec vzero() { return make_double2(0.0, 0.0); }
  // accumulator element r of lane l is row l/16 + 4r: A row i carries field
  // column 4(i%4) + i/4, which puts column 4(l/16)+r in that element
  __device__ static inline int arow(int i) { return 4 * (i & 3) + (i >> 2); }
};
template <>
struct Mma<Bf3> {
  typedef floatx4 acc;
  typedef float4 vec;
  typedef float stage;
  struct pk {
    shortx4 h, m, l;
  };
  static constexpr int VE = 4, KCH = 16, KSTEP = 16, NV = 4;
  __device__ static inline acc zero() { return acc{0.f, 0.f, 0.f, 0.f}; }
  __device__ static inline vec vzero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
  // the BF16 accumulator layout is the FP32 one (element r of lane l is row 4(l/16)+r)
  __device__ static inline int arow(int i) { return i; }
  __device__ static inline short bits(__bf16 x) { return __builtin_bit_cast(short, x); }
  __device__ static inline pk prep(const stage *v) {
    pk o;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float x = v[j];
      const __bf16 h = (__bf16)x;
      const float r1 = x - (float)h;
      const __bf16 m = (__bf16)r1;
      const __bf16 l = (__bf16)(r1 - (float)m);
      o.h[j] = bits(h);
      o.m[j] = bits(m);
      o.l[j] = bits(l);
    }
    return o;
  }
  __device__ static inline acc mm1(shortx4 a, shortx4 b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  }
  // the small products first
  __device__ static inline acc mma(const pk &a, const pk &b, acc c) {
    c = mm1(a.l, b.h, c);
    c = mm1(a.m, b.m, c);
    c = mm1(a.h, b.l, c);
    c = mm1(a.m, b.h, c);
    c = mm1(a.h, b.m, c);
    return mm1(a.h, b.h, c);
  }
};


// ---------------------------------------------------------------------------
// forward: out(2a+p, m, k) = sum_x P_p(a, x) fold_p(x, k)
//   table pfg[m][p][x][a] (a padded to AP), fold from FP64 input
// ---------------------------------------------------------------------------
template <typename T, int AT>  // 16-row tiles per wave: AP = 4 * AT * 16
__global__ void __launch_bounds__(256) k_leg_fwd_mf(const double2 *__restrict__ in, double2 *__restrict__ out,
                                                    const typename Mma<T>::stage *__restrict__ pfg, const double *__restrict__ w,
                                                     int nrdim, int npdim, int nz, int nr, int nrh, int nrc,
                                                     int npc, int nyt) {
  constexpr int AP = 4 * AT * 16;
  constexpr int APP = AP + 16;
  typedef Mma<T> M;
  typedef typename M::vec vec;
  typedef typename M::stage S;
  constexpr int VE = M::VE;
  constexpr int KC_ = M::KCH;
  __shared__ S Bs[2][KC_][BJP];
  __shared__ S As[2][KC_][APP];
  int m, yt;
  xcd_map(blockIdx.x, nyt, m, yt);
  if (m >= npdim) return;
  const int k0 = yt * KT;
  const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
  const size_t zs = (size_t)nrdim * npdim;
  double2 *ocol = out + (size_t)nrdim * m;
  if (m >= npc) {  // inactive column: zero
    for (int idx = tid; idx < nrdim * KT; idx += 256) {
      int i = idx % nrdim, kl = idx / nrdim;
      if (k0 + kl < nz) ocol[(size_t)(k0 + kl) * zs + i] = make_double2(0.0, 0.0);
    }
    return;
  }
  const int nn = nrc - m > 0 ? nrc - m : 0;
  const int Mmax = (nn + 1) / 2;
  const double2 *icol = in + (size_t)nrdim * m;
  const S *Tg = pfg + (size_t)m * 2 * nrh * AP;
  typename M::acc acc[AT][4][2];
#pragma unroll
  for (int s = 0; s < AT; ++s)
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int p = 0; p < 2; ++p) acc[s][t][p] = M::zero();

  // Software pipeline: the next chunk's table and field loads are issued
  // into registers before the current chunk's MFMA work and stored to LDS
  // after it.  Table: 2 KC_ AP / VE 16-byte loads per chunk; field: KC_ KT points,
  // each the pair x, nr-1-x and the weight, folded when stored.
  constexpr int NTA = (2 * KC_ * (AP / VE) + 255) / 256;
  constexpr int NFA = (KC_ * KT + 255) / 256;
  vec rt[NTA];
  double2 ra[NFA], rb[NFA];
  double rw[NFA];
  auto fetch = [&](int x0) {
#pragma unroll
    for (int q = 0; q < NTA; ++q) {
      const int idx = tid + 256 * q;
      vec v = M::vzero();
      if (idx < 2 * KC_ * (AP / VE)) {
        const int a = VE * (idx % (AP / VE)), x = (idx / (AP / VE)) % KC_, p = idx / ((AP / VE) * KC_);
        const int xg = x0 + x;
        if (xg < nrh) v = *reinterpret_cast<const vec *>(Tg + ((size_t)p * nrh + xg) * AP + a);
      }
      rt[q] = v;
    }
#pragma unroll
    for (int q = 0; q < NFA; ++q) {
      const int idx = tid + 256 * q;
      const int x = idx % KC_, kl = idx / KC_;
      const int xg = x0 + x, kg = k0 + kl;
      ra[q] = rb[q] = make_double2(0.0, 0.0);
      rw[q] = 0.0;
      if (idx < KC_ * KT && xg < nrh && kg < nz) {
        ra[q] = icol[(size_t)kg * zs + xg];
        rb[q] = icol[(size_t)kg * zs + (nr - 1 - xg)];
        rw[q] = w[xg];
      }
    }
  };
  fetch(0);
  for (int x0 = 0; x0 < nrh; x0 += KC_) {
#pragma unroll
    for (int q = 0; q < NTA; ++q) {
      const int idx = tid + 256 * q;
      if (idx < 2 * KC_ * (AP / VE)) {
        const int a = VE * (idx % (AP / VE)), x = (idx / (AP / VE)) % KC_, p = idx / ((AP / VE) * KC_);
        *reinterpret_cast<vec *>(&As[p][x][a]) = rt[q];
      }
    }
#pragma unroll
    for (int q = 0; q < NFA; ++q) {
      const int idx = tid + 256 * q;
      if (idx < KC_ * KT) {
        const int x = idx % KC_, kl = idx / KC_;
        const double2 a = ra[q], b = rb[q];
        const double wx = rw[q];
        Bs[0][x][2 * kl] = (S)((a.x + b.x) * wx);
        Bs[0][x][2 * kl + 1] = (S)((a.y + b.y) * wx);
        Bs[1][x][2 * kl] = (S)((a.x - b.x) * wx);
        Bs[1][x][2 * kl + 1] = (S)((a.y - b.y) * wx);
      }
    }
    __syncthreads();
    if (x0 + KC_ < nrh) fetch(x0 + KC_);
#pragma unroll 2
    for (int kk = 0; kk < KC_; kk += M::KSTEP) {
      const int xr = kk + M::NV * (lane / 16);
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        typename M::pk fa[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          S v[M::NV];
#pragma unroll
          for (int j = 0; j < M::NV; ++j) v[j] = Bs[p][xr + j][16 * t + M::arow(lane % 16)];
          fa[t] = M::prep(v);
        }
#pragma unroll
        for (int s = 0; s < AT; ++s) {
          const int at = wv + 4 * s;
          if (at * 16 >= Mmax) continue;
          S v[M::NV];
#pragma unroll
          for (int j = 0; j < M::NV; ++j) v[j] = As[p][xr + j][at * 16 + lane % 16];
          const typename M::pk fb = M::prep(v);
#pragma unroll
          for (int t = 0; t < 4; ++t) acc[s][t][p] = M::mma(fa[t], fb, acc[s][t][p]);
        }
      }
    }
    __syncthreads();
  }
  // epilogue: lane holds (re, im) of slices kl, kl+1 for rows 2a, 2a+1
#pragma unroll
  for (int s = 0; s < AT; ++s) {
    const int a = (wv + 4 * s) * 16 + lane % 16;
    const int n0 = 2 * a;
    if (n0 >= nrdim) continue;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int kl = 8 * t + 2 * (lane / 16) + h;
        const int kg = k0 + kl;
        if (kg >= nz) continue;
        double2 ve = (n0 < nn) ? make_double2(acc[s][t][0][2 * h], acc[s][t][0][2 * h + 1]) : make_double2(0.0, 0.0);
        double2 vo = (n0 + 1 < nn) ? make_double2(acc[s][t][1][2 * h], acc[s][t][1][2 * h + 1])
                                   : make_double2(0.0, 0.0);
        double2 *o = ocol + (size_t)kg * zs + n0;
        o[0] = ve;
        if (n0 + 1 < nrdim) o[1] = vo;
      }
    }
  }
  // rows beyond the tiles: zero
  for (int idx = tid; idx < (nrdim - 2 * AP > 0 ? nrdim - 2 * AP : 0) * KT; idx += 256) {
    int span = nrdim - 2 * AP;
    int i = 2 * AP + idx % span, kl = idx / span;
    if (k0 + kl < nz) ocol[(size_t)(k0 + kl) * zs + i] = make_double2(0.0, 0.0);
  }
}

// ---------------------------------------------------------------------------
// backward: f(x) = be + bo, f(-x) = be - bo, b_p(x, k) = sum_a P(x, 2a+p) e(2a+p, k)
//   table pf32[m][n][x]
// PAIRS: the FP32 sums themselves are written, (be, bo) of radial node x as
// one float4 at pairs[(k npdim + m) pstride + x], x < nr / 2: half the bytes
// of the two FP64 rows, which a consumer forms exactly as be + bo and be - bo
// in FP64.  Only columns m < npc are written.
// ---------------------------------------------------------------------------
template <typename T, int XT, bool PAIRS>  // 16-column tiles of x per wave: XP = 4 * XT * 16 >= nrh
__global__ void __launch_bounds__(256) k_leg_bwd_mf(const double2 *__restrict__ in, double2 *__restrict__ out,
                                                    float4 *__restrict__ pairs, const typename Mma<T>::stage *__restrict__ ptab,
                                                     int nrdim, int npdim, int nz, int nr, int nrh, int ne, int nrc,
                                                     int npc, int nyt, int pstride) {
  constexpr int XP = 4 * XT * 16;
  constexpr int XPP = XP + 16;
  typedef Mma<T> M;
  typedef typename M::vec vec;
  typedef typename M::stage S;
  constexpr int VE = M::VE;
  constexpr int KC_ = M::KCH;
  static_assert(!PAIRS || sizeof(S) == 4, "radial pairs are FP32 sums");
  __shared__ S Es[2][KC_][BJP];
  __shared__ S Ps[2][KC_][XPP];
  int m, yt;
  xcd_map(blockIdx.x, nyt, m, yt);
  if (m >= npdim) return;
  const int k0 = yt * KT;
  const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
  const size_t zs = (size_t)nrdim * npdim;
  double2 *ocol = out + (size_t)nrdim * m;
  if (m >= npc) {
    if (PAIRS) return;  // columns m >= npc are not written
    for (int idx = tid; idx < nrdim * KT; idx += 256) {
      int i = idx % nrdim, kl = idx / nrdim;
      if (k0 + kl < nz) ocol[(size_t)(k0 + kl) * zs + i] = make_double2(0.0, 0.0);
    }
    return;
  }
  const int nn = nrc - m > 0 ? nrc - m : 0;
  const int Me = (nn + 1) / 2;
  const double2 *icol = in + (size_t)nrdim * m;
  const S *P = ptab + (size_t)m * ne * nrh;
  typename M::acc acc[XT][4][2];
#pragma unroll
  for (int s = 0; s < XT; ++s)
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int p = 0; p < 2; ++p) acc[s][t][p] = M::zero();

  // MFMA work of one staged chunk (both parities)
  auto mfma_chunk = [&]() {
#pragma unroll 2
    for (int kk = 0; kk < KC_; kk += M::KSTEP) {
      const int ar = kk + M::NV * (lane / 16);
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        typename M::pk fa[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          S v[M::NV];
#pragma unroll
          for (int j = 0; j < M::NV; ++j) v[j] = Es[p][ar + j][16 * t + M::arow(lane % 16)];
          fa[t] = M::prep(v);
        }
#pragma unroll
        for (int s = 0; s < XT; ++s) {
          const int xt = wv + 4 * s;
          if (xt * 16 >= nrh) continue;
          S v[M::NV];
#pragma unroll
          for (int j = 0; j < M::NV; ++j) v[j] = Ps[p][ar + j][xt * 16 + lane % 16];
          const typename M::pk fb = M::prep(v);
#pragma unroll
          for (int t = 0; t < 4; ++t) acc[s][t][p] = M::mma(fa[t], fb, acc[s][t][p]);
        }
      }
    }
  };
  // field chunk rows 2 a0 .. 2 a0 + 2 KC_ - 1 of 32 slices (FP64 -> T)
  auto field_at = [&](int a0, int idx) {
    const int r = idx % (2 * KC_), kl = idx / (2 * KC_);
    const int n = 2 * a0 + r, kg = k0 + kl;
    return (n < nn && kg < nz) ? icol[(size_t)kg * zs + n] : make_double2(0.0, 0.0);
  };
  auto field_put = [&](int idx, double2 v) {
    const int r = idx % (2 * KC_), kl = idx / (2 * KC_);
    Es[r & 1][r >> 1][2 * kl] = (S)v.x;
    Es[r & 1][r >> 1][2 * kl + 1] = (S)v.y;
  };
  if ((nrh % VE) == 0) {
    // software pipeline: the next chunk's global loads are in flight while
    // the current chunk runs on MFMA; VE consecutive x per 16-byte load
    constexpr int NTB = (2 * KC_ * (XP / VE) + 255) / 256;
    constexpr int NFB = (2 * KC_ * KT + 255) / 256;
    vec rt[NTB];
    double2 rf[NFB];
    auto fetch = [&](int a0) {
#pragma unroll
      for (int q = 0; q < NTB; ++q) {
        const int idx = tid + 256 * q;
        vec v = M::vzero();
        if (idx < 2 * KC_ * (XP / VE)) {
          const int x = VE * (idx % (XP / VE)), a = (idx / (XP / VE)) % KC_, p = idx / ((XP / VE) * KC_);
          const int n = 2 * (a0 + a) + p;
          if (x < nrh && n < nn) v = *reinterpret_cast<const vec *>(P + (size_t)n * nrh + x);
        }
        rt[q] = v;
      }
#pragma unroll
      for (int q = 0; q < NFB; ++q) {
        const int idx = tid + 256 * q;
        rf[q] = (idx < 2 * KC_ * KT) ? field_at(a0, idx) : make_double2(0.0, 0.0);
      }
    };
    fetch(0);
    for (int a0 = 0; a0 < Me; a0 += KC_) {
#pragma unroll
      for (int q = 0; q < NTB; ++q) {
        const int idx = tid + 256 * q;
        if (idx < 2 * KC_ * (XP / VE)) {
          const int x = VE * (idx % (XP / VE)), a = (idx / (XP / VE)) % KC_, p = idx / ((XP / VE) * KC_);
          *reinterpret_cast<vec *>(&Ps[p][a][x]) = rt[q];
        }
      }
#pragma unroll
      for (int q = 0; q < NFB; ++q) {
        const int idx = tid + 256 * q;
        if (idx < 2 * KC_ * KT) field_put(idx, rf[q]);
      }
      __syncthreads();
      if (a0 + KC_ < Me) fetch(a0 + KC_);
      mfma_chunk();
      __syncthreads();
    }
  } else {
    for (int a0 = 0; a0 < Me; a0 += KC_) {
      for (int idx = tid; idx < 2 * KC_ * XP; idx += 256) {
        int x = idx % XP, a = (idx / XP) % KC_, p = idx / (XP * KC_);
        int n = 2 * (a0 + a) + p;
        Ps[p][a][x] = (x < nrh && n < nn) ? P[(size_t)n * nrh + x] : (S)0;
      }
      for (int idx = tid; idx < 2 * KC_ * KT; idx += 256) field_put(idx, field_at(a0, idx));
      __syncthreads();
      mfma_chunk();
      __syncthreads();
    }
  }
#pragma unroll
  for (int s = 0; s < XT; ++s) {
    const int x = (wv + 4 * s) * 16 + lane % 16;
    if (x >= nrh) continue;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int kg = k0 + 8 * t + 2 * (lane / 16) + h;
        if (kg >= nz) continue;
        if (PAIRS) {
          pairs[((size_t)kg * npdim + m) * pstride + x] =
              make_float4((float)acc[s][t][0][2 * h], (float)acc[s][t][0][2 * h + 1], (float)acc[s][t][1][2 * h],
                          (float)acc[s][t][1][2 * h + 1]);
          continue;
        }
        double ber = acc[s][t][0][2 * h], bei = acc[s][t][0][2 * h + 1];
        double bor = acc[s][t][1][2 * h], boi = acc[s][t][1][2 * h + 1];
        double2 *o = ocol + (size_t)kg * zs;
        o[x] = make_double2(ber + bor, bei + boi);
        o[nr - 1 - x] = make_double2(ber - bor, bei - boi);
      }
    }
  }
  if (PAIRS) return;
  for (int idx = tid; idx < (nrdim - nr) * KT; idx += 256) {
    int span = nrdim - nr;
    int i = nr + idx % span, kl = idx / span;
    if (k0 + kl < nz) ocol[(size_t)(k0 + kl) * zs + i] = make_double2(0.0, 0.0);
  }
}

// ---------------------------------------------------------------------------
// vec2tp projection on MFMA, one parity per workgroup:
//   psi = V.X1 + D.X2,   X1 = -i m eo(ur),  X2 = -oe(up)
//   chi = D.X3 + V.X4 + T.X5,  X3 = i kv oe(ur),  X4 = m kv eo(up),  X5 = -eo(uz)
// eo/oe: even/odd folds b(x) +- b(-x) chosen by the row parity (eomul /
// oemul, mlegs_scalar_ops.f90:2067-2145).  Only rows < nrchop - m are
// computed: the three rows of the radial offset are chopped afterwards by
// the reference.  Tables tg[m][p][x][a] (V, D, T).
// ---------------------------------------------------------------------------
template <typename T, int AT>
__global__ void __launch_bounds__(256) k_v2t_mf(const double2 *__restrict__ ur, const double2 *__restrict__ up,
                                                const double2 *__restrict__ uz, double2 *__restrict__ psi,
                                                double2 *__restrict__ chi, const T *__restrict__ tgV,
                                                const T *__restrict__ tgD, const T *__restrict__ tgT,
                                                 const double *__restrict__ ak, int nrdim, int npdim, int nz, int nr,
                                                 int nrh, int nrc0, int npc, int nzc, int nzcu, int nyt, int zlo,
                                                 int zhi, int mbeg, int mend, int asplit, int tpitch, int ksplit,
                                                 size_t zso, size_t kstride) {
  // AP: the table columns (radial rows a of one parity) this workgroup
  // stages; tpitch: the tables' row pitch (4 16 tiles of the widest kernel)
  constexpr int AP = 4 * AT * 16;
  constexpr int APP = AP + 16;
  constexpr int KV = 8;   // contraction chunk for this kernel
  typedef Mma<T> MM;
  typedef typename MM::vec vec;
  constexpr int VE = MM::VE;
  __shared__ T X[5][KV][BJP];
  __shared__ T Tb[3][KV][APP];
  // both parities and every kappa tile of a column m on one XCD: the two
  // parity workgroups read the same input columns and write interleaved
  // rows of the same output lines, which then meet in that XCD's L2
  // this launch's columns: mbeg .. mend-1.  asplit workgroups share the
  // rows of one (m, parity, kappa tile): workgroup sp takes the 16-row tiles
  // wv + 4 (sp AT + s), s < AT, i.e. table columns acol0 .. acol0 + AP - 1
  // ksplit > 1: workgroup ks contracts the radial points of its share
  // xb .. xe - 1 only and writes its partial sums to psi/chi + ks kstride
  // (slice stride zso); a second kernel adds the shares
  int m, y2;
  xcd_map(blockIdx.x, 2 * nyt * asplit * ksplit, m, y2);
  m += mbeg;
  const int p = y2 & 1, yt = (y2 >> 1) % nyt, sp = ((y2 >> 1) / nyt) % asplit, ks = ((y2 >> 1) / nyt) / asplit;
  if (m >= mend) return;
  const int xshare = ((((nrh + KV - 1) / KV) + ksplit - 1) / ksplit) * KV;
  const int xb = ks * xshare, xe = (xb + xshare < nrh) ? xb + xshare : nrh;
  psi += (size_t)ks * kstride;
  chi += (size_t)ks * kstride;
  const int k0 = yt * KT;
  const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
  const size_t zs = (size_t)nrdim * npdim;
  // every (m, parity, kappa tile) writes all its rows: computed values on
  // the active set, zeros elsewhere (columns m >= npc, rows past the
  // chopped range, kappa slices outside the chopped set or in zlo..zhi)
  const int nn = (m < npc && nrc0 - m > 0) ? nrc0 - m : 0;
  const int M = (nn - p + 1) / 2;
  const bool band_tile = (k0 >= zlo && k0 + KT - 1 <= zhi);
  const bool active = M > 0 && !band_tile;
  const double mv = (double)m;
  const double2 *cr = ur + (size_t)nrdim * m;
  const double2 *cp = up + (size_t)nrdim * m;
  const double2 *cz = uz + (size_t)nrdim * m;
  const int acol0 = AP * sp;
  const size_t tb = ((size_t)m * 2 + p) * nrh * tpitch + acol0;
  const double seo = (p == 0) ? 1.0 : -1.0;
  typename MM::acc aps[AT][4], ach[AT][4];
#pragma unroll
  for (int s = 0; s < AT; ++s)
#pragma unroll
    for (int t = 0; t < 4; ++t) aps[s][t] = ach[s][t] = MM::zero();

  // Software pipeline as in the forward kernel: raw loads of the next
  // chunk (three tables; ur, up, uz at x and nr-1-x) are held in registers
  // during the current chunk's MFMA work and combined when stored.  The
  // loop starts one chunk early so the fetch code appears once.
  static_assert(KV * KT == 256, "one field point per thread and chunk");
  constexpr int NTV = (3 * KV * (AP / VE) + 255) / 256;
  vec rt[NTV];
  double2 rra, rrb, rpa, rpb, rza, rzb;
  double rkv = 0.0;
  const int fx = tid % KV, fkl = tid / KV;
  if (active) {
    for (int x0 = xb - KV; x0 < xe; x0 += KV) {
      if (x0 >= xb) {
  #pragma unroll
        for (int j = 0; j < NTV; ++j) {
          const int idx = tid + 256 * j;
          if (idx < 3 * KV * (AP / VE)) {
            const int a = VE * (idx % (AP / VE)), x = (idx / (AP / VE)) % KV, q = idx / ((AP / VE) * KV);
            *reinterpret_cast<vec *>(&Tb[q][x][a]) = rt[j];
          }
        }
        const double kv = rkv;
        const double2 reo = make_double2(rra.x + seo * rrb.x, rra.y + seo * rrb.y);  // eo(ur)
        const double2 roe = make_double2(rra.x - seo * rrb.x, rra.y - seo * rrb.y);  // oe(ur)
        const double2 peo = make_double2(rpa.x + seo * rpb.x, rpa.y + seo * rpb.y);
        const double2 poe = make_double2(rpa.x - seo * rpb.x, rpa.y - seo * rpb.y);
        const double2 zeo = make_double2(rza.x + seo * rzb.x, rza.y + seo * rzb.y);
        const double2 x1 = make_double2(mv * reo.y, -mv * reo.x);           // -i m eo(ur)
        const double2 x2 = make_double2(-poe.x, -poe.y);                    // -oe(up)
        const double2 x3 = make_double2(-kv * roe.y, kv * roe.x);           // i kv oe(ur)
        const double2 x4 = make_double2(mv * kv * peo.x, mv * kv * peo.y);  // m kv eo(up)
        const double2 x5 = make_double2(-zeo.x, -zeo.y);                    // -eo(uz)
        X[0][fx][2 * fkl] = (T)x1.x; X[0][fx][2 * fkl + 1] = (T)x1.y;
        X[1][fx][2 * fkl] = (T)x2.x; X[1][fx][2 * fkl + 1] = (T)x2.y;
        X[2][fx][2 * fkl] = (T)x3.x; X[2][fx][2 * fkl + 1] = (T)x3.y;
        X[3][fx][2 * fkl] = (T)x4.x; X[3][fx][2 * fkl + 1] = (T)x4.y;
        X[4][fx][2 * fkl] = (T)x5.x; X[4][fx][2 * fkl + 1] = (T)x5.y;
        __syncthreads();
      }
      const int xn = x0 + KV;  // chunk fetched now, stored next iteration
      if (xn < xe) {
  #pragma unroll
        for (int j = 0; j < NTV; ++j) {
          const int idx = tid + 256 * j;
          vec v = MM::vzero();
          if (idx < 3 * KV * (AP / VE)) {
            const int a = VE * (idx % (AP / VE)), x = (idx / (AP / VE)) % KV, q = idx / ((AP / VE) * KV);
            const int xg = xn + x;
            const T *src = (q == 0) ? tgV : (q == 1) ? tgD : tgT;
            if (xg < xe) v = *reinterpret_cast<const vec *>(src + tb + (size_t)xg * tpitch + a);
          }
          rt[j] = v;
        }
        const int xg = xn + fx, kg = k0 + fkl;
        const double2 z2 = make_double2(0.0, 0.0);
        rra = rrb = rpa = rpb = rza = rzb = z2;
        rkv = 0.0;
        if (xg < xe && kg < nz) {
          const size_t oa = (size_t)kg * zs + xg, ob = (size_t)kg * zs + (nr - 1 - xg);
          rra = cr[oa];
          rrb = cr[ob];
          rpa = cp[oa];
          rpb = cp[ob];
          rza = cz[oa];
          rzb = cz[ob];
          rkv = ak[kg];
        }
      }
      if (x0 < xb) continue;
  #pragma unroll
      for (int kk = 0; kk < KV; kk += 4) {
        const int xr = kk + lane / 16;
        T f[5][4];
  #pragma unroll
        for (int q = 0; q < 5; ++q)
  #pragma unroll
          for (int t = 0; t < 4; ++t) f[q][t] = X[q][xr][16 * t + MM::arow(lane % 16)];
  #pragma unroll
        for (int s = 0; s < AT; ++s) {
          const int at = wv + 4 * (sp * AT + s);
          if (at * 16 >= M) continue;
          const int al = (wv + 4 * s) * 16 + lane % 16;  // local table column
          T bv = Tb[0][xr][al];
          T bd = Tb[1][xr][al];
          T bt = Tb[2][xr][al];
  #pragma unroll
          for (int t = 0; t < 4; ++t) {
            aps[s][t] = MM::op(f[0][t], bv, aps[s][t]);
            aps[s][t] = MM::op(f[1][t], bd, aps[s][t]);
            ach[s][t] = MM::op(f[2][t], bd, ach[s][t]);
            ach[s][t] = MM::op(f[3][t], bv, ach[s][t]);
            ach[s][t] = MM::op(f[4][t], bt, ach[s][t]);
          }
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int s = 0; s < AT; ++s) {
    const int a = (wv + 4 * (sp * AT + s)) * 16 + lane % 16;
    const int n = 2 * a + p;
    if (n >= nrdim) continue;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int kg = k0 + 8 * t + 2 * (lane / 16) + h;
        if (kg >= nz) continue;
        const bool keep = active && a < M && ((kg + 1 <= nzc) || (kg + 1 >= nzcu)) && !(kg >= zlo && kg <= zhi);
        size_t o = (size_t)kg * zso + (size_t)nrdim * m + n;
        psi[o] = keep ? make_double2(aps[s][t][2 * h], aps[s][t][2 * h + 1]) : make_double2(0.0, 0.0);
        chi[o] = keep ? make_double2(ach[s][t][2 * h], ach[s][t][2 * h + 1]) : make_double2(0.0, 0.0);
      }
    }
  }
  // rows of this parity past the tiles: zero
  if (sp != 0) return;  // the first of the row-sharing workgroups zeroes the rest
  const int A_END = 4 * AT * asplit * 16;
  const int nrest = (nrdim - p - 2 * A_END + 1) / 2;
  for (int idx = tid; idx < (nrest > 0 ? nrest : 0) * KT; idx += 256) {
    const int n = 2 * (A_END + idx % nrest) + p, kg = k0 + idx / nrest;
    if (kg < nz && n < nrdim) {
      const size_t o = (size_t)kg * zso + (size_t)nrdim * m + n;
      psi[o] = make_double2(0.0, 0.0);
      chi[o] = make_double2(0.0, 0.0);
    }
  }
}

// the shares of a split column-0 projection: psi(n, 0, kappa) = sum over the
// ks shares in order (slice stride zs in psi and chi, nrdim in the shares)
__global__ void k_v2t_col0_sum(double2 *__restrict__ psi, double2 *__restrict__ chi, const double2 *__restrict__ pp,
                               const double2 *__restrict__ pc, int ksplit, size_t kstride, int nrdim, int nz,
                               size_t zs) {
  const size_t n = (size_t)nrdim * nz;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    double2 a = pp[t], b = pc[t];
    for (int s = 1; s < ksplit; ++s) {
      const double2 u = pp[(size_t)s * kstride + t], v = pc[(size_t)s * kstride + t];
      a.x += u.x;
      a.y += u.y;
      b.x += v.x;
      b.y += v.y;
    }
    const size_t o = (t / nrdim) * zs + t % nrdim;
    psi[o] = a;
    chi[o] = b;
  }
}

// table re-layouts for the forward and projection kernels: dst[m][p][x][a]
template <typename T>
__global__ void k_tbl_pxa(const double *src, T *dst, int nrh, int rows_per_m, int nrows, int AP, int npchop) {
  size_t n = (size_t)npchop * 2 * nrh * AP;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    int a = (int)(t % AP);
    int x = (int)((t / AP) % nrh);
    int p = (int)((t / ((size_t)AP * nrh)) % 2);
    int m = (int)(t / ((size_t)AP * nrh * 2));
    int r = 2 * a + p;
    dst[t] = (r < nrows) ? (T)src[((size_t)m * rows_per_m + r) * nrh + x] : (T)0;
  }
}

}  // namespace

namespace mlx {

int mfma_tiles(const KitDev &k) {
  // 16-row tiles per wave needed to cover ceil(nrchop/2) rows with 4 waves
  int M = (k.nrchop + 1) / 2;
  int at = (M + 63) / 64;
  return at;
}

bool mfma_ok(const KitDev &k, int nroff) {
  int at = mfma_tiles(k);
  int xt = (k.nrh + 63) / 64;
  return nroff == 0 && at >= 1 && at <= 2 && xt >= 1 && xt <= 2;
}

void mfma_tables(mlx_ctx *c) {
  KitDev &k = c->kit;
  if (!mfma_ok(k, 0)) return;
  int AP = 4 * mfma_tiles(k) * 16;
  size_t n = (size_t)k.npchop * 2 * k.nrh * AP;
  k.mfma_ap = AP;
  auto alloc = [&](auto *&p) {
    if (p) (void)hipFree(p);
    if (hipMalloc(&p, sizeof(*p) * n) != hipSuccess) MLX_THROW(MLX_E_ALLOC, "kit: table allocation failed");
  };
  alloc(k.pfg32);
  alloc(k.tVg32);
  alloc(k.tDg32);
  alloc(k.tTg32);
  alloc(k.pfg64);
  alloc(k.tVg64);
  alloc(k.tDg64);
  alloc(k.tTg64);
  int nb = (int)((n + 255) / 256 > 8192 ? 8192 : (n + 255) / 256);
  auto relayout = [&](auto *dst, const double *src, int rows) {
    hipLaunchKernelGGL(k_tbl_pxa, dim3(nb), dim3(256), 0, c->stream, src, dst, k.nrh, rows, rows, AP, k.npchop);
  };
  relayout(k.pfg32, k.pf, k.ne);
  relayout(k.tVg32, k.tV, k.nt);
  relayout(k.tDg32, k.tD, k.nt);
  relayout(k.tTg32, k.tT, k.nt);
  relayout(k.pfg64, k.pf, k.ne);
  relayout(k.tVg64, k.tV, k.nt);
  relayout(k.tDg64, k.tD, k.nt);
  relayout(k.tTg64, k.tT, k.nt);
  HIPCHK(hipGetLastError());
}

namespace {
// the operand type of the MFMA path at the context's current precision:
// 0 FP32, 1 FP64, 2 split-BF16
int mfma_kind(const mlx_ctx *c) {
  if (c->prec == MLX_PREC_FP64_MFMA) return 1;
  if (c->prec == MLX_PREC_FP32) return 0;
  if (c->prec == MLX_PREC_BF16X3) return 2;
  MLX_THROW(MLX_E_ARG, "MFMA contractions: no kernels for this precision");
}
bool mfma_fp64(const mlx_ctx *c) { return mfma_kind(c) == 1; }
}  // namespace

// ncols > 0: columns m >= ncols are written as zero without being read
// (ztail = false: they get no workgroups and are left as they are)
void leg_forward_mfma(mlx_ctx *c, const double2 *in, double2 *out, int ncols, bool ztail) {
  const KitDev &k = c->kit;
  const int nyt = (k.nzdim + KT - 1) / KT;
  const int npc = (ncols > 0 && ncols < k.npchop) ? ncols : k.npchop;
  dim3 grid = xcd_grid(ztail ? k.npdim : npc, nyt);
  const int at = mfma_tiles(k);
  const bool f64 = mfma_fp64(c);
  ProbeScope probe(c, MLX_PROBE_LEG_FWD, leg_bytes(k, 0, f64 ? 8 : 4, npc), leg_flops(k, 0, npc));
#define MLX_FWD(T_, AT_, TB_)                                                                                  \
  probed_launch(probe, k_leg_fwd_mf<T_, AT_>, grid, dim3(256), 0, c->stream, in, out, TB_, k.w, k.nrdim, k.npdim, \
                k.nzdim, k.nr, k.nrh, k.nrchop, npc, nyt)
  if (f64) {
    if (at == 1) MLX_FWD(double, 1, k.pfg64);
    else MLX_FWD(double, 2, k.pfg64);
  } else if (mfma_kind(c) == 2) {
    if (at == 1) MLX_FWD(Bf3, 1, k.pfg32);
    else MLX_FWD(Bf3, 2, k.pfg32);
  } else {
    if (at == 1) MLX_FWD(float, 1, k.pfg32);
    else MLX_FWD(float, 2, k.pfg32);
  }
#undef MLX_FWD
  HIPCHK(hipGetLastError());
}

void leg_backward_mfma(mlx_ctx *c, const double2 *in, double2 *out, int ncols, bool ztail) {
  const KitDev &k = c->kit;
  const int nyt = (k.nzdim + KT - 1) / KT;
  const int npc = (ncols > 0 && ncols < k.npchop) ? ncols : k.npchop;
  // without the zero tail only the computed columns get workgroups
  dim3 grid = xcd_grid(ztail ? k.npdim : npc, nyt);
  const int xt = (k.nrh + 63) / 64;
  const bool f64 = mfma_fp64(c);
  ProbeScope probe(c, MLX_PROBE_LEG_BWD, leg_bytes(k, 0, f64 ? 8 : 4, npc), leg_flops(k, 0, npc));
#define MLX_BWD(T_, XT_, TB_)                                                                                    \
  probed_launch(probe, k_leg_bwd_mf<T_, XT_, false>, grid, dim3(256), 0, c->stream, in, out, nullptr, TB_, k.nrdim, \
                k.npdim, k.nzdim, k.nr, k.nrh, k.ne, k.nrchop, npc, nyt, 0)
  if (f64) {
    if (xt == 1) MLX_BWD(double, 1, k.pf);
    else MLX_BWD(double, 2, k.pf);
  } else if (mfma_kind(c) == 2) {
    if (xt == 1) MLX_BWD(Bf3, 1, k.pf32);
    else MLX_BWD(Bf3, 2, k.pf32);
  } else {
    if (xt == 1) MLX_BWD(float, 1, k.pf32);
    else MLX_BWD(float, 2, k.pf32);
  }
#undef MLX_BWD
  HIPCHK(hipGetLastError());
}

void leg_backward_pairs_f32(mlx_ctx *c, const double2 *in, float4 *pairs, int ncols) {
  const KitDev &k = c->kit;
  if ((k.nr & 1) || pair_stride(k) > k.nrdim) MLX_THROW(MLX_E_ARG, "rtrans_backward: radial pairs need an even nr");
  const int nyt = (k.nzdim + KT - 1) / KT;
  const int npc = (ncols > 0 && ncols < k.npchop) ? ncols : k.npchop;
  dim3 grid = xcd_grid(npc, nyt);
  int xt = (k.nrh + 63) / 64;
  // input: the active rows; output: one float4 per radial pair
  const double bytes = leg_bytes(k, 0, 4, npc) - 8.0 * k.nr * npc * k.nzdim;
  ProbeScope probe(c, MLX_PROBE_LEG_BWD, bytes, leg_flops(k, 0, npc));
  const int kind = mfma_kind(c);
  if (kind == 1) MLX_THROW(MLX_E_ARG, "rtrans_backward: radial pairs are FP32 sums");
#define MLX_PAIRS(T_, XT_)                                                                                       \
  probed_launch(probe, k_leg_bwd_mf<T_, XT_, true>, grid, dim3(256), 0, c->stream, in, nullptr, pairs, k.pf32, k.nrdim, \
                k.npdim, k.nzdim, k.nr, k.nrh, k.ne, k.nrchop, npc, nyt, pair_stride(k))
  if (kind == 2) {
    if (xt == 1) MLX_PAIRS(Bf3, 1);
    else MLX_PAIRS(Bf3, 2);
  } else {
    if (xt == 1) MLX_PAIRS(float, 1);
    else MLX_PAIRS(float, 2);
  }
#undef MLX_PAIRS
  HIPCHK(hipGetLastError());
}

// Every entry of psi and chi is written (zeros outside the computed set).
// ncols > 0: columns m >= ncols are zero; kappa slices zlo..zhi are zero
// (zlo > zhi: none), for a caller that de-aliases the result.
// prec_m0: precision of column m = 0 (an MFMA precision; -1: c->prec).  When
// it differs from c->prec, column 0 runs as a launch of its own on the
// context's second stream, concurrently with the other columns' launch.
void vec2tp_project_mfma(mlx_ctx *c, const double2 *ur, const double2 *up, const double2 *uz, double2 *psi,
                         double2 *chi, int ncols, int zlo, int zhi, int prec_m0) {
  const KitDev &k = c->kit;
  const int nyt = (k.nzdim + KT - 1) / KT;
  const int npc = (ncols > 0 && ncols < k.npchop) ? ncols : k.npchop;
  int nzk = k.nzdim;
  for (int kk = 0; kk < k.nzdim; ++kk)
    if (kk >= zlo && kk <= zhi) --nzk;
  const int at = mfma_tiles(k);
  const bool f64 = mfma_fp64(c);
  if (prec_m0 == MLX_PREC_FP64) prec_m0 = MLX_PREC_FP64_MFMA;  // column 0 stays on MFMA
  // the projection has no split-BF16 kernel: BF16X3 runs it in FP32
  if (prec_m0 == MLX_PREC_BF16X3) prec_m0 = MLX_PREC_FP32;
  const int prec_rest = c->prec == MLX_PREC_BF16X3 ? MLX_PREC_FP32 : c->prec;
  const bool split = prec_m0 >= 0 && prec_m0 != prec_rest;
  bool f64_0 = f64;
  if (split) {
    if (prec_m0 != MLX_PREC_FP32 && prec_m0 != MLX_PREC_FP64_MFMA)
      MLX_THROW(MLX_E_ARG, "vec2tp: no projection kernels for the m = 0 column's precision");
    f64_0 = prec_m0 == MLX_PREC_FP64_MFMA;
  }
  // algorithmic bytes: column 0's three table slabs at its own width
  const double tb = f64 ? 8.0 : 4.0, tb0 = f64_0 ? 8.0 : 4.0;
  const double bytes = v2t_bytes(k, 0, (int)tb, npc, nzk, true) + 3.0 * k.nrh * nrcs_of(k, 0, 0) * (tb0 - tb);
  ProbeScope probe(c, MLX_PROBE_V2T, bytes, 5.0 * leg_flops(k, 0, npc) * nzk / (double)k.nzdim);
  // rs: row splits; the launch of column 0 alone spreads its rows over
  // `at` workgroups of one tile per wave (more workgroups for its long
  // column, run beside the other columns' launch)
  // ks > 1: the x contraction in ks shares written to (po, pc), slice
  // stride nrdim, share stride kst
  auto launch = [&](bool d, int mb, int me, hipStream_t s, int rs, int ks = 1, double2 *po = nullptr,
                    double2 *pc = nullptr, size_t kst = 0) {
    dim3 grid = xcd_grid(me - mb, 2 * nyt * rs * ks);
    double2 *o0 = ks > 1 ? po : psi, *o1 = ks > 1 ? pc : chi;
    const size_t zso = ks > 1 ? (size_t)k.nrdim : (size_t)k.nrdim * k.npdim;
#define MLX_V2T(T_, AT_, V_, D_, T3_)                                                                    \
  hipLaunchKernelGGL((k_v2t_mf<T_, AT_>), grid, dim3(256), 0, s, ur, up, uz, o0, o1, V_, D_, T3_, k.ak, k.nrdim, \
                     k.npdim, k.nzdim, k.nr, k.nrh, k.nrchop, npc, k.chopzl, k.chopzu, nyt, zlo, zhi, mb, me, rs, \
                     k.mfma_ap, ks, zso, kst)
    const int at_l = rs > 1 ? 1 : at;
    if (d) {
      if (at_l == 1) MLX_V2T(double, 1, k.tVg64, k.tDg64, k.tTg64);
      else MLX_V2T(double, 2, k.tVg64, k.tDg64, k.tTg64);
    } else {
      if (at_l == 1) MLX_V2T(float, 1, k.tVg32, k.tDg32, k.tTg32);
      else MLX_V2T(float, 2, k.tVg32, k.tDg32, k.tTg32);
    }
#undef MLX_V2T
    HIPCHK(hipGetLastError());
  };
  probe.mark_start(c->stream);
  if (!split) {
    launch(f64, 0, k.npdim, c->stream, 1);
  } else {
    // fork: column 0 on the second stream once the inputs are ready, the
    // rest on the context stream; join before anything reads psi or chi
    aux_stream_setup(c);
    // column 0 as 4 shares of its radial points (a quarter of the
    // contraction's chain each), added in order by a second kernel
    constexpr int KS = 4;
    const size_t kst = (size_t)k.nrdim * k.nzdim;
    lazy_alloc(c, reinterpret_cast<void **>(&c->v2t_part), sizeof(double2) * 2 * KS * kst,
               "vector_projection: column-0 shares");
    double2 *po = c->v2t_part, *pc = c->v2t_part + KS * kst;
    HIPCHK(hipEventRecord(c->ev_fork, c->stream));
    HIPCHK(hipStreamWaitEvent(c->aux, c->ev_fork, 0));
    launch(f64_0, 0, 1, c->aux, at, KS, po, pc, kst);
    hipLaunchKernelGGL(k_v2t_col0_sum, dim3((unsigned)((kst + 255) / 256 < 1024 ? (kst + 255) / 256 : 1024)),
                       dim3(256), 0, c->aux, psi, chi, po, pc, KS, kst, k.nrdim, k.nzdim,
                       (size_t)k.nrdim * k.npdim);
    HIPCHK(hipGetLastError());
    launch(f64, 1, k.npdim, c->stream, 1);
    HIPCHK(hipEventRecord(c->ev_join, c->aux));
    HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  }
  probe.mark_stop(c->stream);
}

}  // namespace mlx
