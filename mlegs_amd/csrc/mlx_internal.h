// Internal declarations of the MI355X spectral-transform library.
//
// Device data layout (HBM):
//   field  e(nrdim, npdim, nzdim) complex<double>, i fastest -- the reference's
//          global layout, so host upload/download is a plain copy.
//   kit    x, w, lnx, r (nr); ak (nz); lognorm[m][n] (ne per m);
//          pf[m][n][x] (nrh x ne per m, x fastest: the contraction dimension
//          of both Legendre directions is contiguous);
//          band rows of del2 / xxdx per (m, n); projection tables of vec2tp
//          V, D, T [m][n][x].
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/mlegs_amd.h"

struct MlxError {
  int code;
  std::string msg;
};

#define MLX_THROW(code, msg) throw MlxError{(code), (msg)}
#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess)                                                          \
      throw MlxError{MLX_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)};   \
  } while (0)

// ---------------------------------------------------------------------------
// kit
// ---------------------------------------------------------------------------
struct FftPlan {
  int n;        // transform length
  int npass;    // number of radix passes
  int radix[16];
};

struct KitDev {
  // sizes (mlegs_spectfm_init.f90:111-145)
  int nr, np, nz, nrchop, npchop, nzchop, hyperpow;
  int nrh, nrdim, npdim, nzdim, nph, ne, chopzl, chopzu;
  int nt;  // rows per m of the projection tables (nrchop + 3)
  // nrdim_ref = nr + max(hyperpow, 3), the reference's first dimension and the
  // one the C ABI's host arrays use; nrdim (the device column length) is it
  // rounded up to 8 rows, so every (m, kappa) column starts 128-byte aligned.
  // Rows nrdim_ref .. nrdim-1 are zero padding.
  int nrdim_ref;
  double ell, zlen;
  // device arrays
  double *x = nullptr, *w = nullptr, *lnx = nullptr, *r = nullptr, *ak = nullptr;
  double *lognorm = nullptr;  // [m][n]
  double *pf = nullptr;       // [m][n][x]
  float *pf32 = nullptr;      // FP32 copy of pf (same layout)
  double *at0 = nullptr, *at1 = nullptr;
  double *d2c = nullptr;      // [m][n][5]: del2 band rows (offsets -2..2), /ell^2, no ak
  double *xxc = nullptr;      // [m][n][2]: xxdx band rows (offsets -1, +1)
  double *tV = nullptr, *tD = nullptr, *tT = nullptr;  // [m][n][x], nt rows per m
  float *tV32 = nullptr, *tD32 = nullptr, *tT32 = nullptr;
  // MFMA-friendly FP32 tables [m][parity][x][a] (a padded to mfma_ap)
  float *pfg32 = nullptr, *tVg32 = nullptr, *tDg32 = nullptr, *tTg32 = nullptr;
  // the same re-layouts in FP64 for the FP64 MFMA kernels
  double *pfg64 = nullptr, *tVg64 = nullptr, *tDg64 = nullptr, *tTg64 = nullptr;
  int mfma_ap = 0;
  double2 *tw_nph = nullptr, *tw_nz = nullptr, *tw_half = nullptr;
  FftPlan plan_nph, plan_nz;
  // host mirrors
  std::vector<double> lognorm_h, at1_h, at0_h, ak_h, x_h, w_h;
  bool ready = false;
};

// ---------------------------------------------------------------------------
// fields
// ---------------------------------------------------------------------------
struct mlx_field {
  double2 *e = nullptr;   // nrdim*npdim*nzdim complex
  double *sc = nullptr;   // device scalars: sc[0] = ln
  char space[4] = {'P', 'P', 'P', 0};
  int off[3] = {0, 0, 0};
};

// cached LU factors of (del2 + alpha) per (m, |kappa|) system (layout in solve.hip)
struct LuCache {
  double alpha;
  int kind;   // 0: del2 + alpha; 1: idel2 bordered (alpha = 0); 2: idel2, (0,0) excluded;
              // plus 4: the transposed systems
  int nroff;  // radial chop offset the systems were built for
  int seg = 0;               // rows per lane segment of the solve
  double *sol = nullptr;     // factor rows (segment layout)
  uint64_t *pvw = nullptr;   // row-interchange words, one per lane segment
  int *anyp = nullptr;       // per system: any interchange
  double fac_bytes = 0.0;    // factor bytes one solve reads
  int64_t nsys = 0, npiv = 0; // systems, systems with interchanges
  std::vector<uint8_t> pivh;   // per system (kk * npchop + m): interchanges
  uint64_t stamp = 0;
};

// Factors of ihelmp's banded systems (del2^(P/2) + beta del2 + alpha, band
// kl = ku = P, partial pivoting) for one (P, alpha, beta, radial offset):
// one system per (m, |kappa|), in blocks of 32 consecutive |kappa| values
// kk of one column m (block b = m * ceil(nkk / 32) + kk / 32 of nblk, place
// i = kk % 32; nsys counts the blocks' slots); entry r of row j of a
// system's LAPACK band storage (LDAB = 3P + 1 entries: U's diagonal and 2P
// superdiagonals at r = 2P - q, L's P multipliers at r = 2P + 1 ..) at
// ab[((j * nblk + b) * (3P + 2) + r') * 32 + i], r' = r (r <= 2P) or r + 1:
// the 32 systems of a block are adjacent for every entry; the row
// interchange of row j as the offset to the pivot row at
// pv[(j * nblk + b) * 32 + i].
struct HypLu {
  int power = 0;
  double alpha = 0.0, beta = 0.0;
  int nroff = 0;
  int nkk = 0, nsys = 0;
  double *ab = nullptr;
  int8_t *pv = nullptr;
  double2 *sink = nullptr;  // the solve's work rows for lanes and rows outside any system
  uint64_t stamp = 0;
};

// the state of a vortex run: psi chi nlpsi nlchi psi_prev chi_prev nlpsi_prev
// nlchi_prev uz, the SVV gains and the non-finite flag on the device
struct mlx_vortex {
  mlx_field *f[9] = {nullptr};
  double *gains = nullptr;      // device [2]
  int *flag = nullptr;          // device non-finite flag
  bool prev_stale = false;      // psi_prev/chi_prev equal psi/chi but were not copied
};

// Kernel probes: while on, HIP events are recorded on the context stream
// around each launch of the probed kernels, with the launch's algorithmic
// bytes and flops (the data it must move once / the arithmetic it must do).
struct ProbeSet {
  bool on = false;
  int cap = 0;
  std::vector<hipEvent_t> ev[MLX_PROBE_COUNT];
  int used[MLX_PROBE_COUNT] = {};
  double bytes[MLX_PROBE_COUNT] = {};
  double flops[MLX_PROBE_COUNT] = {};
};

// device list of GEMM work items (m, parity, first row, row count)
struct WorkTable {
  int4 *d = nullptr;
  int n = 0;
};

constexpr int kPartMax = 2 * 2 * 4096;  // two fields, (total, tail), up to 4096 workgroups each
// LAPACK-band LU factors of the transposed (del2 + alpha) systems, one set
// per alpha, for the adjoint of the Helmholtz solve

struct mlx_ctx {
  std::map<long, WorkTable> work;
  mlx_params p;
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  KitDev kit;
  int prec = MLX_PREC_FP64;  // precision of the contraction being issued
  int prec_ctx = MLX_PREC_FP64;  // the context's precision (mlx_set_precision)
  // per contraction site (MLX_STAGE_*): -1 follows prec_ctx
  int stage_prec[MLX_STAGE_COUNT] = {-1, -1, -1, -1};
  std::string err;
  size_t nelem = 0;
  // scratch buffers of one field each, and small device scratch
  double2 *scr[2] = {nullptr, nullptr};
  double *dsmall = nullptr;     // 4096 doubles of reduction scratch
  double *dpart = nullptr;      // kPartMax doubles: partial sums of the SVV reduction
  double *hsmall = nullptr;     // pinned host mirror (4096 doubles)
  std::vector<mlx_field *> pool;  // temporaries (allocated on demand)
  std::vector<LuCache> lu;
  std::vector<HypLu> hyp;  // ihelmp factors, at most kHypEntries (least recently used evicted)
  uint64_t lu_clock = 0;
  ProbeSet probe;
  // second stream (default priority) for work that overlaps the context
  // stream inside one operation, with its fork and join events; created on
  // first use
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // partial sums of the split column-0 projection (allocated on first use)
  double2 *v2t_part = nullptr;
};

// A buffer the context allocates on first use (*p of bytes).  Under stream
// capture nothing may be allocated: a step run once before the capture
// allocates it.
inline void lazy_alloc(mlx_ctx *c, void **p, size_t bytes, const char *what) {
  if (*p) return;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  HIPCHK(hipStreamIsCapturing(c->stream, &st));
  if (st != hipStreamCaptureStatusNone)
    MLX_THROW(MLX_E_ARG, std::string(what) + ": first use under stream capture (run one step before capturing)");
  if (hipMalloc(p, bytes) != hipSuccess) {
    *p = nullptr;
    MLX_THROW(MLX_E_ALLOC, std::string(what) + ": allocation failed");
  }
}

inline void aux_stream_setup(mlx_ctx *c) {
  if (c->aux) return;
  HIPCHK(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
}

// The precision of a contraction site: its own setting, else the context's.
static inline int stage_precision(const mlx_ctx *c, int stage) {
  return c->stage_prec[stage] >= 0 ? c->stage_prec[stage] : c->prec_ctx;
}
// Issues the contractions of one site at its precision (c->prec) and
// restores the previous precision afterwards.
struct PrecScope {
  mlx_ctx *c;
  int saved;
  PrecScope(mlx_ctx *c_, int stage) : c(c_), saved(c_->prec) { c->prec = stage_precision(c, stage); }
  ~PrecScope() { c->prec = saved; }
};

// Records the opening event on construction and the closing one on
// destruction when probing is on and the probe has room left.
// The launches of a probed region go through probed_launch, which has the
// dispatch itself stamp the events (hipExtLaunchKernel): the first launch
// sets the start, every launch the stop, so the interval is the first
// kernel's start to the last one's end, with no event packets in between.
struct ProbeScope {
  mlx_ctx *c;
  int id;
  int slot = -1;
  bool started = false;
  ProbeScope(mlx_ctx *c_, int id_, double bytes, double flops) : c(c_), id(id_) {
    ProbeSet &p = c->probe;
    if (!p.on || p.used[id] >= p.cap) return;
    slot = p.used[id]++;
    p.bytes[id] += bytes;
    p.flops[id] += flops;
  }
  hipEvent_t start_event() {
    if (slot < 0 || started) return nullptr;
    started = true;
    return c->probe.ev[id][2 * slot];
  }
  hipEvent_t stop_event() const { return slot >= 0 ? c->probe.ev[id][2 * slot + 1] : nullptr; }
  // explicit interval on stream s around launches that are not probed_launch
  // (several streams inside): the start before the first, the stop after the
  // join
  void mark_start(hipStream_t s) {
    if (slot < 0 || started) return;
    started = true;
    (void)hipEventRecord(c->probe.ev[id][2 * slot], s);
  }
  void mark_stop(hipStream_t s) const {
    if (slot >= 0) (void)hipEventRecord(c->probe.ev[id][2 * slot + 1], s);
  }
  ~ProbeScope() {
    if (slot >= 0 && !started) {  // nothing launched: an empty interval
      (void)hipEventRecord(c->probe.ev[id][2 * slot], c->stream);
      (void)hipEventRecord(c->probe.ev[id][2 * slot + 1], c->stream);
    }
  }
};

// (an ordinary launch when the probe is off, so that a step can be
// captured into a HIP graph)
template <typename... Args, typename F>
inline void probed_launch(ProbeScope &ps, F kernel, dim3 grid, dim3 block, unsigned shm, hipStream_t s,
                          Args... args) {
  if (ps.slot < 0)
    hipLaunchKernelGGL(kernel, grid, block, shm, s, args...);
  else
    hipExtLaunchKernelGGL(kernel, grid, block, shm, s, ps.start_event(), ps.stop_event(), 0u, args...);
}

// active (n, m) pairs for a radial chop offset: sum over m < npchop of nrcs(m)
static inline double active_pairs(const KitDev &k, int nroff, int ncols = -1) {
  double s = 0;
  const int nc = (ncols > 0 && ncols < k.npchop) ? ncols : k.npchop;
  for (int m = 0; m < nc; ++m) {
    int v = k.nrchop + nroff - m;
    s += v > 0 ? v : 0;
  }
  return s;
}

// ---------------------------------------------------------------------------
// launch helpers (defined in the .hip files)
// ---------------------------------------------------------------------------
namespace mlx {

// kit.hip
void kit_build(mlx_ctx *c);
void kit_free(KitDev &k);

// fft.hip
// dir -1 forward (PPP->PFP), +1 backward; out may equal in; rscale: +1 multiply
// input rows i < nr by r_i (forward), -1 divide output rows by r_i (backward)
// mlim > 0 (backward, fast path only): modes m >= mlim are taken as zero
// without being read
// phys_rows (backward, fast path): only rows i < nr are synthesized and
// written (rows nr.. of the output are left as they are)
void fft_phi(mlx_ctx *c, double2 *out, const double2 *in, int dir, int rscale = 0, int mlim = -1,
             bool phys_rows = false);
// backward phi transform of leg_backward_pairs_f32's radial pairs into rows
// < nr of out (fast path only)
// f32out: the rows are stored in FP32 (float pair e of out's buffer for
// entry e), for fft_phi_cross's FP32 input
void fft_phi_pairs(mlx_ctx *c, double2 *out, const float4 *pairs, int rscale, int mlim, bool f32out = false);
bool fft_phi_fast(const KitDev &k);  // the four-step phi kernels cover np/2
bool fft_z_fast(const KitDev &k);    // the four-step z kernels cover nz
// v x w (v_z + ub) to PFP with the r weights of vector_projection (modes
// m < mmax written; mmax <= 0: all); false when
// the fused form is not built for this np (the caller then runs vecprod and
// three phi transforms)
// f32in: v0..w2 are stored in FP32 (as fft_phi_pairs's f32out writes them)
bool fft_phi_cross(mlx_ctx *c, double2 *o0, double2 *o1, double2 *o2, const double2 *v0, const double2 *v1,
                   const double2 *v2, const double2 *w0, const double2 *w1, const double2 *w2, const double2 *ub,
                   double *ln0 = nullptr, double *ln1 = nullptr, double *ln2 = nullptr, int mmax = -1,
                   bool f32in = false);
bool fft_phi_cross_ok(const KitDev &k);  // the fused kernel is built for this np
// dir -1 forward (/nz), +1 backward.  ncol: the (row, column) lines the caller
// needs transformed (the launch's algorithmic work).  act_nrc > 0: only rows
// i < act_nrc - j of column j need transforming; ncols > 0: only columns
// j < ncols.  Lines outside are left as they are (the caller guarantees they
// are zero or never read); the generic kernel transforms every line.
// zin_lo..zin_hi: line entries known to be zero (not read); zout_lo..zout_hi:
// outputs the caller does not use (not written) -- four-step kernels only,
// the generic kernel reads and writes everything
void fft_z(mlx_ctx *c, double2 *e, int dir, double ncol, int act_nrc = 0, int ncols = 0, int zin_lo = 1,
           int zin_hi = 0, int zout_lo = 1, int zout_hi = 0);
// the same transform of the nf (<= kZBatch) fields e[0..nf) in one launch
// (ncol: per field)
constexpr int kZBatch = 6;
void fft_z_batch(mlx_ctx *c, double2 *const *e, int nf, int dir, double ncol, int act_nrc = 0, int ncols = 0,
                 int zin_lo = 1, int zin_hi = 0, int zout_lo = 1, int zout_hi = 0);

// legendre.hip
// ncols > 0 (MFMA path): only columns m < ncols are transformed, the others
// are written as zero (their input is zero or their output unused)
// ztail = false (MFMA path): columns m >= ncols are left unwritten
void leg_forward(mlx_ctx *c, const double2 *in, double2 *out, int nroff, bool zero = true, int ncols = -1,
                 bool ztail = true);
void leg_backward(mlx_ctx *c, const double2 *in, double2 *out, int nroff, bool zero = true, int ncols = -1,
                  bool ztail = true);
// legendre_mfma.hip (MFMA in FP32 or FP64 by c->prec; valid when mfma_ok)
bool mfma_ok(const KitDev &k, int nroff);
void mfma_tables(mlx_ctx *c);
void leg_forward_mfma(mlx_ctx *c, const double2 *in, double2 *out, int ncols = -1, bool ztail = true);
// ztail = false: columns m >= ncols are left unwritten
void leg_backward_mfma(mlx_ctx *c, const double2 *in, double2 *out, int ncols = -1, bool ztail = true);
// the FP32 even/odd sums (be, bo) of columns m < ncols as float4 radial pairs
// (layout at k_leg_bwd_f32), for fft_phi_pairs; a line of pairs (one z slice
// and column) takes pair_stride float4 (256-byte aligned lines)
inline int pair_stride(const KitDev &k) { return (k.nrh + 15) / 16 * 16; }
void leg_backward_pairs_f32(mlx_ctx *c, const double2 *in, float4 *pairs, int ncols);
// prec_m0: precision of column m = 0 (-1: the context's current one)
void vec2tp_project_mfma(mlx_ctx *c, const double2 *ur, const double2 *up, const double2 *uz, double2 *psi,
                         double2 *chi, int ncols = -1, int zlo = 1, int zhi = 0, int prec_m0 = -1);
void vec2tp_project(mlx_ctx *c, const double2 *ur, const double2 *up, const double2 *uz,
                    double2 *psi, double2 *chi, int nroff, int ncols = -1,
                    int zlo = 1, int zhi = 0, int prec_m0 = -1);

// spectral.hip
void k_ln_column(mlx_ctx *c, double2 *e, const double *ln, double sign);
void k_abcn_rhs_g(mlx_ctx *c, double2 *sh, double *shln, const double2 *s, const double *sln, const double2 *nl,
                  const double *nlln, const double2 *nlp, const double *nlpln, const double2 *svis, double dt,
                  double a, int ab2, double g);
// ihelmp's banded solves (power 4, 6 or 8) and its log-term step, FP64, on
// factors kept per (power, alpha, beta, nroff); a factorization that meets a
// zero pivot sets *sing (device int) to 1 and throws before anything is solved
// pcut > 0: only columns m < pcut are solved; slices zlo..zhi are not solved
// (their right-hand sides are zero or are zeroed afterwards)
void k_ihelmp(mlx_ctx *c, double2 *e, double *sln, int power, double alpha, double beta, int nroff, int *sing,
              int pcut = -1, int zlo = 1, int zhi = 0);
// fftreat's physical-space step on the PFF lines of columns m < npc
void k_fftreat_mid(mlx_ctx *c, double2 *e, int npc);
void k_chop(mlx_ctx *c, double2 *e, const char *space, int nroff);
void k_dealias(mlx_ctx *c, double2 *e, const char *space);
// svv_filter on e0 (gain g0) and, when e1 is given, on e1 (gain g1) in the
// same two launches
// part_ws, out_ws (optional): the partial sums (2 per workgroup and field,
// at most kSvvParts workgroups per field) and the per-field results (3 per
// field) in memory the caller provides instead of the context's
constexpr int kSvvParts = 4096;
void k_svv(mlx_ctx *c, double2 *e0, double *g0, double2 *e1 = nullptr, double *g1 = nullptr, int *flag = nullptr,
           int pcut = -1, int zlo = 1, int zhi = 0, double *part_ws = nullptr, double *out_ws = nullptr);
// noalloc.hip: looks up this device's static scratch array (once per device;
// the device must be current)
void noalloc_setup(int device);
void k_calcat(mlx_ctx *c, const double2 *e, const double *tbl, int nrc, double2 *out);
// zeroat1 on e (rows < nrc) and, when e2 is given, on e2 in the same launch;
// lnz / lnz2 (optional): log coefficients set to zero by that launch
void k_zeroat1(mlx_ctx *c, double2 *e, int nrc, double2 *e2 = nullptr, int nrc2 = 0, double *lnz = nullptr,
               double *lnz2 = nullptr);
void k_delsqp(mlx_ctx *c, double2 *e, double *ln, int nroff, int inverse);
void k_bandmul(mlx_ctx *c, double2 *e, double *ln, int op, int nroff);  // op 0 del2, 1 del2h, 2 xxdx
void k_helm_pre(mlx_ctx *c, double2 *e, double *ln_out, const double *ln_in, double alpha);
// pcut > 0: columns m >= pcut and kappa slices zlo..zhi are left unsolved
// (for a caller that de-aliases the result)
// e2 (optional): a second field solved with the same factors in the same launch
void k_helm_solve(mlx_ctx *c, double2 *e, double alpha, int kind, int nroff, int pcut = -1, int zlo = 1,
                  int zhi = 0, double2 *e2 = nullptr);
void lu_cache_free(mlx_ctx *c);  // solve.hip
void hyp_cache_free(mlx_ctx *c);  // spectral.hip
void k_idel2_post(mlx_ctx *c, const double2 *e, double *ln);
void k_idel2_preln00(mlx_ctx *c, double2 *e, double *ln, double preln, int nroff);
void k_axpby(mlx_ctx *c, double2 *dst, double *dln, double a, const double2 *src, const double *sln, double b);
void k_read_field(mlx_ctx *c, const double2 *src, double *part);
void k_write_field(mlx_ctx *c, double2 *dst, double v);
void k_copy_field(mlx_ctx *c, double2 *dst, const double2 *src);
void k_abcn_rhs(mlx_ctx *c, double2 *sh, double *shln, const double2 *s, const double *sln,
                const double2 *nl, const double *nlln, const double2 *nlp, const double *nlpln,
                const double2 *svis, double dt, double a, int ab2);
// abcn's right-hand side in place on e (= s), with svis = del2(s) formed per column
void k_abcn_rhs_d2(mlx_ctx *c, double2 *e, double *ln, const double2 *nl, const double *nlln, const double2 *nlp,
                   const double *nlpln, double dt, double a, int nroff, int pcut = -1,
                   int zlo = 1, int zhi = 0,
                   bool helm_pre = false);
// abcn's hyperviscous right-hand side in place on e (= s), with svis =
// helmp(s, power, 0, b) formed per column; act: s, N, N_p chopped (rows past
// the system's rows zero), only the active rows read and written
void k_abcn_rhs_hyp(mlx_ctx *c, double2 *e, double *ln, const double2 *nl, const double *nlln, const double2 *nlp,
                    const double *nlpln, double dt, double a, double g, double b, int power, int nroff, int pcut,
                    int zlo, int zhi, bool act);
void k_scale_rows_r(mlx_ctx *c, double2 *e, int inverse);
void k_tp2vec_combine(mlx_ctx *c, double2 *ur, double2 *up, const double2 *psi, const double2 *uz, int nroff);
// mode 0: (P, C) = (psi, chi); mode 1: P = -del2(p_in = chi) over the rows of
// chi's radial offset md_off, C = c_in = psi
void k_tp2vec_front(mlx_ctx *c, double2 *ur, double2 *up, double2 *uz, const double2 *p_in, const double2 *c_in,
                    const double *p_ln, const double *c_ln, int mode, int md_off, double *ln_r = nullptr,
                    double *ln_p = nullptr, double *ln_z = nullptr, int ncols = -1, int zlo = 1,
                    int zhi = 0);
// mode 0 into (vr, vp, vz) and mode 1 (chi's offset md_off) into (wr, wp, wz)
// in one launch
void k_tp2vec_front2(mlx_ctx *c, double2 *vr, double2 *vp, double2 *vz, double *lvr, double *lvp, double *lvz,
                     double2 *wr, double2 *wp, double2 *wz, double *lwr, double *lwp, double *lwz,
                     const double2 *psi, const double2 *chi, const double *psi_ln, const double *chi_ln,
                     int md_off, int ncols, int zlo, int zhi);
void k_vecprod(mlx_ctx *c, double2 *vr, double2 *vp, double2 *vz, const double2 *ur, const double2 *up,
               const double2 *uz);
void k_vec2tp_up_ln(mlx_ctx *c, double2 *up, const double2 *inf, double *psiln);
void k_nonfinite(mlx_ctx *c, const double2 *a, const double2 *b, int *flag);
void k_energy(mlx_ctx *c, const double2 *vr, const double2 *vp, const double2 *vz, double *out);
void k_vormag(mlx_ctx *c, double2 *out, const double2 *wr, const double2 *wp, const double2 *wz);
void k_set_scalar(mlx_ctx *c, double *dst, double v);
// adjoint kernels (adjoint.hip)
void k_adj_phi_scale(mlx_ctx *c, double2 *e, int mode);
void k_adj_wscale(mlx_ctx *c, double2 *e, int mode);
void k_ln_adjoint(mlx_ctx *c, const double2 *e, double *ln_bar, double sign);
void k_adj_scale(mlx_ctx *c, double2 *e, double s);
void k_adj_project(mlx_ctx *c, double2 *e);
void k_adj_diag(mlx_ctx *c, double2 *e, int wmode, int rmode, int phimode, double s);
void k_adj_sum_mask(mlx_ctx *c, double2 *sb, double *sbl, const double2 *g, const double *gl);
void k_del2_t(mlx_ctx *c, double2 *out, const double2 *in, bool with_ak);
void k_lap_ln_t(mlx_ctx *c, const double2 *e, double *ln_bar, double scale);
void k_helm_solve_t(mlx_ctx *c, double2 *e, double alpha, int kind = 0, int nroff = 0, int pcut = -1, int zlo = 1,
                    int zhi = 0, double2 *e2 = nullptr);
void k_abcn_adj_out(mlx_ctx *c, double2 *y, const double2 *t, double2 *nlbar, double2 *nlpbar, double c1, double c2,
                    double a, double c3);
void k_abcn_adj_ln(mlx_ctx *c, const double2 *y, double *ysc, double *nlln, double *nlpln, double dt, double c2);
// ncols > 0, zlo..zhi: columns m >= ncols and slices zlo..zhi taken as zero
void k_tp2vec_t(mlx_ctx *c, double2 *psib, double2 *chib, const double2 *urb, const double2 *upb, const double2 *uzb,
                double *psib_ln, double *chib_ln, int ncols = -1, int zlo = 1, int zhi = 0);
void k_v2t_t(mlx_ctx *c, double2 *urb, double2 *upb, double2 *uzb, const double2 *psib, const double2 *chib,
             int ncols = -1, int zlo = 1, int zhi = 0);
void k_zeroat1_t(mlx_ctx *c, double2 *e);
void k_v2t_up_ln_t(mlx_ctx *c, double2 *upF, const double2 *upP);
void k_energy_grad(mlx_ctx *c, double2 *vr, double2 *vp, double2 *vz);
void k_vecprod_adj(mlx_ctx *c, double2 *a1, double2 *a2, double2 *a3, const double2 *v1, const double2 *v2,
                   const double2 *v3, const double2 *u1, const double2 *u2, const double2 *u3, double2 *ub1,
                   double2 *ub2, double2 *ub3, bool f32 = false, const double2 *bg = nullptr);
// one int set by a kernel: clears flags in code a caller may capture into a
// HIP graph (no memset nodes there)
void k_set_flag(mlx_ctx *c, int *dst, int v);
void k_copy_scalar(mlx_ctx *c, double *dst, const double *src, double scale);

}  // namespace mlx

// Algorithmic work of one Legendre contraction of a whole field (one
// direction): the parity-folded real table times complex data, FMA = 2 flops,
// 4 (nr/2) nz S flops; bytes = physical columns m < npchop read (or written)
// once, spectral coefficients written (or read) once, table read once at
// tbytes per entry.
// ncols > 0: only columns m < ncols are transformed
static inline double leg_flops(const KitDev &k, int nroff, int ncols = -1) {
  return 4.0 * k.nrh * k.nzdim * active_pairs(k, nroff, ncols);
}
static inline double leg_bytes(const KitDev &k, int nroff, int tbytes, int ncols = -1) {
  const double S = active_pairs(k, nroff, ncols);
  const int nc = (ncols > 0 && ncols < k.npchop) ? ncols : k.npchop;
  return 16.0 * k.nr * nc * k.nzdim + 16.0 * S * k.nzdim + (double)tbytes * k.nrh * S;
}
// vec2tp projection: three physical fields in, two spectral fields out,
// tables V, D, T
// ncols > 0: columns m < ncols only; nzk: kappa slices computed (default all);
// full: both output fields written in full (zeros outside the computed set)
static inline double v2t_bytes(const KitDev &k, int nroff, int tbytes, int ncols = -1, int nzk = -1,
                               bool full = false) {
  const double S = active_pairs(k, nroff, ncols);
  const int nc = (ncols > 0 && ncols < k.npchop) ? ncols : k.npchop;
  const double zk = (nzk >= 0) ? nzk : k.nzdim;
  const double out = full ? 32.0 * k.nrdim * k.npdim * k.nzdim : 32.0 * S * zk;
  return 48.0 * k.nr * nc * zk + out + 3.0 * tbytes * k.nrh * S;
}

// XCD-aware work mapping.  Workgroups are dispatched round-robin over the 8
// XCDs (linear id L runs on XCD L % 8), each with its own L2.  The nyt
// workgroups of one column u (they all read the same table slice) get ids
// with the same L % 8, so the slice is fetched from HBM into one L2 only.
constexpr int kXcd = 8;
__device__ inline void xcd_map(int L, int nyt, int &u, int &y) {
  const int x = L % kXcd, q = L / kXcd;
  y = q % nyt;
  u = (q / nyt) * kXcd + x;
}
inline dim3 xcd_grid(int nu, int nyt) { return dim3(((nu + kXcd - 1) / kXcd) * kXcd * nyt); }

// FP32 storage of a field's physical samples (the advection term's velocity
// and vorticity): entry (row i, column j, slice k) at float pair
// (k npdim + j) es32 + i of the field's buffer, es32 = nrdim rounded up to 16
// so that every column starts on a 128-byte line (it fits: the FP64 field
// holds 2 nrdim npdim nz float pairs)
__host__ __device__ inline size_t f32_index(size_t k, size_t j, size_t i, size_t nrdim, size_t npdim) {
  const size_t es32 = (nrdim + 15) & ~(size_t)15;
  return (k * npdim + j) * es32 + i;
}

// Streaming (non-temporal) access to field data that a kernel reads or writes
// once: the lines bypass the normal cache retention, so a pass over a 140 MB
// field does not evict what the next kernels reuse (tables, twiddles).
typedef double mlx_dv2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_nt(const double2 *p) {
  const mlx_dv2 v = __builtin_nontemporal_load(reinterpret_cast<const mlx_dv2 *>(p));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st_nt(double2 *p, double2 x) {
  const mlx_dv2 v = {x.x, x.y};
  __builtin_nontemporal_store(v, reinterpret_cast<mlx_dv2 *>(p));
}

// helpers shared by host code
static inline int nrcs_of(const KitDev &k, int m, int nroff) {
  int nrc = k.nrchop + nroff;
  int v = nrc - m;
  return v > 0 ? v : 0;
}
