/*
 * mlegs_amd.h -- C ABI of the MI355X spectral-transform time step.
 *
 * The MLegS reference (UCBCFD/MLegS v1.1.3, Modern Fortran + MPI) exposes its
 * hot path as Fortran generic interfaces on type(scalar) and the global
 * transform kit `tfm`.  Each entry point below replaces one of them; the
 * citation after each declaration names the interface it stands in for.
 * A Fortran host binds these with iso_c_binding (mlegs_amd/fortran/mlegs_amd.f90) and
 * Python binds them with ctypes (mlegs_amd/_lib.py).
 *
 * Conventions
 *   - Every function returns 0 (MLX_OK) or a non-zero MLX_E_* code; the
 *     message of the last failure is mlx_last_error(ctx).  The library never
 *     exits the process.  Error classes follow the reference's error flags
 *     (src/modules/mlegs_envir.f90:39-45).
 *   - Device memory is owned by the library behind opaque handles; host
 *     buffers are borrowed for the duration of the call.
 *   - Host field buffers use exactly the reference's global layout: complex(8)
 *     e(nrdim, npdim, nzdim), column-major (i fastest), interleaved re/im
 *     doubles (src/modules/mlegs_scalar.f90:15-49, docs/tutorial/
 *     transformation.md:58-86).
 *   - Work is enqueued on the context's HIP stream in call order; only
 *     downloads, mlx_sync and functions returning host scalars block.
 *   - One context per host thread; contexts share no mutable state.
 *   - Only the radial chopping offset (nrchop_offset) is supported; the
 *     reference's path never sets the azimuthal or axial offsets.
 *   - Hyperviscosity: with hyperpow = 4, 6 or 8 the kit sizes nrdim as
 *     nr + hyperpow and febe / abcn / the vortex step take the reference's
 *     hyperviscous branches through helmp / ihelmp
 *     (src/submodules/mlegs_scalar_ops.f90:1181-1195, 1226-1255).
 */
#ifndef MLEGS_AMD_H
#define MLEGS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (mirror the reference's error flags 201-206) */
#define MLX_OK          0
#define MLX_E_LAYOUT    201  /* size / layout / space mismatch            */
#define MLX_E_ALLOC     202  /* device allocation failure                  */
#define MLX_E_ARG       203  /* invalid argument                           */
#define MLX_E_FILE      204  /* field file open / read / write failure     */
#define MLX_E_NONFINITE 205  /* non-finite state (check_stability)         */
#define MLX_E_HIP       206  /* HIP runtime failure                        */

/* precision of the Legendre contractions (trans and vec2tp) */
#define MLX_PREC_FP64   0    /* FP64 operands, FP64 accumulation (vector
                                FMA), every column computed: the plain path */
#define MLX_PREC_FP32   1    /* FP32 operands, FP32 MFMA accumulation       */
#define MLX_PREC_BF16X3 2    /* 3-term split-BF16 on MFMA: each FP32-staged
                                operand split into three BF16 terms, six
                                products accumulated in FP32
                                (v_mfma_f32_16x16x16_bf16); the vec2tp
                                projection runs in FP32                    */
#define MLX_PREC_FP64_MFMA 3 /* FP64 operands, FP64 MFMA accumulation
                                (v_mfma_f64_16x16x4_f64)                    */

/* contraction sites whose precision can be set apart from the context's
 * (mlx_set_stage_precision) */
#define MLX_STAGE_SYNTH 0    /* backward radial transforms of tp2vec and
                                tp2curlvec (the advection term's six)      */
#define MLX_STAGE_UP    1    /* vec2tp's radial transforms of r u_phi (to
                                FFF for its value at r = inf, and back)    */
#define MLX_STAGE_PROJ  2    /* vec2tp's five projection contractions,
                                columns m >= 1                              */
#define MLX_STAGE_PROJ0 3    /* the same for column m = 0, where idel2's
                                kappa = 0 Laplacian systems amplify the
                                contraction's rounding; a setting that
                                differs from MLX_STAGE_PROJ's runs the column
                                concurrently on a second stream (an MFMA
                                precision: plain FP64 there means FP64 MFMA
                                while the other columns run on MFMA)        */
#define MLX_STAGE_COUNT 4

/* kernel probes (measurement): see mlx_probe_start */
#define MLX_PROBE_FFT_PHI   0  /* phi FFT, one launch per direction       */
#define MLX_PROBE_FFT_Z     1  /* z FFT                                     */
#define MLX_PROBE_LEG_FWD   2  /* forward Legendre contraction             */
#define MLX_PROBE_LEG_BWD   3  /* backward Legendre contraction            */
#define MLX_PROBE_V2T       4  /* vec2tp projection (5 contractions)       */
#define MLX_PROBE_LU_SOLVE  5  /* banded LU solve of ihelm / idel2         */
#define MLX_PROBE_PHI_CROSS 6  /* v x w formed inside three phi transforms */
#define MLX_PROBE_TP_FRONT  7  /* tp2vec / tp2curlvec column pass          */
#define MLX_PROBE_ABCN_RHS  8  /* AB2/CN right-hand side with del2         */
#define MLX_PROBE_COUNT     9

typedef struct mlx_ctx mlx_ctx;
typedef struct mlx_field mlx_field;

/* Run parameters: the module variables of src/modules/mlegs_base.f90:7-41. */
typedef struct mlx_params {
  int32_t nr, np, nz;
  int32_t nrchop, npchop, nzchop;
  double ell, zlen;
  double dt, ti, totaltime;
  int32_t ni, totaln;
  double visc;
  int32_t hyperpow;
  double hypervisc;
  int32_t is_svv;
  double svv_cutoff, svv_target, svv_strength, svv_relax;
} mlx_params;

/* ---- context and kit ---------------------------------------------------- */

/* Create a context on HIP device `device`; stream may be NULL (the context
 * creates its own) or a hipStream_t.  Replaces the process-wide globals of
 * mlegs_base / mlegs_envir (src/modules/mlegs_base.f90:7-53). */
int mlx_ctx_create(const mlx_params* p, int device, void* stream, mlx_ctx** out);
int mlx_ctx_destroy(mlx_ctx* ctx);
const char* mlx_last_error(mlx_ctx* ctx);
int mlx_sync(mlx_ctx* ctx);

/* Build the transform kit on the GPU in FP64: Gauss-Legendre nodes and
 * weights, log-norms, the normalised P_n^m table, at0/at1, band-operator
 * coefficients and projection tables.  Replaces tfm%init
 * (src/submodules/mlegs_spectfm_init.f90:6-154). */
int mlx_kit_init(mlx_ctx* ctx);

/* Kit sizes: out[0..8] = nrdim, npdim, nzdim, nrh, ne, chopzl, chopzu, nph, npchop
 * (src/modules/mlegs_spectfm.f90:35-62). */
int mlx_kit_dims(mlx_ctx* ctx, int32_t* out9);

/* Copy one kit array to the host.  which: "x","w","ln","r","ak","lognorm",
 * "pf","at0","at1" with the reference's shapes (column-major); n = number
 * of doubles the host buffer holds.  Replaces direct reads of tfm%...
 * (src/modules/mlegs_spectfm.f90:17-28). */
int mlx_kit_get(mlx_ctx* ctx, const char* which, double* host, int64_t n);

/* Legendre-contraction precision used by mlx_trans and mlx_vec2tp:
 * MLX_PREC_FP64, MLX_PREC_FP32, MLX_PREC_FP64_MFMA or MLX_PREC_BF16X3.
 * Every precision but MLX_PREC_FP64 runs the MFMA kernels, and the vortex
 * step then skips the modes de-aliasing zeroes. */
int mlx_set_precision(mlx_ctx* ctx, int prec);
/* Precision of one contraction site (MLX_STAGE_*) apart from the context's;
 * prec = -1 makes the site follow mlx_set_precision again.  No counterpart
 * in the reference (its contractions are all FP64). */
int mlx_set_stage_precision(mlx_ctx* ctx, int stage, int prec);

/* ---- fields --------------------------------------------------------------- */

/* Allocate a scalar of the kit's global size, space "PPP", ln = 0, offsets 0.
 * Replaces scalar%init (src/submodules/mlegs_scalar_init.f90:6-67). */
int mlx_field_create(mlx_ctx* ctx, mlx_field** out);
/* scalar%dealloc (src/submodules/mlegs_scalar_init.f90:70-81). */
int mlx_field_destroy(mlx_ctx* ctx, mlx_field* f);
/* Deep copy dst = src: data, ln, offsets, space
 * (src/submodules/mlegs_scalar_init.f90:106-140). */
int mlx_field_copy(mlx_ctx* ctx, mlx_field* dst, const mlx_field* src);
/* Upload / download the whole scalar in the reference global layout
 * (replaces scalar%disassemble / %assemble, mlegs_scalar_dist.f90:70-368).
 * host holds 2*nrdim*npdim*nzdim doubles; space is 3 chars; offsets[3] are
 * (nrchop, npchop, nzchop) offsets. */
int mlx_field_upload(mlx_ctx* ctx, mlx_field* f, const double* host,
                     int64_t nrdim, int64_t npdim, int64_t nzdim,
                     const char* space, double ln, const int32_t* offsets);
int mlx_field_download(mlx_ctx* ctx, const mlx_field* f, double* host,
                       int64_t nrdim, int64_t npdim, int64_t nzdim,
                       char* space4, double* ln, int32_t* offsets);
/* Metadata access without moving the data. */
int mlx_field_set_meta(mlx_ctx* ctx, mlx_field* f, const char* space, double ln, const int32_t* offsets);
int mlx_field_get_meta(mlx_ctx* ctx, const mlx_field* f, char* space4, double* ln, int32_t* offsets);
/* scalar%chop_offset (src/submodules/mlegs_scalar_init.f90:84-103). */
int mlx_chop_offset(mlx_ctx* ctx, mlx_field* f, int32_t iof1, int32_t iof2, int32_t iof3);
/* dst = a*dst + b*src on data and ln (the `s%e = a*s%e + b*t%e` statements
 * of the reference's drivers, e.g. vortical_flow_3d.f90:136-137, 379). */
int mlx_field_axpby(mlx_ctx* ctx, mlx_field* dst, double a, const mlx_field* src, double b);

/* ---- transforms and masks ------------------------------------------------- */

/* trans(s, space, tfm) (src/modules/mlegs_scalar.f90:147-155;
 * src/submodules/mlegs_scalar_ops.f90:157-235). */
int mlx_trans(mlx_ctx* ctx, mlx_field* f, const char* space);
/* chop (mlegs_scalar.f90:116-123; mlegs_scalar_ops.f90:6-41). */
int mlx_chop(mlx_ctx* ctx, mlx_field* f);
/* dealias (mlegs_scalar.f90:126-133; mlegs_scalar_ops.f90:43-70). */
int mlx_dealias(mlx_ctx* ctx, mlx_field* f);
/* svv_filter(s, tfm, gain) (mlegs_scalar.f90:136-144; mlegs_scalar_ops.f90:72-155);
 * *gain is read and updated. */
int mlx_svv_filter(mlx_ctx* ctx, mlx_field* f, double* gain);
/* svv_filter without device allocation: the same filter and gain update as
 * mlx_svv_filter, with the reduction's partial sums, results and gain in a
 * static device array of 8200 doubles (64 KiB) that each GPU holds until
 * the process ends, instead of buffers of the context.  No counterpart in
 * the reference; it is for callers that cannot allocate device memory
 * (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Limit: nz <= 4096 (MLX_E_ARG beyond).  Calls that overlap on
 * one GPU (other threads or streams) share the array and race. */
int mlx_svv_filter_noalloc(mlx_ctx* ctx, mlx_field* f, double* gain);
/* calcat0 / calcat1: values at r=0 / r=inf of m=0 per z slice; out holds
 * 2*nz doubles (mlegs_scalar_ops.f90:237-309). */
int mlx_calcat0(mlx_ctx* ctx, const mlx_field* f, double* out);
int mlx_calcat1(mlx_ctx* ctx, const mlx_field* f, double* out);
/* zeroat1 (mlegs_scalar.f90:180-187; mlegs_scalar_ops.f90:311-325). */
int mlx_zeroat1(mlx_ctx* ctx, mlx_field* f);

/* ---- linear operators ---------------------------------------------------- */

/* delsqp / idelsqp (mlegs_scalar_ops.f90:327-416). */
int mlx_delsqp(mlx_ctx* ctx, mlx_field* f);
int mlx_idelsqp(mlx_ctx* ctx, mlx_field* f);
/* xxdx, del2h, del2 (mlegs_scalar_ops.f90:418-573). */
int mlx_xxdx(mlx_ctx* ctx, mlx_field* f);
int mlx_del2h(mlx_ctx* ctx, mlx_field* f);
int mlx_del2(mlx_ctx* ctx, mlx_field* f);
/* idel2 without / with a prescribed log coefficient
 * (mlegs_scalar.f90:250-263; idel2_proln :671-760, idel2_preln :575-669). */
int mlx_idel2(mlx_ctx* ctx, mlx_field* f);
int mlx_idel2_preln(mlx_ctx* ctx, mlx_field* f, double preln);
/* ihelm(s, alpha, tfm) (mlegs_scalar.f90:277-285; mlegs_scalar_ops.f90:791-854). */
int mlx_ihelm(mlx_ctx* ctx, mlx_field* f, double alpha);
/* helmp(s, power, alpha, beta, tfm) (mlegs_scalar.f90:288-297;
 * mlegs_scalar_ops.f90:856-898): s <- del2^(power/2) s + beta del2 s + alpha s,
 * ln <- alpha ln; power 4, 6 or 8. */
int mlx_helmp(mlx_ctx* ctx, mlx_field* f, int power, double alpha, double beta);
/* ihelmp(s, power, alpha, beta, tfm) (mlegs_scalar.f90:300-309;
 * mlegs_scalar_ops.f90:900-1000): the inverse, per (m, kappa) a banded solve
 * with kl = ku = power and partial pivoting; ln <- ln / alpha.  A column whose
 * factorization meets a zero pivot is left unsolved and the call returns
 * MLX_E_ARG (the reference stops); febe, abcn and the vortex calls report it
 * the same way when hyperpow != 0. */
int mlx_ihelmp(mlx_ctx* ctx, mlx_field* f, int power, double alpha, double beta);

/* fftreat(s, tfm) (mlegs_scalar.f90:190-197; mlegs_scalar_ops.f90:1002-1063):
 * far-field smoothing.  delsqp, radial backward transform, (1-x)^2
 * weighting, rows >= 3 nr / 4 + 4 of columns m >= 1 zeroed, five passes of
 * smooth (:2149-2175) over rows >= 3 nr / 4, weighting removed, radial
 * forward transform, idelsqp; ln is kept, zeroat1 applied. */
int mlx_fftreat(mlx_ctx* ctx, mlx_field* f);

/* ---- time steppers (hyperpow = 0, or 4, 6, 8 through helmp / ihelmp) ---- */

/* fefe(s, nl, dt, tfm) (mlegs_scalar.f90:316-325; mlegs_scalar_ops.f90:1065-1096):
 * forward Euler on both terms, s <- s + dt (nl + visc del2 s). */
int mlx_fefe(mlx_ctx* ctx, mlx_field* s, const mlx_field* nl, double dt);
/* abab(s, s_p, nl, nl_p, dt, tfm, is_2nd_svis_p) (mlegs_scalar.f90:332-342;
 * mlegs_scalar_ops.f90:1098-1155): Adams-Bashforth on both terms; s_p <- s,
 * nl_p <- nl.  is_2nd_svis_p != 0: s_p holds the previous diffusive term. */
int mlx_abab(mlx_ctx* ctx, mlx_field* s, mlx_field* s_p, const mlx_field* nl, mlx_field* nl_p, double dt,
             int32_t is_2nd_svis_p);
/* febe(s, nl, dt, tfm) (mlegs_scalar.f90:350-359; mlegs_scalar_ops.f90:1157-1198). */
int mlx_febe(mlx_ctx* ctx, mlx_field* s, const mlx_field* nl, double dt);
/* abcn(s, s_p, nl, nl_p, dt, tfm) (mlegs_scalar.f90:367-376;
 * mlegs_scalar_ops.f90:1200-1262): s advances, s_p <- s, nl_p <- nl. */
int mlx_abcn(mlx_ctx* ctx, mlx_field* s, mlx_field* s_p, const mlx_field* nl, mlx_field* nl_p, double dt);

/* ---- vector-field operators ---------------------------------------------- */

/* vecprod: v <- v x u (mlegs_scalar.f90:381-389; mlegs_scalar_ops.f90:1264-1306). */
int mlx_vecprod(mlx_ctx* ctx, mlx_field* vr, mlx_field* vp, mlx_field* vz,
                const mlx_field* ur, const mlx_field* up, const mlx_field* uz);
/* vec2tp: PPP vector -> (psi, chi) in FFF (mlegs_scalar.f90:394-402;
 * mlegs_scalar_ops.f90:1308-1453). */
int mlx_vec2tp(mlx_ctx* ctx, const mlx_field* vr, const mlx_field* vp, const mlx_field* vz,
               mlx_field* psi, mlx_field* chi);
/* tp2vec: (psi, chi) in FFF -> PPP vector (mlegs_scalar.f90:407-415;
 * mlegs_scalar_ops.f90:1455-1545). */
int mlx_tp2vec(mlx_ctx* ctx, const mlx_field* psi, const mlx_field* chi,
               mlx_field* vr, mlx_field* vp, mlx_field* vz);
/* tp2curlvec (mlegs_scalar.f90:420-428; mlegs_scalar_ops.f90:1547-1560). */
int mlx_tp2curlvec(mlx_ctx* ctx, const mlx_field* psi, const mlx_field* chi,
                   mlx_field* wr, mlx_field* wp, mlx_field* wz);

/* ---- the 3D vortex step (src/apps/vortical_flow_3d.f90) ------------------ */

/* advection_rhs(psi, chi, nlpsi, nlchi, uz) (vortical_flow_3d.f90:353-395);
 * psi and chi are de-aliased in place as in the reference. */
int mlx_advection_rhs(mlx_ctx* ctx, mlx_field* psi, mlx_field* chi,
                      mlx_field* nlpsi, mlx_field* nlchi, const mlx_field* uz);
/* check_stability (vortical_flow_3d.f90:397-409): returns MLX_E_NONFINITE if
 * any entry of psi or chi is NaN/Inf. */
int mlx_check_stability(mlx_ctx* ctx, const mlx_field* psi, const mlx_field* chi);

/* Vortex run state: psi, chi, nlpsi, nlchi, the AB2 history and the SVV
 * gains kept on the device (the local variables of vortical_flow_3d.f90:26-31). */
typedef struct mlx_vortex mlx_vortex;
/* Takes copies of psi, chi (FFF) and uz (PPP). */
int mlx_vortex_create(mlx_ctx* ctx, const mlx_field* psi, const mlx_field* chi,
                      const mlx_field* uz, mlx_vortex** out);
int mlx_vortex_destroy(mlx_ctx* ctx, mlx_vortex* v);
/* FE-BE + Richardson bootstrap (vortical_flow_3d.f90:116-147). */
int mlx_vortex_bootstrap(mlx_ctx* ctx, mlx_vortex* v);
/* nsteps iterations of the AB2/CN loop body (vortical_flow_3d.f90:160-181)
 * with step size dt; the non-finite check is accumulated on the device and
 * reported (MLX_E_NONFINITE) when check != 0 at the end of the call. */
int mlx_vortex_step(mlx_ctx* ctx, mlx_vortex* v, int32_t nsteps, double dt, int32_t check);
/* Borrow the state fields (0: psi, 1: chi, 2: nlpsi, 3: nlchi, 4: psi_prev,
 * 5: chi_prev, 6: nlpsi_prev, 7: nlchi_prev, 8: uz) and read the SVV gains. */
int mlx_vortex_field(mlx_ctx* ctx, mlx_vortex* v, int32_t which, mlx_field** out);
int mlx_vortex_gains(mlx_ctx* ctx, mlx_vortex* v, double* gain_psi, double* gain_chi);

/* save_vort_mag's field (vortical_flow_3d.f90:411-446): out (PPP, ln 0) =
 * |curl u| of the flow (psi, chi) (FFF) at each point, computed on the real
 * and imaginary parts separately (the two azimuthal points of a packed PPP
 * entry).  Write it with mlx_field_save as the reference's msave does. */
int mlx_vorticity_magnitude(mlx_ctx* ctx, const mlx_field* psi, const mlx_field* chi, mlx_field* out);

/* ---- field files: msave / mload (src/submodules/mlegs_scalar_io.f90) ----- */

/* msave(s, fn, is_binary, is_global) (mlegs_scalar_io.f90:6-101,
 * mlegs_scalar.f90:431-439).  is_global != 0: one file `path` with the global
 * array (nr+3, npdim, nzdim), the reference's layout; is_global == 0: the
 * per-process file "<path>_0" with the global size, block size and block
 * start (0,0,0) in front (one process holds the whole field).  is_binary: a
 * Fortran stream file (int32 sizes, complex128 data, real64 ln, int32 chop
 * offsets, 3-character space); else ASCII: a "(3(1X,I10))" size record, one
 * record of 1PE24.15E3 (re, im) pairs per row, a blank record between z
 * slices, then "ln offsets" and "space".  Errors: MLX_E_FILE. */
int mlx_field_save(mlx_ctx* ctx, const mlx_field* f, const char* path, int32_t is_binary, int32_t is_global);
/* mload(fn, s, is_binary, is_global) (mlegs_scalar_io.f90:103-248,
 * mlegs_scalar.f90:442-450): reads a file of the format above into f, with
 * its space, ln and chop offsets.  The sizes in the file must equal the
 * kit's (MLX_E_FILE "mloadc: size inconsistency ..." otherwise, where the
 * reference stops); only radial offsets are accepted (MLX_E_ARG). */
int mlx_field_load(mlx_ctx* ctx, mlx_field* f, const char* path, int32_t is_binary, int32_t is_global);
/* The same file formats for a host array e(n1, n2, n3) (complex128,
 * column-major, interleaved re/im), with no context and no GPU.  The
 * message of the last failure on the calling thread is mlx_file_last_error(). */
int mlx_file_write(const char* path, const double* host, int64_t n1, int64_t n2, int64_t n3,
                   const char* space, double ln, const int32_t* offsets, int32_t is_binary, int32_t is_global);
int mlx_file_read(const char* path, double* host, int64_t n1, int64_t n2, int64_t n3,
                  char* space4, double* ln, int32_t* offsets, int32_t is_binary, int32_t is_global);
const char* mlx_file_last_error(void);

/* ---- adjoints (no reference counterpart) --------------------------------- */

/* Transposes of forward operators with respect to the Euclidean inner
 * product of the stored arrays (real and imaginary parts of every entry,
 * plus the log coefficient ln), for reverse-mode gradients.  They are exact
 * transposes on real (Hermitian) fields in the chopped spectral support;
 * fields with chopping offsets are rejected.
 *   mlx_trans_adjoint: g holds a cotangent of trans(x, g's space) for x in
 *     space `from` ("FFF" with g in PPP, or "PPP" with g in FFF); on return
 *     g holds the cotangent of x, in `from` (transposes of
 *     rtrans_forward / rtrans_backward, mlegs_scalar_ops.f90:1852-2008, and
 *     of the phi and z FFTs, :1567-1848);
 *   mlx_vecprod_adjoint: at the point (v, u) of v <- v x u (vector_product,
 *     :1264-1306), a = (ar, ap, az) holds the cotangent of v x u on entry and
 *     that of v on return; (ubr, ubp, ubz) receive that of u; all PPP;
 *   mlx_abcn_adjoint: the transpose of abcn's map (s, nl, nl_p) -> s
 *     (:1200-1262, with ihelm :791-854; hyperpow = 0 only): sbar holds the
 *     cotangent of the new s on entry and that of the old s on return;
 *     nlbar and nlpbar receive those of nl and nl_p.  Supported on the
 *     chopped support without the Nyquist column and slice;
 *   mlx_adjoint_project: zero an FFF field outside that support;
 *   mlx_febe_adjoint: the same for febe's (s, nl) -> s (:1157-1198);
 *   mlx_zeroat1_adjoint: in place, the transpose of zeroat1 (:311-325);
 *   mlx_tp2vec_adjoint / mlx_tp2curlvec_adjoint: the cotangents of (vr, vp,
 *     vz) in PPP (overwritten) to those of (psi, chi), data and log
 *     coefficients (vector_reconstruction :1455-1545, curl :1547-1560);
 *   mlx_vec2tp_adjoint: the cotangents of (psi, chi) in FFF (overwritten)
 *     to those of (vr, vp, vz) in PPP (vector_projection :1308-1453), the
 *     log coefficients of psi and chi excluded (advection_rhs sets them to
 *     zero);
 *   mlx_advection_adjoint: the vector-Jacobian product of advection_rhs at
 *     (psi, chi) with background uz (vortical_flow_3d.f90:353-395): the
 *     cotangents of (nlpsi, nlchi) in FFF (overwritten) to those of (psi,
 *     chi) in psib, chib; psi and chi are de-aliased in place, as
 *     advection_rhs does;
 *   mlx_vortex_reverse_step: one reverse step of mlx_vortex_step's map
 *     (s' = zeroat1(dealias(abcn(s, N, N_prev))), N' = N(s'), N_prev' = N;
 *     :160-181): from the cotangents (sb, nb, npb) of the state after the
 *     step, whose psi and chi are p1, c1, to those of the state before it,
 *     in place (p1, c1 de-aliased in place); vw (optional, vw_fmt != 0): the
 *     six fields v, w that mlx_vortex_step_record kept for that state;
 *   mlx_vortex_step_record: mlx_vortex_step without the stability check,
 *     copying psi and chi after step i into rec[2 i] and rec[2 i + 1] and,
 *     with rec_vw, the step's v (without uz) and w into rec_vw[6 i .. 6 i +
 *     5], *vw_fmt = 2 (stored in FP32), 1 (FP64) or 0 (not kept: the
 *     kit's step does not keep them);
 *   mlx_energy_grad: in place, a PPP velocity (vr, vp, vz) to the gradient of
 *     its kinetic energy (mlx_energy's quadrature).
 * These linearise and transpose in plain FP64 whatever the context's
 * precision; the Nyquist column and slice are outside their support. */
int mlx_trans_adjoint(mlx_ctx* ctx, mlx_field* g, const char* from);
int mlx_vecprod_adjoint(mlx_ctx* ctx, mlx_field* ar, mlx_field* ap, mlx_field* az, const mlx_field* vr,
                        const mlx_field* vp, const mlx_field* vz, const mlx_field* ur, const mlx_field* up,
                        const mlx_field* uz, mlx_field* ubr, mlx_field* ubp, mlx_field* ubz);
int mlx_abcn_adjoint(mlx_ctx* ctx, mlx_field* sbar, mlx_field* nlbar, mlx_field* nlpbar, double dt);
int mlx_adjoint_project(mlx_ctx* ctx, mlx_field* f);
int mlx_febe_adjoint(mlx_ctx* ctx, mlx_field* sbar, mlx_field* nlbar, double dt);
int mlx_zeroat1_adjoint(mlx_ctx* ctx, mlx_field* f);
int mlx_tp2vec_adjoint(mlx_ctx* ctx, mlx_field* vrb, mlx_field* vpb, mlx_field* vzb, mlx_field* psib,
                       mlx_field* chib);
int mlx_tp2curlvec_adjoint(mlx_ctx* ctx, mlx_field* wrb, mlx_field* wpb, mlx_field* wzb, mlx_field* psib,
                           mlx_field* chib);
int mlx_vec2tp_adjoint(mlx_ctx* ctx, mlx_field* psib, mlx_field* chib, mlx_field* vrb, mlx_field* vpb,
                       mlx_field* vzb);
int mlx_advection_adjoint(mlx_ctx* ctx, mlx_field* psi, mlx_field* chi, const mlx_field* uz, mlx_field* nlpsib,
                          mlx_field* nlchib, mlx_field* psib, mlx_field* chib);
int mlx_vortex_reverse_step(mlx_ctx* ctx, mlx_field* p1, mlx_field* c1, const mlx_field* uz, mlx_field* sbpsi,
                            mlx_field* sbchi, mlx_field* nbpsi, mlx_field* nbchi, mlx_field* npbpsi,
                            mlx_field* npbchi, double dt, mlx_field** vw, int32_t vw_fmt);
int mlx_vortex_step_record(mlx_ctx* ctx, mlx_vortex* v, int32_t nsteps, double dt, mlx_field** rec,
                           mlx_field** rec_vw, int32_t* vw_fmt);
int mlx_energy_grad(mlx_ctx* ctx, mlx_field* vr, mlx_field* vp, mlx_field* vz);

/* ---- ensembles (no reference counterpart) -------------------------------- */

/* n independent vortex runs on one GPU.  The reference runs one kit per
 * process (src/submodules/mlegs_spectfm_init.f90:15-16); an ensemble holds n
 * contexts, member i built from p[i] (its own kit, fields and stream, every
 * parameter free: sizes, ell, visc, dt) at precision prec (MLX_PREC_*).
 * Members are stepped by replaying a HIP graph of each member's captured
 * AB2/CN step (the step of mlx_vortex_step, dt = p[i].dt, no per-step host
 * check), launched as one parent graph or on several streams.
 *   mlx_ensemble_ctx: member i's context, for creating and filling its fields
 *     with the ordinary entry points;
 *   mlx_ensemble_set: member i's state (psi, chi in FFF, uz in PPP, fields of
 *     member i's context; copied, as mlx_vortex_create does);
 *   mlx_ensemble_bootstrap: the FE-BE + Richardson bootstrap of every member
 *     (mlx_vortex_bootstrap);
 *   mlx_ensemble_step: nsteps AB2/CN steps of every member.  mode -1 tunes
 *     the launch configuration on the first call that replays graphs (the
 *     tuning replays are steps of the run) and keeps it; mode 0 replays one
 *     parent graph per round; mode 1 replays the member graphs on nstreams
 *     streams; mode 2 does so with one host thread feeding each stream;
 *     reps (1 or 4; modes 0 to 2) is the count of steps per
 *     graph.  The first step of a member is eager, and so are the steps
 *     left over when nsteps is not a multiple of reps;
 *   mlx_ensemble_vortex: member i's run, to read fields with mlx_vortex_field
 *     (on member i's context) after mlx_ensemble_sync;
 *   mlx_ensemble_check: index of the first member whose state became
 *     non-finite since its bootstrap, or -1;
 *   mlx_ensemble_info: info[0..4] = mode, streams, steps per graph, phases
 *     (graphs per member: the order of the step's buffer permutation),
 *     members; tune (36 doubles, optional) = ms per step of all members for
 *     each candidate [mode 0, 1, 2][streams 1, 2, 4, 8, 16, 32][steps per
 *     graph 1, 4], 0 where not tried.
 * Results of a member equal those of its own single run: every launch
 * configuration issues the same kernels in the same order per member. */
typedef struct mlx_ensemble mlx_ensemble;
int mlx_ensemble_create(const mlx_params* p, int32_t n, int device, int32_t prec, mlx_ensemble** out);
int mlx_ensemble_destroy(mlx_ensemble* e);
const char* mlx_ensemble_last_error(mlx_ensemble* e);
int mlx_ensemble_ctx(mlx_ensemble* e, int32_t i, mlx_ctx** ctx);
int mlx_ensemble_set(mlx_ensemble* e, int32_t i, const mlx_field* psi, const mlx_field* chi, const mlx_field* uz);
int mlx_ensemble_bootstrap(mlx_ensemble* e);
int mlx_ensemble_step(mlx_ensemble* e, int32_t nsteps, int32_t mode, int32_t nstreams, int32_t reps);
int mlx_ensemble_sync(mlx_ensemble* e);
int mlx_ensemble_vortex(mlx_ensemble* e, int32_t i, mlx_vortex** v);
int mlx_ensemble_check(mlx_ensemble* e, int32_t* bad);
int mlx_ensemble_info(mlx_ensemble* e, int32_t* info, double* tune);

/* ---- diagnostics (no reference counterpart; FP64 reductions) ------------- */

/* Kinetic energy of a PPP velocity field by FP64 physical quadrature:
 * 0.5 * sum_{i,j,k} w_i ell^2/(1-x_i)^2 (2pi/np)(zlen/nz) (vr^2+vp^2+vz^2). */
int mlx_energy(mlx_ctx* ctx, const mlx_field* vr, const mlx_field* vp, const mlx_field* vz, double* out);

/* ---- measurement --------------------------------------------------------- */

/* Time `iters` launches of one kernel family on the context stream with HIP
 * events (after one warm-up launch) and return the mean duration in ms.
 * which: "leg_fwd", "leg_bwd", "fft_phi_fwd", "fft_phi_bwd", "fft_z_fwd",
 * "fft_z_bwd", "v2t" (the vec2tp projection with the field as all three
 * components), "lu_solve" (the ihelm solve of the AB2/CN step), "phi_cross"
 * (v x w formed inside the three forward phi transforms, the field as all
 * seven physical inputs), "copy" (a plain copy of the field, 32 bytes per
 * complex entry of HBM traffic), "read" and "write" (a read-only and a
 * write-only pass, 16 bytes per entry).  The field is used as operand and
 * left in an unspecified state of the same space. */
int mlx_time_kernel(mlx_ctx* ctx, mlx_field* f, const char* which, int32_t iters, double* ms);

/* Seeded synthetic spectral field on the host (no context, no GPU): writes
 * 2*nrdim*npdim*nzdim doubles in the reference's global FFF layout.
 * chops = (nrchop, npchop, nzchop, np, nz).  c(n,m,kappa) = a * (g1 + i g2)/sqrt(2)
 * with g1, g2 ~ N(0,1) from SplitMix64 keyed by (seed, field_id, n, m, kappa)
 * and Box-Muller, a = amp*exp(-(n+m+|kappa|)/scale) (kind 0, perturbation) or
 * exp(-n/scale) (kind 1, transform-workload field); zero outside the chops,
 * m = 0 plane Hermitian in kappa with real (n,0,0), Nyquist m and kappa zero.
 * No counterpart in the reference (its drivers call the compiler's rand()). */
int mlx_seed_spectral(double* host, int64_t nrdim, int64_t npdim, int64_t nzdim, const int32_t* chops,
                      uint64_t seed, int32_t field_id, double amp, double scale, int32_t kind);

/* Measurement probes, no counterpart in the reference.  After
 * mlx_probe_start, every launch of a probed kernel (MLX_PROBE_*) is timed by
 * HIP events that the dispatch itself stamps (its start and its end, on the
 * context stream), up to max_launches per probe, and its
 * algorithmic bytes and flops are summed.  mlx_probe_stop synchronises,
 * switches probing off and writes, per probe i, out[4i..4i+3] = launches
 * recorded, total kernel time in ms, algorithmic bytes, algorithmic flops. */
int mlx_probe_start(mlx_ctx* ctx, int32_t max_launches);
int mlx_probe_stop(mlx_ctx* ctx, double* out);

#ifdef __cplusplus
}
#endif
#endif /* MLEGS_AMD_H */
