// C ABI of the library: context, kit, fields, and the reference's scalar
// operations composed from the kernels of kit.hip, fft.hip, legendre.hip and
// spectral.hip.  Each entry point follows the control flow of the reference
// routine named in include/mlegs_amd.h; every device step is enqueued on the
// context stream.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "mlx_internal.h"

using namespace mlx;

namespace {

int space_id(const char *s) {
  if (!s) return -1;
  if (!std::strncmp(s, "PPP", 3)) return 0;
  if (!std::strncmp(s, "PFP", 3)) return 1;
  if (!std::strncmp(s, "FFP", 3)) return 2;
  if (!std::strncmp(s, "FFF", 3)) return 3;
  return -1;
}
void set_space(mlx_field *f, const char *s) { std::memcpy(f->space, s, 3), f->space[3] = 0; }
bool is_fff(const mlx_field *f) { return !std::strncmp(f->space, "FFF", 3); }
bool is_ppp(const mlx_field *f) { return !std::strncmp(f->space, "PPP", 3); }

void need_kit(mlx_ctx *c) {
  if (!c->kit.ready) MLX_THROW(MLX_E_ARG, "transformation kit is not initialized");
}

mlx_field *field_new(mlx_ctx *c) {
  mlx_field *f = new mlx_field();
  if (hipMalloc(&f->e, sizeof(double2) * c->nelem) != hipSuccess) {
    delete f;
    MLX_THROW(MLX_E_ALLOC, "scalar_init: device allocation failed");
  }
  if (hipMalloc(&f->sc, sizeof(double) * 8) != hipSuccess) {
    (void)hipFree(f->e);
    delete f;
    MLX_THROW(MLX_E_ALLOC, "scalar_init: device allocation failed");
  }
  HIPCHK(hipMemsetAsync(f->e, 0, sizeof(double2) * c->nelem, c->stream));
  HIPCHK(hipMemsetAsync(f->sc, 0, sizeof(double) * 8, c->stream));
  return f;
}

void field_free(mlx_field *f) {
  if (!f) return;
  if (f->e) (void)hipFree(f->e);
  if (f->sc) (void)hipFree(f->sc);
  delete f;
}

// deep copy (mlegs_scalar_init.f90:106-140)
void fcopy(mlx_ctx *c, mlx_field *d, const mlx_field *s) {
  if (d == s) return;
  HIPCHK(hipMemcpyAsync(d->e, s->e, sizeof(double2) * c->nelem, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d->sc, s->sc, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  std::memcpy(d->space, s->space, 4);
  std::memcpy(d->off, s->off, sizeof(d->off));
}

// dst takes src's data (buffers are exchanged; src becomes scratch)
void fmove(mlx_field *d, mlx_field *s) {
  std::swap(d->e, s->e);
  std::swap(d->sc, s->sc);
  std::memcpy(d->space, s->space, 4);
  std::memcpy(d->off, s->off, sizeof(d->off));
}

mlx_field *tmp(mlx_ctx *c, size_t i) {
  if (c->pool.size() <= i) c->pool.resize(i + 1, nullptr);
  if (!c->pool[i]) c->pool[i] = field_new(c);
  return c->pool[i];
}

void swap_scr(mlx_ctx *c, mlx_field *f) { std::swap(f->e, c->scr[0]); }

// ---------------------------------------------------------------------------
// trans (mlegs_scalar_ops.f90:157-235)
// ---------------------------------------------------------------------------
// rscale < 0: the backward phi transform also divides rows i < nr by r_i
void trans(mlx_ctx *c, mlx_field *f, const char *space, int rscale = 0) {
  need_kit(c);
  int cur = space_id(f->space);
  if (cur < 0) MLX_THROW(MLX_E_LAYOUT, "trans: scalar space info corrupted (only accepting PPP, PFP, FFP and FFF)");
  int nw = space_id(space);
  if (nw < 0) MLX_THROW(MLX_E_ARG, "trans: only taking PPP, PFP, FFP and FFF for spectral transformation");
  const KitDev &k = c->kit;
  if (f->off[1] != 0 || f->off[2] != 0) MLX_THROW(MLX_E_ARG, "trans: only radial chopping offsets are supported");
  const int start = cur;
  if (cur < nw) {
    while (true) {
      if (cur == 0) {
        if (k.np > 1) fft_phi(c, f->e, f->e, -1);
        set_space(f, "PFP");
      }
      if (cur == 1) {
        k_ln_column(c, f->e, f->sc, -1.0);
        leg_forward(c, f->e, c->scr[0], f->off[0]);
        swap_scr(c, f);
        set_space(f, "FFP");
      }
      if (cur == 2) {
        // after rtrans_forward in this call the inactive lines are zero: only
        // the active (n, m) lines need the z transform
        const bool act = start <= 1;
        if (k.nz > 1)
          fft_z(c, f->e, -1, active_pairs(k, f->off[0]), act ? k.nrchop + f->off[0] : 0, act ? k.npchop : 0);
        set_space(f, "FFF");
      }
      if (cur == 3) break;
      ++cur;
      if (cur == nw) break;
    }
  }
  if (cur > nw) {
    while (true) {
      if (cur == 3) {
        // when rtrans_backward follows, it reads the active (n, m) lines only
        const bool act = nw <= 1;
        if (k.nz > 1)
          fft_z(c, f->e, +1, active_pairs(k, f->off[0]), act ? k.nrchop + f->off[0] : 0, act ? k.npchop : 0);
        set_space(f, "FFP");
      }
      if (cur == 2) {
        leg_backward(c, f->e, c->scr[0], f->off[0]);
        swap_scr(c, f);
        k_ln_column(c, f->e, f->sc, +1.0);
        set_space(f, "PFP");
      }
      if (cur == 1) {
        if (k.np > 1) fft_phi(c, f->e, f->e, +1, rscale);
        else if (rscale < 0) k_scale_rows_r(c, f->e, 1);
        set_space(f, "PPP");
      }
      if (cur == 0) break;
      --cur;
      if (cur == nw) break;
    }
  }
}

void chop(mlx_ctx *c, mlx_field *f) { k_chop(c, f->e, f->space, f->off[0]); }
void dealias(mlx_ctx *c, mlx_field *f) {
  need_kit(c);
  k_dealias(c, f->e, f->space);
}

void to_fff(mlx_ctx *c, mlx_field *f) {
  if (!is_fff(f)) trans(c, f, "FFF");
}

void zeroat1(mlx_ctx *c, mlx_field *f) {
  to_fff(c, f);
  k_zeroat1(c, f->e, c->kit.nrchop + f->off[0]);
}

// zeroat1 on two fields in one launch; ln_zero: their log coefficients are
// then set to zero by the same launch
void zeroat1_pair(mlx_ctx *c, mlx_field *a, mlx_field *b, bool ln_zero = false) {
  to_fff(c, a);
  to_fff(c, b);
  k_zeroat1(c, a->e, c->kit.nrchop + a->off[0], b->e, c->kit.nrchop + b->off[0], ln_zero ? a->sc : nullptr,
            ln_zero ? b->sc : nullptr);
}

void bandmul(mlx_ctx *c, mlx_field *f, int op) {
  need_kit(c);
  to_fff(c, f);
  k_bandmul(c, f->e, f->sc, op, f->off[0]);
}

void ihelm(mlx_ctx *c, mlx_field *f, double alpha) {
  need_kit(c);
  to_fff(c, f);
  if (std::fabs(alpha) < 5e-14) MLX_THROW(MLX_E_ARG, "ihelm: alpha equals to zero. Inversion of 0*identity is impossible");
  k_helm_pre(c, f->e, f->sc, f->sc, alpha);
  k_helm_solve(c, f->e, alpha, 0, f->off[0]);
}

// helmp / ihelmp (mlegs_scalar_ops.f90:856-1000): the powered Helmholtz
// operator del2^(power/2) + beta del2 + alpha and its inverse
void check_power(const char *name, int power) {
  if (power % 2 != 0 || power < 4)
    MLX_THROW(MLX_E_ARG, std::string(name) + ": even power greater than or equal to 4");
  if (power > 8)
    MLX_THROW(MLX_E_ARG, std::string(name) + ": power must be less than or equal to 8 (supported power = 4, 6 or 8)");
}

void helmp(mlx_ctx *c, mlx_field *f, int power, double alpha, double beta) {
  need_kit(c);
  check_power("helmp", power);
  to_fff(c, f);
  // s2 = del2(s), sp = del2^(power/2)(s); s <- sp + beta s2 + alpha s,
  // ln <- alpha ln (del2's results carry no log term)
  mlx_field *s2 = tmp(c, 7), *sp = tmp(c, 8);
  fcopy(c, s2, f);
  bandmul(c, s2, 0);
  fcopy(c, sp, s2);
  for (int n = 1; n < power / 2; ++n) bandmul(c, sp, 0);
  k_axpby(c, f->e, f->sc, alpha, s2->e, s2->sc, beta);
  k_axpby(c, f->e, nullptr, 1.0, sp->e, nullptr, 1.0);
}

// ihelmp's singular-system flag (device int): set by a column whose
// factorization meets a zero pivot, cleared and read by the entry points
int *sing_flag(mlx_ctx *c) { return reinterpret_cast<int *>(c->dsmall + 4000 + 40); }
void clear_sing(mlx_ctx *c) { k_set_flag(c, sing_flag(c), 0); }
void check_sing(mlx_ctx *c) {
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, sing_flag(c), sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (h) MLX_THROW(MLX_E_ARG, "ihelmp: zero pivot in a banded system (singular); the column is left unsolved");
}

// The modes de-aliasing zeroes (dealias, mlegs_scalar_ops.f90:305-352):
// columns m >= pcut and kappa slices zlo..zhi.  When the advection term's
// inputs are de-aliased and its result is de-aliased anyway, the MFMA path
// transforms columns m < pcut only and the projection skips the zlo..zhi
// slices (on = false: every column and slice, as without de-aliasing).
struct DealiasLimits {
  bool on = false;
  int pcut = -1, zlo = 1, zhi = 0;
};

// dl (the time loop, which de-aliases f right after): the systems of the
// modes de-aliasing zeroes are not solved
void ihelmp(mlx_ctx *c, mlx_field *f, int power, double alpha, double beta, const DealiasLimits *dl = nullptr) {
  need_kit(c);
  check_power("ihelmp", power);
  to_fff(c, f);
  if (std::fabs(alpha) < 5e-14)
    MLX_THROW(MLX_E_ARG, "ihelmp: alpha equals to zero. Inversion of 0*identity is impossible");
  const bool lim = dl && dl->on;
  k_ihelmp(c, f->e, f->sc, power, alpha, beta, f->off[0], sing_flag(c), lim ? dl->pcut : -1, lim ? dl->zlo : 1,
           lim ? dl->zhi : 0);
}

// fftreat (mlegs_scalar_ops.f90:1002-1063): far-field smoothing.  delsqp
// (log term left out), rtrans_backward, the physical-space weighting,
// zeroing and smoothing, rtrans_forward, idelsqp; then the original log
// coefficient is restored and zeroat1 applied.
void fftreat(mlx_ctx *c, mlx_field *f) {
  need_kit(c);
  to_fff(c, f);
  const KitDev &k = c->kit;
  if (k.nrchop + f->off[0] > k.nrdim_ref) MLX_THROW(MLX_E_ARG, "chop_index: chopping in r too large");
  mlx_field *so = tmp(c, 15);
  fcopy(c, so, f);
  k_set_scalar(c, so->sc, 0.0);
  k_delsqp(c, so->e, so->sc, so->off[0], 0);
  leg_backward(c, so->e, c->scr[0], so->off[0], true);
  swap_scr(c, so);
  k_fftreat_mid(c, so->e, k.npchop);
  leg_forward(c, so->e, c->scr[0], so->off[0], true);
  swap_scr(c, so);
  k_delsqp(c, so->e, so->sc, so->off[0], 1);
  k_copy_scalar(c, so->sc, f->sc, 1.0);
  k_zeroat1(c, so->e, k.nrchop + so->off[0]);
  set_space(so, "FFF");
  fmove(f, so);
}

void idel2_proln(mlx_ctx *c, mlx_field *f) {
  need_kit(c);
  to_fff(c, f);
  k_helm_solve(c, f->e, 0.0, 1, f->off[0]);
  k_idel2_post(c, f->e, f->sc);
}

void check_step_inputs(mlx_ctx *c, std::initializer_list<const mlx_field *> fs) {
  for (const mlx_field *f : fs)
    if (!is_fff(f)) MLX_THROW(MLX_E_LAYOUT, "fefe: all input scalars must be in FFF for time stepping");
}

// The hyperviscous branches of febe and abcn (mlegs_scalar_ops.f90:1181-1195,
// 1226-1230 and 1241-1255): the diffusive operator visc del2 +
// hypervisc del^p as g (del^p + b del2), g = hypervisc (-1)^(p/2+1),
// b = visc / g, inverted by ihelmp with alpha = -1/(dt g) (febe) or
// -2/(dt g) (abcn, whose right-hand side adds dt/2 g helmp(s, p, 0, b))
// dl: as abcn's (the time loop de-aliases s right after)
void step_hyper(mlx_ctx *c, mlx_field *s, const mlx_field *nl, const mlx_field *nl_p, double dt,
                const DealiasLimits *dl = nullptr) {
  const int pw = c->p.hyperpow;
  const double g = c->p.hypervisc * (((pw / 2 + 1) % 2) ? -1.0 : 1.0);
  const double b = c->p.visc / g;
  if (nl_p) {
    // helmp(s, pw, 0, b) and the right-hand side in one column pass
    const double a = -2.0 / (dt * g);
    const bool lim = dl && dl->on;
    k_abcn_rhs_hyp(c, s->e, s->sc, nl->e, nl->sc, nl_p->e, nl_p->sc, dt, a, g, b, pw, s->off[0],
                   lim ? dl->pcut : -1, lim ? dl->zlo : 1, lim ? dl->zhi : 0, lim);
    ihelmp(c, s, pw, a, b, dl);
  } else {
    const double a = -1.0 / (dt * g);
    k_abcn_rhs_g(c, s->e, s->sc, s->e, s->sc, nl->e, nl->sc, nullptr, nullptr, nullptr, dt, a, 0, g);
    ihelmp(c, s, pw, a, b);
  }
}

// The diffusive term visc del2 s (or the hyperviscous visc del2 + hypervisc
// del^p) of fefe / abab into sv, a copy of s (mlegs_scalar_ops.f90:1077-1090).
// Inviscid: the data is zeroed and the log term left as the copy's, as the
// reference does.
void diffusive_term(mlx_ctx *c, mlx_field *sv, const mlx_field *s) {
  fcopy(c, sv, s);
  const int pw = c->p.hyperpow;
  if (pw == 0) {
    if (c->p.visc < 5e-14) {
      HIPCHK(hipMemsetAsync(sv->e, 0, sizeof(double2) * c->nelem, c->stream));
    } else {
      bandmul(c, sv, 0);
      k_axpby(c, sv->e, sv->sc, c->p.visc, sv->e, sv->sc, 0.0);
    }
  } else {
    const double g = c->p.hypervisc * (((pw / 2 + 1) % 2) ? -1.0 : 1.0);
    helmp(c, sv, pw, 0.0, c->p.visc / g);
    k_axpby(c, sv->e, sv->sc, g, sv->e, sv->sc, 0.0);
  }
}

// fefe (mlegs_scalar_ops.f90:1065-1096): s <- s + dt (nl + D s)
void fefe(mlx_ctx *c, mlx_field *s, const mlx_field *nl, double dt) {
  check_step_inputs(c, {s, nl});
  mlx_field *sv = tmp(c, 16);
  diffusive_term(c, sv, s);
  k_axpby(c, sv->e, sv->sc, 1.0, nl->e, nl->sc, 1.0);
  k_axpby(c, s->e, s->sc, 1.0, sv->e, sv->sc, dt);
}

// abab (mlegs_scalar_ops.f90:1098-1155): s <- s + dt (1.5 (nl + D s) -
// 0.5 (nl_p + D s_p)), then s_p <- s, nl_p <- nl.  is_2nd_svis_p: s_p already
// holds D s_p.  Inviscid with s_p given: the reference zeroes D s (not
// D s_p), so s_p itself enters as its diffusive term; kept.
void abab(mlx_ctx *c, mlx_field *s, mlx_field *s_p, const mlx_field *nl, mlx_field *nl_p, double dt,
          bool is_2nd_svis_p) {
  check_step_inputs(c, {s, nl, s_p, nl_p});
  mlx_field *sv = tmp(c, 16), *svp = tmp(c, 17);
  diffusive_term(c, sv, s);
  fcopy(c, svp, s_p);
  if (!is_2nd_svis_p) {
    const int pw = c->p.hyperpow;
    if (pw == 0) {
      if (!(c->p.visc < 5e-14)) {
        bandmul(c, svp, 0);
        k_axpby(c, svp->e, svp->sc, c->p.visc, svp->e, svp->sc, 0.0);
      }
    } else {
      const double g = c->p.hypervisc * (((pw / 2 + 1) % 2) ? -1.0 : 1.0);
      helmp(c, svp, pw, 0.0, c->p.visc / g);
      k_axpby(c, svp->e, svp->sc, g, svp->e, svp->sc, 0.0);
    }
  }
  // sv <- 1.5 (nl + sv) - 0.5 (nl_p + svp)
  k_axpby(c, sv->e, sv->sc, 1.0, nl->e, nl->sc, 1.0);
  k_axpby(c, svp->e, svp->sc, 1.0, nl_p->e, nl_p->sc, 1.0);
  k_axpby(c, sv->e, sv->sc, 1.5, svp->e, svp->sc, -0.5);
  k_axpby(c, s->e, s->sc, 1.0, sv->e, sv->sc, dt);
  fcopy(c, s_p, s);
  fcopy(c, nl_p, nl);
}

// febe (mlegs_scalar_ops.f90:1157-1198)
void febe(mlx_ctx *c, mlx_field *s, const mlx_field *nl, double dt) {
  check_step_inputs(c, {s, nl});
  if (c->p.hyperpow != 0) {
    step_hyper(c, s, nl, nullptr, dt);
    return;
  }
  if (c->p.visc < 5e-14) MLX_THROW(MLX_E_ARG, "febe: inviscid case and no linear term in rhs. semi-implicit time adv is impossible");
  double a = -1.0 / (dt * c->p.visc);
  k_abcn_rhs(c, s->e, s->sc, s->e, s->sc, nl->e, nl->sc, nullptr, nullptr, nullptr, dt, a, 0);
  ihelm(c, s, a);
}

DealiasLimits dealias_limits(mlx_ctx *c) {
  const KitDev &k = c->kit;
  DealiasLimits d;
  // the FP64 contraction kernels compute every column and slice: the
  // restrictions hold only when every site of the advection term runs on MFMA
  for (int s = 0; s < MLX_STAGE_COUNT; ++s)
    if (stage_precision(c, s) == MLX_PREC_FP64) return d;
  if (!mfma_ok(k, 0) || k.np <= 1) return d;
  d.on = true;
  d.pcut = k.np / 3 + 1 > 1 ? k.np / 3 + 1 : 1;
  if (d.pcut > k.npchop) d.pcut = k.npchop;
  if (k.nz > 1) {
    const int zcut = (k.nz / 3 + 1 > 1) ? k.nz / 3 + 1 : 1;
    d.zlo = zcut;
    d.zhi = k.nz - zcut;  // zupper - 2 with zupper = nz - zcut + 2
  }
  return d;
}

// abcn (mlegs_scalar_ops.f90:1200-1262); copy_prev = false skips s_p <- s
// (s_p is never read by the stepper) and exchanges nl_p <- nl by buffer swap
// when nl may be overwritten afterwards (swap_nl).
// dl (the time loop, which de-aliases s right after): the modes de-aliasing
// zeroes are neither formed nor solved
void abcn(mlx_ctx *c, mlx_field *s, mlx_field *s_p, mlx_field *nl, mlx_field *nl_p, double dt, bool copy_prev,
          bool swap_nl, const DealiasLimits *dl = nullptr) {
  check_step_inputs(c, {s, nl, s_p, nl_p});
  if (c->p.hyperpow != 0) {
    step_hyper(c, s, nl, nl_p, dt, dl);
    if (copy_prev) fcopy(c, s_p, s);
    if (swap_nl) fmove(nl_p, nl);
    else fcopy(c, nl_p, nl);
    return;
  }
  if (c->p.visc < 5e-14) MLX_THROW(MLX_E_ARG, "abcn: inviscid case and no linear term in rhs. semi-implicit time adv is impossible");
  double a = -2.0 / (dt * c->p.visc);
  const bool lim = dl && dl->on;
  if (lim && std::fabs(a) < 5e-14)
    MLX_THROW(MLX_E_ARG, "ihelm: alpha equals to zero. Inversion of 0*identity is impossible");
  // lim: ihelm's log-term step rides on the log-coefficient update
  k_abcn_rhs_d2(c, s->e, s->sc, nl->e, nl->sc, nl_p->e, nl_p->sc, dt, a, s->off[0], lim ? dl->pcut : -1,
                lim ? dl->zlo : 1, lim ? dl->zhi : 0, lim);
  if (lim) {
    k_helm_solve(c, s->e, a, 0, s->off[0], dl->pcut, dl->zlo, dl->zhi);
  } else {
    ihelm(c, s, a);
  }
  if (copy_prev) fcopy(c, s_p, s);
  if (swap_nl) fmove(nl_p, nl);
  else fcopy(c, nl_p, nl);
}

// abcn on psi and chi with the de-aliasing restrictions (the time loop): the
// two right-hand sides, then both Helmholtz solves in one launch (same alpha,
// same factors, read once for both fields)
void abcn_pair(mlx_ctx *c, mlx_field *s1, mlx_field *s1p, mlx_field *nl1, mlx_field *nl1p, mlx_field *s2,
               mlx_field *s2p, mlx_field *nl2, mlx_field *nl2p, double dt, const DealiasLimits &dl) {
  if (!dl.on || s1->off[0] != s2->off[0] || c->p.hyperpow != 0) {
    abcn(c, s1, s1p, nl1, nl1p, dt, false, true, &dl);
    abcn(c, s2, s2p, nl2, nl2p, dt, false, true, &dl);
    return;
  }
  check_step_inputs(c, {s1, nl1, s1p, nl1p, s2, nl2, s2p, nl2p});
  if (c->p.visc < 5e-14) MLX_THROW(MLX_E_ARG, "abcn: inviscid case and no linear term in rhs. semi-implicit time adv is impossible");
  const double a = -2.0 / (dt * c->p.visc);
  if (std::fabs(a) < 5e-14) MLX_THROW(MLX_E_ARG, "ihelm: alpha equals to zero. Inversion of 0*identity is impossible");
  k_abcn_rhs_d2(c, s1->e, s1->sc, nl1->e, nl1->sc, nl1p->e, nl1p->sc, dt, a, s1->off[0], dl.pcut, dl.zlo, dl.zhi,
                true);
  k_abcn_rhs_d2(c, s2->e, s2->sc, nl2->e, nl2->sc, nl2p->e, nl2p->sc, dt, a, s2->off[0], dl.pcut, dl.zlo, dl.zhi,
                true);
  k_helm_solve(c, s1->e, a, 0, s1->off[0], dl.pcut, dl.zlo, dl.zhi, s2->e);
  fmove(nl1p, nl1);
  fmove(nl2p, nl2);
}

// tp2vec_back's z transforms of the nf fields f[0..nf) in one launch
void tp2vec_back_z(mlx_ctx *c, mlx_field *const *f, int nf, int ncols, int zlo, int zhi) {
  const KitDev &k = c->kit;
  if (k.nz <= 1) return;
  const int nc = (ncols > 0 && ncols < k.npchop) ? ncols : k.npchop;
  double2 *e[kZBatch];
  for (int q = 0; q < nf; ++q) e[q] = f[q]->e;
  fft_z_batch(c, e, nf, +1, active_pairs(k, 0, nc), k.nrchop, nc, zlo, zhi);
}

// Back half of tp2vec: the three FFF components (offset 0, no log term) to
// PPP; ur and up are divided by r in the last phi transform.  ncols > 0: the
// components are zero in columns m >= ncols, which the transforms skip.
// The front kernel has zeroed the log terms, so trans's log-column step
// (adding 0 * ln x) is left out; the other steps are trans's FFF -> PPP.
// phys_rows: rows nr.. of the outputs are not needed (the caller's cross
// product takes them as zero) and are not synthesized
// f32_ok: the caller reads the outputs through fft_phi_cross's FP32 input;
// when the radial pairs path runs (FP32 sums) the outputs are then stored in
// FP32.  Returns whether they were.
// z_done: the caller has run the z transforms (tp2vec_back_z)
bool tp2vec_back(mlx_ctx *c, mlx_field *ur, mlx_field *up, mlx_field *uz, int ncols = -1, int zlo = 1,
                 int zhi = 0, bool phys_rows = false, bool f32_ok = false, bool z_done = false) {
  const KitDev &k = c->kit;
  PrecScope ps(c, MLX_STAGE_SYNTH);
  const int rs[3] = {-1, -1, 0};
  const int nc = (ncols > 0 && ncols < k.npchop) ? ncols : k.npchop;
  // limited columns and the four-step phi kernels: the radial transform
  // leaves columns m >= nc unwritten and the phi transform takes them as zero
  const bool lean = nc < k.npchop && fft_phi_fast(k);
  // physical rows only, on the MFMA path: the radial transform hands the
  // phi transform its FP32 even/odd sums (half the bytes of the FP64 rows,
  // which the phi transform forms from them exactly)
  const bool pairs = lean && phys_rows && (c->prec == MLX_PREC_FP32 || c->prec == MLX_PREC_BF16X3) && mfma_ok(k, 0);
  // the de-aliased kappa slices zlo..zhi of the input are zero
  mlx_field *const three[3] = {ur, up, uz};
  if (!z_done) tp2vec_back_z(c, three, 3, ncols, zlo, zhi);
  int q = 0;
  for (mlx_field *f : three) {
    f->off[0] = f->off[1] = f->off[2] = 0;
    if (pairs) {
      float4 *pr = reinterpret_cast<float4 *>(c->scr[0]);
      leg_backward_pairs_f32(c, f->e, pr, nc);
      fft_phi_pairs(c, f->e, pr, rs[q], nc, f32_ok);
      set_space(f, "PPP");
      ++q;
      continue;
    }
    leg_backward(c, f->e, c->scr[0], 0, true, nc, !lean);
    swap_scr(c, f);
    if (k.np > 1) fft_phi(c, f->e, f->e, +1, rs[q], lean ? nc : -1, phys_rows && fft_phi_fast(k));
    else if (rs[q] < 0) k_scale_rows_r(c, f->e, 1);
    set_space(f, "PPP");
    ++q;
  }
  return pairs && f32_ok;
}

// tp2vec (vector_reconstruction, mlegs_scalar_ops.f90:1455-1545): the +3-offset
// band products, the combination and the chop run in one column pass
// ncols > 0: psi and chi are zero in columns m >= ncols (de-aliased)
void tp2vec(mlx_ctx *c, const mlx_field *psi, const mlx_field *chi, mlx_field *vr, mlx_field *vp, mlx_field *vz,
            int ncols = -1, int zlo = 1, int zhi = 0) {
  need_kit(c);
  if (!is_fff(psi) || !is_fff(chi)) MLX_THROW(MLX_E_LAYOUT, "vector_projection: psi and chi must be in FFF");
  mlx_field *ur = tmp(c, 0), *up = tmp(c, 1), *uz = tmp(c, 2);
  // the z transform of tp2vec_back takes zlo..zhi as zero (four-step kernels)
  const bool zskip = fft_z_fast(c->kit);
  k_tp2vec_front(c, ur->e, up->e, uz->e, psi->e, chi->e, psi->sc, chi->sc, 0, 0, ur->sc, up->sc, uz->sc, ncols,
                 zskip ? zlo : 1, zskip ? zhi : 0);
  tp2vec_back(c, ur, up, uz, ncols, zlo, zhi);
  fmove(vr, ur);
  fmove(vp, up);
  fmove(vz, uz);
}

// tp2curlvec (curl_vector_reconstruction, :1547-1560): curl V = tp2vec(-del2 chi, psi),
// with -del2 chi formed inside the same column pass
void tp2curlvec(mlx_ctx *c, const mlx_field *psi, const mlx_field *chi, mlx_field *wr, mlx_field *wp,
                mlx_field *wz, int ncols = -1, int zlo = 1, int zhi = 0) {
  need_kit(c);
  if (!is_fff(psi) || !is_fff(chi)) MLX_THROW(MLX_E_LAYOUT, "vector_projection: psi and chi must be in FFF");
  mlx_field *ur = tmp(c, 0), *up = tmp(c, 1), *uz = tmp(c, 2);
  // del2 of md = chi runs over the rows of chi's own radial offset
  const bool zskip = fft_z_fast(c->kit);
  k_tp2vec_front(c, ur->e, up->e, uz->e, chi->e, psi->e, chi->sc, psi->sc, 1, chi->off[0], ur->sc, up->sc,
                 uz->sc, ncols, zskip ? zlo : 1, zskip ? zhi : 0);
  tp2vec_back(c, ur, up, uz, ncols, zlo, zhi);
  fmove(wr, ur);
  fmove(wp, up);
  fmove(wz, uz);
}

void vecprod(mlx_ctx *c, mlx_field *vr, mlx_field *vp, mlx_field *vz, const mlx_field *ur, const mlx_field *up,
             const mlx_field *uz) {
  for (const mlx_field *f : {(const mlx_field *)vr, (const mlx_field *)vp, (const mlx_field *)vz, ur, up, uz})
    if (!is_ppp(f)) MLX_THROW(MLX_E_LAYOUT, "vector_product: to compute v x u, all inputs must be in PPP");
  k_vecprod(c, vr->e, vp->e, vz->e, ur->e, up->e, uz->e);
  for (mlx_field *f : {vr, vp, vz}) {
    k_set_scalar(c, f->sc, 0.0);
    f->off[0] = f->off[1] = f->off[2] = 0;
  }
}

// ---------------------------------------------------------------------------
// adjoints: transposes with respect to the Euclidean inner products of the
// stored arrays (real and imaginary parts of every entry) plus the log
// coefficient; see adjoint.hip for the transform identities
// ---------------------------------------------------------------------------
void adj_no_offsets(const mlx_field *f, const char *what) {
  if (f->off[0] || f->off[1] || f->off[2])
    MLX_THROW(MLX_E_ARG, std::string(what) + ": chopping offsets are not supported by the adjoint");
}

// The transpose of trans without its log term (for cotangents whose log
// coefficient is not wanted), with a row factor r^rmode on the physical
// side; the diagonal factors run as one pass.  to_fff: g in PPP holds the
// cotangent of r^rmode B(x) and returns that of x, in FFF; otherwise g in
// FFF holds the cotangent of F(r^rmode y) and returns that of y, in PPP.
// lim (optional, with the four-step phi kernels): the FFF side is de-aliased
// (columns m >= pcut and slices zlo..zhi zero on input, or not needed on
// output: to_fff leaves them unwritten).  The nf (<= kZBatch) fields gs[q]
// (row factors rmodes[q]) share one launch of the z transform.
void trans_adjoint_nolog_n(mlx_ctx *c, mlx_field *const *gs, int nf, bool to_fff, const int *rmodes,
                           const DealiasLimits *lim = nullptr) {
  const KitDev &k = c->kit;
  const int nc = lim ? lim->pcut : -1;
  const int zlo = lim ? lim->zlo : 1, zhi = lim ? lim->zhi : 0;
  const int ncz = lim ? lim->pcut : 0;
  double2 *e[kZBatch];
  auto z = [&](int dir) {
    if (k.nz <= 1) return;
    for (int q = 0; q < nf; ++q) e[q] = gs[q]->e;
    if (lim) {
      if (dir < 0) fft_z_batch(c, e, nf, -1, active_pairs(k, 0, nc), k.nrchop, ncz, 1, 0, zlo, zhi);
      else fft_z_batch(c, e, nf, +1, active_pairs(k, 0, nc), k.nrchop, ncz, zlo, zhi);
    } else {
      fft_z_batch(c, e, nf, dir, (double)k.nrdim * k.npdim);
    }
  };
  if (to_fff) {
    // B^T = Z_b^T L_b^T Phi_b^T (L_b^T = L_f W^-1, Z_b^T = nz Z_f)
    for (int q = 0; q < nf; ++q) {
      mlx_field *g = gs[q];
      if (k.np > 1) fft_phi(c, g->e, g->e, -1);
      k_adj_diag(c, g->e, -1, rmodes[q], k.np > 1 ? +1 : 0, k.nz > 1 ? (double)k.nz : 1.0);
      leg_forward(c, g->e, c->scr[0], 0, true, nc, !lim);
      swap_scr(c, g);
    }
    z(-1);
    for (int q = 0; q < nf; ++q) set_space(gs[q], "FFF");
  } else {
    // F^T = Phi_f^T L_f^T Z_f^T (L_f^T = W L_b, Z_f^T = Z_b / nz)
    z(+1);
    for (int q = 0; q < nf; ++q) {
      mlx_field *g = gs[q];
      leg_backward(c, g->e, c->scr[0], 0, true, nc, !lim);
      swap_scr(c, g);
      k_adj_diag(c, g->e, +1, rmodes[q], k.np > 1 ? -1 : 0, k.nz > 1 ? 1.0 / k.nz : 1.0);
      if (k.np > 1) fft_phi(c, g->e, g->e, +1, 0, lim ? lim->pcut : -1);
      set_space(g, "PPP");
    }
  }
}

void trans_adjoint_nolog(mlx_ctx *c, mlx_field *g, bool to_fff, int rmode, const DealiasLimits *lim = nullptr) {
  trans_adjoint_nolog_n(c, &g, 1, to_fff, &rmode, lim);
}

// g holds a cotangent of trans(x, space of g) for x in space `from` (FFF or
// PPP, the other end); on return g holds the cotangent of x, in `from`
void trans_adjoint(mlx_ctx *c, mlx_field *g, const char *from) {
  need_kit(c);
  adj_no_offsets(g, "trans adjoint");
  const KitDev &k = c->kit;
  const bool to_fff = !std::strncmp(from, "FFF", 3);
  if (to_fff ? !is_ppp(g) : !(std::strncmp(from, "PPP", 3) == 0 && is_fff(g)))
    MLX_THROW(MLX_E_ARG, "trans adjoint: the cotangent must be in PPP (from FFF) or FFF (from PPP)");
  if (to_fff) {
    // backward transform B = Phi_b Lambda L_b Z_b; B^T = Z_b^T L_b^T Lambda^T Phi_b^T
    if (k.np > 1) {
      fft_phi(c, g->e, g->e, -1);
      k_adj_phi_scale(c, g->e, +1);
    }
    k_ln_adjoint(c, g->e, g->sc, +1.0);
    k_adj_wscale(c, g->e, -1);
    leg_forward(c, g->e, c->scr[0], 0);
    swap_scr(c, g);
    if (k.nz > 1) {
      fft_z(c, g->e, -1, (double)k.nrdim * k.npdim);
      k_adj_scale(c, g->e, (double)k.nz);
    }
    set_space(g, "FFF");
  } else {
    // forward transform F = Z_f L_f Lambda^- Phi_f; F^T = Phi_f^T Lambda^-T L_f^T Z_f^T
    if (k.nz > 1) {
      fft_z(c, g->e, +1, (double)k.nrdim * k.npdim);
      k_adj_scale(c, g->e, 1.0 / k.nz);
    }
    leg_backward(c, g->e, c->scr[0], 0);
    swap_scr(c, g);
    k_adj_wscale(c, g->e, +1);
    k_ln_adjoint(c, g->e, g->sc, -1.0);
    if (k.np > 1) {
      k_adj_phi_scale(c, g->e, -1);
      fft_phi(c, g->e, g->e, +1);
    }
    set_space(g, "PPP");
  }
}

// abcn (hyperpow = 0) maps (s, nl, nl_p) to s' = (del2 + a)^-1 E with
// R = a (s + dt (1.5 nl - 0.5 nl_p) + dt/2 visc del2 s), L = s.ln + dt (1.5
// nl.ln - 0.5 nl_p.ln), E = R - lap_ln(L), s'.ln = L, a = -2 / (dt visc)
// (mlegs_scalar_ops.f90:1200-1262, ihelm :791-854), on the chopped support
// (rows n < nrcs(m), m < npchop, the kept z slices; Nyquist column and
// slice excluded).  sbar: the cotangent of s' on entry, of s on return;
// nlbar, nlpbar: the cotangents of nl and nl_p (overwritten).
// projected: sbar is already zero outside that support.  nf = 2: a second
// cotangent (sbar[1], nlbar[1], nlpbar[1]) through the same map, its solve
// in the same launch (the same factors).
void abcn_adjoint(mlx_ctx *c, mlx_field *const *sbar, mlx_field *const *nlbar, mlx_field *const *nlpbar, int nf,
                  double dt, bool projected) {
  need_kit(c);
  for (int f = 0; f < nf; ++f) {
    for (mlx_field *x : {sbar[f], nlbar[f], nlpbar[f]}) adj_no_offsets(x, "abcn adjoint");
    if (!is_fff(sbar[f])) MLX_THROW(MLX_E_LAYOUT, "abcn adjoint: the cotangent must be in FFF");
  }
  if (c->p.hyperpow != 0) MLX_THROW(MLX_E_ARG, "abcn adjoint: the hyperviscous stepper has no adjoint here");
  if (c->p.visc < 5e-14) MLX_THROW(MLX_E_ARG, "abcn: inviscid case and no linear term in rhs. semi-implicit time adv is impossible");
  const double a = -2.0 / (dt * c->p.visc), c3 = a * dt / 2.0 * c->p.visc;
  if (!projected)
    for (int f = 0; f < nf; ++f) k_adj_project(c, sbar[f]->e);
  // y <- (del2 + a)^-T y
  k_helm_solve_t(c, sbar[0]->e, a, 0, 0, -1, 1, 0, nf > 1 ? sbar[1]->e : nullptr);
  for (int f = 0; f < nf; ++f) {
    mlx_field *y = sbar[f], *nl = nlbar[f], *nlp = nlpbar[f];
    // the log terms: L = y.ln - lap_ln^T y; nl.ln = 1.5 dt L, nl_p.ln =
    // -0.5 dt L; s.ln = L + (a dt/2 visc) lap_ln^T y (del2's log rows)
    k_abcn_adj_ln(c, y->e, y->sc, nl->sc, nlp->sc, dt, c3);
    // nl = 1.5 a dt y, nl_p = -0.5 a dt y, s = a y + (a dt/2 visc) del2^T y
    mlx_field *t = tmp(c, 3);
    k_del2_t(c, t->e, y->e, true);
    k_abcn_adj_out(c, y->e, t->e, nl->e, nlp->e, 1.5 * a * dt, -0.5 * a * dt, a, c3);
    for (mlx_field *x : {nl, nlp}) x->off[0] = x->off[1] = x->off[2] = 0;
    set_space(y, "FFF");
    set_space(nl, "FFF");
    set_space(nlp, "FFF");
  }
}

void abcn_adjoint(mlx_ctx *c, mlx_field *sbar, mlx_field *nlbar, mlx_field *nlpbar, double dt) {
  abcn_adjoint(c, &sbar, &nlbar, &nlpbar, 1, dt, false);
}

// febe (hyperpow = 0) maps (s, nl) to s' = (del2 + a)^-1 (a (s + dt nl) -
// lap_ln(L)), s'.ln = L = s.ln + dt nl.ln, a = -1 / (dt visc)
// (mlegs_scalar_ops.f90:1157-1198).  sbar: the cotangent of s' on entry, of
// s on return; nlbar: the cotangent of nl (overwritten).
void febe_adjoint(mlx_ctx *c, mlx_field *sbar, mlx_field *nlbar, double dt) {
  need_kit(c);
  for (mlx_field *f : {sbar, nlbar}) adj_no_offsets(f, "febe adjoint");
  if (!is_fff(sbar)) MLX_THROW(MLX_E_LAYOUT, "febe adjoint: the cotangent must be in FFF");
  if (c->p.hyperpow != 0) MLX_THROW(MLX_E_ARG, "febe adjoint: the hyperviscous stepper has no adjoint here");
  if (c->p.visc < 5e-14) MLX_THROW(MLX_E_ARG, "febe: inviscid case and no linear term in rhs. semi-implicit time adv is impossible");
  const double a = -1.0 / (dt * c->p.visc);
  mlx_field *y = sbar;
  k_adj_project(c, y->e);
  k_helm_solve_t(c, y->e, a);
  k_lap_ln_t(c, y->e, y->sc, -1.0);
  fcopy(c, nlbar, y);
  k_adj_scale(c, nlbar->e, a * dt);
  k_copy_scalar(c, nlbar->sc, y->sc, dt);
  k_adj_scale(c, y->e, a);
  k_adj_project(c, y->e);
  set_space(nlbar, "FFF");
}

// Runs the enclosed operations with every contraction in FP64 (on FP64
// MFMA where the kit's shape allows it, whose transforms equal the plain
// FP64 kernels' bit for bit) and restores the context's precision settings
// afterwards: the adjoints linearise and transpose in FP64 whatever
// precision the forward run used.
struct Fp64Scope {
  mlx_ctx *c;
  int prec, prec_ctx, stage[MLX_STAGE_COUNT];
  explicit Fp64Scope(mlx_ctx *c_) : c(c_), prec(c_->prec), prec_ctx(c_->prec_ctx) {
    std::memcpy(stage, c->stage_prec, sizeof(stage));
    c->prec = c->prec_ctx = MLX_PREC_FP64_MFMA;
    for (int &v : c->stage_prec) v = -1;
  }
  ~Fp64Scope() {
    c->prec = prec;
    c->prec_ctx = prec_ctx;
    std::memcpy(c->stage_prec, stage, sizeof(stage));
  }
};

// host-side metadata of a cotangent (no offsets, its space); zero_ln: also
// set its log coefficient to zero
void clear_meta(mlx_ctx *c, mlx_field *f, const char *space, bool zero_ln = false) {
  if (zero_ln) k_set_scalar(c, f->sc, 0.0);
  f->off[0] = f->off[1] = f->off[2] = 0;
  set_space(f, space);
}

// the transpose of y -> Z_f Phi_f (r^rmode y) (the phi and z transforms
// without the radial one): gs[q] (row factor rmodes[q]) in PFF on entry, PPP
// on return; one launch of the z transform for the nf fields
void pff_adjoint_n(mlx_ctx *c, mlx_field *const *gs, int nf, const int *rmodes, const DealiasLimits *lim = nullptr) {
  const KitDev &k = c->kit;
  if (k.nz > 1) {
    double2 *e[kZBatch];
    for (int q = 0; q < nf; ++q) e[q] = gs[q]->e;
    if (lim)
      fft_z_batch(c, e, nf, +1, (double)k.nr * lim->pcut, 0, lim->pcut, lim->zlo, lim->zhi);
    else
      fft_z_batch(c, e, nf, +1, (double)k.nrdim * k.npdim);
  }
  for (int q = 0; q < nf; ++q) {
    mlx_field *g = gs[q];
    k_adj_diag(c, g->e, 0, rmodes[q], k.np > 1 ? -1 : 0, k.nz > 1 ? 1.0 / k.nz : 1.0);
    if (k.np > 1) fft_phi(c, g->e, g->e, +1, 0, lim ? lim->pcut : -1);
    set_space(g, "PPP");
  }
}

// tp2vec's transpose (vector_reconstruction, mlegs_scalar_ops.f90:1455-1545):
// urb, upb, uzb: the cotangents of vr, vp, vz in PPP (overwritten); psib,
// chib: the cotangents of psi and chi (data and log coefficients) on return.
// The chop of the intermediate fields restricts everything to rows n <
// nrcs(m) of the kept slices.
// tp2vec_adjoint_transforms: the transforms' part for one or two triples of
// cotangents (vr = B(ur_s) / r, vp = B(up_s) / r, vz = B(uz_s); the log
// terms of ur_s, up_s, uz_s are zero, so no cotangent of theirs is formed),
// nf = 3 or 6 fields in one z launch
void tp2vec_adjoint_transforms(mlx_ctx *c, mlx_field *const *fs, int nf, const DealiasLimits *lim) {
  for (int q = 0; q < nf; ++q)
    if (!is_ppp(fs[q])) MLX_THROW(MLX_E_LAYOUT, "vector_reconstruction adjoint: the cotangents must be in PPP");
  const int rm[6] = {-1, -1, 0, -1, -1, 0};
  trans_adjoint_nolog_n(c, fs, nf, true, rm, lim);
}

// pre: the cotangents have been through tp2vec_adjoint_transforms (FFF)
void tp2vec_adjoint(mlx_ctx *c, mlx_field *urb, mlx_field *upb, mlx_field *uzb, mlx_field *psib, mlx_field *chib,
                    const DealiasLimits *lim = nullptr, bool pre = false) {
  need_kit(c);
  if (!pre) {
    mlx_field *const three[3] = {urb, upb, uzb};
    tp2vec_adjoint_transforms(c, three, 3, lim);
  }
  if (lim)
    k_tp2vec_t(c, psib->e, chib->e, urb->e, upb->e, uzb->e, psib->sc, chib->sc, lim->pcut, lim->zlo, lim->zhi);
  else
    k_tp2vec_t(c, psib->e, chib->e, urb->e, upb->e, uzb->e, psib->sc, chib->sc);
  clear_meta(c, urb, "FFF");
  clear_meta(c, upb, "FFF");
  clear_meta(c, uzb, "FFF");
  for (mlx_field *f : {psib, chib}) {
    f->off[0] = f->off[1] = f->off[2] = 0;
    set_space(f, "FFF");
  }
}

// tp2curlvec's transpose (curl V = tp2vec(-del2 chi, psi), :1547-1560)
void tp2curlvec_adjoint(mlx_ctx *c, mlx_field *wrb, mlx_field *wpb, mlx_field *wzb, mlx_field *psib,
                        mlx_field *chib, const DealiasLimits *lim = nullptr, bool pre = false) {
  mlx_field *md = tmp(c, 21);
  // md = -del2 chi has a zero log coefficient: its cotangent is dropped
  tp2vec_adjoint(c, wrb, wpb, wzb, md, psib, lim, pre);
  k_del2_t(c, chib->e, md->e, true);
  k_adj_scale(c, chib->e, -1.0);
  k_set_scalar(c, chib->sc, 0.0);
  k_lap_ln_t(c, md->e, chib->sc, -1.0);
  chib->off[0] = chib->off[1] = chib->off[2] = 0;
  set_space(chib, "FFF");
}

// vec2tp's transpose (vector_projection, :1308-1453) with respect to the data
// of its result (advection_rhs sets the result's log coefficients to zero,
// so theirs are not part of it).  psib, chib: the cotangents of psi and chi
// in FFF (overwritten); vrb, vpb, vzb: the cotangents of vr, vp, vz in PPP.
// lim (optional): psib and chib are de-aliased, and so is what the caller
// makes of the result, which lets every transform skip the zero modes
void vec2tp_adjoint(mlx_ctx *c, mlx_field *psib, mlx_field *chib, mlx_field *vrb, mlx_field *vpb, mlx_field *vzb,
                    const DealiasLimits *lim = nullptr) {
  need_kit(c);
  for (mlx_field *f : {psib, chib}) {
    adj_no_offsets(f, "vec2tp adjoint");
    if (!is_fff(f)) MLX_THROW(MLX_E_LAYOUT, "vector_projection adjoint: the cotangents must be in FFF");
  }
  // zeroat1 and the chop to offset 0
  k_zeroat1_t(c, psib->e);
  k_zeroat1_t(c, chib->e);
  chop(c, psib);
  chop(c, chib);
  // idel2_proln on chi's own rows (offset 0: the projection's three extra
  // rows are left as they are by the solve and chopped after it), the (0,0)
  // system bordered
  const int nc = lim ? lim->pcut : -1, zlo = lim ? lim->zlo : 1, zhi = lim ? lim->zhi : 0;
  k_helm_solve_t(c, chib->e, 0.0, 1, 0, nc, zlo, zhi);
  // the five projection products: cotangents of ur = F(r vr), up, uz = F(vz)
  // in PFF
  mlx_field *upP = tmp(c, 20);
  k_v2t_t(c, vrb->e, upP->e, vzb->e, psib->e, chib->e, nc, zlo, zhi);
  // up = L_b Z_f L_f Phi_f (r vp) + psi.ln (1 + x), psi.ln = -1/2 Re
  // calcat1(Z_f L_f Phi_f (r vp))(0): L_b^T = L_f W^-1 on the PFF cotangent,
  // then the calcat term, then the transpose of the forward transform
  fcopy(c, vpb, upP);
  k_adj_wscale(c, vpb->e, -1);
  leg_forward(c, vpb->e, c->scr[0], 0, true, nc, true);
  swap_scr(c, vpb);
  k_v2t_up_ln_t(c, vpb->e, upP->e);
  trans_adjoint_nolog(c, vpb, false, +1, lim);
  // ur = Z_f Phi_f (r vr), uz = Z_f Phi_f vz
  mlx_field *const rz[2] = {vrb, vzb};
  const int rzm[2] = {+1, 0};
  pff_adjoint_n(c, rz, 2, rzm, lim);
  for (mlx_field *f : {vrb, vpb, vzb}) clear_meta(c, f, "PPP", true);
}

// advection_rhs's vector-Jacobian product at (psi, chi) (vortical_flow_3d.f90:
// 353-395): N = dealias(vec2tp((tp2vec(psi, chi) + U) x tp2curlvec(psi, chi)))
// with psi and chi de-aliased first.  nlpsib, nlchib: the cotangents of N
// (overwritten); psib, chib: the cotangents of psi and chi on return.  The
// operators run in FP64.
// psi and chi are de-aliased in place, as advection_rhs does.  vw
// (optional): v and w at (psi, chi) as the forward step left them (v
// without the background flow), in FP32 (vw_fmt 2) or FP64 (1); without
// them they are formed here in FP64.
void advection_adjoint(mlx_ctx *c, mlx_field *psi, mlx_field *chi, const mlx_field *uz, mlx_field *nlpsib,
                       mlx_field *nlchib, mlx_field *psib, mlx_field *chib, mlx_field *const *vw = nullptr,
                       int vw_fmt = 0) {
  need_kit(c);
  if (!is_fff(psi) || !is_fff(chi)) MLX_THROW(MLX_E_LAYOUT, "advection_rhs adjoint: psi and chi must be in FFF");
  if (!is_ppp(uz)) MLX_THROW(MLX_E_LAYOUT, "advection_rhs adjoint: uz must be in PPP");
  for (const mlx_field *f : {psi, chi}) adj_no_offsets(f, "advection_rhs adjoint");
  Fp64Scope fp64(c);
  const DealiasLimits dl = dealias_limits(c);
  const bool given = vw && (vw_fmt == 1 || vw_fmt == 2);
  const double2 *v[3], *w[3];
  if (given) {
    for (int q = 0; q < 3; ++q) {
      v[q] = vw[q]->e;
      w[q] = vw[3 + q]->e;
    }
  } else {
    dealias(c, psi);
    dealias(c, chi);
    // v and w at (psi, chi): the fused column pass and the backward
    // transforms of advection_rhs, skipping what de-aliasing zeroes, in FP64
    const int ncols = dl.on ? dl.pcut : -1;
    const bool zskip = fft_z_fast(c->kit);
    mlx_field *vr = tmp(c, 24), *vp = tmp(c, 25), *vz = tmp(c, 26);
    mlx_field *wr = tmp(c, 27), *wp = tmp(c, 28), *wz = tmp(c, 29);
    k_tp2vec_front2(c, vr->e, vp->e, vz->e, vr->sc, vp->sc, vz->sc, wr->e, wp->e, wz->e, wr->sc, wp->sc, wz->sc,
                    psi->e, chi->e, psi->sc, chi->sc, chi->off[0], ncols, zskip ? dl.zlo : 1, zskip ? dl.zhi : 0);
    mlx_field *const six[6] = {vr, vp, vz, wr, wp, wz};
    tp2vec_back_z(c, six, 6, ncols, dl.zlo, dl.zhi);
    tp2vec_back(c, vr, vp, vz, ncols, dl.zlo, dl.zhi, true, false, true);
    tp2vec_back(c, wr, wp, wz, ncols, dl.zlo, dl.zhi, true, false, true);
    v[0] = vr->e, v[1] = vp->e, v[2] = vz->e;
    w[0] = wr->e, w[1] = wp->e, w[2] = wz->e;
  }
  set_space(nlpsib, "FFF");
  set_space(nlchib, "FFF");
  dealias(c, nlpsib);
  dealias(c, nlchib);
  // the transforms below skip the de-aliased modes (the four-step phi
  // kernels take the mode limit)
  const DealiasLimits *lim = (dl.on && fft_phi_fast(c->kit)) ? &dl : nullptr;
  mlx_field *ar = tmp(c, 30), *ap = tmp(c, 31), *az = tmp(c, 32);
  vec2tp_adjoint(c, nlpsib, nlchib, ar, ap, az, lim);
  // (v + U) x w: v_b = w x a, w_b = a x (v + U)
  mlx_field *br = tmp(c, 33), *bp = tmp(c, 34), *bz = tmp(c, 35);
  k_vecprod_adj(c, ar->e, ap->e, az->e, v[0], v[1], v[2], w[0], w[1], w[2], br->e, bp->e, bz->e,
                given && vw_fmt == 2, uz->e);
  set_space(br, "PPP");
  set_space(bp, "PPP");
  set_space(bz, "PPP");
  // the six cotangents' transforms, their z transforms in one launch
  mlx_field *const six[6] = {ar, ap, az, br, bp, bz};
  tp2vec_adjoint_transforms(c, six, 6, lim);
  tp2vec_adjoint(c, ar, ap, az, psib, chib, lim, true);
  mlx_field *pb = tmp(c, 36), *cb = tmp(c, 37);
  tp2curlvec_adjoint(c, br, bp, bz, pb, cb, lim, true);
  k_axpby(c, psib->e, psib->sc, 1.0, pb->e, pb->sc, 1.0);
  k_axpby(c, chib->e, chib->sc, 1.0, cb->e, cb->sc, 1.0);
  dealias(c, psib);
  dealias(c, chib);
}

// dst = PFP form of the PPP field src (optionally r-weighted): metadata and log
// term of src, then the phi transform from src into dst
void phi_forward_from(mlx_ctx *c, mlx_field *dst, const mlx_field *src, int rscale) {
  const KitDev &k = c->kit;
  std::memcpy(dst->off, src->off, sizeof(dst->off));
  k_copy_scalar(c, dst->sc, src->sc, 1.0);
  if (k.np > 1) {
    fft_phi(c, dst->e, src->e, -1, rscale);
  } else {
    HIPCHK(hipMemcpyAsync(dst->e, src->e, sizeof(double2) * c->nelem, hipMemcpyDeviceToDevice, c->stream));
    if (rscale > 0) k_scale_rows_r(c, dst->e, 0);
  }
  set_space(dst, "PFP");
}

// vec2tp (vector_projection, mlegs_scalar_ops.f90:1308-1453)
// pre_dealias: the caller de-aliases psi and chi afterwards (advection_rhs).
// idel2 and chop act on each (m, kappa) column separately and zeroat1 leaves a
// zero slice at zero, so de-aliasing right after the projection gives the same
// result and lets the banded solve skip the de-aliased (all-zero) systems.
// the projection from ur = r vr, up = r vp, uz = vz already in PFP
// dl.on (with pre_dealias): the result is de-aliased by the projection
// itself, which computes columns m < pcut and slices outside zlo..zhi only;
// the transforms before it then skip the columns it does not read.
// ln_zero: psi.ln and chi.ln are set to zero at the end (advection_rhs's
// nl.ln = 0), by the closing zeroat1 launch
void vec2tp_pfp(mlx_ctx *c, mlx_field *ur, mlx_field *up, mlx_field *uz, mlx_field *psi, mlx_field *chi,
                bool pre_dealias, DealiasLimits dl = DealiasLimits(), bool ln_zero = false) {
  const KitDev &k = c->kit;
  if (!pre_dealias) dl = DealiasLimits();
  const int nc = dl.on ? dl.pcut : k.npchop;
  // psi's log term is set by k_vec2tp_up_ln and chi's by idel2 below
  // the projection reads columns m < npchop (nc) only
  // outputs in the kappa slices the projection skips are not written
  const int zlo = dl.on ? dl.zlo : 1, zhi = dl.on ? dl.zhi : 0;
  // ur's and uz's z transforms (the same lines) in one launch
  if (k.nz > 1) {
    double2 *const e2[2] = {ur->e, uz->e};
    fft_z_batch(c, e2, 2, -1, (double)k.nr * nc, 0, nc, 1, 0, zlo, zhi);
  }
  // up to FFF as trans does it (PFP -> FFP -> FFF), over columns m < nc.
  // With the de-aliasing limits the columns m >= nc of up are read by
  // nothing below (the z transforms, the radial transforms and the
  // projection take columns m < nc, calcat and the log step column 0), so
  // the radial transforms leave them unwritten.
  k_ln_column(c, up->e, up->sc, -1.0);
  PrecScope ps_up(c, MLX_STAGE_UP);
  leg_forward(c, up->e, c->scr[0], up->off[0], true, nc, !dl.on);
  swap_scr(c, up);
  if (k.nz > 1)
    fft_z(c, up->e, -1, active_pairs(k, up->off[0], nc), k.nrchop + up->off[0], nc, 1, 0, zlo, zhi);
  set_space(up, "FFF");
  double2 *inf = reinterpret_cast<double2 *>(c->dsmall + 2048);
  k_calcat(c, up->e, k.at1, k.nrchop + up->off[0], inf);
  leg_backward(c, up->e, c->scr[0], up->off[0], true, nc, !dl.on);
  swap_scr(c, up);
  k_vec2tp_up_ln(c, up->e, inf, psi->sc);
  {
    const int prec_m0 = stage_precision(c, MLX_STAGE_PROJ0);
    PrecScope ps_proj(c, MLX_STAGE_PROJ);
    vec2tp_project(c, ur->e, up->e, uz->e, psi->e, chi->e, 3, nc, zlo, zhi, prec_m0);
  }
  set_space(psi, "FFF");
  set_space(chi, "FFF");
  if (pre_dealias && !dl.on) {
    dealias(c, psi);
    dealias(c, chi);
  }
  // with dl on, the projection has written zeros outside the chopped set, and
  // a solve over offset-0 systems keeps them: chop would change nothing
  const bool chopped = dl.on && chi->off[0] == 0;
  if (dl.on) {
    // the projection left the de-aliased modes zero: their systems (zero
    // right-hand sides) are not visited
    to_fff(c, chi);
    k_helm_solve(c, chi->e, 0.0, 1, chi->off[0], dl.pcut, dl.zlo, dl.zhi);
    k_idel2_post(c, chi->e, chi->sc);
  } else {
    idel2_proln(c, chi);
  }
  psi->off[0] = 0;
  chi->off[0] = 0;
  if (!chopped) {
    chop(c, psi);
    chop(c, chi);
  }
  zeroat1_pair(c, psi, chi, ln_zero);
}

void vec2tp(mlx_ctx *c, const mlx_field *vr, const mlx_field *vp, const mlx_field *vz, mlx_field *psi,
            mlx_field *chi, bool pre_dealias = false) {
  need_kit(c);
  for (const mlx_field *f : {vr, vp, vz})
    if (!is_ppp(f)) MLX_THROW(MLX_E_LAYOUT, "vector_projection: vr, vp, vz must be in PPP");
  mlx_field *ur = tmp(c, 0), *up = tmp(c, 1), *uz = tmp(c, 2);
  // ur = r vr, up = r vp, uz = vz, each taken to PFP by an out-of-place phi
  // transform that applies the r weight on its input
  phi_forward_from(c, ur, vr, 1);
  phi_forward_from(c, up, vp, 1);
  phi_forward_from(c, uz, vz, 0);
  vec2tp_pfp(c, ur, up, uz, psi, chi, pre_dealias);
}

// advection_rhs (vortical_flow_3d.f90:353-395).  dealiased: psi and chi are
// already de-aliased (the stepper de-aliases them before svv_filter and
// zeroat1, which keep a zero mode zero), so the opening dealias is a no-op.
// Returns how v and w (v without the background flow) are left in the
// temporaries 4..9: 2 in FP32, 1 in FP64, 0 not at all (overwritten).
int advection_rhs(mlx_ctx *c, mlx_field *psi, mlx_field *chi, mlx_field *nlpsi, mlx_field *nlchi,
                  const mlx_field *uz, bool dealiased = false) {
  if (!is_fff(psi) || !is_fff(chi)) MLX_THROW(MLX_E_LAYOUT, "advection_rhs: psi and chi must be in FFF");
  if (!is_ppp(uz)) MLX_THROW(MLX_E_LAYOUT, "advection_rhs: uz must be in PPP");
  if (!dealiased) {
    dealias(c, psi);
    dealias(c, chi);
  }
  mlx_field *vr = tmp(c, 4), *vp = tmp(c, 5), *vz = tmp(c, 6);
  mlx_field *wr = tmp(c, 7), *wp = tmp(c, 8), *wz = tmp(c, 9);
  // psi and chi are de-aliased here and so is the result: the MFMA path
  // skips the columns and slices de-aliasing zeroes
  const DealiasLimits dl = dealias_limits(c);
  bool f32phys = false;
  {
    // tp2vec and tp2curlvec: both front passes in one launch, then the six
    // backward transforms
    const int ncols = dl.on ? dl.pcut : -1;
    const bool zskip = fft_z_fast(c->kit);
    k_tp2vec_front2(c, vr->e, vp->e, vz->e, vr->sc, vp->sc, vz->sc, wr->e, wp->e, wz->e, wr->sc, wp->sc, wz->sc,
                    psi->e, chi->e, psi->sc, chi->sc, chi->off[0], ncols, zskip ? dl.zlo : 1, zskip ? dl.zhi : 0);
    // the six physical fields are read only by the fused cross-product
    // transform below: with FP32 radial sums they are stored in FP32
    const bool cross = fft_phi_cross_ok(c->kit);
    // the six z transforms in one launch
    mlx_field *const six[6] = {vr, vp, vz, wr, wp, wz};
    tp2vec_back_z(c, six, 6, ncols, dl.zlo, dl.zhi);
    f32phys = tp2vec_back(c, vr, vp, vz, ncols, dl.zlo, dl.zhi, true, cross, true);
    if (tp2vec_back(c, wr, wp, wz, ncols, dl.zlo, dl.zhi, true, cross, true) != f32phys)
      MLX_THROW(MLX_E_ARG, "advection_rhs: velocity and vorticity stored at different precisions");
  }
  // de-aliased inside vec2tp before idel2; chop and zeroat1 keep zero modes
  // zero, so the closing dealias of the reference is already satisfied
  mlx_field *ur = tmp(c, 0), *up = tmp(c, 1), *uzp = tmp(c, 2);
  if (fft_phi_cross(c, ur->e, up->e, uzp->e, vr->e, vp->e, vz->e, wr->e, wp->e, wz->e, uz->e, ur->sc, up->sc,
                    uzp->sc, dl.on ? dl.pcut : -1, f32phys)) {
    // vz + uz, v x w, and the phi transforms of r (v x w)_r, r (v x w)_phi,
    // (v x w)_z in one pass: log terms zero (set by the kernel), no offsets,
    // as vecprod leaves them
    for (mlx_field *f : {ur, up, uzp}) {
      f->off[0] = f->off[1] = f->off[2] = 0;
      set_space(f, "PFP");
    }
    vec2tp_pfp(c, ur, up, uzp, nlpsi, nlchi, true, dl, true);
    return f32phys ? 2 : 1;
  }
  k_axpby(c, vz->e, vz->sc, 1.0, uz->e, nullptr, 1.0);
  vecprod(c, vr, vp, vz, wr, wp, wz);
  vec2tp(c, vr, vp, vz, nlpsi, nlchi, true);
  k_set_scalar(c, nlpsi->sc, 0.0);
  k_set_scalar(c, nlchi->sc, 0.0);
  return 0;
}

// flag_dev (optional): the non-finite check of the field, on the same pass
// dl (the time loop, right after de-aliasing f): the reduction skips the
// de-aliased modes, which are zero
// f2 (optional): a second field filtered in the same launches, gain gain2
void svv(mlx_ctx *c, mlx_field *f, double *gain_dev, int *flag_dev = nullptr, const DealiasLimits *dl = nullptr,
         mlx_field *f2 = nullptr, double *gain2 = nullptr) {
  if (!c->p.is_svv) return;
  if (!is_fff(f) || (f2 && !is_fff(f2))) MLX_THROW(MLX_E_LAYOUT, "svv_filter: scalar must be in FFF space");
  need_kit(c);
  const bool lim = dl && dl->on;
  k_svv(c, f->e, gain_dev, f2 ? f2->e : nullptr, gain2, flag_dev, lim ? dl->pcut : -1, lim ? dl->zlo : 1,
        lim ? dl->zhi : 0);
}

void check_finite(mlx_ctx *c, const mlx_field *a, const mlx_field *b, int *flag_dev) {
  k_nonfinite(c, a->e, b->e, flag_dev);
}

template <class F>
int guarded(mlx_ctx *c, F &&fn) {
  if (!c) return MLX_E_ARG;
  try {
    fn();
    return MLX_OK;
  } catch (const MlxError &e) {
    c->err = e.msg;
    return e.code;
  } catch (const std::exception &e) {
    c->err = e.what();
    return MLX_E_ARG;
  }
}

}  // namespace

extern "C" {

int mlx_ctx_create(const mlx_params *p, int device, void *stream, mlx_ctx **out) {
  if (!p || !out) return MLX_E_ARG;
  *out = nullptr;
  mlx_ctx *c = new mlx_ctx();
  c->p = *p;
  c->device = device;
  int rc = guarded(c, [&] {
    HIPCHK(hipSetDevice(device));
    // this device's address of the static scratch of the no-allocation
    // entry point (looked up once per device)
    noalloc_setup(device);
    if (stream) {
      c->stream = reinterpret_cast<hipStream_t>(stream);
    } else {
      HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      c->own_stream = true;
    }
    HIPCHK(hipMalloc(&c->dsmall, sizeof(double) * 4096));
    HIPCHK(hipMemsetAsync(c->dsmall, 0, sizeof(double) * 4096, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMalloc(&c->dpart, sizeof(double) * kPartMax));
    HIPCHK(hipHostMalloc(&c->hsmall, sizeof(double) * 4096));
  });
  if (rc != MLX_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return MLX_OK;
}

int mlx_ctx_destroy(mlx_ctx *c) {
  if (!c) return MLX_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (mlx_field *f : c->pool) field_free(f);
  c->pool.clear();
  lu_cache_free(c);
  hyp_cache_free(c);
  for (auto &v : c->probe.ev)
    for (hipEvent_t e : v) (void)hipEventDestroy(e);
  for (auto &kv : c->work)
    if (kv.second.d) (void)hipFree(kv.second.d);
  if (c->aux) (void)hipStreamSynchronize(c->aux);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->aux) (void)hipStreamDestroy(c->aux);
  if (c->v2t_part) (void)hipFree(c->v2t_part);
  for (double2 *s : c->scr)
    if (s) (void)hipFree(s);
  kit_free(c->kit);
  if (c->dsmall) (void)hipFree(c->dsmall);
  if (c->dpart) (void)hipFree(c->dpart);
  if (c->hsmall) (void)hipHostFree(c->hsmall);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MLX_OK;
}

const char *mlx_last_error(mlx_ctx *c) { return c ? c->err.c_str() : "null context"; }

int mlx_sync(mlx_ctx *c) {
  return guarded(c, [&] { HIPCHK(hipStreamSynchronize(c->stream)); });
}

int mlx_kit_init(mlx_ctx *c) {
  return guarded(c, [&] {
    const mlx_params &p = c->p;
    auto ok235 = [](int n) {
      while (n % 2 == 0 || n % 3 == 0 || n % 5 == 0) {
        if (n % 2 == 0) n /= 2;
        if (n % 3 == 0) n /= 3;
        if (n % 5 == 0) n /= 5;
      }
      return n == 1;
    };
    // tfm_kit_init checks (mlegs_spectfm_init.f90:35-62)
    if (!(p.nr > 0 && p.nr % 2 == 0)) MLX_THROW(MLX_E_ARG, "tfm_kit_init: nr must be even");
    if (!(p.np > 0 && (p.np == 1 || p.np % 2 == 0))) MLX_THROW(MLX_E_ARG, "tfm_kit_init: np must be even");
    if (p.np != 1 && !ok235(p.np)) MLX_THROW(MLX_E_ARG, "tfm_kit_init: np must only have factors of 2, 3 and 5");
    if (!(p.nz > 0 && (p.nz == 1 || p.nz % 2 == 0))) MLX_THROW(MLX_E_ARG, "tfm_kit_init: nz must be even");
    if (p.nz != 1 && !ok235(p.nz)) MLX_THROW(MLX_E_ARG, "tfm_kit_init: nz must only have factors of 2, 3 and 5");
    if (p.nrchop > p.nr) MLX_THROW(MLX_E_ARG, "tfm_kit_init: nrchop must be smaller than or equal to nr");
    if (p.np != 1 && p.npchop * 2 > p.np + 2) MLX_THROW(MLX_E_ARG, "tfm_kit_init: npchop <= np/2 + 1 must be satisfied");
    if (p.nz != 1 && p.nzchop * 2 > p.nz + 2) MLX_THROW(MLX_E_ARG, "tfm_kit_init: nzchop <= nz/2 + 1 must be satisfied");
    if (p.nrchop < 1 || p.npchop < 1 || p.nzchop < 1) MLX_THROW(MLX_E_ARG, "tfm_kit_init: chops must be positive");
    // np = 1 (the reference's 1D runs on its 3D kit): one azimuthal column
    if (p.npchop > p.np / 2 + 1) MLX_THROW(MLX_E_ARG, "chop_index: chopping in p too large");
    HIPCHK(hipSetDevice(c->device));
    kit_build(c);
    const KitDev &k = c->kit;
    c->nelem = (size_t)k.nrdim * k.npdim * k.nzdim;
    // a buffer sized by an earlier kit is allocated again on first use
    if (c->v2t_part) {
      HIPCHK(hipStreamSynchronize(c->stream));
      if (c->aux) HIPCHK(hipStreamSynchronize(c->aux));
      (void)hipFree(c->v2t_part);
      c->v2t_part = nullptr;
    }
    for (double2 *&s : c->scr) {
      if (s) (void)hipFree(s);
      if (hipMalloc(&s, sizeof(double2) * c->nelem) != hipSuccess) MLX_THROW(MLX_E_ALLOC, "scratch allocation failed");
      HIPCHK(hipMemsetAsync(s, 0, sizeof(double2) * c->nelem, c->stream));
    }
  });
}

int mlx_kit_dims(mlx_ctx *c, int32_t *o) {
  return guarded(c, [&] {
    need_kit(c);
    const KitDev &k = c->kit;
    int32_t v[9] = {k.nrdim_ref, k.npdim, k.nzdim, k.nrh, k.ne, k.chopzl, k.chopzu, k.nph, k.npchop};
    std::memcpy(o, v, sizeof(v));
  });
}

int mlx_kit_get(mlx_ctx *c, const char *which, double *host, int64_t n) {
  return guarded(c, [&] {
    need_kit(c);
    const KitDev &k = c->kit;
    const double *src = nullptr;
    size_t cnt = 0;
    std::string w(which ? which : "");
    if (w == "x") src = k.x, cnt = k.nr;
    else if (w == "w") src = k.w, cnt = k.nr;
    else if (w == "ln") src = k.lnx, cnt = k.nr;
    else if (w == "r") src = k.r, cnt = k.nr;
    else if (w == "ak") src = k.ak, cnt = k.nz;
    else if (w == "lognorm") src = k.lognorm, cnt = (size_t)k.ne * k.npchop;
    else if (w == "pf") src = k.pf, cnt = (size_t)k.nrh * k.ne * k.npchop;
    else if (w == "at0") src = k.at0, cnt = k.nrchop;
    else if (w == "at1") src = k.at1, cnt = k.nrchop;
    else MLX_THROW(MLX_E_ARG, "mlx_kit_get: unknown array " + w);
    if ((size_t)n < cnt) MLX_THROW(MLX_E_ARG, "mlx_kit_get: host buffer too small");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(host, src, sizeof(double) * cnt, hipMemcpyDeviceToHost));
  });
}

int mlx_set_precision(mlx_ctx *c, int prec) {
  return guarded(c, [&] {
    if (prec != MLX_PREC_FP64 && prec != MLX_PREC_FP32 && prec != MLX_PREC_FP64_MFMA && prec != MLX_PREC_BF16X3)
      MLX_THROW(MLX_E_ARG, "mlx_set_precision: unknown precision");
    c->prec = c->prec_ctx = prec;
  });
}

int mlx_set_stage_precision(mlx_ctx *c, int stage, int prec) {
  return guarded(c, [&] {
    if (stage < 0 || stage >= MLX_STAGE_COUNT) MLX_THROW(MLX_E_ARG, "mlx_set_stage_precision: unknown stage");
    if (prec != -1 && prec != MLX_PREC_FP64 && prec != MLX_PREC_FP32 && prec != MLX_PREC_FP64_MFMA &&
        prec != MLX_PREC_BF16X3)
      MLX_THROW(MLX_E_ARG, "mlx_set_stage_precision: unknown precision");
    c->stage_prec[stage] = prec;
  });
}

int mlx_field_create(mlx_ctx *c, mlx_field **out) {
  return guarded(c, [&] {
    need_kit(c);
    *out = field_new(c);
  });
}

int mlx_field_destroy(mlx_ctx *c, mlx_field *f) {
  return guarded(c, [&] {
    HIPCHK(hipStreamSynchronize(c->stream));
    field_free(f);
  });
}

int mlx_field_copy(mlx_ctx *c, mlx_field *d, const mlx_field *s) {
  return guarded(c, [&] { fcopy(c, d, s); });
}

// chop_index's radial check (mlegs_scalar_ops.f90:2053), made when an offset
// is set: nrchop + offset rows must fit the field's column
static void check_nroff(mlx_ctx *c, int32_t a) {
  need_kit(c);
  if (c->kit.nrchop + a > c->kit.nrdim_ref) MLX_THROW(MLX_E_ARG, "chop_index: chopping in r too large");
}

static void check_dims(mlx_ctx *c, int64_t a, int64_t b, int64_t d) {
  const KitDev &k = c->kit;
  if (a != k.nrdim_ref || b != k.npdim || d != k.nzdim)
    MLX_THROW(MLX_E_LAYOUT, "scalar dimension incompatible with the dimension defined in the kit");
}

int mlx_field_upload(mlx_ctx *c, mlx_field *f, const double *host, int64_t nrdim, int64_t npdim, int64_t nzdim,
                     const char *space, double ln, const int32_t *offsets) {
  return guarded(c, [&] {
    need_kit(c);
    check_dims(c, nrdim, npdim, nzdim);
    if (space_id(space) < 0) MLX_THROW(MLX_E_ARG, "upload: space must be PPP, PFP, FFP or FFF");
    // host columns of nrdim_ref rows into device columns of nrdim rows (the
    // padding rows stay zero)
    const KitDev &k = c->kit;
    if (k.nrdim != k.nrdim_ref) HIPCHK(hipMemsetAsync(f->e, 0, sizeof(double2) * c->nelem, c->stream));
    HIPCHK(hipMemcpy2DAsync(f->e, sizeof(double2) * k.nrdim, host, sizeof(double2) * k.nrdim_ref,
                            sizeof(double2) * k.nrdim_ref, (size_t)k.npdim * k.nzdim, hipMemcpyHostToDevice,
                            c->stream));
    c->hsmall[0] = ln;
    HIPCHK(hipMemcpyAsync(f->sc, c->hsmall, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (offsets && (offsets[1] != 0 || offsets[2] != 0))
      MLX_THROW(MLX_E_ARG, "upload: only radial chopping offsets are supported on this path");
    if (offsets) check_nroff(c, offsets[0]);
    set_space(f, space);
    for (int i = 0; i < 3; ++i) f->off[i] = offsets ? offsets[i] : 0;
  });
}

int mlx_field_download(mlx_ctx *c, const mlx_field *f, double *host, int64_t nrdim, int64_t npdim, int64_t nzdim,
                       char *space4, double *ln, int32_t *offsets) {
  return guarded(c, [&] {
    need_kit(c);
    check_dims(c, nrdim, npdim, nzdim);
    const KitDev &k = c->kit;
    if (host)
      HIPCHK(hipMemcpy2DAsync(host, sizeof(double2) * k.nrdim_ref, f->e, sizeof(double2) * k.nrdim,
                              sizeof(double2) * k.nrdim_ref, (size_t)k.npdim * k.nzdim, hipMemcpyDeviceToHost,
                              c->stream));
    HIPCHK(hipMemcpyAsync(c->hsmall, f->sc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (ln) *ln = c->hsmall[0];
    if (space4) std::memcpy(space4, f->space, 4);
    if (offsets)
      for (int i = 0; i < 3; ++i) offsets[i] = f->off[i];
  });
}

int mlx_field_set_meta(mlx_ctx *c, mlx_field *f, const char *space, double ln, const int32_t *offsets) {
  return guarded(c, [&] {
    if (space_id(space) < 0) MLX_THROW(MLX_E_ARG, "set_meta: space must be PPP, PFP, FFP or FFF");
    if (offsets && (offsets[1] != 0 || offsets[2] != 0))
      MLX_THROW(MLX_E_ARG, "set_meta: only radial chopping offsets are supported on this path");
    if (offsets) check_nroff(c, offsets[0]);
    set_space(f, space);
    k_set_scalar(c, f->sc, ln);
    if (offsets)
      for (int i = 0; i < 3; ++i) f->off[i] = offsets[i];
  });
}

int mlx_field_get_meta(mlx_ctx *c, const mlx_field *f, char *space4, double *ln, int32_t *offsets) {
  return mlx_field_download(c, f, nullptr, c->kit.nrdim_ref, c->kit.npdim, c->kit.nzdim, space4, ln, offsets);
}

int mlx_chop_offset(mlx_ctx *c, mlx_field *f, int32_t a, int32_t b, int32_t d) {
  return guarded(c, [&] {
    if (b != 0 || d != 0) MLX_THROW(MLX_E_ARG, "chop_offset: only radial offsets are supported on this path");
    check_nroff(c, a);
    f->off[0] = a;
    f->off[1] = b;
    f->off[2] = d;
  });
}

int mlx_field_axpby(mlx_ctx *c, mlx_field *d, double a, const mlx_field *s, double b) {
  return guarded(c, [&] { k_axpby(c, d->e, d->sc, a, s->e, s->sc, b); });
}

int mlx_trans(mlx_ctx *c, mlx_field *f, const char *space) {
  return guarded(c, [&] { trans(c, f, space); });
}
int mlx_chop(mlx_ctx *c, mlx_field *f) {
  return guarded(c, [&] {
    need_kit(c);
    chop(c, f);
  });
}
int mlx_dealias(mlx_ctx *c, mlx_field *f) {
  return guarded(c, [&] { dealias(c, f); });
}

int mlx_svv_filter(mlx_ctx *c, mlx_field *f, double *gain) {
  return guarded(c, [&] {
    if (!c->p.is_svv) return;
    double *g = c->dsmall + 4000 + 8;
    c->hsmall[0] = *gain;
    HIPCHK(hipMemcpyAsync(g, c->hsmall, sizeof(double), hipMemcpyHostToDevice, c->stream));
    svv(c, f, g);
    HIPCHK(hipMemcpyAsync(c->hsmall, g, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *gain = c->hsmall[0];
  });
}

static void calcat_host(mlx_ctx *c, const mlx_field *f, const double *tbl, double *out) {
  need_kit(c);
  const mlx_field *src = f;
  if (!is_fff(f)) {
    mlx_field *st = tmp(c, 11);
    fcopy(c, st, f);
    trans(c, st, "FFF");
    src = st;
  }
  double2 *d = reinterpret_cast<double2 *>(c->dsmall + 2048);
  k_calcat(c, src->e, tbl, c->kit.nrchop + src->off[0], d);
  HIPCHK(hipMemcpyAsync(out, d, sizeof(double2) * c->kit.nzdim, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
}

int mlx_calcat0(mlx_ctx *c, const mlx_field *f, double *out) {
  return guarded(c, [&] { calcat_host(c, f, c->kit.at0, out); });
}
int mlx_calcat1(mlx_ctx *c, const mlx_field *f, double *out) {
  return guarded(c, [&] { calcat_host(c, f, c->kit.at1, out); });
}
int mlx_zeroat1(mlx_ctx *c, mlx_field *f) {
  return guarded(c, [&] {
    need_kit(c);
    zeroat1(c, f);
  });
}

int mlx_delsqp(mlx_ctx *c, mlx_field *f) {
  return guarded(c, [&] {
    need_kit(c);
    to_fff(c, f);
    k_delsqp(c, f->e, f->sc, f->off[0], 0);
  });
}
int mlx_idelsqp(mlx_ctx *c, mlx_field *f) {
  return guarded(c, [&] {
    need_kit(c);
    to_fff(c, f);
    k_delsqp(c, f->e, f->sc, f->off[0], 1);
  });
}
int mlx_xxdx(mlx_ctx *c, mlx_field *f) {
  return guarded(c, [&] { bandmul(c, f, 2); });
}
int mlx_del2h(mlx_ctx *c, mlx_field *f) {
  return guarded(c, [&] { bandmul(c, f, 1); });
}
int mlx_del2(mlx_ctx *c, mlx_field *f) {
  return guarded(c, [&] { bandmul(c, f, 0); });
}
int mlx_fftreat(mlx_ctx *c, mlx_field *f) {
  return guarded(c, [&] { fftreat(c, f); });
}
int mlx_idel2(mlx_ctx *c, mlx_field *f) {
  return guarded(c, [&] { idel2_proln(c, f); });
}
int mlx_idel2_preln(mlx_ctx *c, mlx_field *f, double preln) {
  return guarded(c, [&] {
    need_kit(c);
    to_fff(c, f);
    k_idel2_preln00(c, f->e, f->sc, preln, f->off[0]);
    k_helm_solve(c, f->e, 0.0, 2, f->off[0]);
  });
}
int mlx_ihelm(mlx_ctx *c, mlx_field *f, double alpha) {
  return guarded(c, [&] { ihelm(c, f, alpha); });
}

int mlx_helmp(mlx_ctx *c, mlx_field *f, int power, double alpha, double beta) {
  return guarded(c, [&] { helmp(c, f, power, alpha, beta); });
}
int mlx_ihelmp(mlx_ctx *c, mlx_field *f, int power, double alpha, double beta) {
  return guarded(c, [&] {
    clear_sing(c);
    ihelmp(c, f, power, alpha, beta);
    check_sing(c);
  });
}

int mlx_febe(mlx_ctx *c, mlx_field *s, const mlx_field *nl, double dt) {
  return guarded(c, [&] {
    need_kit(c);
    if (c->p.hyperpow != 0) clear_sing(c);
    febe(c, s, nl, dt);
    if (c->p.hyperpow != 0) check_sing(c);
  });
}
int mlx_fefe(mlx_ctx *c, mlx_field *s, const mlx_field *nl, double dt) {
  return guarded(c, [&] {
    need_kit(c);
    fefe(c, s, nl, dt);
  });
}
int mlx_abab(mlx_ctx *c, mlx_field *s, mlx_field *s_p, const mlx_field *nl, mlx_field *nl_p, double dt,
             int32_t is_2nd_svis_p) {
  return guarded(c, [&] {
    need_kit(c);
    abab(c, s, s_p, nl, nl_p, dt, is_2nd_svis_p != 0);
  });
}
int mlx_abcn(mlx_ctx *c, mlx_field *s, mlx_field *s_p, const mlx_field *nl, mlx_field *nl_p, double dt) {
  return guarded(c, [&] {
    need_kit(c);
    if (c->p.hyperpow != 0) clear_sing(c);
    abcn(c, s, s_p, const_cast<mlx_field *>(nl), nl_p, dt, true, false);
    if (c->p.hyperpow != 0) check_sing(c);
  });
}

int mlx_vecprod(mlx_ctx *c, mlx_field *vr, mlx_field *vp, mlx_field *vz, const mlx_field *ur, const mlx_field *up,
                const mlx_field *uz) {
  return guarded(c, [&] {
    need_kit(c);
    vecprod(c, vr, vp, vz, ur, up, uz);
  });
}

int mlx_trans_adjoint(mlx_ctx *c, mlx_field *g, const char *from) {
  return guarded(c, [&] { trans_adjoint(c, g, from); });
}

int mlx_vecprod_adjoint(mlx_ctx *c, mlx_field *ar, mlx_field *ap, mlx_field *az, const mlx_field *vr,
                        const mlx_field *vp, const mlx_field *vz, const mlx_field *ur, const mlx_field *up,
                        const mlx_field *uz, mlx_field *ubr, mlx_field *ubp, mlx_field *ubz) {
  return guarded(c, [&] {
    need_kit(c);
    for (const mlx_field *f : {(const mlx_field *)ar, (const mlx_field *)ap, (const mlx_field *)az, vr, vp, vz, ur,
                               up, uz})
      if (!is_ppp(f)) MLX_THROW(MLX_E_LAYOUT, "vector_product adjoint: all inputs must be in PPP");
    k_vecprod_adj(c, ar->e, ap->e, az->e, vr->e, vp->e, vz->e, ur->e, up->e, uz->e, ubr->e, ubp->e, ubz->e);
    for (mlx_field *f : {ar, ap, az, ubr, ubp, ubz}) {
      k_set_scalar(c, f->sc, 0.0);
      f->off[0] = f->off[1] = f->off[2] = 0;
      set_space(f, "PPP");
    }
  });
}

int mlx_abcn_adjoint(mlx_ctx *c, mlx_field *sbar, mlx_field *nlbar, mlx_field *nlpbar, double dt) {
  return guarded(c, [&] { abcn_adjoint(c, sbar, nlbar, nlpbar, dt); });
}

int mlx_febe_adjoint(mlx_ctx *c, mlx_field *sbar, mlx_field *nlbar, double dt) {
  return guarded(c, [&] { febe_adjoint(c, sbar, nlbar, dt); });
}

int mlx_zeroat1_adjoint(mlx_ctx *c, mlx_field *f) {
  return guarded(c, [&] {
    need_kit(c);
    if (!is_fff(f)) MLX_THROW(MLX_E_LAYOUT, "zeroat1 adjoint: the field must be in FFF");
    adj_no_offsets(f, "zeroat1 adjoint");
    k_zeroat1_t(c, f->e);
  });
}

int mlx_tp2vec_adjoint(mlx_ctx *c, mlx_field *vrb, mlx_field *vpb, mlx_field *vzb, mlx_field *psib,
                       mlx_field *chib) {
  return guarded(c, [&] {
    Fp64Scope fp64(c);
    tp2vec_adjoint(c, vrb, vpb, vzb, psib, chib);
  });
}

int mlx_tp2curlvec_adjoint(mlx_ctx *c, mlx_field *wrb, mlx_field *wpb, mlx_field *wzb, mlx_field *psib,
                           mlx_field *chib) {
  return guarded(c, [&] {
    need_kit(c);
    Fp64Scope fp64(c);
    tp2curlvec_adjoint(c, wrb, wpb, wzb, psib, chib);
  });
}

int mlx_vec2tp_adjoint(mlx_ctx *c, mlx_field *psib, mlx_field *chib, mlx_field *vrb, mlx_field *vpb,
                       mlx_field *vzb) {
  return guarded(c, [&] {
    Fp64Scope fp64(c);
    vec2tp_adjoint(c, psib, chib, vrb, vpb, vzb);
  });
}

int mlx_advection_adjoint(mlx_ctx *c, mlx_field *psi, mlx_field *chi, const mlx_field *uz, mlx_field *nlpsib,
                          mlx_field *nlchib, mlx_field *psib, mlx_field *chib) {
  return guarded(c, [&] { advection_adjoint(c, psi, chi, uz, nlpsib, nlchib, psib, chib); });
}

// one reverse step of the vortex run (the map of mlx_vortex_step: s' =
// zeroat1(dealias(abcn(s, N, N_prev))), N' = N(s'), N_prev' = N): from the
// cotangents (sb, nb, npb) of the state after the step, whose psi and chi
// are p1, c1, to those of the state before it, in place
int mlx_vortex_reverse_step(mlx_ctx *c, mlx_field *p1, mlx_field *c1, const mlx_field *uz, mlx_field *sbpsi,
                            mlx_field *sbchi, mlx_field *nbpsi, mlx_field *nbchi, mlx_field *npbpsi,
                            mlx_field *npbchi, double dt, mlx_field **vw, int32_t vw_fmt) {
  return guarded(c, [&] {
    if (vw && vw_fmt)
      for (int q = 0; q < 6; ++q)
        if (!vw[q]) MLX_THROW(MLX_E_ARG, "mlx_vortex_reverse_step: null v or w field");
    mlx_field *sb[2] = {sbpsi, sbchi}, *nb[2] = {nbpsi, nbchi}, *npb[2] = {npbpsi, npbchi};
    mlx_field *g[2] = {tmp(c, 38), tmp(c, 39)};
    advection_adjoint(c, p1, c1, uz, nb[0], nb[1], g[0], g[1], vw, vw ? vw_fmt : 0);
    for (int i = 0; i < 2; ++i) {
      // sb + g through zeroat1^T and dealias, and onto abcn's support (the
      // three projections commute with zeroat1^T)
      k_adj_sum_mask(c, sb[i]->e, sb[i]->sc, g[i]->e, g[i]->sc);
      k_zeroat1_t(c, sb[i]->e);
    }
    // abcn^T of both fields: nb receives the cotangent of N, npb's old value
    // is the cotangent of the history N_prev' = N (added), the new npb that
    // of N_prev
    mlx_field *b[2] = {tmp(c, 40), tmp(c, 41)};
    abcn_adjoint(c, sb, nb, b, 2, dt, true);
    for (int i = 0; i < 2; ++i) {
      k_axpby(c, nb[i]->e, nb[i]->sc, 1.0, npb[i]->e, npb[i]->sc, 1.0);
      fmove(npb[i], b[i]);
    }
  });
}

int mlx_energy_grad(mlx_ctx *c, mlx_field *vr, mlx_field *vp, mlx_field *vz) {
  return guarded(c, [&] {
    need_kit(c);
    for (const mlx_field *f : {vr, vp, vz})
      if (!is_ppp(f)) MLX_THROW(MLX_E_LAYOUT, "energy gradient: velocity components must be in PPP");
    k_energy_grad(c, vr->e, vp->e, vz->e);
  });
}

int mlx_adjoint_project(mlx_ctx *c, mlx_field *f) {
  return guarded(c, [&] {
    need_kit(c);
    if (!is_fff(f)) MLX_THROW(MLX_E_LAYOUT, "adjoint project: the field must be in FFF");
    adj_no_offsets(f, "adjoint project");
    k_adj_project(c, f->e);
  });
}

int mlx_vec2tp(mlx_ctx *c, const mlx_field *vr, const mlx_field *vp, const mlx_field *vz, mlx_field *psi,
               mlx_field *chi) {
  return guarded(c, [&] {
    if (psi == chi || psi == vr || psi == vp || psi == vz || chi == vr || chi == vp || chi == vz)
      MLX_THROW(MLX_E_ARG, "vec2tp: outputs must be distinct from the inputs");
    vec2tp(c, vr, vp, vz, psi, chi);
  });
}

int mlx_tp2vec(mlx_ctx *c, const mlx_field *psi, const mlx_field *chi, mlx_field *vr, mlx_field *vp,
               mlx_field *vz) {
  return guarded(c, [&] {
    const mlx_field *ins[2] = {psi, chi};
    for (const mlx_field *o : {(const mlx_field *)vr, (const mlx_field *)vp, (const mlx_field *)vz})
      if (o == ins[0] || o == ins[1]) MLX_THROW(MLX_E_ARG, "tp2vec: outputs must be distinct from the inputs");
    tp2vec(c, psi, chi, vr, vp, vz);
  });
}

int mlx_tp2curlvec(mlx_ctx *c, const mlx_field *psi, const mlx_field *chi, mlx_field *wr, mlx_field *wp,
                   mlx_field *wz) {
  return guarded(c, [&] {
    const mlx_field *ins[2] = {psi, chi};
    for (const mlx_field *o : {(const mlx_field *)wr, (const mlx_field *)wp, (const mlx_field *)wz})
      if (o == ins[0] || o == ins[1]) MLX_THROW(MLX_E_ARG, "tp2curlvec: outputs must be distinct from the inputs");
    tp2curlvec(c, psi, chi, wr, wp, wz);
  });
}

int mlx_vorticity_magnitude(mlx_ctx *c, const mlx_field *psi, const mlx_field *chi, mlx_field *out) {
  return guarded(c, [&] {
    need_kit(c);
    if (!is_fff(psi) || !is_fff(chi)) MLX_THROW(MLX_E_LAYOUT, "save_vort_mag: psi and chi must be in FFF space");
    mlx_field *wr = tmp(c, 7), *wp = tmp(c, 8), *wz = tmp(c, 9);
    if (out == psi || out == chi || out == wr || out == wp || out == wz)
      MLX_THROW(MLX_E_ARG, "vorticity_magnitude: the output must be distinct from psi and chi");
    tp2curlvec(c, psi, chi, wr, wp, wz);
    k_vormag(c, out->e, wr->e, wp->e, wz->e);
    k_set_scalar(c, out->sc, 0.0);
    set_space(out, "PPP");
    for (int i = 0; i < 3; ++i) out->off[i] = 0;
  });
}

int mlx_advection_rhs(mlx_ctx *c, mlx_field *psi, mlx_field *chi, mlx_field *nlpsi, mlx_field *nlchi,
                      const mlx_field *uz) {
  return guarded(c, [&] {
    need_kit(c);
    advection_rhs(c, psi, chi, nlpsi, nlchi, uz);
  });
}

int mlx_check_stability(mlx_ctx *c, const mlx_field *psi, const mlx_field *chi) {
  return guarded(c, [&] {
    int *flag = reinterpret_cast<int *>(c->dsmall + 4000 + 16);
    HIPCHK(hipMemsetAsync(flag, 0, sizeof(int), c->stream));
    check_finite(c, psi, chi, flag);
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (h) MLX_THROW(MLX_E_NONFINITE, "ERROR: non-finite vortex state");
  });
}

int mlx_vortex_create(mlx_ctx *c, const mlx_field *psi, const mlx_field *chi, const mlx_field *uz,
                      mlx_vortex **out) {
  return guarded(c, [&] {
    need_kit(c);
    if (!is_fff(psi) || !is_fff(chi) || !is_ppp(uz))
      MLX_THROW(MLX_E_LAYOUT, "vortex: psi and chi must be in FFF and uz in PPP");
    mlx_vortex *v = new mlx_vortex();
    try {
      for (int i = 0; i < 9; ++i) v->f[i] = field_new(c);
      fcopy(c, v->f[0], psi);
      fcopy(c, v->f[1], chi);
      for (int i = 2; i < 8; ++i) set_space(v->f[i], "FFF");
      fcopy(c, v->f[8], uz);
      HIPCHK(hipMalloc(&v->gains, sizeof(double) * 2));
      HIPCHK(hipMalloc(&v->flag, sizeof(int)));
      HIPCHK(hipMemsetAsync(v->gains, 0, sizeof(double) * 2, c->stream));
      HIPCHK(hipMemsetAsync(v->flag, 0, sizeof(int), c->stream));
    } catch (...) {
      for (mlx_field *f : v->f) field_free(f);
      delete v;
      throw;
    }
    *out = v;
  });
}

int mlx_vortex_destroy(mlx_ctx *c, mlx_vortex *v) {
  return guarded(c, [&] {
    if (!v) return;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (mlx_field *f : v->f) field_free(f);
    if (v->gains) (void)hipFree(v->gains);
    if (v->flag) (void)hipFree(v->flag);
    delete v;
  });
}

// vortical_flow_3d.f90:116-147
int mlx_vortex_bootstrap(mlx_ctx *c, mlx_vortex *v) {
  return guarded(c, [&] {
    need_kit(c);
    mlx_field *psi = v->f[0], *chi = v->f[1], *nlpsi = v->f[2], *nlchi = v->f[3];
    mlx_field *psi_prev = v->f[4], *chi_prev = v->f[5], *nlpsi_prev = v->f[6], *nlchi_prev = v->f[7];
    mlx_field *uz = v->f[8];
    double dt = c->p.dt;
    if (c->p.hyperpow != 0) clear_sing(c);
    advection_rhs(c, psi, chi, nlpsi, nlchi, uz);
    fcopy(c, psi_prev, psi);
    fcopy(c, nlpsi_prev, nlpsi);
    fcopy(c, chi_prev, chi);
    fcopy(c, nlchi_prev, nlchi);
    mlx_field *psi_rich = tmp(c, 12), *chi_rich = tmp(c, 13);
    fcopy(c, psi_rich, psi);
    fcopy(c, chi_rich, chi);
    advection_rhs(c, psi_rich, chi_rich, nlpsi, nlchi, uz);
    febe(c, psi_rich, nlpsi, dt);
    febe(c, chi_rich, nlchi, dt);
    for (int i = 0; i < 2; ++i) {
      advection_rhs(c, psi, chi, nlpsi, nlchi, uz);
      febe(c, psi, nlpsi, dt / 2.0);
      febe(c, chi, nlchi, dt / 2.0);
    }
    // psi%e = 2 psi%e - psi_rich%e (ln untouched)
    k_axpby(c, psi->e, nullptr, 2.0 / 1.0, psi_rich->e, nullptr, -1.0 / 1.0);
    k_axpby(c, chi->e, nullptr, 2.0 / 1.0, chi_rich->e, nullptr, -1.0 / 1.0);
    dealias(c, psi);
    dealias(c, chi);
    svv(c, psi, v->gains);
    svv(c, chi, v->gains + 1);
    zeroat1(c, psi);
    zeroat1(c, chi);
    advection_rhs(c, psi, chi, nlpsi, nlchi, uz, true);
    v->prev_stale = false;
    if (c->p.hyperpow != 0) check_sing(c);
  });
}

// vortical_flow_3d.f90:160-181.  rec (optional): 2 nsteps fields that
// receive psi and chi after each step; rec_vw (optional): 6 nsteps fields
// that receive v (without the background flow) and w at them, where the
// step leaves them (*vw_fmt: 2 FP32, 1 FP64, 0 not recorded)
void vortex_steps(mlx_ctx *c, mlx_vortex *v, int32_t nsteps, double dt, int32_t check, mlx_field **rec,
                  mlx_field **rec_vw = nullptr, int *vw_fmt = nullptr) {
  {
    need_kit(c);
    mlx_field *psi = v->f[0], *chi = v->f[1], *nlpsi = v->f[2], *nlchi = v->f[3];
    mlx_field *psi_prev = v->f[4], *chi_prev = v->f[5], *nlpsi_prev = v->f[6], *nlchi_prev = v->f[7];
    mlx_field *uz = v->f[8];
    const DealiasLimits dl = dealias_limits(c);
    if (c->p.hyperpow != 0 && nsteps > 0) clear_sing(c);
    for (int it = 0; it < nsteps; ++it) {
      abcn_pair(c, psi, psi_prev, nlpsi, nlpsi_prev, chi, chi_prev, nlchi, nlchi_prev, dt, dl);
      v->prev_stale = true;
      // With the de-aliasing restrictions the right-hand side and the solve
      // leave the de-aliased modes untouched, and the step's other inputs
      // (the nonlinear terms) are zero there, so after the first step of a
      // call psi and chi stay de-aliased and the dealias pass is a no-op.
      // The first step keeps it: psi and chi come from the caller.
      if (!dl.on || it == 0) {
        dealias(c, psi);
        dealias(c, chi);
      }
      // with SVV on, its reduction pass also checks psi and chi for
      // non-finite entries: the filter and zeroat1 keep a finite field
      // finite and a non-finite one non-finite, so the flag is the one the
      // check after the step would raise
      svv(c, psi, v->gains, v->flag, &dl, chi, v->gains + 1);
      zeroat1_pair(c, psi, chi);
      const int fmt = advection_rhs(c, psi, chi, nlpsi, nlchi, uz, true);
      if (!c->p.is_svv) check_finite(c, psi, chi, v->flag);
      // psi and chi are needed only where v and w are not kept
      if (rec && !(rec_vw && fmt)) {
        fcopy(c, rec[2 * it], psi);
        fcopy(c, rec[2 * it + 1], chi);
      }
      if (rec_vw) {
        if (vw_fmt) *vw_fmt = fmt;
        // the temporaries' buffers go to the record (the record's come back
        // as the next step's temporaries): no copy
        for (int q = 0; fmt && q < 6; ++q) fmove(rec_vw[6 * it + q], tmp(c, 4 + q));
      }
    }
    if (check) {
      int h = 0;
      HIPCHK(hipMemcpyAsync(&h, v->flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      if (h) MLX_THROW(MLX_E_NONFINITE, "ERROR: non-finite vortex state");
      if (c->p.hyperpow != 0 && nsteps > 0) check_sing(c);
    }
  }
}

int mlx_vortex_step(mlx_ctx *c, mlx_vortex *v, int32_t nsteps, double dt, int32_t check) {
  return guarded(c, [&] { vortex_steps(c, v, nsteps, dt, check, nullptr); });
}

int mlx_vortex_step_record(mlx_ctx *c, mlx_vortex *v, int32_t nsteps, double dt, mlx_field **rec,
                           mlx_field **rec_vw, int32_t *vw_fmt) {
  return guarded(c, [&] {
    if (!rec && nsteps > 0) MLX_THROW(MLX_E_ARG, "mlx_vortex_step_record: no fields to record into");
    for (int i = 0; i < 2 * nsteps; ++i)
      if (!rec[i]) MLX_THROW(MLX_E_ARG, "mlx_vortex_step_record: null field");
    for (int i = 0; rec_vw && i < 6 * nsteps; ++i)
      if (!rec_vw[i]) MLX_THROW(MLX_E_ARG, "mlx_vortex_step_record: null v or w field");
    int fmt = 0;
    vortex_steps(c, v, nsteps, dt, 0, rec, rec_vw, &fmt);
    if (vw_fmt) *vw_fmt = rec_vw ? fmt : 0;
  });
}

int mlx_vortex_field(mlx_ctx *c, mlx_vortex *v, int32_t which, mlx_field **out) {
  return guarded(c, [&] {
    if (which < 0 || which > 8) MLX_THROW(MLX_E_ARG, "mlx_vortex_field: index out of range");
    if (v->prev_stale) {
      // psi_prev <- psi and chi_prev <- chi after abcn (mlegs_scalar_ops.f90:1256)
      fcopy(c, v->f[4], v->f[0]);
      fcopy(c, v->f[5], v->f[1]);
      v->prev_stale = false;
    }
    *out = v->f[which];
  });
}

int mlx_vortex_gains(mlx_ctx *c, mlx_vortex *v, double *gp, double *gc) {
  return guarded(c, [&] {
    HIPCHK(hipMemcpyAsync(c->hsmall, v->gains, sizeof(double) * 2, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (gp) *gp = c->hsmall[0];
    if (gc) *gc = c->hsmall[1];
  });
}

int mlx_energy(mlx_ctx *c, const mlx_field *vr, const mlx_field *vp, const mlx_field *vz, double *out) {
  return guarded(c, [&] {
    need_kit(c);
    for (const mlx_field *f : {vr, vp, vz})
      if (!is_ppp(f)) MLX_THROW(MLX_E_LAYOUT, "energy: velocity components must be in PPP");
    double *d = c->dsmall + 4000 + 24;
    k_energy(c, vr->e, vp->e, vz->e, d);
    HIPCHK(hipMemcpyAsync(c->hsmall, d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *out = c->hsmall[0];
  });
}

int mlx_time_kernel(mlx_ctx *c, mlx_field *f, const char *which, int32_t iters, double *ms) {
  return guarded(c, [&] {
    need_kit(c);
    std::string w(which ? which : "");
    auto launch = [&]() {
      if (w == "leg_fwd") leg_forward(c, f->e, c->scr[0], 0, false);
      else if (w == "leg_bwd") leg_backward(c, f->e, c->scr[0], 0, false);
      else if (w == "fft_phi_fwd") fft_phi(c, f->e, f->e, -1);
      else if (w == "fft_phi_bwd") fft_phi(c, f->e, f->e, +1);
      else if (w == "fft_z_fwd") fft_z(c, f->e, -1, active_pairs(c->kit, 0));
      else if (w == "fft_z_bwd") fft_z(c, f->e, +1, active_pairs(c->kit, 0));
      else if (w == "v2t") vec2tp_project(c, f->e, f->e, f->e, c->scr[0], c->scr[1], 3);
      else if (w == "lu_solve") k_helm_solve(c, f->e, -2.0 / (c->p.dt * c->p.visc), 0, 0);
      else if (w == "phi_cross") {
        if (!fft_phi_cross(c, tmp(c, 0)->e, tmp(c, 1)->e, tmp(c, 2)->e, f->e, f->e, f->e, f->e, f->e, f->e, f->e))
          MLX_THROW(MLX_E_ARG, "mlx_time_kernel: phi_cross is not built for this np");
      }
      else if (w == "copy") k_copy_field(c, c->scr[0], f->e);
      else if (w == "read") k_read_field(c, f->e, reinterpret_cast<double *>(c->scr[0]));
      else if (w == "write") k_write_field(c, c->scr[0], 0.0);
      else MLX_THROW(MLX_E_ARG, "mlx_time_kernel: unknown kernel " + w);
    };
    if (iters < 1) MLX_THROW(MLX_E_ARG, "mlx_time_kernel: iters must be positive");
    launch();
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, c->stream));
    for (int i = 0; i < iters; ++i) launch();
    HIPCHK(hipEventRecord(b, c->stream));
    HIPCHK(hipEventSynchronize(b));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = (double)t / iters;
  });
}

int mlx_probe_start(mlx_ctx *c, int32_t max_launches) {
  return guarded(c, [&] {
    if (max_launches < 1) MLX_THROW(MLX_E_ARG, "mlx_probe_start: max_launches must be positive");
    ProbeSet &p = c->probe;
    for (int i = 0; i < MLX_PROBE_COUNT; ++i) {
      while ((int)p.ev[i].size() < 2 * max_launches) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        p.ev[i].push_back(e);
      }
      p.used[i] = 0;
      p.bytes[i] = 0.0;
      p.flops[i] = 0.0;
    }
    p.cap = max_launches;
    p.on = true;
  });
}

int mlx_probe_stop(mlx_ctx *c, double *out) {
  return guarded(c, [&] {
    ProbeSet &p = c->probe;
    p.on = false;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < MLX_PROBE_COUNT; ++i) {
      double tot = 0.0;
      for (int j = 0; j < p.used[i]; ++j) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, p.ev[i][2 * j], p.ev[i][2 * j + 1]));
        tot += t;
      }
      out[4 * i] = p.used[i];
      out[4 * i + 1] = tot;
      out[4 * i + 2] = p.bytes[i];
      out[4 * i + 3] = p.flops[i];
    }
  });
}

}  // extern "C"
