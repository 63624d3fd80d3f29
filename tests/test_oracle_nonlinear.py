"""The oracle's nonlinear half against mathematics it does not contain.

  advection_rhs (vortical_flow_3d.f90:353-395)
      on a non-columnar flow (q-vortex + an m = 1, 2 / kappa = 1 perturbation
      of finite amplitude + an axial background U_z with a uniform part): the
      curl of tp2vec(N) equals curl[(U + V) x curl V] from sympy at the
      collocation points.  The projection removes gradients only, so the
      curl fixes the sign and order of the cross product, the U_z term, the
      r weights and the three contractions of vec2tp.
  febe (mlegs_scalar_ops.f90:1157-1198)
      s_t = nu del^2 s + N with N(t) manufactured from a closed-form s(r, t)
      that carries a time-dependent log term: errors fall as dt.
  Richardson bootstrap + abcn (vortical_flow_3d.f90:116-181,
  mlegs_scalar_ops.f90:1200-1262)
      the same forced problem through VortexRun (its advection term replaced
      by the forcing at the times the driver evaluates it): errors fall as
      dt^2.  A swapped 1.5 / -0.5 weighting (planted in a copy on the CPU)
      loses it.
"""
import math

import numpy as np
import pytest
import sympy as sp

from oracle import kit as K, scalar_ops as so, vortex as V
from mlegs_amd.params import Params
from mlegs_amd import inputs

from test_oracle_ops import _sample

Q = 4.0


# ---------------------------------------------------------------------------
# advection_rhs against sympy
# ---------------------------------------------------------------------------
R, PH, Z = sp.symbols('r p z', positive=True)
EPS = 0.3


def _curl(a):
    ar, ap, az = a
    return [sp.diff(az, PH) / R - sp.diff(ap, Z),
            sp.diff(ar, Z) - sp.diff(az, R),
            (sp.diff(R * ap, R) - sp.diff(ar, PH)) / R]


def _tp_velocity(psi, chi):
    """u = curl(psi z) + curl curl(chi z) (tp2vec, mlegs_scalar_ops.f90:1455-1545)."""
    lap_h = sp.diff(R * sp.diff(chi, R), R) / R + sp.diff(chi, PH, 2) / R ** 2
    return [sp.diff(psi, PH) / R + sp.diff(chi, R, Z),
            -sp.diff(psi, R) + sp.diff(chi, PH, Z) / R,
            -lap_h]


def _perturbation():
    psi_p = EPS * R * sp.exp(-R ** 2) * sp.cos(PH) * sp.cos(Z)
    chi_p = EPS * R ** 2 * sp.exp(-R ** 2) * sp.sin(2 * PH) * sp.sin(Z) + EPS * 0.5 * sp.exp(-R ** 2) * sp.cos(Z)
    return psi_p, chi_p


def _uz_background():
    return -0.5 + 0.2 * sp.exp(-R ** 2) * (1 + R * sp.cos(PH))


def _fff(kit, expr):
    s = so.Scalar(kit, 'PPP')
    s.e = _sample(kit, sp.lambdify((R, PH, Z), expr, 'numpy'))
    so.trans(s, 'FFF', kit)
    return s


def _advection_curl(kit, p, with_uz=True):
    """Oracle: curl tp2vec(N) of the test flow; sympy: curl[(U + V) x curl V]."""
    dims = (kit.nrdim, kit.npdim, kit.nzdim)
    psi_p, chi_p = _perturbation()
    # q-vortex part as the driver builds it (idelsqp of the sources)
    fields = []
    for which, pert in (('psi', psi_p), ('chi', chi_p)):
        s = so.Scalar(kit, 'PPP')
        s.e = inputs.qvortex_ppp(kit.x, kit.r, dims, p.np, p.nz, p.nr, Q, which)
        so.idelsqp(s, kit)
        so.zeroat1(s, kit)
        s.e = s.e + _fff(kit, pert).e
        fields.append(s)
    psi, chi = fields
    uz = so.Scalar(kit, 'PPP')
    if with_uz:
        uz.e = _sample(kit, sp.lambdify((R, PH, Z), _uz_background(), 'numpy'))
    nlp, nlc = V.advection_rhs(psi, chi, uz, kit)
    assert nlp.ln == 0.0 and nlc.ln == 0.0
    got = [f.e[:p.nr, :p.np // 2] for f in so.tp2curlvec(nlp.copy(), nlc.copy(), kit)]

    # sympy: the q-vortex velocity in closed form plus the perturbation's
    vq = [sp.Integer(0), (1 - sp.exp(-R ** 2)) / R, sp.exp(-R ** 2) / Q]
    vp = _tp_velocity(psi_p, chi_p)
    v = [a + b for a, b in zip(vq, vp)]
    w = _curl(v)
    u = [v[0], v[1], v[2] + _uz_background()]
    f = [u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]]
    want = [_sample(kit, sp.lambdify((R, PH, Z), c, 'numpy'))[:p.nr, :p.np // 2] for c in _curl(f)]
    return got, want


@pytest.mark.parametrize('ell', [1.7, 2.5])
def test_advection_rhs_curl_matches_sympy(ell):
    """Closed forms are compared at r > 0.05: nearer the axis their terms
    (1 - e^{-r^2})/r and the like lose digits to cancellation in FP64.  The
    oracle's side converges spectrally in nr (max error 0.15, 4e-3, 6e-4,
    2e-6 of the scale at nr = 32, 48, 64, 96 for L = 1)."""
    p = Params(nr=128, np=16, nz=12, nrchop=128, npchop=9, nzchop=7, ell=ell, zlen=2 * math.pi, hyperpow=0)
    kit = K.Kit3D(p)
    got, want = _advection_curl(kit, p)
    far = kit.r > 0.05
    scale = max(np.abs(wc).max() for wc in want)
    errs = [np.abs((g - wc)[far]).max() for g, wc in zip(got, want)]
    assert scale > 1.0  # a flow with advection: the test has something to check
    assert max(errs) < 5e-7 * scale, (errs, scale)
    # the axial background matters at this tolerance: without it the same
    # comparison is off by far more
    got0, _ = _advection_curl(kit, p, with_uz=False)
    assert max(np.abs((g - wc)[far]).max() for g, wc in zip(got0, want)) > 1e-2 * scale


def test_advection_rhs_curl_check_sees_a_flipped_cross_product():
    """The check above fails for v x w taken the other way round (a planted
    mistake in a copy of the oracle's vecprod, on the CPU)."""
    p = Params(nr=48, np=16, nz=12, nrchop=48, npchop=9, nzchop=7, ell=1.0, zlen=2 * math.pi, hyperpow=0)
    kit = K.Kit3D(p)
    psi_p, chi_p = _perturbation()
    psi, chi = _fff(kit, psi_p), _fff(kit, chi_p)
    uz = so.Scalar(kit, 'PPP')
    orig = so.vecprod

    def flipped(vr, vp, vz, ur, up, uz_, kit_):
        orig(vr, vp, vz, ur, up, uz_, kit_)
        for s in (vr, vp, vz):
            s.e = -s.e

    good = V.advection_rhs(psi.copy(), chi.copy(), uz, kit)
    so.vecprod = flipped
    try:
        bad = V.advection_rhs(psi.copy(), chi.copy(), uz, kit)
    finally:
        so.vecprod = orig
    g0 = so.tp2curlvec(good[0].copy(), good[1].copy(), kit)
    g1 = so.tp2curlvec(bad[0].copy(), bad[1].copy(), kit)
    d = max(np.abs(a.e - b.e).max() for a, b in zip(g0, g1))
    assert d > 1e-3 * max(np.abs(a.e).max() for a in g0)


# ---------------------------------------------------------------------------
# forced diffusion: febe, Richardson + abcn with N != 0 and a log term
# ---------------------------------------------------------------------------
NU = 0.1
T_END = 0.8
ELL = 1.5


def _kit2d(dt):
    p = Params(nr=48, np=4, nz=1, nrchop=48, npchop=2, nzchop=1, ell=ELL, zlen=1.0, hyperpow=0, visc=NU, dt=dt,
               is_svv=False)
    return K.Kit3D(p), p


def _logfun(r):
    """The log term's radial function: -log(1 - x) = log((r^2 + L^2) / (2 L^2))."""
    return np.log((r ** 2 + ELL ** 2) / (2 * ELL ** 2))


def _lap_logfun(r):
    return 4 * ELL ** 2 / (r ** 2 + ELL ** 2) ** 2


# exact solutions s(r, t) = a(t) logfun(r) + b(t) exp(-r^2) + c(t) r^2 exp(-r^2)
SOLS = {
    'psi': (lambda t: 0.4 * math.cos(2 * t), lambda t: -0.8 * math.sin(2 * t),
            lambda t: 1 + math.sin(3 * t), lambda t: 3 * math.cos(3 * t),
            lambda t: 0.5 * t, lambda t: 0.5),
    'chi': (lambda t: 0.0, lambda t: 0.0,
            lambda t: math.exp(-t), lambda t: -math.exp(-t),
            lambda t: 0.3 * math.cos(t), lambda t: -0.3 * math.sin(t)),
}


def _profiles(r):
    g = np.exp(-r ** 2)
    h = r ** 2 * g
    lap_g = (4 * r ** 2 - 4) * g
    lap_h = (4 - 12 * r ** 2 + 4 * r ** 4) * g
    return g, h, lap_g, lap_h


def _field(kit, which, t, forcing):
    """The exact s(t) (forcing False) or N(t) = s_t - nu del^2 s (True), in FFF."""
    a, da, b, db, c, dc = SOLS[which]
    r = kit.r
    g, h, lap_g, lap_h = _profiles(r)
    if forcing:
        smooth = db(t) * g + dc(t) * h - NU * (a(t) * _lap_logfun(r) + b(t) * lap_g + c(t) * lap_h)
        ln = da(t)
    else:
        smooth = b(t) * g + c(t) * h
        ln = a(t)
    s = so.Scalar(kit, 'PPP')
    vals = smooth + ln * _logfun(r)
    s.e[:kit.nr, :kit.np // 2, 0] = (vals + 1j * vals)[:, None]
    s.ln = ln
    so.trans(s, 'FFF', kit)
    return s


def _err(kit, s, which, t):
    e = _field(kit, which, t, False)
    num = np.abs(s.e - e.e).max() + abs(s.ln - e.ln)
    return num / (np.abs(e.e).max() + abs(e.ln))


def test_forced_field_has_its_log_term():
    """The manufactured psi carries a time-dependent log coefficient, so the
    steppers' N.ln bookkeeping is exercised."""
    kit, p = _kit2d(0.1)
    s = _field(kit, 'psi', 0.3, False)
    assert abs(s.ln - 0.4 * math.cos(0.6)) < 1e-15
    n = _field(kit, 'psi', 0.3, True)
    assert abs(n.ln + 0.8 * math.sin(0.6)) < 1e-15
    # log(1-x) removed by trans: the FFF coefficients decay (no log residue)
    assert np.abs(s.e[30:48, 0, 0]).max() < 1e-8 * np.abs(s.e[:, 0, 0]).max()


def _febe_error(dt, which):
    kit, p = _kit2d(dt)
    s = _field(kit, which, 0.0, False)
    n = int(round(T_END / dt))
    for i in range(n):
        so.febe(s, _field(kit, which, i * dt, True), dt, kit, p)
    return _err(kit, s, which, n * dt)


@pytest.mark.parametrize('which', ['psi', 'chi'])
def test_febe_forced_first_order(which):
    e = [_febe_error(dt, which) for dt in (0.1, 0.05, 0.025)]
    assert e[2] < 6e-2
    assert 1.7 < e[0] / e[1] < 2.3 and 1.8 < e[1] / e[2] < 2.2, e


class _Forcing:
    """Stands in for advection_rhs in VortexRun: the driver evaluates its
    nonlinear term at t = 0 (three times), dt/2 and dt in the bootstrap
    (vortical_flow_3d.f90:116-147) and at t_{n+1} after each AB2/CN step."""

    def __init__(self, kit, dt):
        self.kit, self.dt, self.calls = kit, dt, 0

    def time(self, i):
        return [0.0, 0.0, 0.0, 0.5 * self.dt][i] if i < 4 else (i - 3) * self.dt

    def __call__(self, psi, chi, uz, kit):
        t = self.time(self.calls)
        self.calls += 1
        return _field(self.kit, 'psi', t, True), _field(self.kit, 'chi', t, True)


def _ab2cn_error(dt, monkeypatch, weights=None):
    kit, p = _kit2d(dt)
    forcing = _Forcing(kit, dt)
    monkeypatch.setattr(V, 'advection_rhs', forcing)
    if weights is not None:
        orig = so.abcn

        def abcn_w(s, s_p, nl, nl_p, dt_, kit_, p_):
            a, b = weights
            nl2 = nl.copy()
            nl2.e = (a * nl.e + b * nl_p.e) / 1.0
            nl2.ln = a * nl.ln + b * nl_p.ln
            zero = nl_p.copy()
            zero.e = zero.e * 0.0
            zero.ln = 0.0
            # abcn forms 1.5 N - 0.5 N_p: feed it N' = (a N + b N_p)/1.5, N_p = 0
            nl2.e = nl2.e / 1.5
            nl2.ln = nl2.ln / 1.5
            orig(s, s_p, nl2, zero, dt_, kit_, p_)
            nl_p.assign(nl)
        monkeypatch.setattr(so, 'abcn', abcn_w)
    run = V.VortexRun(kit, p, _field(kit, 'psi', 0.0, False), _field(kit, 'chi', 0.0, False),
                      so.Scalar(kit, 'PPP'))
    run.bootstrap()
    n = int(round(T_END / dt))
    for _ in range(n - 1):
        run.step()
    assert forcing.calls == 4 + n
    return max(_err(kit, run.psi, 'psi', n * dt), _err(kit, run.chi, 'chi', n * dt))


def test_richardson_ab2cn_forced_second_order(monkeypatch):
    e = [_ab2cn_error(dt, monkeypatch) for dt in (0.1, 0.05, 0.025)]
    assert e[2] < 2e-3
    assert e[0] / e[1] > 3.4 and e[1] / e[2] > 3.6, e


def test_second_order_check_sees_swapped_ab2_weights(monkeypatch):
    """-0.5 N + 1.5 N_p in place of 1.5 N - 0.5 N_p (a planted mistake on the
    CPU) loses the second order."""
    e = [_ab2cn_error(dt, monkeypatch, weights=(-0.5, 1.5)) for dt in (0.1, 0.05, 0.025)]
    assert not (e[0] / e[1] > 3.4 and e[1] / e[2] > 3.6), e


# ---------------------------------------------------------------------------
# explicit steppers, by self-convergence at T = 0.1 (differences of runs at
# dt, dt/2, dt/4, dt/8; the spatial error of the nr = 16 grid cancels)
# ---------------------------------------------------------------------------
def _kit_explicit(dt):
    p = Params(nr=16, np=4, nz=1, nrchop=16, npchop=2, nzchop=1, ell=ELL, zlen=1.0, hyperpow=0, visc=NU, dt=dt,
               is_svv=False)
    return K.Kit3D(p), p


def _explicit_run(mode, dt, which, t_end=0.1):
    kit, p = _kit_explicit(dt)
    n = int(round(t_end / dt))
    if mode == 'fefe':
        s = _field(kit, which, 0.0, False)
        for i in range(n):
            so.fefe(s, _field(kit, which, i * dt, True), dt, kit, p)
        return s
    # two-level start from the exact s(0), s(dt)
    s_p, s = _field(kit, which, 0.0, False), _field(kit, which, dt, False)
    nl_p = _field(kit, which, 0.0, True)
    for i in range(1, n):
        nl = _field(kit, which, i * dt, True)
        if mode == 'abab':
            so.abab(s, s_p, nl, nl_p, dt, kit, p)
        else:
            # the caller hands in nu del^2 s_{n-1} (is_2nd_svis_p)
            svp = s_p.copy()
            so.del2(svp, kit)
            svp.e = svp.e * NU
            s_p = s.copy()
            so.abab(s, svp, nl, nl_p, dt, kit, p, is_2nd_svis_p=True)
    return s


def _self_convergence(mode, which):
    runs = [_explicit_run(mode, dt, which) for dt in (8e-4, 4e-4, 2e-4, 1e-4)]
    d = [np.abs(a.e - b.e).max() + abs(a.ln - b.ln) for a, b in zip(runs, runs[1:])]
    return d[0] / d[1], d[1] / d[2]


@pytest.mark.parametrize('which', ['psi', 'chi'])
def test_fefe_forced_first_order(which):
    """Forward Euler on both terms (fefe, mlegs_scalar_ops.f90:1065-1096),
    with the log term."""
    r1, r2 = _self_convergence('fefe', which)
    assert 1.9 < r1 < 2.1 and 1.9 < r2 < 2.1, (r1, r2)
    kit, p = _kit_explicit(1e-4)
    assert _err(kit, _explicit_run('fefe', 1e-4, which), which, 0.1) < 1e-4


@pytest.mark.parametrize('which', ['psi', 'chi'])
def test_abab_is_second_order_given_the_previous_diffusive_term(which):
    """AB2 on both terms (mlegs_scalar_ops.f90:1098-1155) when the caller
    passes nu del^2 s_{n-1} itself (is_2nd_svis_p): second order."""
    r1, r2 = _self_convergence('abab2', which)
    assert r1 > 3.8 and r2 > 3.8, (r1, r2)


@pytest.mark.parametrize('which', ['psi', 'chi'])
def test_abab_history_quirk_is_first_order(which):
    """abab stores s_p = s after the update (mlegs_scalar_ops.f90:1151), so on
    the next call the diffusive history is the current state and the term
    1.5 D(s_n) - 0.5 D(s_p) reduces to D(s_n): forward Euler on diffusion,
    AB2 on the nonlinear term, first order overall.  A dropped 0.5 or a
    history written before the update would change these ratios."""
    r1, r2 = _self_convergence('abab', which)
    assert 1.9 < r1 < 2.1 and 1.9 < r2 < 2.1, (r1, r2)


def test_abab_keeps_the_history():
    kit, p = _kit_explicit(1e-4)
    s, s_p = _field(kit, 'psi', 1e-4, False), _field(kit, 'psi', 0.0, False)
    nl, nl_p = _field(kit, 'psi', 1e-4, True), _field(kit, 'psi', 0.0, True)
    so.abab(s, s_p, nl, nl_p, 1e-4, kit, p)
    assert np.array_equal(s_p.e, s.e) and s_p.ln == s.ln
    assert np.array_equal(nl_p.e, nl.e) and nl_p.ln == nl.ln
