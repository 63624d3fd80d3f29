"""The oracle's vector-field operators and time stepping against closed forms.

  tp2vec / tp2curlvec (mlegs_scalar_ops.f90:1455-1560)  the q-vortex
      u_phi = (1 - e^{-r^2})/r, u_z = e^{-r^2}/q, w_z = 2 e^{-r^2},
      w_phi = 2 r e^{-r^2}/q  (docs/tutorial/vector_field.md:92-106)
  save_vort_mag's |w|        2 e^{-r^2} sqrt(1 + r^2/q^2) of the same vortex
  (vortical_flow_3d.f90:411-446)
  vec2tp (:1308-1453)        inverse of tp2vec on smooth fields; a gradient
                             field projects to zero
  vecprod (:1264-1306)       numpy's cross product on both packed phi points
  febe / abcn + bootstrap    the columnar vortex diffuses as a Gaussian
  (vortical_flow_3d.f90)     (its advection term is a pure gradient): errors
                             fall as dt for FE/BE and dt^2 for AB2/CN
"""
import math

import numpy as np
import pytest
import sympy as sp

from oracle import kit as K, scalar_ops as so, vortex as V
from mlegs_amd.params import Params
from mlegs_amd import inputs

from test_oracle_ops import _sample, _rel

Q = 4.0
L = 2.0


def _kit(**kw):
    d = dict(nr=64, np=8, nz=4, nrchop=64, npchop=5, nzchop=2, ell=L, zlen=2 * math.pi, hyperpow=0,
             visc=1e-3, dt=1e-3, is_svv=False)
    d.update(kw)
    p = Params(**d)
    return K.Kit3D(p), p


def _qvortex(kit, p):
    dims = (kit.nrdim, kit.npdim, kit.nzdim)
    out = []
    for which in ('psi', 'chi'):
        s = so.Scalar(kit, 'PPP')
        s.e = inputs.qvortex_ppp(kit.x, kit.r, dims, p.np, p.nz, p.nr, Q, which)
        so.idelsqp(s, kit)
        so.zeroat1(s, kit)
        out.append(s)
    return out


def _radial(s, kit):
    """The packed PPP values as (nr, np/2, nz) complex; both phi points must match."""
    return s.e[:kit.nr, :kit.np // 2, :]


def test_qvortex_velocity_and_vorticity():
    kit, p = _kit()
    psi, chi = _qvortex(kit, p)
    # the log coefficient is (1/2) of the source integral: -2 e^{-r^2} -> -1/2
    assert abs(psi.ln + 0.5) < 1e-13 and abs(chi.ln + 0.25 / Q) < 1e-13
    r = kit.r[:, None, None]
    vr, vp, vz = so.tp2vec(psi.copy(), chi.copy(), kit)
    c = 1 + 1j
    assert np.abs(_radial(vr, kit)).max() < 1e-14
    assert np.abs(_radial(vp, kit) - c * (1 - np.exp(-r ** 2)) / r).max() < 1e-13
    assert np.abs(_radial(vz, kit) - c * np.exp(-r ** 2) / Q).max() < 1e-13
    wr, wp, wz = so.tp2curlvec(psi.copy(), chi.copy(), kit)
    assert np.abs(_radial(wr, kit)).max() < 1e-14
    assert np.abs(_radial(wp, kit) - c * 2 * r * np.exp(-r ** 2) / Q).max() < 5e-12
    assert np.abs(_radial(wz, kit) - c * 2 * np.exp(-r ** 2)).max() < 5e-12


def test_vorticity_magnitude_qvortex():
    """save_vort_mag's field (vortical_flow_3d.f90:411-446) of the q-vortex:
    |w| = 2 e^{-r^2} sqrt(1 + r^2/q^2) at both packed phi points (real and
    imaginary parts), from the closed forms above."""
    kit, p = _kit()
    psi, chi = _qvortex(kit, p)
    for s in (psi, chi):
        so.trans(s, 'FFF', kit)
    vm = V.vorticity_magnitude(psi, chi, kit)
    assert vm.space == 'PPP' and vm.ln == 0.0
    r = kit.r[:, None, None]
    want = 2 * np.exp(-r ** 2) * np.sqrt(1 + r ** 2 / Q ** 2)
    got = _radial(vm, kit)
    assert np.abs(got.real - want).max() < 5e-12
    assert np.abs(got.imag - want).max() < 5e-12
    with pytest.raises(ValueError):
        V.vorticity_magnitude(vm, chi, kit)


def test_vec2tp_inverts_tp2vec_on_smooth_fields():
    kit, p = _kit()
    w = ((kit.r[:, None, None] ** 2 + L ** 2) / (2 * L ** 2)) ** 2   # (1-x)^-2

    def tp(f):
        s = so.Scalar(kit, 'PPP')
        s.e = _sample(kit, f)
        s.e[:kit.nr] *= w
        so.idelsqp(s, kit)
        so.zeroat1(s, kit)
        return s

    psi = tp(lambda r, ph, z: np.exp(-r ** 2) * (2 + r * np.cos(ph) * np.cos(z) + r ** 2 * np.sin(2 * ph)))
    chi = tp(lambda r, ph, z: np.exp(-r ** 2) * (0.3 + r ** 2 * np.cos(2 * ph) * np.sin(z) + r * np.sin(ph)))
    vr, vp, vz = so.tp2vec(psi.copy(), chi.copy(), kit)
    ps, cs = so.vec2tp(vr, vp, vz, kit)
    assert _rel(ps.e, psi.e) < 1e-11 and _rel(cs.e, chi.e) < 1e-10
    assert abs(ps.ln - psi.ln) < 1e-12 and abs(cs.ln - chi.ln) < 1e-12


def test_vec2tp_removes_gradients():
    kit, p = _kit()
    r, ph, z = sp.symbols('r p z', positive=True)
    pot = sp.exp(-r ** 2) * (r * sp.cos(ph) * sp.cos(z) + r ** 2 * sp.sin(2 * ph) + 1)
    comps = [sp.diff(pot, r), sp.diff(pot, ph) / r, sp.diff(pot, z)]
    vs = []
    for c in comps:
        s = so.Scalar(kit, 'PPP')
        s.e = _sample(kit, sp.lambdify((r, ph, z), c, 'numpy'))
        vs.append(s)
    scale = max(np.abs(s.e).max() for s in vs)
    ps, cs = so.vec2tp(*vs, kit)
    assert np.abs(ps.e).max() < 1e-13 * scale and np.abs(cs.e).max() < 1e-13 * scale
    assert abs(ps.ln) < 1e-14


def test_vecprod_matches_cross_product():
    kit, p = _kit(nr=16, nrchop=16, np=12, npchop=4, nz=6, nzchop=3)
    rng = np.random.default_rng(9)
    sl = (slice(0, kit.nr), slice(0, kit.np // 2), slice(None))
    sc = []
    for _ in range(6):
        s = so.Scalar(kit, 'PPP')
        s.e[sl] = rng.standard_normal(s.e[sl].shape) + 1j * rng.standard_normal(s.e[sl].shape)
        sc.append(s)
    a = np.stack([s.e[sl] for s in sc[:3]], -1)
    b = np.stack([s.e[sl] for s in sc[3:]], -1)
    ref = np.cross(a.real, b.real) + 1j * np.cross(a.imag, b.imag)
    so.vecprod(*sc, kit)
    for c in range(3):
        np.testing.assert_allclose(sc[c].e[sl], ref[..., c], rtol=0, atol=1e-14)
    s = so.Scalar(kit, 'FFF')
    with pytest.raises(ValueError):
        so.vecprod(s, *sc[1:], kit)


def test_columnar_vortex_has_no_advection():
    """(U + u) x w of a columnar vortex is radial and r-dependent only, a gradient."""
    kit, p = _kit()
    psi, chi = _qvortex(kit, p)
    nlp, nlc = V.advection_rhs(psi, chi, so.Scalar(kit, 'PPP'), kit)
    assert np.abs(nlp.e).max() < 1e-14 and np.abs(nlc.e).max() < 1e-14
    assert nlp.ln == 0.0 and nlc.ln == 0.0


def test_check_stability_flags_nonfinite():
    kit, p = _kit(nr=16, nrchop=16)
    a, b = so.Scalar(kit, 'FFF'), so.Scalar(kit, 'FFF')
    V.check_stability(a, b)
    b.e[3, 1, 2] = np.nan
    with pytest.raises(FloatingPointError):
        V.check_stability(a, b)


NU = 0.1
T_END = 0.8


def _diffused(kit, t):
    a = 1 + 4 * NU * t
    r = kit.r
    return 2 * np.exp(-r ** 2 / a) / a, np.exp(-r ** 2 / a) / (Q * a)


def _vortex_error(dt):
    kit, p = _kit(np=4, npchop=2, nz=1, nzchop=1, zlen=1.0, visc=NU, dt=dt)
    psi, chi = _qvortex(kit, p)
    run = V.VortexRun(kit, p, psi, chi, so.Scalar(kit, 'PPP'))
    run.bootstrap()
    n = int(round(T_END / dt))
    for _ in range(n - 1):
        run.step()
    assert run.curr_n == n
    vr, vp, vz = so.tp2vec(run.psi.copy(), run.chi.copy(), kit)
    wr, wp, wz = so.tp2curlvec(run.psi.copy(), run.chi.copy(), kit)
    wz_ex, uz_ex = _diffused(kit, n * dt)
    return max(np.abs(wz.e[:kit.nr, 0, 0].real - wz_ex).max() / 2, np.abs(vz.e[:kit.nr, 0, 0].real - uz_ex).max() * Q)


def test_ab2cn_second_order_in_time():
    """FE-BE + Richardson bootstrap, then AB2/CN (vortical_flow_3d.f90:116-181)."""
    e = [_vortex_error(dt) for dt in (0.05, 0.025, 0.0125)]
    assert e[2] < 2e-6
    assert e[0] / e[1] > 3.0 and e[1] / e[2] > 3.5


def test_febe_first_order_in_time():
    """Backward Euler on pure diffusion (febe, mlegs_scalar_ops.f90:1157-1198)."""
    errs = []
    for dt in (0.1, 0.05, 0.025):
        kit, p = _kit(np=4, npchop=2, nz=1, nzchop=1, zlen=1.0, visc=NU, dt=dt)
        s = so.Scalar(kit, 'PPP')
        s.e[:kit.nr, :2, 0] = (2 * np.exp(-kit.r ** 2))[:, None] * (1 + 1j)
        so.trans(s, 'FFF', kit)
        nl = so.Scalar(kit, 'FFF')
        n = int(round(T_END / dt))
        for _ in range(n):
            so.febe(s, nl, dt, kit, p)
        so.trans(s, 'PPP', kit)
        ex, _ = _diffused(kit, n * dt)
        errs.append(np.abs(s.e[:kit.nr, 0, 0].real - ex).max())
    assert 1.7 < errs[0] / errs[1] < 2.3 and 1.8 < errs[1] / errs[2] < 2.2


def test_steppers_reject_inviscid_and_physical_input():
    kit, p = _kit(nr=16, nrchop=16)
    s, nl = so.Scalar(kit, 'FFF'), so.Scalar(kit, 'FFF')
    with pytest.raises(ValueError):
        so.febe(s, nl, 1e-3, kit, p.replace(visc=0.0))
    with pytest.raises(ValueError):
        so.abcn(s, s.copy(), so.Scalar(kit, 'PPP'), nl.copy(), 1e-3, kit, p)


def test_energy_closed_forms():
    """1/2 int |u|^2 over the domain: u_r = e^{-r^2} gives
    1/2 * 1/4 * 2 pi * zlen (int_0^inf e^{-2 r^2} r dr = 1/4); the q-vortex
    axial flow e^{-r^2}/q gives the same over q^2; a cos(phi) cos(z)
    modulation halves it twice."""
    kit, p = _kit(nr=64, nrchop=64, np=8, npchop=5, nz=4, nzchop=3, zlen=3.0)
    zero = so.Scalar(kit, 'PPP')
    u = so.Scalar(kit, 'PPP')
    u.e = _sample(kit, lambda r, ph, z: np.exp(-r ** 2) + 0 * ph + 0 * z)
    want = 0.5 * 0.25 * 2 * math.pi * 3.0
    assert abs(V.energy(u, zero, zero, kit) - want) < 1e-13 * want
    uz = so.Scalar(kit, 'PPP')
    uz.e = _sample(kit, lambda r, ph, z: np.exp(-r ** 2) / Q + 0 * ph + 0 * z)
    assert abs(V.energy(zero, zero, uz, kit) - want / Q ** 2) < 1e-13 * want
    m = so.Scalar(kit, 'PPP')
    m.e = _sample(kit, lambda r, ph, z: np.exp(-r ** 2) * np.cos(ph) * np.cos(2 * math.pi * z / 3.0))
    assert abs(V.energy(zero, m, zero, kit) - want / 4) < 1e-13 * want
    with pytest.raises(ValueError):
        V.energy(so.Scalar(kit, 'FFF'), zero, zero, kit)
