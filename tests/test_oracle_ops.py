"""The oracle's field operations against closed forms, brute force and
operator identities (no reference code is run).

  FFTs (mlegs_scalar_ops.f90:1567-1848)      explicit DFT sums on small grids
  trans (:157-235)                           single-mode synthesis with scipy's
                                             associated Legendre functions; the
                                             closed form of docs/tutorial/operation.md:146
  del2 / del2h / xxdx / delsqp (:327-573)    analytic derivatives (sympy) at the
                                             collocation points; operation.md:150
  ihelm / idel2 (:575-854)                   operator residuals
  helmp / ihelmp (:856-1000)                 helmp against the dense band products
                                             ihelmp assembles; helmp(ihelmp(s)) = s
  chop / dealias (:6-70)                     index rules enumerated by hand
"""
import cmath
import math

import numpy as np
import pytest
import scipy.special as sps
import sympy as sp

from oracle import kit as K, scalar_ops as so
from mlegs_amd.params import Params
from mlegs_amd import inputs


def _kit(**kw):
    d = dict(nr=16, np=12, nz=6, nrchop=14, npchop=5, nzchop=3, ell=1.3, zlen=2.5, hyperpow=0)
    d.update(kw)
    return K.Kit3D(Params(**d))


def _pbar(n, m, x):
    """Fully normalised associated Legendre function with the Condon-Shortley phase (scipy)."""
    lognrm = 0.5 * (math.log((2 * n + 1) / 2.0) + math.lgamma(n - m + 1) - math.lgamma(n + m + 1))
    return sps.lpmv(m, n, x) * math.exp(lognrm)


def _sample(kit, f):
    """PPP array of a real function f(r, phi, z): Re = even phi points, Im = odd ones."""
    nr, nph = kit.nr, kit.np // 2
    r = kit.r[:, None, None]
    ph = 2 * math.pi / kit.np * np.arange(kit.np)
    z = kit.zlen / kit.nz * np.arange(kit.nz)
    e = np.zeros((kit.nrdim, kit.npdim, kit.nzdim), dtype=np.complex128)
    ev = f(r, ph[None, 0::2, None], z[None, None, :])
    od = f(r, ph[None, 1::2, None], z[None, None, :])
    e[:nr, :nph, :] = np.broadcast_to(ev, (nr, nph, kit.nz)) + 1j * np.broadcast_to(od, (nr, nph, kit.nz))
    return e


def _phys(kit, e):
    return e[:kit.nr, :kit.np // 2, :]


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


# ---------------------------------------------------------------------------
# FFTs
# ---------------------------------------------------------------------------
def test_horizontal_fft_bruteforce():
    kit = _kit()
    rng = np.random.default_rng(3)
    s = so.Scalar(kit, 'PPP')
    s.e[:kit.nr, :kit.np // 2, :] = rng.standard_normal((kit.nr, 6, 6)) + 1j * rng.standard_normal((kit.nr, 6, 6))
    f = np.zeros((kit.nr, kit.np, kit.nz))
    f[:, 0::2, :] = s.e[:kit.nr, :6, :].real
    f[:, 1::2, :] = s.e[:kit.nr, :6, :].imag
    so.horizontal_fft_forward(s, kit)
    N = kit.np
    for i in (0, 5, 15):
        for k in (0, 4):
            for m in range(N // 2 + 1):
                c = sum(f[i, j, k] * cmath.exp(-2j * math.pi * j * m / N) for j in range(N)) / N
                assert abs(s.e[i, m, k] - c) < 1e-14
    # Hermitian synthesis: f_j = Re c0 + 2 sum_{0<m<N/2} Re(c_m e^{i m phi_j}) + Re c_{N/2} (-1)^j
    c = s.e.copy()
    so.horizontal_fft_backward(s, kit)
    for i in (0, 9):
        for k in (1, 5):
            for j in range(N):
                v = c[i, 0, k].real + c[i, N // 2, k].real * (-1) ** j
                v += 2 * sum((c[i, m, k] * cmath.exp(2j * math.pi * j * m / N)).real for m in range(1, N // 2))
                got = s.e[i, j // 2, k].real if j % 2 == 0 else s.e[i, j // 2, k].imag
                assert abs(got - v) < 1e-13
                assert abs(got - f[i, j, k]) < 1e-13


def test_vertical_fft_bruteforce():
    kit = _kit(nz=10, nzchop=4)
    rng = np.random.default_rng(4)
    s = so.Scalar(kit, 'FFP')
    s.e[:] = rng.standard_normal(s.e.shape) + 1j * rng.standard_normal(s.e.shape)
    g = s.e.copy()
    so.vertical_fft_forward(s, kit)
    N = kit.nz
    for (i, j) in ((0, 0), (3, 2), (18, 6)):
        for kap in range(N):
            c = sum(g[i, j, k] * cmath.exp(-2j * math.pi * k * kap / N) for k in range(N)) / N
            assert abs(s.e[i, j, kap] - c) < 1e-14
    so.vertical_fft_backward(s, kit)
    assert np.abs(s.e - g).max() < 1e-14


# ---------------------------------------------------------------------------
# transforms
# ---------------------------------------------------------------------------
@pytest.mark.parametrize('n,m,kap', [(0, 0, 0), (3, 0, 0), (4, 1, 0), (6, 2, 1), (9, 4, -2), (2, 2, 2)])
def test_trans_single_mode_synthesis(n, m, kap):
    """One coefficient c at (n, m, kappa) synthesises 2 Re(c P_n^m(x) e^{i(m phi + kappa z)})."""
    kit = _kit()
    s = so.Scalar(kit, 'FFF')
    k = kap % kit.nz
    c = 0.7 - 0.4j
    if m == 0 and kap == 0:
        c = 0.7
        s.e[n - m, 0, 0] = c
    elif m == 0:
        s.e[n, 0, k] = c
        s.e[n, 0, (-kap) % kit.nz] = np.conj(c)
    else:
        s.e[n - m, m, k] = c
    so.trans(s, 'PPP', kit)
    nr = kit.nr
    P = np.array([_pbar(n, m, xv) for xv in kit.x])
    ph = 2 * math.pi / kit.np * np.arange(kit.np)
    z = kit.zlen / kit.nz * np.arange(kit.nz)
    akv = 2 * math.pi / kit.zlen * kap
    if m == 0 and kap == 0:
        ref = np.broadcast_to(c * P[:, None, None], (nr, kit.np, kit.nz))
    else:
        ref = 2 * np.real(c * P[:, None, None] * np.exp(1j * (m * ph[None, :, None] + akv * z[None, None, :])))
    got = np.zeros((nr, kit.np, kit.nz))
    got[:, 0::2, :] = s.e[:nr, :kit.np // 2, :].real
    got[:, 1::2, :] = s.e[:nr, :kit.np // 2, :].imag
    assert np.abs(got - ref).max() < 1e-13 * max(1.0, np.abs(ref).max())


def test_trans_closed_form_operation_md():
    """s_2^1 = s_3^2 = 1 at L = 1 (docs/tutorial/operation.md:146) and its Laplacian (:150)."""
    kit = _kit(nr=32, np=8, nz=1, nrchop=32, npchop=3, nzchop=1, ell=1.0, zlen=1.0)
    s = so.Scalar(kit, 'FFF')
    s.e[1, 1, 0] = 1.0   # n = 2, m = 1
    s.e[1, 2, 0] = 1.0   # n = 3, m = 2
    lap = s.copy()
    so.trans(s, 'PPP', kit)

    def f(r, ph, z):
        return 2 * (math.sqrt(5 / 12) * (-6 * r * (r * r - 1) / (r * r + 1) ** 2) * np.cos(ph)
                    + math.sqrt(7 / 240) * (60 * r * r * (r * r - 1) / (r * r + 1) ** 3) * np.cos(2 * ph))

    def lf(r, ph, z):
        return (48 * math.sqrt(15) * r * (r * r - 1) * ((r * r + 1) * np.cos(ph) - 2 * math.sqrt(7) * r * np.cos(2 * ph))
                / (r * r + 1) ** 5)

    ref = _phys(kit, _sample(kit, f))
    assert np.abs(_phys(kit, s.e) - ref).max() < 1e-13
    so.del2(lap, kit)
    so.trans(lap, 'PPP', kit)
    ref = _phys(kit, _sample(kit, lf))
    assert np.abs(_phys(kit, lap.e) - ref).max() < 1e-12


@pytest.mark.parametrize('ln', [0.0, 0.37])
def test_trans_round_trip(ln):
    kit = _kit(nr=24, np=16, nz=8, nrchop=20, npchop=7, nzchop=4, ell=2.0, zlen=3.0)
    dims = (kit.nrdim, kit.npdim, kit.nzdim)
    e = inputs.transform_field(dims, (kit.nrchop, kit.npchop, kit.nzchop, kit.np, kit.nz), field_id=7, decay=6.0)
    so_ = so.Scalar(kit, 'FFF')
    so_.e = e.copy()
    so_.ln = ln
    so.trans(so_, 'PPP', kit)
    assert so_.space == 'PPP'
    so.trans(so_, 'FFF', kit)
    assert _rel(so_.e, e) < 1e-13
    assert so_.ln == ln


def test_trans_state_machine_errors():
    kit = _kit()
    s = so.Scalar(kit, 'FFF')
    with pytest.raises(ValueError):
        so.trans(s, 'PFF', kit)
    s.space = 'XYZ'
    with pytest.raises(ValueError):
        so.trans(s, 'PPP', kit)


# ---------------------------------------------------------------------------
# band operators against analytic derivatives
# ---------------------------------------------------------------------------
_r, _p, _z = sp.symbols('r p z', positive=True)
_F = sp.exp(-_r ** 2) * (1 + _r ** 2 * sp.cos(2 * _p) * sp.cos(_z) + _r * sp.sin(_p) * sp.sin(_z)
                          + sp.Rational(1, 3) * _r ** 3 * sp.cos(3 * _p))
_LAPH = sp.diff(_F, _r, 2) + sp.diff(_F, _r) / _r + sp.diff(_F, _p, 2) / _r ** 2
_LAP = _LAPH + sp.diff(_F, _z, 2)
_XXDX = _r * sp.diff(_F, _r)


def _lamb(expr):
    return sp.lambdify((_r, _p, _z), sp.simplify(expr), 'numpy')


ELL = 2.0


def _bandkit():
    return _kit(nr=64, np=8, nz=4, nrchop=64, npchop=5, nzchop=2, ell=ELL, zlen=2 * math.pi)


@pytest.mark.parametrize('op', ['del2', 'del2h', 'xxdx', 'delsqp'])
def test_band_operators_vs_analytic(op):
    kit = _bandkit()
    s = so.Scalar(kit, 'PPP')
    s.e = _sample(kit, _lamb(_F))
    so.trans(s, 'FFF', kit)
    getattr(so, op)(s, kit)
    so.trans(s, 'PPP', kit)
    if op == 'delsqp':
        # (1-x)^-2 del_perp^2 with (1-x)^-2 = ((r^2+L^2)/(2L^2))^2
        expr = _LAPH * ((_r ** 2 + ELL ** 2) / (2 * ELL ** 2)) ** 2
    else:
        expr = {'del2': _LAP, 'del2h': _LAPH, 'xxdx': _XXDX}[op]
    ref = _phys(kit, _sample(kit, _lamb(expr)))
    # spectral truncation of the Gaussian at NR = 64, L = 2 leaves ~1e-12
    assert _rel(_phys(kit, s.e), ref) < 1e-10, op


def test_log_term_derivatives():
    """f = ln * (-log(1-x)) is represented by s.ln alone; xxdx and del2h then act
    on log(1 + r^2/L^2) + log 2 (mlegs_scalar_ops.f90:452-456, 563-566)."""
    kit = _bandkit()
    s = so.Scalar(kit, 'FFF')
    s.ln = 1.0
    x = s.copy()
    so.xxdx(x, kit)
    so.trans(x, 'PPP', kit)
    r = kit.r
    # -log(1-x) = log((r^2+L^2)/(2L^2)); r d/dr of it = 2 r^2/(r^2+L^2)
    ref = 2 * r ** 2 / (r ** 2 + ELL ** 2)
    got = x.e[:kit.nr, 0, 0]
    assert np.abs(got.real - ref).max() < 1e-12 and np.abs(got.imag - ref).max() < 1e-12
    h = s.copy()
    so.del2h(h, kit)
    so.trans(h, 'PPP', kit)
    # del_perp^2 log(r^2+L^2) = 4 L^2/(r^2+L^2)^2
    ref = 4 * ELL ** 2 / (r ** 2 + ELL ** 2) ** 2
    got = h.e[:kit.nr, 0, 0]
    assert np.abs(got.real - ref).max() < 1e-12


# ---------------------------------------------------------------------------
# solves
# ---------------------------------------------------------------------------
def _solvekit():
    return _kit(nr=24, np=12, nz=8, nrchop=22, npchop=6, nzchop=4, ell=1.5, zlen=3.0)


def _field(kit, fid, ln=0.0):
    dims = (kit.nrdim, kit.npdim, kit.nzdim)
    s = so.Scalar(kit, 'FFF')
    s.e = inputs.transform_field(dims, (kit.nrchop, kit.npchop, kit.nzchop, kit.np, kit.nz), field_id=fid, decay=5.0)
    s.ln = ln
    return s


@pytest.mark.parametrize('alpha', [-7.3, 2.5, -1e4])
def test_ihelm_residual(alpha):
    """(del2 + alpha) ihelm(s, alpha) = s, log term included."""
    kit = _solvekit()
    s = _field(kit, 11, ln=0.4)
    x = s.copy()
    so.ihelm(x, alpha, kit)
    d = x.copy()
    so.del2(d, kit)
    res = d.e + alpha * x.e
    assert _rel(res, s.e) < 1e-12
    assert abs(alpha * x.ln - s.ln) < 1e-14


def test_ihelm_rejects_zero_alpha():
    kit = _solvekit()
    with pytest.raises(ValueError):
        so.ihelm(_field(kit, 1), 0.0, kit)


@pytest.mark.parametrize('power', [4, 6, 8])
def test_helmp_matches_band_products(power):
    """helmp (repeated del2 on the field) equals the dense product of the
    nn x nn truncations that ihelmp assembles (leg_del2^(p/2) + beta leg_del2
    + alpha), column by column, for a chopped field without log term."""
    kit = _solvekit()
    s = _field(kit, 21)
    alpha, beta = 3.7, -0.45
    x = s.copy()
    so.helmp(x, power, alpha, beta, kit)
    nrc, npc, nzc, nzcu, nrcs = so.chop_index(s, kit)
    for mm in range(npc):
        nn = nrcs[mm]
        for k in so.zlist(kit, nzc, nzcu):
            W = so.helmp_band(kit, kit.m[mm], kit.ak[k], nn, power, alpha, beta)
            want = W @ s.e[:nn, mm, k]
            assert np.allclose(x.e[:nn, mm, k], want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    assert x.ln == 0.0


@pytest.mark.parametrize('power,alpha,beta', [(4, -2.0e3, 0.8), (6, 55.0, -1.5), (8, -1.0e6, 0.0)])
def test_ihelmp_residual(power, alpha, beta):
    """helmp(ihelmp(s)) = s, the log term and the (0, 0) correction included."""
    kit = _solvekit()
    s = _field(kit, 22, ln=0.3)
    x = s.copy()
    so.ihelmp(x, power, alpha, beta, kit)
    assert abs(alpha * x.ln - s.ln) < 1e-14
    y = x.copy()
    so.helmp(y, power, alpha, beta, kit)
    assert _rel(y.e, s.e) < 1e-10
    assert abs(y.ln - s.ln) < 1e-14


def test_helmp_power_checks():
    kit = _solvekit()
    for pw in (2, 5, 10):
        with pytest.raises(ValueError):
            so.helmp(_field(kit, 1), pw, 1.0, 0.0, kit)
        with pytest.raises(ValueError):
            so.ihelmp(_field(kit, 1), pw, 1.0, 0.0, kit)
    with pytest.raises(ValueError):
        so.ihelmp(_field(kit, 1), 4, 0.0, 1.0, kit)


def test_idel2_proln_residual():
    """del2 idel2(s) = s; the (m, k) = (0, 0) system is bordered by the log term."""
    kit = _solvekit()
    s = _field(kit, 12)
    x = s.copy()
    so.idel2_proln(x, kit)
    assert x.ln != 0.0
    d = x.copy()
    so.del2(d, kit)
    assert _rel(d.e, s.e) < 1e-11


def test_idel2_preln_residual():
    kit = _solvekit()
    s = _field(kit, 13)
    x = s.copy()
    so.idel2_preln(x, kit, 0.25)
    assert abs(x.ln - 0.25) < 1e-13
    d = x.copy()
    so.del2(d, kit)
    # the (0,0) system trades its last equation for the prescribed log term
    mask = np.ones(s.e.shape, bool)
    n00 = kit.nrchop
    mask[n00 - 1, 0, 0] = False
    assert _rel(d.e[mask], s.e[mask]) < 1e-11


def test_repeated_slice_when_nzchop_is_half():
    """nzchop = nz/2 + 1 makes chopzl = chopzu: the two kk-loops of del2 (:546-557)
    both visit that slice, so del2 acts on it twice."""
    kit = _kit(nr=16, np=8, nz=8, nrchop=14, npchop=3, nzchop=5, ell=1.0, zlen=2.0)
    assert kit.chopzl == kit.chopzu == 5
    ks = so.zlist(kit, kit.chopzl, kit.chopzu)
    assert ks.count(4) == 2 and len(ks) == 9
    s = so.Scalar(kit, 'FFF')
    rng = np.random.default_rng(2)
    s.e[:14, 0, 4] = rng.standard_normal(14)
    x = s.copy()
    so.del2(x, kit)
    D = kit.leg_del2(0, kit.ak[4], 14)
    np.testing.assert_allclose(x.e[:14, 0, 4], D @ (D @ s.e[:14, 0, 4]), rtol=1e-13)


# ---------------------------------------------------------------------------
# masks
# ---------------------------------------------------------------------------
def test_chop_rule_enumerated():
    """Zero rows i >= nrchop + off - m, columns m >= npchop and 1-based slices
    nzchop+1 .. nz-nzchop+1 (chop, :6-41; chop_index, :2012-2063)."""
    kit = _kit(nr=12, np=10, nz=10, nrchop=10, npchop=4, nzchop=3)
    for off in (0, 3):
        s = so.Scalar(kit, 'FFF')
        s.e[:] = 1.0
        s.chop_offset(off)
        so.chop(s, kit)
        for i in range(kit.nrdim):
            for j in range(kit.npdim):
                for k in range(kit.nzdim):
                    keep = (j < 4 and i < 10 + off - j) and not (3 + 1 <= k + 1 <= 10 - 3 + 1)
                    assert s.e[i, j, k] == (1.0 if keep else 0.0), (off, i, j, k)


def test_dealias_rule_enumerated():
    """2/3 rule: zero m > np/3 and 1-based slices zcut < k < nz - zcut + 2 with
    zcut = nz/3 + 1 (:43-70)."""
    kit = _kit(nr=12, np=18, nz=12, nrchop=10, npchop=10, nzchop=7)
    s = so.Scalar(kit, 'FFF')
    s.e[:] = 1.0
    so.dealias(s, kit)
    for j in range(kit.npdim):
        for k in range(kit.nzdim):
            zcut = 12 // 3 + 1
            keep = (j <= 18 // 3) and not (zcut < k + 1 < 12 - zcut + 2)
            assert np.all(s.e[:, j, k] == (1.0 if keep else 0.0)), (j, k)
    # a physical-space axis is left alone
    s = so.Scalar(kit, 'PPP')
    s.e[:] = 1.0
    so.dealias(s, kit)
    assert np.all(s.e == 1.0)


def test_svv_filter_properties():
    """svv_filter (:72-155): the gain relaxes to clamp(tail/target - 1, 0, 1);
    only entries with quality q > cutoff are damped, by exp(-strength ((q-c)/(1-c))^8)."""
    kit = _kit(nr=16, np=12, nz=6, nrchop=16, npchop=7, nzchop=4)
    p = Params(nr=16, np=12, nz=6, nrchop=16, npchop=7, nzchop=4, is_svv=True,
               svv_cutoff=0.5, svv_target=1e-3, svv_strength=0.5, svv_relax=1.0)
    rng = np.random.default_rng(5)
    s = so.Scalar(kit, 'FFF')
    s.e = rng.standard_normal(s.e.shape) + 1j * rng.standard_normal(s.e.shape)
    e0 = s.e.copy()
    g = so.svv_filter(s, kit, 0.0, p)
    assert g == 1.0   # tail fraction far above target saturates the feedback
    ratio = np.abs(s.e) / np.abs(e0)
    i = np.arange(s.e.shape[0])[:, None, None] / 15.0
    j = np.arange(s.e.shape[1])[None, :, None] / 6.0
    kmax = np.abs(kit.ak).max()
    k = np.abs(kit.ak)[None, None, :] / kmax
    q = np.minimum(1.0, np.maximum(np.maximum(i, j), k))
    expect = np.where(q > 0.5, np.exp(-0.5 * ((q - 0.5) / 0.5) ** 8), 1.0)
    np.testing.assert_allclose(ratio, np.broadcast_to(expect, ratio.shape), rtol=1e-14)
    # disabled: identity and unchanged gain
    s.e = e0.copy()
    assert so.svv_filter(s, kit, 0.3, p.replace(is_svv=False)) == 0.3
    assert np.array_equal(s.e, e0)


def test_calcat_values_at_origin_and_infinity():
    """calcat0 / calcat1 (:237-309) give f(r=0) and f(r=inf) of the m = 0 column:
    e^{-r^2} -> (1, 0); r^2/(r^2+L^2) = (1+x)/2 -> (0, 1)."""
    kit = _kit(nr=64, np=4, nz=1, nrchop=64, npchop=2, nzchop=1, ell=ELL, zlen=1.0)
    for f, at0, at1, tol in ((lambda r, p, z: np.exp(-r ** 2), 1.0, 0.0, 1e-12),
                             (lambda r, p, z: r ** 2 / (r ** 2 + ELL ** 2), 0.0, 1.0, 1e-13)):
        s = so.Scalar(kit, 'PPP')
        s.e = _sample(kit, f)
        so.trans(s, 'FFF', kit)
        c0, c1 = so.calcat0(s, kit), so.calcat1(s, kit)
        assert abs(c0[0] - at0) < tol and abs(c1[0] - at1) < tol
    # zeroat1 removes the value at infinity through the n = 0 coefficient
    so.zeroat1(s, kit)
    assert abs(so.calcat1(s, kit)[0]) < 1e-14


def test_svv_gain_on_a_two_entry_field():
    """Tail ratio b^2/(a^2+b^2) of a field with one entry below and one at the
    top of the radial range; gain and damping by hand (:72-155)."""
    kit = _kit(nr=16, np=4, nz=1, nrchop=16, npchop=2, nzchop=1)
    p = Params(nr=16, np=4, nz=1, nrchop=16, npchop=2, nzchop=1, is_svv=True,
               svv_cutoff=0.75, svv_target=6e-3, svv_strength=0.4, svv_relax=0.25)
    s = so.Scalar(kit, 'FFF')
    s.e[0, 0, 0] = 1.0
    s.e[15, 0, 0] = 0.1j          # q = 15/15 = 1
    ratio = 0.01 / 1.01
    feedback = min(max(ratio / 6e-3 - 1.0, 0.0), 1.0)
    gain = 0.75 * 0.2 + 0.25 * feedback
    g = so.svv_filter(s, kit, 0.2, p)
    assert abs(g - gain) < 1e-15
    assert s.e[0, 0, 0] == 1.0
    assert abs(s.e[15, 0, 0] - 0.1j * math.exp(-0.4 * gain)) < 1e-16


# ---------------------------------------------------------------------------
# fftreat / smooth (mlegs_scalar_ops.f90:1002-1063, 2149-2175)
# ---------------------------------------------------------------------------
def test_smooth_conserves_and_stays_local():
    """smooth is linear, conserves the sum of the line (every source's
    weights add to one), moves each entry at most one place, and damps an
    alternating line."""
    rng = np.random.default_rng(3)
    for ni in (3, 4, 7, 68):
        a = rng.standard_normal(ni) + 1j * rng.standard_normal(ni)
        out = so.smooth(a)
        assert abs(out.sum() - a.sum()) < 1e-13 * np.abs(a).sum()
        b = rng.standard_normal(ni)
        assert np.allclose(so.smooth(2.0 * a + 3.0 * b), 2.0 * out + 3.0 * so.smooth(b), atol=1e-14, rtol=0)
        for k in range(1, ni - 1):  # an interior source reaches k-1 .. k+1 only
            e = np.zeros(ni)
            e[k] = 1.0
            o = so.smooth(e)
            assert np.all(o[:max(k - 1, 0)] == 0.0) and np.all(o[k + 2:] == 0.0)
            assert abs(o[k - 1] - o[k + 1]) < 1e-16 and o[k] >= 0.5 - 1e-15
    alt = (-1.0) ** np.arange(40)
    assert np.linalg.norm(so.smooth(alt)[1:-3]) < 0.6 * np.linalg.norm(alt[1:-3])
    with pytest.raises(ValueError):
        so.smooth(np.ones(2))


def _fft_kit():
    return _kit(nr=32, np=12, nz=6, nrchop=30, npchop=5, nzchop=3, ell=1.0, zlen=2.5)


def test_fftreat_leaves_a_far_decaying_field_alone():
    """A field whose values past r ~ 2 are below 1e-9 of its size has nothing
    for the far-field treatment to zero or smooth: fftreat returns it, with
    its log coefficient."""
    kit = _fft_kit()
    s = so.Scalar(kit, 'PPP')
    s.e = _sample(kit, lambda r, ph, z: np.exp(-4 * r ** 2) * (1 + r * np.cos(ph) * np.cos(2 * np.pi * z / 2.5)
                                                              + r ** 2 * np.sin(2 * ph)))
    so.trans(s, 'FFF', kit)
    so.chop(s, kit)
    s.ln = 0.3
    so.zeroat1(s, kit)
    t = s.copy()
    so.fftreat(t, kit)
    assert t.space == 'FFF' and t.ln == 0.3
    # what remains is the field's own far tail (2e-8 of it at this resolution)
    assert _rel(t.e, s.e) < 1e-7


def test_fftreat_is_linear_and_acts_on_the_far_field():
    kit = _fft_kit()
    rng = np.random.default_rng(5)
    dims = (kit.nrdim, kit.npdim, kit.nzdim)
    ch = (kit.nrchop, kit.npchop, kit.nzchop, kit.np, kit.nz)
    f1 = so.Scalar(kit, 'FFF')
    f1.e = inputs.transform_field(dims, ch, field_id=21, decay=20.0)
    f1.ln = 0.2
    f2 = so.Scalar(kit, 'FFF')
    f2.e = inputs.transform_field(dims, ch, field_id=22, decay=20.0)
    f2.ln = -0.1
    c = so.Scalar(kit, 'FFF')
    c.e = 2.0 * f1.e - 3.0 * f2.e
    c.ln = 2.0 * f1.ln - 3.0 * f2.ln
    for s in (f1, f2, c):
        so.fftreat(s, kit)
    assert _rel(c.e, 2.0 * f1.e - 3.0 * f2.e) < 1e-12
    assert abs(c.ln - (2.0 * f1.ln - 3.0 * f2.ln)) < 1e-15
    # a field with slowly decaying coefficients has far-field content: changed
    g = so.Scalar(kit, 'FFF')
    g.e = inputs.transform_field(dims, ch, field_id=21, decay=20.0)
    g.ln = 0.2
    assert _rel(f1.e, g.e) > 1e-3
