"""The oracle's transform kit against values fixed by mathematics.

Golden vectors (tests/golden/kit_golden.npz, written by make_golden.py with
mpmath from an explicit-sum formula, independent of the oracle's recurrence)
pin gauss_legendre (mlegs_spectfm_init.f90:181-214), leg_lognorm (:218-250)
and leg_tbl (:254-300); quadrature sum rules and orthonormality pin them
again without any fixture.
"""
import math
import os

import numpy as np
import pytest

from oracle import kit as K
from mlegs_amd.params import Params

GOLD = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'kit_golden.npz'))


@pytest.mark.parametrize('nr', [32, 256])
def test_gauss_nodes_vs_mpmath(nr):
    x, w = K.gauss_legendre(nr)
    assert np.all(np.diff(x) > 0) and x[0] < 0 < x[-1]
    np.testing.assert_allclose(x, GOLD['gl%d_x' % nr], rtol=0, atol=2.3e-16)
    np.testing.assert_allclose(w, GOLD['gl%d_w' % nr], rtol=2e-12, atol=0)


@pytest.mark.parametrize('nr', [16, 32, 64, 256])
def test_gauss_sum_rules(nr):
    """sum w x^(2k) = 2/(2k+1), exact for degree 2k <= 2nr-1; odd moments vanish."""
    x, w = K.gauss_legendre(nr)
    for k in range(0, min(nr, 40)):
        assert abs(np.sum(w * x ** (2 * k)) - 2.0 / (2 * k + 1)) < 2e-13 * 2.0 / (2 * k + 1) + 1e-15, k
        assert abs(np.sum(w * x ** (2 * k + 1))) < 1e-14
    xn, wn = np.polynomial.legendre.leggauss(nr)
    np.testing.assert_allclose(x, xn, atol=3e-16, rtol=0)


def _colerr(a, b, floor=1e-300):
    """Worst relative error per table column (entries over the column's largest).

    A column of P_NR^0 at the NR Gauss nodes is zero up to rounding; the floor
    keeps such a column on an absolute scale."""
    scale = np.maximum(np.abs(b).max(axis=0, keepdims=True), floor)
    return float((np.abs(a - b) / scale).max())


def test_legendre_table_nr32_vs_mpmath():
    x, _ = K.gauss_legendre(32)
    t = K.leg_tbl(x[:16], 46, list(range(9)))
    assert _colerr(t, GOLD['tbl32'], floor=0.1) < 2e-15


def test_legendre_table_nr256_spot_vs_mpmath():
    x, _ = K.gauss_legendre(256)
    ms = [int(v) for v in GOLD['spot256_m']]
    t = K.leg_tbl(x[:128], 270, ms)
    for a in range(len(ms)):
        for b, off in enumerate(GOLD['spot256_n']):
            ref = GOLD['spot256'][:, a, b]
            assert _colerr(t[:, off, a][:, None], ref[:, None]) < 2e-13, (ms[a], off)


def test_lognorm_vs_mpmath():
    ln = K.leg_lognorm(46, list(range(9)))
    np.testing.assert_allclose(ln, GOLD['lognorm32'], rtol=1e-14, atol=1e-14)
    ms = [int(v) for v in GOLD['lognorm256_m']]
    ln = K.leg_lognorm(270, ms)
    np.testing.assert_allclose(ln, GOLD['lognorm256'], rtol=1e-14, atol=0)


def test_values_near_poles_vs_mpmath():
    a1 = K.leg_tbl(np.array([1.0 - 1e-15]), 256, [0])[0, :, 0]
    a0 = K.leg_tbl(np.array([-1.0 + 1e-15]), 256, [0])[0, :, 0]
    np.testing.assert_allclose(a1, GOLD['at1_256'], rtol=1e-14)
    np.testing.assert_allclose(a0, GOLD['at0_256'], rtol=1e-14)
    # P_n^0(+-1) = (+-1)^n sqrt((2n+1)/2) up to the 1e-15 offset
    n = np.arange(256)
    np.testing.assert_allclose(a1, np.sqrt((2 * n + 1) / 2.0), rtol=1e-10)
    np.testing.assert_allclose(a0, (-1.0) ** n * np.sqrt((2 * n + 1) / 2.0), rtol=1e-10)


@pytest.mark.parametrize('m', [0, 1, 5, 20])
def test_legendre_orthonormal_by_quadrature(m):
    """sum_i w_i P_n^m(x_i) P_n'^m(x_i) = delta_nn' while n + n' <= 2 nr - 1."""
    nr = 64
    x, w = K.gauss_legendre(nr)
    ne = nr - m
    half = K.leg_tbl(x[: nr // 2], ne, [m])[:, :, 0]
    n = m + np.arange(ne)
    # P_n^m(-x) = (-1)^(n+m) P_n^m(x)
    full = np.concatenate([half, half[::-1] * ((-1.0) ** (n + m))[None, :]], axis=0)
    g = full.T @ (w[:, None] * full)
    ok = (n[:, None] + n[None, :]) <= 2 * nr - 1
    assert np.abs((g - np.eye(ne))[ok]).max() < 1e-13


def test_kit_geometry():
    """r = L sqrt((1+x)/(1-x)), ln = -log(1-x), ak and index sets (mlegs_spectfm_init.f90:111-139)."""
    p = Params(nr=16, np=12, nz=10, nrchop=14, npchop=5, nzchop=4, ell=1.7, zlen=3.0)
    k = K.Kit3D(p)
    np.testing.assert_allclose(k.r, 1.7 * np.sqrt((1 + k.x) / (1 - k.x)), rtol=1e-15)
    # the map sends x -> r with r(x_i) r(-x_i) = L^2
    np.testing.assert_allclose(k.r * k.r[::-1], 1.7 ** 2, rtol=1e-13)
    np.testing.assert_allclose(k.lnx, -np.log(1 - k.x), rtol=1e-15)
    assert (k.nrdim, k.npdim, k.nzdim) == (19, 7, 10)
    assert (k.chopzl, k.chopzu) == (4, 8)
    kap = np.array([0, 1, 2, 3, 4, 5, -4, -3, -2, -1])
    np.testing.assert_allclose(k.ak, 2 * math.pi / 3.0 * kap, rtol=1e-15)
    assert list(k.chops) == [14, 13, 12, 11, 10]


@pytest.mark.parametrize('bad', [
    dict(nr=15), dict(np=14), dict(nz=14), dict(nrchop=40), dict(npchop=10), dict(nzchop=7)])
def test_kit_validation_rejects(bad):
    """tfm_kit_init's checks (mlegs_spectfm_init.f90:35-62)."""
    d = dict(nr=32, np=16, nz=8, nrchop=32, npchop=9, nzchop=5, ell=1.0, zlen=1.0)
    d.update(bad)
    with pytest.raises(ValueError):
        K.Kit3D(Params(**d))
