"""Banded solves (ihelm, idel2, idel2 with a preset ln) at every radial size
class of the segmented solve kernel, through the C ABI (marked gpu).

The solve splits a system's rows over 32 lane segments of S rows, S chosen
from {4, 8, 9, 12, 16, 17} by the number of radial rows, and runs systems
with and without row interchanges in separate kernels.  Each case below lands
in a different S, partly filled last segments included, with nzchop = nz/2 + 1
(the duplicated Nyquist slice, solved twice) and a radial chopping offset of
3 on one case.  The oracle solves every system with zgbsv
(mlegs_bndmat_ops.f90:369-448 semantics).  FP64 on both sides.
"""
import math

import numpy as np
import pytest

from conftest import rel_err
import _cases as C

pytestmark = pytest.mark.gpu

# (nr, nrchop, offset): rows per system nrc = nrchop + offset
SIZES = [
    (32, 30, 0),     # S = 4, 8 segments in use
    (250, 250, 0),   # S = 8, last segment partly filled
    (256, 254, 3),   # S = 9 (nrc = 257), offset rows
    (300, 290, 0),   # S = 12
    (420, 420, 0),   # S = 16
    (530, 530, 0),   # S = 17
]


def _say(*a):
    print('[case]', *a, flush=True)


@pytest.fixture(scope='module', params=SIZES, ids=['nr%d' % s[0] for s in SIZES])
def setup(request):
    import mlegs_amd as M
    from oracle import kit as K
    nr, nrchop, off = request.param
    p = C.params(dict(nr=nr, np=8, nz=8, nrchop=nrchop, npchop=5, nzchop=5, ell=1.3, zlen=3.0,
                      visc=1e-3, dt=1e-3, hyperpow=0, hypervisc=0.0, is_svv=False))
    _say('kit', nr, nrchop, off)
    tfm = M.Kit(p, device=0)
    okit = K.Kit3D(p)
    yield p, off, tfm, okit
    tfm.close()


def _rhs(tfm, p, off, seed):
    """Seeded coefficients, decaying with the radial index as a smooth
    field's do, on the rows nrchop + offset - m of every active column."""
    rng = np.random.default_rng(seed)
    e = np.zeros(tfm.dims, dtype=np.complex128)
    for m in range(p.npchop):
        n = p.nrchop + off - m
        decay = np.exp(-8.0 * np.arange(n) / n)[:, None]
        e[:n, m, :] = decay * (rng.standard_normal((n, p.nz)) + 1j * rng.standard_normal((n, p.nz)))
    return e


@pytest.mark.parametrize('op', ['ihelm', 'ihelm_stiff', 'idel2', 'idel2_preln'])
def test_banded_solve_sizes(setup, op):
    import mlegs_amd as M
    from oracle import scalar_ops as so
    p, off, tfm, okit = setup
    _say('solve', op, p.nr, p.nrchop, off)
    e = _rhs(tfm, p, off, 77)
    g, o = C.pair(tfm, okit, e, 'FFF', ln=0.3, offsets=(off, 0, 0))
    if op == 'ihelm':
        M.ihelm(g, -2.5, tfm); so.ihelm(o, -2.5, okit)
    elif op == 'ihelm_stiff':  # the Crank-Nicolson operator's scale: no interchanges
        M.ihelm(g, -4e6, tfm); so.ihelm(o, -4e6, okit)
    elif op == 'idel2':
        M.idel2(g, tfm); so.idel2_proln(o, okit)
    else:
        M.idel2(g, tfm, preln=0.5); so.idel2_preln(o, okit, 0.5)
    # The Laplacian's kappa = 0 systems (idel2: alpha = 0) have condition
    # numbers from 1e6 up to ~1e12 (the bordered (0,0) one) at these sizes.
    # Two backward-stable FP64 solves with different rounding (LAPACK's
    # zgbtrf/zgbtrs and this factorisation with fused multiply-adds) agree
    # there only to about cond * eps, so those systems are checked by their
    # backward error ||A x - b|| / (||A|| ||x|| + ||b||), which a correct solve
    # keeps at rounding level (below n eps) whatever the conditioning, and by
    # the forward bound cond * eps.  Every other system is compared at FP64
    # tolerance.
    ge, oe = g.e, o.e
    strict = np.ones(ge.shape[1:], dtype=bool)
    if op.startswith('idel2'):
        for m in range(p.npchop):
            for k in range(p.nz):
                if okit.ak[k] != 0.0:
                    continue
                strict[m, k] = False
                nn = p.nrchop + off - m
                rhs = e[:nn, m, k]
                if m == 0 and op == 'idel2_preln':
                    A, rhs = _preln_system(okit, p, off, rhs, 0.5)
                else:
                    A = _dense_system(okit, p, off, m, k, bordered=(m == 0))
                err = rel_err(ge[:nn, m, k], oe[:nn, m, k])
                bound = 64 * np.finfo(float).eps * np.linalg.cond(A)
                x = ge[:nn, m, k]
                eta = np.linalg.norm(A @ x - rhs) / (np.linalg.norm(A, 2) * np.linalg.norm(x) + np.linalg.norm(rhs))
                _say('kappa=0 system', m, k, 'err %.2e forward bound %.2e backward error %.2e' % (err, bound, eta))
                assert err < bound
                assert eta < nn * np.finfo(float).eps
    tol = 1e-11 if op.startswith('idel2') else 1e-12
    err = rel_err(ge[:, strict], oe[:, strict])
    _say('other systems err %.2e' % err)
    assert err < tol
    # ln follows the bordered system's first coefficient
    if op == 'idel2':
        assert abs(g.ln - o.ln) <= 64 * np.finfo(float).eps * np.linalg.cond(
            _dense_system(okit, p, off, 0, 0, bordered=True)) * abs(o.ln)
    else:
        assert abs(g.ln - o.ln) <= 1e-12 * max(1.0, abs(o.ln))


def _dense_system(okit, p, off, m, k, bordered):
    """The (del2 - kappa^2) band matrix of column m, slice k, as a dense array
    (plus the log-term border of idel2's (0,0) system)."""
    from oracle.scalar_ops import _ab
    nn = p.nrchop + off - m
    ab = _ab(okit.band_del2h(okit.m[m], nn), -(okit.ak[k] ** 2))
    if bordered:
        ell2 = okit.ell ** 2
        ln = okit.lognorm[:, 0]
        ab[2, 0] += 4.0 / 3.0 / ell2
        ab[3, 0] -= 2.0 / ell2 * np.exp(ln[0] - ln[1])
        ab[4, 0] += 2.0 / 3.0 / ell2 * np.exp(ln[0] - ln[2])
    A = np.zeros((nn, nn))
    for j in range(nn):
        for i in range(max(0, j - 2), min(nn, j + 3)):
            A[i, j] = ab[2 + i - j, j]
    return A


def _preln_system(okit, p, off, b, preln):
    """idel2_preln's (0,0) system: the first equation prescribes the log
    coefficient, the Laplacian's rows shift down by one
    (mlegs_scalar_ops.f90:575-669)."""
    nn = p.nrchop + off
    ell2 = okit.ell ** 2
    ln = okit.lognorm[:, 0]
    A = okit.leg_del2(okit.m[0], okit.ak[0], nn)
    P = np.zeros((nn, nn))
    P[1:nn, :] = A[0:nn - 1, :]
    P[0, 0] = 1.0
    P[1, 0] += 4.0 / 3.0 / ell2
    P[2, 0] -= 2.0 / 1.0 / ell2 * np.exp(ln[0] - ln[1])
    P[3, 0] += 2.0 / 3.0 / ell2 * np.exp(ln[0] - ln[2])
    rhs = np.empty_like(b)
    rhs[1:] = b[:-1]
    rhs[0] = preln / np.exp(ln[0])
    return P, rhs
