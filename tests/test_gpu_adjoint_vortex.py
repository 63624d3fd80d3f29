"""Adjoints of the vortex step's operators and the gradient of the vortex run.

The reference has no adjoint.  Each transpose is pinned by the identity that
defines it, <A x, y> = <x, A^T y>, in the Euclidean inner product of the
stored arrays plus the log coefficients, on seeded inputs in the supported
set (real fields in the chopped support without the Nyquist column and
slice), to 1e-12 relative.  Maps through idel2 (vec2tp, advection_rhs) are
checked twice: with chi's cotangent zero on the kappa = 0 slice, to 1e-12,
and in full, to 1e-10 (vec2tp) and 1e-9 (advection_rhs), as idel2's
kappa = 0 Laplacian systems (condition numbers 1e6 to 1e12) amplify the
rounding of both sides.  advection_rhs is quadratic plus
linear in (psi, chi), so its central difference (N(s + d) - N(s - d)) / 2 is
its exact directional derivative: the vector-Jacobian product is pinned
against it.  The checkpointed gradient of the vortex run (bootstrap and AB2/CN
steps) is pinned by central differences along seeded directions.  A
transposed band, a dropped log row, a wrong sign of i m or i kappa, or a
missing calcat / zeroat1 term shows as an O(1) failure.
"""
import math

import numpy as np
import pytest

import _cases as C

pytestmark = pytest.mark.gpu

K2 = dict(nr=32, np=48, nz=1, nrchop=32, npchop=25, nzchop=1, ell=1.0, zlen=1.0,
          visc=5e-3, dt=1e-3, hyperpow=0, hypervisc=0.0, is_svv=False)


def _say(*a):
    print('[case]', *a, flush=True)


@pytest.fixture(scope='module', params=['C1', 'ODD', 'POW2', 'K2'])
def kit(request):
    import mlegs_amd as M
    cfg = {'C1': C.C1, 'ODD': C.ODD, 'POW2': C.POW2, 'K2': K2}[request.param]
    p = C.params(cfg, is_svv=False)
    _say('kit', request.param)
    tfm = M.Kit(p, device=0)
    yield request.param, p, tfm
    tfm.close()


def _ppp(tfm, p, seed, amp=1.0):
    rng = np.random.default_rng(700 + seed)
    e = np.zeros(tfm.dims, dtype=np.complex128)
    e[:p.nr, :p.np // 2, :] = amp * (rng.standard_normal((p.nr, p.np // 2, tfm.dims[2])) +
                                     1j * rng.standard_normal((p.nr, p.np // 2, tfm.dims[2])))
    return e


def _adm(tfm, p, seed, ln=0.0, dealiased=False):
    """A seeded real field in the supported set (chopped, Hermitian, no
    Nyquist column or slice), optionally de-aliased."""
    import mlegs_amd as M
    g = M.Scalar(tfm).upload(C.spectral(tfm.dims, p, seed), space='FFF', ln=ln)
    M.adjoint_project(g, tfm)
    if dealiased:
        M.dealias(g, tfm)
    return g


def _ip(a, b, la=0.0, lb=0.0):
    return float(np.sum(a.real * b.real + a.imag * b.imag)) + la * lb


def _ips(xs, ys):
    """sum over fields of the inner product of the data and the log coefficients."""
    return sum(_ip(x.e, y.e, x.ln, y.ln) for x, y in zip(xs, ys))


def _close(lhs, rhs, tol=1e-12):
    return abs(lhs - rhs) <= tol * max(abs(lhs), abs(rhs), 1e-300)


def _no_k0(tfm, g):
    """g with its kappa = 0 slice zeroed: idel2's ill-conditioned systems
    (kappa = 0) then enter neither side of an identity."""
    import mlegs_amd as M
    e = g.e
    e[:, :, 0] = 0.0
    return M.Scalar(tfm).upload(e, space='FFF', ln=0.0)


def test_tp2vec_adjoint(kit):
    """<tp2vec(psi, chi), y> = <(psi, chi), tp2vec^T y>, with log terms."""
    import mlegs_amd as M
    name, p, tfm = kit
    x = (_adm(tfm, p, 70, 0.3), _adm(tfm, p, 71, -0.2))
    v = [M.Scalar(tfm) for _ in range(3)]
    M.tp2vec(x[0], x[1], *v, tfm)
    y = [_ppp(tfm, p, 72 + i) for i in range(3)]
    lhs = sum(_ip(a.e, b) for a, b in zip(v, y))
    yb = [M.Scalar(tfm).upload(a, space='PPP') for a in y]
    g = (M.Scalar(tfm, 'FFF'), M.Scalar(tfm, 'FFF'))
    M.tp2vec_adjoint(*yb, *g, tfm)
    rhs = _ips(x, g)
    _say('  tp2vec: %.15e %.15e' % (lhs, rhs))
    assert _close(lhs, rhs)


def test_tp2curlvec_adjoint(kit):
    import mlegs_amd as M
    name, p, tfm = kit
    x = (_adm(tfm, p, 80, 0.25), _adm(tfm, p, 81, 0.4))
    w = [M.Scalar(tfm) for _ in range(3)]
    M.tp2curlvec(x[0], x[1], *w, tfm)
    y = [_ppp(tfm, p, 82 + i) for i in range(3)]
    lhs = sum(_ip(a.e, b) for a, b in zip(w, y))
    yb = [M.Scalar(tfm).upload(a, space='PPP') for a in y]
    g = (M.Scalar(tfm, 'FFF'), M.Scalar(tfm, 'FFF'))
    M.tp2curlvec_adjoint(*yb, *g, tfm)
    rhs = _ips(x, g)
    _say('  tp2curlvec: %.15e %.15e' % (lhs, rhs))
    assert _close(lhs, rhs)


def test_vec2tp_adjoint(kit):
    """<vec2tp(v), y> = <v, vec2tp^T y> on the data of (psi, chi) (the log
    coefficients of the result are not part of the map)."""
    import mlegs_amd as M
    name, p, tfm = kit
    x = [_ppp(tfm, p, 90 + i) for i in range(3)]
    xs = [M.Scalar(tfm).upload(a, space='PPP') for a in x]
    out = (M.Scalar(tfm, 'FFF'), M.Scalar(tfm, 'FFF'))
    M.vec2tp(*xs, *out, tfm)
    y0 = (_adm(tfm, p, 93), _adm(tfm, p, 94))
    for y, tol in ((y0, 1e-10), ((y0[0], _no_k0(tfm, y0[1])), 1e-12)):
        lhs = _ip(out[0].e, y[0].e) + _ip(out[1].e, y[1].e)
        g = [M.Scalar(tfm) for _ in range(3)]
        M.vec2tp_adjoint(y[0].copy(), y[1].copy(), *g, tfm)
        rhs = sum(_ip(a, b.e) for a, b in zip(x, g))
        _say('  vec2tp: %.15e %.15e (tol %g)' % (lhs, rhs, tol))
        assert _close(lhs, rhs, tol)


def test_zeroat1_adjoint(kit):
    import mlegs_amd as M
    name, p, tfm = kit
    x = _adm(tfm, p, 100, 0.1)
    y = _adm(tfm, p, 101, 0.2)
    fx = x.copy()
    M.zeroat1(fx, tfm)
    lhs = _ip(fx.e, y.e, fx.ln, y.ln)
    gy = y.copy()
    M.zeroat1_adjoint(gy, tfm)
    rhs = _ip(x.e, gy.e, x.ln, gy.ln)
    _say('  zeroat1: %.15e %.15e' % (lhs, rhs))
    assert _close(lhs, rhs)


def test_febe_adjoint(kit):
    """febe is linear in (s, nl): <febe(s, nl), y> against its cotangents."""
    import mlegs_amd as M
    name, p, tfm = kit
    s, nl = _adm(tfm, p, 110, 0.3), _adm(tfm, p, 111, -0.15)
    y = _adm(tfm, p, 112, 0.45)
    fs = s.copy()
    M.febe(fs, nl, p.dt, tfm)
    lhs = _ip(fs.e, y.e, fs.ln, y.ln)
    gy, gn = y.copy(), M.Scalar(tfm, 'FFF')
    M.febe_adjoint(gy, gn, p.dt, tfm)
    rhs = _ips((s, nl), (gy, gn))
    _say('  febe: %.15e %.15e' % (lhs, rhs))
    assert _close(lhs, rhs)


def test_advection_adjoint(kit):
    """advection_rhs is quadratic plus linear in (psi, chi): the central
    difference (N(s + d) - N(s - d)) / 2 is N'(s) d exactly; its inner product
    with seeded cotangents must equal <d, N'(s)^T y>."""
    import mlegs_amd as M
    name, p, tfm = kit
    s = (_adm(tfm, p, 120, 0.2, True), _adm(tfm, p, 121, -0.1, True))
    d = (_adm(tfm, p, 122, 0.05, True), _adm(tfm, p, 123, 0.07, True))
    uz = M.Scalar(tfm).upload(_ppp(tfm, p, 124, 0.3), space='PPP')
    y = (_adm(tfm, p, 125), _adm(tfm, p, 126))

    def N(sign):
        a, b = s[0].copy(), s[1].copy()
        a.axpby(1.0, d[0], sign)
        b.axpby(1.0, d[1], sign)
        na, nb = M.Scalar(tfm, 'FFF'), M.Scalar(tfm, 'FFF')
        M.advection_rhs(a, b, na, nb, uz, tfm)
        return na.e, nb.e

    plus, minus = N(1.0), N(-1.0)
    for yy, tol in ((y, 1e-9), ((y[0], _no_k0(tfm, y[1])), 1e-12)):
        lhs = sum(_ip(0.5 * (a - b), c.e) for a, b, c in zip(plus, minus, yy))
        g = (M.Scalar(tfm, 'FFF'), M.Scalar(tfm, 'FFF'))
        yb = (yy[0].copy(), yy[1].copy())
        M.advection_adjoint(s[0], s[1], uz, yb[0], yb[1], g[0], g[1], tfm)
        rhs = _ips(d, g)
        _say('  advection: %.15e %.15e (tol %g)' % (lhs, rhs, tol))
        assert _close(lhs, rhs, tol)


def _qvortex(tfm, p, amp, q=4.0):
    """psi, chi of one q-vortex plus a seeded perturbation of amplitude amp."""
    import mlegs_amd as M
    from mlegs_amd import inputs
    out = []
    for which, fid in (('psi', 1), ('chi', 2)):
        s = M.Scalar(tfm).upload(inputs.qvortex_ppp(tfm.x, tfm.r, tfm.dims, p.np, p.nz, p.nr, q, which), space='PPP')
        M.trans(s, 'FFF', tfm)
        M.idelsqp(s, tfm)
        M.zeroat1(s, tfm)
        if amp:
            pert = M.Scalar(tfm).upload(inputs.perturbation(tfm.dims, C.chops(p), amp, field_id=fid), space='FFF')
            s.axpby(1.0, pert, 1.0)
        M.dealias(s, tfm)
        M.zeroat1(s, tfm)
        out.append(s)
    return out


def _chain_fd(cfg, nsteps, ndir, segment):
    import mlegs_amd as M
    from mlegs_amd import adjoint as A, inputs
    p = C.params(cfg, is_svv=False)
    tfm = M.Kit(p, device=0)
    uz = M.Scalar(tfm).upload(0.3 * np.real(_ppp(tfm, p, 130)).astype(np.complex128), space='PPP')
    chain = A.VortexChain(tfm, uz, segment=segment)
    q0 = _qvortex(tfm, p, 1e-2)
    base = chain.forward(*_qvortex(tfm, p, 0.0), nsteps)
    timing = {}
    J, g = chain.gradient(q0[0], q0[1], nsteps, base, timing)
    rows = []
    eps = 1e-6
    for k in range(ndir):
        dirs = [M.Scalar(tfm).upload(inputs.perturbation(tfm.dims, C.chops(p), 1.0, field_id=2000 + 2 * k + i),
                                     space='FFF', ln=0.05 * (k % 3 - 1)) for i in range(2)]
        Jpm = []
        for sgn in (1.0, -1.0):
            a, b = q0[0].copy(), q0[1].copy()
            a.axpby(1.0, dirs[0], sgn * eps)
            b.axpby(1.0, dirs[1], sgn * eps)
            Jpm.append(chain.perturbation_energy(chain.forward(a, b, nsteps), base, grad=False))
        fd = (Jpm[0] - Jpm[1]) / (2 * eps)
        ad = A.inner(g, dirs)
        rows.append((fd, ad, math.sqrt(A.inner(g, g) * A.inner(dirs, dirs))))
    tfm.close()
    return J, rows, timing


def test_vortex_chain_gradient_c1():
    """The bootstrap and 7 AB2/CN steps at C1 with a background axial flow,
    checkpoints every 3 steps: the gradient of the perturbation energy
    against central differences (eps = 1e-6) on 4 seeded directions."""
    J, rows, timing = _chain_fd(C.C1, 7, 4, 3)
    rel = max(abs(fd - ad) / max(abs(fd), abs(ad), 1e-30) for fd, ad, _ in rows)
    scaled = max(abs(fd - ad) / sc for fd, ad, sc in rows)
    _say('  J %.6e, FD vs adjoint: worst rel. %.2e, scaled %.2e' % (J, rel, scaled), timing)
    for fd, ad, _ in rows:
        _say('    fd %.12e adjoint %.12e' % (fd, ad))
    assert rel < 1e-5 and scaled < 1e-8


def test_vortex_chain_gradient_pow2():
    """The same at POW2 (power-of-two phi and z lengths: the forward keeps v
    and w for the reverse steps, and the adjoint transforms skip the
    de-aliased modes): 6 steps, checkpoints every 4, 3 directions."""
    J, rows, timing = _chain_fd(C.POW2, 6, 3, 4)
    rel = max(abs(fd - ad) / max(abs(fd), abs(ad), 1e-30) for fd, ad, _ in rows)
    scaled = max(abs(fd - ad) / sc for fd, ad, sc in rows)
    _say('  J %.6e, FD vs adjoint: worst rel. %.2e, scaled %.2e' % (J, rel, scaled), timing)
    for fd, ad, _ in rows:
        _say('    fd %.12e adjoint %.12e' % (fd, ad))
    assert rel < 1e-5 and scaled < 1e-8
