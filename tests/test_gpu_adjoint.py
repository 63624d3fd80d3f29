"""Adjoint operators (mlx_*_adjoint) and reverse-mode gradients.

The reference has no adjoint, so nothing here is compared with it.  Each
transpose is pinned by the identity that defines it, <A x, y> = <x, A^T y>
in the Euclidean inner product of the stored arrays (plus the log
coefficient), on seeded inputs, to 1e-12 relative; the checkpointed reverse
sweep of the nonlinear model step is pinned by central differences
(eps = 1e-6) along seeded directions; and the forward transform is run
through torch.autograd.gradcheck as a test-only autograd.Function over the C
ABI.  A transposed operand, a dropped weight or log-term row, or a wrong
phi-mode factor shows as an O(1) failure of the first identity.
"""
import ctypes
import math

import numpy as np
import pytest

import _cases as C

pytestmark = pytest.mark.gpu

C5 = dict(nr=128, np=64, nz=64, nrchop=128, npchop=33, nzchop=33, ell=1.0, zlen=2 * math.pi,
          visc=1e-3, dt=1e-3, hyperpow=0, hypervisc=0.0, is_svv=False)
K2 = dict(nr=32, np=48, nz=1, nrchop=32, npchop=25, nzchop=1, ell=1.0, zlen=1.0,
          visc=5e-3, dt=1e-3, hyperpow=0, hypervisc=0.0, is_svv=False)


def _say(*a):
    print('[case]', *a, flush=True)


@pytest.fixture(scope='module', params=['C1', 'ODD', 'K2'])
def kit(request):
    import mlegs_amd as M
    cfg = {'C1': C.C1, 'ODD': C.ODD, 'K2': K2}[request.param]
    p = C.params(cfg, is_svv=False)
    _say('kit', request.param)
    tfm = M.Kit(p, device=0)
    yield request.param, p, tfm
    tfm.close()


def _ppp(tfm, p, seed, amp=1.0):
    rng = np.random.default_rng(900 + seed)
    e = np.zeros(tfm.dims, dtype=np.complex128)
    e[:p.nr, :p.np // 2, :] = amp * (rng.standard_normal((p.nr, p.np // 2, tfm.dims[2])) +
                                     1j * rng.standard_normal((p.nr, p.np // 2, tfm.dims[2])))
    return e


def _ip(a, b, la=0.0, lb=0.0):
    return float(np.sum(a.real * b.real + a.imag * b.imag)) + la * lb


def test_trans_backward_adjoint(kit):
    """<B x, y> = <x, B^T y>, B = trans FFF -> PPP, x real (Hermitian) with a log term."""
    import mlegs_amd as M
    name, p, tfm = kit
    x = C.spectral(tfm.dims, p, 5)
    y = _ppp(tfm, p, 6)
    gx = M.Scalar(tfm).upload(x, space='FFF', ln=0.37)
    M.trans(gx, 'PPP', tfm)
    lhs = _ip(gx.e, y)
    gy = M.Scalar(tfm).upload(y, space='PPP', ln=0.0)
    M.trans_adjoint(gy, 'FFF', tfm)
    assert gy.space == 'FFF'
    rhs = _ip(x, gy.e, 0.37, gy.ln)
    _say('  B: <Bx,y> %.15e <x,BTy> %.15e' % (lhs, rhs))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_trans_forward_adjoint(kit):
    """<F x, y> = <x, F^T y>, F = trans PPP -> FFF, x real samples with a log
    term, y any complex FFF array."""
    import mlegs_amd as M
    name, p, tfm = kit
    x = _ppp(tfm, p, 7)
    rng = np.random.default_rng(77)
    y = rng.standard_normal(tfm.dims) + 1j * rng.standard_normal(tfm.dims)
    gx = M.Scalar(tfm).upload(x, space='PPP', ln=-0.21)
    M.trans(gx, 'FFF', tfm)
    lhs = _ip(gx.e, y, gx.ln, 0.8)
    gy = M.Scalar(tfm).upload(y, space='FFF', ln=0.8)
    M.trans_adjoint(gy, 'PPP', tfm)
    assert gy.space == 'PPP'
    rhs = _ip(x, gy.e, -0.21, gy.ln)
    _say('  F: <Fx,y> %.15e <x,FTy> %.15e' % (lhs, rhs))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_vecprod_adjoint(kit):
    """d(v x u) = dv x u + v x du against (v_bar, u_bar) = (u x a, a x v)."""
    import mlegs_amd as M
    name, p, tfm = kit
    v = [_ppp(tfm, p, 10 + i) for i in range(3)]
    u = [_ppp(tfm, p, 20 + i) for i in range(3)]
    dv = [_ppp(tfm, p, 30 + i) for i in range(3)]
    du = [_ppp(tfm, p, 40 + i) for i in range(3)]
    a = [_ppp(tfm, p, 50 + i) for i in range(3)]
    up = lambda arrs: [M.Scalar(tfm).upload(x, space='PPP') for x in arrs]
    # tangent: dv x u + v x du
    t1, t2 = up(dv), up(v)
    M.vecprod(*t1, *up(u), tfm)
    M.vecprod(*t2, *up(du), tfm)
    lhs = sum(_ip(x.e + y.e, z) for x, y, z in zip(t1, t2, a))
    ga = up(a)
    ub = up([np.zeros(tfm.dims, dtype=np.complex128)] * 3)
    M.vecprod_adjoint(*ga, *up(v), *up(u), *ub, tfm)
    rhs = sum(_ip(x, g.e) for x, g in zip(dv, ga)) + sum(_ip(x, g.e) for x, g in zip(du, ub))
    _say('  vecprod: %.15e %.15e' % (lhs, rhs))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_abcn_adjoint(kit):
    """abcn is linear in (s, nl, nl_p): <abcn(s, nl, nl_p), y> against the
    cotangents, every field in the chopped support, with log terms."""
    import mlegs_amd as M
    name, p, tfm = kit
    dt = p.dt
    s, nl, nlp = (C.spectral(tfm.dims, p, 60 + i) for i in range(3))
    ls, lnl, lnlp = 0.3, -0.2, 0.15
    y = C.spectral(tfm.dims, p, 64)
    ly = 0.45
    gs = M.Scalar(tfm).upload(s, space='FFF', ln=ls)
    gn = M.Scalar(tfm).upload(nl, space='FFF', ln=lnl)
    gnp = M.Scalar(tfm).upload(nlp, space='FFF', ln=lnlp)
    M.abcn(gs, gs.copy(), gn, gnp, dt, tfm)
    lhs = _ip(gs.e, y, gs.ln, ly)
    gy = M.Scalar(tfm).upload(y, space='FFF', ln=ly)
    a, b = M.Scalar(tfm, 'FFF'), M.Scalar(tfm, 'FFF')
    M.abcn_adjoint(gy, a, b, dt, tfm)
    rhs = _ip(s, gy.e, ls, gy.ln) + _ip(nl, a.e, lnl, a.ln) + _ip(nlp, b.e, lnlp, b.ln)
    _say('  abcn: %.15e %.15e' % (lhs, rhs))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_abcn_adjoint_rejects_hyperviscosity():
    import mlegs_amd as M
    p = C.params(C.SMALL, hyperpow=4, hypervisc=1e-6, is_svv=False)
    tfm = M.Kit(p, device=0)
    a, b, c = (M.Scalar(tfm, 'FFF') for _ in range(3))
    with pytest.raises(M.MlxError):
        M.abcn_adjoint(a, b, c, 1e-3, tfm)
    with pytest.raises(M.MlxError):
        M.trans_adjoint(a, 'FFF', tfm)  # a cotangent of FFF -> PPP must be in PPP
    tfm.close()


def _chain_case(cfg, nsteps, ndir, seed0, segment=4):
    import mlegs_amd as M
    from mlegs_amd import adjoint as A
    p = C.params(cfg, is_svv=False)
    tfm = M.Kit(p, device=0)
    chops = C.chops(p)
    from mlegs_amd import inputs
    x0 = [M.Scalar(tfm).upload(inputs.perturbation(tfm.dims, chops, 0.5, field_id=seed0 + i), space='FFF',
                               ln=0.1 * (i + 1)) for i in range(3)]
    chain = A.ModelChain(tfm, p.dt, segment=segment)
    timing = {}
    J, g = chain.gradient(x0, nsteps, timing)
    rows = []
    eps = 1e-6
    for d in range(ndir):
        dirs = [M.Scalar(tfm).upload(inputs.perturbation(tfm.dims, chops, 1.0, field_id=1000 + 3 * d + i),
                                     space='FFF', ln=0.05 * (d % 5 - 2)) for i in range(3)]
        plus = [x.copy() for x in x0]
        minus = [x.copy() for x in x0]
        for a, b, dd in zip(plus, minus, dirs):
            a.axpby(1.0, dd, eps)
            b.axpby(1.0, dd, -eps)
        fd = (chain.objective(chain.forward(plus, nsteps)) - chain.objective(chain.forward(minus, nsteps))) / (2 * eps)
        ad = A.inner(g, dirs)
        rows.append((fd, ad, math.sqrt(A.inner(g, g) * A.inner(dirs, dirs))))
    tfm.close()
    return J, rows, timing


def _fd_errors(rows):
    """Worst |fd - adjoint| relative to the derivative and to the Cauchy-Schwarz
    scale |g| |d| (central differences at eps = 1e-6 carry ~1e-10 |J| / eps
    of rounding, which dominates along directions with a small derivative)."""
    rel = max(abs(fd - ad) / max(abs(fd), abs(ad), 1e-30) for fd, ad, _ in rows)
    scaled = max(abs(fd - ad) / sc for fd, ad, sc in rows)
    return rel, scaled


def test_model_chain_gradient_c1():
    J, rows, timing = _chain_case(C.C1, 9, 6, 7)
    rel, scaled = _fd_errors(rows)
    _say('  J %.6e, FD vs adjoint: worst rel. %.2e, scaled %.2e' % (J, rel, scaled), timing)
    for fd, ad, _ in rows:
        _say('    fd %.12e adjoint %.12e' % (fd, ad))
    assert rel < 1e-5 and scaled < 1e-8


def test_model_chain_gradient_c5():
    """The C5 size (128 x 64 x 64), 10 steps, 32 seeded directions."""
    J, rows, timing = _chain_case(C5, 10, 32, 11, segment=5)
    rel, scaled = _fd_errors(rows)
    _say('  J %.6e, over %d directions: worst rel. %.2e, scaled %.2e' % (J, len(rows), rel, scaled), timing)
    assert rel < 1e-5 and scaled < 1e-8


def test_autograd_function_gradcheck():
    """trans PPP -> FFF as a torch.autograd.Function whose backward is
    mlx_trans_adjoint; torch.autograd.gradcheck builds the Jacobian by finite
    differences on every real input (a small kit)."""
    torch = pytest.importorskip('torch')
    import mlegs_amd as M
    p = C.params(dict(nr=8, np=8, nz=4, nrchop=8, npchop=5, nzchop=3, ell=1.3, zlen=2.0, visc=1e-3, dt=1e-3,
                      hyperpow=0, hypervisc=0.0, is_svv=False))
    tfm = M.Kit(p, device=0)
    dims = tfm.dims

    class Forward(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x):
            e = (x[..., 0] + 1j * x[..., 1]).detach().numpy()
            g = M.Scalar(tfm).upload(e, space='PPP')
            M.trans(g, 'FFF', tfm)
            o = g.e
            return torch.from_numpy(np.stack([o.real, o.imag], axis=-1))

        @staticmethod
        def backward(ctx, gout):
            e = (gout[..., 0] + 1j * gout[..., 1]).detach().numpy()
            g = M.Scalar(tfm).upload(e, space='FFF')
            M.trans_adjoint(g, 'PPP', tfm)
            o = g.e
            return torch.from_numpy(np.stack([o.real, o.imag], axis=-1))

    rng = np.random.default_rng(3)
    x = np.zeros(dims + (2,))
    x[:p.nr, :p.np // 2] = rng.standard_normal((p.nr, p.np // 2, dims[2], 2))
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    ok = torch.autograd.gradcheck(Forward.apply, (xt,), eps=1e-6, atol=1e-9, rtol=1e-7)
    tfm.close()
    assert ok
