"""Legendre contractions on FP64 MFMA (v_mfma_f64_16x16x4_f64) and the
per-site precision policy, against the FP64 oracle.

MLX_PREC_FP64_MFMA runs the MFMA kernels of the FP32 path with FP64
operands and FP64 accumulation, so it must agree with the oracle to the FP64
tolerances of the plain path (1e-12 for transforms, 1e-11 to 1e-10 for the
vector-field operators and the step).  A wrong lane layout of the FP64
instruction would show here as O(1) errors.

The site policy (mlx_set_stage_precision): with the projection of vec2tp in
FP64 and the advection term's backward transforms in FP32, the nonlinear
term of the q-vortex step loses the large error that the FP32 projection
puts into its (m = 0, kappa = 0) column, where idel2's kappa = 0 Laplacian
systems amplify it (rel. error of N_chi from ~1e-2 to ~1e-4).
"""
import numpy as np
import pytest

from conftest import phys, rel_err
import _cases as C

pytestmark = pytest.mark.gpu

TOL_T = 1e-12


def _say(*a):
    print('[case]', *a, flush=True)


@pytest.fixture(scope='module', params=['C1', 'ODD', 'POW2'])
def setup(request):
    import mlegs_amd as M
    from oracle import kit as K
    p = C.params(getattr(C, request.param))
    _say('kit', request.param)
    tfm = M.Kit(p, device=0, precision=M.PREC_FP64_MFMA)
    okit = K.Kit3D(p)
    yield request.param, p, tfm, okit
    tfm.close()


def test_trans_fp64_mfma(setup):
    import mlegs_amd as M
    from oracle import scalar_ops as so
    name, p, tfm, okit = setup
    _say('trans fp64 mfma', name)
    e = C.spectral(tfm.dims, p, 11)
    g, o = C.pair(tfm, okit, e, 'FFF', ln=0.37)
    M.trans(g, 'PPP', tfm)
    so.trans(o, 'PPP', okit)
    err_b = rel_err(phys(g.e, p.nr, p.np // 2), phys(o.e, p.nr, p.np // 2))
    M.trans(g, 'FFF', tfm)
    so.trans(o, 'FFF', okit)
    err_f = rel_err(g.e, o.e)
    _say('  backward %.2e forward %.2e' % (err_b, err_f))
    assert err_b < TOL_T and err_f < TOL_T


def test_vec2tp_fp64_mfma(setup):
    import mlegs_amd as M
    from oracle import scalar_ops as so
    name, p, tfm, okit = setup
    _say('vec2tp fp64 mfma', name)
    vs = [C.physical(tfm.dims, p, 60 + i) for i in range(3)]
    g = [M.Scalar(tfm).upload(v, space='PPP') for v in vs]
    o = []
    for v in vs:
        s = so.Scalar(okit, 'PPP')
        s.e = v.copy()
        o.append(s)
    gpsi, gchi = M.Scalar(tfm), M.Scalar(tfm)
    M.vec2tp(g[0], g[1], g[2], gpsi, gchi, tfm)
    opsi, ochi = so.vec2tp(o[0], o[1], o[2], okit)
    ep, ec = rel_err(gpsi.e, opsi.e), rel_err(gchi.e, ochi.e)
    _say('  psi %.2e chi %.2e' % (ep, ec))
    assert ep < 1e-11 and ec < 1e-10


def test_advection_rhs_fp64_mfma(setup):
    """The step's nonlinear term with the de-aliasing restrictions and the
    fused cross-product transform, FP64 MFMA contractions."""
    import mlegs_amd as M
    from oracle import scalar_ops as so, vortex as V
    name, p, tfm, okit = setup
    _say('advection_rhs fp64 mfma', name)
    (gpsi, opsi), (gchi, ochi) = C.qvortex(tfm, okit, p, amp=0.5)
    ub = C.physical(tfm.dims, p, 31)
    guz = M.Scalar(tfm).upload(ub, space='PPP')
    ouz = so.Scalar(okit, 'PPP')
    ouz.e = ub.copy()
    gnp, gnc = M.Scalar(tfm), M.Scalar(tfm)
    M.advection_rhs(gpsi, gchi, gnp, gnc, guz, tfm)
    onp, onc = V.advection_rhs(opsi, ochi, ouz, okit)
    ep, ec = rel_err(gnp.e, onp.e), rel_err(gnc.e, onc.e)
    _say('  vs oracle: nlpsi %.2e nlchi %.2e' % (ep, ec))
    assert ep < 1e-10 and ec < 1e-10


def test_advection_rhs_fused_matches_composed(setup):
    """The fused path of the step's nonlinear term (both tp2vec front passes
    in one kernel, the de-aliasing restrictions, v x w inside the forward phi
    transforms) against the same operators composed from the public entry
    points (tp2vec, tp2curlvec, vecprod, vec2tp, dealias), both on FP64 MFMA:
    the same arithmetic, so agreement to rounding."""
    import mlegs_amd as M
    from oracle import scalar_ops as so
    name, p, tfm, okit = setup
    _say('fused vs composed fp64 mfma', name)
    (gpsi, _), (gchi, _) = C.qvortex(tfm, okit, p, amp=0.5)
    guz = M.Scalar(tfm).upload(C.physical(tfm.dims, p, 31), space='PPP')
    gnp, gnc = M.Scalar(tfm), M.Scalar(tfm)
    M.advection_rhs(gpsi, gchi, gnp, gnc, guz, tfm)
    vr, vp, vz = M.Scalar(tfm), M.Scalar(tfm), M.Scalar(tfm)
    wr, wp, wz = M.Scalar(tfm), M.Scalar(tfm), M.Scalar(tfm)
    M.tp2vec(gpsi, gchi, vr, vp, vz, tfm)
    M.tp2curlvec(gpsi, gchi, wr, wp, wz, tfm)
    vz.axpby(1.0, guz, 1.0)
    M.vecprod(vr, vp, vz, wr, wp, wz, tfm)
    cp, cc = M.Scalar(tfm), M.Scalar(tfm)
    M.vec2tp(vr, vp, vz, cp, cc, tfm)
    M.dealias(cp, tfm)
    M.dealias(cc, tfm)
    fp, fc = rel_err(gnp.e, cp.e), rel_err(gnc.e, cc.e)
    _say('  nlpsi %.2e nlchi %.2e' % (fp, fc))
    assert fp < 1e-12 and fc < 1e-10


def test_vortex_fp64_mfma(setup):
    import mlegs_amd as M
    from oracle import scalar_ops as so, vortex as V
    name, p, tfm, okit = setup
    _say('vortex fp64 mfma', name)
    (gpsi, opsi), (gchi, ochi) = C.qvortex(tfm, okit, p)
    run = M.VortexRun(tfm, gpsi, gchi, M.Scalar(tfm))
    orun = V.VortexRun(okit, p, opsi, ochi, so.Scalar(okit, 'PPP'))
    run.bootstrap()
    orun.bootstrap()
    run.step(3)
    for _ in range(3):
        orun.step()
    for nm, ob in (('psi', orun.psi), ('chi', orun.chi)):
        err = rel_err(run.field(nm).e, ob.e)
        _say('  %s %.2e' % (nm, err))
        assert err < 1e-10, nm


def _nl_errors(tfm, okit, p, stages):
    """Errors of the q-vortex nonlinear term against the oracle with the
    given site precisions (context precision FP32)."""
    import mlegs_amd as M
    from oracle import vortex as V, scalar_ops as so
    # the inputs are built at the fixture's FP64 MFMA precision
    (gpsi, opsi), (gchi, ochi) = C.qvortex(tfm, okit, p, amp=1e-4)
    tfm.set_precision(M.PREC_FP32)
    for st, pr in stages.items():
        tfm.set_stage_precision(st, pr)
    try:
        gnp, gnc = M.Scalar(tfm), M.Scalar(tfm)
        M.advection_rhs(gpsi, gchi, gnp, gnc, M.Scalar(tfm), tfm)
        onp, onc = V.advection_rhs(opsi, ochi, so.Scalar(okit, 'PPP'), okit)
        return rel_err(gnp.e, onp.e), rel_err(gnc.e, onc.e)
    finally:
        for st in stages:
            tfm.set_stage_precision(st, -1)
        tfm.set_precision(M.PREC_FP64_MFMA)


def test_stage_precision_projection_fp64(setup):
    import mlegs_amd as M
    name, p, tfm, okit = setup
    _say('stage precision', name)
    F = M.PREC_FP64_MFMA
    ep32, ec32 = _nl_errors(tfm, okit, p, {})
    ep0, ec0 = _nl_errors(tfm, okit, p, {M.STAGE_PROJ0: F})
    epm, ecm = _nl_errors(tfm, okit, p, {M.STAGE_PROJ: F, M.STAGE_PROJ0: F})
    ep64, ec64 = _nl_errors(tfm, okit, p, {M.STAGE_PROJ: F, M.STAGE_PROJ0: F, M.STAGE_SYNTH: F, M.STAGE_UP: F})
    epx, ecx = _nl_errors(tfm, okit, p, {M.STAGE_PROJ: F, M.STAGE_SYNTH: F, M.STAGE_UP: F})
    _say('  fp32: %.2e %.2e; m=0 fp64: %.2e %.2e; proj fp64: %.2e %.2e; all fp64: %.2e %.2e; '
         'all but m=0 fp64: %.2e %.2e' % (ep32, ec32, ep0, ec0, epm, ecm, ep64, ec64, epx, ecx))
    # every site in FP64 through the policy is the FP64 MFMA path
    assert ep64 < 1e-10 and ec64 < 1e-10
    # FP32 rounding is visible, and the projection's site settings take effect
    assert ec32 > 1e-9
    assert ec0 != ec32 and ecm != ec0 and ecx > 1e-9
    for a, b in ((ep0, ec0), (epm, ecm)):
        assert a < 1e-5 and b < 1e-3


def test_stage_precision_rejects_bad_arguments():
    import mlegs_amd as M
    p = C.params(C.SMALL)
    tfm = M.Kit(p, device=0)
    with pytest.raises(M.MlxError):
        tfm.set_stage_precision(7, M.PREC_FP32)
    with pytest.raises(M.MlxError):
        tfm.set_stage_precision(M.STAGE_PROJ, 9)
    tfm.close()
