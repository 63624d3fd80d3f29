"""The benchmark's workload size, NR = NP = NZ = 256 (NRCHOP = 256,
NPCHOP = NZCHOP = 129), q-vortex + seeded 1e-4 perturbation, against the
oracle: the nonlinear term and one AB2/CN step from it
(vortical_flow_3d.f90:160-181, 353-395) in the plain FP64 path and on FP64
MFMA, and the mixed preset's nonlinear term at its per-field gates.  The
oracle takes several seconds per nonlinear term at this size.

Tolerances (FP64): N_psi and the step 1e-10; N_chi 1e-9.  N_chi is idel2 of
the projection, and idel2's kappa = 0 Laplacian systems at nr = 256 have
condition numbers far above 1e6: two backward-stable solves of them agree
only to about cond * eps, which puts N_chi's difference at 2e-10 of its
norm (its (0, 0) column, 1e-4 of the products it is formed from).
"""
import math

import pytest

from conftest import rel_err
import _cases as C

pytestmark = pytest.mark.gpu


def _say(*a):
    print('[case]', *a, flush=True)


@pytest.fixture(scope='module')
def c3():
    import mlegs_amd as M
    from oracle import kit as K, scalar_ops as so, vortex as V
    p = M.Params(nr=256, np=256, nz=256, nrchop=256, npchop=129, nzchop=129, ell=1.0, zlen=2 * math.pi,
                 visc=5e-4, dt=5e-4, hyperpow=0, is_svv=True)
    _say('256^3 kit and oracle')
    tfm = M.Kit(p, device=0, precision=M.PREC_FP64)
    okit = K.Kit3D(p)
    (gpsi, opsi), (gchi, ochi) = C.qvortex(tfm, okit, p, amp=1e-4)
    _say('oracle nonlinear term')
    onp, onc = V.advection_rhs(opsi, ochi, so.Scalar(okit, 'PPP'), okit)
    yield p, tfm, okit, (gpsi, gchi), (opsi, ochi), (onp, onc)
    tfm.close()


def _gpu_nl(tfm, gpsi, gchi):
    import mlegs_amd as M
    gp, gc = gpsi.copy(), gchi.copy()
    gnp, gnc = M.Scalar(tfm), M.Scalar(tfm)
    M.advection_rhs(gp, gc, gnp, gnc, M.Scalar(tfm), tfm)
    return gp, gc, gnp, gnc


@pytest.mark.parametrize('prec', ['fp64', 'fp64m'])
def test_c3_nonlinear_term_and_step_fp64(c3, prec):
    import mlegs_amd as M
    from oracle import scalar_ops as so
    p, tfm, okit, (gpsi, gchi), (opsi, ochi), (onp, onc) = c3
    tfm.set_precision(M.PREC_FP64 if prec == 'fp64' else M.PREC_FP64_MFMA)
    try:
        gp, gc, gnp, gnc = _gpu_nl(tfm, gpsi, gchi)
        ep, ec = rel_err(gnp.e, onp.e), rel_err(gnc.e, onc.e)
        _say('%s nonlinear term: psi %.2e chi %.2e' % (prec, ep, ec))
        assert ep < 1e-10 and ec < 1e-9
        # one AB2/CN step with N_prev = N (abcn, mlegs_scalar_ops.f90:1200-1262)
        errs = []
        for g, o, gn, on in ((gp, opsi, gnp, onp), (gc, ochi, gnc, onc)):
            gs, gsp, gnp_ = g.copy(), M.Scalar(tfm, space='FFF'), gn.copy()
            M.abcn(gs, gsp, gn, gnp_, p.dt, tfm)
            os_, osp, onp_ = o.copy(), so.Scalar(okit, 'FFF'), on.copy()
            so.abcn(os_, osp, on, onp_, p.dt, okit, p)
            errs.append(rel_err(gs.e, os_.e))
        _say('%s AB2/CN step: psi %.2e chi %.2e' % (prec, errs[0], errs[1]))
        assert max(errs) < 1e-10
    finally:
        tfm.set_precision(M.PREC_FP64)


def test_c3_nonlinear_term_mixed(c3):
    import mlegs_amd as M
    p, tfm, okit, (gpsi, gchi), _, (onp, onc) = c3
    tfm.set_precision(M.PREC_FP32)
    tfm.set_stage_precision(M.STAGE_PROJ0, M.PREC_FP64_MFMA)
    try:
        _, _, gnp, gnc = _gpu_nl(tfm, gpsi, gchi)
    finally:
        tfm.set_stage_precision(M.STAGE_PROJ0, -1)
        tfm.set_precision(M.PREC_FP64)
    ep, ec = rel_err(gnp.e, onp.e), rel_err(gnc.e, onc.e)
    _say('mixed nonlinear term: psi %.2e chi %.2e' % (ep, ec))
    assert ep < 1e-5 and ec < 1e-3
