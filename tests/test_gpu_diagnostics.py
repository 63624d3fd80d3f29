"""Device diagnostics: check_stability's non-finite flag in the vortex step
(vortical_flow_3d.f90:397-409) and the FP64 kinetic-energy reduction.

A NaN is put into the data of psi from the host; every kernel runs to its
end as usual (NaN is ordinary data to them) and the step reports
MLX_E_NONFINITE, with SVV on (the flag rides on SVV's reduction pass) and
with SVV off (a separate check kernel).
"""
import math

import numpy as np
import pytest

from conftest import rel_err
import _cases as C

pytestmark = pytest.mark.gpu


def _say(*a):
    print('[case]', *a, flush=True)


@pytest.mark.parametrize('svv', [True, False])
@pytest.mark.parametrize('prec', ['fp64', 'fp32'])
def test_vortex_step_reports_nonfinite_state(svv, prec):
    import mlegs_amd as M
    from oracle import kit as K
    _say('nonfinite svv=%s prec=%s' % (svv, prec))
    p = C.params(C.C1, is_svv=svv)
    tfm = M.Kit(p, device=0, precision=M.PREC_FP64 if prec == 'fp64' else M.PREC_FP32)
    okit = K.Kit3D(p)
    (gpsi, _), (gchi, _) = C.qvortex(tfm, okit, p)
    run = M.VortexRun(tfm, gpsi, gchi, M.Scalar(tfm))
    run.bootstrap()
    run.step(2, check=True)  # finite: no error
    psi = run.field('psi')
    e = psi.e
    e[5, 2, 3] = np.nan
    psi.upload(e, space='FFF', ln=psi.ln)
    with pytest.raises(M.MlxError) as ex:
        run.step(1, check=True)
    assert ex.value.code == M._lib.MLX_E_NONFINITE
    # the same state through the public check
    with pytest.raises(M.MlxError) as ex2:
        M.check_stability(run.field('psi'), run.field('chi'), tfm)
    assert ex2.value.code == M._lib.MLX_E_NONFINITE
    tfm.close()


def test_check_stability_finite_and_inf():
    import mlegs_amd as M
    p = C.params(C.SMALL)
    tfm = M.Kit(p, device=0)
    a = M.Scalar(tfm).upload(C.spectral(tfm.dims, p, 3), space='FFF')
    b = M.Scalar(tfm).upload(C.spectral(tfm.dims, p, 4), space='FFF')
    M.check_stability(a, b, tfm)
    e = b.e
    e[0, 0, 0] = np.inf
    b.upload(e, space='FFF')
    with pytest.raises(M.MlxError) as ex:
        M.check_stability(a, b, tfm)
    assert ex.value.code == M._lib.MLX_E_NONFINITE
    tfm.close()


def test_energy_closed_form_and_oracle():
    """1/2 int e^{-2 r^2} r dr dphi dz = pi zlen / 4 (int_0^inf e^{-2r^2} r dr
    = 1/4); a seeded velocity against the oracle's quadrature."""
    import mlegs_amd as M
    from oracle import kit as K, scalar_ops as so, vortex as V
    p = C.params(C.C1)
    tfm = M.Kit(p, device=0)
    okit = K.Kit3D(p)
    zero = M.Scalar(tfm)
    r = tfm.r
    u = np.zeros(tfm.dims, dtype=np.complex128)
    g = np.exp(-r ** 2)
    u[:p.nr, :p.np // 2, :] = (g + 1j * g)[:, None, None]
    gu = M.Scalar(tfm).upload(u, space='PPP')
    want = math.pi * p.zlen / 4
    got = M.energy(gu, zero, zero, tfm)
    _say('energy %.16e want %.16e' % (got, want))
    assert abs(got - want) < 1e-12 * want
    vs = [C.physical(tfm.dims, p, 70 + i) for i in range(3)]
    gv = [M.Scalar(tfm).upload(v, space='PPP') for v in vs]
    ov = []
    for v in vs:
        s = so.Scalar(okit, 'PPP')
        s.e = v.copy()
        ov.append(s)
    eg, eo = M.energy(*gv, tfm), V.energy(*ov, okit)
    _say('seeded: gpu %.16e oracle %.16e' % (eg, eo))
    assert abs(eg - eo) < 1e-12 * abs(eo)
    with pytest.raises(M.MlxError):
        M.energy(M.Scalar(tfm).upload(u, space='FFF'), zero, zero, tfm)
    tfm.close()


def test_energy_of_the_qvortex_velocity():
    """tp2vec of the q-vortex, then the energy of its axial flow e^{-r^2}/q:
    pi zlen / (4 q^2), the chain the step's diagnostics use."""
    import mlegs_amd as M
    from oracle import kit as K
    p = C.params(C.C1)
    tfm = M.Kit(p, device=0)
    okit = K.Kit3D(p)
    (gpsi, _), (gchi, _) = C.qvortex(tfm, okit, p, amp=0.0)
    vr, vp, vz = M.Scalar(tfm), M.Scalar(tfm), M.Scalar(tfm)
    M.tp2vec(gpsi, gchi, vr, vp, vz, tfm)
    zero = M.Scalar(tfm)
    got = M.energy(zero, zero, vz, tfm)
    want = math.pi * p.zlen / 4 / 16.0
    _say('axial energy %.16e want %.16e' % (got, want))
    assert abs(got - want) < 1e-10 * want
    tfm.close()
