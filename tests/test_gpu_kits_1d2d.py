"""The polar (NZ = 1) and radial (NP = NZ = 1) kits of the reference's 2D and
1D programs, against the oracle.

The reference's 2D and 1D programs (scalar_diffusion_2d.f90,
time_integration_*.f90, wave_propagation_1d.f90) run on the 3D kit with
NZ = 1 and NP = 1 (mlegs_spectfm_init.f90:36-46 admits np = 1 and nz = 1);
trans skips the phi and z FFTs there (mlegs_scalar_ops.f90:188-226), and
dealias skips a direction of length 1 (:50-58).  Every operator on the path
of those programs is compared with the oracle on seeded inputs: the
transforms, del2 / idel2 / ihelm / helmp / ihelmp, the four steppers
(fefe, abab, febe, abcn) and, on the polar kit, the vector-field operators
and the q-vortex step.  Tolerances are the FP64 ones of the 3D tests.
"""
import numpy as np
import pytest

from conftest import rel_err
import _cases as C

pytestmark = pytest.mark.gpu

# input_2d.params of the reference (NR = 32, NP = 48, NZ = 1)
K2D = dict(nr=32, np=48, nz=1, nrchop=32, npchop=25, nzchop=1, ell=1.0, zlen=1.0,
           visc=5e-3, dt=1e-3, hyperpow=0, hypervisc=0.0, is_svv=True)
# input_1d.params of the reference at NR = 64 (NP = NZ = 1, inviscid)
K1D = dict(nr=64, np=1, nz=1, nrchop=64, npchop=1, nzchop=1, ell=10.0, zlen=1.0,
           visc=0.0, dt=1e-4, hyperpow=0, hypervisc=0.0, is_svv=False)
# a viscous radial kit with hyperviscosity
K1DV = dict(nr=48, np=1, nz=1, nrchop=44, npchop=1, nzchop=1, ell=2.0, zlen=1.0,
            visc=1e-2, dt=1e-3, hyperpow=4, hypervisc=1e-5, is_svv=False)
# a polar kit with hyperviscosity and chopped rows
K2DH = dict(nr=40, np=32, nz=1, nrchop=36, npchop=12, nzchop=1, ell=1.5, zlen=1.0,
            visc=1e-3, dt=2e-3, hyperpow=6, hypervisc=1e-6, is_svv=False)

TOL = 1e-12


def _say(*a):
    print('[case]', *a, flush=True)


@pytest.fixture(scope='module', params=['K2D', 'K1D', 'K1DV', 'K2DH'])
def kit(request):
    import mlegs_amd as M
    from oracle import kit as K
    p = C.params(globals()[request.param])
    _say('kit', request.param)
    tfm = M.Kit(p, device=0)
    okit = K.Kit3D(p)
    yield request.param, p, tfm, okit
    tfm.close()


def _fff_pair(tfm, okit, p, seed, ln=0.0):
    return C.pair(tfm, okit, C.spectral(tfm.dims, p, seed), 'FFF', ln=ln)


def _ppp(tfm, p, seed):
    """Seeded PPP data: a real radial profile on the radial kit, packed
    columns < np/2 on the polar kit."""
    rng = np.random.default_rng(500 + seed)
    e = np.zeros(tfm.dims, dtype=np.complex128)
    if p.np == 1:
        e[:p.nr, 0, 0] = rng.standard_normal(p.nr) + 1j * rng.standard_normal(p.nr)
    else:
        e[:p.nr, :p.np // 2, :] = rng.standard_normal((p.nr, p.np // 2, 1)) + \
            1j * rng.standard_normal((p.nr, p.np // 2, 1))
    return e


def test_dims(kit):
    name, p, tfm, okit = kit
    assert tfm.dims == (okit.nrdim, okit.npdim, okit.nzdim)
    assert tfm.dims[1] == p.np // 2 + 1 and tfm.dims[2] == 1


def test_trans_round_trip(kit):
    import mlegs_amd as M
    from oracle import scalar_ops as so
    name, p, tfm, okit = kit
    _say('trans', name)
    g, o = C.pair(tfm, okit, _ppp(tfm, p, 1), 'PPP')
    M.trans(g, 'FFF', tfm)
    so.trans(o, 'FFF', okit)
    ef = rel_err(g.e, o.e)
    g2, o2 = _fff_pair(tfm, okit, p, 2, ln=0.25)
    M.trans(g2, 'PPP', tfm)
    so.trans(o2, 'PPP', okit)
    eb = rel_err(g2.e[:p.nr], o2.e[:p.nr])
    M.trans(g2, 'FFF', tfm)
    so.trans(o2, 'FFF', okit)
    er = rel_err(g2.e, o2.e)
    _say('  forward %.2e backward %.2e round trip %.2e' % (ef, eb, er))
    assert ef < TOL and eb < TOL and er < TOL
    assert g2.space == 'FFF'


def test_laplacians(kit):
    import mlegs_amd as M
    from oracle import scalar_ops as so
    name, p, tfm, okit = kit
    _say('del2 idel2 ihelm', name)
    for op, args in (('del2', ()), ('del2h', ()), ('idel2', ()), ('ihelm', (-3.5,)), ('xxdx', ())):
        g, o = _fff_pair(tfm, okit, p, 3, ln=0.4)
        getattr(M, op)(g, *args, tfm) if args else getattr(M, op)(g, tfm)
        if op == 'idel2':
            so.idel2_proln(o, okit)
        elif args:
            getattr(so, op)(o, *args, okit)
        else:
            getattr(so, op)(o, okit)
        err = rel_err(g.e, o.e)
        _say('  %s %.2e ln %.3e / %.3e' % (op, err, g.ln, o.ln))
        assert err < 1e-11, op
        assert abs(g.ln - o.ln) <= 1e-11 * max(1.0, abs(o.ln)), op


def test_helmp_ihelmp(kit):
    """helmp to rounding.  ihelmp's systems (del2^(p/2) + 0.2 del2 - 40) are
    ill-conditioned (solutions differ from the oracle's by 3e-12 at power 4
    and 3e-10 at power 6 on the polar kit), so the solution is compared at
    1e-8, and so is the residual: the oracle's helmp of the GPU solution
    gives back the input (6e-10 at power 6, the cancellation in helmp of a
    large solution)."""
    import mlegs_amd as M
    from oracle import scalar_ops as so
    name, p, tfm, okit = kit
    _say('helmp ihelmp', name)
    for power in (4, 6):
        g, o = _fff_pair(tfm, okit, p, 4)
        M.helmp(g, power, 0.7, 0.3, tfm)
        so.helmp(o, power, 0.7, 0.3, okit)
        e1 = rel_err(g.e, o.e)
        g, o = _fff_pair(tfm, okit, p, 5)
        s0 = o.copy()
        M.ihelmp(g, power, -40.0, 0.2, tfm)
        so.ihelmp(o, power, -40.0, 0.2, okit)
        e2 = rel_err(g.e, o.e)
        back = o.copy()
        back.e = g.e.copy()
        back.ln = g.ln
        so.helmp(back, power, -40.0, 0.2, okit)
        e3 = rel_err(back.e, s0.e)
        _say('  power %d helmp %.2e ihelmp %.2e residual %.2e' % (power, e1, e2, e3))
        assert e1 < 1e-11 and e2 < 1e-8 and e3 < 1e-8


@pytest.mark.parametrize('stepper', ['fefe', 'abab', 'abab2', 'febe', 'abcn'])
def test_steppers(kit, stepper):
    import mlegs_amd as M
    from oracle import scalar_ops as so
    name, p, tfm, okit = kit
    _say('stepper', stepper, name)
    dt = p.dt
    gs, os_ = _fff_pair(tfm, okit, p, 6, ln=0.3)
    gsp, osp = _fff_pair(tfm, okit, p, 7, ln=0.2)
    gn, on = _fff_pair(tfm, okit, p, 8, ln=-0.1)
    gnp, onp = _fff_pair(tfm, okit, p, 9, ln=0.05)
    if stepper in ('febe', 'abcn') and p.visc < 5e-14 and p.hyperpow == 0:
        # no linear term to treat implicitly (mlegs_scalar_ops.f90:1172-1174)
        with pytest.raises(ValueError):
            getattr(so, stepper)(*((os_, on) if stepper == 'febe' else (os_, osp, on, onp)), dt, okit, p)
        with pytest.raises(M.MlxError) as ex:
            getattr(M, stepper)(*((gs, gn) if stepper == 'febe' else (gs, gsp, gn, gnp)), dt, tfm)
        assert 'inviscid' in str(ex.value)
        return
    for _ in range(3):
        if stepper == 'fefe':
            M.fefe(gs, gn, dt, tfm)
            so.fefe(os_, on, dt, okit, p)
        elif stepper in ('abab', 'abab2'):
            is2 = stepper == 'abab2'
            M.abab(gs, gsp, gn, gnp, dt, tfm, is_2nd_svis_p=is2)
            so.abab(os_, osp, on, onp, dt, okit, p, is_2nd_svis_p=is2)
        elif stepper == 'febe':
            M.febe(gs, gn, dt, tfm)
            so.febe(os_, on, dt, okit, p)
        else:
            M.abcn(gs, gsp, gn, gnp, dt, tfm)
            so.abcn(os_, osp, on, onp, dt, okit, p)
    err = rel_err(gs.e, os_.e)
    _say('  s %.2e ln %.6e / %.6e' % (err, gs.ln, os_.ln))
    assert err < 1e-11
    assert abs(gs.ln - os_.ln) <= 1e-11 * max(1.0, abs(os_.ln))
    if stepper in ('abab', 'abab2', 'abcn'):
        assert rel_err(gsp.e, osp.e) < 1e-11 and rel_err(gnp.e, onp.e) < 1e-11


def test_steppers_reject_non_fff(kit):
    import mlegs_amd as M
    name, p, tfm, okit = kit
    g = M.Scalar(tfm).upload(_ppp(tfm, p, 3), space='PPP')
    n = M.Scalar(tfm).upload(C.spectral(tfm.dims, p, 4), space='FFF')
    for fn in (lambda: M.fefe(g, n, 1e-3, tfm), lambda: M.abab(g, n.copy(), n, n.copy(), 1e-3, tfm)):
        with pytest.raises(M.MlxError) as ex:
            fn()
        assert 'FFF' in str(ex.value)


def test_calcat0_and_dealias(kit):
    import mlegs_amd as M
    from oracle import scalar_ops as so
    name, p, tfm, okit = kit
    g, o = _fff_pair(tfm, okit, p, 10, ln=0.1)
    a, b = M.calcat0(g, tfm), so.calcat0(o, okit)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))
    M.dealias(g, tfm)
    so.dealias(o, okit)
    assert rel_err(g.e, o.e) < TOL


def _polar(kit):
    name, p, tfm, okit = kit
    if p.np == 1:
        pytest.skip('the vector-field operators act on the polar kit')
    return name, p, tfm, okit


def test_polar_vector_fields(kit):
    import mlegs_amd as M
    from oracle import scalar_ops as so
    name, p, tfm, okit = _polar(kit)
    _say('tp2vec vec2tp', name)
    (gpsi, opsi), (gchi, ochi) = C.qvortex(tfm, okit, p, amp=0.3)
    vr, vp, vz = M.Scalar(tfm), M.Scalar(tfm), M.Scalar(tfm)
    M.tp2vec(gpsi, gchi, vr, vp, vz, tfm)
    ov = so.tp2vec(opsi, ochi, okit)
    errs = [rel_err(g.e[:p.nr], o.e[:p.nr]) for g, o in zip((vr, vp, vz), ov)]
    cp, cc = M.Scalar(tfm), M.Scalar(tfm)
    M.vec2tp(vr, vp, vz, cp, cc, tfm)
    op_, oc = so.vec2tp(*ov, okit)
    errs += [rel_err(cp.e, op_.e), rel_err(cc.e, oc.e)]
    _say('  ' + ' '.join('%.2e' % e for e in errs))
    # chi's (m, kappa) = (0, 0) column goes through idel2's kappa = 0 system,
    # the same 1e-9 bound as the 3D projection at 256^3
    assert max(errs[:4]) < 1e-12 and errs[4] < 1e-9


def test_polar_vortex_steps(kit):
    import mlegs_amd as M
    from oracle import scalar_ops as so, vortex as V
    name, p, tfm, okit = _polar(kit)
    _say('2D q-vortex steps', name)
    (gpsi, opsi), (gchi, ochi) = C.qvortex(tfm, okit, p, amp=0.2)
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
