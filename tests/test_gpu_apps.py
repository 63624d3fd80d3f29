"""The programs of the reference's tutorial gate, built with amdflang against
the library (mlegs_amd/bin), run the way the reference's validator runs
them (tools/validate_tutorials.py:217-471): the same input files in one run
directory, in the same order (each later program reads what an earlier one
wrote), every exit status zero, SVV on and off giving different vorticity
fields, and the three invalid inputs rejected.

Beyond the exit status, every program's output is checked: the fields the
transform / Laplacian / vector-field tutorials write against the oracle on
the same inputs, the Lamb-Oseen origin values of both time-integration
tutorials against the oracle's replay of the same operator sequence and
against the closed form 1 / (2 pi (1 + 2 visc t)) (first order and second
order convergence), and the final fields of the 1D wave, 2D diffusion and 2D
transport programs against the oracle's replay.
"""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, 'mlegs_amd', 'bin')


def _say(*a):
    print('[case]', *a, flush=True)


def input_text(nr, np, nz, nrchop, npchop, nzchop, ell, dt, totaltime, totaln, visc, hyperpow=0,
               hypervisc=0.0, fields=False, field_interval=1, logs=False, log_interval=1, stability=True,
               svv_cutoff=0.75, svv_target=2.0e-2):
    """An input.params file in the reference's layout (mlegs_misc_io.f90:307-408)."""
    tf = lambda b: 'T' if b else 'F'
    return f"""!!! COMPUTATIONAL DOMAIN INFO !!!
# ---------- NR ----------- NP ----------- NZ ----------------------------------
             {nr}             {np}              {nz}
# ------ NRCHOP ------- NPCHOP ------- NZCHOP ----------------------------------
             {nrchop}             {npchop}              {nzchop}
# --------- ELL --------- ZLEN ------ ZLENxPI ---(IF ZLENxPI==T, ZLEN=ZLEN*PI)--
           {ell:.6E}           1.D0              F
#
!!! TIME STEPPING INFO !!!
# ---------- DT ----------- TI --------- TOTT ----------- NI --------- TOTN ----
          {dt:.6E}           0.D0          {totaltime:.6E}              0             {totaln}
#
!!! FIELD PROPERTY INFO !!!
# -------- VISC ----- HYPERPOW ---- HYPERVISC ----------------------------------
          {visc:.6E}              {hyperpow}           {hypervisc:.6E}
#
!!! FIELD SAVE INFO !!!
# --------------------- FLDDIR -------------------------------------------------
                  ./output/fld
# ---- ISFLDSAV -- FLDSAVINTVL ---(IF ISFLDSAV!=T, FIELDS ARE NOT SAVED)--------
              {tf(fields)}              {field_interval}
#
!!! DATA LOGGING INFO !!!
# --------------------- LOGDIR -------------------------------------------------
                  ./output/dat
# ---- ISLOGSAV -- LOGSAVINTVL ---(IF ISLOGSAV!=T, LOGS ARE NOT GENERATED)------
              {tf(logs)}              {log_interval}
!!! STABILITY CONTROL INFO (OPTIONAL) !!!
# ------- IS_SVV ------- CUTOFF ------- TARGET ------ STRENGTH ------- RELAX ---
              {tf(stability)}          {svv_cutoff:.6E}          {svv_target:.6E}          0.12          0.25
/* ------------------------------ END OF INPUT ------------------------------ */
"""


# the validator's configurations (validate_tutorials.py:222-327)
GATE = {
    'input.params': dict(nr=32, np=16, nz=8, nrchop=32, npchop=9, nzchop=5, ell=4.0, dt=1.0e-2, totaltime=2.0e-2,
                         totaln=2, visc=1.0e-4, hyperpow=8, hypervisc=5.0e-7),
    'input_1d.params': dict(nr=32, np=1, nz=1, nrchop=32, npchop=1, nzchop=1, ell=1.0, dt=1.0e-2, totaltime=2.0e-2,
                            totaln=2, visc=0.0),
    'input_2d.params': dict(nr=32, np=48, nz=1, nrchop=32, npchop=25, nzchop=1, ell=1.0, dt=1.0e-2,
                            totaltime=2.0e-2, totaln=2, visc=5.0e-3),
    'input_tutorial.params': dict(nr=32, np=48, nz=1, nrchop=32, npchop=25, nzchop=1, ell=1.0, dt=2.0e-1,
                                  totaltime=4.0e-1, totaln=2, visc=1.0e-1, fields=True, field_interval=1,
                                  stability=False),
    'input_svv_off.params': dict(nr=32, np=16, nz=8, nrchop=32, npchop=9, nzchop=5, ell=4.0, dt=5.0e-3,
                                 totaltime=1.0e-1, totaln=20, visc=1.0e-10, hyperpow=8, hypervisc=1.0e-10,
                                 fields=True, field_interval=20, stability=False, svv_cutoff=0.5,
                                 svv_target=1.0e-12),
    'input_svv_on.params': dict(nr=32, np=16, nz=8, nrchop=32, npchop=9, nzchop=5, ell=4.0, dt=5.0e-3,
                                totaltime=1.0e-1, totaln=20, visc=1.0e-10, hyperpow=8, hypervisc=1.0e-10,
                                fields=True, field_interval=20, stability=True, svv_cutoff=0.5,
                                svv_target=1.0e-12),
}
VECTOR = dict(nr=200, np=128, nz=8, nrchop=200, npchop=65, nzchop=5, ell=5.0, dt=1.0e-1, totaltime=1.0e-1,
              totaln=1, visc=1.0e-3, fields=True, field_interval=1, stability=False)
REJECTED = {
    'zero enabled hyperviscosity': dict(nr=32, np=16, nz=8, nrchop=32, npchop=9, nzchop=5, ell=4.0, dt=1.0e-2,
                                        totaltime=1.0e-2, totaln=1, visc=1.0e-4, hyperpow=8, hypervisc=0.0),
    'unsupported hyperviscosity power': dict(nr=32, np=16, nz=8, nrchop=32, npchop=9, nzchop=5, ell=4.0,
                                             dt=1.0e-2, totaltime=1.0e-2, totaln=1, visc=1.0e-4, hyperpow=10,
                                             hypervisc=1.0e-7),
    'invalid SVV cutoff': dict(nr=32, np=16, nz=8, nrchop=32, npchop=9, nzchop=5, ell=4.0, dt=1.0e-2,
                               totaltime=1.0e-2, totaln=1, visc=1.0e-4, hyperpow=8, hypervisc=1.0e-7,
                               svv_cutoff=2.0),
}


def _params(cfg):
    from mlegs_amd.params import Params
    keys = ('nr', 'np', 'nz', 'nrchop', 'npchop', 'nzchop', 'ell', 'dt', 'visc', 'hyperpow')
    d = {k: cfg[k] for k in keys if k in cfg}
    d.setdefault('hyperpow', 0)
    return Params(zlen=1.0, hypervisc=cfg.get('hypervisc', 0.0), is_svv=cfg.get('stability', True), **d)


def _run(app, cwd, config=None, check=True, timeout=120):
    _say('run', app, config or '')
    args = [os.path.join(BIN, app)] + ([config] if config else [])
    r = subprocess.run(args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=timeout)
    if check and r.returncode != 0:
        raise AssertionError('%s exited %d:\n%s' % (app, r.returncode, r.stdout[-3000:]))
    return r


def _write(run_dir, name, cfg):
    with open(os.path.join(run_dir, name), 'w') as f:
        f.write(input_text(**cfg))


def _read(run_dir, rel, dims):
    import mlegs_amd as M
    return M.file_read(os.path.join(run_dir, rel), dims)


def _log(run_dir, rel):
    rows = [ln.split() for ln in open(os.path.join(run_dir, rel)) if ln.strip()]
    return np.array([[float(x) for x in r] for r in rows])


@pytest.fixture(scope='module')
def run_dir(tmp_path_factory):
    d = str(tmp_path_factory.mktemp('gate'))
    os.makedirs(os.path.join(d, 'output', 'fld'))
    os.makedirs(os.path.join(d, 'output', 'dat'))
    for name, cfg in GATE.items():
        _write(d, name, cfg)
    return d


# --------------------------------------------------------------------------
# the gate, in the validator's order
# --------------------------------------------------------------------------
def test_gate_transform_tutorials(run_dir):
    import mlegs_amd as M
    from oracle import kit as K, scalar_ops as so
    for app in ('barebone_template', 'backward_trans', 'forward_trans', 'laplacian', 'inverse_laplacian'):
        _run(app, run_dir)
    p = _params(GATE['input_2d.params'])
    okit = K.Kit3D(p)
    dims = (okit.nrdim, okit.npdim, okit.nzdim)
    dat = 'output/dat/'
    e0, sp0, _, _ = _read(run_dir, dat + 'scalar_FFF_original.dat', dims)
    want = np.zeros(dims, dtype=np.complex128)
    want[1, 1, 0] = want[1, 2, 0] = 1.0
    assert sp0 == 'FFF' and np.array_equal(e0, want)
    # backward_trans
    o = so.Scalar(okit, 'FFF')
    o.e = want.copy()
    so.trans(o, 'PPP', okit)
    ep, spp, _, _ = _read(run_dir, dat + 'scalar_PPP.dat', dims)
    err_b = rel_err(ep[:p.nr], o.e[:p.nr])
    # forward_trans gives back the coefficients
    ef, spf, _, _ = _read(run_dir, dat + 'scalar_FFF.dat', dims)
    err_f = np.abs(ef - want).max()
    # laplacian of the forward_trans output, and its inverse
    ol = so.Scalar(okit, 'FFF')
    ol.e = ef.copy()
    so.del2(ol, okit)
    el, _, lnl, _ = _read(run_dir, dat + 'scalar_FFF_Laplacian.dat', dims)
    err_l = rel_err(el, ol.e)
    elp, spl, _, _ = _read(run_dir, dat + 'scalar_PPP_Laplacian.dat', dims)
    so.trans(ol, 'PPP', okit)
    err_lp = rel_err(elp[:p.nr], ol.e[:p.nr])
    oi = so.Scalar(okit, 'FFF')
    oi.e = el.copy()
    oi.ln = lnl
    so.idel2_proln(oi, okit)
    ei, _, _, _ = _read(run_dir, dat + 'scalar_FFF_from_inv.dat', dims)
    err_i = rel_err(ei, oi.e)
    with open(os.path.join(run_dir, dat + 'coords_r.dat')) as fh:
        head = [int(v) for v in fh.readline().split()]
        r = np.array([float(v) for v in fh.read().split()])
    _say('  backward %.2e forward %.2e del2 %.2e del2 PPP %.2e idel2 %.2e' % (err_b, err_f, err_l, err_lp, err_i))
    assert spp == 'PPP' and spf == 'FFF' and spl == 'PPP'
    assert err_b < 1e-13 and err_f < 1e-13 and err_l < 1e-13 and err_lp < 1e-13 and err_i < 1e-12
    assert head == [p.nr, 1, 1] and r.shape == (p.nr,)
    assert np.abs(r - okit.r[:p.nr]).max() < 1e-13 * okit.r[:p.nr].max()


def test_gate_time_integration_tutorials(run_dir):
    """Both run with the gate's two-step input; their saved fields after
    step 1 and 2 against the oracle's replay."""
    for app in ('time_integration_first', 'time_integration_second'):
        _run(app, run_dir)
        p = _params(GATE['input_tutorial.params'])
        want = lamb_oseen_replay(p, 2, second=(app.endswith('second')))
        dims = want[0][0].shape
        for n, (e_want, _) in enumerate(want, start=1):
            e, sp, _, _ = _read(run_dir, 'output/fld/sfld%06d_PPP' % n, dims)
            err = rel_err(e[:p.nr], e_want[:p.nr])
            _say('  %s step %d: %.2e' % (app, n, err))
            assert sp == 'PPP' and err < 1e-12


def test_gate_vector_field_tutorials(run_dir):
    import mlegs_amd as M
    from oracle import kit as K, scalar_ops as so
    _write(run_dir, 'input_tutorial.params', VECTOR)
    _run('vecfld_reconstruction', run_dir)
    _run('tp_project', run_dir)
    p = _params(VECTOR)
    okit = K.Kit3D(p)
    dims = (okit.nrdim, okit.npdim, okit.nzdim)
    fld = 'output/fld/'
    psi, chi = vortex_at(okit, p, 1.0, (1.0, 1.0))
    errs = {}
    for nm, ob in (('psi', psi), ('chi', chi)):
        e, _, ln, _ = _read(run_dir, fld + nm + '000000_FFF', dims)
        errs[nm] = rel_err(e, ob.e)
    ov = so.tp2vec(psi, chi, okit)
    got = []
    for nm, ob in zip(('vr', 'vp', 'vz'), ov):
        e, sp, _, _ = _read(run_dir, fld + nm + '000000_PPP', dims)
        errs[nm] = rel_err(e[:p.nr], ob.e[:p.nr])
        s = so.Scalar(okit, 'PPP')
        s.e = e
        got.append(s)
    opsi, ochi = so.vec2tp(*got, okit)
    for nm, ob in (('pjpsi', opsi), ('pjchi', ochi)):
        e, _, _, _ = _read(run_dir, fld + nm + '000000_FFF', dims)
        errs[nm] = rel_err(e, ob.e)
    _say('  ' + ' '.join('%s %.2e' % kv for kv in errs.items()))
    # pjchi's (m, kappa) = (0, 0) column passes idel2's kappa = 0 system: the
    # 1e-9 bound of the projection at 256^3
    assert max(v for k, v in errs.items() if k != 'pjchi') < 1e-11 and errs['pjchi'] < 1e-9


def test_gate_1d_2d_programs(run_dir):
    for app in ('wave_propagation_1d', 'scalar_diffusion_2d', 'scalar_transport_2d'):
        r = _run(app, run_dir)
        assert 'Program Finished' in r.stdout
        assert os.path.isfile(os.path.join(run_dir, 'output', 'dat', 'sval_origin.dat'))


def test_gate_vortex_svv_and_rejections(run_dir):
    _run('vortical_flow_3d', run_dir)
    states = []
    field = os.path.join(run_dir, 'output', 'fld', 'vormagfld000020_PPP')
    for cfg in ('input_svv_off.params', 'input_svv_on.params'):
        shutil.copyfile(os.path.join(run_dir, cfg), os.path.join(run_dir, 'input.params'))
        if os.path.exists(field):
            os.remove(field)
        _run('vortical_flow_3d', run_dir)
        assert os.path.isfile(field)
        states.append(open(field, 'rb').read())
    assert states[0] != states[1], 'SVV enabled/disabled fields are identical'
    for label, cfg in REJECTED.items():
        _write(run_dir, 'input.params', cfg)
        r = _run('vortical_flow_3d', run_dir, check=False)
        _say('  rejected %s: rc %d' % (label, r.returncode))
        assert r.returncode != 0, label


# --------------------------------------------------------------------------
# longer runs with logging on, against the oracle and the closed form
# --------------------------------------------------------------------------
LAMB = dict(nr=32, np=48, nz=1, nrchop=32, npchop=25, nzchop=1, ell=1.0, dt=2.0e-1, totaltime=1.0e1, totaln=50,
            visc=1.0e-1, fields=True, field_interval=10, logs=True, log_interval=1, stability=False)


@pytest.mark.parametrize('app', ['time_integration_first', 'time_integration_second'])
def test_lamb_oseen_origin_values(tmp_path, app):
    """The tutorial problem (docs/tutorial/time_integration.md:40-54): visc
    0.1 to t = 10 at dt = 0.2, 0.4 and 1.0.  Each logged w(0, t) equals the
    oracle's replay of the operator sequence; at t = 10 the error against
    the closed form falls as dt (first) or dt^2 (second)."""
    d = str(tmp_path)
    os.makedirs(os.path.join(d, 'output', 'fld'))
    os.makedirs(os.path.join(d, 'output', 'dat'))
    _write(d, 'input_tutorial.params', LAMB)
    _run(app, d)
    second = app.endswith('second')
    p = _params(LAMB)
    errs = []
    for fn, dt, n in (('sval_origin_dt2e-1.dat', 0.2, 50), ('sval_origin_dt4e-1.dat', 0.4, 25),
                      ('sval_origin_dt1e-0.dat', 1.0, 10)):
        log = _log(d, 'output/dat/' + fn)
        assert log.shape[0] == n - (1 if second else 0)
        p.dt = dt
        want = lamb_oseen_replay(p, n, second)
        wv = np.array([w[1] for w in want])[(1 if second else 0):]
        rel = np.abs(log[:, 2] - wv).max() / np.abs(wv).max()
        t_end = log[-1, 1]
        exact = 1.0 / (2 * math.pi * (1 + 2 * p.visc * t_end))
        errs.append(abs(log[-1, 2] - exact) / exact)
        _say('  %s dt %.1f: vs oracle %.2e, t=%.1f vs closed form %.3e' % (app, dt, rel, t_end, errs[-1]))
        assert rel < 1e-12 and abs(t_end - 10.0) < 1e-9
    ratio = errs[1] / errs[0]
    if second:
        assert 3.0 < ratio < 5.0 and errs[0] < 1e-3, errs
    else:
        assert 1.6 < ratio < 2.4 and errs[0] < 2e-2, errs


def test_wave_diffusion_transport_fields(tmp_path):
    """wave_propagation_1d, scalar_diffusion_2d and scalar_transport_2d at 20
    steps: the saved final field against the oracle's replay."""
    runs = (('wave_propagation_1d', 'input_1d.params', dict(GATE['input_1d.params'], totaln=20, totaltime=0.2)),
            ('scalar_diffusion_2d', 'input_2d.params', dict(GATE['input_2d.params'], totaln=20, totaltime=0.2)),
            ('scalar_transport_2d', 'input_2d.params', dict(GATE['input_2d.params'], totaln=20, totaltime=0.2)))
    for app, name, cfg in runs:
        d = os.path.join(str(tmp_path), app)
        os.makedirs(os.path.join(d, 'output', 'fld'))
        os.makedirs(os.path.join(d, 'output', 'dat'))
        cfg = dict(cfg, fields=True, field_interval=20, logs=True, log_interval=1)
        _write(d, name, cfg)
        _run(app, d)
        p = _params(cfg)
        want = {'wave_propagation_1d': wave_replay, 'scalar_diffusion_2d': diffusion_replay,
                'scalar_transport_2d': transport_replay}[app](p, 20)
        e, sp, _, _ = _read(d, 'output/fld/sfld000020_PPP', want.shape)
        err = rel_err(e[:p.nr], want[:p.nr])
        log = _log(d, 'output/dat/sval_origin.dat')
        _say('  %s: field vs oracle %.2e, %d log rows' % (app, err, log.shape[0]))
        assert sp == 'PPP' and err < 1e-11 and log.shape[0] == 20


# --------------------------------------------------------------------------
# oracle replays of the programs' operator sequences
# --------------------------------------------------------------------------
def _ppp_of(okit, p, fn_r_phi):
    """A PPP array from f(x, y) at the collocation points, the packed real
    and imaginary parts at angles p(2j-1) and p(2j) (1-based)."""
    e = np.zeros((okit.nrdim, okit.npdim, okit.nzdim), dtype=np.complex128)
    r = okit.r[:p.nr]
    ph = 2 * math.pi / p.np * np.arange(p.np + 1)
    for j in range(p.np // 2):
        e[:p.nr, j, :] = (fn_r_phi(r, ph[2 * j]) + 1j * fn_r_phi(r, ph[2 * j + 1]))[:, None]
    return e


def _zero_like(s):
    z = s.copy()
    z.e = np.zeros_like(z.e)
    return z


def lamb_oseen_replay(p, nsteps, second):
    """(PPP field, origin value) after each step 1..nsteps of the
    time-integration tutorials' first run."""
    from oracle import kit as K, scalar_ops as so
    okit = K.Kit3D(p)
    w = so.Scalar(okit, 'PPP')
    w.e = _ppp_of(okit, p, lambda r, ph: np.exp(-r ** 2 / 2.0) / (2 * math.pi))
    out = []

    def record(s):
        c = s.copy()
        so.trans(c, 'PPP', okit)
        v = so.calcat0(c, okit)[0].real
        out.append((c.e, v))

    dt = p.dt
    if not second:
        for _ in range(nsteps):
            so.trans(w, 'FFF', okit)
            so.febe(w, _zero_like(w), dt, okit, p)
            record(w)
        return out
    nlw = _zero_like(w)
    w_prev, nlw_prev = w.copy(), nlw.copy()
    w_rich = w.copy()
    so.trans(w_rich, 'FFF', okit)
    so.febe(w_rich, _zero_like(w_rich), dt, okit, p)
    so.trans(w, 'FFF', okit)
    for _ in range(2):
        so.febe(w, _zero_like(w), dt / 2, okit, p)
    w.e = 2.0 * w.e - w_rich.e
    record(w)
    so.trans(w_prev, 'FFF', okit)
    so.trans(nlw_prev, 'FFF', okit)
    for _ in range(nsteps - 1):
        so.abcn(w, w_prev, _zero_like(w), nlw_prev, dt, okit, p)
        record(w)
    return out


def wave_replay(p, nsteps):
    from oracle import kit as K, scalar_ops as so
    okit = K.Kit3D(p)
    s = so.Scalar(okit, 'PPP')
    s.e[:p.nr, 0, 0] = np.exp(-(okit.r[:p.nr] / 4.0) ** 2)
    dsdt = so.Scalar(okit, 'PPP')
    for _ in range(nsteps):
        so.trans(s, 'FFF', okit)
        d2 = s.copy()
        so.del2(d2, okit)
        so.trans(dsdt, 'FFF', okit)
        so.fefe(s, dsdt, p.dt, okit, p)
        so.fefe(dsdt, d2, p.dt, okit, p)
    so.trans(s, 'PPP', okit)
    return s.e


def _plateaus(r, ph):
    x, y = r * np.cos(ph), r * np.sin(ph)
    a = np.maximum(1.0 - np.hypot(x - 0.5, y - 0.5) ** 8, 0.0)
    return np.maximum(np.maximum(a, 0.5 * (1.0 - np.hypot(x + 0.5, y + 0.5) ** 8)), 0.0)


def diffusion_replay(p, nsteps):
    from oracle import kit as K, scalar_ops as so
    okit = K.Kit3D(p)
    s = so.Scalar(okit, 'PPP')
    s.e = _ppp_of(okit, p, _plateaus)
    for _ in range(nsteps):
        so.trans(s, 'FFF', okit)
        so.febe(s, _zero_like(s), p.dt, okit, p)
    so.trans(s, 'PPP', okit)
    return s.e


def transport_replay(p, nsteps):
    from oracle import kit as K, scalar_ops as so
    okit = K.Kit3D(p)
    r = okit.r[:p.nr]
    s = so.Scalar(okit, 'PPP')
    t = 0.0
    for _ in range(nsteps):
        t += p.dt
        nls = s.copy()
        so.trans(nls, 'FFF', okit)
        m = np.arange(min(okit.npdim, p.npchop))
        nls.e[:, :len(m), :] *= (1j * m)[None, :, None]
        so.trans(nls, 'PPP', okit)
        nls.e[:p.nr] = -nls.e[:p.nr] * ((1.0 - np.exp(-r) ** 2) / r ** 2)[:, None, None]
        src = _ppp_of(okit, p, lambda rr, ph: np.maximum(
            1.0 - (4 * np.hypot(rr * np.cos(ph) - 0.5, rr * np.sin(ph) - 0.5)) ** 8, 0.0))
        nls.e = -nls.e + src / 2.0 * (1.0 - math.cos(math.pi * t))
        so.trans(s, 'FFF', okit)
        so.trans(nls, 'FFF', okit)
        so.febe(s, nls, p.dt, okit, p)
        so.trans(s, 'PPP', okit)
    return s.e


def vortex_at(okit, p, q, centre):
    """psi, chi of vecfld_reconstruction's q-vortex at `centre` (FFF)."""
    from oracle import scalar_ops as so
    xo, yo = centre
    x = okit.x[:p.nr]
    out = []
    for amp in (2.0, 1.0 / q):
        f = lambda r, ph: -np.exp(-((r * np.cos(ph) - xo) ** 2 + (r * np.sin(ph) - yo) ** 2)) * amp / (1 - x) ** 2
        e = np.zeros((okit.nrdim, okit.npdim, okit.nzdim), dtype=np.complex128)
        r = okit.r[:p.nr]
        ph = 2 * math.pi / p.np * np.arange(p.np + 1)
        for j in range(p.np // 2):
            e[:p.nr, j, :] = (f(r, ph[2 * j]) + 1j * f(r, ph[2 * j + 1]))[:, None]
        s = so.Scalar(okit, 'PPP')
        s.e = e
        so.trans(s, 'FFF', okit)
        so.idelsqp(s, okit)
        so.zeroat1(s, okit)
        out.append(s)
    return out
