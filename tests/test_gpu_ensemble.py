"""Ensembles of 2D Lamb-Oseen members stepped by HIP-graph replay
(mlx_ensemble_*), against each member's own single run.

Every launch configuration issues the same kernels in the same order per
member, so each member's psi and chi after the ensemble's steps must equal
those of a VortexRun of the same member stepped alone, to the last bit
(checked here at 1e-15 relative).  The configurations covered: the tuned
one, one parent graph per round, member graphs on 4 streams with four
steps per graph, and step counts that leave eager remainders.  The
unperturbed member is also checked against the Lamb-Oseen closed form
w(0, t) = 1 / (2 pi (1 + 2 visc t)).
"""
import math

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _say(*a):
    print('[case]', *a, flush=True)


def _members(n, nr_max=128):
    from mlegs_amd import ensemble as E
    out = []
    mid = 0
    while len(out) < n:
        p = E.member_params(mid)
        if p.nr <= nr_max:
            out.append((mid, p))
        mid += 1
    return out


def _single(mid, p, nsteps):
    import mlegs_amd as M
    from mlegs_amd import ensemble as E
    tfm = M.Kit(p, device=0)
    psi, chi, uz = E.lamb_oseen_tp(tfm, mid)
    run = M.VortexRun(tfm, psi, chi, uz)
    run.bootstrap()
    run.step(nsteps)
    out = (run.field('psi').e, run.field('chi').e)
    del run
    tfm.close()
    return out


def _ensemble(members, nsteps_list, **step_kw):
    from mlegs_amd import ensemble as E
    ens = E.Ensemble([p for _, p in members])
    for i, (mid, p) in enumerate(members):
        psi, chi, uz = E.lamb_oseen_tp(ens.kits[i], mid)
        ens.set(i, psi, chi, uz)
    ens.bootstrap()
    for n in nsteps_list:
        ens.step(n, **step_kw)
    ens.sync()
    assert ens.check() == -1
    res = [(ens.field(i, 'psi').e, ens.field(i, 'chi').e) for i in range(len(members))]
    info = ens.info()
    ens.close()
    return res, info


@pytest.fixture(scope='module')
def members():
    return _members(6)


@pytest.fixture(scope='module')
def singles(members):
    return {mid: _single(mid, p, 40) for mid, p in members}


@pytest.mark.parametrize('cfg', ['tuned', 'parent', 'streams4_reps4', 'remainder'])
def test_ensemble_matches_single_runs(members, singles, cfg):
    _say('ensemble', cfg, [(mid, p.nr) for mid, p in members])
    kw, steps = {
        'tuned': (dict(mode=-1), [1, 39]),
        'parent': (dict(mode=0, reps=1), [1, 39]),
        'streams4_reps4': (dict(mode=1, streams=4, reps=4), [1, 39]),
        'remainder': (dict(mode=1, streams=2, reps=1), [1, 10, 29]),
    }[cfg]
    res, info = _ensemble(members, steps, **kw)
    _say('  info', {k: v for k, v in info.items() if k != 'tune_ms'})
    worst = 0.0
    for (mid, p), (gpsi, gchi) in zip(members, res):
        spsi, schi = singles[mid]
        e = max(rel_err(gpsi, spsi), rel_err(gchi, schi))
        worst = max(worst, e)
        assert e <= 1e-15, (mid, e)
    _say('  worst rel. difference %.2e' % worst)
    if cfg == 'parent':
        assert info['mode'] == 0
    if cfg == 'streams4_reps4':
        assert info['mode'] == 1 and info['streams'] == 4 and info['steps_per_graph'] == 4


def test_lamb_oseen_member_closed_form():
    """An unperturbed member (amp 0) through 200 graph-replayed steps: the
    vorticity at the origin, -del2h(psi) at r = 0, follows
    1 / (2 pi (1 + 2 visc t))."""
    import mlegs_amd as M
    from mlegs_amd import ensemble as E
    p = E.member_params(0).replace(nr=64, np=64, nrchop=64, npchop=33, ell=1.0, visc=1e-2, is_svv=False)
    ens = E.Ensemble([p, p.replace(visc=1e-3)])
    for i in range(2):
        psi, chi, uz = E.lamb_oseen_tp(ens.kits[i], 0, amp=0.0)
        ens.set(i, psi, chi, uz)
    ens.bootstrap()
    ens.step(200)
    ens.sync()
    t = 201 * p.dt  # bootstrap step + 200
    for i, visc in enumerate((1e-2, 1e-3)):
        tfm = ens.kits[i]
        w = ens.field(i, 'psi').copy()
        M.del2h(w, tfm)
        w0 = -M.calcat0(w, tfm)[0].real
        want = 1.0 / (2 * math.pi * (1 + 2 * visc * t))
        _say('  visc %.0e: w(0, %.2f) = %.12f closed form %.12f' % (visc, t, w0, want))
        assert abs(w0 - want) < 1e-6 * want
    ens.close()


def test_ensemble_errors():
    import mlegs_amd as M
    from mlegs_amd import ensemble as E
    p = E.member_params(0).replace(nr=64, np=64, nrchop=64, npchop=33)
    ens = E.Ensemble([p, p])
    with pytest.raises(M.MlxError):
        ens.bootstrap()  # no member state yet
    with pytest.raises(M.MlxError):
        ens.step(1)
    psi, chi, uz = E.lamb_oseen_tp(ens.kits[0], 0)
    with pytest.raises(M.MlxError):
        ens.set(0, uz, chi, uz)  # psi must be FFF
    ens.set(0, psi, chi, uz)
    ens.set(1, *E.lamb_oseen_tp(ens.kits[1], 1))
    ens.bootstrap()
    with pytest.raises(M.MlxError):
        ens.step(-1)
    with pytest.raises(M.MlxError):
        ens.step(3, mode=1, streams=0)
    with pytest.raises(M.MlxError):
        ens.step(3, mode=0, reps=3)
    ens.close()
    with pytest.raises(M.MlxError):
        E.Ensemble([p.replace(np=7)])
