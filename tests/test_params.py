"""``input.params`` reading (mlegs_misc_io.f90:307-388) and timestep_set
(:390-408), including the rejection cases of the reference's validator
(tools/validate_tutorials.py:438-455)."""
import math

import pytest

from mlegs_amd.params import InputError, Params, read_input, timestep_set


def _text(nr=32, np=16, nz=8, nrchop=32, npchop=9, nzchop=5, ell='4.D0', zlen='2.D0', zpi='T',
          dt='1.D-2', ti='0.D0', tott='5.D1', ni=0, totn=5000, visc='1.D-4', hyperpow=8, hypervisc='5.D-7',
          flddir='./output/fld', isfld='T', fldint=100, logdir='./output/dat', islog='F', logint=1,
          svv=('T', '0.75', '2.D-2', '0.12', '0.25')):
    """A file in the reference's fixed-line layout (the shape of its input.params)."""
    lines = [
        '!!! COMPUTATIONAL DOMAIN INFO !!!',
        '# ---------- NR ----------- NP ----------- NZ ---------',
        '            %s            %s            %s' % (nr, np, nz),
        '# ------ NRCHOP ------- NPCHOP ------- NZCHOP ---------',
        '            %s             %s             %s' % (nrchop, npchop, nzchop),
        '# --------- ELL --------- ZLEN ------ ZLENxPI ---------',
        '           %s           %s              %s' % (ell, zlen, zpi),
        '#',
        '!!! TIME STEPPING INFO !!!',
        '# ---------- DT ----------- TI --------- TOTT ----------- NI --------- TOTN ----',
        '          %s           %s           %s              %s           %s' % (dt, ti, tott, ni, totn),
        '#',
        '!!! FIELD PROPERTY INFO !!!',
        '# -------- VISC ----- HYPERPOW ---- HYPERVISC ---------',
        '          %s              %s          %s' % (visc, hyperpow, hypervisc),
        '#',
        '!!! FIELD SAVE INFO !!!',
        '# --------------------- FLDDIR ------------------------',
        '                  %s' % flddir,
        '# ---- ISFLDSAV -- FLDSAVINTVL -------------------------',
        '              %s            %s' % (isfld, fldint),
        '#',
        '!!! DATA LOGGING INFO !!!',
        '# --------------------- LOGDIR ------------------------',
        '                  %s' % logdir,
        '# ---- ISLOGSAV -- LOGSAVINTVL -------------------------',
        '              %s              %s' % (islog, logint),
    ]
    if svv is not None:
        lines += ['!!! STABILITY CONTROL INFO (OPTIONAL) !!!',
                  '# ------- IS_SVV ------- CUTOFF ------- TARGET ------ STRENGTH ------- RELAX ---',
                  '              %s          %s          %s          %s          %s' % svv,
                  '/* ------------------------------ END OF INPUT ------------------------------ */']
    return '\n'.join(lines) + '\n'


def _read(tmp_path, text, **kw):
    fn = tmp_path / 'input.params'
    fn.write_text(text)
    return read_input(str(fn), params=Params(), make_dirs=False, **kw)


def test_reads_3d_file(tmp_path):
    p = _read(tmp_path, _text(nr=200, np=128, nz=128, nrchop=200, npchop=65, nzchop=65))
    assert (p.nr, p.np, p.nz, p.nrchop, p.npchop, p.nzchop) == (200, 128, 128, 200, 65, 65)
    assert p.ell == 4.0 and abs(p.zlen - 2 * math.pi) < 1e-15
    assert (p.dt, p.ti, p.totaltime, p.ni, p.totaln) == (1e-2, 0.0, 50.0, 0, 5000)
    assert (p.visc, p.hyperpow, p.hypervisc) == (1e-4, 8, 5e-7)
    assert p.flddir == './output/fld/' and p.isfldsav and p.fldsavintvl == 100
    assert p.logdir == './output/dat/' and not p.islogsav and p.logsavintvl == 1
    assert p.is_svv and (p.svv_cutoff, p.svv_target, p.svv_strength, p.svv_relax) == (0.75, 2e-2, 0.12, 0.25)


def test_reads_2d_and_1d_files(tmp_path):
    p = _read(tmp_path, _text(nr=32, np=48, nz=1, nrchop=32, npchop=25, nzchop=1, ell='1.D0', zlen='1.D0',
                              zpi='F', hyperpow=0, hypervisc='0.D0', svv=None))
    assert (p.nz, p.nzchop, p.zlen) == (1, 1, 1.0)
    assert p.is_svv  # the optional block is absent: the default stays
    p = _read(tmp_path, _text(nr=32, np=1, nz=1, nrchop=32, npchop=1, nzchop=1, hyperpow=0, hypervisc='0.D0'))
    assert (p.np, p.npchop) == (1, 1)


def test_svv_block_switches_filter_off(tmp_path):
    p = _read(tmp_path, _text(svv=('F', '0.6', '1.D-3', '0.2', '0.5')))
    assert not p.is_svv and p.svv_cutoff == 0.6 and p.svv_relax == 0.5


@pytest.mark.parametrize('label,kw', [
    ('zero enabled hyperviscosity', dict(hyperpow=8, hypervisc='0.0')),
    ('unsupported hyperviscosity power', dict(hyperpow=10, hypervisc='1.D-7')),
    ('invalid SVV cutoff', dict(svv=('T', '2.0', '2.D-2', '0.12', '0.25'))),
    ('odd nr', dict(nr=31)),
    ('npchop above np/2+1', dict(npchop=10)),
    ('nrchop above nr', dict(nrchop=40)),
    ('negative viscosity', dict(visc='-1.D-3')),
    ('truncated svv block', dict(svv=('T', '0.75', '2.D-2', '', ''))),
])
def test_rejections(tmp_path, label, kw):
    with pytest.raises(InputError):
        _read(tmp_path, _text(**kw))


def test_missing_and_truncated_files(tmp_path):
    with pytest.raises(InputError):
        read_input(str(tmp_path / 'nope.params'), make_dirs=False)
    with pytest.raises(InputError):
        _read(tmp_path, '\n'.join(_text().splitlines()[:9]))


def test_output_directories_created(tmp_path):
    d = tmp_path / 'run'
    fn = tmp_path / 'input.params'
    fn.write_text(_text(flddir=str(d / 'fld'), logdir=str(d / 'dat')))
    p = read_input(str(fn), params=Params())
    assert (d / 'fld').is_dir() and (d / 'dat').is_dir()
    assert p.flddir.endswith('/')


def test_timestep_set_picks_longer_criterion():
    """TOTT and TOTN*DT: the longer one wins (mlegs_misc_io.f90:390-408)."""
    p = Params(dt=1e-2, totaltime=50.0, totaln=100, ni=3, ti=0.5)
    assert timestep_set(p) == (3, 0.5)
    assert p.totaltime == 50.0 and p.totaln == 5000
    p = Params(dt=1e-2, totaltime=0.5, totaln=100)
    timestep_set(p)
    assert p.totaltime == 1.0 and p.totaln == 100
    with pytest.raises(InputError):
        timestep_set(Params(dt=0.0))
