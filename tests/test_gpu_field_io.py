"""Field files and the vorticity-magnitude field through the C ABI (marked gpu).

msave / mload (src/submodules/mlegs_scalar_io.f90:6-248) on device fields:
binary files round-trip bit for bit with space, ln and chop offsets; ASCII
files hold 16 significant digits and load back exactly what the host reader
(tests/test_field_io.py pins it against the Fortran runtime) reads.
save_vort_mag's field (vortical_flow_3d.f90:411-446) is compared with the
oracle's (pinned by the q-vortex closed form in test_oracle_vortex.py).
"""
import os

import numpy as np
import pytest

from conftest import phys, rel_err
import _cases as C

pytestmark = pytest.mark.gpu


def _say(*a):
    print('[case]', *a, flush=True)


@pytest.fixture(scope='module', params=['C1', 'ODD'])
def setup(request):
    import mlegs_amd as M
    from oracle import kit as K
    cfg = getattr(C, request.param)
    p = C.params(cfg)
    _say('kit', request.param, p.nr, p.np, p.nz)
    tfm = M.Kit(p, device=0)
    okit = K.Kit3D(p)
    yield request.param, p, tfm, okit
    tfm.close()


@pytest.mark.parametrize('binary', [True, False], ids=['binary', 'ascii'])
@pytest.mark.parametrize('glob', [True, False], ids=['global', 'per_process'])
def test_msave_mload_round_trip(setup, tmp_path, binary, glob):
    import mlegs_amd as M
    name, p, tfm, okit = setup
    _say('msave/mload', name, 'binary' if binary else 'ascii', 'global' if glob else 'per-process')
    e = C.spectral(tfm.dims, p, 5)
    a = M.Scalar(tfm).upload(e, space='FFF', ln=0.375, offsets=(3, 0, 0))
    fn = str(tmp_path / 'scalar_FFF.dat')
    M.msave(a, fn, is_binary=binary, is_global=glob)
    assert os.path.exists(fn if glob else fn + '_0')
    b = M.mload(fn, M.Scalar(tfm), is_binary=binary, is_global=glob)
    assert b.meta() == ('FFF', 0.375, (3, 0, 0))
    got = b.e
    if binary:
        assert np.array_equal(got, e)
    else:
        host, sp, ln, off = M.file_read(fn, tfm.dims, is_global=glob)
        assert np.array_equal(got, host) and (sp, ln, off) == ('FFF', 0.375, (3, 0, 0))
        assert rel_err(got, e) < 1e-15


def test_mload_of_a_host_file_and_errors(setup, tmp_path):
    import mlegs_amd as M
    name, p, tfm, okit = setup
    _say('mload host file', name)
    e = C.physical(tfm.dims, p, 9)
    fn = str(tmp_path / 'scalar_PPP.dat')
    M.file_write(fn, e, 'PPP', 0.0, (0, 0, 0))
    s = M.mload(fn, M.Scalar(tfm))
    assert s.space == 'PPP'
    host, _, _, _ = M.file_read(fn, tfm.dims)
    assert np.array_equal(s.e, host)
    # a file of another size: the reference's "size inconsistency" stop
    bad = str(tmp_path / 'bad.dat')
    M.file_write(bad, np.zeros((tfm.dims[0] + 1, tfm.dims[1], tfm.dims[2]), dtype=np.complex128), 'FFF')
    with pytest.raises(M.MlxError) as ei:
        M.mload(bad, s)
    assert ei.value.code == 204 and 'size inconsistency' in str(ei.value)
    assert np.array_equal(s.e, host)  # the field is left as it was
    # azimuthal / axial chop offsets are not on this path
    M.file_write(bad, e, 'FFF', 0.0, (0, 1, 0))
    with pytest.raises(M.MlxError) as ei:
        M.mload(bad, s)
    assert ei.value.code == 203


def test_vorticity_magnitude(setup, tmp_path):
    import mlegs_amd as M
    from oracle import vortex as V
    name, p, tfm, okit = setup
    _say('vorticity magnitude', name)
    (gpsi, opsi), (gchi, ochi) = C.qvortex(tfm, okit, p, amp=1e-3)
    out = M.Scalar(tfm)
    M.vorticity_magnitude(gpsi, gchi, out, tfm)
    ref = V.vorticity_magnitude(opsi, ochi, okit)
    nr, nph = p.nr, p.np // 2
    assert out.meta() == ('PPP', 0.0, (0, 0, 0))
    err = rel_err(phys(out.e, nr, nph), phys(ref.e, nr, nph))
    _say('  rel err %.2e' % err)
    assert err < 1e-11
    # written as save_vort_mag does: ASCII, global
    fn = str(tmp_path / 'vormagfld000000_PPP')
    M.msave(out, fn)
    host, sp, ln, off = M.file_read(fn, tfm.dims)
    assert sp == 'PPP' and rel_err(phys(host, nr, nph), phys(ref.e, nr, nph)) < 1e-11
    with pytest.raises(M.MlxError):
        M.vorticity_magnitude(out, gchi, M.Scalar(tfm), tfm)  # psi must be in FFF
