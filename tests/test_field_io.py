"""Field files of msave / mload (src/submodules/mlegs_scalar_io.f90:6-248):
the library's host writer and reader (mlx_file_write / mlx_file_read, no
GPU) against the Fortran runtime executing the same format statements.

Pins:
- binary: a Fortran stream read of the library's file followed by msave's
  stream writes reproduces the file byte for byte;
- ASCII: the data records the library writes equal, field for field, the
  records Fortran's "1PE24.15E3" edit descriptor produces for the same
  values (including -0, subnormals and 1e308), and the size
  record equals "(3(1X,I10))";
- ASCII input: Fortran's formatted read with mload's statements and the
  library's reader recover the same doubles and the same trailer;
- the reference tutorials' own reader (np.loadtxt(skiprows=1, max_rows=NR),
  docs/tutorial/transformation.md:252, 354) reads the first z slice as
  alternating real and imaginary columns.
The Fortran check program is tests/golden/fio_check.f90 (compiled here with
the image's amdflang; the test skips if it is absent).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import mlegs_amd as M

FC = os.environ.get('FC_MLEGS', '/opt/rocm/bin/amdflang')
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope='module')
def fio(tmp_path_factory):
    if not (os.path.exists(FC) or shutil.which(FC)):
        pytest.skip('no Fortran compiler for the format check')
    d = tmp_path_factory.mktemp('fio')
    exe = str(d / 'fio_check')
    r = subprocess.run([FC, '-O1', os.path.join(HERE, 'golden', 'fio_check.f90'), '-o', exe],
                       capture_output=True, text=True, cwd=str(d))
    if r.returncode != 0:
        pytest.skip('fio_check did not compile: %s' % r.stderr[-400:])
    return exe


def _field(shape, seed):
    rng = np.random.default_rng(seed)
    mag = 10.0 ** rng.uniform(-300, 300, size=shape)
    e = (rng.standard_normal(shape) * mag + 1j * rng.standard_normal(shape) * mag).astype(np.complex128)
    flat = e.reshape(-1, order='F')
    specials = [0.0, -0.0, 5e-324, -2.2250738585072014e-308, 1.0e308, 1.0, -1.0 / 3.0, 0.1]
    for q, v in enumerate(specials[:flat.size]):
        flat[q] = complex(v, specials[-1 - q])
    return flat.reshape(shape, order='F')


SHAPES = [(7, 5, 3), (4, 1, 1), (3, 2, 4), (11, 6, 2)]


def _run(fio, d, shape):
    r = subprocess.run([fio, str(d)] + [str(x) for x in shape], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize('shape', SHAPES, ids=['x'.join(map(str, s)) for s in SHAPES])
def test_formats_against_fortran(fio, tmp_path, shape):
    e = _field(shape, 7 + len(shape))
    off = (3, 0, 0)
    M.file_write(str(tmp_path / 'c.bin'), e, 'PFP', -0.125, off, is_binary=True)
    M.file_write(str(tmp_path / 'c.txt'), e, 'PFP', -0.125, off, is_binary=False)
    _run(fio, tmp_path, shape)
    # binary: Fortran stream in, msave's stream out == the library's bytes
    assert (tmp_path / 'f.bin').read_bytes() == (tmp_path / 'c.bin').read_bytes()
    # ASCII: the size record and every data record, field for field
    ours = (tmp_path / 'c.txt').read_text().split('\n')
    theirs = (tmp_path / 'f.txt').read_text().split('\n')
    assert ours[0][:33] == theirs[0][:33]
    w = 48 * shape[1]
    drec = lambda lines: [ln[:w] for ln in lines[1:] if len(ln.rstrip()) == w]
    d_ours, d_theirs = drec(ours), drec(theirs)
    assert len(d_ours) == shape[0] * shape[2]
    assert d_ours == d_theirs
    # one blank record between z slices, as mload skips
    body = ours[1:1 + shape[2] * (shape[0] + 1) - 1]
    for k in range(shape[2] - 1):
        assert body[(k + 1) * (shape[0] + 1) - 1].strip() == ''
    # ASCII input: Fortran's formatted read and the library's reader agree
    raw = (tmp_path / 'r.bin').read_bytes()
    n = e.size
    b = np.frombuffer(raw[:16 * n], dtype=np.complex128).reshape(shape, order='F')
    ln_f = np.frombuffer(raw[16 * n:16 * n + 8], dtype=np.float64)[0]
    off_f = tuple(np.frombuffer(raw[16 * n + 8:16 * n + 20], dtype=np.int32))
    sp_f = raw[16 * n + 20:16 * n + 23].decode()
    x, sp, ln, off2 = M.file_read(str(tmp_path / 'c.txt'), shape)
    xs = np.stack([x.real, x.imag]).reshape(-1)
    bs = np.stack([b.real, b.imag]).reshape(-1)
    assert np.array_equal(xs.view(np.uint64), bs.view(np.uint64))  # bit for bit, -0 included
    assert (sp, ln, off2) == (sp_f, ln_f, off_f) == ('PFP', -0.125, off)
    # 16 significant digits: the text round trip is within half a unit of the 16th
    rel = np.abs(x - e) / np.maximum(np.abs(e), 1e-300)
    assert np.max(rel[np.abs(e) > 1e-290]) <= 1e-15


def test_tutorial_reader(tmp_path):
    """The reference tutorials read msave's ASCII output with
    np.loadtxt(fn, skiprows=1, max_rows=NR) and take re, im from alternating
    columns (transformation.md:252, 354-360)."""
    shape = (9, 4, 3)
    e = _field(shape, 3)
    fn = str(tmp_path / 'scalar_FFF.dat')
    M.file_write(fn, e, 'FFF')
    d = np.loadtxt(fn, skiprows=1, max_rows=shape[0])
    assert d.shape == (shape[0], 2 * shape[1])
    got = d[:, 0::2] + 1j * d[:, 1::2]
    assert np.allclose(got, e[:, :, 0], rtol=1e-15, atol=0)


@pytest.mark.parametrize('binary', [False, True], ids=['ascii', 'binary'])
def test_per_process_files(tmp_path, binary):
    """is_global = .false.: the file is fn_0 with the global size, block size
    and block start (0, 0, 0) in front (mlegs_scalar_io.f90:74-99)."""
    shape = (5, 3, 2)
    e = _field(shape, 11)
    fn = str(tmp_path / 'blk')
    M.file_write(fn, e, 'FFF', 0.5, (0, 0, 0), is_binary=binary, is_global=False)
    assert os.path.exists(fn + '_0') and not os.path.exists(fn)
    x, sp, ln, off = M.file_read(fn, shape, is_binary=binary, is_global=False)
    assert (sp, ln, off) == ('FFF', 0.5, (0, 0, 0))
    if binary:
        raw = open(fn + '_0', 'rb').read()
        hdr = np.frombuffer(raw[:36], dtype=np.int32)
        assert tuple(hdr) == shape + shape + (0, 0, 0)
        assert np.array_equal(x, e)
    else:
        lines = open(fn + '_0').read().split('\n')
        for q, want in enumerate((shape, shape, (0, 0, 0))):
            assert tuple(int(v) for v in lines[q].split()) == want
        assert np.allclose(x, e, rtol=1e-15, atol=0)


@pytest.mark.parametrize('binary', [False, True], ids=['ascii', 'binary'])
def test_load_errors(tmp_path, binary):
    e = _field((5, 3, 2), 5)
    fn = str(tmp_path / 'f')
    M.file_write(fn, e, 'FFF', is_binary=binary)
    with pytest.raises(M.MlxError) as ei:
        M.file_read(fn, (5, 3, 3), is_binary=binary)
    assert ei.value.code == 204 and 'size inconsistency' in str(ei.value)
    with pytest.raises(M.MlxError) as ei:
        M.file_read(str(tmp_path / 'missing'), (5, 3, 2), is_binary=binary)
    assert ei.value.code == 204 and 'cannot open' in str(ei.value)
    if not binary:
        # the largest double has no 16-digit form that reads back: its text
        # overflows, which the reader reports (the Fortran runtime stops)
        M.file_write(fn, np.full((5, 3, 2), 1.7976931348623157e308, dtype=np.complex128), 'FFF')
        with pytest.raises(M.MlxError) as ei:
            M.file_read(fn, (5, 3, 2))
        assert ei.value.code == 204 and 'out of range' in str(ei.value)
        M.file_write(fn, e, 'FFF')
    # a truncated file
    raw = open(fn, 'rb').read()
    open(fn, 'wb').write(raw[:len(raw) // 2])
    with pytest.raises(M.MlxError) as ei:
        M.file_read(fn, (5, 3, 2), is_binary=binary)
    assert ei.value.code == 204
