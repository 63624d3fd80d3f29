"""The C ABI: the shared library loads and exports every function that
include/mlegs_amd.h declares; the ctypes table and error codes agree with the
header.  No compute call is made (this runs without a GPU)."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'mlegs_amd.h')


def _declared():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(mlx_[a-z0-9_]+)\s*\(', text)))


def _defines():
    out = {}
    for name, val in re.findall(r'#define\s+(MLX_[A-Z0-9_]+)\s+(\d+)', open(HEADER).read()):
        out[name] = int(val)
    return out


@pytest.fixture(scope='module')
def lib():
    from mlegs_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from mlegs_amd import build
        build.build()
    return ctypes.CDLL(_lib.LIB_PATH)


def test_header_declares_the_boundary():
    names = _declared()
    for required in ('mlx_ctx_create', 'mlx_kit_init', 'mlx_field_upload', 'mlx_field_download', 'mlx_trans',
                     'mlx_ihelm', 'mlx_idel2', 'mlx_abcn', 'mlx_febe', 'mlx_vec2tp', 'mlx_tp2vec',
                     'mlx_tp2curlvec', 'mlx_vecprod', 'mlx_svv_filter', 'mlx_energy', 'mlx_last_error'):
        assert required in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    from mlegs_amd import _lib
    assert sorted(_lib.SIGNATURES) == _declared()
    d = _defines()
    for name in ('MLX_OK', 'MLX_E_LAYOUT', 'MLX_E_ALLOC', 'MLX_E_ARG', 'MLX_E_FILE', 'MLX_E_NONFINITE', 'MLX_E_HIP'):
        assert getattr(_lib, name) == d[name]
    assert (_lib.PREC_FP64, _lib.PREC_FP32, _lib.PREC_BF16X3, _lib.PREC_FP64_MFMA) == (
        d['MLX_PREC_FP64'], d['MLX_PREC_FP32'], d['MLX_PREC_BF16X3'], d['MLX_PREC_FP64_MFMA'])
    assert (_lib.STAGE_SYNTH, _lib.STAGE_UP, _lib.STAGE_PROJ, _lib.STAGE_PROJ0) == (
        d['MLX_STAGE_SYNTH'], d['MLX_STAGE_UP'], d['MLX_STAGE_PROJ'], d['MLX_STAGE_PROJ0'])


def test_params_struct_layout(tmp_path):
    """mlx_params field order, offsets and size as a C compiler lays out the header."""
    from mlegs_amd import _lib
    text = open(HEADER).read()
    body = re.search(r'typedef struct\s+mlx_params\s*\{(.*?)\}\s*mlx_params;', text, re.S).group(1)
    body = re.sub(r'/\*.*?\*/', '', body, flags=re.S)
    fields = re.findall(r'(int32_t|double|int)\s+([a-z0-9_,\s]+);', body)
    names = []
    for _, group in fields:
        names += [g.strip() for g in group.split(',')]
    assert names == [f[0] for f in _lib.MlxParams._fields_]
    src = tmp_path / 'layout.c'
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "mlegs_amd.h"\nint main(void){\n' +
                   ''.join('printf("%%zu\\n", offsetof(mlx_params, %s));\n' % n for n in names) +
                   'printf("%zu\\n", sizeof(mlx_params)); return 0; }\n')
    exe = tmp_path / 'layout'
    subprocess.run(['gcc', '-I', os.path.dirname(HEADER), str(src), '-o', str(exe)], check=True, timeout=120)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [getattr(_lib.MlxParams, n).offset for n in names] + [ctypes.sizeof(_lib.MlxParams)]
    assert got == want


def test_missing_library_is_an_error(tmp_path):
    """No CPU fallback: without the library the package refuses to load it."""
    code = ('import sys; sys.path.insert(0, %r); from mlegs_amd import _lib; _lib.LIB_PATH = %r\n'
            'try:\n    _lib.load()\nexcept ImportError as e:\n    print("IMPORTERROR", e)\n' % (ROOT, str(tmp_path / 'none.so')))
    out = subprocess.run([sys.executable, '-c', code], capture_output=True, text=True, timeout=120)
    assert 'IMPORTERROR' in out.stdout, out.stdout + out.stderr


def test_product_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, 'mlegs_amd')
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(('.py', '.hip', '.h', '.cpp', '.f90')):
                text = open(os.path.join(dirpath, f)).read()
                assert not re.search(r'^\s*(from|import)\s+oracle\b', text, re.M), f
                assert 'oracle/' not in text or f.endswith('.md'), f
