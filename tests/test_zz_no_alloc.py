"""mlx_svv_filter_noalloc (the entry point that allocates no device memory)
against mlx_svv_filter through the C ABI: the same kernels on the same
input, so the filtered field and the updated gain must be bit-identical.
Two calls in a row, the second on a smaller kit (the static scratch keeps
the first call's contents), then a third on the first kit again."""
import ctypes

import numpy as np
import pytest

import _cases as C

pytestmark = pytest.mark.gpu


def _say(*a):
    print('[case]', *a, flush=True)


def _call(tfm, fn, e, gain):
    import mlegs_amd as M
    from mlegs_amd import _lib
    s = M.Scalar(tfm).upload(e, space='FFF', ln=0.25)
    g = ctypes.c_double(gain)
    _lib.check(tfm.ctx, getattr(tfm.lib, fn)(tfm.ctx, s.h, ctypes.byref(g)))
    return s.e, g.value


def _field(tfm, p, seed):
    e = C.spectral(tfm.dims, p, seed, decay=3.0)
    # a heavy tail, so the gain feedback is active
    rng = np.random.default_rng(seed)
    e = e + 0.3 * (rng.standard_normal(e.shape) + 1j * rng.standard_normal(e.shape))
    return e


@pytest.mark.parametrize('gain0', [0.0, 0.4])
def test_noalloc_svv_matches_svv_filter(gain0):
    import mlegs_amd as M
    big = C.params(C.C1, is_svv=True)
    small = C.params(C.SMALL, is_svv=True)
    kb = M.Kit(big, device=0)
    ks = M.Kit(small, device=0)
    for tfm, p, seed in ((kb, big, 5), (ks, small, 6), (kb, big, 7)):
        _say('noalloc svv nr=%d np=%d nz=%d gain0=%g' % (p.nr, p.np, p.nz, gain0))
        e = _field(tfm, p, seed)
        ref, gref = _call(tfm, 'mlx_svv_filter', e, gain0)
        got, ggot = _call(tfm, 'mlx_svv_filter_noalloc', e, gain0)
        _say('  gain %.17g vs %.17g' % (ggot, gref))
        assert gref > 0.0  # the filter acted
        assert ggot == gref
        assert np.array_equal(got, ref)
        assert not np.array_equal(got, e)
    kb.close()
    ks.close()


def test_noalloc_svv_rejects_physical_input():
    import mlegs_amd as M
    p = C.params(C.SMALL, is_svv=True)
    tfm = M.Kit(p, device=0)
    s = M.Scalar(tfm)
    g = ctypes.c_double(0.0)
    assert tfm.lib.mlx_svv_filter_noalloc(tfm.ctx, s.h, ctypes.byref(g)) == M._lib.MLX_E_LAYOUT
    tfm.close()
