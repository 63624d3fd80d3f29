"""The seeded synthetic inputs shared by the GPU path and the oracle."""
import numpy as np
import pytest

from mlegs_amd import inputs


def test_splitmix64_known_answers():
    """SplitMix64 (Steele, Lea, Flood 2014) from state 0: the generator's published
    first outputs; state advances by the golden gamma 0x9E3779B97F4A7C15."""
    g = 0x9E3779B97F4A7C15
    got = [int(inputs.splitmix64(np.uint64((i * g) % 2 ** 64))) for i in range(3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_gaussian_pair_statistics():
    n = np.arange(200000, dtype=np.int64)
    z = np.zeros_like(n)
    g1, g2 = inputs.gaussian_pair(inputs.DEFAULT_SEED, 0, n, z, z)
    for g in (g1, g2):
        assert abs(g.mean()) < 0.01 and abs(g.std() - 1) < 0.01
    assert abs(np.corrcoef(g1, g2)[0, 1]) < 0.01
    a, b = inputs.gaussian_pair(inputs.DEFAULT_SEED, 1, n, z, z)
    assert abs(np.corrcoef(g1, a)[0, 1]) < 0.01   # field ids are independent streams


def test_spectral_field_is_admissible_and_deterministic():
    nr, np_, nz, nrchop, npchop, nzchop = 16, 12, 8, 14, 7, 5
    dims = (nr + 3, np_ // 2 + 1, nz)
    chops = (nrchop, npchop, nzchop, np_, nz)
    c = inputs.transform_field(dims, chops, field_id=4)
    assert np.array_equal(c, inputs.transform_field(dims, chops, field_id=4))
    assert not np.array_equal(c, inputs.transform_field(dims, chops, field_id=5))
    for i in range(dims[0]):
        for m in range(dims[1]):
            if i >= max(nrchop - m, 0) or m >= npchop:
                assert np.all(c[i, m, :] == 0)
    kap = inputs.kappa_index(nz)
    assert list(kap) == [0, 1, 2, 3, 4, -3, -2, -1]
    for k in range(1, nz // 2):
        np.testing.assert_array_equal(c[:, 0, nz - k], np.conj(c[:, 0, k]))
    assert np.all(c[:, 0, 0].imag == 0)
    assert np.all(c[:, :, nz // 2] == 0)       # kappa Nyquist
    assert np.all(c[:, np_ // 2, :] == 0)      # m Nyquist


def test_perturbation_amplitude_envelope():
    dims, chops = (67, 25, 32), (64, 25, 17, 48, 32)
    c = inputs.perturbation(dims, chops, 1e-5, scale=4.0)
    n = np.arange(dims[0])[:, None, None] + np.arange(dims[1])[None, :, None]
    env = 1e-5 * np.exp(-(n + np.arange(dims[1])[None, :, None] + np.abs(inputs.kappa_index(32))[None, None, :]) / 4.0)
    ratio = np.abs(c[env > 0]) / env[env > 0]
    assert ratio.max() < 7.0 and np.median(ratio[ratio > 0]) > 0.3


def test_qvortex_source_terms():
    x = np.linspace(-0.9, 0.9, 8)
    r = np.sqrt((1 + x) / (1 - x))
    psi = inputs.qvortex_ppp(x, r, (11, 3, 2), 4, 2, 8, 4.0, 'psi')
    chi = inputs.qvortex_ppp(x, r, (11, 3, 2), 4, 2, 8, 4.0, 'chi')
    np.testing.assert_allclose(psi[:8, 0, 0].real, -2 * np.exp(-r ** 2) / (1 - x) ** 2)
    np.testing.assert_allclose(chi[:8, 1, 1].imag, -np.exp(-r ** 2) / 4.0 / (1 - x) ** 2)
    assert np.all(psi[8:] == 0) and np.all(psi[:, 2] == 0)


def test_fortran_order_round_trip():
    a = np.arange(24).reshape(2, 3, 4).astype(np.complex128)
    f = inputs.to_fortran_bytes(a)
    assert f.flags.f_contiguous
    np.testing.assert_array_equal(inputs.from_fortran(f.ravel(order='F'), (2, 3, 4)), a)


@pytest.mark.parametrize('kind', [0, 1])
def test_native_generator_matches_numpy(kind):
    """mlx_seed_spectral (host C++ in the library, no GPU needed) and the numpy
    definition agree."""
    import os
    from mlegs_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from mlegs_amd import build
        build.build()
    dims, chops = (19, 7, 8), (16, 7, 5, 12, 8)
    if kind == 0:
        ref = inputs.perturbation(dims, chops, 1e-3, scale=4.0, field_id=6)
        got = inputs.seed_spectral_native(dims, chops, field_id=6, amp=1e-3, scale=4.0, kind=0)
    else:
        ref = inputs.transform_field(dims, chops, field_id=6, decay=5.0)
        got = inputs.seed_spectral_native(dims, chops, field_id=6, scale=5.0, kind=1)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-300)
    assert np.array_equal(got == 0, ref == 0)
