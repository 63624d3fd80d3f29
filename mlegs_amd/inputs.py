"""Deterministic synthetic inputs shared by the GPU path and its checker.

Every seeded field comes from one counter-based generator: SplitMix64
keyed by (seed, field id, n, m, kappa), then Box-Muller in FP64.  Arrays are
produced on the host in the reference's global layout e(nrdim, npdim, nzdim)
(complex128, index order (i, j, k) with i fastest in memory once flattened in
Fortran order) and uploaded unchanged, so both sides start from bit-identical
inputs.
"""
import numpy as np

DEFAULT_SEED = 20261202
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    """SplitMix64 finaliser on a uint64 array (wrapping arithmetic)."""
    with np.errstate(over='ignore'):
        z = (x + np.uint64(0x9E3779B97F4A7C15)).astype(np.uint64)
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)).astype(np.uint64)
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)).astype(np.uint64)
        return (z ^ (z >> np.uint64(31))).astype(np.uint64)


def _key(seed, field_id, n, m, kap):
    with np.errstate(over='ignore'):
        h = splitmix64(np.uint64(seed) ^ (np.uint64(field_id) << np.uint64(48)))
        h = splitmix64(h ^ n.astype(np.uint64))
        h = splitmix64(h ^ (m.astype(np.uint64) << np.uint64(20)))
        h = splitmix64(h ^ (kap.astype(np.int64).astype(np.uint64) << np.uint64(40)))
    return h


def gaussian_pair(seed, field_id, n, m, kap):
    """Two independent N(0,1) samples per (n, m, kappa) via Box-Muller."""
    h1 = _key(seed, field_id, n, m, kap)
    h2 = splitmix64(h1)
    u1 = ((h1 >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)
    u2 = ((h2 >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)
    rad = np.sqrt(-2.0 * np.log(u1))
    return rad * np.cos(2.0 * np.pi * u2), rad * np.sin(2.0 * np.pi * u2)


def kappa_index(nz):
    """Signed axial wavenumber index per FFF slice k (mlegs_spectfm_init.f90:132-134)."""
    k = np.arange(nz)
    return np.where(k <= nz // 2, k, k - nz)


def spectral_field(dims, chops, amp_fn, seed=DEFAULT_SEED, field_id=0):
    """Seeded admissible FFF coefficients.

    dims = (nrdim, npdim, nzdim); chops = (nrchop, npchop, nzchop, np, nz).
    c(n, m, kappa) = amp_fn(n, m, kappa) * (g1 + i g2)/sqrt(2) on the active
    set (row i < nrchop - m, m < npchop, every slice), then made admissible:
    the m = 0 plane is Hermitian in kappa with real (n, 0, 0) entries, and
    the Nyquist column m = np/2 and slice kappa = nz/2 are zero.
    """
    nrdim, npdim, nzdim = dims
    nrchop, npchop, nzchop, np_, nz = chops
    i = np.arange(nrdim)[:, None, None]
    m = np.arange(npdim)[None, :, None]
    kap = kappa_index(nzdim)[None, None, :] if nzdim > 1 else np.zeros((1, 1, 1), dtype=np.int64)
    n = m + i
    I, M, K = np.broadcast_arrays(i, m, kap)
    N = M + I
    g1, g2 = gaussian_pair(seed, field_id, N.astype(np.int64), M.astype(np.int64), K.astype(np.int64))
    c = amp_fn(N, M, np.abs(K)) * (g1 + 1j * g2) / np.sqrt(2.0)
    active = (I < np.maximum(nrchop - M, 0)) & (M < npchop)
    c = np.where(active, c, 0.0)
    if nzdim > 1:
        # Hermitian m = 0 plane: c(n,0,-kappa) = conj c(n,0,kappa)
        for k in range(1, nzdim // 2):
            c[:, 0, nzdim - k] = np.conj(c[:, 0, k])
        c[:, :, nzdim // 2] = 0.0
    c[:, 0, 0] = c[:, 0, 0].real
    if np_ > 1 and npdim - 1 == np_ // 2:
        c[:, npdim - 1, :] = 0.0
    del n
    return np.asarray(c, dtype=np.complex128)


def perturbation(dims, chops, amplitude, scale=4.0, seed=DEFAULT_SEED, field_id=0):
    """amplitude * exp(-(n+|m|+|kappa|)/scale) * (g1 + i g2)/sqrt(2), admissible."""
    return spectral_field(dims, chops, lambda N, M, K: amplitude * np.exp(-(N + M + K) / scale),
                          seed=seed, field_id=field_id)


def transform_field(dims, chops, seed=DEFAULT_SEED, field_id=0, decay=32.0):
    """|c| ~ N(0,1) exp(-n/decay): the transform-only workload's input."""
    return spectral_field(dims, chops, lambda N, M, K: np.exp(-N / decay), seed=seed, field_id=field_id)


def qvortex_ppp(x, r, dims, np_, nz, nr, q, which):
    """PPP samples of the centred q-vortex source terms.

    which='psi': -2 exp(-r^2)/(1-x)^2  (so that idelsqp gives psi with
    del_perp^2 psi = -2 exp(-r^2));  which='chi': -exp(-r^2)/q/(1-x)^2.
    Pattern of qvort_dist_tp (vortical_flow_3d.f90:258-326) for one vortex at
    the axis, without noise.
    """
    nrdim, npdim, nzdim = dims
    out = np.zeros(dims, dtype=np.complex128)
    amp = 2.0 if which == 'psi' else 1.0 / q
    f = -np.exp(-(r[:nr] ** 2)) * amp / (1.0 - x[:nr]) ** 2
    out[:nr, : np_ // 2, :] = (f + 1j * f)[:, None, None]
    return out


def to_fortran_bytes(a):
    """Flatten a (nrdim, npdim, nzdim) complex array in the reference's memory order."""
    return np.asfortranarray(a, dtype=np.complex128)


def from_fortran(buf, dims):
    return np.asarray(buf, dtype=np.complex128).reshape(dims, order='F')


def seed_spectral_native(dims, chops, seed=DEFAULT_SEED, field_id=0, amp=1.0, scale=4.0, kind=0):
    """The same field from the library's host generator (mlx_seed_spectral).

    kind 0: perturbation(amp, scale); kind 1: transform_field(decay=scale)."""
    import ctypes
    from . import _lib
    lib = _lib.load()
    out = np.zeros(2 * int(np.prod(dims)), dtype=np.float64)
    ch = (ctypes.c_int32 * 5)(*[int(v) for v in chops])
    rc = lib.mlx_seed_spectral(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), dims[0], dims[1], dims[2], ch,
                               int(seed), int(field_id), float(amp), float(scale), int(kind))
    if rc != 0:
        raise ValueError('mlx_seed_spectral: invalid arguments (code %d)' % rc)
    return out.view(np.complex128).reshape(dims, order='F')
