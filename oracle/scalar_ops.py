"""Scalar-field operations of the oracle (test infrastructure only).

Restates ``src/submodules/mlegs_scalar_ops.f90`` for one process (every
``loc_st`` is 0 and every ``exchange`` is a no-op on one address space).
Fields use the reference's global layout e(nrdim, npdim, nzdim), complex128,
here as numpy arrays of that shape (0-based indices).

Banded solves use scipy's LAPACK ``gbsv`` (partial pivoting), the same
factorisation as the reference's zgbtrf/zgbtrs (mlegs_bndmat_ops.f90:369-430).

Every z-loop visits the reference's index lists exactly, including the quirk
that slice k = nzchop is visited twice when nzchop = nz/2+1 (chopzl = chopzu):
operators that are not idempotent (del2, del2h, xxdx, ihelm, idel2) are then
applied twice to that slice, as the reference does.
"""
import numpy as np
import scipy.fft as sfft
from scipy.linalg import solve_banded
from scipy.linalg.lapack import zgbsv

# threads of the oracle's FFTs (scipy.fft workers); the CPU baseline sets it
FFT_WORKERS = 1

PI = np.pi
IU = 1j


class Scalar:
    """type(scalar) on one rank (mlegs_scalar.f90:15-49)."""

    def __init__(self, kit, space='PPP'):
        self.e = np.zeros((kit.nrdim, kit.npdim, kit.nzdim), dtype=np.complex128)
        self.ln = 0.0
        self.nrchop_offset = 0
        self.npchop_offset = 0
        self.nzchop_offset = 0
        self.space = space

    def copy(self):
        """Deep copy '=' (mlegs_scalar_init.f90:106-140)."""
        o = Scalar.__new__(Scalar)
        o.e = self.e.copy()
        o.ln = self.ln
        o.nrchop_offset = self.nrchop_offset
        o.npchop_offset = self.npchop_offset
        o.nzchop_offset = self.nzchop_offset
        o.space = self.space
        return o

    def assign(self, other):
        self.e = other.e.copy()
        self.ln = other.ln
        self.nrchop_offset = other.nrchop_offset
        self.npchop_offset = other.npchop_offset
        self.nzchop_offset = other.nzchop_offset
        self.space = other.space

    def chop_offset(self, iof1, iof2=0, iof3=0):
        """scalar_chop_offset (mlegs_scalar_init.f90:84-103)."""
        self.nrchop_offset = iof1
        self.npchop_offset = iof2
        self.nzchop_offset = iof3


# ----------------------------------------------------------------------------
# index helpers
# ----------------------------------------------------------------------------
def chop_index(s, kit):
    """chop_index (mlegs_scalar_ops.f90:2012-2063), 3D kit."""
    if s.npchop_offset != 0 or s.nzchop_offset != 0:
        raise ValueError('chop_index: only radial chopping offsets are used on this path')
    nrc = kit.chops[0] + s.nrchop_offset
    npc = kit.chopp + s.npchop_offset
    nzc = kit.chopzl + s.nzchop_offset
    nzcu = kit.chopzu - s.nzchop_offset
    if nrc > kit.nrdim:
        raise ValueError('chop_index: chopping in r too large')
    if npc > kit.npdim:
        raise ValueError('chop_index: chopping in p too large')
    nrcs = np.zeros(kit.npchop, dtype=np.int64)
    for mm in range(npc):
        nrcs[mm] = max(min(nrc, nrc - kit.m[mm]), 0)
    return nrc, npc, nzc, nzcu, nrcs


def zlist(kit, nzc, nzcu):
    """0-based z slices in the order of the two kk-loops of del2 (:546-557)."""
    nz = kit.nzdim
    ks = list(range(0, min(nz, nzc)))
    lo = max(nzcu, 1)
    if lo <= nz:
        ks += list(range(lo - 1, nz))
    return ks


# ----------------------------------------------------------------------------
# masks
# ----------------------------------------------------------------------------
def chop(s, kit):
    """chop (mlegs_scalar_ops.f90:6-41)."""
    nrc, npc, nzc, nzcu, nrcs = chop_index(s, kit)
    if s.space[0] == 'F':
        for mm in range(s.e.shape[1]):
            rc = nrcs[mm] + 1 if mm < len(nrcs) else 1
            if 1 <= rc:
                s.e[rc - 1:, mm, :] = 0.0
            else:
                s.e[:, mm, :] = 0.0
    if s.space[1] == 'F':
        s.e[:, npc:, :] = 0.0
    if s.space[2] == 'F':
        if nzc < nzcu:
            s.e[:, :, nzc:nzcu - 1] = 0.0


def dealias(s, kit):
    """2/3 rule in phi and z (mlegs_scalar_ops.f90:43-70)."""
    np_, nz = kit.np, kit.nz
    if np_ > 1 and s.space[1] == 'F':
        pcut = max(np_ // 3 + 1, 1)
        s.e[:, pcut:, :] = 0.0
    if nz > 1 and s.space[2] == 'F':
        zcut = max(nz // 3 + 1, 1)
        zupper = nz - zcut + 2
        # 1-based slices zcut < k < zupper  ->  0-based zcut .. zupper-2
        if zupper - 1 > zcut:
            s.e[:, :, zcut:zupper - 1] = 0.0


def svv_quality(kit, shape):
    """q = min(1, max(q_r, q_p, q_z)) per entry (mlegs_scalar_ops.f90:95-115)."""
    nrchop, np_ = kit.nrchop, kit.np
    i = np.arange(shape[0], dtype=np.float64)
    q_r = np.minimum(1.0, np.maximum(i, 0) / max(float(nrchop - 1), 1.0))
    j = np.arange(shape[1], dtype=np.float64)
    q_p = np.minimum(1.0, np.abs(j) / max(float(np_ // 2), 1.0))
    kmax = max(np.max(np.abs(kit.ak)), 1.0)
    q_z = np.minimum(1.0, np.abs(kit.ak[: shape[2]]) / kmax)
    q = np.maximum(np.maximum(q_r[:, None, None], q_p[None, :, None]), q_z[None, None, :])
    return np.minimum(1.0, q)


def svv_filter(s, kit, gain, p):
    """Dynamic spectral vanishing-viscosity filter (mlegs_scalar_ops.f90:72-155).

    Returns the updated gain.
    """
    if not p.is_svv:
        return gain
    if s.space != 'FFF':
        raise ValueError('svv_filter: scalar must be in FFF space')
    a2 = np.abs(s.e) ** 2
    total = float(np.sum(a2))
    cutoff = min(max(p.svv_cutoff, 0.0), 0.99)
    target = max(p.svv_target, 1e-12)
    q = svv_quality(kit, s.e.shape)
    tail = float(np.sum(a2[q >= cutoff]))
    if total <= np.finfo(np.float64).tiny:
        return gain
    ratio = tail / total
    feedback = min(max(ratio / target - 1.0, 0.0), 1.0)
    relax = min(max(p.svv_relax, 0.0), 1.0)
    gain = (1.0 - relax) * gain + relax * feedback
    strength = min(max(p.svv_strength, 0.0) * gain, 1.0)
    if strength <= 0.0:
        return gain
    sel = q > cutoff
    shape = (q[sel] - cutoff) / (1.0 - cutoff)
    s.e[sel] *= np.exp(-strength * shape ** 8)
    return gain


# ----------------------------------------------------------------------------
# FFT kernels (FFTE sign/scale: forward e^{-}, /N; backward e^{+}, unscaled)
# ----------------------------------------------------------------------------
def horizontal_fft_forward(s, kit):
    """Real FFT along phi of packed rows (mlegs_scalar_ops.f90:1567-1641)."""
    nph = kit.np // 2
    f = np.empty((s.e.shape[0], 2 * nph, s.e.shape[2]))
    f[:, 0::2, :] = s.e[:, :nph, :].real
    f[:, 1::2, :] = s.e[:, :nph, :].imag
    c = sfft.rfft(f, axis=1, workers=FFT_WORKERS) / (2 * nph)
    s.e[:, :nph + 1, :] = c


def horizontal_fft_backward(s, kit):
    """Hermitian synthesis along phi, packed output (mlegs_scalar_ops.f90:1645-1717).

    Column np/2+1 is FFTE's scratch in the reference; here it is set to 0.
    """
    nph = kit.np // 2
    f = sfft.irfft(s.e[:, :nph + 1, :], n=2 * nph, axis=1, workers=FFT_WORKERS) * (2 * nph)
    s.e[:, :nph, :] = f[:, 0::2, :] + IU * f[:, 1::2, :]
    s.e[:, nph, :] = 0.0


def vertical_fft_forward(s, kit):
    """(1/nz) sum e^{-2 pi i jk/nz} along z for every (row, col) (:1721-1783)."""
    s.e[:, :, :] = sfft.fft(s.e, axis=2, workers=FFT_WORKERS) / kit.nz


def vertical_fft_backward(s, kit):
    """sum e^{+2 pi i jk/nz} along z for every (row, col) (:1787-1848)."""
    s.e[:, :, :] = sfft.ifft(s.e, axis=2, workers=FFT_WORKERS) * kit.nz


# ----------------------------------------------------------------------------
# radial (mapped Legendre) transforms
# ----------------------------------------------------------------------------
def _nrcs_rtrans(s, kit):
    nrc = kit.chops[0] + s.nrchop_offset
    npc = kit.chopp + s.npchop_offset
    if nrc > kit.nrdim:
        raise ValueError('rtrans: chopping in r too large')
    if npc > kit.npdim:
        raise ValueError('rtrans: chopping in p too large')
    nrcs = np.array([max(min(nrc, nrc - kit.m[mm]), 0) for mm in range(npc)], dtype=np.int64)
    return npc, nrcs


def rtrans_forward(s, kit):
    """Gauss quadrature against P_n^m, parity folded (mlegs_scalar_ops.f90:1852-1927)."""
    npc, nrcs = _nrcs_rtrans(s, kit)
    nr, nrh = kit.nr, kit.nr // 2
    se = np.zeros_like(s.e)
    w = kit.w[:nrh, None]
    for mm in range(npc):
        a = s.e[:nrh, mm, :]
        b = s.e[nr - 1:nrh - 1:-1, mm, :]
        be = (a + b) * w
        bo = (a - b) * w
        nn = nrcs[mm]
        pf = kit.pf[:, :, mm]
        if nn >= 1:
            se[0:nn:2, mm, :] = pf[:, 0:nn:2].T @ be
        if nn >= 2:
            se[1:nn:2, mm, :] = pf[:, 1:nn:2].T @ bo
    s.e = se


def rtrans_backward(s, kit):
    """Synthesis sum_n e_n P_n^m(x), unfolded to +-x (mlegs_scalar_ops.f90:1931-2008)."""
    npc, nrcs = _nrcs_rtrans(s, kit)
    nr, nrh = kit.nr, kit.nr // 2
    se = np.zeros_like(s.e)
    for mm in range(npc):
        nn = nrcs[mm]
        pf = kit.pf[:, :, mm]
        be = pf[:, 0:nn:2] @ s.e[0:nn:2, mm, :] if nn >= 1 else 0.0
        bo = pf[:, 1:nn:2] @ s.e[1:nn:2, mm, :] if nn >= 2 else 0.0
        se[:nrh, mm, :] = be + bo
        se[nr - 1:nrh - 1:-1, mm, :] = be - bo
    s.e = se


_SPACES = {'PPP': 0, 'PFP': 1, 'FFP': 2, 'FFF': 3}


def trans(s, space, kit):
    """State machine PPP <-> PFP <-> FFP <-> FFF (mlegs_scalar_ops.f90:157-235)."""
    if s.space not in _SPACES:
        raise ValueError('trans: scalar space info corrupted (only accepting PPP, PFP, FFP and FFF)')
    if space not in _SPACES:
        raise ValueError('trans: only taking PPP, PFP, FFP and FFF for spectral transformation')
    cur, new = _SPACES[s.space], _SPACES[space]
    nr = kit.nr
    if cur < new:
        while True:
            if cur == 0:
                if kit.np > 1:
                    horizontal_fft_forward(s, kit)
                s.space = 'PFP'
            if cur == 1:
                s.e[:nr, 0, :] -= s.ln * kit.lnx[:nr, None]
                rtrans_forward(s, kit)
                s.space = 'FFP'
            if cur == 2:
                if kit.nz > 1:
                    vertical_fft_forward(s, kit)
                s.space = 'FFF'
            if cur == 3:
                break
            cur += 1
            if cur == new:
                break
    if cur > new:
        while True:
            if cur == 3:
                if kit.nz > 1:
                    vertical_fft_backward(s, kit)
                s.space = 'FFP'
            if cur == 2:
                rtrans_backward(s, kit)
                s.e[:nr, 0, :] += s.ln * kit.lnx[:nr, None]
                s.space = 'PFP'
            if cur == 1:
                if kit.np > 1:
                    horizontal_fft_backward(s, kit)
                s.space = 'PPP'
            if cur == 0:
                break
            cur -= 1
            if cur == new:
                break


# ----------------------------------------------------------------------------
# values at the origin / infinity
# ----------------------------------------------------------------------------
def _calcat(s, kit, tbl):
    nrc, npc, nzc, nzcu, nrcs = chop_index(s, kit)
    st = s.copy()
    trans(st, 'FFF', kit)
    n = min(st.e.shape[0], nrc)
    return tbl[:n] @ st.e[:n, 0, :]


def calcat0(s, kit):
    """f(r=0) of the m=0 column per z slice (mlegs_scalar_ops.f90:237-272)."""
    return _calcat(s, kit, kit.at0)


def calcat1(s, kit):
    """f(r=inf) of the m=0 column per z slice (mlegs_scalar_ops.f90:274-309)."""
    return _calcat(s, kit, kit.at1)


def zeroat1(s, kit):
    """Remove the value at infinity through the n=0 coefficient (:311-325)."""
    if s.space != 'FFF':
        trans(s, 'FFF', kit)
    at1 = calcat1(s, kit)
    s.e[0, 0, :] -= at1 / kit.at1[0]


# ----------------------------------------------------------------------------
# diagonal operators
# ----------------------------------------------------------------------------
def delsqp(s, kit):
    """(1-x)^-2 del_perp^2: diagonal -n(n+1)/ell^2 (mlegs_scalar_ops.f90:327-366)."""
    nrc, npc, nzc, nzcu, nrcs = chop_index(s, kit)
    if s.space != 'FFF':
        trans(s, 'FFF', kit)
    so = s.copy()
    so.ln = 0.0
    ell = kit.ell
    for mm in range(npc):
        nn = nrcs[mm]
        n = kit.m[mm] + np.arange(nn)
        so.e[:nn, mm, :] = -s.e[:nn, mm, :] * (n * (n + 1.0))[:, None] / (ell ** 2)
    so.e[0, 0, 0] = s.ln / ell ** 2 / np.exp(kit.lognorm[0, 0])
    s.assign(so)


def idelsqp(s, kit):
    """Inverse of delsqp with the log term (mlegs_scalar_ops.f90:368-416)."""
    nrc, npc, nzc, nzcu, nrcs = chop_index(s, kit)
    if s.space != 'FFF':
        trans(s, 'FFF', kit)
    so = s.copy()
    ell = kit.ell
    so.ln = float(np.real(s.e[0, 0, 0])) * ell ** 2 * np.exp(kit.lognorm[0, 0])
    for mm in range(npc):
        nn = nrcs[mm]
        n = (kit.m[mm] + np.arange(nn)).astype(np.float64)
        fac = np.where(n == 0, 0.0, -(ell ** 2) / np.where(n == 0, 1.0, n) / (n + 1.0))
        so.e[:nn, mm, :] = s.e[:nn, mm, :] * fac[:, None]
        if kit.m[mm] == 0 and nn >= 1:
            so.e[0, mm, :] = 0.0
    so.e[0, 0, :] = 0.0
    s.assign(so)


def smooth(a):
    """One smoothing pass along the first axis (mlegs_scalar_ops.f90:2149-2175):
    out(1) = a(1); a(ni) spread over out(ni-2:ni) with 0.2, 0.3, 0.5; each
    interior a(i) over out(i-1:i+1) with (1-f)/2, f, (1-f)/2,
    f = (1 + (ni - i)/(ni - 1))/2 (1-based i)."""
    ni = a.shape[0]
    if ni < 3:
        raise ValueError('smooth: the input length must be longer than or equal to 3.')
    out = np.zeros_like(a)
    out[0] = a[0]
    for q, fae in enumerate((0.2, 0.3, 0.5)):
        out[ni - 3 + q] += a[ni - 1] * fae
    for i1 in range(2, ni):  # 1-based i = 2 .. ni-1
        f = (1.0 + (ni - i1) / (ni - 1.0)) * 0.5
        i = i1 - 1
        out[i - 1] += a[i] * ((1.0 - f) / 2.0)
        out[i] += a[i] * f
        out[i + 1] += a[i] * ((1.0 - f) / 2.0)
    return out


def fftreat(s, kit):
    """Far-field smoothing (mlegs_scalar_ops.f90:1002-1063): delsqp of s
    without its log term, radial synthesis, (1-x)^2 weighting, rows >= ns0 of
    columns m >= 1 zeroed (every slice), smooth five times over rows >= ns
    on the chopped slices, weighting removed, radial analysis, idelsqp; the
    log coefficient is restored and zeroat1 applied.  ns = 3 nr / 4 (integer),
    ns0 = min(ns + 4, nr), 1-based; the smoothed lines run to the column's
    end (row nrdim), past the physical rows."""
    nrc, npc, nzc, nzcu, nrcs = chop_index(s, kit)
    nr = kit.nr
    ns = nr * 3 // 4
    ns0 = min(ns + 4, nr)
    so = s.copy()
    ln = s.ln
    so.ln = 0.0
    delsqp(so, kit)
    rtrans_backward(so, kit)
    w = ((1.0 - kit.x[:nr]) ** 2)[:, None]
    act = [kk for kk in range(kit.nzdim) if (kk + 1 <= nzc) or (kk + 1 >= nzcu)]
    for mm in range(npc):
        so.e[:nr, mm, act] = so.e[:nr, mm, act] * w
    so.e[ns0 - 1:, 1:npc, :] = 0.0
    for mm in range(npc):
        line = so.e[ns - 1:, mm, act]
        for _ in range(5):
            line = smooth(line)
        so.e[ns - 1:, mm, act] = line
    for mm in range(npc):
        so.e[:nr, mm, act] = so.e[:nr, mm, act] / w
    so.space = 'FFF'
    rtrans_forward(so, kit)
    idelsqp(so, kit)
    so.ln = ln
    zeroat1(so, kit)
    s.assign(so)


# ----------------------------------------------------------------------------
# banded operators
# ----------------------------------------------------------------------------
def _slices(ks):
    """z slices of a kk-loop list and how often each is visited."""
    u, c = np.unique(np.asarray(ks, dtype=np.int64), return_counts=True)
    return u, c


def _band_apply(band, X):
    """Y[i] = sum_d band[d + h, i] X[i + d] along the first axis of X."""
    h = band.shape[0] // 2
    nn = X.shape[0]
    Y = np.zeros_like(X)
    for d in range(-h, h + 1):
        c = band[d + h]
        if d >= 0:
            if nn - d > 0:
                Y[:nn - d] += c[:nn - d, None] * X[d:]
        elif nn + d > 0:
            Y[-d:] += c[-d:, None] * X[:nn + d]
    return Y


def _apply_per_slice(so, mm, nn, ks, fn):
    """Apply fn to the (nn x slices) block of column mm once per visit of each slice."""
    u, c = _slices(ks)
    for rep in range(1, int(c.max()) + 1 if len(c) else 1):
        sel = u[c >= rep]
        so.e[:nn, mm, sel] = fn(so.e[:nn, mm, sel], sel)


def xxdx(s, kit):
    """r d/dr (mlegs_scalar_ops.f90:418-463)."""
    nrc, npc, nzc, nzcu, nrcs = chop_index(s, kit)
    if s.space != 'FFF':
        trans(s, 'FFF', kit)
    so = s.copy()
    so.ln = 0.0
    ks = zlist(kit, nzc, nzcu)
    for mm in range(npc):
        nn = nrcs[mm]
        if nn == 0:
            continue
        B = kit.band_xxdx(kit.m[mm], nn)
        _apply_per_slice(so, mm, nn, ks, lambda X, sel: _band_apply(B, X))
    so.e[0, 0, 0] += 1.0 / np.exp(kit.lognorm[0, 0]) * s.ln
    so.e[1, 0, 0] += 1.0 / np.exp(kit.lognorm[1, 0]) * s.ln
    s.assign(so)


def _lap_ln(so, kit, ln, sign=1.0):
    ell2 = kit.ell ** 2
    so.e[0, 0, 0] += sign * 4.0 / 3.0 / ell2 / np.exp(kit.lognorm[0, 0]) * ln
    so.e[1, 0, 0] -= sign * 2.0 / 1.0 / ell2 / np.exp(kit.lognorm[1, 0]) * ln
    so.e[2, 0, 0] += sign * 2.0 / 3.0 / ell2 / np.exp(kit.lognorm[2, 0]) * ln


def del2h(s, kit):
    """Horizontal Laplacian (mlegs_scalar_ops.f90:465-518)."""
    nrc, npc, nzc, nzcu, nrcs = chop_index(s, kit)
    if s.space != 'FFF':
        trans(s, 'FFF', kit)
    so = s.copy()
    so.ln = 0.0
    ks = zlist(kit, nzc, nzcu)
    for mm in range(npc):
        nn = nrcs[mm]
        if nn == 0:
            continue
        H = kit.band_del2h(kit.m[mm], nn)
        _apply_per_slice(so, mm, nn, ks, lambda X, sel: _band_apply(H, X))
    _lap_ln(so, kit, s.ln)
    s.assign(so)


def del2(s, kit):
    """Laplacian (mlegs_scalar_ops.f90:520-573): del2h - ak^2 per slice."""
    nrc, npc, nzc, nzcu, nrcs = chop_index(s, kit)
    if s.space != 'FFF':
        trans(s, 'FFF', kit)
    so = s.copy()
    so.ln = 0.0
    ks = zlist(kit, nzc, nzcu)
    for mm in range(npc):
        nn = nrcs[mm]
        if nn == 0:
            continue
        H = kit.band_del2h(kit.m[mm], nn)
        _apply_per_slice(so, mm, nn, ks,
                         lambda X, sel: _band_apply(H, X) - (kit.ak[sel] ** 2)[None, :] * X)
    _lap_ln(so, kit, s.ln)
    s.assign(so)


def _ab(band, shift):
    """LAPACK band storage (kl = ku = 2) of the diagonals plus shift on the main one."""
    nn = band.shape[1]
    ab = np.zeros((5, nn))
    for d in range(-2, 3):
        # A[i, i + d] sits at ab[2 - d, i + d]
        c = band[d + 2]
        if d >= 0 and nn - d > 0:
            ab[2 - d, d:] = c[:nn - d]
        elif d < 0 and nn + d > 0:
            ab[2 - d, :nn + d] = c[-d:]
    ab[2] = ab[2] + shift
    return ab


def _band_solve(A, b, kl=2, ku=2):
    """LAPACK gbsv on a dense matrix with known bandwidths."""
    n = A.shape[0]
    ab = np.zeros((kl + ku + 1, n))
    for k in range(-kl, ku + 1):
        d = np.diagonal(A, k)
        j0 = max(k, 0)
        ab[ku - k, j0:j0 + len(d)] = d
    return solve_banded((kl, ku), ab, b)


def _gbsv(ab, b):
    """LAPACK zgbsv (kl = ku = 2, partial pivoting), as zgbtrf + zgbtrs."""
    abx = np.zeros((7, ab.shape[1]), dtype=np.complex128)
    abx[2:] = ab
    _, _, x, info = zgbsv(2, 2, abx, b)
    if info != 0:
        raise ValueError('lurc: lu factorization resulted in failure')
    return x


def _helm_solves(so, kit, mm, nn, ks, alpha):
    """(del2 + alpha) x = b on every visit of every slice of column mm."""
    H = kit.band_del2h(kit.m[mm], nn)
    for k in ks:
        ab = _ab(H, -(kit.ak[k] ** 2))
        if alpha != 0.0:
            ab[2] = ab[2] + alpha
        so.e[:nn, mm, k] = _gbsv(ab, so.e[:nn, mm, k])


def ihelm(s, alpha, kit):
    """Solve (del2 + alpha) x = s per (m, k) (mlegs_scalar_ops.f90:791-854)."""
    nrc, npc, nzc, nzcu, nrcs = chop_index(s, kit)
    if s.space != 'FFF':
        trans(s, 'FFF', kit)
    if abs(alpha) < 5e-14:
        raise ValueError('ihelm: alpha equals to zero. Inversion of 0*identity is impossible')
    so = s.copy()
    so.ln = s.ln / alpha
    _lap_ln(so, kit, so.ln, sign=-1.0)
    ks = zlist(kit, nzc, nzcu)
    for mm in range(npc):
        nn = nrcs[mm]
        if nn == 0:
            continue
        _helm_solves(so, kit, mm, nn, ks, alpha)
    s.assign(so)


def _check_power(name, power):
    if power % 2 != 0 or power < 4:
        raise ValueError('%s: even power greater than or equal to 4' % name)
    if power > 8:
        raise ValueError('%s: power must be less than or equal to 8 (supported power = 4, 6 or 8)' % name)


def helmp(s, power, alpha, beta, kit):
    """del2^(power/2) + beta del2 + alpha (mlegs_scalar_ops.f90:856-898).

    As the reference: so%e = sp%e + beta*s2%e + alpha*s%e over every entry,
    with s2 = del2(s) and sp = del2^(power/2)(s) (del2's copy semantics:
    entries del2 does not visit keep the value of s), so%ln = alpha*s%ln.
    """
    _check_power('helmp', power)
    if s.space != 'FFF':
        trans(s, 'FFF', kit)
    s2 = s.copy()
    del2(s2, kit)
    sp = s2.copy()
    for _ in range(power // 2 - 1):
        del2(sp, kit)
    so = s.copy()
    so.e = sp.e + beta * s2.e + alpha * s.e
    so.ln = alpha * s.ln
    s.assign(so)


def helmp_band(kit, mval, akval, nn, power, alpha, beta):
    """ihelmp's system matrix of one (m, kappa) (mlegs_scalar_ops.f90:950-958):
    leg_del2^(power/2) + beta leg_del2 (products of the nn x nn truncations),
    banded with kl = ku = power, plus alpha on the diagonal."""
    D = kit.leg_del2(mval, akval, nn)
    H = D.copy()
    for _ in range(power // 2 - 1):
        H = D @ H
    W = H + beta * D
    W[np.abs(np.subtract.outer(np.arange(nn), np.arange(nn))) > power] = 0.0
    W[np.arange(nn), np.arange(nn)] += alpha
    return W


def ihelmp(s, power, alpha, beta, kit):
    """Solve (del2^(power/2) + beta del2 + alpha) x = s per (m, k)
    (mlegs_scalar_ops.f90:900-1000)."""
    _check_power('ihelmp', power)
    nrc, npc, nzc, nzcu, nrcs = chop_index(s, kit)
    if s.space != 'FFF':
        trans(s, 'FFF', kit)
    if abs(alpha) < 5e-14:
        raise ValueError('ihelmp: alpha equals to zero. Inversion of 0*identity is impossible')
    so = s.copy()
    so.ln = s.ln / alpha
    ell2 = kit.ell ** 2
    bl2 = np.zeros(nrc)
    bl2[0] = 4.0 / 3.0 / ell2 / np.exp(kit.lognorm[0, 0]) * so.ln
    bl2[1] = -2.0 / ell2 / np.exp(kit.lognorm[1, 0]) * so.ln
    bl2[2] = 2.0 / 3.0 / ell2 / np.exp(kit.lognorm[2, 0]) * so.ln
    bl = bl2.copy()
    for _ in range(power // 2 - 1):
        bl = kit.leg_del2(kit.m[0], kit.ak[0], nrc) @ bl
    so.e[:nrc, 0, 0] = so.e[:nrc, 0, 0] - bl - beta * bl2
    ks = zlist(kit, nzc, nzcu)
    for mm in range(npc):
        nn = nrcs[mm]
        if nn == 0:
            continue
        for k in ks:
            W = helmp_band(kit, kit.m[mm], kit.ak[k], nn, power, alpha, beta)
            so.e[:nn, mm, k] = _band_solve(W, so.e[:nn, mm, k], power, power)
    s.assign(so)


def idel2_proln(s, kit):
    """Inverse Laplacian, log coefficient from the (m,k)=(0,0) solve (:671-760)."""
    nrc, npc, nzc, nzcu, nrcs = chop_index(s, kit)
    if s.space != 'FFF':
        trans(s, 'FFF', kit)
    so = s.copy()
    ell2 = kit.ell ** 2
    ln = kit.lognorm[:, 0]
    ks = zlist(kit, nzc, nzcu)
    for mm in range(npc):
        nn = nrcs[mm]
        if nn == 0:
            continue
        first00 = (mm == 0 and ks and ks[0] == 0)
        if first00:
            # bordered system: column 0 carries the log term's Laplacian
            ab = _ab(kit.band_del2h(0, nn), -(kit.ak[0] ** 2))
            ab[2, 0] += 4.0 / 3.0 / ell2
            if nn > 1:
                ab[3, 0] -= 2.0 / 1.0 / ell2 * np.exp(ln[0] - ln[1])
            if nn > 2:
                ab[4, 0] += 2.0 / 3.0 / ell2 * np.exp(ln[0] - ln[2])
            so.e[:nn, 0, 0] = _gbsv(ab, so.e[:nn, 0, 0])
            so.ln = float(np.real(so.e[0, 0, 0])) * np.exp(ln[0])
            _helm_solves(so, kit, mm, nn, ks[1:], 0.0)
        else:
            _helm_solves(so, kit, mm, nn, ks, 0.0)
    s.assign(so)


def idel2_preln(s, kit, preln):
    """Inverse Laplacian with a prescribed log coefficient (:575-669)."""
    nrc, npc, nzc, nzcu, nrcs = chop_index(s, kit)
    if s.space != 'FFF':
        trans(s, 'FFF', kit)
    so = s.copy()
    ell2 = kit.ell ** 2
    ln = kit.lognorm[:, 0]
    ks = zlist(kit, nzc, nzcu)
    for mm in range(npc):
        nn = nrcs[mm]
        if nn == 0:
            continue
        for idx, k in enumerate(ks):
            if not (mm == 0 and idx == 0 and k == 0):
                _helm_solves(so, kit, mm, nn, [k], 0.0)
                continue
            # the first equation prescribes the log term; the rest shift down a row
            A = kit.leg_del2(kit.m[mm], kit.ak[k], nn)
            P = np.zeros((nn, nn))
            P[1:nn, :] = A[0:nn - 1, :]
            P[0, 0] = 1.0
            P[1, 0] += 4.0 / 3.0 / ell2
            P[2, 0] -= 2.0 / 1.0 / ell2 * np.exp(ln[0] - ln[1])
            P[3, 0] += 2.0 / 3.0 / ell2 * np.exp(ln[0] - ln[2])
            b = so.e[:nn, mm, k].copy()
            rhs = np.empty_like(b)
            rhs[1:] = b[:-1]
            rhs[0] = preln / np.exp(ln[0])
            so.e[:nn, mm, k] = _band_solve(P, rhs, kl=3, ku=2)
            so.ln = float(np.real(so.e[0, mm, k])) * np.exp(ln[0])
    s.assign(so)


# ----------------------------------------------------------------------------
# time steppers
# ----------------------------------------------------------------------------
def _need_fff(*ss):
    for s in ss:
        if s.space != 'FFF':
            raise ValueError('fefe: all input scalars must be in FFF for time stepping')


def _diffusive(s, kit, p):
    """visc del2 s, or visc del2 s + hypervisc del^p s (mlegs_scalar_ops.f90:1077-1090);
    inviscid: the data zeroed, the copy's log term kept."""
    sv = s.copy()
    if p.hyperpow == 0:
        if p.visc < 5e-14:
            sv.e = np.zeros_like(sv.e)
        else:
            del2(sv, kit)
            sv.e = sv.e * p.visc
            sv.ln = sv.ln * p.visc
    else:
        g = p.hypervisc * (-1.0) ** (p.hyperpow // 2 + 1)
        helmp(sv, p.hyperpow, 0.0, p.visc / g, kit)
        sv.e = sv.e * g
        sv.ln = sv.ln * g
    return sv


def fefe(s, nl, dt, kit, p):
    """Forward Euler on both terms (mlegs_scalar_ops.f90:1065-1096)."""
    _need_fff(s, nl)
    sv = _diffusive(s, kit, p)
    s.e = s.e + dt * (nl.e + sv.e)
    s.ln = s.ln + dt * (nl.ln + sv.ln)


def abab(s, s_p, nl, nl_p, dt, kit, p, is_2nd_svis_p=False):
    """Adams-Bashforth on both terms (mlegs_scalar_ops.f90:1098-1155).  In the
    inviscid case the reference zeroes the current diffusive term twice and
    never the previous one, so s_p enters as its own diffusive term."""
    _need_fff(s, nl, s_p, nl_p)
    sv = _diffusive(s, kit, p)
    svp = s_p.copy()
    if not is_2nd_svis_p:
        if p.hyperpow == 0:
            if not p.visc < 5e-14:
                del2(svp, kit)
                svp.e = svp.e * p.visc
                svp.ln = svp.ln * p.visc
        else:
            g = p.hypervisc * (-1.0) ** (p.hyperpow // 2 + 1)
            helmp(svp, p.hyperpow, 0.0, p.visc / g, kit)
            svp.e = svp.e * g
            svp.ln = svp.ln * g
    s.e = s.e + dt * (1.5 * (nl.e + sv.e) - 0.5 * (nl_p.e + svp.e))
    s.ln = s.ln + dt * (1.5 * (nl.ln + sv.ln) - 0.5 * (nl_p.ln + svp.ln))
    s_p.assign(s)
    nl_p.assign(nl)


def febe(s, nl, dt, kit, p):
    """Forward/backward Euler (mlegs_scalar_ops.f90:1157-1198)."""
    _need_fff(s, nl)
    if p.hyperpow != 0:
        # (:1181-1195) ihelmp with alpha = -1/(dt g), beta = visc/g,
        # g = hypervisc (-1)^(hyperpow/2+1)
        g = p.hypervisc * (-1.0) ** (p.hyperpow // 2 + 1)
        a, b = -1.0 / (dt * g), p.visc / g
        sh = s.copy()
        sh.e = a * (s.e + dt * nl.e)
        sh.ln = a * (s.ln + dt * nl.ln)
        ihelmp(sh, p.hyperpow, a, b, kit)
        s.assign(sh)
        return
    if p.visc < 5e-14:
        raise ValueError('febe: inviscid case and no linear term in rhs. semi-implicit time adv is impossible')
    sh = s.copy()
    sh.e = s.e + dt * nl.e
    sh.ln = s.ln + dt * nl.ln
    a = -1.0 / (dt * p.visc)
    sh.ln = a * sh.ln
    sh.e = a * sh.e
    ihelm(sh, a, kit)
    s.assign(sh)


def abcn(s, s_p, nl, nl_p, dt, kit, p):
    """Adams-Bashforth / Crank-Nicolson (mlegs_scalar_ops.f90:1200-1262)."""
    _need_fff(s, nl, s_p, nl_p)
    sh = s.copy()
    sh.e = s.e + dt * (1.5 * nl.e - 0.5 * nl_p.e)
    sh.ln = s.ln + dt * (1.5 * nl.ln - 0.5 * nl_p.ln)
    svis = s.copy()
    if p.hyperpow != 0:
        # (:1226-1230, 1241-1255) svis = g helmp(s, p, 0, b), then ihelmp
        g = p.hypervisc * (-1.0) ** (p.hyperpow // 2 + 1)
        b = p.visc / g
        helmp(svis, p.hyperpow, 0.0, b, kit)
        svis.e = svis.e * g
        a = -2.0 / (dt * g)
        sh.ln = a * (sh.ln + dt / 2.0 * svis.ln)
        sh.e = a * (sh.e + dt / 2.0 * svis.e)
        ihelmp(sh, p.hyperpow, a, b, kit)
        s.assign(sh)
        s_p.assign(s)
        nl_p.assign(nl)
        return
    if p.visc < 5e-14:
        raise ValueError('abcn: inviscid case and no linear term in rhs. semi-implicit time adv is impossible')
    del2(svis, kit)
    svis.e = svis.e * p.visc
    a = -2.0 / (dt * p.visc)
    sh.ln = a * (sh.ln + dt / 2.0 * svis.ln)
    sh.e = a * (sh.e + dt / 2.0 * svis.e)
    ihelm(sh, a, kit)
    s.assign(sh)
    s_p.assign(s)
    nl_p.assign(nl)


# ----------------------------------------------------------------------------
# vector-field operations
# ----------------------------------------------------------------------------
def vecprod(vr, vp, vz, ur, up, uz, kit):
    """v <- v x u on the two packed phi points (mlegs_scalar_ops.f90:1264-1306)."""
    for s in (vr, vp, vz, ur, up, uz):
        if s.space != 'PPP':
            raise ValueError('vector_product: to compute v x u, all inputs must be in PPP')
    nr, nph = kit.nr, kit.np // 2
    sl = (slice(0, nr), slice(0, nph), slice(None))
    a1, a2, a3 = vr.e[sl].real, vp.e[sl].real, vz.e[sl].real
    b1, b2, b3 = ur.e[sl].real, up.e[sl].real, uz.e[sl].real
    c1, c2, c3 = vr.e[sl].imag, vp.e[sl].imag, vz.e[sl].imag
    d1, d2, d3 = ur.e[sl].imag, up.e[sl].imag, uz.e[sl].imag
    outs = []
    for re, im in ((a2 * b3 - a3 * b2, c2 * d3 - c3 * d2),
                   (a3 * b1 - a1 * b3, c3 * d1 - c1 * d3),
                   (a1 * b2 - a2 * b1, c1 * d2 - c2 * d1)):
        o = Scalar(kit, 'PPP')
        o.e[sl] = re + IU * im
        outs.append(o)
    vr.assign(outs[0])
    vp.assign(outs[1])
    vz.assign(outs[2])


def tp2vec(psi, chi, kit):
    """(psi, chi) -> (vr, vp, vz) in PPP (vector_reconstruction, :1455-1545)."""
    if psi.space != 'FFF' or chi.space != 'FFF':
        raise ValueError('vector_projection: psi and chi must be in FFF')
    nr = kit.nr
    r = kit.ell * np.sqrt((1.0 + kit.x) / (1.0 - kit.x))
    ur = chi.copy(); ur.chop_offset(3)
    up = psi.copy(); up.chop_offset(3)
    uz = chi.copy(); uz.chop_offset(3)
    nrc, npc, nzc, nzcu, nrcs = chop_index(ur, kit)
    xxdx(ur, kit)
    xxdx(up, kit)
    for mm in range(npc):
        nn = nrcs[mm]
        mv = kit.m[mm]
        for k in range(kit.nzdim):
            if (k + 1 <= nzc) or (k + 1 >= nzcu):
                kv = kit.ak[k]
                ur.e[:nn, mm, k] = IU * mv * psi.e[:nn, mm, k] + IU * kv * ur.e[:nn, mm, k]
                up.e[:nn, mm, k] = -up.e[:nn, mm, k] - mv * kv * uz.e[:nn, mm, k]
    ur.chop_offset(0); up.chop_offset(0)
    chop(ur, kit); chop(up, kit)
    trans(ur, 'PPP', kit); trans(up, 'PPP', kit)
    ur.e[:nr, :, :] /= r[:nr, None, None]
    up.e[:nr, :, :] /= r[:nr, None, None]
    del2h(uz, kit)
    uz.e = -uz.e
    uz.chop_offset(0)
    chop(uz, kit)
    trans(uz, 'PPP', kit)
    return ur, up, uz


def tp2curlvec(psi, chi, kit):
    """curl V = tp2vec(-del2 chi, psi) (curl_vector_reconstruction, :1547-1560)."""
    md = chi.copy()
    del2(md, kit)
    md.e = -md.e
    return tp2vec(md, psi, kit)


def vec2tp_tables(kit, mm, nn):
    """V, D, T projection tables of vector_projection (:1387-1411) for one m."""
    nrh = kit.nr // 2
    mv = kit.m[mm]
    fff = (1.0 - kit.x[:nrh] ** 2) / kit.w[:nrh]
    pfd = kit.pf[:, :nn + 1, mm] @ kit.leg_xxdx(mv, nn + 1)
    v = np.zeros((nn, nrh)); d = np.zeros((nn, nrh)); t = np.zeros((nn, nrh))
    for i in range(1, nn + 1):
        n = max(1, mv + i - 1)
        v[i - 1] = kit.pf[:, i - 1, mm] / n / (n + 1) / fff
        d[i - 1] = pfd[:, i - 1] / n / (n + 1) / fff
        t[i - 1] = kit.pf[:, i - 1, mm] * kit.w[:nrh]
    return v, d, t


def _eomul(a, b):
    """Parity-folded product, odd rows <- even part (mlegs_scalar_ops.f90:2067-2104).

    b holds the nr collocation values along its first axis (further axes are
    independent columns)."""
    nj = b.shape[0]
    njh = nj // 2
    be = b[:njh] + b[nj - 1:njh - 1:-1]
    bo = b[:njh] - b[nj - 1:njh - 1:-1]
    c = np.zeros((a.shape[0],) + b.shape[1:], dtype=np.complex128)
    c[0::2] = a[0::2] @ be
    if a.shape[0] > 1:
        c[1::2] = a[1::2] @ bo
    return c


def _oemul(a, b):
    """Parity-folded product, odd rows <- odd part (mlegs_scalar_ops.f90:2108-2145)."""
    nj = b.shape[0]
    njh = nj // 2
    be = b[:njh] + b[nj - 1:njh - 1:-1]
    bo = b[:njh] - b[nj - 1:njh - 1:-1]
    c = np.zeros((a.shape[0],) + b.shape[1:], dtype=np.complex128)
    c[0::2] = a[0::2] @ bo
    if a.shape[0] > 1:
        c[1::2] = a[1::2] @ be
    return c


def vec2tp(vr, vp, vz, kit):
    """Toroidal-poloidal projection of a PPP vector field (vector_projection, :1308-1453).

    Returns (psi, chi) in FFF.
    """
    for s in (vr, vp, vz):
        if s.space != 'PPP':
            raise ValueError('vector_projection: vr, vp, vz must be in PPP')
    nr = kit.nr
    r = kit.ell * np.sqrt((1.0 + kit.x) / (1.0 - kit.x))
    psi = Scalar(kit, 'FFF')
    chi = Scalar(kit, 'FFF')
    ur = vr.copy(); up = vp.copy(); uz = vz.copy()
    ur.e[:nr, :, :] *= r[:nr, None, None]
    up.e[:nr, :, :] *= r[:nr, None, None]
    trans(ur, 'PFP', kit)
    if kit.nz > 1:
        vertical_fft_forward(ur, kit)
    ur.space = 'PFF'
    trans(up, 'FFF', kit)
    inf = calcat1(up, kit)
    psi.ln = -0.5 * float(np.real(inf[0]))
    rtrans_backward(up, kit)
    up.space = 'PFF'
    up.e[:nr, 0, 0] += psi.ln * (1.0 + kit.x)
    trans(uz, 'PFP', kit)
    if kit.nz > 1:
        vertical_fft_forward(uz, kit)
    uz.space = 'PFF'
    ur.chop_offset(3); up.chop_offset(3); uz.chop_offset(3)
    nrc, npc, nzc, nzcu, nrcs = chop_index(ur, kit)
    # a slice visited twice is recomputed from the same inputs: once suffices
    ks = np.unique(np.asarray(zlist(kit, nzc, nzcu), dtype=np.int64))
    kv = kit.ak[ks][None, :]
    for mm in range(npc):
        nn = nrcs[mm]
        mv = kit.m[mm]
        if nn == 0:
            continue
        v, d, t = vec2tp_tables(kit, mm, nn)
        r_ = ur.e[:nr, mm][:, ks]
        p_ = up.e[:nr, mm][:, ks]
        z_ = uz.e[:nr, mm][:, ks]
        zero = np.zeros((nn, len(ks)), complex)
        w1 = _eomul(v, r_) if mv != 0 else zero
        w2 = _oemul(d, p_)
        psi.e[:nn, mm, ks] = -IU * mv * w1 - w2
        w1 = _oemul(d, r_)
        w2 = _eomul(v, p_) if mv != 0 else zero
        c = _eomul(t, z_)
        chi.e[:nn, mm, ks] = IU * kv * w1 + mv * kv * w2 - c
    idel2_proln(chi, kit)
    psi.chop_offset(0); chi.chop_offset(0)
    chop(psi, kit); chop(chi, kit)
    zeroat1(psi, kit); zeroat1(chi, kit)
    return psi, chi
