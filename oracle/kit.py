"""Transform kit of the oracle (test infrastructure only).

Restates ``src/submodules/mlegs_spectfm_init.f90`` (tfm_kit_init for the 3D
kit, gauss_legendre, leg_lognorm, leg_tbl) and the closed-form band operators
of ``src/submodules/mlegs_spectfm_diff.f90``.

Table accuracy.  The reference evaluates its P_n^m table with a 50-digit
FM recurrence on Q_n = P_n^m / (2n-1)!! and then multiplies by
exp(log_fact(n) + lognorm) formed from FP64 inputs, which leaves a relative
error of about |log_fact + lognorm| * 1.1e-16 (1e-13 at n ~ 270).  This
oracle uses an 80-bit normalised three-term recurrence, accurate to ~1e-17
relative; agreement with the reference's table is therefore ~1e-13 and not
1e-16 (documented, not replicated).
"""
import numpy as np

PI = np.pi


def validate(p):
    """Size checks of tfm_kit_init (mlegs_spectfm_init.f90:35-62)."""
    def _235(n):
        while n % 2 == 0 or n % 3 == 0 or n % 5 == 0:
            if n % 2 == 0:
                n //= 2
            if n % 3 == 0:
                n //= 3
            if n % 5 == 0:
                n //= 5
        return n == 1

    if not (p.nr > 0 and p.nr % 2 == 0):
        raise ValueError('tfm_kit_init: nr must be even')
    if not (p.np > 0 and (p.np == 1 or p.np % 2 == 0)):
        raise ValueError('tfm_kit_init: np must be even')
    if p.np != 1 and not _235(p.np):
        raise ValueError('tfm_kit_init: np must only have factors of 2, 3 and 5')
    if not (p.nz > 0 and (p.nz == 1 or p.nz % 2 == 0)):
        raise ValueError('tfm_kit_init: nz must be even')
    if p.nz != 1 and not _235(p.nz):
        raise ValueError('tfm_kit_init: nz must only have factors of 2, 3 and 5')
    if p.nrchop > p.nr:
        raise ValueError('tfm_kit_init: nrchop must be smaller than or equal to nr')
    if p.np != 1 and p.npchop * 2 > p.np + 2:
        raise ValueError('tfm_kit_init: npchop <= np/2 + 1 must be satisfied')
    if p.nz != 1 and p.nzchop * 2 > p.nz + 2:
        raise ValueError('tfm_kit_init: nzchop <= nz/2 + 1 must be satisfied')


def gauss_legendre(n):
    """Newton iteration on P_n from cos((i-1/4)pi/(n+1/2)), tolerance 1e-15.

    Follows mlegs_spectfm_init.f90:181-214 (x ascending, x(1) ~ -1); the
    iteration runs per root until |z - z1| <= 1e-15 as the reference does.
    """
    m = int(np.ceil((n + 1) / 2.0))
    i = np.arange(1, m + 1, dtype=np.float64)
    z = np.cos(PI * (i - 0.25) / (n + 0.5))
    z1 = z + 1.0
    pp = np.zeros_like(z)
    active = np.abs(z - z1) > 1e-15
    while np.any(active):
        za = z[active]
        p1 = np.ones_like(za)
        p2 = np.zeros_like(za)
        for j in range(1, n + 1):
            p3 = p2
            p2 = p1
            p1 = ((2 * j - 1) * za * p2 - (j - 1) * p3) / j
        ppa = n * (za * p1 - p2) / (za * za - 1)
        pp[active] = ppa
        z1[active] = za
        z[active] = za - p1 / ppa
        active = np.abs(z - z1) > 1e-15
    x = np.zeros(n)
    w = np.zeros(n)
    wv = 2.0 / ((1 - z * z) * pp * pp)
    # sequential writes as in the reference: for even n the last root
    # (i = n/2 + 1) overwrites the middle pair
    for ii in range(m):
        x[ii] = -z[ii]
        x[n - 1 - ii] = z[ii]
        w[ii] = wv[ii]
        w[n - 1 - ii] = wv[ii]
    return x, w


def leg_lognorm(ndim, ms):
    """0.5*log((2n+1)/2 (n-m)!/(n+m)!) by recurrence (mlegs_spectfm_init.f90:218-250)."""
    me = int(max(abs(int(v)) for v in ms))
    wk = np.zeros(me + 1)
    wk[0] = np.log(0.5)
    for m in range(1, me + 1):
        wk[m] = wk[m - 1] + np.log(2.0 * m + 1.0) - np.log(2.0 * m * (2.0 * m - 1.0) ** 2)
    out = np.zeros((ndim, len(ms)))
    for mm, mv in enumerate(ms):
        m = abs(int(mv))
        out[0, mm] = wk[m]
        for nn in range(2, ndim + 1):
            n = m + nn - 1
            out[nn - 1, mm] = (out[nn - 2, mm] + np.log(2.0 * n + 1.0) - np.log(2.0 * n - 1.0)
                               + np.log(1.0 * (n - m)) - np.log(1.0 * (n + m)))
    return 0.5 * out


def leg_tbl(x, ne, ms):
    """Fully normalised P_n^m(x), n = m..m+ne-1, Condon-Shortley phase.

    Same quantity as leg_tbl (mlegs_spectfm_init.f90:254-300): orthonormal on
    [-1,1].  Evaluated here with the normalised recurrence in 80-bit floats:
      P_m^m = -sqrt((2m+1)/(2m)) sqrt(1-x^2) P_{m-1}^{m-1},  P_0^0 = 1/sqrt(2)
      P_{m+1}^m = sqrt(2m+3) x P_m^m
      P_n^m = a (x P_{n-1}^m - b P_{n-2}^m),
        a = sqrt((4n^2-1)/(n^2-m^2)),  b = sqrt(((n-1)^2-m^2)/(4(n-1)^2-1)).
    Returns tbl[len(x), ne, len(ms)] in float64.
    """
    ld = np.longdouble
    xl = np.asarray(x, dtype=ld)
    s = np.sqrt(ld(1) - xl * xl)
    out = np.zeros((len(x), ne, len(ms)), dtype=np.float64)
    mmax = int(max(ms))
    pmm = np.full(len(x), ld(1) / np.sqrt(ld(2)))
    seeds = {0: pmm.copy()}
    for m in range(1, mmax + 1):
        pmm = -np.sqrt(ld(2 * m + 1) / ld(2 * m)) * s * pmm
        seeds[m] = pmm.copy()
    for mm, mv in enumerate(ms):
        m = int(mv)
        p0 = seeds[m]
        out[:, 0, mm] = p0.astype(np.float64)
        if ne == 1:
            continue
        p1 = np.sqrt(ld(2 * m + 3)) * xl * p0
        out[:, 1, mm] = p1.astype(np.float64)
        for nn in range(2, ne):
            n = m + nn
            a = np.sqrt(ld(4 * n * n - 1) / ld(n * n - m * m))
            b = np.sqrt(ld((n - 1) ** 2 - m * m) / ld(4 * (n - 1) ** 2 - 1))
            p2 = a * (xl * p1 - b * p0)
            out[:, nn, mm] = p2.astype(np.float64)
            p0, p1 = p1, p2
    return out


class Kit3D:
    """tfm_kit_3d (mlegs_spectfm.f90:55-62) built as tfm_kit_init does (:6-154)."""

    def __init__(self, p):
        validate(p)
        self.p = p
        nr, np_, nz = p.nr, p.np, p.nz
        self.nr, self.np, self.nz = nr, np_, nz
        self.nrchop, self.npchop, self.nzchop = p.nrchop, p.npchop, p.nzchop
        self.ell, self.zlen = p.ell, p.zlen
        self.x, self.w = gauss_legendre(nr)
        self.lnx = -np.log(1.0 - self.x)                       # :67-68
        self.r = p.ell * np.sqrt((1.0 + self.x) / (1.0 - self.x))  # :112
        self.chops = np.array([max(p.nrchop - i, 0) for i in range(p.npchop)])  # :114-115
        self.nrdim = nr + max(3, p.hyperpow)                   # :117
        self.p_pts = np.array([2 * PI / np_ * i for i in range(np_ + 1)])  # :119-120
        self.m = np.arange(p.npchop)                           # :122-123
        self.npdim = np_ // 2 + 1                              # :125
        self.chopp = p.npchop
        self.z = np.array([p.zlen / nz * i for i in range(nz + 1)])  # :129-130
        ak = np.array([2.0 * PI / p.zlen * i for i in range(-nz, 0)], dtype=np.float64)
        ak[: nz // 2 + 1] = [2.0 * PI / p.zlen * i for i in range(0, nz // 2 + 1)]  # :132-134
        self.ak = ak
        self.nzdim = nz
        self.chopzl = p.nzchop                                 # :138
        self.chopzu = nz - p.nzchop + 2                        # :139
        self.ne = p.nrchop + 14
        self.lognorm = leg_lognorm(self.ne, self.m)            # :141-142
        self.pf = leg_tbl(self.x[: nr // 2], self.ne, self.m)  # :144-145
        self.at0 = leg_tbl(np.array([-1.0 + 1e-15]), p.nrchop, [0])[0, :, 0]  # :147-148
        self.at1 = leg_tbl(np.array([1.0 - 1e-15]), p.nrchop, [0])[0, :, 0]   # :149-150

    # ------------------------------------------------------------------
    # band operators (mlegs_spectfm_diff.f90), returned as dense nn x nn
    # matrices in the normalised basis: out = M @ coeffs.
    # ------------------------------------------------------------------
    def _scale(self, g, am):
        ln = self.lognorm[: g.shape[0], am]
        return g * np.exp(ln[None, :] - ln[:, None])

    def leg_xxdx(self, mval, nn):
        """r d/dr as (1-x^2) d/dx, tridiagonal (mlegs_spectfm_diff.f90:6-43)."""
        am = abs(int(mval))
        g = np.zeros((nn, nn))
        for i in range(1, nn + 1):
            n = am + (i - 1)
            if i - 1 >= 1:
                g[i - 1, i - 2] = -(n - 1.0) * (n - am) / (2.0 * n - 1.0)
            if i + 1 <= nn:
                g[i - 1, i] = (n + 2.0) * (n + am + 1.0) / (2.0 * n + 3.0)
        return self._scale(g, am)

    def _del2_core(self, am, nn):
        key = (am, nn)
        cache = self.__dict__.setdefault('_core_cache', {})
        if key in cache:
            return cache[key].copy()
        g = np.zeros((nn, nn))
        for i in range(1, nn + 1):
            n = float(am + (i - 1))
            if i - 2 >= 1:
                g[i - 1, i - 3] = -(n - am - 1.0) * (n - am) * (n - 2.0) * (n - 1.0) / (2.0 * n - 3.0) / (2.0 * n - 1.0)
            if i - 1 >= 1:
                g[i - 1, i - 2] = 2.0 * n * (n - am) * (n - 1.0) / (2.0 * n - 1.0)
            g[i - 1, i - 1] = (-2.0 * n * (n + 1.0) * (3.0 * n * n + 3.0 * n - am * am - 2.0)) / (2.0 * n - 1.0) / (2.0 * n + 3.0)
            if i + 1 <= nn:
                g[i - 1, i] = 2.0 * (n + 1.0) * (n + am + 1.0) * (n + 2.0) / (2.0 * n + 3.0)
            if i + 2 <= nn:
                g[i - 1, i + 1] = -(n + am + 1.0) * (n + am + 2.0) * (n + 3.0) * (n + 2.0) / (2.0 * n + 3.0) / (2.0 * n + 5.0)
        g = g / self.ell ** 2
        cache[key] = g
        return g.copy()

    def leg_del2h(self, mval, nn):
        """1/r d/dr(r d/dr) - m^2/r^2, pentadiagonal (mlegs_spectfm_diff.f90:45-96)."""
        am = abs(int(mval))
        return self._scale(self._del2_core(am, nn), am)

    def leg_del2(self, mval, akval, nn):
        """del2h - ak^2, pentadiagonal (mlegs_spectfm_diff.f90:98-152)."""
        am = abs(int(mval))
        g = self._del2_core(am, nn)
        g[np.arange(nn), np.arange(nn)] -= akval ** 2
        return self._scale(g, am)

    # ------------------------------------------------------------------
    # the same operators as diagonals: band[d + h, i] = M[i, i + d]
    # (zero where i + d is outside 0..nn-1), h = half-bandwidth.  The
    # scaling exp(lognorm_j - lognorm_i) touches only band entries, and
    # del2 = del2h - ak^2 I because the scaling leaves the diagonal alone.
    # ------------------------------------------------------------------
    def _band_of(self, dense, h):
        nn = dense.shape[0]
        band = np.zeros((2 * h + 1, nn))
        for d in range(-h, h + 1):
            diag = np.diagonal(dense, d)
            if d >= 0:
                band[d + h, :len(diag)] = diag
            else:
                band[d + h, -d:] = diag
        return band

    def band_xxdx(self, mval, nn):
        """Diagonals (offsets -1..1) of leg_xxdx."""
        key = ('x', abs(int(mval)), nn)
        cache = self.__dict__.setdefault('_band_cache', {})
        if key not in cache:
            cache[key] = self._band_of(self.leg_xxdx(mval, nn), 1)
        return cache[key]

    def band_del2h(self, mval, nn):
        """Diagonals (offsets -2..2) of leg_del2h."""
        key = ('h', abs(int(mval)), nn)
        cache = self.__dict__.setdefault('_band_cache', {})
        if key not in cache:
            cache[key] = self._band_of(self.leg_del2h(mval, nn), 2)
        return cache[key]
