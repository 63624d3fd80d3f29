"""Writes tests/golden/kit_golden.npz: high-precision values of the transform
kit's tables, computed with mpmath independently of the oracle.

  gl{32,256}_x, _w     Gauss-Legendre nodes/weights (50-digit Newton on P_n)
  tbl32                P_n^m(x_i) fully normalised, Condon-Shortley phase, at the
                       16 negative nodes of NR=32 (each node rounded to the nearest
                       double first, as a table built from FP64 nodes sees it),
                       n = m..m+45, m = 0..8
                       (the NR=32, NRCHOP=32, NPCHOP=9 kit: ne = NRCHOP+14)
  spot256_m, _n, spot256  the same at the 128 negative nodes of NR=256 for
                       m in {0, 1, 64, 128} and a few degrees per m
  lognorm32            0.5*log((2n+1)/2 (n-m)!/(n+m)!) for the NR=32 kit
  lognorm256_m, lognorm256  the same for m in spot256_m, n = m..m+269
  at0_256, at1_256     P_n^0(-/+(1 - 1e-15)), n = 0..255 (the double nearest 1-1e-15)

The associated Legendre values come from the explicit sum
  P_n^m(x) = (-1)^m (1-x^2)^(m/2) 2^-n sum_k (-1)^k C(n,k) C(2n-2k,n) (n-2k)!/(n-2k-m)! x^(n-2k-m)
with exact integer coefficients, evaluated at 600 digits (the sum cancels
heavily for |x| < 1); for small degrees they are cross-checked against
mpmath.legenp.  Run:  python tests/golden/make_golden.py
"""
import math
import os

import mpmath as mp
import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'kit_golden.npz')


def gauss_nodes(n):
    """All n nodes (ascending) and weights, 50 digits."""
    with mp.workdps(60):
        xs, ws = [], []
        for i in range(1, n // 2 + 1):
            z = mp.cos(mp.pi * (i - mp.mpf(1) / 4) / (n + mp.mpf(1) / 2))
            for _ in range(200):
                p1, p2 = mp.mpf(1), mp.mpf(0)
                for j in range(1, n + 1):
                    p3 = p2
                    p2 = p1
                    p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j
                pp = n * (z * p1 - p2) / (z * z - 1)
                dz = p1 / pp
                z = z - dz
                if abs(dz) < mp.mpf(10) ** -55:
                    break
            p1, p2 = mp.mpf(1), mp.mpf(0)
            for j in range(1, n + 1):
                p3 = p2
                p2 = p1
                p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j
            pp = n * (z * p1 - p2) / (z * z - 1)
            xs.append(z)
            ws.append(2 / ((1 - z * z) * pp * pp))
        x = [-v for v in xs] + [v for v in reversed(xs)]
        w = list(ws) + list(reversed(ws))
        return x, w


_coef_cache = {}


def _coefs(n, m):
    key = (n, m)
    if key not in _coef_cache:
        cs = []
        for k in range(0, (n - m) // 2 + 1):
            p = n - 2 * k
            if p - m < 0:
                break
            c = (-1) ** k * math.comb(n, k) * math.comb(2 * n - 2 * k, n) * math.factorial(p) // math.factorial(p - m)
            cs.append((p - m, c))
        _coef_cache[key] = cs
    return _coef_cache[key]


def pbar(n, m, x):
    """Fully normalised P_n^m(x), Condon-Shortley phase, as an mpf."""
    with mp.workdps(600):
        x = mp.mpf(x)
        s = mp.mpf(0)
        for p, c in _coefs(n, m):
            s += c * x ** p
        val = (-1) ** m * (1 - x * x) ** (mp.mpf(m) / 2) * s / mp.mpf(2) ** n
        norm = mp.sqrt(mp.mpf(2 * n + 1) / 2 * mp.factorial(n - m) / mp.factorial(n + m))
        return +(val * norm)


def lognorm(n, m):
    with mp.workdps(50):
        return 0.5 * mp.log(mp.mpf(2 * n + 1) / 2 * mp.factorial(n - m) / mp.factorial(n + m))


def main():
    out = {}
    for nr in (32, 256):
        x, w = gauss_nodes(nr)
        out['gl%d_x' % nr] = np.array([float(v) for v in x])
        out['gl%d_w' % nr] = np.array([float(v) for v in w])

    # cross-check the explicit sum against mpmath's hypergeometric legenp
    with mp.workdps(40):
        for n, m, xv in ((2, 1, 0.3), (7, 3, -0.71), (20, 0, 0.2), (33, 8, -0.95)):
            a = pbar(n, m, xv)
            b = mp.legenp(n, m, mp.mpf(xv)) * mp.sqrt(mp.mpf(2 * n + 1) / 2 * mp.factorial(n - m) / mp.factorial(n + m))
            assert abs(a - b) < mp.mpf(10) ** -30 * max(1, abs(b)), (n, m, a, b)

    # the NR=32 kit, full table at the 16 negative nodes
    x32 = [float(v) for v in gauss_nodes(32)[0][:16]]
    ne, ms = 32 + 14, list(range(9))
    tbl = np.zeros((16, ne, len(ms)))
    for mi, m in enumerate(ms):
        for nn in range(ne):
            for i, xv in enumerate(x32):
                tbl[i, nn, mi] = float(pbar(m + nn, m, xv))
    out['tbl32'] = tbl
    out['lognorm32'] = np.array([[float(lognorm(m + nn, m)) for m in ms] for nn in range(ne)])

    # NR=256 spot columns
    x256 = [float(v) for v in gauss_nodes(256)[0][:128]]
    sm = [0, 1, 64, 128]
    sn = [0, 1, 7, 100, 255 - 128, 269]   # offsets n - m
    spot = np.zeros((128, len(sm), len(sn)))
    for a, m in enumerate(sm):
        for b, off in enumerate(sn):
            for i, xv in enumerate(x256):
                spot[i, a, b] = float(pbar(m + off, m, xv))
    out['spot256_m'] = np.array(sm)
    out['spot256_n'] = np.array(sn)
    out['spot256'] = spot
    out['lognorm256_m'] = np.array(sm)
    out['lognorm256'] = np.array([[float(lognorm(m + nn, m)) for m in sm] for nn in range(256 + 14)])

    xe = mp.mpf(float(1.0 - 1e-15))
    out['at1_256'] = np.array([float(pbar(n, 0, xe)) for n in range(256)])
    out['at0_256'] = np.array([float(pbar(n, 0, -xe)) for n in range(256)])
    np.savez_compressed(OUT, **out)
    print('wrote', OUT, {k: v.shape for k, v in out.items()})


if __name__ == '__main__':
    main()
