"""Shared set-up for the GPU parity tests: the same seeded inputs go to the
HIP path (mlegs_amd) and to the oracle."""
import math

import numpy as np

from mlegs_amd import inputs
from mlegs_amd.params import Params

# the reference's CPU-runnable 3D vortex size (NR=64, NP=48, NZ=32)
C1 = dict(nr=64, np=48, nz=32, nrchop=64, npchop=25, nzchop=17, ell=1.0, zlen=2 * math.pi,
          visc=1e-3, dt=1e-3, hyperpow=0, hypervisc=0.0, is_svv=False)
# the validator's 3D size (tools/validate_tutorials.py), chops below the maximum
SMALL = dict(nr=32, np=16, nz=8, nrchop=30, npchop=8, nzchop=4, ell=2.0, zlen=3.0,
             visc=5e-3, dt=2e-3, hyperpow=0, hypervisc=0.0, is_svv=True)
# radix-5/3 lengths (np/2 = 20 = 4*5, nz = 30 = 2*3*5)
ODD = dict(nr=40, np=40, nz=30, nrchop=36, npchop=21, nzchop=16, ell=1.5, zlen=4.0,
           visc=1e-2, dt=1e-3, hyperpow=0, hypervisc=0.0, is_svv=True)


# power-of-two phi and z lengths (np/2 = 32, nz = 16): the four-step FFT kernels;
# nzchop = nz/2 + 1 repeats the Nyquist slice in the kk-loops
POW2 = dict(nr=32, np=64, nz=16, nrchop=30, npchop=20, nzchop=9, ell=1.2, zlen=2.5,
            visc=2e-3, dt=1e-3, hyperpow=0, hypervisc=0.0, is_svv=True)


def params(cfg, **kw):
    d = dict(cfg)
    d.update(kw)
    return Params(**d)


def chops(p):
    return (p.nrchop, p.npchop, p.nzchop, p.np, p.nz)


def spectral(kit_dims, p, seed_id, decay=8.0, admissible=True):
    """Seeded FFF coefficients; admissible ones are chopped and Hermitian."""
    c = inputs.transform_field(kit_dims, chops(p), field_id=seed_id, decay=decay)
    if not admissible:
        rng = np.random.default_rng(seed_id)
        c = c + 1e-3 * (rng.standard_normal(kit_dims) + 1j * rng.standard_normal(kit_dims))
    return c


def physical(kit_dims, p, seed_id):
    """A seeded PPP array: rows < nr, packed columns < np/2 (padding zero)."""
    rng = np.random.default_rng(1000 + seed_id)
    e = np.zeros(kit_dims, dtype=np.complex128)
    e[:p.nr, :p.np // 2, :] = rng.standard_normal((p.nr, p.np // 2, kit_dims[2])) + \
        1j * rng.standard_normal((p.nr, p.np // 2, kit_dims[2]))
    return e


def pair(tfm, okit, e, space, ln=0.0, offsets=(0, 0, 0)):
    """(device Scalar, oracle Scalar) holding the same data."""
    import mlegs_amd as M
    from oracle import scalar_ops as so
    g = M.Scalar(tfm).upload(e, space=space, ln=ln, offsets=offsets)
    o = so.Scalar(okit, space)
    o.e = np.array(e, dtype=np.complex128)
    o.ln = ln
    o.nrchop_offset = offsets[0]
    return g, o


def qvortex(tfm, okit, p, q=4.0, amp=1e-5):
    """Single centred q-vortex psi, chi (FFF) plus the seeded perturbation,
    built identically on the device and in the oracle."""
    import mlegs_amd as M
    from oracle import scalar_ops as so
    dims = tfm.dims
    outs = []
    for which, fid in (('psi', 1), ('chi', 2)):
        ppp_g = inputs.qvortex_ppp(tfm.x, tfm.r, dims, p.np, p.nz, p.nr, q, which)
        ppp_o = inputs.qvortex_ppp(okit.x, okit.r, dims, p.np, p.nz, p.nr, q, which)
        pert = inputs.perturbation(dims, chops(p), amp, field_id=fid)
        g = M.Scalar(tfm).upload(ppp_g, space='PPP')
        M.trans(g, 'FFF', tfm)
        M.idelsqp(g, tfm)
        M.zeroat1(g, tfm)
        gp = M.Scalar(tfm).upload(pert, space='FFF')
        g.axpby(1.0, gp, 1.0)
        M.dealias(g, tfm)
        M.zeroat1(g, tfm)
        o = so.Scalar(okit, 'PPP')
        o.e = ppp_o
        so.trans(o, 'FFF', okit)
        so.idelsqp(o, okit)
        so.zeroat1(o, okit)
        o.e = o.e + pert
        so.dealias(o, okit)
        so.zeroat1(o, okit)
        outs.append((g, o))
    return outs
