"""MI355X-native spectral-transform time step with the MLegS operator API.

Host-side mirror of the reference's Fortran surface (UCBCFD/MLegS v1.1.3):
``Kit`` is the transform kit ``tfm`` (src/modules/mlegs_spectfm.f90:55-63),
``Scalar`` is ``type(scalar)`` (src/modules/mlegs_scalar.f90:15-49), and the
module functions carry the names, argument order and ``space`` strings of
the generic interfaces in src/modules/mlegs_scalar.f90:96-450.  Every call
goes through the C ABI of ``libmlegs_amd.so`` to hand-written HIP kernels;
errors raise ``MlxError`` (the reference stops the program).
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import MlxError, PREC_FP64, PREC_FP32, PREC_BF16X3, PREC_FP64_MFMA  # noqa: F401
from ._lib import STAGE_SYNTH, STAGE_UP, STAGE_PROJ, STAGE_PROJ0  # noqa: F401
from .params import Params, read_input, timestep_set, InputError  # noqa: F401

__all__ = ['Kit', 'Scalar', 'trans', 'chop', 'dealias', 'svv_filter', 'calcat0', 'calcat1', 'zeroat1', 'fftreat',
           'delsqp', 'idelsqp', 'xxdx', 'del2h', 'del2', 'idel2', 'ihelm', 'helmp', 'ihelmp', 'fefe', 'abab', 'febe', 'abcn', 'vecprod',
           'vec2tp', 'tp2vec', 'tp2curlvec', 'advection_rhs', 'check_stability', 'VortexRun', 'energy',
           'trans_adjoint', 'vecprod_adjoint', 'abcn_adjoint', 'adjoint_project', 'febe_adjoint', 'zeroat1_adjoint',
           'tp2vec_adjoint', 'tp2curlvec_adjoint', 'vec2tp_adjoint', 'advection_adjoint', 'energy_grad',
           'Params', 'read_input', 'timestep_set', 'MlxError', 'InputError']


def _params_struct(p):
    s = _lib.MlxParams()
    for name, _ in _lib.MlxParams._fields_:
        v = getattr(p, name)
        setattr(s, name, int(v) if isinstance(v, bool) else v)
    return s


class Kit:
    """The transform kit on one GPU (tfm_kit_3d + tfm%init)."""

    def __init__(self, params, device=0, precision=PREC_FP64, _handle=None):
        self.lib = _lib.load()
        self.params = params
        # _handle: an existing context with its kit built (an ensemble
        # member's), used without taking ownership
        self._own = _handle is None
        if self._own:
            self._ctx = ctypes.c_void_p()
            ps = _params_struct(params)
            rc = self.lib.mlx_ctx_create(ctypes.byref(ps), int(device), None, ctypes.byref(self._ctx))
            if rc != 0:
                raise MlxError(rc, 'mlx_ctx_create failed')
            self._check(self.lib.mlx_kit_init(self._ctx))
        else:
            self._ctx = _handle
        d = (ctypes.c_int32 * 9)()
        self._check(self.lib.mlx_kit_dims(self._ctx, d))
        self.nrdim, self.npdim, self.nzdim, self.nrh, self.ne, self.chopzl, self.chopzu, self.nph, _ = list(d)
        self.dims = (self.nrdim, self.npdim, self.nzdim)
        self.nr, self.np, self.nz = params.nr, params.np, params.nz
        self.nrchop, self.npchop, self.nzchop = params.nrchop, params.npchop, params.nzchop
        self.ell, self.zlen = params.ell, params.zlen
        self.set_precision(precision)
        self.x = self.get('x', self.nr)
        self.w = self.get('w', self.nr)
        self.r = self.get('r', self.nr)
        self.lnx = self.get('ln', self.nr)
        self.ak = self.get('ak', self.nz)

    def _check(self, rc):
        _lib.check(self._ctx, rc)

    @property
    def ctx(self):
        return self._ctx

    def set_precision(self, prec):
        self._check(self.lib.mlx_set_precision(self._ctx, int(prec)))
        self.precision = prec

    def set_stage_precision(self, stage, prec):
        """Precision of one contraction site (STAGE_*); prec = -1 follows set_precision."""
        self._check(self.lib.mlx_set_stage_precision(self._ctx, int(stage), int(prec)))

    def get(self, which, n):
        buf = np.zeros(n, dtype=np.float64)
        self._check(self.lib.mlx_kit_get(self._ctx, which.encode(), buf.ctypes.data_as(_lib._DP), n))
        return buf

    def table(self):
        """pf(nr/2, nrchop+14, npchop) in the reference's shape."""
        n = self.nrh * self.ne * self.npchop
        return self.get('pf', n).reshape((self.nrh, self.ne, self.npchop), order='F')

    def lognorm(self):
        return self.get('lognorm', self.ne * self.npchop).reshape((self.ne, self.npchop), order='F')

    def sync(self):
        self._check(self.lib.mlx_sync(self._ctx))

    def close(self):
        if self._ctx and self._own:
            self.lib.mlx_ctx_destroy(self._ctx)
        self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def time_kernel(self, s, which, iters):
        ms = ctypes.c_double()
        self._check(self.lib.mlx_time_kernel(self._ctx, s.h, which.encode(), int(iters), ctypes.byref(ms)))
        return ms.value

    PROBES = ('fft_phi', 'fft_z', 'leg_fwd', 'leg_bwd', 'v2t', 'lu_solve', 'phi_cross', 'tp_front', 'abcn_rhs')

    def probe_start(self, max_launches=4096):
        """Bracket every launch of the probed kernels with HIP events from now on."""
        self._check(self.lib.mlx_probe_start(self._ctx, int(max_launches)))

    def probe_stop(self):
        """{probe: (launches, total ms, algorithmic bytes, algorithmic flops)} since probe_start."""
        out = (ctypes.c_double * (4 * len(self.PROBES)))()
        self._check(self.lib.mlx_probe_stop(self._ctx, out))
        return {nm: (int(out[4 * i]), out[4 * i + 1], out[4 * i + 2], out[4 * i + 3])
                for i, nm in enumerate(self.PROBES)}


class Scalar:
    """type(scalar) on the device: data e(nrdim, npdim, nzdim), ln, chop offsets, space."""

    def __init__(self, tfm, space='PPP', handle=None):
        self.tfm = tfm
        self.lib = tfm.lib
        if handle is None:
            h = ctypes.c_void_p()
            tfm._check(self.lib.mlx_field_create(tfm.ctx, ctypes.byref(h)))
            self.h = h
            self._owned = True
            if space != 'PPP':
                self.set_meta(space=space)
        else:
            self.h = handle
            self._owned = False

    def __del__(self):
        try:
            if self._owned and self.h and self.tfm._ctx:
                self.lib.mlx_field_destroy(self.tfm.ctx, self.h)
        except Exception:
            pass

    # -- metadata -----------------------------------------------------------
    def meta(self):
        sp = ctypes.create_string_buffer(4)
        ln = ctypes.c_double()
        off = (ctypes.c_int32 * 3)()
        self.tfm._check(self.lib.mlx_field_get_meta(self.tfm.ctx, self.h, sp, ctypes.byref(ln), off))
        return sp.value.decode(), ln.value, tuple(off)

    @property
    def space(self):
        return self.meta()[0]

    @space.setter
    def space(self, v):
        self.set_meta(space=v)

    @property
    def ln(self):
        return self.meta()[1]

    @ln.setter
    def ln(self, v):
        self.set_meta(ln=v)

    def set_meta(self, space=None, ln=None, offsets=None):
        sp0, ln0, off0 = self.meta()
        sp = (space if space is not None else sp0).encode()
        lnv = ln if ln is not None else ln0
        off = (ctypes.c_int32 * 3)(*(offsets if offsets is not None else off0))
        self.tfm._check(self.lib.mlx_field_set_meta(self.tfm.ctx, self.h, sp, float(lnv), off))

    def chop_offset(self, iof1, iof2=0, iof3=0):
        self.tfm._check(self.lib.mlx_chop_offset(self.tfm.ctx, self.h, iof1, iof2, iof3))

    # -- data ---------------------------------------------------------------
    def upload(self, e, space=None, ln=0.0, offsets=(0, 0, 0)):
        a = np.asfortranarray(e, dtype=np.complex128)
        if a.shape != self.tfm.dims:
            raise ValueError('scalar shape %s does not match the kit %s' % (a.shape, self.tfm.dims))
        off = (ctypes.c_int32 * 3)(*offsets)
        sp = (space or self.space).encode()
        self.tfm._check(self.lib.mlx_field_upload(self.tfm.ctx, self.h, a.ctypes.data_as(_lib._DP),
                                                  *self.tfm.dims, sp, float(ln), off))
        return self

    def download(self):
        out = np.zeros(self.tfm.dims, dtype=np.complex128, order='F')
        sp = ctypes.create_string_buffer(4)
        ln = ctypes.c_double()
        off = (ctypes.c_int32 * 3)()
        self.tfm._check(self.lib.mlx_field_download(self.tfm.ctx, self.h, out.ctypes.data_as(_lib._DP),
                                                    *self.tfm.dims, sp, ctypes.byref(ln), off))
        return out

    @property
    def e(self):
        return self.download()

    def copy(self):
        o = Scalar(self.tfm)
        self.tfm._check(self.lib.mlx_field_copy(self.tfm.ctx, o.h, self.h))
        return o

    def assign(self, other):
        self.tfm._check(self.lib.mlx_field_copy(self.tfm.ctx, self.h, other.h))

    def axpby(self, a, other, b):
        """self = a*self + b*other (data and ln)."""
        self.tfm._check(self.lib.mlx_field_axpby(self.tfm.ctx, self.h, float(a), other.h, float(b)))


def _c(tfm, rc):
    tfm._check(rc)


def trans(s, space, tfm):
    _c(tfm, tfm.lib.mlx_trans(tfm.ctx, s.h, space.encode()))


def chop(s, tfm):
    _c(tfm, tfm.lib.mlx_chop(tfm.ctx, s.h))


def dealias(s, tfm):
    _c(tfm, tfm.lib.mlx_dealias(tfm.ctx, s.h))


def svv_filter(s, tfm, gain):
    g = ctypes.c_double(gain)
    _c(tfm, tfm.lib.mlx_svv_filter(tfm.ctx, s.h, ctypes.byref(g)))
    return g.value


def _calcat(fn, s, tfm):
    out = np.zeros(2 * tfm.nzdim)
    _c(tfm, fn(tfm.ctx, s.h, out.ctypes.data_as(_lib._DP)))
    return out[0::2] + 1j * out[1::2]


def calcat0(s, tfm):
    return _calcat(tfm.lib.mlx_calcat0, s, tfm)


def calcat1(s, tfm):
    return _calcat(tfm.lib.mlx_calcat1, s, tfm)


def zeroat1(s, tfm):
    _c(tfm, tfm.lib.mlx_zeroat1(tfm.ctx, s.h))


def delsqp(s, tfm):
    _c(tfm, tfm.lib.mlx_delsqp(tfm.ctx, s.h))


def idelsqp(s, tfm):
    _c(tfm, tfm.lib.mlx_idelsqp(tfm.ctx, s.h))


def xxdx(s, tfm):
    _c(tfm, tfm.lib.mlx_xxdx(tfm.ctx, s.h))


def del2h(s, tfm):
    _c(tfm, tfm.lib.mlx_del2h(tfm.ctx, s.h))


def del2(s, tfm):
    _c(tfm, tfm.lib.mlx_del2(tfm.ctx, s.h))


def idel2(s, tfm, preln=None):
    if preln is None:
        _c(tfm, tfm.lib.mlx_idel2(tfm.ctx, s.h))
    else:
        _c(tfm, tfm.lib.mlx_idel2_preln(tfm.ctx, s.h, float(preln)))


def fftreat(s, tfm):
    _c(tfm, tfm.lib.mlx_fftreat(tfm.ctx, s.h))


def ihelm(s, alpha, tfm):
    _c(tfm, tfm.lib.mlx_ihelm(tfm.ctx, s.h, float(alpha)))


def helmp(s, power, alpha, beta, tfm):
    """del2^(power/2) + beta del2 + alpha (mlegs_scalar_ops.f90:856-898)."""
    _c(tfm, tfm.lib.mlx_helmp(tfm.ctx, s.h, int(power), float(alpha), float(beta)))


def ihelmp(s, power, alpha, beta, tfm):
    """Inverse of helmp by banded solves (mlegs_scalar_ops.f90:900-1000)."""
    _c(tfm, tfm.lib.mlx_ihelmp(tfm.ctx, s.h, int(power), float(alpha), float(beta)))


def fefe(s, nl, dt, tfm):
    _c(tfm, tfm.lib.mlx_fefe(tfm.ctx, s.h, nl.h, float(dt)))


def abab(s, s_p, nl, nl_p, dt, tfm, is_2nd_svis_p=False):
    _c(tfm, tfm.lib.mlx_abab(tfm.ctx, s.h, s_p.h, nl.h, nl_p.h, float(dt), int(bool(is_2nd_svis_p))))


def febe(s, nl, dt, tfm):
    _c(tfm, tfm.lib.mlx_febe(tfm.ctx, s.h, nl.h, float(dt)))


def abcn(s, s_p, nl, nl_p, dt, tfm):
    _c(tfm, tfm.lib.mlx_abcn(tfm.ctx, s.h, s_p.h, nl.h, nl_p.h, float(dt)))


def vecprod(vr, vp, vz, ur, up, uz, tfm):
    _c(tfm, tfm.lib.mlx_vecprod(tfm.ctx, vr.h, vp.h, vz.h, ur.h, up.h, uz.h))


def trans_adjoint(g, from_space, tfm):
    """g <- T^T g for T = trans(., g.space) applied to a field in from_space
    ('FFF' with g in PPP or 'PPP' with g in FFF); see mlx_trans_adjoint."""
    _c(tfm, tfm.lib.mlx_trans_adjoint(tfm.ctx, g.h, from_space.encode()))


def vecprod_adjoint(ar, ap, az, vr, vp, vz, ur, up, uz, ubr, ubp, ubz, tfm):
    """At (v, u): a <- u x a (the cotangent of v), ub <- a x v (that of u)."""
    _c(tfm, tfm.lib.mlx_vecprod_adjoint(tfm.ctx, ar.h, ap.h, az.h, vr.h, vp.h, vz.h, ur.h, up.h, uz.h, ubr.h, ubp.h,
                                        ubz.h))


def abcn_adjoint(sbar, nlbar, nlpbar, dt, tfm):
    """The transpose of abcn's (s, nl, nl_p) -> s: sbar in/out, nlbar, nlpbar out."""
    _c(tfm, tfm.lib.mlx_abcn_adjoint(tfm.ctx, sbar.h, nlbar.h, nlpbar.h, float(dt)))


def adjoint_project(f, tfm):
    _c(tfm, tfm.lib.mlx_adjoint_project(tfm.ctx, f.h))


def febe_adjoint(sbar, nlbar, dt, tfm):
    """The transpose of febe's (s, nl) -> s: sbar in/out, nlbar out."""
    _c(tfm, tfm.lib.mlx_febe_adjoint(tfm.ctx, sbar.h, nlbar.h, float(dt)))


def zeroat1_adjoint(f, tfm):
    _c(tfm, tfm.lib.mlx_zeroat1_adjoint(tfm.ctx, f.h))


def tp2vec_adjoint(vrb, vpb, vzb, psib, chib, tfm):
    """Cotangents of (vr, vp, vz) in PPP (overwritten) -> those of (psi, chi)."""
    _c(tfm, tfm.lib.mlx_tp2vec_adjoint(tfm.ctx, vrb.h, vpb.h, vzb.h, psib.h, chib.h))


def tp2curlvec_adjoint(wrb, wpb, wzb, psib, chib, tfm):
    _c(tfm, tfm.lib.mlx_tp2curlvec_adjoint(tfm.ctx, wrb.h, wpb.h, wzb.h, psib.h, chib.h))


def vec2tp_adjoint(psib, chib, vrb, vpb, vzb, tfm):
    """Cotangents of (psi, chi) in FFF (overwritten) -> those of (vr, vp, vz) in PPP."""
    _c(tfm, tfm.lib.mlx_vec2tp_adjoint(tfm.ctx, psib.h, chib.h, vrb.h, vpb.h, vzb.h))


def advection_adjoint(psi, chi, uz, nlpsib, nlchib, psib, chib, tfm):
    """advection_rhs's vector-Jacobian product at (psi, chi): the cotangents
    of (nlpsi, nlchi) (overwritten) -> those of (psi, chi); psi and chi are
    de-aliased in place, as advection_rhs does."""
    _c(tfm, tfm.lib.mlx_advection_adjoint(tfm.ctx, psi.h, chi.h, uz.h, nlpsib.h, nlchib.h, psib.h, chib.h))


def energy_grad(vr, vp, vz, tfm):
    """In place: a PPP velocity -> the gradient of its kinetic energy."""
    _c(tfm, tfm.lib.mlx_energy_grad(tfm.ctx, vr.h, vp.h, vz.h))


def vec2tp(vr, vp, vz, psi, chi, tfm):
    _c(tfm, tfm.lib.mlx_vec2tp(tfm.ctx, vr.h, vp.h, vz.h, psi.h, chi.h))


def tp2vec(psi, chi, vr, vp, vz, tfm):
    _c(tfm, tfm.lib.mlx_tp2vec(tfm.ctx, psi.h, chi.h, vr.h, vp.h, vz.h))


def tp2curlvec(psi, chi, wr, wp, wz, tfm):
    _c(tfm, tfm.lib.mlx_tp2curlvec(tfm.ctx, psi.h, chi.h, wr.h, wp.h, wz.h))


def advection_rhs(psi, chi, nlpsi, nlchi, uz, tfm):
    _c(tfm, tfm.lib.mlx_advection_rhs(tfm.ctx, psi.h, chi.h, nlpsi.h, nlchi.h, uz.h))


def check_stability(psi, chi, tfm):
    _c(tfm, tfm.lib.mlx_check_stability(tfm.ctx, psi.h, chi.h))


def energy(vr, vp, vz, tfm):
    out = ctypes.c_double()
    _c(tfm, tfm.lib.mlx_energy(tfm.ctx, vr.h, vp.h, vz.h, ctypes.byref(out)))
    return out.value


def vorticity_magnitude(psi, chi, out, tfm):
    """save_vort_mag's field (vortical_flow_3d.f90:411-446): out (PPP) =
    |curl u| at each point, on the real and imaginary parts separately."""
    _c(tfm, tfm.lib.mlx_vorticity_magnitude(tfm.ctx, psi.h, chi.h, out.h))


def msave(s, fn, is_binary=False, is_global=True):
    """msave(s, fn, is_binary, is_global) (mlegs_scalar_io.f90:6-101): the
    field as a global file `fn` or the per-process file `fn`_0."""
    tfm = s.tfm
    _c(tfm, tfm.lib.mlx_field_save(tfm.ctx, s.h, os.fsencode(fn), int(bool(is_binary)), int(bool(is_global))))


def mload(fn, s, is_binary=False, is_global=True):
    """mload(fn, s, is_binary, is_global) (mlegs_scalar_io.f90:103-248): data,
    space, ln and chop offsets of s from a file msave wrote."""
    tfm = s.tfm
    _c(tfm, tfm.lib.mlx_field_load(tfm.ctx, s.h, os.fsencode(fn), int(bool(is_binary)), int(bool(is_global))))
    return s


def file_write(fn, e, space='FFF', ln=0.0, offsets=(0, 0, 0), is_binary=False, is_global=True):
    """The msave file formats for a host array e(n1, n2, n3) (no GPU)."""
    lib = _lib.load()
    a = np.asfortranarray(e, dtype=np.complex128)
    if a.ndim != 3:
        raise ValueError('file_write: a 3D array is expected')
    off = (ctypes.c_int32 * 3)(*offsets)
    rc = lib.mlx_file_write(os.fsencode(fn), a.ctypes.data_as(_lib._DP), *a.shape, space.encode(), float(ln), off,
                            int(bool(is_binary)), int(bool(is_global)))
    if rc:
        raise _lib.MlxError(rc, lib.mlx_file_last_error().decode())


def file_read(fn, shape, is_binary=False, is_global=True):
    """Read a file of the msave formats into a host array of `shape`;
    returns (e, space, ln, offsets)."""
    lib = _lib.load()
    out = np.zeros(shape, dtype=np.complex128, order='F')
    sp = ctypes.create_string_buffer(4)
    ln = ctypes.c_double()
    off = (ctypes.c_int32 * 3)()
    rc = lib.mlx_file_read(os.fsencode(fn), out.ctypes.data_as(_lib._DP), *shape, sp, ctypes.byref(ln), off,
                           int(bool(is_binary)), int(bool(is_global)))
    if rc:
        raise _lib.MlxError(rc, lib.mlx_file_last_error().decode())
    return out, sp.value.decode(), ln.value, tuple(off)


class VortexRun:
    """The 3D vortex driver's state on the device (vortical_flow_3d.f90:26-206)."""

    FIELDS = ('psi', 'chi', 'nlpsi', 'nlchi', 'psi_prev', 'chi_prev', 'nlpsi_prev', 'nlchi_prev', 'uz')

    def __init__(self, tfm, psi, chi, uz):
        self.tfm = tfm
        self.h = ctypes.c_void_p()
        _c(tfm, tfm.lib.mlx_vortex_create(tfm.ctx, psi.h, chi.h, uz.h, ctypes.byref(self.h)))

    def __del__(self):
        try:
            if self.h and self.tfm._ctx:
                self.tfm.lib.mlx_vortex_destroy(self.tfm.ctx, self.h)
        except Exception:
            pass

    def bootstrap(self):
        _c(self.tfm, self.tfm.lib.mlx_vortex_bootstrap(self.tfm.ctx, self.h))

    def step(self, nsteps=1, dt=None, check=True):
        dt = self.tfm.params.dt if dt is None else dt
        _c(self.tfm, self.tfm.lib.mlx_vortex_step(self.tfm.ctx, self.h, int(nsteps), float(dt), int(bool(check))))

    def field(self, name):
        i = self.FIELDS.index(name)
        h = ctypes.c_void_p()
        _c(self.tfm, self.tfm.lib.mlx_vortex_field(self.tfm.ctx, self.h, i, ctypes.byref(h)))
        return Scalar(self.tfm, handle=h)

    def gains(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        _c(self.tfm, self.tfm.lib.mlx_vortex_gains(self.tfm.ctx, self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value
