"""Ensembles of independent 2D vortex runs on one GPU (mlx_ensemble_* of the
C ABI), and the Lamb-Oseen member recipe.

The reference runs one transform kit per process
(src/submodules/mlegs_spectfm_init.f90:15-16).  ``Ensemble`` holds n member
contexts, each with its own kit (every parameter free per member), and steps
them together by replaying each member's captured AB2/CN step as a HIP
graph; a member's results equal those of its own single ``VortexRun``.

The member recipe is the reference's closed-form tutorial case, the
Lamb-Oseen vortex w0 = exp(-r^2/2) / (2 pi) of unit circulation
(docs/tutorial/time_integration.md:40-54), in toroidal-poloidal form:
chi = 0, psi with del_perp^2 psi = -w0 (psi = idelsqp(-w0 / (1 - x)^2), the
pattern of the reference's q-vortex set-up, src/apps/vortical_flow_3d.f90:
258-326), uz = 0, plus a seeded perturbation on psi and chi.
"""
import ctypes
import math

import numpy as np

from . import _lib, inputs
from . import Kit, Scalar, MlxError, PREC_FP64, trans, idelsqp, zeroat1, dealias, _params_struct
from .params import Params

NR_CHOICES = (64, 96, 128, 192, 256)


def member_dt(nr, ell, dt_max=1.0e-2):
    """The member's time step: dt_max * min(1, 64 ell / NR).  The mapped
    collocation points crowd the axis with a spacing of order ell / NR, and
    the explicit advection term's limit scales with it: at dt = 1e-2 a member
    with NR = 128, ell = 0.83, visc = 1.3e-4 loses stability after 600 steps,
    while 5e-3 (this rule gives 4.1e-3) runs its 2000 steps and beyond."""
    return dt_max * min(1.0, 64.0 * ell / nr)


def member_params(member_id, seed=2026, dt=None):
    """Parameters of ensemble member `member_id`: NR uniform over
    {64, 96, 128, 192, 256}, NP = NR, NZ = 1, full chops, ell ~ U[0.5, 2],
    visc ~ logU[1e-4, 1e-2] (drawn from a generator keyed by (seed,
    member_id)), dt = member_dt(NR, ell) unless given, SVV on and always
    active: cutoff 0.5, target tail ratio 1e-12, the reference validator's
    SVV-on setting (tools/validate_tutorials.py:308-327).  With the default
    controls (0.75, 2e-2) the filter never engages here, as the vortex holds
    nearly all the energy, and members with visc near 1e-4 and ell / NR small
    lose stability once the sheared perturbation outruns the grid: at
    t = 1.6 for NR = 96, ell = 0.63, visc = 1.1e-4, whatever the time step;
    with cutoff 0.75 and target 1e-12 that member still fails at t = 2.5."""
    rng = np.random.default_rng([seed, member_id])
    nr = int(NR_CHOICES[rng.integers(len(NR_CHOICES))])
    ell = float(rng.uniform(0.5, 2.0))
    visc = float(10.0 ** rng.uniform(-4.0, -2.0))
    if dt is None:
        dt = member_dt(nr, ell)
    return Params(nr=nr, np=nr, nz=1, nrchop=nr, npchop=nr // 2 + 1, nzchop=1, ell=ell, zlen=1.0, dt=dt,
                  visc=visc, hyperpow=0, hypervisc=0.0, is_svv=True, svv_cutoff=0.5,
                  svv_target=1e-12)


def lamb_oseen_tp(tfm, member_id, amp=1.0e-3, seed=inputs.DEFAULT_SEED):
    """(psi, chi, uz) of the perturbed Lamb-Oseen member on kit `tfm`:
    psi = idelsqp(-w0 / (1 - x)^2) + pert, chi = pert', uz = 0; de-aliased,
    zeroat1 applied."""
    p = tfm.params
    dims = tfm.dims
    x, r = tfm.x, tfm.r
    f = -np.exp(-r[:p.nr] ** 2 / 2.0) / (2.0 * math.pi) / (1.0 - x[:p.nr]) ** 2
    src = np.zeros(dims, dtype=np.complex128)
    src[:p.nr, :p.np // 2, :] = (f + 1j * f)[:, None, None]
    chops = (p.nrchop, p.npchop, p.nzchop, p.np, p.nz)
    psi = Scalar(tfm).upload(src, space='PPP')
    trans(psi, 'FFF', tfm)
    idelsqp(psi, tfm)
    zeroat1(psi, tfm)
    chi = Scalar(tfm).upload(np.zeros(dims, dtype=np.complex128), space='FFF')
    for s, fid in ((psi, 2 * member_id + 1), (chi, 2 * member_id + 2)):
        pert = Scalar(tfm).upload(inputs.perturbation(dims, chops, amp, seed=seed, field_id=fid), space='FFF')
        s.axpby(1.0, pert, 1.0)
        dealias(s, tfm)
        zeroat1(s, tfm)
    uz = Scalar(tfm).upload(np.zeros(dims, dtype=np.complex128), space='PPP')
    return psi, chi, uz


class Ensemble:
    """n independent vortex runs on one GPU, stepped by graph replay."""

    def __init__(self, params_list, device=0, precision=PREC_FP64):
        self.lib = _lib.load()
        self.params = list(params_list)
        n = len(self.params)
        arr = (_lib.MlxParams * n)(*[_params_struct(p) for p in self.params])
        self.h = ctypes.c_void_p()
        rc = self.lib.mlx_ensemble_create(ctypes.cast(arr, ctypes.c_void_p), n, int(device), int(precision),
                                          ctypes.byref(self.h))
        if rc != 0:
            raise MlxError(rc, 'mlx_ensemble_create failed')
        self.kits = []
        for i, p in enumerate(self.params):
            c = ctypes.c_void_p()
            self._check(self.lib.mlx_ensemble_ctx(self.h, i, ctypes.byref(c)))
            self.kits.append(Kit(p, device=device, precision=precision, _handle=c))
        self._runs = {}

    def _check(self, rc):
        if rc != 0:
            raise MlxError(rc, self.lib.mlx_ensemble_last_error(self.h).decode())

    def __len__(self):
        return len(self.params)

    def set(self, i, psi, chi, uz):
        self._check(self.lib.mlx_ensemble_set(self.h, int(i), psi.h, chi.h, uz.h))

    def bootstrap(self):
        self._check(self.lib.mlx_ensemble_bootstrap(self.h))

    def step(self, nsteps, mode=-1, streams=8, reps=1):
        """nsteps AB2/CN steps of every member; mode -1 tunes the launch
        configuration once, 0 replays one parent graph, 1 the member graphs
        on `streams` streams, 2 the same with one host thread per stream;
        reps = steps per graph (1 or 4)."""
        self._check(self.lib.mlx_ensemble_step(self.h, int(nsteps), int(mode), int(streams), int(reps)))

    def sync(self):
        self._check(self.lib.mlx_ensemble_sync(self.h))

    def field(self, i, name):
        """Member i's field `name` (VortexRun.FIELDS) as a Scalar of its kit."""
        from . import VortexRun
        v = ctypes.c_void_p()
        self._check(self.lib.mlx_ensemble_vortex(self.h, int(i), ctypes.byref(v)))
        tfm = self.kits[i]
        h = ctypes.c_void_p()
        _lib.check(tfm.ctx, self.lib.mlx_vortex_field(tfm.ctx, v, VortexRun.FIELDS.index(name), ctypes.byref(h)))
        return Scalar(tfm, handle=h)

    def check(self):
        """Index of the first member whose state became non-finite, or -1."""
        bad = ctypes.c_int32()
        self._check(self.lib.mlx_ensemble_check(self.h, ctypes.byref(bad)))
        return bad.value

    def info(self):
        v = (ctypes.c_int32 * 5)()
        t = (ctypes.c_double * 36)()
        self._check(self.lib.mlx_ensemble_info(self.h, v, t))
        tune = np.array(list(t)).reshape(3, 6, 2)
        return dict(mode=v[0], streams=v[1], steps_per_graph=v[2], phases=v[3], members=v[4], tune_ms=tune)

    def close(self):
        if self.h:
            for k in self.kits:
                k._ctx = ctypes.c_void_p()
            self.lib.mlx_ensemble_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
