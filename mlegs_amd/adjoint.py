"""Reverse-mode gradients with checkpoint-and-recompute, through operators
with hand-written adjoints (the transforms, the vector product, the
toroidal-poloidal reconstruction and projection, advection_rhs, abcn, febe,
zeroat1, and the self-adjoint projections dealias and chop).  The reference
has no adjoint.

``VortexChain`` differentiates the vortex run of vortical_flow_3d.f90:116-206
(the FE-BE + Richardson bootstrap, then AB2/CN steps with de-aliasing,
zeroat1 and advection_rhs) with respect to the initial psi and chi.

``ModelChain`` is a smaller nonlinear time step on
three scalars (a, b, c) that exercises every one of those operators the
way the vortex step does (backward transforms, a quadratic product of
physical fields, forward transforms, de-aliasing, AB2/CN):

    v = (B a, B b, B c)            B = trans(., 'PPP') from FFF
    u = (v_b, v_c, v_a)
    n = chop(dealias(F(v x u)))    F = trans(., 'FFF') from PPP
    s <- abcn(s, s_p, n_s, n_s_prev, dt)   for s in (a, b, c)

with the AB2 history of the first step equal to its own term.  The
objective is J = 1/2 sum_s (|s.e|^2 + s.ln^2) after nsteps steps, the
Euclidean norm of the stored arrays.  ``gradient`` stores the state every
``segment`` steps on the device, then walks the segments backwards,
recomputing each one's states before its reverse steps.
"""
import ctypes
import time

import numpy as np

from . import (Scalar, VortexRun, trans, dealias, chop, zeroat1, vecprod, abcn, febe, advection_rhs, tp2vec, energy,
               trans_adjoint, vecprod_adjoint, abcn_adjoint, adjoint_project, febe_adjoint, zeroat1_adjoint,
               advection_adjoint, tp2vec_adjoint, energy_grad, PREC_FP64, STAGE_SYNTH, STAGE_UP, STAGE_PROJ,
               STAGE_PROJ0)


def inner(x, y):
    """Euclidean inner product of two lists of scalars: sum Re(conj(x) y) + x.ln y.ln."""
    s = 0.0
    for a, b in zip(x, y):
        s += float(np.sum(a.e.real * b.e.real + a.e.imag * b.e.imag)) + a.ln * b.ln
    return s


def copy_all(xs):
    return [x.copy() for x in xs]


class ModelChain:
    """The nonlinear model step above on kit `tfm` (hyperpow = 0)."""

    def __init__(self, tfm, dt, segment=5):
        self.tfm = tfm
        self.dt = float(dt)
        self.segment = int(segment)

    # -- forward ----------------------------------------------------------
    def _physical(self, state):
        v = []
        for s in state:
            x = s.copy()
            trans(x, 'PPP', self.tfm)
            v.append(x)
        return v

    def _term(self, v):
        """n = chop(dealias(F(v x u))), u = (v_b, v_c, v_a)."""
        t = self.tfm
        w = copy_all(v)
        u = [v[1], v[2], v[0]]
        vecprod(w[0], w[1], w[2], u[0], u[1], u[2], t)
        for x in w:
            trans(x, 'FFF', t)
            dealias(x, t)
            chop(x, t)
        return w

    def step(self, state, hist):
        """One step: returns (new state, new history); hist None on the first."""
        t = self.tfm
        n = self._term(self._physical(state))
        if hist is None:
            hist = copy_all(n)
        new = []
        for s, ns, hp in zip(state, n, hist):
            x, xp, h = s.copy(), s.copy(), hp.copy()
            abcn(x, xp, ns, h, self.dt, t)
            new.append(x)
        return new, n

    def forward(self, x0, nsteps):
        state, hist = copy_all(x0), None
        for _ in range(nsteps):
            state, hist = self.step(state, hist)
        return state

    @staticmethod
    def objective(state):
        return 0.5 * inner(state, state)

    # -- reverse ----------------------------------------------------------
    def _reverse_step(self, state_k, sbar, carry, first):
        """Cotangents of step k's inputs from those of its outputs.
        sbar: of the new state (overwritten with those of state_k);
        carry: of the history this step writes (n_k), or None.
        Returns the cotangent of the history it read (n_{k-1}), or None on
        the first step (whose history is its own term)."""
        t = self.tfm
        nbar, npbar = [], []
        for sb in sbar:
            a, b = Scalar(t, 'FFF'), Scalar(t, 'FFF')
            abcn_adjoint(sb, a, b, self.dt, t)
            nbar.append(a)
            npbar.append(b)
        if carry is not None:
            for a, c in zip(nbar, carry):
                a.axpby(1.0, c, 1.0)
        if first:
            for a, b in zip(nbar, npbar):
                a.axpby(1.0, b, 1.0)
        # through chop, dealias (projections) and F
        for a in nbar:
            chop(a, t)
            dealias(a, t)
            trans_adjoint(a, 'PPP', t)
        # through the product at (v, u)
        v = self._physical(state_k)
        u = [v[1], v[2], v[0]]
        ub = [Scalar(t, 'PPP') for _ in range(3)]
        vecprod_adjoint(nbar[0], nbar[1], nbar[2], v[0], v[1], v[2], u[0], u[1], u[2], ub[0], ub[1], ub[2], t)
        # u = (v_b, v_c, v_a)
        nbar[1].axpby(1.0, ub[0], 1.0)
        nbar[2].axpby(1.0, ub[1], 1.0)
        nbar[0].axpby(1.0, ub[2], 1.0)
        # through B
        for sb, vb in zip(sbar, nbar):
            trans_adjoint(vb, 'FFF', t)
            sb.axpby(1.0, vb, 1.0)
            adjoint_project(sb, t)
        return None if first else npbar

    def gradient(self, x0, nsteps, timing=None):
        """(J, grad): the objective after nsteps steps from x0 and its
        gradient with respect to x0 (cotangents of the three fields' data and
        log coefficients).  timing: a dict that receives the forward and
        reverse seconds."""
        seg = max(1, self.segment)
        t0 = time.perf_counter()
        # forward sweep: checkpoints at steps 0, seg, 2 seg, ...
        ckpt = {}
        state, hist = copy_all(x0), None
        for k in range(nsteps):
            if k % seg == 0:
                ckpt[k] = (copy_all(state), copy_all(hist) if hist is not None else None)
            state, hist = self.step(state, hist)
        self.tfm.sync()
        t1 = time.perf_counter()
        J = self.objective(state)
        t1b = time.perf_counter()
        sbar = copy_all(state)  # dJ/ds = s
        carry = None
        for k0 in sorted(ckpt, reverse=True):
            k1 = min(k0 + seg, nsteps)
            st, hi = ckpt[k0]
            states = [copy_all(st)]
            for _ in range(k0, k1 - 1):
                nxt, hi = self.step(states[-1], hi)
                states.append(nxt)
            for k in range(k1 - 1, k0 - 1, -1):
                carry = self._reverse_step(states[k - k0], sbar, carry, first=(k == 0))
        self.tfm.sync()
        t2 = time.perf_counter()
        if timing is not None:
            timing['forward_s'] = t1 - t0
            timing['reverse_s'] = t2 - t1b
        return J, sbar


def _fff(tfm):
    return Scalar(tfm, 'FFF')


def _add(a, b):
    """a += b (data and ln); returns a."""
    a.axpby(1.0, b, 1.0)
    return a


class VortexChain:
    """The vortex run on kit `tfm` with background axial flow `uz` (PPP),
    SVV off and hyperpow = 0, and its reverse-mode gradient.

    Forward: the bootstrap of vortical_flow_3d.f90:116-147 (advection_rhs
    at the initial state q, which de-aliases q in place; r = febe(q, N(q),
    dt); h1 = febe(q, N(q), dt/2); h2 = febe(h1, N(h1), dt/2); s0 = zeroat1(
    dealias(2 h2 - r)) with h2's log coefficient; N0 = N(s0), and N(q) as the
    AB2 history), then the AB2/CN steps of :160-181 (mlx_vortex_step):
    s' = zeroat1(dealias(abcn(s, N, N_prev))), N' = N(s').  The steps run
    at the kit's precision settings (`configure` re-applies them after an
    FP64 objective); the adjoints run in FP64.

    gradient(): checkpoints (psi, chi, N, N_prev) every `segment` steps on
    the device, then walks the segments backwards, recomputing each one's
    states before its reverse steps."""

    def __init__(self, tfm, uz, segment=20, configure=None):
        p = tfm.params
        if p.is_svv:
            raise ValueError('VortexChain: the SVV gain depends on the state; run with SVV off')
        if p.hyperpow != 0:
            raise ValueError('VortexChain: the hyperviscous stepper has no adjoint')
        self.tfm, self.uz = tfm, uz
        self.segment = max(1, int(segment))
        self.dt = float(p.dt)
        self.configure = configure

    # -- precision ----------------------------------------------------------
    def _fp64(self):
        t = self.tfm
        t.set_precision(PREC_FP64)
        for st in (STAGE_SYNTH, STAGE_UP, STAGE_PROJ, STAGE_PROJ0):
            t.set_stage_precision(st, -1)

    def _restore(self):
        if self.configure is not None:
            self.configure(self.tfm)

    # -- forward ------------------------------------------------------------
    def bootstrap(self, psi0, chi0):
        """The bootstrap from (psi0, chi0); returns its states."""
        t, dt, uz = self.tfm, self.dt, self.uz
        p, c = psi0.copy(), chi0.copy()
        n1, n2 = _fff(t), _fff(t)
        advection_rhs(p, c, n1, n2, uz, t)
        q = (p.copy(), c.copy())
        nq = (n1.copy(), n2.copy())
        rp, rc = p.copy(), c.copy()
        febe(rp, n1, dt, t)
        febe(rc, n2, dt, t)
        febe(p, n1, dt / 2.0, t)
        febe(c, n2, dt / 2.0, t)
        h1 = (p.copy(), c.copy())
        advection_rhs(p, c, n1, n2, uz, t)
        febe(p, n1, dt / 2.0, t)
        febe(c, n2, dt / 2.0, t)
        lp, lc = p.ln, c.ln
        p.axpby(2.0, rp, -1.0)
        c.axpby(2.0, rc, -1.0)
        p.ln, c.ln = lp, lc
        for s in (p, c):
            dealias(s, t)
            zeroat1(s, t)
        advection_rhs(p, c, n1, n2, uz, t)
        return {'q': q, 'h1': h1, 's0': (p.copy(), c.copy()), 'X0': (p, c, n1, n2, nq[0].copy(), nq[1].copy())}

    def _run(self, X):
        run = VortexRun(self.tfm, X[0], X[1], self.uz)
        self._load(run, X)
        return run

    @staticmethod
    def _load(run, X):
        for name, f in zip(('psi', 'chi', 'nlpsi', 'nlchi', 'nlpsi_prev', 'nlchi_prev'), X):
            run.field(name).assign(f)

    @staticmethod
    def _save(run):
        return tuple(run.field(n).copy() for n in ('psi', 'chi', 'nlpsi', 'nlchi', 'nlpsi_prev', 'nlchi_prev'))

    def forward(self, psi0, chi0, nsteps):
        """(psi, chi) after the bootstrap and nsteps steps (the production
        path: mlx_vortex_bootstrap and mlx_vortex_step)."""
        run = VortexRun(self.tfm, psi0, chi0, self.uz)
        run.bootstrap()
        if nsteps:
            run.step(nsteps, check=False)
        run.step(0, check=True)
        return run.field('psi').copy(), run.field('chi').copy()

    # -- objective ----------------------------------------------------------
    def perturbation_energy(self, s, base, grad=True):
        """J = 1/2 int |u(s) - u(base)|^2 (u = tp2vec, mlx_energy's quadrature)
        in FP64, and its gradient with respect to s (psi, chi)."""
        t = self.tfm
        self._fp64()
        try:
            dp, dc = s[0].copy(), s[1].copy()
            dp.axpby(1.0, base[0], -1.0)
            dc.axpby(1.0, base[1], -1.0)
            vr, vp, vz = Scalar(t), Scalar(t), Scalar(t)
            tp2vec(dp, dc, vr, vp, vz, t)
            J = energy(vr, vp, vz, t)
            if not grad:
                return J
            energy_grad(vr, vp, vz, t)
            gp, gc = _fff(t), _fff(t)
            tp2vec_adjoint(vr, vp, vz, gp, gc, t)
            return J, gp, gc
        finally:
            self._restore()

    # -- reverse ------------------------------------------------------------
    def _scratch(self, nseg):
        """Device fields reused by every gradient (no allocation per step)."""
        t = self.tfm
        if len(getattr(self, '_states', ())) < nseg:
            self._states = [(_fff(t), _fff(t)) for _ in range(nseg)]
            self._vw = [tuple(Scalar(t) for _ in range(6)) for _ in range(nseg)]
        if not hasattr(self, '_g'):
            self._g = [_fff(t), _fff(t)]

    def _advection_t(self, s, nb):
        """Cotangents of (psi, chi) from those of N(s) (nb overwritten), in
        buffers that the next call reuses."""
        g = self._g
        advection_adjoint(s[0], s[1], self.uz, nb[0], nb[1], g[0], g[1], self.tfm)
        return g

    def _reverse_step(self, j, sb, nb, npb):
        """From the cotangents (sb, nb, npb) of the state after a step, whose
        psi, chi (and v, w) are the j-th recorded ones, to those of the state
        before it, in place (mlx_vortex_reverse_step)."""
        t = self.tfm
        s1 = self._states[j]
        vw = (ctypes.c_void_p * 6)(*[f.h for f in self._vw[j]]) if self._vw_fmt else None
        t._check(t.lib.mlx_vortex_reverse_step(t.ctx, s1[0].h, s1[1].h, self.uz.h, sb[0].h, sb[1].h, nb[0].h,
                                                nb[1].h, npb[0].h, npb[1].h, ctypes.c_double(self.dt), vw,
                                                int(self._vw_fmt)))

    def _record(self, run, n):
        """n steps of the run; psi, chi and the step's v, w after each into
        the state buffers."""
        t = self.tfm
        arr = (ctypes.c_void_p * (2 * n))(*[f.h for st in self._states[:n] for f in st])
        avw = (ctypes.c_void_p * (6 * n))(*[f.h for st in self._vw[:n] for f in st])
        fmt = ctypes.c_int32()
        t._check(t.lib.mlx_vortex_step_record(t.ctx, run.h, int(n), ctypes.c_double(self.dt), arr, avw,
                                               ctypes.byref(fmt)))
        self._vw_fmt = fmt.value

    def _reverse_bootstrap(self, B, sb, nb, npb):
        t, dt = self.tfm, self.dt
        g = self._advection_t(B['s0'], nb)
        h2b, rb = [], []
        for i in (0, 1):
            _add(sb[i], g[i])
            zeroat1_adjoint(sb[i], t)
            dealias(sb[i], t)
            ln = sb[i].ln
            h = sb[i].copy()
            h.axpby(2.0, sb[i], 0.0)
            h.ln = ln
            r = sb[i].copy()
            r.axpby(-1.0, sb[i], 0.0)
            r.ln = 0.0
            h2b.append(h)
            rb.append(r)
        nh = [_fff(t), _fff(t)]
        for i in (0, 1):
            febe_adjoint(h2b[i], nh[i], dt / 2.0, t)
        g = self._advection_t(B['h1'], nh)
        nq1, nq2 = [_fff(t), _fff(t)], [_fff(t), _fff(t)]
        for i in (0, 1):
            _add(h2b[i], g[i])
            dealias(h2b[i], t)
            febe_adjoint(h2b[i], nq1[i], dt / 2.0, t)
            febe_adjoint(rb[i], nq2[i], dt, t)
            _add(_add(nq1[i], nq2[i]), npb[i])
        g = self._advection_t(B['q'], nq1)
        out = []
        for i in (0, 1):
            qb = _add(_add(h2b[i], rb[i]), g[i])
            dealias(qb, t)
            out.append(qb)
        return out

    def gradient(self, psi0, chi0, nsteps, base, timing=None):
        """(J, (grad_psi, grad_chi)): the perturbation energy against `base`
        after the bootstrap and nsteps steps from (psi0, chi0), and its
        gradient with respect to psi0 and chi0 (data and log coefficients).
        timing: a dict that receives the forward and reverse seconds."""
        t, seg = self.tfm, self.segment
        t0 = time.perf_counter()
        B = self.bootstrap(psi0, chi0)
        run = self._run(B['X0'])
        ckpt = {0: tuple(f.copy() for f in B['X0'])}
        for k in range(nsteps):
            if k and k % seg == 0:
                ckpt[k] = self._save(run)
            run.step(1, check=False)
        run.step(0, check=True)
        sK = (run.field('psi').copy(), run.field('chi').copy())
        t.sync()
        t1 = time.perf_counter()
        J, gp, gc = self.perturbation_energy(sK, base)
        sb, nb, npb = [gp, gc], [_fff(t), _fff(t)], [_fff(t), _fff(t)]
        self._scratch(seg)
        t.sync()
        t2 = time.perf_counter()
        for k0 in sorted(ckpt, reverse=True):
            k1 = min(k0 + seg, nsteps)
            self._load(run, ckpt[k0])
            self._record(run, k1 - k0)
            for k in range(k1 - 1, k0 - 1, -1):
                self._reverse_step(k - k0, sb, nb, npb)
        g = self._reverse_bootstrap(B, sb, nb, npb)
        t.sync()
        t3 = time.perf_counter()
        if timing is not None:
            timing['forward_s'] = t1 - t0
            timing['reverse_s'] = t3 - t2
        return J, g
