"""Reverse-mode gradients through steps built from the operators that have
hand-written adjoints (mlx_trans_adjoint, mlx_vecprod_adjoint,
mlx_abcn_adjoint, and the self-adjoint projections dealias and chop), with
checkpoint-and-recompute.

The reference has no adjoint.  ``ModelChain`` is a nonlinear time step on
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
import time

import numpy as np

from . import (Scalar, trans, dealias, chop, vecprod, abcn, trans_adjoint, vecprod_adjoint, abcn_adjoint,
               adjoint_project)


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
