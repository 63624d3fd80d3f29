"""3D vortex time step of the oracle (test infrastructure only).

Restates the stepping of ``src/apps/vortical_flow_3d.f90``: advection_rhs
(:353-395), the FE-BE + Richardson bootstrap (:116-147), the AB2/CN loop body
(:160-181), check_stability (:397-409) and the vorticity-magnitude field of
save_vort_mag (:411-446).
"""
import numpy as np

from . import scalar_ops as so


def advection_rhs(psi, chi, uz, kit):
    """N = P_TP[(U+V) x curl V], de-aliased before and after (:353-395).

    psi and chi are de-aliased in place, as the reference does.
    Returns (nlpsi, nlchi).
    """
    for s in (psi, chi):
        if s.space != 'FFF':
            raise ValueError('advection_rhs: psi and chi must be in FFF')
    if uz.space != 'PPP':
        raise ValueError('advection_rhs: uz must be in PPP')
    so.dealias(psi, kit)
    so.dealias(chi, kit)
    vr, vp, vz = so.tp2vec(psi, chi, kit)
    vz.e = vz.e + uz.e
    wr, wp, wz = so.tp2curlvec(psi, chi, kit)
    so.vecprod(vr, vp, vz, wr, wp, wz, kit)
    nlpsi, nlchi = so.vec2tp(vr, vp, vz, kit)
    nlpsi.ln = 0.0
    nlchi.ln = 0.0
    so.dealias(nlpsi, kit)
    so.dealias(nlchi, kit)
    return nlpsi, nlchi


def vorticity_magnitude(psi, chi, kit):
    """The field save_vort_mag writes (:411-446): |curl V| from tp2curlvec,
    on the real and imaginary parts of each packed PPP entry separately (the
    azimuthal points 2j-1 and 2j).  Returns a PPP scalar with ln = 0."""
    for s in (psi, chi):
        if s.space != 'FFF':
            raise ValueError('save_vort_mag: psi and chi must be in FFF')
    wr, wp, wz = so.tp2curlvec(psi, chi, kit)
    out = so.Scalar(kit, 'PPP')
    re = np.sqrt(wr.e.real ** 2 + wp.e.real ** 2 + wz.e.real ** 2)
    im = np.sqrt(wr.e.imag ** 2 + wp.e.imag ** 2 + wz.e.imag ** 2)
    out.e = re + 1j * im
    return out


def energy(vr, vp, vz, kit):
    """Kinetic energy 1/2 int |u|^2 r dr dphi dz of a PPP velocity by physical
    quadrature: r dr = L^2 dx / (1 - x)^2 on the Gauss-Legendre nodes, the
    packed phi points (real and imaginary parts) at 2 pi / np, the z points
    at zlen / nz.  No counterpart in the reference (a diagnostic it lacks,
    docs/release_notes_v1.1.x.md:48-50)."""
    for s in (vr, vp, vz):
        if s.space != 'PPP':
            raise ValueError('energy: velocity components must be in PPP')
    nr, nph = kit.nr, kit.np // 2
    wq = kit.w * kit.ell ** 2 / (1.0 - kit.x) ** 2
    q = sum(np.abs(s.e[:nr, :nph, :]) ** 2 for s in (vr, vp, vz))
    cell = (2.0 * np.pi / kit.np) * (kit.zlen / kit.nz)
    return 0.5 * cell * float(np.einsum('i,ijk->', wq, q))


def check_stability(psi, chi):
    """Non-finite check (:397-409); raises instead of MPI_abort(205)."""
    if not (np.all(np.isfinite(psi.e)) and np.all(np.isfinite(chi.e))):
        raise FloatingPointError('non-finite vortex state')


class VortexRun:
    """State of the vortex driver: psi, chi, their AB2 history and SVV gains."""

    def __init__(self, kit, p, psi, chi, uz):
        self.kit, self.p = kit, p
        self.psi, self.chi, self.uz = psi, chi, uz
        self.gain_psi = 0.0
        self.gain_chi = 0.0
        self.curr_n = 0

    def bootstrap(self):
        """FE-BE + Richardson extrapolation to the first step (:116-147)."""
        kit, p = self.kit, self.p
        dt = p.dt
        psi, chi, uz = self.psi, self.chi, self.uz
        nlpsi, nlchi = advection_rhs(psi, chi, uz, kit)
        self.psi_prev, self.nlpsi_prev = psi.copy(), nlpsi.copy()
        self.chi_prev, self.nlchi_prev = chi.copy(), nlchi.copy()
        psi_rich, chi_rich = psi.copy(), chi.copy()
        nlpsi, nlchi = advection_rhs(psi_rich, chi_rich, uz, kit)
        so.febe(psi_rich, nlpsi, dt, kit, p)
        so.febe(chi_rich, nlchi, dt, kit, p)
        for _ in range(2):
            nlpsi, nlchi = advection_rhs(psi, chi, uz, kit)
            so.febe(psi, nlpsi, dt / 2.0, kit, p)
            so.febe(chi, nlchi, dt / 2.0, kit, p)
        psi.e = 2.0 / 1.0 * psi.e - 1.0 / 1.0 * psi_rich.e
        chi.e = 2.0 / 1.0 * chi.e - 1.0 / 1.0 * chi_rich.e
        so.dealias(psi, kit)
        so.dealias(chi, kit)
        self.gain_psi = so.svv_filter(psi, kit, self.gain_psi, p)
        self.gain_chi = so.svv_filter(chi, kit, self.gain_chi, p)
        so.zeroat1(psi, kit)
        so.zeroat1(chi, kit)
        self.nlpsi, self.nlchi = advection_rhs(psi, chi, uz, kit)
        self.curr_n = 1

    def step(self, dt_step=None):
        """One AB2/CN step of the main loop (:160-181)."""
        kit, p = self.kit, self.p
        dt = p.dt if dt_step is None else dt_step
        so.abcn(self.psi, self.psi_prev, self.nlpsi, self.nlpsi_prev, dt, kit, p)
        so.abcn(self.chi, self.chi_prev, self.nlchi, self.nlchi_prev, dt, kit, p)
        so.dealias(self.psi, kit)
        so.dealias(self.chi, kit)
        self.gain_psi = so.svv_filter(self.psi, kit, self.gain_psi, p)
        self.gain_chi = so.svv_filter(self.chi, kit, self.gain_chi, p)
        so.zeroat1(self.psi, kit)
        so.zeroat1(self.chi, kit)
        self.nlpsi, self.nlchi = advection_rhs(self.psi, self.chi, self.uz, kit)
        check_stability(self.psi, self.chi)
        self.curr_n += 1
