"""CPU FP64 restatement of the MLegS spectral-transform time step.

TEST INFRASTRUCTURE ONLY.  This package is the checker for the HIP path in
``mlegs_amd``: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product never calls
it and has no CPU fallback.

It restates, in numpy/scipy, the algorithm of the reference (UCBCFD/MLegS
v1.1.3); every function names the reference file:line it follows.  The
reference itself cannot be built in this image (no gfortran, no Fortran MPI
module, empty LAPACK submodule), so this restatement is pinned instead by
multiprecision (mpmath) fixtures, the closed forms printed in the reference's
tutorials, and identities of the method (see ``tests/test_oracle_*.py``).

Modules
  kit          Gauss-Legendre nodes, log-norms, P_n^m tables, at0/at1, band
               operator coefficients (tfm_kit_init, leg_xxdx/del2h/del2).
  scalar_ops   type(scalar) and the field operations on the path: trans and
               its FFT/Legendre kernels, chop, dealias, svv_filter, calcat0/1,
               zeroat1, delsqp/idelsqp, xxdx, del2h, del2, idel2, ihelm,
               febe, abcn, vecprod, tp2vec, tp2curlvec, vec2tp.
  vortex       advection_rhs, the Richardson bootstrap and the AB2/CN loop of
               the 3D vortex driver, check_stability.
"""
from . import kit, scalar_ops, vortex  # noqa: F401
