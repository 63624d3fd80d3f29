"""The Fortran host (mlegs_amd/fortran) end to end: the vortex driver built
with amdflang runs its bootstrap and AB2/CN steps through the C ABI, and its
final psi and chi equal the same sequence of operators called from Python
(the same library, so agreement is to FP64 rounding)."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, 'mlegs_amd', 'bin', 'vortical_flow_3d')

PARAMS = """!!! COMPUTATIONAL DOMAIN INFO !!!
# ---------- NR ----------- NP ----------- NZ ---------
             48             32             16
# ------ NRCHOP ------- NPCHOP ------- NZCHOP ---------
             48             17              9
# --------- ELL --------- ZLEN ------ ZLENxPI ---------
           2.D0           2.D0              T
#
!!! TIME STEPPING INFO !!!
# ---------- DT ----------- TI --------- TOTT ----------- NI --------- TOTN ----
          1.D-3           0.D0           3.D-3              0              3
#
!!! FIELD PROPERTY INFO !!!
# -------- VISC ----- HYPERPOW ---- HYPERVISC ---------
          1.D-3              0           0.D0
#
!!! FIELD SAVE INFO !!!
# --------------------- FLDDIR ------------------------
                  ./out/fld
# ---- ISFLDSAV -- FLDSAVINTVL -------------------------
              F            100
#
!!! DATA LOGGING INFO !!!
# --------------------- LOGDIR ------------------------
                  ./out/dat
# ---- ISLOGSAV -- LOGSAVINTVL -------------------------
              F              1
!!! STABILITY CONTROL INFO (OPTIONAL) !!!
# ------- IS_SVV ------- CUTOFF ------- TARGET ------ STRENGTH ------- RELAX ---
              T          0.75          2.D-2          0.12          0.25
"""


def _load_dump(fn):
    raw = open(fn, 'rb').read()
    n = np.frombuffer(raw[:24], dtype=np.int64)
    ln = np.frombuffer(raw[24:32], dtype=np.float64)[0]
    e = np.frombuffer(raw[32:], dtype=np.complex128).reshape(tuple(n), order='F')
    return e, ln


def _python_run(pfile, mixed=False):
    """vortical_flow_3d's sequence, operator by operator, from Python; mixed:
    FP32 MFMA contractions with the projection's column 0 on FP64 MFMA."""
    import mlegs_amd as M
    from mlegs_amd import inputs
    from mlegs_amd.params import read_input, timestep_set
    p = read_input(pfile, make_dirs=False)
    timestep_set(p)
    tfm = M.Kit(p, device=0, precision=M.PREC_FP32 if mixed else M.PREC_FP64)
    if mixed:
        tfm.set_stage_precision(M.STAGE_PROJ0, M.PREC_FP64_MFMA)
    nr, np_, nz = p.nr, p.np, p.nz
    r, x = tfm.r, tfm.x
    ph = 2 * math.pi / np_ * np.arange(np_ + 1)
    flds = []
    for which, fid in (('psi', 1), ('chi', 2)):
        amp = 2.0 if which == 'psi' else 1.0 / 1.0
        glb = np.zeros(tfm.dims, dtype=np.complex128)
        for xo in (-2, 2):
            for j in range(np_ // 2):
                rr = np.sqrt((r[:nr] * math.cos(ph[2 * j]) - xo) ** 2 + (r[:nr] * math.sin(ph[2 * j])) ** 2)
                ri = np.sqrt((r[:nr] * math.cos(ph[2 * j + 1]) - xo) ** 2 + (r[:nr] * math.sin(ph[2 * j + 1])) ** 2)
                val = -np.exp(-rr ** 2) * amp / (1 - x[:nr]) ** 2 + 1j * (-np.exp(-ri ** 2) * amp / (1 - x[:nr]) ** 2)
                glb[:nr, j, :nz] += val[:, None]
        s = M.Scalar(tfm).upload(glb, space='PPP')
        M.trans(s, 'FFF', tfm)
        M.idelsqp(s, tfm)
        pert = inputs.seed_spectral_native(tfm.dims, (p.nrchop, p.npchop, p.nzchop, np_, nz), seed=20261202,
                                           field_id=fid, amp=1e-10, scale=4.0, kind=0)
        s.axpby(1.0, M.Scalar(tfm).upload(pert, space='FFF'), 1.0)
        M.zeroat1(s, tfm)
        flds.append(s)
    psi, chi = flds
    uzh = np.zeros(tfm.dims, dtype=np.complex128)
    uzh[:nr, :np_ // 2, :nz] = -0.5 - 0.5j
    uz = M.Scalar(tfm).upload(uzh, space='PPP')
    nlpsi, nlchi = M.Scalar(tfm, 'FFF'), M.Scalar(tfm, 'FFF')
    M.advection_rhs(psi, chi, nlpsi, nlchi, uz, tfm)
    psi_prev, nlpsi_prev, chi_prev, nlchi_prev = psi.copy(), nlpsi.copy(), chi.copy(), nlchi.copy()
    psi_rich, chi_rich = psi.copy(), chi.copy()
    M.advection_rhs(psi_rich, chi_rich, nlpsi, nlchi, uz, tfm)
    M.febe(psi_rich, nlpsi, p.dt, tfm)
    M.febe(chi_rich, nlchi, p.dt, tfm)
    for _ in range(2):
        M.advection_rhs(psi, chi, nlpsi, nlchi, uz, tfm)
        M.febe(psi, nlpsi, p.dt / 2, tfm)
        M.febe(chi, nlchi, p.dt / 2, tfm)
    for s, rich in ((psi, psi_rich), (chi, chi_rich)):
        ln = s.ln
        s.axpby(2.0, rich, -1.0)
        s.ln = ln
    gains = [0.0, 0.0]
    M.dealias(psi, tfm); M.dealias(chi, tfm)
    gains[0] = M.svv_filter(psi, tfm, gains[0]); gains[1] = M.svv_filter(chi, tfm, gains[1])
    M.zeroat1(psi, tfm); M.zeroat1(chi, tfm)
    M.advection_rhs(psi, chi, nlpsi, nlchi, uz, tfm)
    n = 1
    while n < p.ni + p.totaln:
        n += 1
        dts = p.dt if n - p.ni != p.totaln else p.totaltime - (p.totaln - 1) * p.dt
        M.abcn(psi, psi_prev, nlpsi, nlpsi_prev, dts, tfm)
        M.abcn(chi, chi_prev, nlchi, nlchi_prev, dts, tfm)
        M.dealias(psi, tfm); M.dealias(chi, tfm)
        gains[0] = M.svv_filter(psi, tfm, gains[0]); gains[1] = M.svv_filter(chi, tfm, gains[1])
        M.zeroat1(psi, tfm); M.zeroat1(chi, tfm)
        M.advection_rhs(psi, chi, nlpsi, nlchi, uz, tfm)
        M.check_stability(psi, chi, tfm)
    out = (psi.e, psi.ln, chi.e, chi.ln, n)
    tfm.close()
    return out


def test_fortran_driver_matches_python_sequence(tmp_path):
    print('[case] fortran vortical_flow_3d 48x32x16, 3 steps', flush=True)
    assert os.path.exists(DRIVER), 'build the library first (python -m mlegs_amd.build)'
    pfile = tmp_path / 'input.params'
    pfile.write_text(PARAMS)
    r = subprocess.run([DRIVER, str(pfile), str(tmp_path / 'dump')], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert 'time stepping #          3' in r.stdout and 'Program Finished' in r.stdout
    fpsi, fpsi_ln = _load_dump(tmp_path / 'dump_psi.bin')
    fchi, fchi_ln = _load_dump(tmp_path / 'dump_chi.bin')
    ppsi, ppsi_ln, pchi, pchi_ln, n = _python_run(str(pfile))
    assert n == 3
    assert rel_err(fpsi, ppsi) < 1e-13 and rel_err(fchi, pchi) < 1e-13
    assert abs(fpsi_ln - ppsi_ln) <= 1e-13 * max(1.0, abs(ppsi_ln))
    assert abs(fchi_ln - pchi_ln) <= 1e-13 * max(1.0, abs(pchi_ln))


def test_fortran_driver_mixed_precision(tmp_path):
    """precision 'mixed': the driver's run equals the same sequence from Python
    in that preset, and stays within 1e-6 of the FP64 run."""
    print('[case] fortran vortical_flow_3d 48x32x16, 3 steps, mixed', flush=True)
    assert os.path.exists(DRIVER)
    pfile = tmp_path / 'input.params'
    pfile.write_text(PARAMS)
    r = subprocess.run([DRIVER, str(pfile), str(tmp_path / 'dump'), 'mixed'], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert 'mixed T' in r.stdout and 'Program Finished' in r.stdout
    fpsi, _ = _load_dump(tmp_path / 'dump_psi.bin')
    fchi, _ = _load_dump(tmp_path / 'dump_chi.bin')
    mpsi, _, mchi, _, n = _python_run(str(pfile), mixed=True)
    assert n == 3
    assert rel_err(fpsi, mpsi) < 1e-13 and rel_err(fchi, mchi) < 1e-13
    dpsi, _, dchi, _, _ = _python_run(str(pfile))
    assert 0 < rel_err(fchi, dchi) < 1e-6 and rel_err(fpsi, dpsi) < 1e-6


def test_fortran_driver_rejects_bad_input(tmp_path):
    """read_input's checks stop the program with a non-zero status (as the reference's error stop)."""
    assert os.path.exists(DRIVER)
    pfile = tmp_path / 'input.params'
    pfile.write_text(PARAMS.replace('          1.D-3              0           0.D0',
                                    '          1.D-3             10           1.D-7'))
    r = subprocess.run([DRIVER, str(pfile)], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert 'hyperpow' in (r.stdout + r.stderr)


def test_fortran_driver_saves_vorticity_magnitude(tmp_path):
    """ISFLDSAV = T: the driver writes save_vort_mag's field at the start, after
    the bootstrap and every FLDSAVINTVL steps (vortical_flow_3d.f90:112, 157,
    194-197) as <FLDDIR>vormagfld<n, 6 digits>_PPP in the msave ASCII format;
    the last one equals |curl u| of the final psi, chi it dumps."""
    import mlegs_amd as M
    from mlegs_amd.params import read_input, timestep_set
    print('[case] fortran vortical_flow_3d 48x32x16, field saves', flush=True)
    assert os.path.exists(DRIVER)
    pfile = tmp_path / 'input.params'
    pfile.write_text(PARAMS.replace('              F            100', '              T              3'))
    r = subprocess.run([DRIVER, str(pfile), str(tmp_path / 'dump')], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    fld = tmp_path / 'out' / 'fld'
    names = sorted(os.listdir(fld))
    assert names == ['vormagfld000000_PPP', 'vormagfld000001_PPP', 'vormagfld000003_PPP']
    assert 'Field Data Saved at Time Step #000003' in r.stdout
    p = read_input(str(pfile), make_dirs=False)
    timestep_set(p)
    tfm = M.Kit(p, device=0)
    fpsi, fpsi_ln = _load_dump(tmp_path / 'dump_psi.bin')
    fchi, fchi_ln = _load_dump(tmp_path / 'dump_chi.bin')
    psi = M.Scalar(tfm).upload(fpsi, space='FFF', ln=fpsi_ln)
    chi = M.Scalar(tfm).upload(fchi, space='FFF', ln=fchi_ln)
    vm = M.Scalar(tfm)
    M.vorticity_magnitude(psi, chi, vm, tfm)
    want = vm.e
    got, sp, ln, off = M.file_read(str(fld / 'vormagfld000003_PPP'), tfm.dims)
    assert sp == 'PPP' and ln == 0.0 and off == (0, 0, 0)
    assert rel_err(got, want) < 1e-15
    # the tutorials' reader takes the first z slice
    d = np.loadtxt(str(fld / 'vormagfld000003_PPP'), skiprows=1, max_rows=tfm.dims[0])
    assert d.shape == (tfm.dims[0], 2 * tfm.dims[1])
    tfm.close()
