#!/usr/bin/env python
"""Benchmark of the spectral-transform time step on one MI355X.

Workload (one "step" = one AB2/CN time step of the 3D vortex driver,
vortical_flow_3d.f90:160-181): NR = NP = NZ = 256, NRCHOP = 256,
NPCHOP = NZCHOP = 129, L = 1, ZLEN = 2 pi, VISC = 5e-4, dt = 5e-4, SVV on;
initial state = centred q = 4 vortex + seeded 1e-4 perturbation (synthetic,
SplitMix64/Box-Muller).  The FE-BE + Richardson bootstrap runs untimed, then
--warmup steps, then --steps timed steps between device synchronisations.

Precisions (--prec) of the Legendre contractions; everything else is FP64:
  mixed  (default) FP32 MFMA, except the m = 0 column of vec2tp's projection
         in FP64 MFMA (on a second stream, concurrent with the rest)
  mixedp FP32 MFMA, except vec2tp's whole projection in FP64 MFMA
  bf16   3-term split-BF16 MFMA (six products, FP32 accumulation), the
         projection in FP32 except its m = 0 column in FP64 MFMA
  fp32   FP32 MFMA everywhere
  fp64m  FP64 MFMA everywhere
  fp64   the plain FP64 path: FP64 vector-FMA contractions over every column,
         no de-aliasing restrictions (the reference's operation set)

Printed: one JSON line with
  value            steps/s of the selected precision
  fp64_steps_per_s the same-GPU plain FP64 path over the same steps
  gate             the accuracy gate of the metric: the selected precision's
                   psi and chi after a --traj-steps trajectory against the
                   plain FP64 path's, norm-wise relative, <= 1e-6
  rel_err_vs_fp64  that trajectory error, the error of one step from
                   identical FP64 states (per field), and per transform
                   direction of one field (the transform-only input)
  ke_history       FP64 kinetic energy of the perturbation, 1/2 int |u - U|^2
                   (U: the same run without perturbation, plain FP64), along
                   the trajectory for the selected precision and plain FP64
  other_precisions steps/s and trajectory error of the other precisions
  roofline         the kernel family with the most kernel time, from a second
                   timed region: --steps more steps of the same run in which
                   every launch of the probed kernel families is timed by HIP
                   events that the dispatch itself stamps (library stream);
                   algorithmic bytes or flops per launch come from the library
                   (fixed per kernel and chop).  The event stamping slows the
                   step, so `value` comes from the first, unprobed region and
                   the probed region's rate is reported as probed_steps_per_s
  transform_c2     the transform-only workload (one 256^3 field, forward +
                   backward 1000 times) per precision of trans: round trips/s,
                   errors against plain FP64 (one backward, one round trip,
                   after the 1000), the contractions' achieved rate
  cpu_baseline     the oracle (numpy/scipy FP64) timing one AB2/CN step at the
                   same size on the host, on the job's threads (`cores`) and
                   on one thread, with the CPU model
  ensemble_c4      64 2D Lamb-Oseen runs of uneven NR stepped together by
                   HIP-graph replay: member-steps/s, tuned launch shape,
                   members checked against their own single runs
  adjoint_c5       the gradient of the perturbation kinetic energy after
                   --adjoint steps at 128 x 64 x 64 (q = 4 vortex + 1e-4
                   perturbation, VISC = 1e-3, dt = 1e-3, SVV off) with respect
                   to the initial psi, chi: forward mixed, adjoints FP64,
                   checkpoints every --adjoint-segment steps; reverse/forward
                   time ratio (adjoint_forward_ratio), central differences on
                   --adjoint-dirs directions, mixed vs FP64-forward gradient
  hyperviscous_p8  the same 256^3 step with HYPERPOW = 8, HYPERVISC = 5e-7 (the
                   reference's shipped setting): steps/s mixed and FP64 over
                   --hyper steps, and the two runs' difference
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'steps/s @256³ mixed-prec vs FP64 at ≤1e-6 rel.err; adjoint/forward time ratio'
PEAK_TFLOPS = {'fp64': 78.6, 'fp32': 157.3, 'bf16': 2500.0}
HBM_GBPS = 8000.0
HBM_BOUND = ('fft_phi', 'fft_z', 'lu_solve', 'phi_cross', 'tp_front', 'abcn_rhs')
TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'traffic_256.json')


def log(*a):
    print('[bench]', *a, file=sys.stderr, flush=True)


def workload(n=256):
    import mlegs_amd as M
    return M.Params(nr=n, np=n, nz=n, nrchop=n, npchop=n // 2 + 1, nzchop=n // 2 + 1, ell=1.0,
                    zlen=2 * math.pi, visc=5e-4, dt=5e-4, hyperpow=0, is_svv=True)


def initial_state(tfm, p, amp=1e-4, q=4.0):
    """psi, chi of the q-vortex (qvort_dist_tp pattern, one vortex) plus the seeded perturbation."""
    import mlegs_amd as M
    from mlegs_amd import inputs
    dims = tfm.dims
    ch = (p.nrchop, p.npchop, p.nzchop, p.np, p.nz)
    out = []
    for which, fid in (('psi', 1), ('chi', 2)):
        s = M.Scalar(tfm).upload(inputs.qvortex_ppp(tfm.x, tfm.r, dims, p.np, p.nz, p.nr, q, which), space='PPP')
        M.trans(s, 'FFF', tfm)
        M.idelsqp(s, tfm)
        M.zeroat1(s, tfm)
        pert = M.Scalar(tfm).upload(inputs.perturbation(dims, ch, amp, field_id=fid), space='FFF')
        s.axpby(1.0, pert, 1.0)
        M.dealias(s, tfm)
        M.zeroat1(s, tfm)
        out.append(s)
    return out


def rel(a, b):
    import numpy as np
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


PRESETS = ('mixed', 'mixedp', 'bf16', 'fp32', 'fp64m', 'fp64')
DTYPES = {'fp64': 'f64', 'fp64m': 'f64-mfma', 'fp32': 'f32-mfma+f64', 'mixed': 'f32-mfma+f64-mfma',
          'mixedp': 'f32-mfma+f64-mfma', 'bf16': 'bf16x3-mfma+f64-mfma', 'bf16x3': 'bf16x3-mfma+f64'}
GATE_TOL = 1e-6


def configure(tfm, name):
    """Contraction precisions of a preset (context precision + site overrides)."""
    import mlegs_amd as M
    base = {'fp64': M.PREC_FP64, 'fp32': M.PREC_FP32, 'fp64m': M.PREC_FP64_MFMA, 'mixed': M.PREC_FP32,
            'mixedp': M.PREC_FP32, 'bf16': M.PREC_BF16X3, 'bf16x3': M.PREC_BF16X3}[name]
    tfm.set_precision(base)
    for st in (M.STAGE_SYNTH, M.STAGE_UP, M.STAGE_PROJ, M.STAGE_PROJ0):
        tfm.set_stage_precision(st, -1)
    if name in ('mixed', 'mixedp', 'bf16'):
        tfm.set_stage_precision(M.STAGE_PROJ0, M.PREC_FP64_MFMA)
    if name == 'mixedp':
        tfm.set_stage_precision(M.STAGE_PROJ, M.PREC_FP64_MFMA)


def timed_run(tfm, name, psi0, chi0, warmup, steps):
    """Bootstrap + warmup untimed, then `steps` timed steps; returns (run, seconds)."""
    import mlegs_amd as M
    configure(tfm, name)
    run = M.VortexRun(tfm, psi0, chi0, M.Scalar(tfm))
    run.bootstrap()
    if warmup:
        run.step(warmup, check=True)
    tfm.sync()
    t0 = time.perf_counter()
    run.step(steps, check=False)
    tfm.sync()
    t1 = time.perf_counter()
    run.step(0, check=True)  # raises on a non-finite state
    return run, t1 - t0


def perturbation_energy(tfm, run, base):
    """1/2 int |u - U|^2 in FP64: tp2vec of (psi - psi_base, chi - chi_base)."""
    import mlegs_amd as M
    dp = run.field('psi').copy()
    dc = run.field('chi').copy()
    dp.axpby(1.0, base[0], -1.0)
    dc.axpby(1.0, base[1], -1.0)
    vr, vp, vz = M.Scalar(tfm), M.Scalar(tfm), M.Scalar(tfm)
    M.tp2vec(dp, dc, vr, vp, vz, tfm)
    return M.energy(vr, vp, vz, tfm)


def trajectory(tfm, name, psi0, chi0, nsteps, every, base_states=None):
    """Bootstrap + nsteps steps in `name`; returns (psi, chi) on the host and
    the perturbation energy at steps 0, every, ... (when base_states holds
    the unperturbed run's (psi, chi) at those steps)."""
    import mlegs_amd as M
    configure(tfm, name)
    run = M.VortexRun(tfm, psi0, chi0, M.Scalar(tfm))
    ke = []
    states = []
    done = 0
    run.bootstrap()
    marks = list(range(0, nsteps + 1, every))
    if marks[-1] != nsteps:
        marks.append(nsteps)
    for mk in marks:
        if mk > done:
            configure(tfm, name)
            run.step(mk - done, check=True)
            done = mk
        if base_states is not None:
            configure(tfm, 'fp64')
            ke.append(perturbation_energy(tfm, run, base_states[len(ke)]))
        else:
            states.append((run.field('psi').copy(), run.field('chi').copy()))
    out = (run.field('psi').e, run.field('chi').e)
    del run
    return out, ke, states, marks


def step_errors(tfm, name, psi0, chi0, presteps=2):
    """From the same FP64 state, one step in `name` against one in FP64.

    Inside one step the transforms only produce the nonlinear terms N for the
    next step, so the error of one step is the error of N (nlpsi, nlchi)."""
    import mlegs_amd as M
    runs = []
    for _ in range(2):
        configure(tfm, 'fp64')
        r = M.VortexRun(tfm, psi0, chi0, M.Scalar(tfm))
        r.bootstrap()
        r.step(presteps, check=True)
        runs.append(r)
    configure(tfm, 'fp64')
    runs[0].step(1)
    configure(tfm, name)
    runs[1].step(1)
    out = {nm: rel(runs[1].field(nm).e, runs[0].field(nm).e) for nm in ('nlpsi', 'nlchi')}
    del runs
    return out


def transform_error(tfm, p, name):
    """Transform-only input (|c| ~ N(0,1) exp(-n/32)): backward and round-trip
    error of trans at the preset's context precision vs plain FP64."""
    import mlegs_amd as M
    from mlegs_amd import inputs
    e = inputs.transform_field(tfm.dims, (p.nrchop, p.npchop, p.nzchop, p.np, p.nz), field_id=3)
    res = {}
    for nm in ('fp64', name):
        configure(tfm, nm)
        g = M.Scalar(tfm).upload(e, space='FFF')
        M.trans(g, 'PPP', tfm)
        b = g.e[:p.nr, :p.np // 2].copy()
        M.trans(g, 'FFF', tfm)
        res[nm] = (b, g.e.copy())
    (b64, f64), (bx, fx) = res['fp64'], res[name]
    return {'backward': rel(bx, b64), 'round_trip': rel(fx, f64)}


# arithmetic type of each probed family's contractions under a preset
def family_precision(name, family):
    if name in ('fp64', 'fp64m'):
        return 'fp64'
    if name in ('bf16', 'bf16x3') and family in ('leg_fwd', 'leg_bwd'):
        return 'bf16'
    if name == 'mixedp' and family == 'v2t':
        return 'fp64'
    return 'fp32' if family in ('leg_fwd', 'leg_bwd', 'v2t') else 'fp64'


def roofline(probes, name):
    """The probe with the most kernel time; achieved = algorithmic work / kernel time.

    The bound is the resource whose peak time is the longer for the launch's
    algorithmic bytes and flops (arithmetic intensity against the ridge
    point peak flops / peak bandwidth); the FFTs and the solve count as
    HBM-bound whatever their flop count."""
    fam = max(probes, key=lambda k: probes[k][1])
    n, ms, nbytes, flops = probes[fam]
    prec = family_precision(name, fam)
    per_launch_s = ms * 1e-3 / n
    t_hbm = nbytes / n / (HBM_GBPS * 1e9)
    t_mfma = flops / n / (PEAK_TFLOPS[prec] * 1e12)
    if fam in HBM_BOUND or t_hbm >= t_mfma:
        achieved = nbytes / n / per_launch_s / 1e9
        peak, unit, bound = HBM_GBPS, 'GB/s', 'hbm'
    else:
        achieved = flops / n / per_launch_s / 1e12
        peak, unit, bound = PEAK_TFLOPS[prec], 'TFLOP/s', 'mfma'
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        entry = t.get('kernels', {}).get(fam)
        # the counters hold for the launch they were taken on: the same
        # family, precision and launch shape (its algorithmic bytes per launch)
        if entry and entry.get('precision') in (prec, 'any'):
            alg = entry.get('algorithmic_bytes_per_launch')
            if alg and abs(alg - nbytes / n) <= 5e-3 * alg:
                traffic = entry.get('hbm_bytes_per_launch')
    return {'bound': bound, 'kernel': fam, 'precision': prec, 'achieved': achieved, 'peak': peak, 'unit': unit,
            'frac': achieved / peak, 'traffic': traffic,
            'launches': n, 'avg_launch_us': 1e3 * ms / n,
            'algorithmic_bytes_per_launch': nbytes / n, 'algorithmic_flops_per_launch': flops / n}


def probe_table(probes):
    out = {}
    for k, (n, ms, nbytes, flops) in probes.items():
        if n == 0:
            continue
        s = ms * 1e-3 / n
        out[k] = {'launches': n, 'avg_us': round(1e6 * s, 2), 'total_ms': round(ms, 3),
                  'GB_per_s': round(nbytes / n / s / 1e9, 1), 'TFLOP_per_s': round(flops / n / s / 1e12, 2)}
    return out


def host_info():
    """CPU model, the host's logical CPUs, and the threads this process may use
    (the job's CPU share: OMP_NUM_THREADS when set, else the affinity mask)."""
    model = None
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else (os.cpu_count() or 1)
    env = os.environ.get('OMP_NUM_THREADS')
    threads = min(int(env), avail) if env and env.isdigit() and int(env) > 0 else avail
    return model, os.cpu_count(), avail, threads


def cpu_step_seconds(okit, p, threads):
    """One AB2/CN step of the oracle at the workload's size with `threads`
    BLAS threads and FFT workers, from a synthetic state whose history
    (psi_prev, N, N_prev) is set directly: the step's cost does not depend on
    the values, so the bootstrap is not run."""
    from threadpoolctl import threadpool_limits
    from oracle import scalar_ops as so, vortex as V
    from mlegs_amd import inputs
    dims = (okit.nrdim, okit.npdim, okit.nzdim)
    ch = (p.nrchop, p.npchop, p.nzchop, p.np, p.nz)
    psi = so.Scalar(okit, 'FFF')
    psi.e = inputs.perturbation(dims, ch, 1e-4, field_id=1)
    chi = so.Scalar(okit, 'FFF')
    chi.e = inputs.perturbation(dims, ch, 1e-4, field_id=2)
    run = V.VortexRun(okit, p, psi, chi, so.Scalar(okit, 'PPP'))
    run.psi_prev, run.chi_prev = psi.copy(), chi.copy()
    run.nlpsi, run.nlchi = so.Scalar(okit, 'FFF'), so.Scalar(okit, 'FFF')
    run.nlpsi_prev, run.nlchi_prev = so.Scalar(okit, 'FFF'), so.Scalar(okit, 'FFF')
    saved = so.FFT_WORKERS
    so.FFT_WORKERS = threads
    try:
        with threadpool_limits(limits=threads):
            t0 = time.perf_counter()
            run.step()
            return time.perf_counter() - t0
    finally:
        so.FFT_WORKERS = saved


def cpu_baseline(p, one_core=True):
    """The oracle (numpy/scipy FP64) on the host's cores: a bounded sample of
    the workload, one AB2/CN step, on all the threads the job may use and on
    one thread."""
    from oracle import kit as K
    model, ncpu, avail, threads = host_info()
    okit = K.Kit3D(p)
    t_all = cpu_step_seconds(okit, p, threads)
    out = {'value': 1.0 / t_all, 'unit': 'steps/s', 'cores': threads, 'kind': 'port',
           'cpu_model': model, 'host_logical_cpus': ncpu, 'affinity_cpus': avail,
           'sample': 'oracle AB2/CN step (numpy/scipy FP64; BLAS threads and scipy.fft workers = %d) at '
                     'NR=NP=NZ=%d, one step timed: %.2f s' % (threads, p.nr, t_all)}
    if one_core:
        t1 = cpu_step_seconds(okit, p, 1)
        out['one_core'] = {'value': 1.0 / t1, 'seconds_per_step': t1}
    return out


def transform_bench(tfm, p, names, n_rt, iters_probe=10):
    """The transform-only workload: one 256^3 field, |c| ~ N(0,1) exp(-n/32),
    forward + backward (FFF -> PPP -> FFF) n_rt times per precision of trans.
    Errors against the plain FP64 path: one backward transform, one round
    trip, and after the n_rt round trips.  A probed region of iters_probe
    round trips gives each contraction's time and achieved rate."""
    import mlegs_amd as M
    from mlegs_amd import inputs
    e = inputs.transform_field(tfm.dims, (p.nrchop, p.npchop, p.nzchop, p.np, p.nz), field_id=3)
    out, ref = {}, None
    for nm in names:
        configure(tfm, nm)
        g = M.Scalar(tfm).upload(e, space='FFF')
        M.trans(g, 'PPP', tfm)
        b1 = g.e[:p.nr, :p.np // 2].copy()
        M.trans(g, 'FFF', tfm)
        r1 = g.e.copy()
        for _ in range(2):
            M.trans(g, 'PPP', tfm)
            M.trans(g, 'FFF', tfm)
        g = M.Scalar(tfm).upload(e, space='FFF')
        tfm.sync()
        t0 = time.perf_counter()
        for _ in range(n_rt):
            M.trans(g, 'PPP', tfm)
            M.trans(g, 'FFF', tfm)
        tfm.sync()
        t = time.perf_counter() - t0
        rn = g.e.copy()
        tfm.probe_start(16 * iters_probe)
        for _ in range(iters_probe):
            M.trans(g, 'PPP', tfm)
            M.trans(g, 'FFF', tfm)
        probes = tfm.probe_stop()
        if nm == 'fp64':
            ref = (b1, r1, rn)
        res = {'round_trips_per_s': n_rt / t, 'ms_per_round_trip': 1e3 * t / n_rt, 'dtype': DTYPES[nm],
               'kernels': probe_table(probes)}
        if ref is not None and nm != 'fp64':
            res['rel_err_vs_fp64'] = {'backward': rel(b1, ref[0]), 'round_trip': rel(r1, ref[1]),
                                      'after_%d_round_trips' % n_rt: rel(rn, ref[2])}
        leg = {k: v for k, v in probes.items() if k in ('leg_fwd', 'leg_bwd') and v[0] > 0}
        if leg:
            res['roofline'] = roofline(leg, nm)
        out[nm] = res
        log('transform', nm, '%.1f round trips/s' % res['round_trips_per_s'])
        del g
    return out


def kernel_sample(tfm, p, name, iters):
    """Each probed kernel family on one field of the workload's size (plus a
    plain copy, read-only and write-only pass), at the precision configure()
    has set."""
    import mlegs_amd as M
    from mlegs_amd import inputs
    e = inputs.transform_field(tfm.dims, (p.nrchop, p.npchop, p.nzchop, p.np, p.nz), field_id=3)
    out = {}
    for w in ('copy', 'read', 'write', 'fft_phi_fwd', 'fft_phi_bwd', 'fft_z_fwd', 'fft_z_bwd', 'leg_fwd', 'leg_bwd',
              'v2t', 'lu_solve', 'phi_cross'):
        g = M.Scalar(tfm).upload(e, space='FFF')
        out[w] = round(1e3 * tfm.time_kernel(g, w, iters), 2)
        log(w, out[w], 'us')
    print(json.dumps({'kernel_sample_us': out, 'iters': iters, 'precision': name}), flush=True)


def ensemble_bench(n_members, steps, n_check, eager_steps=40, mode=-1, streams=4):
    """The 2D Lamb-Oseen ensemble: n_members members (NR uniform over
    {64, 96, 128, 192, 256}, NP = NR, ell ~ U[0.5, 2], visc ~ logU[1e-4,
    1e-2], dt = 1e-2 min(1, 64 ell / NR), seeded 1e-3 perturbation),
    bootstrap, then `steps`
    AB2/CN steps replayed as HIP graphs with the tuned launch configuration
    (the tuning replays are steps of the run and are inside the timed
    region).  Also times the same members stepped eagerly (one launch per
    kernel, the members' streams interleaved) and checks n_check members
    against their own single runs."""
    import mlegs_amd as M
    from mlegs_amd import ensemble as E
    params = [E.member_params(i) for i in range(n_members)]
    log('ensemble: %d members, NR counts %s' % (n_members, {nr: sum(p.nr == nr for p in params)
                                                            for nr in E.NR_CHOICES}))
    ens = E.Ensemble(params)
    for i in range(n_members):
        ens.set(i, *E.lamb_oseen_tp(ens.kits[i], i))
    ens.bootstrap()
    ens.sync()
    # eager reference: every member stepped by its own launches, members interleaved
    runs = []
    t0 = time.perf_counter()
    for i in range(n_members):
        v = ctypes.c_void_p()
        ens._check(ens.lib.mlx_ensemble_vortex(ens.h, i, ctypes.byref(v)))
        runs.append(v)
    for _ in range(eager_steps):
        for i, v in enumerate(runs):
            tfm = ens.kits[i]
            M._lib.check(tfm.ctx, ens.lib.mlx_vortex_step(tfm.ctx, v, 1, float(tfm.params.dt), 0))
    ens.sync()
    t_eager = time.perf_counter() - t0
    log('eager: %.1f member-steps/s' % (n_members * eager_steps / t_eager))
    t0 = time.perf_counter()
    ens.step(steps, mode=mode, streams=streams)
    ens.sync()
    t = time.perf_counter() - t0
    bad = ens.check()
    info = ens.info()
    rate = n_members * steps / t
    log('graphs: %d steps in %.3f s, %.1f member-steps/s, %s' % (steps, t, rate,
                                                                {k: v for k, v in info.items() if k != 'tune_ms'}))
    checks = []
    for i in range(min(n_check, n_members)):
        ge = (ens.field(i, 'psi').e, ens.field(i, 'chi').e)
        tfm = M.Kit(params[i], device=0)
        run = M.VortexRun(tfm, *E.lamb_oseen_tp(tfm, i))
        run.bootstrap()
        run.step(eager_steps + steps, check=False)
        se = (run.field('psi').e, run.field('chi').e)
        checks.append({'member': i, 'nr': params[i].nr, 'psi': rel(ge[0], se[0]), 'chi': rel(ge[1], se[1])})
        del run
        tfm.close()
    log('single-run checks', checks)
    tune = info.pop('tune_ms')
    table = {}
    for md in range(3):
        for si, ns in enumerate((1, 2, 4, 8, 16, 32)):
            for ri, rp in enumerate((1, 4)):
                v = tune[md, si, ri]
                if v > 0:
                    nm = 'parent' if md == 0 else ('streams%d' if md == 1 else 'threads%d') % ns
                    table['%s_steps%d' % (nm, rp)] = round(n_members / (v * 1e-3), 1)
    ens.close()
    return {'metric': 'member-steps/s', 'value': rate, 'members': n_members, 'steps': steps, 'seconds': t,
            'eager_member_steps_per_s': n_members * eager_steps / t_eager, 'speedup_vs_eager':
            rate / (n_members * eager_steps / t_eager), 'launch': info, 'tuning_member_steps_per_s': table,
            'nonfinite_member': bad, 'single_run_check': checks,
            'max_rel_diff_vs_single': max(max(c['psi'], c['chi']) for c in checks) if checks else None,
            'workload': '2D Lamb-Oseen (w0 = exp(-r^2/2)/(2 pi)) in TP form + seeded 1e-3 perturbation; NR in '
                        '{64,96,128,192,256}, NP=NR, NZ=1, ell~U[0.5,2], visc~logU[1e-4,1e-2], '
                        'dt=1e-2*min(1,64*ell/NR), '
                        'FP64, SVV on (cutoff 0.5, target 1e-12)'}


def adjoint_bench(nsteps, ndir, segment=20, amp=1e-4):
    """The adjoint leg: the gradient of the perturbation kinetic energy after
    the bootstrap and nsteps steps (T = nsteps dt) with respect to the initial
    psi and chi, for the q = 4 vortex plus a seeded perturbation of amplitude
    amp at 128 x 64 x 64 (L = 1, ZLEN = 2 pi, VISC = 1e-3, dt = 1e-3, SVV off,
    no background flow), checkpoints every `segment` steps.  The forward runs
    in the mixed preset, the adjoints in FP64.  Times nsteps forward steps
    alone and the reverse sweep (recomputation included); checks the
    gradient of the FP64 forward against central differences (eps = 1e-6)
    on ndir seeded directions, and the mixed-forward gradient against the
    FP64-forward one."""
    import mlegs_amd as M
    from mlegs_amd import adjoint as A, inputs
    from mlegs_amd.params import Params
    p = Params(nr=128, np=64, nz=64, nrchop=128, npchop=33, nzchop=33, ell=1.0, zlen=2 * math.pi, visc=1e-3,
               dt=1e-3, hyperpow=0, hypervisc=0.0, is_svv=False)
    tfm = M.Kit(p, device=0)
    uz = M.Scalar(tfm)
    q0 = initial_state(tfm, p, amp=amp)
    qb = initial_state(tfm, p, amp=0.0)
    out = {'steps': nsteps, 'segment': segment, 'perturbation_amplitude': amp}
    grads = {}
    for name in ('mixed', 'fp64'):
        configure(tfm, name)
        chain = A.VortexChain(tfm, uz, segment=segment, configure=lambda t, n=name: configure(t, n))
        base = chain.forward(qb[0], qb[1], nsteps)
        chain.forward(q0[0], q0[1], 2)  # warm-up: factors and temporaries
        run = M.VortexRun(tfm, q0[0], q0[1], uz)
        run.bootstrap()
        tfm.sync()
        t0 = time.perf_counter()
        run.step(nsteps, check=False)
        tfm.sync()
        t_fwd = time.perf_counter() - t0
        run.step(0, check=True)
        del run
        timing = {}
        J, g = chain.gradient(q0[0], q0[1], nsteps, base, timing)
        grads[name] = g
        out[name] = {'forward_ms_per_step': 1e3 * t_fwd / nsteps,
                     'reverse_ms_per_step': 1e3 * timing['reverse_s'] / nsteps,
                     'adjoint_forward_ratio': timing['reverse_s'] / t_fwd,
                     'checkpointed_forward_ms_per_step': 1e3 * timing['forward_s'] / nsteps, 'objective': J}
        if name == 'fp64' and ndir:
            worst = worst_scaled = 0.0
            eps = 1e-6
            gg = A.inner(g, g)
            for d in range(ndir):
                dirs = [M.Scalar(tfm).upload(inputs.perturbation(tfm.dims, (p.nrchop, p.npchop, p.nzchop, p.np, p.nz),
                                                                 1.0, field_id=3000 + 2 * d + i),
                                             space='FFF', ln=0.05 * (d % 3 - 1)) for i in range(2)]
                Jpm = []
                for sgn in (1.0, -1.0):
                    a, b = q0[0].copy(), q0[1].copy()
                    a.axpby(1.0, dirs[0], sgn * eps)
                    b.axpby(1.0, dirs[1], sgn * eps)
                    Jpm.append(chain.perturbation_energy(chain.forward(a, b, nsteps), base, grad=False))
                fd = (Jpm[0] - Jpm[1]) / (2 * eps)
                ad = A.inner(g, dirs)
                worst = max(worst, abs(fd - ad) / max(abs(fd), abs(ad), 1e-300))
                worst_scaled = max(worst_scaled, abs(fd - ad) / math.sqrt(gg * A.inner(dirs, dirs)))
            out.update({'fd_directions': ndir, 'fd_eps': eps, 'max_rel_diff_vs_fd': worst,
                        'max_diff_vs_fd_over_norm_g_norm_d': worst_scaled})
    dg = [a.copy() for a in grads['mixed']]
    for a, b in zip(dg, grads['fp64']):
        a.axpby(1.0, b, -1.0)
    out['mixed_vs_fp64_gradient_rel'] = math.sqrt(A.inner(dg, dg) / A.inner(grads['fp64'], grads['fp64']))
    # the same over the radial modes n < nrchop / 2: the FP32 sums' noise in
    # the forward states sits in the high modes, where the far-field weight
    # of the energy makes the gradient large
    import numpy as np
    half = p.nrchop // 2
    num = sum(float(np.sum(np.abs(d.e[:half]) ** 2)) for d in dg)
    den = sum(float(np.sum(np.abs(g.e[:half]) ** 2)) for g in grads['fp64'])
    out['mixed_vs_fp64_gradient_rel_low_half'] = math.sqrt(num / den)
    out['adjoint_forward_ratio'] = out['mixed']['adjoint_forward_ratio']
    out['workload'] = ('vortex run (bootstrap + AB2/CN steps, dealias, zeroat1, advection_rhs) at 128x64x64, '
                       'q = 4 vortex + seeded perturbation, objective 1/2 int |u - u_base|^2 at T, gradient '
                       'w.r.t. initial psi, chi; forward mixed preset, adjoint FP64')
    tfm.close()
    log('adjoint', out)
    return out


def hyper_bench(n, steps, warmup=2, hyperpow=8, hypervisc=5e-7):
    """The hyperviscous step: the headline workload with HYPERPOW = 8 and
    HYPERVISC = 5e-7, the reference's shipped setting (input.params:15), so
    that every Helmholtz solve is ihelmp's banded (kl = ku = 8) system.  Steps/s
    in the mixed preset and in FP64, and the two runs' relative difference
    after the timed steps."""
    import mlegs_amd as M
    import numpy as np
    p = M.Params(nr=n, np=n, nz=n, nrchop=n, npchop=n // 2 + 1, nzchop=n // 2 + 1, ell=1.0, zlen=2 * math.pi,
                 visc=5e-4, dt=5e-4, hyperpow=hyperpow, hypervisc=hypervisc, is_svv=True)
    tfm = M.Kit(p, device=0, precision=M.PREC_FP64)
    psi0, chi0 = initial_state(tfm, p)
    out = {'workload': 'the 3D vortex step above with HYPERPOW=%d, HYPERVISC=%g' % (hyperpow, hypervisc),
           'steps': steps}
    final = {}
    for nm in ('mixed', 'fp64'):
        log('hyperviscous run', nm)
        run, t = timed_run(tfm, nm, psi0, chi0, warmup, steps)
        out[nm] = {'steps_per_s': steps / t, 'ms_per_step': 1e3 * t / steps}
        final[nm] = (run.field('psi').download(), run.field('chi').download())
        del run
        log('  %.3f ms/step' % out[nm]['ms_per_step'])
    out['mixed_vs_fp64'] = {'psi': rel(final['mixed'][0], final['fp64'][0]),
                            'chi': rel(final['mixed'][1], final['fp64'][1])}
    tfm.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--prec', default='mixed', choices=list(PRESETS))
    ap.add_argument('--size', type=int, default=256)
    ap.add_argument('--cpu-baseline', type=int, default=1)
    ap.add_argument('--errors', type=int, default=1, help='measure the errors against FP64 (the accuracy gate)')
    ap.add_argument('--traj-steps', type=int, default=500, help='length of the gate trajectory')
    ap.add_argument('--ke-every', type=int, default=50, help='perturbation-energy sampling interval')
    ap.add_argument('--others', type=int, default=1, help='also time the other precisions')
    ap.add_argument('--transform', type=int, default=1000,
                    help='round trips of the transform-only workload per precision (0: skip)')
    ap.add_argument('--cpu-one-core', type=int, default=1, help='also time the CPU baseline on one thread')
    ap.add_argument('--compare-fp64', type=int, default=1,
                    help='also time the plain FP64 path (0: the selected precision only, e.g. under a profiler)')
    ap.add_argument('--ensemble', type=int, default=64, help='members of the 2D ensemble leg (0: skip)')
    ap.add_argument('--ensemble-steps', type=int, default=2000)
    ap.add_argument('--ensemble-check', type=int, default=4, help='members checked against single runs')
    ap.add_argument('--ensemble-only', type=int, default=0, help='run the ensemble leg alone and print it')
    ap.add_argument('--ensemble-mode', type=int, default=-1,
                    help='launch mode of the ensemble leg: -1 tuned, 0 parent graph, 1 streams, 2 threaded streams')
    ap.add_argument('--ensemble-streams', type=int, default=4)
    ap.add_argument('--adjoint', type=int, default=200, help='steps of the adjoint leg (0: skip)')
    ap.add_argument('--adjoint-segment', type=int, default=20, help='checkpoint interval of the adjoint leg')
    ap.add_argument('--adjoint-dirs', type=int, default=32, help='finite-difference directions of the adjoint leg')
    ap.add_argument('--hyper', type=int, default=10, help='steps of the HYPERPOW = 8 leg (0: skip)')
    ap.add_argument('--hyper-only', type=int, default=0, help='run the HYPERPOW = 8 leg alone and print it')
    ap.add_argument('--kernel-sample', type=int, default=0,
                    help='only launch each probed kernel family (and a plain field copy) this many times, '
                         'for hardware-counter runs; prints their mean durations')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py: this build runs on one GPU; --gpus %d is not supported (the multi-GPU z-slab '
              'sharding is a design note, not part of this build)' % args.gpus, file=sys.stderr)
        sys.exit(2)

    import mlegs_amd as M

    if args.ensemble_only:
        print(json.dumps(ensemble_bench(args.ensemble, args.ensemble_steps, args.ensemble_check,
                                        mode=args.ensemble_mode, streams=args.ensemble_streams)), flush=True)
        return
    if args.hyper_only:
        print(json.dumps(hyper_bench(args.size, args.hyper)), flush=True)
        return

    p = workload(args.size)
    name = args.prec
    log('kit', p.nr, p.np, p.nz)
    tfm = M.Kit(p, device=0, precision=M.PREC_FP64)
    if args.kernel_sample:
        configure(tfm, name)
        kernel_sample(tfm, p, name, args.kernel_sample)
        return
    psi0, chi0 = initial_state(tfm, p)

    log(name, 'run: warmup', args.warmup, 'steps', args.steps)
    run, t = timed_run(tfm, name, psi0, chi0, args.warmup, args.steps)
    rate, ms = args.steps / t, 1e3 * t / args.steps
    log('%.3f ms/step' % ms)
    fp64_rate = None
    if name == 'fp64':
        fp64_rate = rate
    elif args.compare_fp64:
        log('plain fp64 run')
        r64, t64 = timed_run(tfm, 'fp64', psi0, chi0, args.warmup, args.steps)
        del r64
        fp64_rate = args.steps / t64

    err, gate, ke, others = None, None, None, None
    if args.errors:
        n = args.traj_steps
        log('gate trajectories: %d steps' % n)
        base0 = initial_state(tfm, p, amp=0.0)
        _, _, base_states, marks = trajectory(tfm, 'fp64', base0[0], base0[1], n, args.ke_every)
        (p64, c64), ke64, _, _ = trajectory(tfm, 'fp64', psi0, chi0, n, args.ke_every, base_states)
        trajs = {}
        for nm in ([name] + ([x for x in PRESETS if x not in (name, 'fp64')] if args.others else [])):
            if nm == 'fp64':
                continue
            (ps, cs), kx, _, _ = trajectory(tfm, nm, psi0, chi0, n, args.ke_every, base_states)
            trajs[nm] = ({'psi': rel(ps, p64), 'chi': rel(cs, c64)}, kx)
            log(nm, 'trajectory error', trajs[nm][0])
        del base_states
        if name == 'fp64':
            tr = {'psi': 0.0, 'chi': 0.0}
            kx = ke64
        else:
            tr, kx = trajs[name]
        err = {'trajectory_%d_steps' % n: tr}
        gate = {'tol': GATE_TOL, 'steps': n, 'fields': ['psi', 'chi'],
                'pass': bool(max(tr.values()) <= GATE_TOL)}
        if name != 'fp64':
            err['1_step'] = step_errors(tfm, name, psi0, chi0)
        err['transform'] = transform_error(tfm, p, name)
        ke = {'steps': marks, 'fp64': ke64, name: kx,
              'max_rel_diff': max(abs(a - b) / abs(b) for a, b in zip(kx, ke64) if b != 0.0)}
        if args.others:
            others = {}
            for nm in PRESETS:
                if nm in (name,):
                    continue
                if nm == 'fp64':
                    others[nm] = {'steps_per_s': fp64_rate, 'trajectory_err': {'psi': 0.0, 'chi': 0.0},
                                  'gate_pass': True}
                    continue
                ro, to = timed_run(tfm, nm, psi0, chi0, args.warmup, args.steps)
                del ro
                te = trajs[nm][0]
                others[nm] = {'steps_per_s': args.steps / to, 'trajectory_err': te,
                              'gate_pass': bool(max(te.values()) <= GATE_TOL), 'dtype': DTYPES[nm]}
                log(nm, others[nm])
        log('errors', err, 'gate', gate)
        if not gate['pass']:
            log('WARNING: %s misses the accuracy gate (trajectory error above %.0e)' % (name, GATE_TOL))

    transform = None
    if args.transform:
        log('transform-only workload: %d round trips per precision' % args.transform)
        transform = transform_bench(tfm, p, ['fp64', 'fp32', 'bf16x3', 'fp64m'], args.transform)

    # the probed region: as many steps again, continuing the run
    configure(tfm, name)
    tfm.sync()
    tfm.probe_start(64 * max(args.steps, 1))
    t0 = time.perf_counter()
    run.step(args.steps, check=False)
    tfm.sync()
    t1 = time.perf_counter()
    probes = tfm.probe_stop()
    run.step(0, check=True)
    probed_rate = args.steps / (t1 - t0)
    roof = roofline(probes, name)
    log('probed: %.3f ms/step; dominant kernel %s %.1f us/launch' % (1e3 * (t1 - t0) / args.steps, roof['kernel'],
                                                                     roof['avg_launch_us']))
    del run
    tfm.close()

    adj = None
    if args.adjoint:
        adj = adjoint_bench(args.adjoint, args.adjoint_dirs, args.adjoint_segment)

    hyp = None
    if args.hyper:
        hyp = hyper_bench(args.size, args.hyper)

    ens = None
    if args.ensemble:
        ens = ensemble_bench(args.ensemble, args.ensemble_steps, args.ensemble_check,
                             mode=args.ensemble_mode, streams=args.ensemble_streams)

    cpu = None
    if args.cpu_baseline:
        log('cpu baseline (oracle step)')
        cpu = cpu_baseline(p, one_core=bool(args.cpu_one_core))

    out = {
        'metric': METRIC, 'value': rate, 'unit': 'steps/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': ms, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': DTYPES[name],
        'data': 'synthetic: centred q=4 vortex + seeded 1e-4 perturbation (SplitMix64/Box-Muller, seed 20261202)',
        'config': {'workload': '3D vortex AB2/CN time step, NR=NP=NZ=%d, NRCHOP=%d, NPCHOP=NZCHOP=%d, L=1, '
                               'ZLEN=2pi, VISC=5e-4, dt=5e-4, SVV on' % (p.nr, p.nrchop, p.npchop),
                   'precision': name},
        'gate': gate,
        'fp64_steps_per_s': fp64_rate, 'speedup_vs_fp64': rate / fp64_rate if fp64_rate else None,
        'rel_err_vs_fp64': err, 'ke_history': ke, 'other_precisions': others,
        'adjoint_forward_ratio': adj['adjoint_forward_ratio'] if adj else None, 'adjoint_c5': adj,
        'probed_steps_per_s': probed_rate,
        'transform_c2': transform, 'ensemble_c4': ens, 'hyperviscous_p8': hyp,
        'kernels': probe_table(probes),
        'roofline': roof, 'cpu_baseline': cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
