#!/usr/bin/env python
"""Benchmark of the spectral-transform time step on one MI355X.

Workload (one "step" = one AB2/CN time step of the 3D vortex driver,
vortical_flow_3d.f90:160-181): NR = NP = NZ = 256, NRCHOP = 256,
NPCHOP = NZCHOP = 129, L = 1, ZLEN = 2 pi, VISC = 5e-4, dt = 5e-4; initial
state = centred q = 4 vortex + seeded 1e-4 perturbation (synthetic).  The
bootstrap (FE-BE + Richardson) runs untimed, then --warmup steps, then
--steps timed steps bracketed by device synchronisation.

Printed: one JSON line with the steps/s of the selected precision (--prec,
default fp32: Legendre contractions on FP32 MFMA, everything else FP64), the
same-GPU FP64 plain path's steps/s, the spectral error of the mixed path
against FP64 after the same steps, the roofline of the dominant kernel
(timed live with HIP events on the library's stream) and a CPU baseline (the
oracle on the host, bounded sample).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'steps/s @256³ mixed-prec vs FP64 at ≤1e-6 rel.err; adjoint/forward time ratio'
PEAK = {'fp64': 78.6e12, 'fp32': 157.3e12, 'bf16x3': 2.5e15}
HBM = 8.0e12


def workload(n=256):
    import mlegs_amd as M
    return M.Params(nr=n, np=n, nz=n, nrchop=n, npchop=n // 2 + 1, nzchop=n // 2 + 1, ell=1.0,
                    zlen=2 * math.pi, visc=5e-4, dt=5e-4, hyperpow=0, is_svv=True)


def initial_state(tfm, p, amp=1e-4, q=4.0):
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


def log(*a):
    print('[bench]', *a, file=sys.stderr, flush=True)


def run_steps(tfm, p, psi0, chi0, warmup, steps):
    import mlegs_amd as M
    run = M.VortexRun(tfm, psi0, chi0, M.Scalar(tfm))
    run.bootstrap()
    tfm.sync()
    if warmup:
        run.step(warmup, check=True)
    tfm.sync()
    t0 = time.perf_counter()
    run.step(steps, check=False)
    tfm.sync()
    t1 = time.perf_counter()
    run.step(0, check=True)  # report a non-finite state, if any
    return run, (t1 - t0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--prec', default='fp64', choices=['fp64', 'fp32', 'bf16x3'])
    ap.add_argument('--size', type=int, default=256)
    ap.add_argument('--fp64-steps', type=int, default=5, help='timed steps of the FP64 reference run')
    ap.add_argument('--kernel-iters', type=int, default=20)
    ap.add_argument('--cpu-baseline', type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py: this build runs on one GPU; --gpus %d is not supported (multi-GPU sharding is a '
              'design note, not part of this build)' % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import mlegs_amd as M

    p = workload(args.size)
    precs = {'fp64': M.PREC_FP64, 'fp32': M.PREC_FP32, 'bf16x3': M.PREC_BF16X3}
    log('kit', p.nr, p.np, p.nz)
    tfm = M.Kit(p, device=0, precision=M.PREC_FP64)
    psi0, chi0 = initial_state(tfm, p)

    # FP64 plain path (reference numerics)
    fp64_rate = None
    run64 = None
    if args.prec != 'fp64' or True:
        tfm.set_precision(M.PREC_FP64)
        k64 = args.fp64_steps if args.prec != 'fp64' else args.steps
        w64 = 1 if args.prec != 'fp64' else args.warmup
        log('fp64 run', w64, k64)
        run64, t64 = run_steps(tfm, p, psi0, chi0, w64, k64)
        fp64_rate = k64 / t64
        ms64 = 1e3 * t64 / k64
    # selected precision
    if args.prec == 'fp64':
        rate, ms, run = fp64_rate, ms64, run64
        nsteps_total = args.warmup + args.steps
        err = {'psi': 0.0, 'chi': 0.0}
    else:
        tfm.set_precision(precs[args.prec])
        log(args.prec, 'run', args.warmup, args.steps)
        run, t = run_steps(tfm, p, psi0, chi0, args.warmup, args.steps)
        rate = args.steps / t
        ms = 1e3 * t / args.steps
        # error vs FP64 after the same number of steps
        tfm.set_precision(M.PREC_FP64)
        ref, _ = run_steps(tfm, p, psi0, chi0, args.warmup, args.steps)
        err = {}
        for nm in ('psi', 'chi'):
            a = run.field(nm).e
            b = ref.field(nm).e
            err[nm] = float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
        tfm.set_precision(precs[args.prec])

    # dominant kernel: the forward Legendre contraction, timed with HIP events
    nr, nz = p.nr, p.nz
    S = sum(max(p.nrchop - m, 0) for m in range(p.npchop))
    f_leg = 4.0 * (nr // 2) * nz * S
    probe = M.Scalar(tfm).upload(np.zeros(tfm.dims, dtype=np.complex128), space='PFP')
    kern = {}
    for w in ('leg_fwd', 'leg_bwd', 'fft_phi_fwd', 'fft_phi_bwd', 'fft_z_fwd', 'fft_z_bwd'):
        kern[w] = tfm.time_kernel(probe, w, args.kernel_iters)
    t_leg = kern['leg_fwd'] * 1e-3
    achieved = f_leg / t_leg
    peak = PEAK[args.prec]
    roofline = {'bound': 'mfma', 'kernel': 'leg_fwd (%s)' % args.prec, 'achieved': achieved / 1e12,
                'peak': peak / 1e12, 'unit': 'TFLOP/s', 'frac': achieved / peak, 'traffic': None,
                'flops_per_launch': f_leg, 'ms_per_launch': kern['leg_fwd']}
    nb = 16.0 * p.nr * (p.np // 2) * p.nz * 2
    kernels_ms = {k: round(v, 4) for k, v in kern.items()}
    fft_bw = {k: nb / (v * 1e-3) / 1e9 for k, v in kern.items() if k.startswith('fft')}

    cpu = None
    if args.cpu_baseline:
        cpu = cpu_baseline(p)

    out = {
        'metric': METRIC, 'value': rate, 'unit': 'steps/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': ms, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': {'fp64': 'f64', 'fp32': 'f32-mfma+f64', 'bf16x3': 'bf16x3-mfma+f64'}[args.prec],
        'data': 'synthetic: centred q=4 vortex + seeded 1e-4 perturbation (SplitMix64/Box-Muller, seed 20261202)',
        'config': {'workload': '3D vortex AB2/CN time step, NR=NP=NZ=%d, NRCHOP=%d, NPCHOP=NZCHOP=%d, L=1, '
                               'VISC=5e-4, dt=5e-4, SVV on' % (p.nr, p.nrchop, p.npchop),
                   'precision': args.prec},
        'fp64_steps_per_s': fp64_rate, 'speedup_vs_fp64': (rate / fp64_rate) if fp64_rate else None,
        'rel_err_vs_fp64': err, 'kernels_ms': kernels_ms,
        'fft_GBps': {k: round(v, 1) for k, v in fft_bw.items()},
        'roofline': roofline, 'cpu_baseline': cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(p):
    """The oracle on the host: one transform round trip of one field at the
    workload's size (the bounded sample), scaled to steps/s by the number of
    transform-equivalents per step (13 Legendre-sized contractions)."""
    import numpy as np
    from oracle import kit as K, scalar_ops as so
    from mlegs_amd import inputs
    t0 = time.perf_counter()
    okit = K.Kit3D(p)
    tk = time.perf_counter() - t0
    e = inputs.transform_field((okit.nrdim, okit.npdim, okit.nzdim), (p.nrchop, p.npchop, p.nzchop, p.np, p.nz))
    s = so.Scalar(okit, 'FFF')
    s.e = e
    t0 = time.perf_counter()
    so.trans(s, 'PPP', okit)
    so.trans(s, 'FFF', okit)
    t = time.perf_counter() - t0
    per_step = t * 13.0 / 2.0
    return {'value': 1.0 / per_step, 'unit': 'steps/s', 'cores': os.cpu_count(), 'kind': 'port',
            'sample': 'oracle (numpy/scipy FP64) one backward+forward transform of one 256^3 field (%.2f s), '
                      'scaled by 6.5 round-trip equivalents per step; kit build %.1f s' % (t, tk)}


if __name__ == '__main__':
    main()
