"""Build in-tree: ``mlegs_amd/libmlegs_amd.so`` for gfx950 and the Fortran host.

Each ``csrc/*.hip`` is compiled to an object in ``mlegs_amd/_build`` (in
parallel), then linked into one shared library exposing the C ABI declared
in ``include/mlegs_amd.h``.  The Fortran module ``fortran/mlegs_amd.f90``
(iso_c_binding interface of that ABI), the vortex driver
``fortran/vortical_flow_3d.f90`` and the programs under ``fortran/apps``
are then compiled with amdflang into ``mlegs_amd/bin/<program>``, which find
the library through their rpath.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
OBJ = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libmlegs_amd.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FC = os.environ.get('FC_MLEGS', '/opt/rocm/bin/amdflang')
FSRC = os.path.join(HERE, 'fortran')
APPS = os.path.join(FSRC, 'apps')
BIN = os.path.join(HERE, 'bin')
ARCH = os.environ.get('MLEGS_AMD_ARCH', 'gfx950')

CFLAGS = [
    '--offload-arch=' + ARCH, '-O3', '-fPIC', '-std=c++17',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-result', '-I' + os.path.join(ROOT, 'include'),
]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + '.o')
    cmd = [HIPCC] + CFLAGS + ['-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s\n%s' % (src, ' '.join(cmd), r.stderr))
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '--offload-arch=' + ARCH, '-shared', '-fPIC', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n%s\n%s' % (' '.join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    build_fortran(verbose)
    return LIB


def build_fortran(verbose=True):
    """The Fortran host module, then every program (the vortex driver and the
    tutorial and 1D/2D programs under fortran/apps) linked against the
    library into mlegs_amd/bin/<program>."""
    os.makedirs(BIN, exist_ok=True)
    mod_obj = os.path.join(OBJ, 'mlegs_amd_f90.o')
    r = subprocess.run([FC, '-O2', '-c', os.path.join(FSRC, 'mlegs_amd.f90'), '-module-dir', OBJ, '-o', mod_obj],
                       capture_output=True, text=True, cwd=OBJ)
    if r.returncode != 0:
        raise RuntimeError('amdflang failed on mlegs_amd.f90:\n%s' % r.stderr)
    apps = [os.path.join(FSRC, 'vortical_flow_3d.f90')]
    apps += sorted(os.path.join(APPS, f) for f in os.listdir(APPS) if f.endswith('.f90'))

    def link(src):
        name = os.path.splitext(os.path.basename(src))[0]
        out = os.path.join(BIN, name)
        cmd = [FC, '-O2', src, mod_obj, '-I', OBJ, '-module-dir', os.path.join(OBJ, 'mod_' + name),
               '-L' + HERE, '-lmlegs_amd', '-Wl,-rpath,$ORIGIN/..', '-o', out]
        os.makedirs(os.path.join(OBJ, 'mod_' + name), exist_ok=True)
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=OBJ)
        if r.returncode != 0:
            raise RuntimeError('amdflang failed:\n%s\n%s' % (' '.join(cmd), r.stderr))
        return out

    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(apps))) as ex:
        outs = list(ex.map(link, apps))
    if verbose:
        for out in outs:
            print('built', out)


if __name__ == '__main__':
    build()
    sys.exit(0)
