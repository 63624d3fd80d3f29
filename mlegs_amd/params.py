"""Run parameters and the ``input.params`` reader.

Mirrors the module variables of the reference's ``mlegs_base``
(src/modules/mlegs_base.f90:7-53) with their defaults, the fixed-line
parser ``read_input`` (src/submodules/mlegs_misc_io.f90:307-388) and
``timestep_set`` (:390-408).  Errors raise ``InputError`` where the
reference calls ``error stop`` (non-zero exit).
"""
import dataclasses
import math
import os
import re


class InputError(ValueError):
    """An ``input.params`` file the reference would reject with ``error stop``."""


@dataclasses.dataclass
class Params:
    """Global run parameters (mlegs_base.f90:7-53)."""
    nr: int = 256
    nrchop: int = 256
    np: int = 128
    npchop: int = 65
    nz: int = 128
    nzchop: int = 65
    ell: float = 4.0
    zlen: float = 2.0 * math.pi
    dt: float = 1.0e-3
    ti: float = 0.0
    totaltime: float = 1.0e2
    ni: int = 0
    totaln: int = 100000
    visc: float = 1.0e-3
    hyperpow: int = 0
    hypervisc: float = 0.0
    is_svv: bool = True
    svv_cutoff: float = 0.75
    svv_target: float = 2.0e-2
    svv_strength: float = 0.12
    svv_relax: float = 0.25
    flddir: str = './output/fld/'
    isfldsav: bool = True
    fldsavintvl: int = 100
    logdir: str = './output/log/'
    islogsav: bool = True
    logsavintvl: int = 1

    def replace(self, **kw):
        return dataclasses.replace(self, **kw)


_DIR_CHARS = re.compile(r'^[A-Za-z0-9_./\-]*$')


def _fortran_real(tok):
    t = tok.strip().rstrip(',')
    return float(t.replace('D', 'E').replace('d', 'e'))


def _fortran_int(tok):
    return int(tok.strip().rstrip(','))


def _fortran_logical(tok):
    t = tok.strip().rstrip(',').upper().lstrip('.')
    if t.startswith('T'):
        return True
    if t.startswith('F'):
        return False
    raise InputError('read_input: invalid logical value %r' % tok)


def _values(line, n):
    toks = line.replace(',', ' ').split()
    if len(toks) < n:
        raise InputError('read_input: expected %d values in %r' % (n, line.rstrip()))
    return toks[:n]


def read_input(fn, params=None, make_dirs=True):
    """Parse a fixed-line ``input.params`` (mlegs_misc_io.f90:307-388).

    Line order: a header line, then (label, values) pairs for NR/NP/NZ,
    chops, ELL/ZLEN/ZLENxPI, a separator, time stepping, viscosity, field
    and log output, and an optional SVV block.  Output directories are
    validated and created with ``mkdir -p`` semantics as the reference does.
    """
    p = params if params is not None else Params()
    try:
        with open(fn, 'r') as f:
            lines = f.read().splitlines()
    except OSError:
        raise InputError('read_input: could not open input file')
    it = iter(lines)

    def nxt():
        try:
            return next(it)
        except StopIteration:
            raise InputError('read_input: unexpected end of input file')

    try:
        nxt()
        nxt(); a = _values(nxt(), 3); p.nr, p.np, p.nz = (_fortran_int(v) for v in a)
        nxt(); a = _values(nxt(), 3); p.nrchop, p.npchop, p.nzchop = (_fortran_int(v) for v in a)
        nxt(); a = _values(nxt(), 3)
        p.ell, p.zlen = _fortran_real(a[0]), _fortran_real(a[1])
        if _fortran_logical(a[2]):
            p.zlen = p.zlen * math.pi
        nxt()
        nxt()
        nxt(); a = _values(nxt(), 5)
        p.dt, p.ti, p.totaltime = (_fortran_real(v) for v in a[:3])
        p.ni, p.totaln = _fortran_int(a[3]), _fortran_int(a[4])
        nxt()
        nxt()
        nxt(); a = _values(nxt(), 3)
        p.visc, p.hyperpow, p.hypervisc = _fortran_real(a[0]), _fortran_int(a[1]), _fortran_real(a[2])
        nxt()
        nxt()
        nxt(); d = nxt().strip()
        d = d if d.endswith('/') else d + '/'
        if not _DIR_CHARS.match(d):
            raise InputError('read_input: invalid field output directory')
        p.flddir = d
        if make_dirs:
            try:
                os.makedirs(d, exist_ok=True)
            except OSError:
                raise InputError('read_input: could not create field output directory')
        nxt(); a = _values(nxt(), 2); p.isfldsav, p.fldsavintvl = _fortran_logical(a[0]), _fortran_int(a[1])
        nxt()
        nxt()
        nxt(); d = nxt().strip()
        d = d if d.endswith('/') else d + '/'
        if not _DIR_CHARS.match(d):
            raise InputError('read_input: invalid log output directory')
        p.logdir = d
        if make_dirs:
            try:
                os.makedirs(d, exist_ok=True)
            except OSError:
                raise InputError('read_input: could not create log output directory')
        nxt(); a = _values(nxt(), 2); p.islogsav, p.logsavintvl = _fortran_logical(a[0]), _fortran_int(a[1])
    except (ValueError, TypeError) as exc:
        if isinstance(exc, InputError):
            raise
        raise InputError('read_input: malformed input file (%s)' % exc)

    # optional SVV block: a header line, a label line and a value line
    rest = list(it)
    if len(rest) >= 2:
        try:
            if len(rest) < 3:
                raise InputError('read_input: invalid SVV controls')
            a = _values(rest[2], 5)
            p.is_svv = _fortran_logical(a[0])
            p.svv_cutoff, p.svv_target, p.svv_strength, p.svv_relax = (_fortran_real(v) for v in a[1:])
        except (ValueError, InputError):
            raise InputError('read_input: invalid SVV controls')

    if p.nr <= 0 or p.nr % 2 != 0 or p.np <= 0 or p.nz <= 0:
        raise InputError('read_input: nr must be positive and even; np and nz must be positive')
    if (p.nrchop <= 0 or p.nrchop > p.nr or p.npchop <= 0 or p.npchop > p.np // 2 + 1
            or p.nzchop <= 0 or p.nzchop > p.nz // 2 + 1):
        raise InputError('read_input: invalid chopping sizes')
    if p.ell <= 0.0 or p.zlen <= 0.0:
        raise InputError('read_input: ell and zlen must be positive')
    if p.dt <= 0.0 or p.totaltime < 0.0 or p.totaln < 0:
        raise InputError('read_input: invalid time stepping values')
    if p.visc < 0.0 or p.hypervisc < 0.0:
        raise InputError('read_input: viscosities must be non-negative')
    if p.hyperpow not in (0, 4, 6, 8):
        raise InputError('read_input: hyperpow must be zero, 4, 6, or 8')
    if p.hyperpow != 0 and p.hypervisc <= 0.0:
        raise InputError('read_input: hypervisc must be positive when enabled')
    if p.fldsavintvl <= 0 or p.logsavintvl <= 0:
        raise InputError('read_input: output intervals must be positive')
    if (p.svv_cutoff <= 0.0 or p.svv_cutoff >= 1.0 or p.svv_target <= 0.0
            or p.svv_strength < 0.0 or p.svv_relax < 0.0 or p.svv_relax > 1.0):
        raise InputError('read_input: invalid SVV controls')
    return p


def timestep_set(p):
    """Pick the longer of TOTT and TOTN*DT as the end criterion (mlegs_misc_io.f90:390-408).

    Returns (curr_n, curr_t) = (ni, ti).
    """
    if p.dt <= 0.0:
        raise InputError('timestep_setup: dt must be positive')
    tott, totn = p.totaltime, p.totaln
    p.totaltime = max(tott, p.dt * totn)
    if p.totaltime == tott:
        p.totaln = int(math.floor(tott / p.dt))
    else:
        p.totaln = totn
    return p.ni, p.ti
