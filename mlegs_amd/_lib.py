"""ctypes binding of ``libmlegs_amd.so`` (the C ABI in include/mlegs_amd.h).

The library is loaded from this package directory only; a missing or stale
library is an error (there is no CPU fallback).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libmlegs_amd.so')

MLX_OK = 0
MLX_E_LAYOUT = 201
MLX_E_ALLOC = 202
MLX_E_ARG = 203
MLX_E_FILE = 204
MLX_E_NONFINITE = 205
MLX_E_HIP = 206

PREC_FP64, PREC_FP32, PREC_BF16X3, PREC_FP64_MFMA = 0, 1, 2, 3
STAGE_SYNTH, STAGE_UP, STAGE_PROJ, STAGE_PROJ0 = 0, 1, 2, 3


class MlxParams(ctypes.Structure):
    _fields_ = [
        ('nr', ctypes.c_int32), ('np', ctypes.c_int32), ('nz', ctypes.c_int32),
        ('nrchop', ctypes.c_int32), ('npchop', ctypes.c_int32), ('nzchop', ctypes.c_int32),
        ('ell', ctypes.c_double), ('zlen', ctypes.c_double),
        ('dt', ctypes.c_double), ('ti', ctypes.c_double), ('totaltime', ctypes.c_double),
        ('ni', ctypes.c_int32), ('totaln', ctypes.c_int32),
        ('visc', ctypes.c_double),
        ('hyperpow', ctypes.c_int32),
        ('hypervisc', ctypes.c_double),
        ('is_svv', ctypes.c_int32),
        ('svv_cutoff', ctypes.c_double), ('svv_target', ctypes.c_double),
        ('svv_strength', ctypes.c_double), ('svv_relax', ctypes.c_double),
    ]


# exported symbols and their signatures (argtypes, restype)
_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int32
_I64 = ctypes.c_int64
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int32)
_S = ctypes.c_char_p

SIGNATURES = {
    'mlx_ctx_create': ([ctypes.POINTER(MlxParams), ctypes.c_int, _P, ctypes.POINTER(_P)], ctypes.c_int),
    'mlx_ctx_destroy': ([_P], ctypes.c_int),
    'mlx_last_error': ([_P], ctypes.c_char_p),
    'mlx_sync': ([_P], ctypes.c_int),
    'mlx_kit_init': ([_P], ctypes.c_int),
    'mlx_kit_dims': ([_P, _IP], ctypes.c_int),
    'mlx_kit_get': ([_P, _S, _DP, _I64], ctypes.c_int),
    'mlx_set_precision': ([_P, ctypes.c_int], ctypes.c_int),
    'mlx_set_stage_precision': ([_P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    'mlx_field_create': ([_P, ctypes.POINTER(_P)], ctypes.c_int),
    'mlx_field_destroy': ([_P, _P], ctypes.c_int),
    'mlx_field_copy': ([_P, _P, _P], ctypes.c_int),
    'mlx_field_upload': ([_P, _P, _DP, _I64, _I64, _I64, _S, _D, _IP], ctypes.c_int),
    'mlx_field_download': ([_P, _P, _DP, _I64, _I64, _I64, ctypes.c_char_p, _DP, _IP], ctypes.c_int),
    'mlx_field_set_meta': ([_P, _P, _S, _D, _IP], ctypes.c_int),
    'mlx_field_get_meta': ([_P, _P, ctypes.c_char_p, _DP, _IP], ctypes.c_int),
    'mlx_chop_offset': ([_P, _P, _I, _I, _I], ctypes.c_int),
    'mlx_field_axpby': ([_P, _P, _D, _P, _D], ctypes.c_int),
    'mlx_trans': ([_P, _P, _S], ctypes.c_int),
    'mlx_chop': ([_P, _P], ctypes.c_int),
    'mlx_dealias': ([_P, _P], ctypes.c_int),
    'mlx_svv_filter': ([_P, _P, _DP], ctypes.c_int),
    'mlx_svv_filter_noalloc': ([_P, _P, _DP], ctypes.c_int),
    'mlx_calcat0': ([_P, _P, _DP], ctypes.c_int),
    'mlx_calcat1': ([_P, _P, _DP], ctypes.c_int),
    'mlx_zeroat1': ([_P, _P], ctypes.c_int),
    'mlx_delsqp': ([_P, _P], ctypes.c_int),
    'mlx_idelsqp': ([_P, _P], ctypes.c_int),
    'mlx_xxdx': ([_P, _P], ctypes.c_int),
    'mlx_del2h': ([_P, _P], ctypes.c_int),
    'mlx_del2': ([_P, _P], ctypes.c_int),
    'mlx_idel2': ([_P, _P], ctypes.c_int),
    'mlx_fftreat': ([_P, _P], ctypes.c_int),
    'mlx_idel2_preln': ([_P, _P, _D], ctypes.c_int),
    'mlx_ihelm': ([_P, _P, _D], ctypes.c_int),
    'mlx_helmp': ([_P, _P, ctypes.c_int, _D, _D], ctypes.c_int),
    'mlx_ihelmp': ([_P, _P, ctypes.c_int, _D, _D], ctypes.c_int),
    'mlx_febe': ([_P, _P, _P, _D], ctypes.c_int),
    'mlx_fefe': ([_P, _P, _P, _D], ctypes.c_int),
    'mlx_abab': ([_P, _P, _P, _P, _P, _D, _I], ctypes.c_int),
    'mlx_abcn': ([_P, _P, _P, _P, _P, _D], ctypes.c_int),
    'mlx_vecprod': ([_P, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    'mlx_vec2tp': ([_P, _P, _P, _P, _P, _P], ctypes.c_int),
    'mlx_tp2vec': ([_P, _P, _P, _P, _P, _P], ctypes.c_int),
    'mlx_tp2curlvec': ([_P, _P, _P, _P, _P, _P], ctypes.c_int),
    'mlx_advection_rhs': ([_P, _P, _P, _P, _P, _P], ctypes.c_int),
    'mlx_check_stability': ([_P, _P, _P], ctypes.c_int),
    'mlx_vortex_create': ([_P, _P, _P, _P, ctypes.POINTER(_P)], ctypes.c_int),
    'mlx_trans_adjoint': ([_P, _P, ctypes.c_char_p], ctypes.c_int),
    'mlx_vecprod_adjoint': ([_P] + [_P] * 12, ctypes.c_int),
    'mlx_abcn_adjoint': ([_P, _P, _P, _P, _D], ctypes.c_int),
    'mlx_adjoint_project': ([_P, _P], ctypes.c_int),
    'mlx_febe_adjoint': ([_P, _P, _P, _D], ctypes.c_int),
    'mlx_zeroat1_adjoint': ([_P, _P], ctypes.c_int),
    'mlx_tp2vec_adjoint': ([_P, _P, _P, _P, _P, _P], ctypes.c_int),
    'mlx_tp2curlvec_adjoint': ([_P, _P, _P, _P, _P, _P], ctypes.c_int),
    'mlx_vec2tp_adjoint': ([_P, _P, _P, _P, _P, _P], ctypes.c_int),
    'mlx_advection_adjoint': ([_P, _P, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    'mlx_energy_grad': ([_P, _P, _P, _P], ctypes.c_int),
    'mlx_vortex_reverse_step': ([_P] * 10 + [_D, _P, _I], ctypes.c_int),
    'mlx_vortex_step_record': ([_P, _P, _I, _D, _P, _P, _P], ctypes.c_int),
    'mlx_ensemble_create': ([ctypes.c_void_p, _I, _I, _I, ctypes.POINTER(_P)], ctypes.c_int),
    'mlx_ensemble_destroy': ([_P], ctypes.c_int),
    'mlx_ensemble_last_error': ([_P], ctypes.c_char_p),
    'mlx_ensemble_ctx': ([_P, _I, ctypes.POINTER(_P)], ctypes.c_int),
    'mlx_ensemble_set': ([_P, _I, _P, _P, _P], ctypes.c_int),
    'mlx_ensemble_bootstrap': ([_P], ctypes.c_int),
    'mlx_ensemble_step': ([_P, _I, _I, _I, _I], ctypes.c_int),
    'mlx_ensemble_sync': ([_P], ctypes.c_int),
    'mlx_ensemble_vortex': ([_P, _I, ctypes.POINTER(_P)], ctypes.c_int),
    'mlx_ensemble_check': ([_P, ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
    'mlx_ensemble_info': ([_P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    'mlx_vortex_destroy': ([_P, _P], ctypes.c_int),
    'mlx_vortex_bootstrap': ([_P, _P], ctypes.c_int),
    'mlx_vortex_step': ([_P, _P, _I, _D, _I], ctypes.c_int),
    'mlx_vortex_field': ([_P, _P, _I, ctypes.POINTER(_P)], ctypes.c_int),
    'mlx_vortex_gains': ([_P, _P, _DP, _DP], ctypes.c_int),
    'mlx_energy': ([_P, _P, _P, _P, _DP], ctypes.c_int),
    'mlx_vorticity_magnitude': ([_P, _P, _P, _P], ctypes.c_int),
    'mlx_field_save': ([_P, _P, _S, _I, _I], ctypes.c_int),
    'mlx_field_load': ([_P, _P, _S, _I, _I], ctypes.c_int),
    'mlx_file_write': ([_S, _DP, _I64, _I64, _I64, _S, _D, _IP, _I, _I], ctypes.c_int),
    'mlx_file_read': ([_S, _DP, _I64, _I64, _I64, ctypes.c_char_p, _DP, _IP, _I, _I], ctypes.c_int),
    'mlx_file_last_error': ([], ctypes.c_char_p),
    'mlx_time_kernel': ([_P, _P, _S, _I, _DP], ctypes.c_int),
    'mlx_probe_start': ([_P, _I], ctypes.c_int),
    'mlx_seed_spectral': ([_DP, _I64, _I64, _I64, _IP, ctypes.c_uint64, _I, ctypes.c_double, ctypes.c_double, _I],
                          ctypes.c_int),
    'mlx_probe_stop': ([_P, _DP], ctypes.c_int),
}

_lib = None


class MlxError(RuntimeError):
    """A non-zero status from the library; ``code`` is the MLX_E_* value."""

    def __init__(self, code, msg):
        super().__init__('%s (code %d)' % (msg, code))
        self.code = code


def load():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError('mlegs_amd: %s is missing; build it with `python -m mlegs_amd.build` '
                          '(there is no CPU fallback)' % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(ctx, rc):
    if rc != MLX_OK:
        lib = load()
        msg = lib.mlx_last_error(ctx)
        raise MlxError(rc, msg.decode() if msg else 'error')
