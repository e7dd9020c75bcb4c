"""ctypes binding of libhip_sys.so, the C ABI declared in include/hip_sys.h.

The library is built in-tree by `python -m rustymath_amd.build`. Loading fails loudly where it is missing: there
is no CPU path behind this package.
"""
import ctypes, os

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, 'libhip_sys.so')

NO_TRANS, TRANS, CONJ_TRANS = 111, 112, 113
UPPER, LOWER, NON_UNIT, UNIT, LEFT, RIGHT = 121, 122, 131, 132, 141, 142

_i32, _i64, _vp = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p
_d, _f = ctypes.c_double, ctypes.c_float

# name: (restype, argtypes), mirroring include/hip_sys.h
SIGNATURES = {
    'hs_init': (_i32, []),
    'hs_finalize': (_i32, []),
    'hs_set_exchange': (_i32, [_i32, _vp]),
    'hs_last_error': (ctypes.c_char_p, []),
    'hs_dgemm': (_i32, [_i32, _i32, _i64, _i64, _i64, _d, _vp, _i64, _vp, _i64, _d, _vp, _i64]),
    'hs_sgemm': (_i32, [_i32, _i32, _i64, _i64, _i64, _f, _vp, _i64, _vp, _i64, _f, _vp, _i64]),
    'hs_dgemm_batched': (_i32, [_i32, _i32, _i64, _i64, _i64, _d, _vp, _i64, _i64, _vp, _i64, _i64, _d, _vp, _i64, _i64,
                                _i64]),
    'hs_sgemm_batched': (_i32, [_i32, _i32, _i64, _i64, _i64, _f, _vp, _i64, _i64, _vp, _i64, _i64, _f, _vp, _i64, _i64,
                                _i64]),
    'hs_dev_dgemm_batched': (_i32, [_i32, _i32, _i64, _i64, _i64, _d, _vp, _i64, _i64, _vp, _i64, _i64, _d, _vp, _i64,
                                    _i64, _i64, _vp]),
    'hs_dev_sgemm_batched': (_i32, [_i32, _i32, _i64, _i64, _i64, _f, _vp, _i64, _i64, _vp, _i64, _i64, _f, _vp, _i64,
                                    _i64, _i64, _vp]),
    'hs_dtrsm': (_i32, [_i32, _i32, _i32, _i32, _i64, _i64, _vp, _i64, _vp, _i64]),
    'hs_strsm': (_i32, [_i32, _i32, _i32, _i32, _i64, _i64, _vp, _i64, _vp, _i64]),
    'hs_dev_dtrsm': (_i32, [_i32, _i32, _i32, _i32, _i64, _i64, _vp, _i64, _vp, _i64, _vp]),
    'hs_dev_strsm': (_i32, [_i32, _i32, _i32, _i32, _i64, _i64, _vp, _i64, _vp, _i64, _vp]),
    'hs_dlu_factor': (_i32, [_vp, _i64, _vp, _vp, _vp, _vp]),
    'hs_slu_factor': (_i32, [_vp, _i64, _vp, _vp, _vp, _vp]),
    'hs_dlu_solve': (_i32, [_vp, _vp, _i64, _vp, _i64, _vp]),
    'hs_slu_solve': (_i32, [_vp, _vp, _i64, _vp, _i64, _vp]),
    'hs_dlu_inverse': (_i32, [_vp, _vp, _i64, _vp]),
    'hs_slu_inverse': (_i32, [_vp, _vp, _i64, _vp]),
    'hs_dinverse': (_i32, [_vp, _i64, _vp, _vp, _vp]),
    'hs_sinverse': (_i32, [_vp, _i64, _vp, _vp, _vp]),
    'hs_dcholesky': (_i32, [_vp, _i64, _vp, _vp]),
    'hs_scholesky': (_i32, [_vp, _i64, _vp, _vp]),
    'hs_dev_dpotrf': (_i32, [_i64, _vp, _i64, _vp, _vp]),
    'hs_dev_spotrf': (_i32, [_i64, _vp, _i64, _vp, _vp]),
    'hs_dsolve_stack': (_i32, [_i64, _i64, _i64, _vp, _vp, _vp, _vp]),
    'hs_ssolve_stack': (_i32, [_i64, _i64, _i64, _vp, _vp, _vp, _vp]),
    'hs_dlu_stack': (_i32, [_i64, _i64, _vp, _vp, _vp]),
    'hs_slu_stack': (_i32, [_i64, _i64, _vp, _vp, _vp]),
    'hs_dgesv_batched': (_i32, [_i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    'hs_sgesv_batched': (_i32, [_i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    'hs_dlu_solve_stack': (_i32, [_i64, _i64, _i64, _vp, _vp, _vp, _vp]),
    'hs_slu_solve_stack': (_i32, [_i64, _i64, _i64, _vp, _vp, _vp, _vp]),
    'hs_dsolve': (_i32, [_vp, _i64, _vp, _i64, _vp, _vp]),
    'hs_ssolve': (_i32, [_vp, _i64, _vp, _i64, _vp, _vp]),
    'hs_dev_dgemm': (_i32, [_i32, _i32, _i64, _i64, _i64, _d, _vp, _i64, _vp, _i64, _d, _vp, _i64, _vp]),
    'hs_dev_sgemm': (_i32, [_i32, _i32, _i64, _i64, _i64, _f, _vp, _i64, _vp, _i64, _f, _vp, _i64, _vp]),
    'hs_dev_dgetrf': (_i32, [_i64, _vp, _i64, _vp, _vp, _vp]),
    'hs_dev_sgetrf': (_i32, [_i64, _vp, _i64, _vp, _vp, _vp]),
    'hs_dev_dgesv': (_i32, [_i64, _vp, _i64, _vp, _vp, _i64, _i64, _vp, _vp]),
    'hs_dev_sgesv': (_i32, [_i64, _vp, _i64, _vp, _vp, _i64, _i64, _vp, _vp]),
    'hs_dev_dgetrs': (_i32, [_i32, _i64, _i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    'hs_dev_sgetrs': (_i32, [_i32, _i64, _i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    'hs_dev_dgetrs_noalloc': (_i32, [_i32, _i64, _i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    'hs_dev_dgetri': (_i32, [_i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    'hs_dev_sgetri': (_i32, [_i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    'hs_dev_dsgesv': (_i32, [_i64, _vp, _i64, _vp, _vp, _i32, _d, _vp, _vp, _vp]),
    'hs_dev_sgemm_bf16_update': (_i32, [_i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _i32, _vp]),
    'hs_dev_dbgesv': (_i32, [_i64, _vp, _i64, _vp, _vp, _i32, _i32, _i32, _d, _vp, _vp, _vp, _vp]),
    'hs_dev_dgesv_batched_plan': (_i32, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i32, _i32, _vp]),
    'hs_dev_sgesv_batched_plan': (_i32, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i32, _i32, _vp]),
    'hs_dev_dsgesv_batched_plan': (_i32, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                          _i32, _d, _i32, _i32, _vp]),
    'hs_dev_gesv_batched_run': (_i32, [_i64, _i32, _vp]),
    'hs_dev_gesv_batched_free': (_i32, [_i64]),
    'hs_profile_panel': (_i32, [_i32, _vp]),
    'hs_profile_gemm': (_i32, [_i32, _vp, _vp, _vp, _vp]),
    'hs_profile_lu': (_i32, [_i32, _vp]),
    'hs_profile_batched': (_i32, [_i32, _vp]),
}

_lib = None


class HipSysError(RuntimeError):
    "A failure status from the HIP library (a HIP runtime error or a device-side wait that ran out)."


def load():
    "The loaded library, with argument types set; raises where it has not been built."
    global _lib
    if _lib is None:
        # torch ships its own libamdhip64.so.7; with one soname the first one loaded serves the whole process, and
        # torch's device code needs its own. Loading torch first makes this library share that runtime (and the
        # same device pointers and streams).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(PATH):
            raise RuntimeError(f'{PATH} is not built: run `python -m rustymath_amd.build` (hipcc, gfx950)')
        lib = ctypes.CDLL(PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args
        _lib = lib
    return _lib


UNSUPPORTED = 100001  # valid arguments beyond what the library handles (include/hip_sys.h)


def check(status, what):
    """Turns a status into Python's errors: ValueError for a bad argument, NotImplementedError for valid arguments the
    library does not handle, HipSysError for a device failure."""
    if status == 0: return
    if status < 0: raise ValueError(f'{what}: argument {-status} is invalid')
    text = load().hs_last_error().decode(errors="replace")
    if status == UNSUPPORTED: raise NotImplementedError(f'{what}: {text}')
    raise HipSysError(f'{what}: {text} (status {status})')
