"""ctypes binding of libmacromax_hip.so (the C ABI declared in include/macromax_hip.h).

The library is the only compute path of this package: if it is missing or cannot be loaded, every entry point
raises. There is no CPU or pure-PyTorch substitute.
"""
from __future__ import annotations

import ctypes
import os
import threading

__all__ = ['lib', 'LIB_PATH', 'check', 'NativeError', 'MMX_C64', 'MMX_C128']

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'libmacromax_hip.so')

MMX_C64 = 0
MMX_C128 = 1

EXPORTED = ('mmx_version', 'mmx_last_error', 'mmx_register_fft_length', 'mmx_plan_create', 'mmx_plan_destroy', 'mmx_plan_set_param',
            'mmx_plan_get_param', 'mmx_plan_set_source_support', 'mmx_fft', 'mmx_born_step', 'mmx_born_step_noalloc',
            'mmx_born_prologue', 'mmx_magnetic_iterate', 'mmx_magnetic_iterate_split',
            'mmx_born_iterate', 'mmx_born_iterate_split',
            'mmx_batch_step', 'mmx_wall_clock', 'mmx_wall_clock_rate_khz', 'mmx_setup_sigma', 'mmx_setup_gain',
            'mmx_norm2', 'mmx_chi_mul',
            'mmx_longitudinal_projection_ft')

# batched-solve state bits and per-problem track record (include/macromax_hip.h, MMX_STATE_* / MMX_TRACK_*)
STATE_PARITY, STATE_DONE, STATE_STALE = 1, 2, 4
# size of the static device array that holds the norm partials of mmx_born_step_noalloc (MMX_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 4 << 20
TRACK = {'prev': 0, 'field_norm': 1, 'residue': 2, 'iter': 3, 'alpha_r': 4, 'alpha_i': 5, 'a': 6, 'c_r': 7,
         'c_i': 8, 'done_clock': 9, 'rejected': 10}
TRACK_LEN = 12

PARAMS = {'g_axis': 1, 'p_rows': 2, 'p_strided': 3, 'p_green': 4, 'xcd_remap': 5, 'lds_budget': 6,
          'big_elems': 7, 'reg_fft': 8, 'transposed': 9, 'scratch_gen': 10, 'radix_order': 11}


class NativeError(RuntimeError):
    """An error reported by libmacromax_hip.so."""


_lib = None
_lock = threading.Lock()


def _declare(L):
    c_int, c_i64, c_p = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p
    L.mmx_version.restype = c_int
    L.mmx_version.argtypes = []
    L.mmx_last_error.restype = ctypes.c_char_p
    L.mmx_last_error.argtypes = []
    L.mmx_register_fft_length.restype = c_int
    L.mmx_register_fft_length.argtypes = [c_i64]
    L.mmx_plan_create.restype = c_int
    L.mmx_plan_create.argtypes = [ctypes.POINTER(c_p), c_int, ctypes.POINTER(c_i64), ctypes.POINTER(ctypes.c_double), c_int]
    L.mmx_plan_destroy.restype = c_int
    L.mmx_plan_destroy.argtypes = [c_p]
    L.mmx_plan_set_param.restype = c_int
    L.mmx_plan_set_param.argtypes = [c_p, c_int, c_i64]
    L.mmx_plan_set_source_support.restype = c_int
    L.mmx_plan_set_source_support.argtypes = [c_p, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)]
    L.mmx_plan_get_param.restype = c_int
    L.mmx_plan_get_param.argtypes = [c_p, c_int, ctypes.POINTER(c_i64)]
    L.mmx_fft.restype = c_int
    L.mmx_fft.argtypes = [c_p, c_p, c_p, c_i64, c_int, c_p]
    L.mmx_born_step.restype = c_int
    L.mmx_born_step.argtypes = [c_p, c_p, c_p, c_p, c_p, c_int, c_int, c_p, c_i64, c_i64, c_i64, c_p, c_p]
    L.mmx_born_step_noalloc.restype = c_int
    L.mmx_born_step_noalloc.argtypes = [c_p, c_p, c_p, c_p, c_p, c_int, c_int, c_p, c_i64, c_i64, c_i64, c_p, c_p]
    L.mmx_magnetic_iterate.restype = c_int
    L.mmx_magnetic_iterate.argtypes = [c_p, c_p, c_p, c_p, ctypes.POINTER(c_p), ctypes.POINTER(c_int),
                                       ctypes.POINTER(c_i64), ctypes.POINTER(ctypes.c_double), c_p, c_p, c_p]
    L.mmx_magnetic_iterate_split.restype = c_int
    L.mmx_magnetic_iterate_split.argtypes = [c_p, c_p, c_p, c_p, ctypes.POINTER(c_p), ctypes.POINTER(c_int),
                                             ctypes.POINTER(c_i64), ctypes.POINTER(ctypes.c_double), c_p, c_p, c_p,
                                             c_int, c_p, c_p]
    L.mmx_born_prologue.restype = c_int
    L.mmx_born_prologue.argtypes = [c_p, c_p, c_p, c_p, c_p, c_int, c_int, c_p, c_i64, c_i64, c_i64, c_p]
    L.mmx_born_iterate.restype = c_int
    L.mmx_born_iterate.argtypes = [c_p, c_p, c_p, c_p, c_p, c_p, c_int, c_int, c_p, c_i64, c_i64, c_i64, c_p, c_p]
    L.mmx_born_iterate_split.restype = c_int
    L.mmx_born_iterate_split.argtypes = [c_p, c_p, c_p, c_p, c_p, c_p, c_int, c_int, c_p, c_i64, c_i64, c_i64, c_p,
                                         c_p, c_int, c_p, c_p]
    L.mmx_batch_step.restype = c_int
    L.mmx_batch_step.argtypes = [c_p, c_p, c_p, c_p, c_p, c_p, c_int, c_int, c_p, c_i64, c_i64, c_i64, c_p, c_p, c_p,
                                 ctypes.c_double, c_i64, c_p]
    L.mmx_wall_clock.restype = c_int
    L.mmx_wall_clock.argtypes = [c_p, c_p]
    L.mmx_wall_clock_rate_khz.restype = c_int
    L.mmx_wall_clock_rate_khz.argtypes = [ctypes.POINTER(c_i64)]
    L.mmx_setup_sigma.restype = c_int
    L.mmx_setup_sigma.argtypes = [c_p, c_p, c_int, ctypes.c_double, c_p, c_p]
    L.mmx_setup_gain.restype = c_int
    L.mmx_setup_gain.argtypes = [c_p, c_p, c_int, c_p, c_p]
    L.mmx_norm2.restype = c_int
    L.mmx_norm2.argtypes = [c_p, c_p, c_i64, c_p, c_p]
    L.mmx_chi_mul.restype = c_int
    L.mmx_chi_mul.argtypes = [c_p, c_p, c_int, c_p, c_p, c_int, c_i64, c_p]
    L.mmx_longitudinal_projection_ft.restype = c_int
    L.mmx_longitudinal_projection_ft.argtypes = [c_p, c_p, c_p, c_p]
    return L


def lib():
    """The loaded library (loaded once). Raises if it has not been built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(f'{LIB_PATH} not found: build it with `python -m macromax_amd.build`.')
                _lib = _declare(ctypes.CDLL(LIB_PATH))
    return _lib


def wall_clock_rate_khz() -> int:
    """Rate of the device wall clock that mmx_wall_clock and the batch controller read."""
    v = ctypes.c_int64()
    check(lib().mmx_wall_clock_rate_khz(ctypes.byref(v)), 'mmx_wall_clock_rate_khz')
    return int(v.value)


def register_fft_length(n: int) -> bool:
    """Whether complex64 axes of n points have register-resident FFT kernels (mmx_register_fft_length)."""
    return bool(lib().mmx_register_fft_length(int(n)))


def check(rc: int, what: str = ''):
    """Raise NativeError for a non-zero return code of the C ABI."""
    if rc != 0:
        msg = lib().mmx_last_error().decode(errors='replace')
        if rc == -2:
            raise ValueError(f'{what}: {msg}')
        raise NativeError(f'{what} failed ({rc}): {msg}')
