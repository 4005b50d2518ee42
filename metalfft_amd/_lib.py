"""ctypes binding of libmifft.so (the C-ABI declared in include/mifft.h).

The library is loaded from this package directory only; a missing or stale
build raises instead of falling back to any other implementation.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmifft.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "mifft.h")

c_int64 = ctypes.c_int64
c_int = ctypes.c_int
c_void_p = ctypes.c_void_p
c_char_p = ctypes.c_char_p
c_size_t = ctypes.c_size_t
fptr = ctypes.POINTER(ctypes.c_float)
fpp = ctypes.POINTER(fptr)
i64p = ctypes.POINTER(ctypes.c_int64)

# mifft_status values (include/mifft.h) and the FFTError case each mirrors.
STATUS_CASES = {
    0: "ok",
    1: "noMetalDevice",
    2: "noCommandQueue",
    3: "libraryBuildFailed",
    4: "kernelNotFound",
    5: "commandBufferFailed",
    6: "invalidInputSize",
    7: "batchInputSize",
    8: "unsupportedFFTSize",
    9: "bufferAllocationFailed",
    10: "precondition",
    11: "invalidArgument",
}

FORWARD = -1
INVERSE = 1


class FFTError(Exception):
    """Mirror of the reference's FFTError
    (Sources/MetalFFT/Internal/MetalContext.swift:6-30).

    ``case`` names the enum case; ``expected`` / ``got`` / ``batch_index``
    carry its payload where it has one."""

    def __init__(self, status: int, message: str, expected=None, got=None, batch_index=None):
        super().__init__(message)
        self.status = status
        self.case = STATUS_CASES.get(status, "unknown")
        self.expected = expected
        self.got = got
        self.batch_index = batch_index


# (name, restype, argtypes) of every exported entry point.
_SIGNATURES = [
    ("mifft_status_string", c_char_p, [c_int]),
    ("mifft_last_error", c_int, [i64p, i64p, i64p]),
    ("mifft_last_error_message", c_char_p, []),
    ("mifft_init", c_int, []),
    ("mifft_shutdown", c_int, []),
    ("mifft_device_name", c_int, [ctypes.c_char_p, c_size_t]),
    ("mifft_plan_create", c_int, [c_int64, ctypes.POINTER(c_void_p)]),
    ("mifft_plan_create_ex", c_int, [c_int64, c_int, c_int, ctypes.POINTER(c_void_p)]),
    ("mifft_plan_destroy", c_int, [c_void_p]),
    ("mifft_plan_size", c_int64, [c_void_p]),
    ("mifft_plan_kernel_name", c_char_p, [c_void_p]),
    ("mifft_kernel_variants", c_int, [c_int64, ctypes.POINTER(c_char_p), c_int]),
    ("mifft_plan_select_kernel", c_int, [c_void_p, c_char_p]),
    ("mifft_execute_device", c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_void_p]),
    ("mifft_forward_real_device", c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_void_p]),
    ("mifft_inverse_real_device", c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int64, c_void_p]),
    ("mifft_forward_host", c_int, [c_void_p, fptr, c_int64, fptr]),
    ("mifft_inverse_host", c_int, [c_void_p, fptr, c_int64, fptr]),
    ("mifft_forward_batch_host", c_int, [c_void_p, fpp, i64p, c_int64, fpp]),
    ("mifft_inverse_batch_host", c_int, [c_void_p, fpp, i64p, c_int64, fpp]),
    ("mifft_convolver_create", c_int, [fptr, c_int64, c_int64, c_int, c_int, ctypes.POINTER(c_void_p)]),
    ("mifft_convolver_destroy", c_int, [c_void_p]),
    ("mifft_convolver_block_size", c_int64, [c_void_p]),
    ("mifft_convolver_hop_size", c_int64, [c_void_p]),
    ("mifft_convolver_kernel_len", c_int64, [c_void_p]),
    ("mifft_convolver_recommended_fft_size", c_int64, [c_int64]),
    ("mifft_convolver_apply_host", c_int, [c_void_p, fptr, c_int64, fptr]),
    ("mifft_convolver_apply_device", c_int, [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_void_p]),
    ("mifft_convolver_set_taps_device", c_int, [c_void_p, c_void_p, c_void_p]),
    ("mifft_convolver_backward_device", c_int,
     [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("mifft_convolver_backward_device_noalloc", c_int,
     [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("mifft_ragged_plan_create", c_int, [i64p, i64p, c_int64, ctypes.POINTER(c_void_p)]),
    ("mifft_ragged_plan_destroy", c_int, [c_void_p]),
    ("mifft_ragged_extent", c_int64, [c_void_p]),
    ("mifft_ragged_elements", c_int64, [c_void_p]),
    ("mifft_ragged_launches", c_int64, [c_void_p, c_int]),
    ("mifft_ragged_execute_device", c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    ("mifft_rda_default_params", c_int, [c_void_p]),
    ("mifft_rda_create", c_int, [c_void_p, ctypes.POINTER(c_void_p)]),
    ("mifft_rda_destroy", c_int, [c_void_p]),
    ("mifft_rda_scales", c_int, [c_void_p, ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_double), i64p]),
    ("mifft_rda_process_device", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("mifft_rda_range_compress_device", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("mifft_rda_corner_turn_device", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("mifft_rda_azimuth_compress_device", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("mifft_correlator_create", c_int, [c_int64, ctypes.POINTER(c_void_p)]),
    ("mifft_correlator_destroy", c_int, [c_void_p]),
    ("mifft_correlator_fft_size", c_int64, [c_void_p]),
    ("mifft_correlator_cross_host", c_int, [c_void_p, fptr, c_int64, fptr, c_int64, fptr]),
    ("mifft_correlator_auto_host", c_int, [c_void_p, fptr, c_int64, fptr]),
    ("mifft_correlator_cross_device", c_int,
     [c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_int64, c_void_p, c_void_p]),
    ("mifft_copy_device", c_int, [c_void_p, c_void_p, c_int64, c_void_p]),
    ("mifft_xspec_create", c_int, [c_int64, c_int64, ctypes.POINTER(c_void_p)]),
    ("mifft_xspec_destroy", c_int, [c_void_p]),
    ("mifft_xspec_frames", c_int64, [c_void_p, c_int64, c_int64]),
    ("mifft_xspec_device", c_int, [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p, c_void_p]),
    ("mifft_stft_device", c_int, [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_void_p,
                                  c_void_p, c_void_p, ctypes.c_float, c_void_p]),
    ("mifft_welch_device", c_int, [c_void_p, c_void_p, c_int64, c_int64, c_void_p,
                                   ctypes.c_double, c_void_p, c_void_p]),
    ("mifft_xspec_host", c_int, [c_void_p, fptr, c_int64, c_int64, fptr, fptr, fptr, fptr, fptr]),
    ("mifft_stft_host", c_int, [c_void_p, fptr, c_int64, c_int64, fptr, fptr, fptr, fptr,
                                ctypes.c_float]),
    ("mifft_welch_host", c_int, [c_void_p, fptr, c_int64, c_int64, fptr, ctypes.c_double, fptr]),
]

_lib = None


def exported_names() -> list[str]:
    return [s[0] for s in _SIGNATURES]


def load():
    """Load libmifft.so from the package directory (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libmifft.so not found at {LIB_PATH}; build it with "
            "`python -m metalfft_amd.build` (hipcc, gfx950)")
    # torch ships its own libamdhip64 / libhsa-runtime64.  Two HIP runtimes
    # in one process cannot both open the GPU, so when torch is present it is
    # loaded first and libmifft's libamdhip64.so.7 dependency binds to the
    # runtime already in the process (same soname).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in _SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status: int) -> None:
    """Raise FFTError for a non-zero mifft_status, with the thread's payload."""
    if status == 0:
        return
    L = load()
    e, g, i = c_int64(-1), c_int64(-1), c_int64(-1)
    L.mifft_last_error(ctypes.byref(e), ctypes.byref(g), ctypes.byref(i))
    msg = (L.mifft_last_error_message() or b"").decode() or L.mifft_status_string(status).decode()
    val = lambda v: None if v.value < 0 else int(v.value)
    raise FFTError(status, msg, val(e), val(g), val(i))
