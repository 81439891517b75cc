"""ctypes binding of libsurfhip.so (the C ABI declared in include/surfhip.h).

The library is loaded from this package's directory; it is built in-tree by
``surfacelets.jl_amd/build.py``.  There is no fallback: if the library is
missing every call raises.
"""
import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libsurfhip.so")

OK, ERR_ARG, ERR_DIM, ERR_DEVICE, ERR_OOM = 0, 1, 2, 3, 4
F32, C64 = 0, 1
MAX_SCALES = 16


class ArgumentError(ValueError):
    """The reference's ArgumentError."""


class DimensionMismatch(ValueError):
    """The reference's DimensionMismatch."""


class DeviceError(RuntimeError):
    """A HIP runtime failure or a device allocation failure."""


class DeviceOutOfMemory(DeviceError, MemoryError):
    """A device allocation the GPU cannot satisfy (SURFHIP_ERR_OOM)."""


class NSParams(ctypes.Structure):
    _fields_ = [
        ("scales", ctypes.c_int32),
        ("direction_levels", ctypes.c_int32 * MAX_SCALES),
        ("bernstein_order", ctypes.c_int32),
        ("halfband_order", ctypes.c_int32),
        ("transition", ctypes.c_double),
        ("tight", ctypes.c_int32),
    ]


class SParams(ctypes.Structure):
    _fields_ = [
        ("scales", ctypes.c_int32),
        ("levels", ((ctypes.c_int32 * 3) * 3) * MAX_SCALES),
        ("pyramid_ratio", ctypes.c_double),
        ("checkerboard_order", ctypes.c_int32),
        ("diamond_length", ctypes.c_int32),
        ("kaiser_beta", ctypes.c_double),
        ("diamond_power", ctypes.c_int32),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int32
_SIGS = {
    "surfhip_last_error": (ctypes.c_size_t, [ctypes.c_char_p, ctypes.c_size_t]),
    "surfhip_version": (ctypes.c_char_p, []),
    "surfhip_ns_nsubbands": (ctypes.c_int, [ctypes.POINTER(NSParams), ctypes.POINTER(_I)]),
    "surfhip_ns_plan_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(NSParams), _I, _I, _I, _I, _P]),
    "surfhip_ns_plan_destroy": (ctypes.c_int, [_P]),
    "surfhip_ns_set_stream": (ctypes.c_int, [_P, _P]),
    "surfhip_ns_plan_tune": (ctypes.c_int, [_P, ctypes.POINTER(_I)]),
    "surfhip_ns_analyze": (ctypes.c_int, [_P, _P, _P]),
    "surfhip_ns_analyze_noalloc": (ctypes.c_int, [_P, _P, _P]),
    "surfhip_ns_synthesize": (ctypes.c_int, [_P, _P, _P]),
    "surfhip_ns_adjoint": (ctypes.c_int, [_P, _P, _P]),
    "surfhip_ns_synthesize_adjoint": (ctypes.c_int, [_P, _P, _P]),
    "surfhip_ns_stream_begin": (ctypes.c_int, [_P, _P]),
    "surfhip_ns_stream_subband": (ctypes.c_int, [_P, _I, _P]),
    "surfhip_ns_stream_inverse_begin": (ctypes.c_int, [_P]),
    "surfhip_ns_stream_accumulate": (ctypes.c_int, [_P, _I, _P]),
    "surfhip_ns_stream_finish": (ctypes.c_int, [_P, _P]),
    "surfhip_ns_mask_host": (ctypes.c_int, [ctypes.POINTER(NSParams), _I, _I, _I, _I, ctypes.POINTER(ctypes.c_float)]),
    "surfhip_ndfb_plan_create": (ctypes.c_int, [ctypes.POINTER(_P), _P, _I, _I, _I, _I, _P]),
    "surfhip_ndfb_mask_host": (ctypes.c_int, [_P, _I, _I, _I, _I, ctypes.POINTER(ctypes.c_float)]),
    "surfhip_ns_group_table": (ctypes.c_int, [_P, _I, _I, ctypes.POINTER(ctypes.c_float)]),
    "surfhip_ns_group_table_host": (ctypes.c_int, [ctypes.POINTER(NSParams), _I, _I, _I, _I, _I, ctypes.POINTER(ctypes.c_float)]),
    "surfhip_ns_set_timing": (ctypes.c_int, [_P, _I]),
    "surfhip_ns_timing": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double),
                                         ctypes.c_char_p, ctypes.c_size_t]),
    "surfhip_s_check_size": (ctypes.c_int, [ctypes.POINTER(SParams), _I, _I, _I]),
    "surfhip_s_counts": (ctypes.c_int, [ctypes.POINTER(SParams), _I, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_int64)]),
    "surfhip_s_layout": (ctypes.c_int, [ctypes.POINTER(SParams), _I, _I, _I, ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_int64), _I]),
    "surfhip_s_plan_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(SParams), _I, _I, _I, _I, _P]),
    "surfhip_s_plan_destroy": (ctypes.c_int, [_P]),
    "surfhip_s_set_stream": (ctypes.c_int, [_P, _P]),
    "surfhip_s_analyze": (ctypes.c_int, [_P, _P, _P]),
    "surfhip_s_synthesize": (ctypes.c_int, [_P, _P, _P]),
    "surfhip_s_adjoint": (ctypes.c_int, [_P, _P, _P]),
    "surfhip_ns_batch_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(NSParams), _I,
                                               ctypes.POINTER(_I), _I, _I, _I, _P]),
    "surfhip_ns_batch_run": (ctypes.c_int, [_P, _I, _P, _P, _P]),
    "surfhip_ns_batch_set_stream": (ctypes.c_int, [_P, _P]),
    "surfhip_ns_batch_info": (ctypes.c_int, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "surfhip_ns_batch_destroy": (ctypes.c_int, [_P]),
    "surfhip_s_index_maps": (ctypes.c_int, [_P, _I, _I, _I, _I, ctypes.POINTER(ctypes.c_int64), ctypes.c_int64,
                                            ctypes.POINTER(ctypes.c_int64)]),
}

_lock = threading.Lock()
_lib = None


def lib():
    """The loaded library; raises if it was not built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIBPATH):
                raise ImportError(f"{LIBPATH} is missing: build it with surfacelets.jl_amd/build.py")
            L = ctypes.CDLL(LIBPATH)
            for name, (res, args) in _SIGS.items():
                if not hasattr(L, name):
                    continue
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def last_error():
    L = lib()
    n = L.surfhip_last_error(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    L.surfhip_last_error(buf, n + 1)
    return buf.value.decode("utf-8", "replace")


def check(status):
    """Raise the exception a status code maps to."""
    if status == OK:
        return
    msg = last_error()
    if status == ERR_ARG:
        raise ArgumentError(msg)
    if status == ERR_DIM:
        raise DimensionMismatch(msg)
    if status == ERR_OOM:
        raise DeviceOutOfMemory(msg)
    raise DeviceError(msg)


def call(name, *args):
    check(getattr(lib(), name)(*args))
