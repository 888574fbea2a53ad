"""ctypes binding of ``libvk_hip.so`` (the C ABI in ``include/vk_hip.h``).

The library is built in-tree by ``verkehrt_amd.build``. Loading fails loudly when it is
missing: there is no CPU fallback for any kernel of the hip variant.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libvk_hip.so")

# vk_type
F64, F32, I64, U64, I32, U8, BOOL = range(7)
ROWS, SPLAT = 0, 1
VALUE_ARRAY, VALUE_SCALAR = 0, 1
OK, ERR_RUNTIME, ERR_USAGE, ERR_DEVICE = 0, 1, 2, 3


class Status(C.Structure):
    _fields_ = [("code", C.c_int32), ("msg", C.c_char * 512)]


class CScalar(C.Structure):
    _fields_ = [("type", C.c_int32), ("bits", C.c_uint64)]


class CArg(C.Structure):
    _fields_ = [("kind", C.c_int32), ("rows", C.c_void_p), ("splat", CScalar)]


class CValue(C.Structure):
    _fields_ = [("kind", C.c_int32), ("array", C.c_void_p), ("scalar", CScalar)]


BREF_INPUT, BREF_STMT, BREF_SCALAR = 0, 1, 2
SRC_INPUT, SRC_KERNEL = 0, 1


class CSrc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("index", C.c_int32)]


class CRKernel(C.Structure):
    _fields_ = [("name", C.c_char_p), ("args", CSrc * 3), ("nargs", C.c_int32)]


class CBref(C.Structure):
    _fields_ = [("kind", C.c_int32), ("index", C.c_int32), ("scalar", CScalar)]


# Every symbol the header declares, with its ctypes signature.
_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_ST = C.POINTER(Status)
SIGNATURES = {
    "vk_ctx_create": (C.c_int, [C.c_int, _PP, _ST]),
    "vk_ctx_destroy": (None, [_P]),
    "vk_ctx_sync": (C.c_int, [_P, _ST]),
    "vk_ctx_stream": (_P, [_P]),
    "vk_ctx_trim": (None, [_P]),
    "vk_ctx_pool_bytes": (C.c_uint64, [_P, C.POINTER(C.c_uint64)]),
    "vk_ctx_set_trace": (None, [_P, C.c_int]),
    "vk_ctx_trace_take": (C.c_size_t, [_P, C.c_char_p, C.c_size_t]),
    "vk_ctx_last_us": (C.c_double, [_P]),
    "vk_array_new": (C.c_int, [_P, C.c_int32, C.c_uint64, _PP, _ST]),
    "vk_array_from_host": (C.c_int, [_P, C.c_int32, _P, C.c_uint64, _PP, _ST]),
    "vk_array_to_host": (C.c_int, [_P, _P, _P, _ST]),
    "vk_array_wrap": (C.c_int, [_P, C.c_int32, _P, C.c_uint64, _PP, _ST]),
    "vk_array_slice": (C.c_int, [_P, C.c_uint64, C.c_uint64, _PP, _ST]),
    "vk_array_retain": (None, [_P]),
    "vk_array_release": (None, [_P]),
    "vk_array_len": (C.c_uint64, [_P]),
    "vk_array_type": (C.c_int32, [_P]),
    "vk_array_bytes": (C.c_uint64, [_P]),
    "vk_array_device_ptr": (_P, [_P]),
    "vk_kernel_count": (C.c_int32, []),
    "vk_kernel_lookup": (C.c_int32, [C.c_char_p]),
    "vk_kernel_name": (C.c_char_p, [C.c_int32]),
    "vk_kernel_class": (C.c_int32, [C.c_int32]),
    "vk_variant_name": (C.c_char_p, []),
    "vk_call": (C.c_int, [_P, C.c_char_p, C.POINTER(CArg), C.c_size_t, C.POINTER(CValue), _ST]),
    "vk_reducer_new": (C.c_int, [_P, C.c_char_p, _PP, _ST]),
    "vk_reducer_update": (C.c_int, [_P, C.POINTER(CArg), C.c_size_t, _ST]),
    "vk_reducer_finish": (C.c_int, [_P, C.POINTER(CScalar), _ST]),
    "vk_reducer_free": (None, [_P]),
    "vk_scanner_new": (C.c_int, [_P, C.c_char_p, _PP, _ST]),
    "vk_scanner_update": (C.c_int, [_P, C.POINTER(CArg), C.c_size_t, _P, _ST]),
    "vk_scanner_free": (None, [_P]),
    "vk_builder_new": (C.c_int, [_P, C.c_int32, _PP, _ST]),
    "vk_builder_len": (C.c_uint64, [_P]),
    "vk_builder_finish": (C.c_int, [_P, _PP, _ST]),
    "vk_builder_free": (None, [_P]),
    "vk_select_tile": (C.c_int, [_P, C.c_char_p, C.POINTER(CArg), C.c_size_t, C.c_uint64, _P, _ST]),
    "vk_elementwise_tile": (C.c_int, [_P, C.c_char_p, C.POINTER(CArg), C.c_size_t, _P, _ST]),
    "vk_region_reduce": (C.c_int, [_P, C.POINTER(CRKernel), C.c_size_t, C.c_char_p, C.POINTER(CSrc), C.c_size_t,
                                   C.POINTER(CArg), C.c_size_t, C.POINTER(CScalar), C.POINTER(C.c_int32), _ST]),
    "vk_select_into": (C.c_int, [_P, C.c_char_p, C.POINTER(CArg), C.c_size_t, _P, C.POINTER(C.c_uint64), _ST]),
    "vk_value_release": (None, [C.POINTER(CValue)]),
    "vk_batch_call": (C.c_int, [_P, C.c_char_p, C.POINTER(CArg), C.c_size_t, C.c_size_t,
                                C.POINTER(CValue), _ST]),
    "vk_bplan_new": (C.c_int, [_P, _PP, C.c_size_t, C.c_size_t, _PP, _ST]),
    "vk_bplan_add": (C.c_int, [_P, C.c_char_p, C.POINTER(CBref), C.c_size_t, C.POINTER(C.c_size_t), _ST]),
    "vk_bplan_compile": (C.c_int, [_P, C.c_int, _ST]),
    "vk_bplan_run": (C.c_int, [_P, _ST]),
    "vk_bplan_result": (C.c_int, [_P, C.c_size_t, C.c_size_t, C.POINTER(CValue), _ST]),
    "vk_bplan_counts": (C.c_int, [_P, C.c_size_t, C.POINTER(C.c_uint64), _ST]),
    "vk_bplan_tuning": (C.c_size_t, [_P, C.c_char_p, C.c_size_t]),
    "vk_bplan_free": (None, [_P]),
}

_lib = None


def lib():
    """The loaded library; raises if the hip backend has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"hip backend not built: {LIB_PATH} is missing (run verkehrt_amd.build.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib
