"""
ctypes binding of ``libvulkpy_hip.so``, the C-ABI HIP library declared in
``include/vulkpy_hip.h``.

This module replaces the pybind11 extension ``vulkpy._vkarray``
(vulkpy/_vkarray.cc:756-898).  ctypes releases the GIL around every foreign
call, as the reference's ``call_guard<gil_scoped_release>`` does for
``submit`` / ``wait`` / ``random_*`` (_vkarray.cc:770-791, 877-878, 894-897).

There is no fallback: if the library is missing the first use raises
``RuntimeError`` naming the build command.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, c_char_p, c_double, c_float, c_int, c_size_t,
                    c_uint16, c_uint32, c_uint64, c_void_p)
from typing import Optional

__all__ = ["lib", "load", "check", "LIB_PATH", "HEADER_PATH",
           "VKH_OK", "VKH_ERROR", "VKH_TIMEOUT", "VKH_INVALID",
           "VKH_OUT_OF_MEMORY", "VKH_NO_DEVICE"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("VULKPY_HIP_LIB", os.path.join(_HERE, "libvulkpy_hip.so"))
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "vulkpy_hip.h")

VKH_OK = 0
VKH_ERROR = 1
VKH_TIMEOUT = 2
VKH_INVALID = 3
VKH_OUT_OF_MEMORY = 4
VKH_NO_DEVICE = 5

# opcode tables (must match the enums of include/vulkpy_hip.h)
BINOP = {"add": 0, "sub": 1, "mul": 2, "div": 3, "max": 4, "min": 5, "pow": 6,
         "rsub": 7, "rdiv": 8, "rpow": 9}
UNOP = {"abs": 0, "sign": 1, "sin": 2, "cos": 3, "tan": 4, "asin": 5, "acos": 6,
        "atan": 7, "sinh": 8, "cosh": 9, "tanh": 10, "asinh": 11, "acosh": 12,
        "atanh": 13, "exp": 14, "log": 15, "exp2": 16, "log2": 17, "sqrt": 18,
        "invsqrt": 19}
REDOP = {"sum": 0, "prod": 1, "maximum": 2, "minimum": 3}
GEMM_MODE = {"fp32": 0, "bf16": 1, "bf16x3": 2}
ACT = {None: 0, "none": 0, "relu": 1, "sigmoid": 2}

P = c_void_p
U64 = c_uint64
_u64p = POINTER(c_uint64)

ADAM_MULTI_MAX = 8  # VKH_ADAM_MULTI_MAX


class AdamParam(ctypes.Structure):
    """struct vkh_adam_param of include/vulkpy_hip.h"""
    _fields_ = [("w", c_void_p), ("g", c_void_p), ("m", c_void_p), ("v", c_void_p),
                ("n", c_uint64), ("lr", c_float), ("eps", c_float), ("beta1", c_double),
                ("beta2", c_double), ("bpow", c_void_p), ("reg_kind", c_int), ("reg_l1", c_float),
                ("reg_s2", c_float)]

# name -> (restype, argtypes); every function returns an int status except
# vkh_last_error.
_SIGNATURES = {
    "vkh_last_error": (c_char_p, []),
    "vkh_abi_version": (c_int, [POINTER(c_int)]),
    "vkh_device_count": (c_int, [POINTER(c_int)]),
    "vkh_gpu_create": (c_int, [c_int, c_float, POINTER(P)]),
    "vkh_gpu_create_on_stream": (c_int, [c_int, P, POINTER(P)]),
    "vkh_gpu_destroy": (c_int, [P]),
    "vkh_gpu_wait": (c_int, [P]),
    "vkh_gpu_can_subgroup_arithmetic": (c_int, [P, POINTER(c_int)]),
    "vkh_gpu_get_stream": (c_int, [P, POINTER(P)]),
    "vkh_gpu_info": (c_int, [P, POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
    "vkh_alloc": (c_int, [P, c_size_t, POINTER(P)]),
    "vkh_free": (c_int, [P, P]),
    "vkh_empty_cache": (c_int, [P]),
    "vkh_memory_stats": (c_int, [P, POINTER(c_size_t), POINTER(c_size_t)]),
    "vkh_upload": (c_int, [P, P, P, c_size_t]),
    "vkh_download": (c_int, [P, P, P, c_size_t]),
    "vkh_host_alloc": (c_int, [c_size_t, POINTER(P)]),
    "vkh_host_free": (c_int, [P]),
    "vkh_host_empty_cache": (c_int, []),
    "vkh_host_stats": (c_int, [POINTER(c_size_t), POINTER(c_size_t)]),
    "vkh_copy": (c_int, [P, P, P, c_size_t]),
    "vkh_fill_f32": (c_int, [P, P, c_float, c_size_t]),
    "vkh_fill_u32": (c_int, [P, P, c_uint32, c_size_t]),
    "vkh_event_record": (c_int, [P, POINTER(P)]),
    "vkh_event_wait": (c_int, [P, U64]),
    "vkh_event_query": (c_int, [P, POINTER(c_int)]),
    "vkh_event_destroy": (c_int, [P]),
    "vkh_timer_record": (c_int, [P, POINTER(P)]),
    "vkh_event_elapsed_ms": (c_int, [P, P, POINTER(c_float)]),
    "vkh_timer_next_kernel": (c_int, [P, POINTER(P), POINTER(P)]),
    "vkh_binary": (c_int, [P, c_int, P, P, P, U64]),
    "vkh_binary_scalar": (c_int, [P, c_int, P, c_float, P, U64]),
    "vkh_binary_bcast": (c_int, [P, c_int, P, P, P, c_int, _u64p, _u64p, _u64p]),
    "vkh_unary": (c_int, [P, c_int, P, P, U64]),
    "vkh_clamp": (c_int, [P, P, P, P, c_float, c_float, P, c_int, _u64p, _u64p, _u64p, _u64p]),
    "vkh_broadcast_to": (c_int, [P, P, P, c_int, _u64p, _u64p]),
    "vkh_onehot": (c_int, [P, P, U64, c_uint32, P]),
    "vkh_reduce_all": (c_int, [P, c_int, P, U64, P, c_int]),
    "vkh_reduce_all_noalloc": (c_int, [P, c_int, P, U64, P, c_int]),
    "vkh_reduce_axis": (c_int, [P, c_int, P, P, U64, U64, U64, c_int, c_float, c_float]),
    "vkh_reduce_family": (c_int, [c_int, U64, U64, U64, POINTER(c_int)]),
    "vkh_reduce_configs": (c_int, [c_int, POINTER(c_int)]),
    "vkh_reduce_config_info": (c_int, [c_int, c_int, POINTER(c_int), POINTER(c_int)]),
    "vkh_reduce_get_config": (c_int, [P, c_int, U64, U64, U64, POINTER(c_int), POINTER(c_int)]),
    "vkh_reduce_set_config": (c_int, [P, c_int, U64, U64, U64, c_int]),
    "vkh_reduce_autotune": (c_int, [P, c_int, U64, U64, U64, c_int, c_int, POINTER(c_int),
                                    POINTER(c_float)]),
    "vkh_gather": (c_int, [P, P, U64, P, P, U64]),
    "vkh_gather_axis": (c_int, [P, P, P, P, U64, U64, U64, U64]),
    "vkh_xoshiro_create": (c_int, [P, c_uint32, U64, POINTER(P)]),
    "vkh_xoshiro_destroy": (c_int, [P]),
    "vkh_xoshiro_size": (c_int, [P, POINTER(c_uint32)]),
    "vkh_xoshiro_state": (c_int, [P, POINTER(P)]),
    "vkh_xoshiro_uint32": (c_int, [P, P, U64]),
    "vkh_xoshiro_float": (c_int, [P, P, U64]),
    "vkh_xoshiro_normal": (c_int, [P, P, U64, c_float, c_float]),
    "vkh_xoshiro_randrange": (c_int, [P, P, U64, c_uint32, c_uint32]),
    "vkh_box_muller": (c_int, [P, P, P, U64, c_float, c_float]),
    "vkh_gemm": (c_int, [P, c_int, c_int, c_int, U64, U64, U64, c_float, P, U64, P, U64,
                         c_float, P, U64, P, c_int, P]),
    "vkh_gemm_reserve": (c_int, [P, c_int, U64, U64, U64]),
    "vkh_gemm_configs": (c_int, [c_int, POINTER(c_int)]),
    "vkh_gemm_config_info": (c_int, [c_int, c_int, POINTER(c_int), POINTER(c_int), POINTER(c_int),
                                     POINTER(c_int), POINTER(c_int)]),
    "vkh_gemm_get_config": (c_int, [P, c_int, c_int, c_int, U64, U64, U64, POINTER(c_int),
                                    POINTER(c_int)]),
    "vkh_gemm_set_config": (c_int, [P, c_int, c_int, c_int, U64, U64, U64, c_int]),
    "vkh_gemm_autotune": (c_int, [P, c_int, c_int, c_int, U64, U64, U64, c_int, POINTER(c_int),
                                  POINTER(c_float)]),
    "vkh_gemm_ex": (c_int, [P, c_int, c_int, c_int, U64, U64, U64, c_float, P, U64, P, U64,
                            c_float, P, U64, P, c_int, P, P, U64, c_int]),
    "vkh_gemm_bf16a": (c_int, [P, c_int, c_int, c_int, U64, U64, U64, c_float, P, U64, P, U64,
                               c_float, P, U64, P, c_int, P, P, U64, c_int, P]),
    "vkh_gemm_bf16io": (c_int, [P, c_int, c_int, c_int, U64, U64, U64, c_float, P, U64, P, U64,
                                c_float, P, U64, P, c_int, P, P, U64, c_int, P, P, P, U64]),
    "vkh_activation":(c_int, [P, c_int, P, P, U64]),
    "vkh_activation_backward": (c_int, [P, c_int, P, P, P, U64]),
    "vkh_softmax": (c_int, [P, P, P, U64, U64]),
    "vkh_softmax_backward": (c_int, [P, P, P, P, U64, U64]),
    "vkh_cross_entropy": (c_int, [P, P, P, P, U64, U64]),
    "vkh_cross_entropy_backward": (c_int, [P, P, P, P, U64, c_float]),
    "vkh_softmax_cross_entropy": (c_int, [P, P, P, P, P, U64, U64]),
    "vkh_mse_rows": (c_int, [P, P, P, P, U64, U64]),
    "vkh_huber_rows": (c_int, [P, P, P, P, U64, U64]),
    "vkh_loss_grad": (c_int, [P, c_int, P, P, P, U64, c_float]),
    "vkh_loss_rows_grad": (c_int, [P, c_int, P, P, P, P, U64, U64, c_float]),
    "vkh_adam": (c_int, [P, P, P, P, P, U64, c_float, c_double, c_double, c_float, c_double, c_double]),
    "vkh_adam_pow_advance": (c_int, [P, P, c_double, c_double]),
    "vkh_adam_device": (c_int, [P, P, P, P, P, U64, c_float, c_double, c_double, c_float, P, P,
                                c_int, c_float, c_float]),
    "vkh_adam_device_multi": (c_int, [P, P, c_int]),
    "vkh_sgd": (c_int, [P, P, P, U64, c_float, c_int, c_float, c_float]),
    "vkh_regularizer_grad": (c_int, [P, c_int, P, P, U64, c_float, c_float, c_int]),
    "vkh_dense1_backward": (c_int, [P, P, P, P, c_int, U64, U64, P, P, P, U64, P, P]),
    "vkh_gemm_colsum_supported": (c_int, [P, c_int, c_int, c_int, U64, U64, U64, POINTER(c_int)]),
    "vkh_gemm_bf16io_colsum": (c_int, [P, c_int, c_int, c_int, U64, U64, U64, c_float, P, U64, P, U64,
                                       c_float, P, U64, P, c_int, P, P, U64, c_int, P, P, P, U64, P,
                                       c_int]),
    "vkh_dense1_backward_supported": (c_int, [P, U64, U64, POINTER(c_int)]),
    "vkh_gemm_bf16io_ext": (c_int, [P, c_int, c_int, c_int, U64, U64, U64, c_float, P, U64, P, U64,
                                    c_float, P, U64, P, c_int, P, P, P, U64, c_int, P, P, P, U64, P,
                                    c_int]),
    "vkh_gemm_autotune_ext": (c_int, [P, c_int, c_int, c_int, U64, U64, U64, c_float, P, U64, P, U64,
                                      c_float, P, U64, P, c_int, P, P, P, U64, c_int, P, P, P, U64, P,
                                      c_int, c_int, POINTER(c_int), POINTER(c_float)]),
    "vkh_adagrad": (c_int, [P, P, P, P, U64, c_float, c_float, c_int, c_float, c_float]),
    "vkh_bias_grad": (c_int, [P, P, P, U64, U64, c_float]),
    "vkh_to_bf16": (c_int, [P, P, P, U64]),
    "vkh_graph_begin": (c_int, [P]),
    "vkh_graph_end": (c_int, [P, POINTER(P)]),
    "vkh_graph_launch": (c_int, [P, P]),
    "vkh_graph_destroy": (c_int, [P]),
}

_lib: Optional[ctypes.CDLL] = None
_load_error: Optional[str] = None


def load() -> ctypes.CDLL:
    """Load the library once and declare every signature."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        _load_error = (f"HIP library not found at {LIB_PATH}; build it with "
                       f"`make -C vulkpy_amd/csrc` or `python -c 'import __graft_entry__ as g; g.build()'`")
        raise RuntimeError(_load_error)
    handle = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    _lib = handle
    return handle


class _Lib:
    """Attribute access to the loaded library (loads on first use)."""

    def __getattr__(self, name):
        return getattr(load(), name)


lib = _Lib()


def last_error() -> str:
    msg = load().vkh_last_error()
    return msg.decode("utf-8", "replace") if msg else ""


def check(status: int, what: str = "") -> None:
    """Raise the Python exception matching a non-zero status."""
    if status == VKH_OK:
        return
    msg = last_error() or what or f"status {status}"
    if status == VKH_INVALID:
        raise ValueError(msg)
    if status == VKH_OUT_OF_MEMORY:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def call(name: str, *args) -> None:
    """Call an int-status function and raise on failure."""
    check(getattr(load(), name)(*args), name)


def shape_array(shape) -> ctypes.Array:
    """uint64 array of a shape tuple (for the broadcast entry points)."""
    arr = (c_uint64 * max(len(shape), 1))()
    for i, v in enumerate(shape):
        arr[i] = int(v)
    return arr


def exported_symbols():
    """Names of the entry points this binding declares."""
    return list(_SIGNATURES)
