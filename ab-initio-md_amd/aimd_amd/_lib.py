"""
ctypes binding of libaimdhf.so, the gfx950 engine (C ABI: include/aimdhf.h).

The library is built in-tree (``make -C ab-initio-md_amd``, or
``__graft_entry__.build()``). There is no fallback: if the library or a GPU
is missing, :func:`lib` raises, so nothing silently runs elsewhere.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaimdhf.so")

AIMDHF_OK = 0
AIMDHF_E_ARG = 1
AIMDHF_E_NOMEM = 2
AIMDHF_E_HIP = 3
AIMDHF_E_UNSUPPORTED = 4
AIMDHF_E_STATE = 5

# every entry point declared in include/aimdhf.h: name -> (restype, argtypes)
_P, _I, _D, _L, _S = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64, ctypes.c_size_t
SIGNATURES = {
    "aimdhf_last_error": (_I, [ctypes.c_char_p, _S]),
    "aimdhf_lmax": (_I, []),
    "aimdhf_plan_create": (_I, [_I, _I, _P, _P, _P, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "aimdhf_plan_destroy": (_I, [_P]),
    "aimdhf_batch_create": (_I, [_I, _P, _P, _I, _I, _P, _P, _P, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "aimdhf_batch_layout": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_L)]),
    "aimdhf_set_geometry": (_I, [_P, _P, _D, _P]),
    "aimdhf_schwarz_bounds": (_I, [_P, _P, _L]),
    "aimdhf_quartet_count": (_I, [_P, ctypes.POINTER(_L)]),
    "aimdhf_quartets": (_I, [_P, _P, _L, ctypes.POINTER(_L)]),
    "aimdhf_jk": (_I, [_P, _I, _P, _P, _P, _P]),
    "aimdhf_jk_delta": (_I, [_P, _I, _P, _P, _P, _D, _P]),
    "aimdhf_set_jk_mode": (_I, [_P, _I, _L]),
    "aimdhf_set_precision": (_I, [_P, _I, _D, _D, _D]),
    "aimdhf_set_streams": (_I, [_P, _I]),
    "aimdhf_tc2_step": (_I, [_I, _P, _P, _P, _D, _D, _P]),
    "aimdhf_tc2_init": (_I, [_I, _P, _P, _P]),
    "aimdhf_batch_purify": (_I, [_I, _I, _P, _P, _P, _I, _D, _P, _P]),
    "aimdhf_batch_diis_solve": (_I, [_I, _I, _P, _P, _P, _P]),
    "aimdhf_batch_set_jk_lanes": (_I, [_P, _I]),
    "aimdhf_tc2_error": (_I, [_I, _P, _P, _P]),
    "aimdhf_jk_info": (_I, [_P, _P, _P, _P, _P, _P]),
    "aimdhf_jk_type_info": (_I, [_P, _P, _P, _P, _P, _P]),
    "aimdhf_set_store_units": (_I, [_P, _I]),
    "aimdhf_profile_replay": (_I, [_P, _I, _I, _P, _P]),
    "aimdhf_one_electron": (_I, [_P, _P, _P, _P, _P, _P]),
    "aimdhf_grad_one": (_I, [_P, _P, _P, _P, _P]),
    "aimdhf_grad_one_terms": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "aimdhf_grad_one_noalloc": (_I, [_P, _P, _P, _P, _P]),
    "aimdhf_grad_two": (_I, [_P, _P, _I, _P, _D, _D, _P, _P]),
    "aimdhf_profile_enable": (_I, [_P, _I]),
    "aimdhf_profile_read": (_I, [_P, _P, _P, _I, _I]),
    "aimdhf_class_counts": (_I, [_P, _P, _P]),
    "aimdhf_pair_npp": (_I, [_P, _P, _L]),
    "aimdhf_grad_quartets": (_I, [_P, _P, _L, ctypes.POINTER(_L)]),
}

_lib = None


class EngineError(RuntimeError):
    """A HIP runtime failure inside the engine."""


def lib() -> ctypes.CDLL:
    """Load libaimdhf.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with `make -C ab-initio-md_amd` "
                          "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().aimdhf_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(code: int) -> None:
    """Map a status code to the exception the reference raises for that condition."""
    if code == AIMDHF_OK:
        return
    msg = last_error()
    if code in (AIMDHF_E_ARG, AIMDHF_E_UNSUPPORTED):
        raise ValueError(msg)
    if code == AIMDHF_E_NOMEM:
        raise MemoryError(msg)
    if code == AIMDHF_E_STATE:
        raise RuntimeError(msg)
    raise EngineError(msg)
