"""ctypes binding of libamigo_amd.so (C ABI in include/amigo_amd.h).

The library is built in-tree by ``amigo_amd.build``.  Loading fails loudly
when it is missing: there is no CPU fallback for the device path.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libamigo_amd.so")

AMIGO_ORDER = {"ND": 0, "NESTED_DISSECTION": 0, "AMD": 1, "MULTI_COLOR": 2, "NATURAL": 3}
AMIGO_ERR_CAPACITY = -4

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)
_lib = None


class AmigoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"amigo_amd error {code}: {msg}")
        self.code = code


def ip(a):
    """int32 numpy array -> int*"""
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_c_int_p)


def dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_c_dbl_p)


def vp(ptr):
    """integer device address -> void*"""
    return ctypes.c_void_p(int(ptr))


# name: (restype, argtypes)
_SIGS = {
    "amigo_last_error": (ctypes.c_char_p, []),
    "amigo_order": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p]),
    "amigo_block_amd": (ctypes.c_int, [ctypes.c_int, _c_int_p, _c_int_p, ctypes.c_int, _c_int_p, _c_int_p]),
    "amigo_etree_post_colcount": (
        ctypes.c_int,
        [ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p],
    ),
    "amigo_ldl_create": (
        ctypes.c_int,
        [
            ctypes.c_int,
            _c_int_p,
            _c_int_p,
            ctypes.c_int,
            ctypes.c_double,
            ctypes.c_double,
            ctypes.c_double,
            ctypes.c_int,
            ctypes.POINTER(ctypes.c_void_p),
        ],
    ),
    "amigo_ldl_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "amigo_ldl_sizes": (ctypes.c_int, [ctypes.c_void_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p]),
    "amigo_ldl_get_symbolic": (
        ctypes.c_int,
        [ctypes.c_void_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p],
    ),
    "amigo_ldl_factor": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "amigo_ldl_factor_shifted": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
    ),
    "amigo_ldl_get_inertia": (ctypes.c_int, [ctypes.c_void_p, _c_int_p, _c_int_p, _c_int_p]),
    "amigo_ldl_pivot_counts": (ctypes.c_int, [ctypes.c_void_p, _c_int_p, _c_int_p]),
    "amigo_ldl_solve": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
    ),
    "amigo_ldl_work": (ctypes.c_int, [ctypes.c_void_p, _c_dbl_p, _c_dbl_p]),
}


def exported_names():
    return list(_SIGS)


def register(name, restype, argtypes):
    _SIGS[name] = (restype, argtypes)
    if _lib is not None:
        f = getattr(_lib, name)
        f.restype = restype
        f.argtypes = argtypes


def lib():
    """Load (once) and return the ctypes library."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m amigo_amd.build` "
                "(the device path has no CPU fallback)"
            )
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().amigo_last_error()
        raise AmigoError(rc, msg.decode() if msg else "")
    return rc
