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
    "amigo_chol_create": (ctypes.c_int, [ctypes.c_int, _c_int_p, _c_int_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "amigo_chol_factor": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "amigo_ldl_create_grouped": (
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
            _c_int_p,
            ctypes.POINTER(ctypes.c_void_p),
        ],
    ),
    "amigo_ldl_create_perm": (
        ctypes.c_int,
        [
            ctypes.c_int,
            _c_int_p,
            _c_int_p,
            ctypes.c_int,
            ctypes.c_double,
            ctypes.c_double,
            ctypes.c_double,
            _c_int_p,
            ctypes.POINTER(ctypes.c_void_p),
        ],
    ),
    "amigo_ldl_create_union": (
        ctypes.c_int,
        [
            ctypes.c_int,
            _c_int_p,
            _c_int_p,
            ctypes.c_int,
            ctypes.POINTER(ctypes.c_void_p),
            ctypes.POINTER(ctypes.c_void_p),
        ],
    ),
    "amigo_ldl_set_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_int_p]),
    "amigo_ldl_set_batch_mask": (ctypes.c_int, [ctypes.c_void_p, _c_int_p, ctypes.c_void_p]),
    "amigo_ldl_get_batch_inertia": (
        ctypes.c_int,
        [ctypes.c_void_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p],
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
    "amigo_ldl_memory": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)]),
    "amigo_ldl_set_graphs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "amigo_ldl_get_stats": (ctypes.c_int, [ctypes.c_void_p, _c_int_p, _c_int_p]),
    "amigo_ldl_set_scaling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "amigo_ldl_set_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "amigo_ldl_get_timing": (
        ctypes.c_int,
        [ctypes.c_void_p, _c_dbl_p, ctypes.POINTER(ctypes.c_longlong), _c_dbl_p, _c_dbl_p],
    ),
}


class IpmProblem(ctypes.Structure):
    """amigo_ipm_problem (include/amigo_ipm.h)."""

    _fields_ = [
        ("n_primal", ctypes.c_int),
        ("n_constraints", ctypes.c_int),
        ("primal_indices", ctypes.c_void_p),
        ("constraint_indices", ctypes.c_void_p),
        ("lbx", ctypes.c_void_p),
        ("ubx", ctypes.c_void_p),
        ("lbh", ctypes.c_void_p),
    ]


class IpmState(ctypes.Structure):
    """amigo_ipm_state (include/amigo_ipm.h)."""

    _fields_ = [
        ("xlam", ctypes.c_void_p),
        ("zl", ctypes.c_void_p),
        ("zu", ctypes.c_void_p),
        ("sl", ctypes.c_void_p),
        ("su", ctypes.c_void_p),
    ]


class IpmBatch(ctypes.Structure):
    """amigo_ipm_batch (include/amigo_ipm.h)."""

    _fields_ = [
        ("num_problems", ctypes.c_int),
        ("pseg", ctypes.c_void_p),
        ("cseg", ctypes.c_void_p),
        ("xseg", ctypes.c_void_p),
        ("nblk_total", ctypes.c_int),
        ("blk_total", ctypes.c_void_p),
        ("first_total", ctypes.c_void_p),
        ("count_total", ctypes.c_void_p),
        ("nblk_primal", ctypes.c_int),
        ("blk_primal", ctypes.c_void_p),
        ("first_primal", ctypes.c_void_p),
        ("count_primal", ctypes.c_void_p),
        ("active", ctypes.c_void_p),
        ("nlaunch_total", ctypes.c_int),
        ("launch_total", ctypes.c_void_p),
        ("nlaunch_primal", ctypes.c_int),
        ("launch_primal", ctypes.c_void_p),
        ("nlaunch_problems", ctypes.c_int),
        ("launch_problems", ctypes.c_void_p),
    ]


_V = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_PP = ctypes.POINTER(IpmProblem)
_PS = ctypes.POINTER(IpmState)
_PB = ctypes.POINTER(IpmBatch)
_SIGS.update(
    {
        "amigo_ipm_last_error": (ctypes.c_char_p, []),
        "amigo_ipm_workspace_doubles": (_I, []),
        "amigo_ipm_residual": (_I, [_PP, _PS, _D, _V, _V, _V, _c_dbl_p, _V]),
        "amigo_ipm_diagonal": (_I, [_PP, _PS, _V, _V]),
        "amigo_ipm_bound_dual_step": (_I, [_PP, _PS, _D, _V, _V, _V, _V]),
        "amigo_ipm_max_step": (_I, [_PP, _PS, _D, _V, _V, _V, _V, _c_dbl_p, _V]),
        "amigo_ipm_apply_step": (_I, [_PP, _PS, _D, _D, _V, _V, _V, _I, _PS, _V]),
        "amigo_ipm_reduce": (_I, [_I, _PP, _PS, _D, _V, _V, _V, _V, _c_dbl_p, _V]),
        "amigo_ipm_reduce_noalloc": (_I, [_I, _PP, _PS, _D, _V, _V, _V, _c_dbl_p, _V]),
        "amigo_ipm_reset_bound_multipliers": (_I, [_PP, _PS, _D, _D, _V]),
        "amigo_ipm_dual_residual_vector": (_I, [_PP, _PS, _V, _V, _I, _V]),
        "amigo_ipm_affine_start_point": (_I, [_PP, _PS, _D, _V, _V, _V, _V, _V]),
        "amigo_ipm_initialize": (_I, [_V, _V, _I, _V, _D, _D, _D, _D, _V, _V, _PS, _I, _I, _V]),
        "amigo_ipm_floor_slacks": (_I, [_PP, _PS, _V, _V, _D, _D, _D, _V, _c_dbl_p, _V]),
        "amigo_csr_spmv": (_I, [_I, _V, _V, _V, _V, _V, _V]),
        "amigo_csr_spmv_masked": (_I, [_I, _V, _V, _V, _V, _V, _V, _V, _V]),
        "amigo_csr_scale_symmetric_masked": (_I, [_I, _V, _V, _V, _V, _V, _V, _V]),
        "amigo_elements_eval_masked": (_I, [_I, _I, _V, _V, _I, _V, _D, _V, _V, _V, _V, _V, _V, _V]),
        "amigo_gather_sum_masked": (_I, [_I, _V, _V, _V, _V, _I, _V, _I, _V, _V]),
        "amigo_gather_sum_listed": (_I, [_V, _V, _V, _V, _I, _V, _V, _I, ctypes.c_longlong, _V]),
        "amigo_csr_scale_symmetric": (_I, [_I, _V, _V, _V, _V, _V]),
        "amigo_ipm_constraint_scaling": (_I, [_I, _V, _V, _V, _D, _D, _V, _V]),
        "amigo_ipm_scale_rows": (_I, [_I, _V, _V, _V, _I, _V]),
        "amigo_ipm_ir_init": (_I, [_PP, _PS, _D, _V, _V, _V, _V, _V, _V, _c_dbl_p, _V]),
        "amigo_ipm_ir_residual": (
            _I,
            [_PP, _PS, _D, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V, _c_dbl_p, _V],
        ),
        "amigo_ipm_ir_update": (_I, [_PP, _PS, _I, _V, _V, _V, _V, _V, _V, _V]),
        "amigo_assembly_last_error": (ctypes.c_char_p, []),
        "amigo_elem_num_inputs": (_I, [_I]),
        "amigo_elem_register": (_I, [_I, _I, _V]),
        "amigo_elements_eval": (_I, [_I, _I, _V, _V, _I, _V, _D, _V, _V, _V, _V]),
        "amigo_gather_sum": (_I, [_I, _V, _V, _V, _V, _I, _V]),
        "amigo_sum_partials": (_I, [_V, _I, _V, _c_dbl_p, _V]),
        "amigo_elements_eval_scaled": (_I, [_I, _I, _V, _V, _I, _V, _D, _V, _V, _V, _V, _V]),
        "amigo_sum_partials_batch": (_I, [_V, _I, _V, _V, _V, _c_dbl_p, _V]),
        "amigo_ipm_grid_blocks": (_I, [_I]),
        "amigo_ipm_batch_residual": (_I, [_PB, _PP, _PS, _V, _V, _V, _V, _V, _c_dbl_p, _V]),
        "amigo_ipm_batch_diagonal": (_I, [_PB, _PP, _PS, _V, _V]),
        "amigo_ipm_batch_bound_dual_step": (_I, [_PB, _PP, _PS, _V, _V, _V, _V, _V]),
        "amigo_ipm_batch_max_step": (_I, [_PB, _PP, _PS, _V, _V, _V, _V, _V, _V, _c_dbl_p, _V]),
        "amigo_ipm_batch_apply_step": (_I, [_PB, _PP, _PS, _V, _V, _V, _V, _V, _PS, _V]),
        "amigo_ipm_batch_reduce": (_I, [_I, _PB, _PP, _PS, _V, _V, _V, _V, _V, _V, _c_dbl_p, _V]),
        "amigo_ipm_batch_reset_bound_multipliers": (_I, [_PB, _PP, _PS, _V, _D, _V]),
        "amigo_ipm_batch_dual_residual_vector": (_I, [_PB, _PP, _PS, _V, _V, _V]),
        "amigo_ipm_batch_initialize": (_I, [_PB, _V, _V, _V, _D, _D, _D, _D, _V, _V, _PS, _I, _I, _V]),
        "amigo_ipm_batch_floor_slacks": (_I, [_PB, _PP, _PS, _V, _V, _V, _V, _D, _V, _V, _c_dbl_p, _V]),
        "amigo_ipm_batch_scale_rows": (_I, [_PB, _V, _V, _V, _I, _V]),
        "amigo_ipm_batch_ir_init": (_I, [_PB, _PP, _PS, _V, _V, _V, _V, _V, _V, _V, _V, _c_dbl_p, _V]),
        "amigo_ipm_batch_ir_residual": (
            _I,
            [_PB, _PP, _PS, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V, _c_dbl_p, _V],
        ),
        "amigo_ipm_batch_ir_update": (_I, [_PB, _PP, _PS, _V, _V, _V, _V, _V, _V, _V]),
        "amigo_ipm_batch_copy_state": (_I, [_PB, _PS, _PS, _I, _V]),
        "amigo_segment_absmax": (_I, [_V, _V, _I, _V, _c_dbl_p, _V]),
    }
)


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


def check(rc, which="ldl"):
    if rc != 0:
        fn = {
            "ldl": "amigo_last_error",
            "ipm": "amigo_ipm_last_error",
            "asm": "amigo_assembly_last_error",
        }[which]
        msg = getattr(lib(), fn)()
        raise AmigoError(rc, msg.decode() if msg else "")
    return rc


def dptr(t):
    """void* of a device tensor (None -> NULL)."""
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)
