"""ctypes wrapper of oracle/_ref/libamigo_ref.so: the reference's own
SparseLDL / SparseCholesky / ordering code compiled from its headers.

ORACLE infrastructure: used only by tests/, smoke() and bench.py's
cpu_baseline leg, as the checker.
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_LIB = os.path.join(HERE, "_ref", "libamigo_ref.so")
REF_INCLUDE = "/root/reference/include"

_P = ctypes.POINTER
_ci = ctypes.c_int
_cd = ctypes.c_double
_lib = None


def build_ref(force=False):
    """Compile the reference harness when the reference tree is present."""
    if not os.path.isdir(REF_INCLUDE):
        return os.path.exists(REF_LIB)
    target = "_ref/libamigo_ref.so"
    cmd = ["make", "-s", "-C", HERE, target]
    if force:
        cmd.insert(2, "-B")
    subprocess.run(cmd, check=True)
    return True


def available():
    return os.path.exists(REF_LIB)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError(REF_LIB)
        os.environ.setdefault("MKL_THREADING_LAYER", "GNU")
        L = ctypes.CDLL(REF_LIB)
        sig = {
            "ref_order": (_ci, [_ci, _ci, _P(_ci), _P(_ci), _P(_ci), _P(_ci)]),
            "ref_block_amd": (_ci, [_ci, _P(_ci), _P(_ci), _ci, _P(_ci), _P(_ci)]),
            "ref_element_csr": (_ci, [_ci, _ci, _P(_ci), _P(_ci), _P(_ci), _P(_ci), _P(_ci), _P(_ci), _ci]),
            "ref_ldl_create": (ctypes.c_void_p, [_ci, _P(_ci), _P(_ci), _ci, _cd, _cd, _cd, _ci]),
            "ref_ldl_destroy": (None, [ctypes.c_void_p]),
            "ref_ldl_num_nodes": (_ci, [ctypes.c_void_p]),
            "ref_ldl_contrib_total": (_ci, [ctypes.c_void_p]),
            "ref_ldl_get_symbolic": (None, [ctypes.c_void_p] + [_P(_ci)] * 7),
            "ref_etree_post_colcount": (None, [ctypes.c_void_p] + [_P(_ci)] * 5),
            "ref_ldl_factor": (_ci, [ctypes.c_void_p, _P(_cd)]),
            "ref_ldl_inertia": (None, [ctypes.c_void_p, _P(_ci), _P(_ci)]),
            "ref_ldl_solve": (None, [ctypes.c_void_p, _ci, _P(_cd), _ci]),
            "ref_ldl_pivot_counts": (None, [ctypes.c_void_p, _P(_ci), _P(_ci)]),
            "ref_set_num_threads": (None, [_ci]),
            "ref_chol_create": (ctypes.c_void_p, [_ci, _P(_ci), _P(_ci), _ci]),
            "ref_chol_destroy": (None, [ctypes.c_void_p]),
            "ref_chol_factor": (_ci, [ctypes.c_void_p, _P(_cd)]),
            "ref_chol_solve": (None, [ctypes.c_void_p, _P(_cd)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ip(a):
    return a.ctypes.data_as(_P(_ci))


def _dp(a):
    return a.ctypes.data_as(_P(_cd))


ORDER = {"ND": 0, "NESTED_DISSECTION": 0, "AMD": 1, "MULTI_COLOR": 2, "NATURAL": 3}


def order(rowp, cols, kind="AMD"):
    rowp = np.ascontiguousarray(rowp, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    n = len(rowp) - 1
    perm = np.zeros(n, np.int32)
    iperm = np.zeros(n, np.int32)
    lib().ref_order(ORDER[kind], n, _ip(rowp), _ip(cols), _ip(perm), _ip(iperm))
    return perm, iperm


def element_csr(nrows, arrays, num_cons):
    """The reference's element-graph CSR (create_csr_from_element_conn as
    reorder_model calls it) of element arrays (E_g x npe_g node indices)
    with num_cons[g] constraint nodes last."""
    ne = np.array([a.shape[0] for a in arrays], np.int32)
    npe = np.array([a.shape[1] for a in arrays], np.int32)
    nc = np.ascontiguousarray(num_cons, np.int32)
    conn = np.ascontiguousarray(np.concatenate([np.ascontiguousarray(a, np.int32).ravel() for a in arrays]))
    rowp = np.zeros(nrows + 1, np.int32)
    cols = np.zeros(1, np.int32)
    nnz = lib().ref_element_csr(nrows, len(arrays), _ip(ne), _ip(npe), _ip(nc), _ip(conn), _ip(rowp), _ip(cols), 0)
    cols = np.zeros(max(nnz, 1), np.int32)
    lib().ref_element_csr(nrows, len(arrays), _ip(ne), _ip(npe), _ip(nc), _ip(conn), _ip(rowp), _ip(cols), nnz)
    return rowp, cols[:nnz]


def block_amd(rowp, cols, mult=()):
    """Raw BlockAMD::amd on the given pattern (multipliers last)."""
    rowp = np.ascontiguousarray(rowp, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    m = np.ascontiguousarray(mult, np.int32)
    n = len(rowp) - 1
    perm = np.zeros(n, np.int32)
    rc = lib().ref_block_amd(n, _ip(rowp), _ip(cols), len(m), _ip(m if m.size else np.zeros(1, np.int32)),
                             _ip(perm))
    if rc != 0:
        raise RuntimeError("reference BlockAMD failed")
    return perm


class RefLDL:
    """The reference SparseLDL (include/sparse_ldl.h) on the host."""

    def __init__(self, rowp, cols, ustab=0.01, pivot_tol=1e-14, delay_growth=2.0,
                 order="NATURAL", cholesky=False):
        self.rowp = np.ascontiguousarray(rowp, np.int32)
        self.cols = np.ascontiguousarray(cols, np.int32)
        self.n = len(self.rowp) - 1
        self.h = lib().ref_ldl_create(self.n, _ip(self.rowp), _ip(self.cols),
                                      1 if cholesky else 0, ustab, pivot_tol,
                                      delay_growth, ORDER[order])

    def __del__(self):
        # (at interpreter exit the module globals may already be gone)
        if getattr(self, "h", None) and lib is not None:
            lib().ref_ldl_destroy(self.h)
            self.h = None

    def symbolic(self):
        L = lib()
        K = L.ref_ldl_num_nodes(self.h)
        C = L.ref_ldl_contrib_total(self.h)
        names = ["node_ptr", "node_to_var", "children_ptr", "children",
                 "contrib_ptr", "contrib_rows", "invperm"]
        arrs = [np.zeros(s, np.int32) for s in (K + 1, self.n, K + 2, K, K + 1, C, self.n)]
        L.ref_ldl_get_symbolic(self.h, *[_ip(a) for a in arrs])
        return dict(zip(names, arrs))

    def etree(self, perm=None, iperm=None):
        n = self.n
        parent = np.zeros(n, np.int32)
        post = np.zeros(n, np.int32)
        cc = np.zeros(n, np.int32)
        if perm is None:
            perm = np.arange(n, dtype=np.int32)
            iperm = perm.copy()
        perm = np.ascontiguousarray(perm, np.int32)
        iperm = np.ascontiguousarray(iperm, np.int32)
        lib().ref_etree_post_colcount(self.h, _ip(perm), _ip(iperm), _ip(parent), _ip(post), _ip(cc))
        return parent, post, cc

    def factor(self, data):
        data = np.ascontiguousarray(data, np.float64)
        return lib().ref_ldl_factor(self.h, _dp(data))

    def inertia(self):
        p, q = _ci(), _ci()
        lib().ref_ldl_inertia(self.h, ctypes.byref(p), ctypes.byref(q))
        return p.value, q.value

    def solve(self, b):
        x = np.array(b, dtype=np.float64, order="F", copy=True)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        lib().ref_ldl_solve(self.h, nrhs, x.ctypes.data_as(_P(_cd)), self.n)
        return x

    def pivot_counts(self):
        K = lib().ref_ldl_num_nodes(self.h)
        a = np.zeros(K, np.int32)
        b = np.zeros(K, np.int32)
        lib().ref_ldl_pivot_counts(self.h, _ip(a), _ip(b))
        return a, b


class RefCholesky:
    """The reference SparseCholesky (include/sparse_cholesky.h:64-376)."""

    def __init__(self, rowp, cols, sqdef_index=-1):
        self.rowp = np.ascontiguousarray(rowp, np.int32)
        self.cols = np.ascontiguousarray(cols, np.int32)
        self.n = len(self.rowp) - 1
        self.h = lib().ref_chol_create(self.n, _ip(self.rowp), _ip(self.cols), sqdef_index)

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:
            lib().ref_chol_destroy(self.h)
            self.h = None

    def factor(self, data):
        data = np.ascontiguousarray(data, np.float64)
        return lib().ref_chol_factor(self.h, _dp(data))

    def solve(self, b):
        x = np.array(b, dtype=np.float64, copy=True)
        lib().ref_chol_solve(self.h, _dp(x))
        return x
