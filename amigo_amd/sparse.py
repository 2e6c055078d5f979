"""Device-resident CSR matrix and the SparseLDL / SparseCholesky drop-ins.

Mirrors the reference's Python surface (amigo/amigo.cpp:604-694):

    ldl = SparseLDL(mat, solver_type="LDL", ustab=0.01, pivot_tol=1e-14,
                    delay_growth=2.0, order="NATURAL")
    ldl.factor()            # -> 0
    ldl.solve(x)            # in place; x is a device tensor (n,) or (n, nrhs)
    ldl.get_inertia()       # -> (npos, nneg)

The symbolic analysis runs on the host in the constructor (as in
include/sparse_ldl.h:65-118); values and factors live in HBM.
"""

import ctypes

import numpy as np
import torch

from . import _lib

LDL = "LDL"
CHOLESKY = "CHOLESKY"

_ORDER = _lib.AMIGO_ORDER


def _order_code(order):
    if isinstance(order, int):
        return order
    key = str(order).upper().split(".")[-1]
    if key == "DEFAULT":
        key = "AMD"
    return _ORDER[key]


def current_stream_ptr():
    return torch.cuda.current_stream().cuda_stream


class CSRMat:
    """CSR matrix with the full symmetric pattern (sorted rows).

    The pattern is kept on the host (int32) and, once ``to_device`` is
    called, also on the device; values live in a float64 device tensor.
    """

    def __init__(self, nrows, ncols, rowp, cols, data=None, device="cuda", sqdef_index=-1):
        # First variable of the negative definite block of a quasi-definite
        # matrix (CSRMat's sqdef_index, include/csr_matrix.h:98-113; -1 = none).
        self.sqdef_index = int(sqdef_index)
        self.nrows = int(nrows)
        self.ncols = int(ncols)
        rowp = np.ascontiguousarray(rowp, dtype=np.int64)
        cols = np.ascontiguousarray(cols, dtype=np.int64)
        if rowp.shape != (self.nrows + 1,) or rowp[0] != 0:
            raise ValueError("rowp must have nrows + 1 entries starting at 0")
        # Sort each row (values follow) so that the pattern meets the
        # solver's requirement.
        order = np.lexsort((cols, np.repeat(np.arange(self.nrows), np.diff(rowp))))
        self.rowp = rowp.astype(np.int32)
        self.cols = cols[order].astype(np.int32)
        self.nnz = int(self.rowp[-1])
        self.device = torch.device(device)
        if data is None:
            vals = np.zeros(self.nnz)
        else:
            vals = np.asarray(data, dtype=np.float64)[order]
        self.data = torch.as_tensor(vals, dtype=torch.float64, device=self.device)
        self._diag = None
        self._rowp_d = None
        self._cols_d = None

    @classmethod
    def from_scipy(cls, A, device="cuda", sqdef_index=-1):
        A = A.tocsr()
        return cls(A.shape[0], A.shape[1], A.indptr, A.indices, A.data, device=device,
                   sqdef_index=sqdef_index)

    def diag_indices(self):
        """CSR position of each diagonal entry (-1 when absent)."""
        if self._diag is None:
            d = np.full(self.nrows, -1, dtype=np.int64)
            for i in range(self.nrows):
                s, e = self.rowp[i], self.rowp[i + 1]
                k = s + np.searchsorted(self.cols[s:e], i)
                if k < e and self.cols[k] == i:
                    d[i] = k
            self._diag = d
        return self._diag

    def pattern_device(self):
        if self._rowp_d is None:
            self._rowp_d = torch.as_tensor(self.rowp, device=self.device)
            self._cols_d = torch.as_tensor(self.cols, device=self.device)
        return self._rowp_d, self._cols_d

    def get_data(self):
        return self.data.cpu().numpy()

    def set_data(self, values):
        self.data.copy_(torch.as_tensor(np.asarray(values, dtype=np.float64)))

    def to_scipy(self):
        from scipy.sparse import csr_matrix

        return csr_matrix(
            (self.get_data(), self.cols, self.rowp), shape=(self.nrows, self.ncols)
        )


class SparseLDL:
    """Multifrontal LDL^T with symbolic analysis on the host and numeric
    factorization / solves on the GPU (include/sparse_ldl.h:19-206)."""

    def __init__(
        self,
        mat,
        solver_type=LDL,
        ustab=0.01,
        pivot_tol=1e-14,
        delay_growth=2.0,
        order="NATURAL",
        precision="fp64",
        groups=None,
        scaling=False,
        perm=None,
        _handle=None,
    ):
        self.mat = mat
        self.n = mat.nrows
        self.solver_type = CHOLESKY if str(solver_type).upper().endswith("CHOLESKY") else LDL
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        h = ctypes.c_void_p()
        gp = None
        if groups is not None:
            self._groups = np.ascontiguousarray(groups, dtype=np.int32)
            gp = _lib.ip(self._groups)
        if _handle is not None:
            h = _handle
        elif perm is not None:
            # Caller's permutation (perm[new] = old) in place of the ordering.
            self._perm = np.ascontiguousarray(perm, dtype=np.int32)
            _lib.check(
                _lib.lib().amigo_ldl_create_perm(
                    self.n,
                    _lib.ip(mat.rowp),
                    _lib.ip(mat.cols),
                    1 if self.solver_type == CHOLESKY else 0,
                    float(ustab),
                    float(pivot_tol),
                    float(delay_growth),
                    _lib.ip(self._perm),
                    ctypes.byref(h),
                )
            )
        if _handle is None and perm is None:
            _lib.check(
                _lib.lib().amigo_ldl_create_grouped(
                    self.n,
                    _lib.ip(mat.rowp),
                    _lib.ip(mat.cols),
                    1 if self.solver_type == CHOLESKY else 0,
                    float(ustab),
                    float(pivot_tol),
                    float(delay_growth),
                    _order_code(order),
                    gp,
                    ctypes.byref(h),
                )
            )
        self._nbatch = 0
        self._h = h
        self._factored = False
        if scaling:
            _lib.check(_lib.lib().amigo_ldl_set_scaling(self._h, 1))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().amigo_ldl_destroy(h)
            except Exception:
                pass
            self._h = None

    # ---- symbolic information -------------------------------------------
    def sizes(self):
        K, C, L, M = (ctypes.c_int() for _ in range(4))
        _lib.check(
            _lib.lib().amigo_ldl_sizes(
                self._h, ctypes.byref(K), ctypes.byref(C), ctypes.byref(L), ctypes.byref(M)
            )
        )
        return {"num_nodes": K.value, "contrib_total": C.value, "num_levels": L.value, "max_front": M.value}

    def get_symbolic(self):
        s = self.sizes()
        K, C, n = s["num_nodes"], s["contrib_total"], self.n
        arrs = [np.zeros(k, np.int32) for k in (K + 1, n, K + 2, K, K + 1, C, n)]
        _lib.check(_lib.lib().amigo_ldl_get_symbolic(self._h, *[_lib.ip(a) for a in arrs]))
        names = ["node_ptr", "node_to_var", "children_ptr", "children", "contrib_ptr", "contrib_rows", "invperm"]
        return dict(zip(names, arrs))

    @classmethod
    def union(cls, mat, blocks, precision="fp64", scaling=False):
        """SparseLDL of the block-diagonal union of the blocks' matrices (mat
        holds that union pattern): the blocks' symbolic plans are
        concatenated, so each block keeps its own fronts."""
        arr = (ctypes.c_void_p * len(blocks))(*[b._h.value for b in blocks])
        h = ctypes.c_void_p()
        _lib.check(
            _lib.lib().amigo_ldl_create_union(
                mat.nrows, _lib.ip(mat.rowp), _lib.ip(mat.cols), len(blocks), arr, ctypes.byref(h)
            )
        )
        b0 = blocks[0]
        return cls(mat, b0.solver_type, precision=precision, scaling=scaling, _handle=h)

    def get_perm(self):
        """The symmetric permutation (perm[new] = old) of the analysis."""
        return np.argsort(self.get_symbolic()["invperm"]).astype(np.int32)

    # ---- block-diagonal batches -----------------------------------------
    def set_batch(self, node_seg):
        """Independent diagonal blocks: block b owns nodes node_seg[b] ..
        node_seg[b+1] (contiguous).  Each factorization then also reports
        the inertia per block, and a non-finite pivot is confined to its
        block (get_batch_inertia) instead of failing the factorization."""
        seg = np.ascontiguousarray(node_seg, dtype=np.int32)
        _lib.check(_lib.lib().amigo_ldl_set_batch(self._h, len(seg) - 1, _lib.ip(seg)))
        self._nbatch = len(seg) - 1

    def set_batch_mask(self, mask=None):
        """Blocks the following factorizations and solves touch (bool per
        block; None = all).  Blocks left out keep their last factor."""
        if mask is None:
            _lib.check(_lib.lib().amigo_ldl_set_batch_mask(self._h, None, _lib.vp(current_stream_ptr())))
            return
        m = np.ascontiguousarray(np.asarray(mask, dtype=bool).astype(np.int32))
        _lib.check(_lib.lib().amigo_ldl_set_batch_mask(self._h, _lib.ip(m), _lib.vp(current_stream_ptr())))

    def get_batch_inertia(self):
        """(npos, nneg, nzero, nnan) arrays, one entry per block; npos counts
        zero pivots as positive, as get_inertia does."""
        B = self._nbatch
        a, b, c, d = (np.zeros(B, np.int32) for _ in range(4))
        _lib.check(
            _lib.lib().amigo_ldl_get_batch_inertia(self._h, _lib.ip(a), _lib.ip(b), _lib.ip(c), _lib.ip(d))
        )
        return a, b, c, d

    # ---- numeric --------------------------------------------------------
    def factor(self, shift=None, precision=None):
        """Numeric factorization of mat.data (+ optional diagonal shift)."""
        prec = precision or self.precision
        if prec not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        code = 1 if prec == "mixed" else 0
        vals = self.mat.data
        if shift is None:
            rc = _lib.lib().amigo_ldl_factor(
                self._h, _lib.vp(vals.data_ptr()), code, _lib.vp(current_stream_ptr())
            )
        else:
            assert shift.dtype == torch.float64 and shift.is_cuda and shift.numel() == self.n
            rc = _lib.lib().amigo_ldl_factor_shifted(
                self._h,
                _lib.vp(vals.data_ptr()),
                _lib.vp(shift.data_ptr()),
                code,
                _lib.vp(current_stream_ptr()),
            )
        _lib.check(rc)
        self._factored = True
        return 0

    def get_inertia(self):
        p, q, z = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(
            _lib.lib().amigo_ldl_get_inertia(self._h, ctypes.byref(p), ctypes.byref(q), ctypes.byref(z))
        )
        return p.value, q.value

    def get_inertia3(self):
        p, q, z = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(
            _lib.lib().amigo_ldl_get_inertia(self._h, ctypes.byref(p), ctypes.byref(q), ctypes.byref(z))
        )
        return p.value - z.value, q.value, z.value

    def pivot_counts(self):
        K = self.sizes()["num_nodes"]
        a = np.zeros(K, np.int32)
        b = np.zeros(K, np.int32)
        _lib.check(_lib.lib().amigo_ldl_pivot_counts(self._h, _lib.ip(a), _lib.ip(b)))
        return a, b

    def solve(self, x):
        """Solve in place.  x: float64 device tensor (n,) or column-major
        (n, nrhs) (i.e. x.t() contiguous), or a writeable float64 NumPy
        array (n,) or (n, nrhs) as the reference binding takes
        (amigo/amigo.cpp:627-689): an F-ordered array is overwritten
        directly, any other layout through an F-ordered copy that is then
        copied back.  Errors follow the binding: TypeError for a dtype
        other than float64, more than two dimensions or a read-only array,
        RuntimeError for any other object; a first dimension other than n
        raises ValueError (the reference reads past the array instead)."""
        if isinstance(x, np.ndarray):
            return self._solve_numpy(x)
        if not isinstance(x, torch.Tensor):
            raise RuntimeError("Expected either a NumPy array or a float64 device tensor")
        if not (x.is_cuda and x.dtype == torch.float64):
            raise TypeError("solve expects a float64 device tensor")
        if x.dim() == 1:
            if not x.is_contiguous():
                raise ValueError("x must be contiguous")
            nrhs, ldx = 1, self.n
        else:
            if x.shape[0] != self.n or x.stride(0) != 1:
                raise ValueError("2-D x must be column-major (n, nrhs)")
            nrhs, ldx = x.shape[1], x.stride(1)
        _lib.check(
            _lib.lib().amigo_ldl_solve(
                self._h, nrhs, _lib.vp(x.data_ptr()), ldx, _lib.vp(current_stream_ptr())
            )
        )
        return x

    def _solve_numpy(self, arr):
        if arr.dtype != np.float64:
            raise TypeError("NumPy array must have dtype=float64")
        if arr.ndim not in (1, 2):
            raise TypeError("NumPy array must be 1D or 2D")
        if not arr.flags.writeable:
            raise TypeError("NumPy array must be writeable")
        if arr.shape[0] != self.n:
            raise ValueError(f"array has {arr.shape[0]} rows, the matrix {self.n}")
        nrhs = 1 if arr.ndim == 1 else arr.shape[1]
        if nrhs == 0:
            return arr
        host = np.asfortranarray(arr)
        dev = torch.from_numpy(host.reshape(-1, order="F")).to(self.mat.device)
        _lib.check(
            _lib.lib().amigo_ldl_solve(self._h, nrhs, _lib.vp(dev.data_ptr()), self.n,
                                       _lib.vp(current_stream_ptr()))
        )
        out = dev.cpu().numpy().reshape(arr.shape, order="F")
        np.copyto(arr, out)
        return arr

    def stats(self):
        """(delayed pivots, largest front) of the last factorization."""
        d, m = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.lib().amigo_ldl_get_stats(self._h, ctypes.byref(d), ctypes.byref(m)))
        return d.value, m.value

    def set_timing(self, on=True, reset=False):
        _lib.check(_lib.lib().amigo_ldl_set_timing(self._h, 2 if (on and reset) else int(bool(on))))

    def timing(self):
        """{class: (ms, launches)} and the flops split (small, big)."""
        ms = (ctypes.c_double * 6)()
        cnt = (ctypes.c_longlong * 6)()
        fs, fb = ctypes.c_double(), ctypes.c_double()
        _lib.check(_lib.lib().amigo_ldl_get_timing(self._h, ms, cnt, ctypes.byref(fs), ctypes.byref(fb)))
        names = ["level", "panel", "schur_update", "solve_fwd", "solve_bwd", "assemble_big"]
        return {k: (ms[i], cnt[i]) for i, k in enumerate(names)}, (fs.value, fb.value)

    def set_graphs(self, on=True):
        """Replay captured HIP graphs of the factor / solve launch sequences."""
        _lib.check(_lib.lib().amigo_ldl_set_graphs(self._h, int(bool(on))))

    def memory_bytes(self):
        """Predicted device bytes of the numeric phase (pools, workspaces)."""
        b = ctypes.c_longlong()
        _lib.check(_lib.lib().amigo_ldl_memory(self._h, ctypes.byref(b)))
        return b.value

    def work(self):
        f, b = ctypes.c_double(), ctypes.c_double()
        _lib.check(_lib.lib().amigo_ldl_work(self._h, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value


class SparseCholesky(SparseLDL):
    """Drop-in for SparseCholesky<T> (include/sparse_cholesky.h:64-376): a
    factorization without pivoting, in natural order, of a symmetric
    quasi-definite matrix [A B^T; B -C] whose first ``mat.sqdef_index``
    variables form the positive definite block A (-1 or n: the whole matrix
    is positive definite).  ``factor()`` returns 0, or a positive value
    when a pivot has the wrong sign for its block, like the reference's
    failing-column return."""

    def __init__(self, mat):
        h = ctypes.c_void_p()
        _lib.check(
            _lib.lib().amigo_chol_create(
                mat.nrows, _lib.ip(mat.rowp), _lib.ip(mat.cols), int(getattr(mat, "sqdef_index", -1)),
                ctypes.byref(h),
            )
        )
        super().__init__(mat, solver_type=CHOLESKY, order="NATURAL", _handle=h)

    def factor(self):
        rc = _lib.lib().amigo_chol_factor(self._h, _lib.vp(self.mat.data.data_ptr()),
                                          _lib.vp(current_stream_ptr()))
        if rc < 0:
            _lib.check(rc)
        self._factored = True
        return int(rc)
