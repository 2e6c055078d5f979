"""Device-resident NLP built from component groups.

A ``ComponentProblem`` plays the role of the reference's OptimizationProblem
(include/optimization_problem.h:14-1373) for the interior-point path: it owns
the KKT pattern (create_matrix, optimization_problem.h:596-682), and
evaluates gradient, Hessian and Lagrangian (:718-818) of

    L(x, lambda) = alpha * f(x) + lambda^T c(x)

with HIP element kernels.  Like the reference, the "gradient" vector holds
dL/dx on primal rows and c(x) on multiplier rows, and the Hessian is the full
KKT matrix [[W, J^T], [J, 0]] (the multiplier diagonal stays in the pattern
so that -delta_c can be added).

Element contributions are summed into the gradient and the CSR values by
gather lists built once on the host, in a fixed order, so assembly is
deterministic and needs no atomics or per-entry searches.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .sparse import CSRMat, current_stream_ptr

ELEM_LINK = 0
ELEM_DI_INTERVAL = 1
ELEM_ELLIPTIC2D = 2
ELEM_ELLIPTIC3D = 3
ELEM_CARTPOLE = 4

NUM_INPUTS = {ELEM_LINK: 2, ELEM_DI_INTERVAL: 8, ELEM_ELLIPTIC2D: 7, ELEM_ELLIPTIC3D: 9, ELEM_CARTPOLE: 14}


def _mask_link():
    M = np.zeros((2, 2), bool)
    M[0, 1] = M[1, 0] = True
    return M


def _mask_di():
    M = np.zeros((8, 8), bool)
    M[2, 2] = M[5, 5] = True
    for q in (0, 1, 3, 4):
        M[6, q] = M[q, 6] = True
    for q in (1, 2, 4, 5):
        M[7, q] = M[q, 7] = True
    return M


def _mask_elliptic(dim):
    nb = 2 * dim
    nc = nb + 3
    M = np.zeros((nc, nc), bool)
    M[0, 0] = True
    M[nb + 1, nb + 1] = True
    for q in range(nb + 2):
        M[nb + 2, q] = M[q, nb + 2] = True
    M[nb + 2, nb + 2] = False
    return M


def _mask_cartpole():
    M = np.zeros((14, 14), bool)
    M[4, 4] = M[9, 9] = True
    for o in (0, 5):
        vi = [o + 1, o + 3, o + 4]
        for p in vi:
            for q in vi:
                M[p, q] = True
        for r in range(4):
            M[10 + r, o + r] = M[o + r, 10 + r] = True
        M[10, o + 2] = M[o + 2, 10] = True
        M[11, o + 3] = M[o + 3, 11] = True
        for r in (12, 13):
            for p in vi:
                M[r, p] = M[p, r] = True
    return M


STRUCTURE = {
    ELEM_LINK: _mask_link(),
    ELEM_DI_INTERVAL: _mask_di(),
    ELEM_ELLIPTIC2D: _mask_elliptic(2),
    ELEM_ELLIPTIC3D: _mask_elliptic(3),
    ELEM_CARTPOLE: _mask_cartpole(),
}


class ElementGroup:
    """E elements of one type: idx (E x NC, -1 = absent input), data (E x nd)."""

    def __init__(self, etype, idx, data):
        self.etype = int(etype)
        self.idx = np.ascontiguousarray(idx, dtype=np.int32)
        self.nc = NUM_INPUTS[self.etype]
        if self.idx.ndim != 2 or self.idx.shape[1] != self.nc:
            raise ValueError(f"element type {etype} needs {self.nc} inputs per element")
        self.E = self.idx.shape[0]
        data = np.asarray(data, dtype=np.float64)
        if data.ndim == 1:
            data = data.reshape(self.E, -1) if self.E else data.reshape(0, 1)
        self.data = np.ascontiguousarray(data)
        self.nd = self.data.shape[1] if self.data.size else 1


class ComponentProblem:
    """NLP on the device: variables x (n), multiplier indicator, bounds."""

    def __init__(self, n, groups, mult_ind, lower, upper, x0, device="cuda", name="problem",
                 var_groups=None):
        self.n = int(n)
        # Model node of every variable (used by the grouped ND ordering).
        self.var_groups = None if var_groups is None else np.asarray(var_groups, np.int32)
        self.name = name
        self.device = torch.device(device)
        self.groups = [g for g in groups if g.E > 0]
        self.mult_ind = np.asarray(mult_ind, dtype=bool)
        self.lower = np.asarray(lower, dtype=np.float64).copy()
        self.upper = np.asarray(upper, dtype=np.float64).copy()
        self.x0 = np.asarray(x0, dtype=np.float64).copy()
        assert self.mult_ind.shape == (self.n,)
        self._build_pattern()
        self._upload()

    # ---- pattern and gather maps ------------------------------------------
    def _build_pattern(self):
        n = self.n
        keys, hpos, gvar, gpos = [], [], [], []
        goff = hoff = 0
        for g in self.groups:
            nc = g.nc
            M = STRUCTURE[g.etype]
            qs, rs = np.nonzero(M)
            rows = g.idx[:, qs]
            cols = g.idx[:, rs]
            e = np.repeat(np.arange(g.E, dtype=np.int64)[:, None], len(qs), axis=1)
            ok = (rows >= 0) & (cols >= 0)
            # Structurally zero multiplier-multiplier couplings are dropped.
            ok &= ~(self.mult_ind[np.maximum(rows, 0)] & self.mult_ind[np.maximum(cols, 0)] & (rows != cols))
            k = rows.astype(np.int64) * n + cols
            keys.append(k[ok])
            hpos.append((hoff + e * nc * nc + (qs * nc + rs)[None, :])[ok])
            present = g.idx >= 0
            ee = np.repeat(np.arange(g.E, dtype=np.int64)[:, None], nc, axis=1)
            gvar.append(g.idx[present].astype(np.int64))
            gpos.append((goff + ee * nc + np.arange(nc)[None, :])[present])
            g.goff, g.hoff = goff, hoff
            goff += g.E * nc
            hoff += g.E * nc * nc
        self.gsize, self.hsize = goff, hoff
        diag = np.arange(n, dtype=np.int64) * (n + 1)
        allk = np.concatenate(keys + [diag]) if keys else diag
        ukeys = np.unique(allk)
        rows = (ukeys // n).astype(np.int64)
        cols = (ukeys % n).astype(np.int32)
        rowp = np.zeros(n + 1, np.int64)
        np.add.at(rowp, rows + 1, 1)
        rowp = np.cumsum(rowp).astype(np.int32)
        self.rowp, self.cols = rowp, cols
        self.nnz = len(cols)
        # Hessian gather: CSR slot -> buffer positions (fixed order).
        if keys:
            hk = np.concatenate(keys)
            hp = np.concatenate(hpos)
            slot = np.searchsorted(ukeys, hk)
            order = np.lexsort((hp, slot))
            self.h_src = hp[order].astype(np.int64)
            self.h_ptr = np.concatenate([[0], np.cumsum(np.bincount(slot, minlength=self.nnz))]).astype(np.int32)
            gv = np.concatenate(gvar)
            gp = np.concatenate(gpos)
            order = np.lexsort((gp, gv))
            self.g_src = gp[order].astype(np.int64)
            self.g_ptr = np.concatenate([[0], np.cumsum(np.bincount(gv, minlength=n))]).astype(np.int32)
        else:
            self.h_src = np.zeros(0, np.int64)
            self.h_ptr = np.zeros(self.nnz + 1, np.int32)
            self.g_src = np.zeros(0, np.int64)
            self.g_ptr = np.zeros(n + 1, np.int32)
        d = np.full(n, -1, np.int64)
        r = np.repeat(np.arange(n), np.diff(rowp))
        isd = cols == r
        d[r[isd]] = np.nonzero(isd)[0]
        self.diag_idx = d

    def _upload(self):
        dev = self.device
        T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
        for g in self.groups:
            g.idx_d = T(g.idx, torch.int32)
            g.data_d = T(g.data if g.data.size else np.zeros((max(g.E, 1), 1)), torch.float64)
        self.h_src_d = T(self.h_src, torch.int64)
        self.h_ptr_d = T(self.h_ptr, torch.int32)
        self.g_src_d = T(self.g_src, torch.int64)
        self.g_ptr_d = T(self.g_ptr, torch.int32)
        self.gbuf = torch.zeros(max(self.gsize, 1), dtype=torch.float64, device=dev)
        self.hbuf = torch.zeros(max(self.hsize, 1), dtype=torch.float64, device=dev)
        nblk = sum((g.E + 255) // 256 for g in self.groups)
        self.lpart = torch.zeros(max(nblk, 1), dtype=torch.float64, device=dev)
        self.lout = torch.zeros(1, dtype=torch.float64, device=dev)
        self.diag_idx_d = T(self.diag_idx, torch.int64)

    # ---- OptimizationProblem surface ---------------------------------------
    def get_num_variables(self):
        return self.n

    def get_multiplier_indicator(self):
        return self.mult_ind

    def create_vector(self):
        return torch.zeros(self.n, dtype=torch.float64, device=self.device)

    def create_matrix(self):
        return CSRMat(self.n, self.n, self.rowp, self.cols, None, device=self.device)

    def update(self, x):
        pass

    def _eval(self, alpha, x, want_g, want_h, want_l):
        L = _lib.lib()
        s = _lib.vp(current_stream_ptr())
        boff = 0
        for g in self.groups:
            gb = _lib.vp(self.gbuf.data_ptr() + 8 * g.goff) if want_g else _lib.vp(0)
            hb = _lib.vp(self.hbuf.data_ptr() + 8 * g.hoff) if want_h else _lib.vp(0)
            lp = _lib.vp(self.lpart.data_ptr() + 8 * boff) if want_l else _lib.vp(0)
            _lib.check(
                L.amigo_elements_eval(
                    g.etype, g.E, _lib.dptr(g.idx_d), _lib.dptr(g.data_d), g.nd, _lib.dptr(x),
                    float(alpha), gb, hb, lp, s,
                ),
                "asm",
            )
            boff += (g.E + 255) // 256
        return boff

    def gradient(self, alpha, x, g):
        """g = [alpha grad f + J^T lambda ; c(x)] (optimization_problem.h:742-762)."""
        self._eval(alpha, x, True, False, False)
        _lib.check(
            _lib.lib().amigo_gather_sum(
                self.n, _lib.dptr(self.g_ptr_d), _lib.dptr(self.g_src_d), _lib.dptr(self.gbuf),
                _lib.dptr(g), 0, _lib.vp(current_stream_ptr()),
            ),
            "asm",
        )

    def hessian(self, alpha, x, mat):
        """KKT values of L (optimization_problem.h:801-818) into mat.data."""
        self._eval(alpha, x, False, True, False)
        _lib.check(
            _lib.lib().amigo_gather_sum(
                self.nnz, _lib.dptr(self.h_ptr_d), _lib.dptr(self.h_src_d), _lib.dptr(self.hbuf),
                _lib.dptr(mat.data), 0, _lib.vp(current_stream_ptr()),
            ),
            "asm",
        )

    def lagrangian(self, alpha, x):
        """alpha f(x) + lambda^T c(x) (optimization_problem.h:718-733)."""
        nb = self._eval(alpha, x, False, False, True)
        out = ctypes.c_double()
        _lib.check(
            _lib.lib().amigo_sum_partials(
                _lib.dptr(self.lpart), nb, _lib.dptr(self.lout), ctypes.byref(out),
                _lib.vp(current_stream_ptr()),
            ),
            "asm",
        )
        return out.value

    # ---- helpers -----------------------------------------------------------
    def primal_indices(self):
        return np.nonzero(~self.mult_ind)[0].astype(np.int32)

    def constraint_indices(self):
        return np.nonzero(self.mult_ind)[0].astype(np.int32)
