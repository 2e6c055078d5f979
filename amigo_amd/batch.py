"""Many independent NLPs solved in lockstep on one GPU.

    bopt = BatchOptimizer([prob_0, ..., prob_{K-1}], order="ND", scaling=True)
    results = bopt.optimize({"record_objective": False})   # K result dicts
    x_k = bopt.get_solution(k)

The K problems are laid end to end as one block-diagonal NLP:

* ``BatchProblem``: the union of the problems' element groups (one group per
  element type, each problem's elements padded to whole blocks of 256 so
  that every block belongs to one problem), their KKT patterns and gather
  lists concatenated with offsets.  One element launch per element type
  evaluates every problem; each block scales the objective by its own
  problem's objective scale.
* ``BatchSolver``: one SparseLDL of the block-diagonal KKT whose symmetric
  permutation is the concatenation of each problem's own ordering, so each
  problem keeps exactly the fronts it has alone.  Every tree level of every
  problem runs in the same launches (one level kernel launch per
  elimination-tree level for the whole batch), and the inertia is counted
  per problem (amigo_ldl_set_batch).
* ``BatchIPM``: the ipm:: kernels over all problems at once
  (amigo_ipm_batch_*), with per-problem scalars and a mask of the problems
  a call touches.  Each problem's reductions sum the same partials in the
  same order as alone, so the batch reproduces every problem's
  single-problem arithmetic bit for bit.
* ``BatchOptimizer``: the interior-point driver of ``optimizer.Optimizer``
  (which restates amigo/algorithm/ipm_driver.py:41-483) with every
  per-problem scalar held in arrays: NLP scaling and least-squares
  multipliers, the LOQO barrier heuristic, inertia correction
  (inertia_correction.py:261-393), iterative refinement
  (linear_solver.py:109-264), the filter line search with second-order
  correction and watchdog (filter_line_search.py:48-438), slack
  safeguards and the convergence tests.  A problem that needs the
  feasibility-restoration phase (feasibility_restoration.py:16-172) leaves
  the batch and is solved again on its own by ``Optimizer``.

The reference solves one problem per InteriorPointOptimizer; it has no
batched counterpart.
"""

import ctypes
import time

import numpy as np
import torch

from . import _lib
from .default_options import get_default_options
from .ipm import RED
from .optimizer import (CONTINUE, CONVERGED, CONVERGED_ACCEPTABLE, DIVERGED, PRECISION_FLOOR,
                        Optimizer)
from .problem import NUM_INPUTS
from .sparse import CSRMat, SparseLDL, current_stream_ptr

NT = 256
ELEM_PAD = -2


def _s():
    return _lib.vp(current_stream_ptr())


def _p(t):
    return _lib.dptr(t)


# ---------------------------------------------------------------------------
# Union problem
# ---------------------------------------------------------------------------
class BatchProblem:
    """K ComponentProblems as one block-diagonal NLP (see the module doc)."""

    def __init__(self, problems, device=None):
        self.problems = list(problems)
        K = self.K = len(self.problems)
        if K == 0:
            raise ValueError("a batch needs at least one problem")
        dev = self.device = torch.device(device) if device is not None else self.problems[0].device
        ns = np.array([p.n for p in self.problems], np.int64)
        xseg = self.xseg = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
        self.n = int(xseg[-1])
        if self.n >= 2**31:
            raise ValueError("batch too large for 32-bit indices")
        self.mult_ind = np.concatenate([p.mult_ind for p in self.problems])
        self.lower = np.concatenate([p.lower for p in self.problems])
        self.upper = np.concatenate([p.upper for p in self.problems])
        self.x0 = np.concatenate([p.x0 for p in self.problems])
        self.name = f"batch_of_{K}"
        self._build_groups()
        self._build_pattern()
        self._upload()
        self.alpha_k = torch.ones(K, dtype=torch.float64, device=dev)
        self._alpha_scalar = 1.0
        self._set_alpha_blocks()

    # ---- element groups: one per (type, occurrence), problems padded -------
    def _build_groups(self):
        K, xseg = self.K, self.xseg
        keys, per = [], []
        for p in self.problems:
            cnt, kl = {}, []
            for g in p.groups:
                j = cnt.get(g.etype, 0)
                cnt[g.etype] = j + 1
                key = (g.etype, j)
                kl.append(key)
                if key not in keys:
                    keys.append(key)
            per.append(dict(zip(kl, p.groups)))
        self.ukeys = keys
        self.ugroups = []
        # (union group, first element, elements) of every problem group,
        # keyed by (problem, key); element positions feed the gather remap.
        self.place = {}
        for gi, key in enumerate(keys):
            etype = key[0]
            nc = NUM_INPUTS[etype]
            idx_parts, data_parts, bprob = [], [], []
            e0 = 0
            nd = None
            for k in range(K):
                g = per[k].get(key)
                if g is None:
                    continue
                nb = (g.E + NT - 1) // NT
                pad = nb * NT - g.E
                idx = g.idx.astype(np.int64)
                idx = np.where(idx >= 0, idx + xseg[k], -1)
                data = g.data if g.data.size else np.zeros((g.E, 1))
                if nd is None:
                    nd = data.shape[1]
                if data.shape[1] != nd:
                    raise ValueError("element data widths differ within one element type")
                if pad:
                    pi = np.full((pad, nc), -1, np.int64)
                    pi[:, 0] = ELEM_PAD
                    idx = np.vstack([idx, pi])
                    data = np.vstack([data, np.zeros((pad, nd))])
                idx_parts.append(idx)
                data_parts.append(data)
                bprob.append(np.full(nb, k, np.int64))
                self.place[(k, key)] = (gi, e0, g)
                e0 += nb * NT
            G = _UGroup(etype, nc, np.vstack(idx_parts).astype(np.int32),
                        np.ascontiguousarray(np.vstack(data_parts)), np.concatenate(bprob))
            self.ugroups.append(G)
        goff = hoff = boff = 0
        for G in self.ugroups:
            G.goff, G.hoff, G.boff = goff, hoff, boff
            goff += G.E * G.nc
            hoff += G.E * G.nc * G.nc
            boff += G.E // NT
        self.gsize, self.hsize, self.nblk = goff, hoff, boff
        # Lagrangian partial lists: problem k's blocks, its groups in its order.
        lidx, lptr = [], [0]
        for k, p in enumerate(self.problems):
            cnt = {}
            for g in p.groups:
                j = cnt.get(g.etype, 0)
                cnt[g.etype] = j + 1
                gi, e0, _ = self.place[(k, (g.etype, j))]
                G = self.ugroups[gi]
                nb = (g.E + NT - 1) // NT
                lidx.append(G.boff + e0 // NT + np.arange(nb))
            lptr.append(lptr[-1] + sum(len(a) for a in lidx[len(lidx) - len(p.groups):]))
        self.lidx = np.concatenate(lidx).astype(np.int32) if lidx else np.zeros(0, np.int32)
        self.lptr = np.asarray(lptr, np.int32)

    def _remap(self, k, p, src, per_elem):
        """Positions in problem k's own g/h buffers -> union buffer positions."""
        out = np.empty_like(src)
        cnt = {}
        for g in p.groups:
            j = cnt.get(g.etype, 0)
            cnt[g.etype] = j + 1
            gi, e0, _ = self.place[(k, (g.etype, j))]
            G = self.ugroups[gi]
            lo = g.goff if per_elem == 1 else g.hoff
            size = g.E * (g.nc if per_elem == 1 else g.nc * g.nc)
            sel = (src >= lo) & (src < lo + size)
            base = G.goff if per_elem == 1 else G.hoff
            stride = g.nc if per_elem == 1 else g.nc * g.nc
            out[sel] = base + e0 * stride + (src[sel] - lo)
        return out

    def _build_pattern(self):
        rowp, cols, hp, hs, gp, gs, dg = [np.zeros(1, np.int64)], [], [np.zeros(1, np.int64)], [], [np.zeros(1, np.int64)], [], []
        nnz0 = 0
        for k, p in enumerate(self.problems):
            x0 = int(self.xseg[k])
            rowp.append(p.rowp[1:].astype(np.int64) + nnz0)
            cols.append(p.cols.astype(np.int64) + x0)
            hp.append(p.h_ptr[1:].astype(np.int64) + hp[-1][-1])
            hs.append(self._remap(k, p, p.h_src, 2))
            gp.append(p.g_ptr[1:].astype(np.int64) + gp[-1][-1])
            gs.append(self._remap(k, p, p.g_src, 1))
            dg.append(np.where(p.diag_idx >= 0, p.diag_idx + nnz0, -1))
            nnz0 += p.nnz
        self.nnz = nnz0
        if nnz0 >= 2**31:
            raise ValueError("batch too large for 32-bit CSR offsets")
        self.rowp = np.concatenate(rowp).astype(np.int32)
        self.cols = np.concatenate(cols).astype(np.int32)
        self.h_ptr = np.concatenate(hp).astype(np.int32)
        self.h_src = np.concatenate(hs).astype(np.int64)
        self.g_ptr = np.concatenate(gp).astype(np.int32)
        self.g_src = np.concatenate(gs).astype(np.int64)
        self.diag_idx = np.concatenate(dg).astype(np.int64)

    def _upload(self):
        dev = self.device
        T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
        for G in self.ugroups:
            G.idx_d = T(G.idx, torch.int32)
            G.data_d = T(G.data, torch.float64)
            G.bprob_d = T(G.bprob, torch.int64)
            G.bprob32_d = T(G.bprob, torch.int32)
            G.alpha_d = torch.ones(G.E // NT, dtype=torch.float64, device=dev)
        self.h_src_d = T(self.h_src, torch.int64)
        self.h_ptr_d = T(self.h_ptr, torch.int32)
        self.g_src_d = T(self.g_src, torch.int64)
        self.g_ptr_d = T(self.g_ptr, torch.int32)
        self.gbuf = torch.zeros(max(self.gsize, 1), dtype=torch.float64, device=dev)
        self.hbuf = torch.zeros(max(self.hsize, 1), dtype=torch.float64, device=dev)
        self.lpart = torch.zeros(max(self.nblk, 1), dtype=torch.float64, device=dev)
        self.lout = torch.zeros(self.K, dtype=torch.float64, device=dev)
        self.lptr_d = T(self.lptr, torch.int32)
        self.lidx_d = T(self.lidx if self.lidx.size else np.zeros(1, np.int32), torch.int32)
        self.diag_idx_d = T(self.diag_idx, torch.int64)
        self._lhost = np.zeros(self.K)
        # Output groups of the masked evaluations: gradient rows and Hessian
        # slots of each problem (contiguous), and the device mask.
        self.xseg64_d = T(self.xseg, torch.int64)
        self.hseg64_d = T(self.rowp[self.xseg].astype(np.int64), torch.int64)
        self._act_d = torch.zeros(self.K, dtype=torch.int32, device=dev)
        self._act_last = None

    # ---- per-problem objective scale ---------------------------------------
    def set_obj_scales(self, alpha):
        """Objective scale of every problem (K values)."""
        self.alpha_k.copy_(torch.as_tensor(np.asarray(alpha, np.float64)))
        self._set_alpha_blocks()

    def _set_alpha_blocks(self):
        for G in self.ugroups:
            G.alpha_d.copy_(self.alpha_k[G.bprob_d])

    # ---- OptimizationProblem surface (per-problem objective scales) --------
    def create_vector(self):
        return torch.zeros(self.n, dtype=torch.float64, device=self.device)

    def create_matrix(self):
        return CSRMat(self.n, self.n, self.rowp, self.cols, None, device=self.device)

    def get_multiplier_indicator(self):
        return self.mult_ind

    def active_dev(self, mask):
        """Device copy of a K-problem mask (None when every problem is in),
        uploaded only when it changes."""
        if mask is None or bool(np.all(mask)):
            return None
        m = np.ascontiguousarray(mask, dtype=np.int32)
        if self._act_last is None or not np.array_equal(m, self._act_last):
            self._act_d.copy_(torch.from_numpy(m))
            self._act_last = m.copy()
        return self._act_d

    def _eval(self, scaled, x, want_g, want_h, want_l, mask=None):
        """scaled: True -> each problem's objective scale; a float -> that
        scale for every problem (as the single-problem call passes it).
        mask: only those problems' elements (the others write nothing)."""
        L = _lib.lib()
        s = _s()
        act = self.active_dev(mask)
        for G in self.ugroups:
            gb = _lib.vp(self.gbuf.data_ptr() + 8 * G.goff) if want_g else _lib.vp(0)
            hb = _lib.vp(self.hbuf.data_ptr() + 8 * G.hoff) if want_h else _lib.vp(0)
            lp = _lib.vp(self.lpart.data_ptr() + 8 * G.boff) if want_l else _lib.vp(0)
            if scaled is True:
                alpha, ab = 1.0, _p(G.alpha_d)
            else:
                alpha, ab = float(scaled), _lib.vp(0)
            if act is None:
                rc = L.amigo_elements_eval_scaled(G.etype, G.E, _p(G.idx_d), _p(G.data_d), G.nd, _p(x),
                                                  alpha, ab, gb, hb, lp, s)
            else:
                rc = L.amigo_elements_eval_masked(G.etype, G.E, _p(G.idx_d), _p(G.data_d), G.nd, _p(x),
                                                  alpha, ab, _p(G.bprob32_d), _p(act), gb, hb, lp, s)
            _lib.check(rc, "asm")

    def gradient(self, scaled, x, g, mask=None):
        """g for the problems of mask (all when None; the other rows keep
        their values)."""
        self._eval(scaled, x, True, False, False, mask)
        act = self.active_dev(mask)
        if act is None:
            rc = _lib.lib().amigo_gather_sum(self.n, _p(self.g_ptr_d), _p(self.g_src_d), _p(self.gbuf),
                                             _p(g), 0, _s())
        else:
            rc = _lib.lib().amigo_gather_sum_masked(self.n, _p(self.g_ptr_d), _p(self.g_src_d), _p(self.gbuf),
                                                    _p(g), 0, _p(self.xseg64_d), self.K, _p(act), _s())
        _lib.check(rc, "asm")

    def hessian(self, scaled, x, mat, mask=None):
        """Hessian values of the problems of mask (all when None; the other
        problems' entries keep their values)."""
        self._eval(scaled, x, False, True, False, mask)
        act = self.active_dev(mask)
        if act is None:
            rc = _lib.lib().amigo_gather_sum(self.nnz, _p(self.h_ptr_d), _p(self.h_src_d), _p(self.hbuf),
                                             _p(mat.data), 0, _s())
        else:
            rc = _lib.lib().amigo_gather_sum_masked(self.nnz, _p(self.h_ptr_d), _p(self.h_src_d),
                                                    _p(self.hbuf), _p(mat.data), 0, _p(self.hseg64_d),
                                                    self.K, _p(act), _s())
        _lib.check(rc, "asm")

    def lagrangian(self, scaled, x):
        """Per-problem alpha_k f_k + lambda_k^T c_k (K values, host)."""
        self._eval(scaled, x, False, False, True)
        _lib.check(_lib.lib().amigo_sum_partials_batch(
            _p(self.lpart), self.K, _p(self.lptr_d), _p(self.lidx_d), _p(self.lout),
            self._lhost.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _s()), "asm")
        return self._lhost.copy()


class _UGroup:
    def __init__(self, etype, nc, idx, data, bprob):
        self.etype, self.nc, self.idx, self.data, self.bprob = etype, nc, idx, data, bprob
        self.E = idx.shape[0]
        self.nd = data.shape[1]


# ---------------------------------------------------------------------------
# Union KKT solver
# ---------------------------------------------------------------------------
_BLOCK_CACHE = {}


def problem_ldl(p, order, ustab=0.01, pivot_tol=1e-14, delay_growth=2.0):
    """Host-side SparseLDL analysis of p as the single-problem solver runs
    it (cached by sparsity structure, ordering and parameters)."""
    vg = getattr(p, "var_groups", None)
    key = (order, ustab, pivot_tol, delay_growth, p.n, p.nnz, hash(p.rowp.tobytes()),
           hash(p.cols.tobytes()), None if vg is None else hash(np.asarray(vg).tobytes()))
    ldl = _BLOCK_CACHE.get(key)
    if ldl is None:
        m = CSRMat(p.n, p.n, p.rowp, p.cols, None, device="cpu")
        ldl = SparseLDL(m, "LDL", ustab=ustab, pivot_tol=pivot_tol, delay_growth=delay_growth,
                        order=order, groups=vg)
        if len(_BLOCK_CACHE) > 4096:
            _BLOCK_CACHE.clear()
        _BLOCK_CACHE[key] = ldl
    return ldl


class BatchSolver:
    """The block-diagonal KKT of a BatchProblem on one SparseLDL, with the
    inertia counted per problem."""

    def __init__(self, bp, order="ND", ustab=0.01, pivot_tol=1e-14, delay_growth=2.0,
                 scaling=False, precision="fp64"):
        self.bp = bp
        self.hess = bp.create_matrix()
        self.nrows = bp.n
        blocks = [problem_ldl(p, order, ustab, pivot_tol, delay_growth) for p in bp.problems]
        self.ldl = SparseLDL.union(self.hess, blocks, precision=precision, scaling=scaling)
        self.precision = precision
        sym = self.ldl.get_symbolic()
        first = sym["node_to_var"][sym["node_ptr"][:-1]]
        owner = np.searchsorted(bp.xseg, first, side="right") - 1
        if np.any(np.diff(owner) < 0):
            raise RuntimeError("batch symbolic analysis interleaved the problems' nodes")
        self.node_seg = np.searchsorted(owner, np.arange(bp.K + 1), side="left").astype(np.int32)
        self.ldl.set_batch(self.node_seg)
        self._mask_all = True
        self.n_factor_problems = 0
        self.shift = torch.zeros(self.nrows, dtype=torch.float64, device=bp.device)
        self._tmp = torch.zeros(self.nrows, dtype=torch.float64, device=bp.device)
        self.n_factor = 0
        self.n_solve = 0

    def _mask(self, mask):
        key = None if mask is None or bool(np.all(mask)) else np.asarray(mask, bool)
        if key is None and self._mask_all:
            return
        self.ldl.set_batch_mask(key)
        self._mask_all = key is None

    def factor(self, scaled, x, diag, post_hessian=None, mask=None):
        self.bp.hessian(scaled, x, self.hess)
        if post_hessian is not None:
            post_hessian(self.hess)
        self.shift.copy_(diag)
        self._mask(mask)
        self.ldl.factor(shift=self.shift)
        self.n_factor += 1

    def assemble_hessian(self, scaled, x, mask=None):
        self.bp.hessian(scaled, x, self.hess, mask)

    def add_diagonal_and_factor(self, diag, mask=None):
        """Factor the problems of mask (all when None); the others keep
        their last factor."""
        self.shift.copy_(diag)
        self._mask(mask)
        self.ldl.factor(shift=self.shift)
        self.n_factor += 1
        self.n_factor_problems += int(self.bp.K if mask is None else np.count_nonzero(mask))

    def refactor_fp64(self, mask):
        """FP64 factorization of the masked problems' current matrices (the
        mixed-precision form's fallback)."""
        self._mask(mask)
        self.ldl.factor(shift=self.shift, precision="fp64")
        self.n_factor += 1

    def inertia(self):
        return self.ldl.get_batch_inertia()

    def solve(self, bx, px, mask=None):
        """px = K^-1 bx for the problems of mask; the others' entries of px
        are bx's."""
        px.copy_(bx)
        self._mask(mask)
        self.ldl.solve(px)
        self.n_solve += 1

    def kkt_matvec(self, x, y, mask=None):
        """y = (H + diag(shift)) x, for the rows of the problems of mask (all
        when None; the other rows of y are not defined)."""
        rowp_d, cols_d = self.hess.pattern_device()
        if mask is None or bool(np.all(mask)):
            _lib.check(_lib.lib().amigo_csr_spmv(self.nrows, _p(rowp_d), _p(cols_d), _p(self.hess.data),
                                                 _p(x), _p(y), _s()), "ipm")
        else:
            if getattr(self, "_row_prob", None) is None:
                K = self.bp.K
                self._row_prob = torch.as_tensor(np.repeat(np.arange(K, dtype=np.int32), np.diff(self.bp.xseg)),
                                                 dtype=torch.int32, device=self.shift.device)
                self._mv_act = torch.zeros(K, dtype=torch.int32, device=self.shift.device)
                self._mv_last = None
            m = np.ascontiguousarray(mask, dtype=np.int32)
            if self._mv_last is None or not np.array_equal(m, self._mv_last):
                self._mv_act.copy_(torch.from_numpy(m))
                self._mv_last = m.copy()
            _lib.check(_lib.lib().amigo_csr_spmv_masked(
                self.nrows, _p(rowp_d), _p(cols_d), _p(self.hess.data), _p(x), _p(y),
                _p(self._row_prob), _p(self._mv_act), _s()), "ipm")
        y.addcmul_(self.shift, x)


# ---------------------------------------------------------------------------
# Batched interior-point kernels
# ---------------------------------------------------------------------------
class BVec:
    """x (primals + multipliers), zl, zu, sl, su of every problem."""

    def __init__(self, n, np_, dev, x=None):
        self.x = x if x is not None else torch.zeros(n, dtype=torch.float64, device=dev)
        self.zl = torch.zeros(np_, dtype=torch.float64, device=dev)
        self.zu = torch.zeros(np_, dtype=torch.float64, device=dev)
        self.sl = torch.ones(np_, dtype=torch.float64, device=dev)
        self.su = torch.ones(np_, dtype=torch.float64, device=dev)
        self._st = _lib.IpmState(self.x.data_ptr(), self.zl.data_ptr(), self.zu.data_ptr(),
                                 self.sl.data_ptr(), self.su.data_ptr())

    def state(self):
        return ctypes.byref(self._st)


class BatchIPM:
    """ipm:: kernels over a BatchProblem (amigo_ipm_batch_*)."""

    def __init__(self, bp, lower, upper):
        self.bp = bp
        K = self.K = bp.K
        dev = self.device = bp.device
        self.n = bp.n
        p_loc, c_loc, pseg, cseg, pprob, cprob = [], [], [0], [0], [], []
        for k, p in enumerate(bp.problems):
            pi = np.nonzero(~p.mult_ind)[0].astype(np.int32)
            ci = np.nonzero(p.mult_ind)[0].astype(np.int32)
            p_loc.append(pi)
            c_loc.append(ci)
            pseg.append(pseg[-1] + len(pi))
            cseg.append(cseg[-1] + len(ci))
            pprob.append(np.full(len(pi), k))
            cprob.append(np.full(len(ci), k))
        self.pseg = np.asarray(pseg, np.int64)
        self.cseg = np.asarray(cseg, np.int64)
        self.np_ = int(self.pseg[-1])
        self.nc_ = int(self.cseg[-1])
        xs = bp.xseg
        pl = np.concatenate(p_loc)
        cl = np.concatenate(c_loc)
        pprob = np.concatenate(pprob).astype(np.int64)
        cprob = np.concatenate(cprob).astype(np.int64)
        self.pidx_g_h = (pl + xs[pprob]).astype(np.int64)
        self.cidx_g_h = (cl + xs[cprob]).astype(np.int64)
        T = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
        self.pidx = T(pl, torch.int32)
        self.cidx = T(cl, torch.int32)
        self.pidx_g = T(self.pidx_g_h, torch.int64)
        self.cidx_g = T(self.cidx_g_h, torch.int64)
        self.xprob = T(np.repeat(np.arange(K), np.diff(xs)), torch.int64)
        self.pprob = T(pprob, torch.int64)
        self.cprob = T(cprob, torch.int64)
        lower = np.asarray(lower, np.float64)
        upper = np.asarray(upper, np.float64)
        self.lbx0_h = lower[self.pidx_g_h].copy()
        self.ubx0_h = upper[self.pidx_g_h].copy()
        self.lbx0 = T(self.lbx0_h)
        self.ubx0 = T(self.ubx0_h)
        self.lbx = self.lbx0.clone()
        self.ubx = self.ubx0.clone()
        self.lbh0 = T(lower[self.cidx_g_h])
        self.lbh = self.lbh0.clone()
        self.is_mult = T(bp.mult_ind, torch.bool)
        self.n_primal_k = np.diff(self.pseg)
        self.n_constr_k = np.diff(self.cseg)
        # Block tables: each problem gets the grid it gets alone.
        L = _lib.lib()
        ct = np.array([L.amigo_ipm_grid_blocks(int(a + b)) for a, b in zip(self.n_primal_k, self.n_constr_k)],
                      np.int32)
        cp = np.array([L.amigo_ipm_grid_blocks(int(a)) for a in self.n_primal_k], np.int32)
        self._tabs = []
        for c in (ct, cp):
            first = np.concatenate([[0], np.cumsum(c)[:-1]]).astype(np.int32)
            blk = np.repeat(np.arange(K, dtype=np.int32), c)
            self._tabs.append((int(c.sum()), T(blk, torch.int32), T(first, torch.int32), T(c, torch.int32)))
        self.pseg_d = T(self.pseg, torch.int32)
        self.cseg_d = T(self.cseg, torch.int32)
        self.xseg_d = T(xs, torch.int32)
        # The Hessian buffer of the NLP scaling (pattern built once, here).
        self._scal_hess = bp.create_matrix()
        self._segmax_d = torch.zeros(K, dtype=torch.float64, device=dev)
        self._segmax_h = np.zeros(K)
        nbmax = max(self._tabs[0][0], self._tabs[1][0])
        self.ws = torch.zeros(4 * nbmax + 16, dtype=torch.float64, device=dev)
        self.out_d = torch.zeros(4 * K, dtype=torch.float64, device=dev)
        self.out_h = np.zeros(4 * K)
        self.act_d = torch.ones(K, dtype=torch.int32, device=dev)
        self._sc = [torch.zeros(K, dtype=torch.float64, device=dev) for _ in range(4)]
        # Host copies of the last uploads: an upload that would write the
        # same bits is skipped (each upload from pageable memory waits for
        # the stream).
        self._act_last = None
        self._sc_last = [None] * 4
        self.constr_scale = torch.ones(self.nc_, dtype=torch.float64, device=dev)
        self.scale_vec = torch.ones(self.n, dtype=torch.float64, device=dev)
        self.scaling_active = False
        self.obj_scale = np.ones(K)
        self._set_info()
        self._batch = _lib.IpmBatch()
        self._fill_batch(None)

    def _set_info(self):
        self._info = _lib.IpmProblem(self.np_, self.nc_, self.pidx.data_ptr(), self.cidx.data_ptr(),
                                     self.lbx.data_ptr(), self.ubx.data_ptr(), self.lbh.data_ptr())
        self.info = ctypes.byref(self._info)

    def _fill_batch(self, mask):
        b = self._batch
        (nt, bt, ft, ct), (npb, bp_, fp, cp) = self._tabs
        b.num_problems = self.K
        b.pseg, b.cseg, b.xseg = self.pseg_d.data_ptr(), self.cseg_d.data_ptr(), self.xseg_d.data_ptr()
        b.nblk_total, b.blk_total, b.first_total, b.count_total = nt, bt.data_ptr(), ft.data_ptr(), ct.data_ptr()
        b.nblk_primal, b.blk_primal, b.first_primal, b.count_primal = npb, bp_.data_ptr(), fp.data_ptr(), cp.data_ptr()
        if mask is None:
            b.active = None
        else:
            m = np.ascontiguousarray(mask, dtype=np.int32)
            if self._act_last is None or not np.array_equal(m, self._act_last):
                self.act_d.copy_(torch.from_numpy(m))
                self._act_last = m.copy()
            b.active = self.act_d.data_ptr()

    def B(self, mask=None):
        self._fill_batch(mask)
        return ctypes.byref(self._batch)

    def sc(self, slot, v):
        """Device copy of K per-problem scalars (uploaded only when its bits
        differ from the slot's last upload)."""
        t = self._sc[slot]
        h = np.array(np.broadcast_to(np.asarray(v, np.float64), (self.K,)))
        last = self._sc_last[slot]
        if last is None or not np.array_equal(h.view(np.int64), last.view(np.int64)):
            t.copy_(torch.from_numpy(h))
            self._sc_last[slot] = h
        return t

    def _o(self):
        return self.out_h.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def _rows(self, nv):
        return self.out_h[: nv * self.K].reshape(self.K, nv).copy()

    # ---- vectors -------------------------------------------------------------
    def create_vec(self, x=None):
        return BVec(self.n, self.np_, self.device, x)

    def copy_vec(self, dst, src, mask=None, what=3):
        if mask is None and what == 3:
            for a, b in ((dst.x, src.x), (dst.zl, src.zl), (dst.zu, src.zu), (dst.sl, src.sl), (dst.su, src.su)):
                a.copy_(b)
            return
        _lib.check(_lib.lib().amigo_ipm_batch_copy_state(self.B(mask), dst.state(), src.state(), what, _s()), "ipm")

    def copy_x(self, dst, src, mask=None):
        """dst[x segment] = src for the masked problems (x-length tensors)."""
        if mask is None:
            dst.copy_(src)
            return
        d = _lib.IpmState(dst.data_ptr(), 0, 0, 0, 0)
        s = _lib.IpmState(src.data_ptr(), 0, 0, 0, 0)
        _lib.check(_lib.lib().amigo_ipm_batch_copy_state(self.B(mask), ctypes.byref(d), ctypes.byref(s), 1, _s()),
                   "ipm")

    # ---- kernels ----------------------------------------------------------
    def residual(self, mu, v, grad, res, mask=None):
        _lib.check(_lib.lib().amigo_ipm_batch_residual(self.B(mask), self.info, v.state(), _p(self.sc(0, mu)),
                                                       _p(grad), _p(res), _p(self.ws), _p(self.out_d), self._o(),
                                                       _s()), "ipm")
        with np.errstate(invalid="ignore"):
            return np.sqrt(self._rows(3)[:, 0])

    def diagonal(self, v, diag, mask=None):
        _lib.check(_lib.lib().amigo_ipm_batch_diagonal(self.B(mask), self.info, v.state(), _p(diag), _s()), "ipm")

    def update(self, mu, v, px, upd, mask=None):
        self.copy_x(upd.x, px, mask)
        _lib.check(_lib.lib().amigo_ipm_batch_bound_dual_step(self.B(mask), self.info, v.state(), _p(self.sc(0, mu)),
                                                              _p(px), _p(upd.zl), _p(upd.zu), _s()), "ipm")

    def max_step(self, tau, v, upd, mask=None):
        _lib.check(_lib.lib().amigo_ipm_batch_max_step(self.B(mask), self.info, v.state(), _p(self.sc(0, tau)),
                                                       _p(upd.x), _p(upd.zl), _p(upd.zu), _p(self.ws),
                                                       _p(self.out_d), self._o(), _s()), "ipm")
        o = self._rows(4)
        return o[:, 0], o[:, 1].astype(np.int64), o[:, 2], o[:, 3].astype(np.int64)

    def apply_step(self, ax, az, v, upd, tmp, mask=None):
        _lib.check(_lib.lib().amigo_ipm_batch_apply_step(self.B(mask), self.info, v.state(), _p(self.sc(0, ax)),
                                                         _p(self.sc(1, az)), _p(upd.x), _p(upd.zl), _p(upd.zu),
                                                         tmp.state(), _s()), "ipm")

    def reduce(self, kind, v, mu=0.0, a=None, b=None, c=None, mask=None):
        _lib.check(_lib.lib().amigo_ipm_batch_reduce(RED[kind], self.B(mask), self.info, v.state(),
                                                     _p(self.sc(0, mu)), _p(a), _p(b), _p(c), _p(self.ws),
                                                     _p(self.out_d), self._o(), _s()), "ipm")
        return self._rows(4)

    def complementarity(self, v, mask=None):
        """(avg, xi) per problem (optimizer.h:287-300)."""
        r = self.reduce("compl", v, mask=mask)
        s, cnt, mn = r[:, 0], r[:, 1], r[:, 2]
        avg = np.where(cnt > 0, s / np.where(cnt > 0, cnt, 1.0), 0.0)
        mn = np.where(cnt == 0, np.finfo(np.float64).max, mn)
        with np.errstate(divide="ignore", invalid="ignore"):
            xi = np.where(avg > 0, np.maximum(0.0, np.minimum(1.0, mn / np.where(avg > 0, avg, 1.0))), 1.0)
        return avg, xi

    def kkt_error_mu(self, mu, v, grad, mask=None):
        r = self.reduce("kkt_err", v, mu, grad, mask=mask)
        return r[:, 0], r[:, 1], r[:, 2]

    def theta(self, v, grad, mask=None):
        return self.reduce("theta", v, 0.0, grad, mask=mask)[:, 0]

    def logsum(self, mu, v, mask=None):
        return self.reduce("logsum", v, mu, mask=mask)[:, 0]

    def dphi(self, mu, v, res, px, diag, mask=None):
        return self.reduce("dphi_kkt", v, mu, res, px, diag, mask=mask)[:, 0]

    def abs_sums(self, v, mask=None):
        r = self.reduce("abs_sums", v, mask=mask)
        return r[:, 0], r[:, 1], r[:, 2], r[:, 3]

    def reset_bound_multipliers(self, mu, kappa, v, mask=None):
        _lib.check(_lib.lib().amigo_ipm_batch_reset_bound_multipliers(self.B(mask), self.info, v.state(),
                                                                      _p(self.sc(0, mu)), float(kappa), _s()), "ipm")

    def dual_residual_vector(self, v, grad, out, mask=None):
        _lib.check(_lib.lib().amigo_ipm_batch_dual_residual_vector(self.B(mask), self.info, v.state(), _p(grad),
                                                                   _p(out), _s()), "ipm")

    def relax_bounds(self, factor, cvt, v):
        if factor <= 0:
            return
        _lib.check(_lib.lib().amigo_ipm_batch_initialize(self.B(None), _p(self.lbx0), _p(self.ubx0), _p(self.pidx),
                                                         float(factor), float(cvt), 1e-2, 1e-2, _p(self.lbx),
                                                         _p(self.ubx), v.state(), 1, 0, _s()), "ipm")

    def initialize_multipliers_and_slacks(self, v):
        _lib.check(_lib.lib().amigo_ipm_batch_initialize(self.B(None), _p(self.lbx0), _p(self.ubx0), _p(self.pidx),
                                                         0.0, 0.0, 1e-2, 1e-2, _p(self.lbx), _p(self.ubx),
                                                         v.state(), 0, 1, _s()), "ipm")

    def floor_slacks(self, v, mu, s_min, slack_move, mask=None):
        _lib.check(_lib.lib().amigo_ipm_batch_floor_slacks(self.B(mask), self.info, v.state(), _p(self.lbx0),
                                                           _p(self.ubx0), _p(self.sc(0, mu)), _p(self.sc(1, s_min)),
                                                           float(slack_move), _p(self.ws), _p(self.out_d),
                                                           self._o(), _s()), "ipm")
        return self._rows(1)[:, 0]

    def segmax(self, v, seg_d):
        """max |v| over each problem's contiguous segment of v (host copy)."""
        h = self._segmax_h
        _lib.check(_lib.lib().amigo_segment_absmax(
            _p(v), _p(seg_d), self.K, _p(self._segmax_d),
            h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _s()), "ipm")
        return h.copy()

    # ---- NLP scaling (optimizer.h:502-598), per problem ----------------------
    def compute_nlp_scaling(self, x, grad, solver, max_gradient=100.0, min_value=1e-8):
        obj_max = self.segmax(grad[self.pidx_g].contiguous(), self.pseg_d)
        has_p = self.n_primal_k > 0
        with np.errstate(divide="ignore"):
            sc = np.where(has_p & (obj_max > max_gradient), max_gradient / np.where(obj_max > 0, obj_max, 1.0), 1.0)
        sc = np.where(sc < min_value, min_value, sc)
        self.obj_scale = sc
        hess = self._scal_hess
        self.bp.hessian(1.0, x, hess)
        rowp_d, _ = hess.pattern_device()
        dc = torch.ones(self.nc_, dtype=torch.float64, device=self.device)
        _lib.check(_lib.lib().amigo_ipm_constraint_scaling(self.nc_, _p(self.cidx_g.to(torch.int32)), _p(rowp_d),
                                                           _p(hess.data), float(max_gradient), float(min_value),
                                                           _p(dc), _s()), "ipm")
        self.constr_scale = dc
        self.lbh.mul_(dc)
        sv = torch.ones(self.n, dtype=torch.float64, device=self.device)
        sv[self.cidx_g] = dc
        self.scale_vec = sv
        self.scaling_active = True
        self._set_info()

    def apply_gradient_scaling(self, grad):
        if not self.scaling_active:
            return
        _lib.check(_lib.lib().amigo_ipm_batch_scale_rows(self.B(None), _p(self.cidx), _p(self.constr_scale),
                                                         _p(grad), 0, _s()), "ipm")

    def apply_hessian_scaling(self, hess, mask=None):
        if not self.scaling_active:
            return
        rowp_d, cols_d = hess.pattern_device()
        act = self.bp.active_dev(mask)
        if act is None:
            rc = _lib.lib().amigo_csr_scale_symmetric(self.n, _p(rowp_d), _p(cols_d), _p(self.scale_vec),
                                                      _p(hess.data), _s())
        else:
            if getattr(self, "_xprob32", None) is None:
                self._xprob32 = self.xprob.to(torch.int32)
            rc = _lib.lib().amigo_csr_scale_symmetric_masked(self.n, _p(rowp_d), _p(cols_d),
                                                             _p(self.scale_vec), _p(hess.data),
                                                             _p(self._xprob32), _p(act), _s())
        _lib.check(rc, "ipm")

    def unscale_multipliers(self, x):
        if self.scaling_active:
            _lib.check(_lib.lib().amigo_ipm_batch_scale_rows(self.B(None), _p(self.cidx), _p(self.constr_scale),
                                                             _p(x), 1, _s()), "ipm")

    # ---- iterative refinement (linear_solver.py:109-264) ---------------------
    def ir_init(self, mu, v, grad, rhs, px, dzl, dzu, mask):
        _lib.check(_lib.lib().amigo_ipm_batch_ir_init(self.B(mask), self.info, v.state(), _p(self.sc(0, mu)),
                                                      _p(grad), _p(rhs), _p(px), _p(dzl), _p(dzu), _p(self.ws),
                                                      _p(self.out_d), self._o(), _s()), "ipm")
        return self._rows(1)[:, 0]

    def ir_residual(self, mu, v, grad, rhs, Kpx, px, dzl, dzu, econd, ezl, ezu, mask):
        _lib.check(_lib.lib().amigo_ipm_batch_ir_residual(self.B(mask), self.info, v.state(), _p(self.sc(0, mu)),
                                                          _p(grad), _p(rhs), _p(Kpx), _p(px), _p(dzl), _p(dzu),
                                                          _p(econd), _p(ezl), _p(ezu), _p(self.ws), _p(self.out_d),
                                                          self._o(), _s()), "ipm")
        o = self._rows(2)
        return o[:, 0], o[:, 1]

    def ir_update(self, v, corr, ezl, ezu, px, dzl, dzu, mask):
        _lib.check(_lib.lib().amigo_ipm_batch_ir_update(self.B(mask), self.info, v.state(), _p(corr), _p(ezl),
                                                        _p(ezu), _p(px), _p(dzl), _p(dzu), _s()), "ipm")


# ---------------------------------------------------------------------------
# Inertia correction, per problem (inertia_correction.py:13-402)
# ---------------------------------------------------------------------------
class BatchIC:
    """Algorithm IC of Wachter & Biegler with one state per problem; every
    method acts on the problems of an index array and follows
    optimizer.InertiaCorrector step for step."""

    NOT_YET, NOT_DEGEN, DEGENERATE = 0, 1, 2
    NO_TEST, T00, T10, T01, T11 = 0, 1, 2, 3, 4

    def __init__(self, K, mu):
        z = lambda: np.zeros(K)
        self.numerical_eps = 1e-12
        self.dx_last, self.dc_last, self.dx_curr, self.dc_curr = z(), z(), z(), z()
        self.last_dw, self.last_dc = z(), z()
        self.dw_init, self.dw_min, self.dw_max = 1e-4, 1e-20, 1e20
        self.kw_inc, self.kw_first_inc, self.kw_dec = 8.0, 100.0, 1.0 / 3
        self.dc_val, self.dc_exp = 1e-8, 0.25
        self.hess_degen = np.zeros(K, np.int64)
        self.jac_degen = np.zeros(K, np.int64)
        self.degen_iters = np.zeros(K, np.int64)
        self.degen_max = 3
        self.test = np.zeros(K, np.int64)
        self.pivtol, self.pivtolmax = np.full(K, 1e-6), 0.1
        self.mu = np.array(mu, np.float64).copy()
        self.history = [[] for _ in range(K)]

    def delta_cd(self, I):
        return self.dc_val * _pow(self.mu[I], self.dc_exp)

    def deltas_wrong_inertia(self, I):
        """_get_deltas_for_wrong_inertia on I; returns success per entry."""
        prev = self.dx_curr[I].copy()
        cur = self.dx_curr[I].copy()
        last = self.dx_last[I]
        z = cur == 0.0
        zl = z & (last == 0.0)
        zn = z & (last != 0.0)
        nz = ~z
        first = nz & ((last == 0.0) | (1e5 * last < cur))
        other = nz & ~first
        cur = np.where(zl, self.dw_init, cur)
        cur = np.where(zn, np.maximum(self.dw_min, last * self.kw_dec), cur)
        cur = np.where(first, cur * self.kw_first_inc, cur)
        cur = np.where(other, cur * self.kw_inc, cur)
        over = cur > self.dw_max
        cur = np.where(over, prev, cur)
        self.dx_curr[I] = cur
        dl = self.dx_last[I]
        self.dx_last[I] = np.where(over, 0.0, dl)
        return ~over

    def perturb_wrong_inertia(self, I):
        self.finalize_test(I)
        ok = self.deltas_wrong_inertia(I)
        res = ok.copy()
        J = I[~ok]
        if J.size:
            c0 = self.dc_curr[J] == 0.0
            J2 = J[c0]
            if J2.size:
                self.dc_curr[J2] = self.delta_cd(J2)
                self.dx_curr[J2] = 0.0
                hd = self.hess_degen[J2]
                self.hess_degen[J2] = np.where(hd == self.DEGENERATE, self.NOT_YET, hd)
                self.test[J2] = self.NO_TEST
                r2 = self.deltas_wrong_inertia(J2)
                pos = {j: t for t, j in enumerate(J2)}
                for t, i in enumerate(I):
                    if i in pos:
                        res[t] = r2[pos[i]]
        return res

    def perturb_singularity(self, I):
        """_perturb_for_singularity, one problem at a time (rare path)."""
        out = np.ones(len(I), bool)
        for t, i in enumerate(I):
            out[t] = self._sing_one(int(i))
        return out

    def _sing_one(self, i):
        one = np.array([i])
        if self.hess_degen[i] == self.NOT_YET or self.jac_degen[i] == self.NOT_YET:
            ts = self.test[i]
            if ts == self.T00:
                if self.jac_degen[i] == self.NOT_YET:
                    self.dc_curr[i] = self.delta_cd(one)[0]
                    self.test[i] = self.T10
                else:
                    if not self.deltas_wrong_inertia(one)[0]:
                        return False
                    self.test[i] = self.T01
            elif ts == self.T10:
                self.dc_curr[i] = 0.0
                if not self.deltas_wrong_inertia(one)[0]:
                    return False
                self.test[i] = self.T01
            elif ts == self.T01:
                self.dc_curr[i] = self.delta_cd(one)[0]
                if not self.deltas_wrong_inertia(one)[0]:
                    return False
                self.test[i] = self.T11
            elif ts == self.T11:
                if not self.deltas_wrong_inertia(one)[0]:
                    return False
        else:
            if self.dc_curr[i] > 0.0:
                if not self.deltas_wrong_inertia(one)[0]:
                    return False
            else:
                self.dc_curr[i] = self.delta_cd(one)[0]
        return True

    def finalize_test(self, I):
        ts = self.test[I]
        hd, jd, di = self.hess_degen[I].copy(), self.jac_degen[I].copy(), self.degen_iters[I].copy()
        a = ts == self.T00
        hd = np.where(a & (hd == self.NOT_YET), self.NOT_DEGEN, hd)
        jd = np.where(a & (jd == self.NOT_YET), self.NOT_DEGEN, jd)
        b = ts == self.T10
        hd = np.where(b & (hd == self.NOT_YET), self.NOT_DEGEN, hd)
        bj = b & (jd == self.NOT_YET)
        di = np.where(bj, di + 1, di)
        jd = np.where(bj & (di >= self.degen_max), self.DEGENERATE, jd)
        c = ts == self.T01
        jd = np.where(c & (jd == self.NOT_YET), self.NOT_DEGEN, jd)
        ch = c & (hd == self.NOT_YET)
        di = np.where(ch, di + 1, di)
        hd = np.where(ch & (di >= self.degen_max), self.DEGENERATE, hd)
        d = ts == self.T11
        di = np.where(d, di + 1, di)
        dd = d & (di >= self.degen_max)
        hd = np.where(dd, self.DEGENERATE, hd)
        jd = np.where(dd, self.DEGENERATE, jd)
        self.hess_degen[I], self.jac_degen[I], self.degen_iters[I] = hd, jd, di
        self.test[I] = self.NO_TEST

    def consider_new_system(self, I):
        self.finalize_test(I)
        dx, dc = self.dx_curr[I], self.dc_curr[I]
        self.dx_last[I] = np.where(dx > 0.0, dx, self.dx_last[I])
        self.dc_last[I] = np.where(dc > 0.0, dc, self.dc_last[I])
        undecided = (self.hess_degen[I] == self.NOT_YET) | (self.jac_degen[I] == self.NOT_YET)
        self.test[I] = np.where(undecided, self.T00, self.NO_TEST)
        self.dc_curr[I] = np.where(self.jac_degen[I] == self.DEGENERATE, self.delta_cd(I), 0.0)
        self.dx_curr[I] = 0.0
        H = I[self.hess_degen[I] == self.DEGENERATE]
        if H.size:
            self.deltas_wrong_inertia(H)

    def increase_quality(self, I):
        ok = self.pivtol[I] < self.pivtolmax
        J = I[ok]
        self.pivtol[J] = np.minimum(self.pivtolmax, _pow(self.pivtol[J], 0.5))
        return ok


class BatchFilter:
    """(phi, theta) filters, one per problem (filter_acceptance.py:13-51)."""

    def __init__(self, K, gamma_phi=1e-8, gamma_theta=1e-5, cap=64):
        self.phi = np.zeros((K, cap))
        self.theta = np.zeros((K, cap))
        self.valid = np.zeros((K, cap), bool)
        self.gamma_phi, self.gamma_theta = gamma_phi, gamma_theta

    def acceptable(self, I, phi, theta):
        bad = self.valid[I] & (phi[:, None] > self.phi[I]) & (theta[:, None] > self.theta[I])
        return ~bad.any(axis=1)

    def add(self, I, phi, theta):
        """Add (phi - gamma_phi theta, (1 - gamma_theta) theta) to the
        filters of the problems I, dropping the entries it dominates."""
        I = np.asarray(I)
        if I.size == 0:
            return
        pe = phi - self.gamma_phi * theta
        te = (1.0 - self.gamma_theta) * theta
        dom = self.valid[I] & (self.phi[I] >= pe[:, None]) & (self.theta[I] >= te[:, None])
        v = self.valid[I] & ~dom
        if (~v).sum(axis=1).min() == 0:
            self._grow()
            v = np.concatenate([v, np.zeros_like(v)], axis=1)
        slot = np.argmax(~v, axis=1)
        v[np.arange(len(I)), slot] = True
        self.valid[I] = v
        self.phi[I, slot] = pe
        self.theta[I, slot] = te

    def _grow(self):
        K, c = self.phi.shape
        for name in ("phi", "theta"):
            a = getattr(self, name)
            setattr(self, name, np.concatenate([a, np.zeros((K, c))], axis=1))
        self.valid = np.concatenate([self.valid, np.zeros((K, c), bool)], axis=1)

    def clear(self, I):
        self.valid[I] = False

    def size(self):
        return self.valid.sum(axis=1)


# ---------------------------------------------------------------------------
# Lockstep driver
# ---------------------------------------------------------------------------
EPS10 = 10.0 * np.finfo(float).eps


def _pow1(x, e):
    try:
        r = x ** e
    except OverflowError:
        return np.inf
    return r if isinstance(r, float) else np.nan


def _pow(a, e):
    """a ** e element by element with the scalar pow the single-problem
    driver uses (numpy's vector pow may differ in the last bit)."""
    return np.array([_pow1(float(x), e) for x in np.asarray(a, np.float64).ravel()]).reshape(np.shape(a))


def _pow_at(a, e, I):
    """_pow at the indices I only (zeros elsewhere): the scalar pow costs a
    Python call per element, so it is taken only where a decision reads it."""
    out = np.zeros(np.shape(a))
    for i in I:
        out[i] = _pow1(float(a[i]), e)
    return out


def _log10(a):
    with np.errstate(all="ignore"):
        return np.array([np.log10(float(x)) for x in np.asarray(a, np.float64).ravel()]).reshape(np.shape(a))


class BatchOptimizer:
    """Interior-point solves of K problems in lockstep (see the module doc).

    Every per-problem decision follows optimizer.Optimizer.optimize; each
    device call covers all problems still in the phase that issues it."""

    def __init__(self, problems, order="ND", scaling=True, verbose=False, precision="fp64"):
        self.problems = list(problems)
        self.order = order
        self.scaling = scaling
        self.precision = precision
        self.fp64_fallbacks = 0
        self.verbose = verbose
        t = time.perf_counter()
        bp = self.bp = BatchProblem(self.problems)
        self.K = bp.K
        self.solver = BatchSolver(bp, order=order, scaling=scaling, precision=precision)
        be = self.be = BatchIPM(bp, bp.lower, bp.upper)
        dev = bp.device
        self.x = torch.as_tensor(bp.x0.copy(), dtype=torch.float64, device=dev)
        self._x0 = self.x.clone()
        self.vars = be.create_vec(self.x)
        self.update = be.create_vec()
        self.temp = be.create_vec()
        mk = lambda: torch.zeros(bp.n, dtype=torch.float64, device=dev)
        self.grad, self.res, self.diag, self.px, self.ir_corr = mk(), mk(), mk(), mk(), mk()
        self._gs, self._sol = mk(), mk()
        self.mult_dev = be.is_mult
        self.primal_dev = ~be.is_mult
        self.n_primal = be.n_primal_k
        self.n_dual = be.n_constr_k
        self.setup_time = time.perf_counter() - t
        self.fallback = np.zeros(self.K, bool)
        self.results = None
        self.profile_phases = False
        self.phase_times = {}

    def _tick(self, name, t0):
        """Accumulate device-synchronized wall time of a phase (when
        profile_phases is set); returns the new start time."""
        if not self.profile_phases:
            return t0
        torch.cuda.synchronize()
        t = time.perf_counter()
        self.phase_times[name] = self.phase_times.get(name, 0.0) + t - t0
        return t

    # ---- helpers ----------------------------------------------------------
    @staticmethod
    def _mask(K, I):
        m = np.zeros(K, bool)
        m[I] = True
        return m

    def _reset(self):
        be = self.be
        self.x.copy_(self._x0)
        be.lbx.copy_(be.lbx0)
        be.ubx.copy_(be.ubx0)
        be.lbh.copy_(be.lbh0)
        be.scaling_active = False
        be.obj_scale = np.ones(self.K)
        be._set_info()
        self.vars.zl.zero_()
        self.vars.zu.zero_()
        self.vars.sl.fill_(1.0)
        self.vars.su.fill_(1.0)
        self.bp.set_obj_scales(np.ones(self.K))

    def _update_gradient(self, x, mask=None):
        """grad = scaled [alpha grad f + J^T lambda; c] at x for the masked
        problems (newton_direction.py:17-27)."""
        out = self.grad if mask is None else self._gs
        self.bp.gradient(True, x, out, mask)
        self.be.apply_gradient_scaling(out)
        if mask is not None:
            self.be.copy_x(self.grad, out, mask)

    def _barrier_objective(self, v, mask):
        """phi_mu = alpha f(x) - mu sum log(gap), per problem
        (filter_line_search.py:19-46)."""
        x = v.x
        ci = self.be.cidx_g
        lam = x[ci].clone()
        x[ci] = 0.0
        f = self.bp.lagrangian(True, x)
        x[ci] = lam
        return f + self.be.logsum(self.mu, v, mask)

    def _mx(self, I):
        return self._mask(self.K, I)

    # ---- initialization (iterate_initialization.py:16-69) -------------------
    def _initialize_iterate(self, options):
        be = self.be
        K = self.K
        be.relax_bounds(1e-8, options["constr_viol_tol"], self.vars)
        self.x[be.cidx_g] = 0.0
        be.initialize_multipliers_and_slacks(self.vars)
        self._update_gradient(self.x)
        nlp_max_grad = options["nlp_scaling_max_gradient"]
        if nlp_max_grad > 0:
            be.compute_nlp_scaling(self.x, self.grad, self.solver, max_gradient=nlp_max_grad)
            self.bp.set_obj_scales(be.obj_scale)
            self._update_gradient(self.x)
        if options["init_affine_step_multipliers"]:
            raise NotImplementedError("affine-step multiplier initialization is not batched")
        if options["init_least_squares_multipliers"]:
            self._least_squares_multipliers()
        self._update_gradient(self.x)

    def _least_squares_multipliers(self, lambda_max=1e3):
        """multiplier_initialization.py:19-58, per problem."""
        be = self.be
        be.dual_residual_vector(self.vars, self.grad, self.res)
        self.res.mul_(-1.0)
        self.diag.zero_()
        self.diag[be.pidx_g] = 1.0
        hs = be.apply_hessian_scaling if be.scaling_active else None
        self.solver.factor(0.0, self.x, self.diag, post_hessian=hs)
        p, q, _, nan = self.solver.inertia()
        # The solver-level inertia record (AmigoSolver.inertia_history)
        # starts with this factorization.
        self._ls_inertia = [(int(p[k]), int(q[k])) if nan[k] == 0 else None for k in range(self.K)]
        self.solver.solve(self.res, self.px)
        lmax = be.segmax(self.px[be.cidx_g].contiguous(), be.cseg_d)
        reject = (self.n_dual > 0) & (lmax > lambda_max)
        keep = torch.as_tensor(~reject, device=self.x.device)[be.cprob]
        ci = be.cidx_g
        self.x[ci] = torch.where(keep, self.px[ci], torch.zeros_like(self.px[ci]))

    # ---- scaling / convergence (optimality_scaling.py, convergence_check.py)
    def _optimality_scaling(self, mask):
        s_max = 100.0
        z_asum, y_asum, n_bounds, n_constraints = self.be.abs_sums(self.vars, mask)
        s_c = np.where(n_bounds == 0, 1.0, np.maximum(s_max, z_asum / np.where(n_bounds == 0, 1.0, n_bounds)) / s_max)
        n_all = n_constraints + n_bounds
        s_d = np.where(n_all == 0, 1.0,
                       np.maximum(s_max, (y_asum + z_asum) / np.where(n_all == 0, 1.0, n_all)) / s_max)
        return s_d, s_c

    def _x_max(self):
        """max |x| of every problem (the divergence test), one native
        segmented reduction over the contiguous per-problem segments."""
        if getattr(self, "_xmax_d", None) is None:
            self._xmax_d = torch.zeros(self.K, dtype=torch.float64, device=self.x.device)
            self._xmax_h = np.zeros(self.K)
        h = self._xmax_h
        _lib.check(_lib.lib().amigo_segment_absmax(
            _p(self.x), _p(self.be.xseg_d), self.K, _p(self._xmax_d),
            h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _s()))
        return h.copy()

    def _check_convergence(self, i, options, A, res_norm, kkt):
        tol = options["convergence_tolerance"]
        d_inf, p_inf, c_inf = kkt
        s_d, s_c = self._optimality_scaling(A)
        err = np.maximum(np.maximum(d_inf / s_d, p_inf), c_inf / s_c)
        status = np.full(self.K, CONTINUE)
        conv = ((err <= tol) & (d_inf <= options["dual_inf_tol"]) & (p_inf <= options["constr_viol_tol"])
                & (c_inf <= options["compl_inf_tol"]))
        status[A & conv] = CONVERGED
        left = A & ~conv
        x_max = self._x_max()
        div = left & (x_max > options["diverging_iterates_tol"])
        status[div] = DIVERGED
        left &= ~div
        is_acc = ((err <= options["acceptable_tol"]) & (d_inf <= options["acceptable_dual_inf_tol"])
                  & (p_inf <= options["acceptable_constr_viol_tol"])
                  & (c_inf <= options["acceptable_compl_inf_tol"]))
        if options["acceptable_iter"] > 0:
            inc = left & is_acc
            self.acc[inc] += 1
            self.acc[left & ~is_acc] = 0
            acc_done = inc & (self.acc >= options["acceptable_iter"])
        else:
            self.acc[left] = 0
            acc_done = np.zeros(self.K, bool)
        status[acc_done] = CONVERGED_ACCEPTABLE
        left &= ~acc_done
        rel = np.abs(res_norm - self.prev_res) / np.maximum(res_norm, 1e-30)
        fl = (rel < 1e-14) & (i > 0)
        self.floor[left & fl] += 1
        self.floor[left & ~fl] = 0
        pf = left & (self.floor >= 3) & is_acc
        status[pf] = PRECISION_FLOOR
        return status, err

    # ---- barrier (barrier_strategy/heuristic.py) ------------------------------
    @staticmethod
    def _loqo(xi, comp, gamma, r, tol, M):
        """The LOQO barrier rule for the problems of the mask M (the other
        entries are not used)."""
        with np.errstate(divide="ignore", invalid="ignore"):
            term = (1 - r) * (1 - xi) / np.where(xi > 1e-10, xi, 1.0)
        p = _pow_at(np.minimum(term, 2.0), 3, np.nonzero(M & (xi > 1e-10))[0])
        factor = np.where(xi > 1e-10, p, 2.0**3)
        return np.maximum(gamma * factor * comp, tol)

    def _barrier_update(self, i, res_norm, options, A):
        tol = options["convergence_tolerance"]
        cit = options["compl_inf_tol"]
        if i == 0:
            return
        M = A & (self.mu > min(tol, cit) / (options["barrier_tol_factor"] + 1.0))
        if not M.any():
            return
        if options["barrier_strategy"] != "heuristic":
            raise NotImplementedError("the batched driver runs the heuristic barrier strategy")
        comp_h, xi_h = self.be.complementarity(self.vars, M)
        mu_floor = min(tol, cit) / (options["barrier_tol_factor"] + 1.0)
        new = self._loqo(xi_h, comp_h, options["heuristic_barrier_gamma"], options["heuristic_barrier_r"],
                         mu_floor, M)
        self.mu[M] = new[M]

    def _increase_barrier_on_rejections(self, I, max_rej, inc, initial):
        for i in I:
            if self.rej[i] >= max_rej:
                new = min(self.mu[i] * inc, initial)
                if new > self.mu[i]:
                    self.mu[i] = new
                self.rej[i] = 0

    # ---- inertia-corrected factorization (inertia_correction.py:261-393) ----
    def _factorize(self, A, options):
        """Returns factorize_ok per problem (for the problems in A)."""
        be, ic, K = self.be, self.ic, self.K
        # Only the problems of A are refactored and read their matrices.
        self.solver.assemble_hessian(True, self.x, A)
        be.apply_hessian_scaling(self.solver.hess, A)
        itol = options.get("inertia_tolerance", 0)
        max_corr = options["max_inertia_corrections"]
        diag = self.diag
        diag[self.primal_dev] += ic.numerical_eps
        reg = diag.clone()
        I = np.nonzero(A)[0]
        ic.consider_new_system(I)
        improved = np.zeros(K, bool)
        ok_out = np.zeros(K, bool)
        pending = A.copy()
        xprob = be.xprob
        for attempt in range(max_corr + 1):
            dx = torch.as_tensor(ic.dx_curr, device=diag.device)[xprob]
            dc = torch.as_tensor(ic.dc_curr, device=diag.device)[xprob]
            diag.copy_(reg)
            diag.add_(torch.where(self.primal_dev & (dx > 0), dx, torch.zeros_like(dx)))
            diag.sub_(torch.where(self.mult_dev & (dc > 0), dc, torch.zeros_like(dc)))
            self.solver.add_diagonal_and_factor(diag, pending)
            p, q, _, nan = self.solver.inertia()
            P = np.nonzero(pending)[0]
            singular = nan[P] > 0
            for t, k in enumerate(P):
                if not singular[t]:
                    ic.history[k].append((int(p[k]), int(q[k])))
            npk, ndk = self.n_primal[P], self.n_dual[P]
            good = (~singular & (np.abs(p[P] - npk) <= itol) & (np.abs(q[P] - ndk) <= itol)
                    & (p[P] + q[P] >= npk + ndk - itol))
            G = P[good]
            ok_out[G] = True
            ic.last_dw[G] = ic.dx_curr[G]
            ic.last_dc[G] = ic.dc_curr[G]
            pending[G] = False
            R = P[~good]
            if R.size == 0:
                break
            sing = singular[~good]
            qr = q[R]
            fail = np.zeros(len(R), bool)
            # singular with multipliers -> perturb for singularity
            s1 = sing & (self.n_dual[R] > 0)
            if s1.any():
                fail[s1] = ~ic.perturb_singularity(R[s1])
            # too few negative pivots -> better pivoting first, then singular
            s2 = ~sing & (qr < self.n_dual[R])
            if s2.any():
                R2 = R[s2]
                assume = np.ones(len(R2), bool)
                ni = ~improved[R2]
                if ni.any():
                    got = ic.increase_quality(R2[ni])
                    improved[R2[ni]] = got
                    tmp = assume[ni]
                    tmp[got] = False
                    assume[ni] = tmp
                if assume.any():
                    r = ~ic.perturb_singularity(R2[assume])
                    f2 = np.zeros(len(R2), bool)
                    f2[assume] = r
                    fail[s2] = f2
            s3 = ~s1 & ~s2
            if s3.any():
                fail[s3] = ~ic.perturb_wrong_inertia(R[s3])
            F = R[fail]
            pending[F] = False
            ic.last_dw[F] = ic.dx_curr[F]
            ic.last_dc[F] = ic.dc_curr[F]
            if not pending.any():
                break
        rest = np.nonzero(pending)[0]
        ic.last_dw[rest] = ic.dx_curr[rest]
        ic.last_dc[rest] = ic.dc_curr[rest]
        return ok_out

    # ---- direction (newton_direction.py:29-195) ------------------------------
    def _solve_with_mu(self, B):
        be = self.be
        be.residual(self.mu, self.vars, self.grad, self.res, B)
        rhs = self.res.clone()
        self.solver.solve(self.res, self._sol, B)
        be.copy_x(self.px, self._sol, B)
        ratio = self._refine(B, rhs)
        if self.precision == "mixed":
            # Refinement that ends above the 1e-10 ratio: FP64 factor of
            # those problems' matrices, a fresh solve and refinement again
            # (solver.AmigoSolver.iterative_refinement, per problem).
            F = B & ~(ratio < 1e-10)
            if F.any():
                self.solver.refactor_fp64(F)
                self.fp64_fallbacks += int(F.sum())
                self.solver.solve(rhs, self._sol, F)
                be.copy_x(self.px, self._sol, F)
                self._refine(F, rhs)
        be.update(self.mu, self.vars, self.px, self.update, B)

    def _refine(self, B, rhs, max_steps=10):
        """FP64 refinement on the 8-block system (linear_solver.py:109-264)."""
        be = self.be
        dev = self.x.device
        dzl = torch.empty(be.np_, dtype=torch.float64, device=dev)
        dzu = torch.empty_like(dzl)
        ezl = torch.empty_like(dzl)
        ezu = torch.empty_like(dzl)
        Kpx = self.solver._tmp
        nrm_rhs = be.ir_init(self.mu, self.vars, self.grad, rhs, self.px, dzl, dzu, B)
        ratio_max, improve, min_steps, max_cond = 1e-10, 0.999999999, 1, 1e6
        ratio_old = np.full(self.K, 1e30)
        final = np.zeros(self.K)
        R = B.copy()
        for step in range(max_steps):
            if not R.any():
                break
            self.solver.kkt_matvec(self.px, Kpx, R)
            nrm_resid, nrm_res = be.ir_residual(self.mu, self.vars, self.grad, rhs, Kpx, self.px, dzl, dzu,
                                                self.res, ezl, ezu, R)
            denom = np.minimum(nrm_res, max_cond * nrm_rhs) + nrm_rhs
            with np.errstate(divide="ignore", invalid="ignore"):
                ratio = np.where(denom > 0, nrm_resid / np.where(denom > 0, denom, 1.0), 0.0)
            final[R] = ratio[R]
            stop = R & (ratio < ratio_max) & (step >= min_steps)
            stop |= R & ~stop & (step >= min_steps) & (ratio > improve * ratio_old)
            R &= ~stop
            ratio_old = np.where(R, ratio, ratio_old)
            if not R.any():
                break
            self.solver.solve(self.res, self.ir_corr, R)
            be.ir_update(self.vars, self.ir_corr, ezl, ezu, self.px, dzl, dzu, R)
        return final

    # ---- filter line search (filter_line_search.py:48-280) -------------------
    def _line_search(self, S, alpha_x, alpha_z, tau, options, use_soc, phi_cur, wd):
        """S: problems searching; wd: those using the watchdog reference.
        Returns (alpha ratio, steps, accepted, last filter rejection)."""
        be, K = self.be, self.K
        max_iters = options["max_line_search_iterations"]
        alpha_red = options["backtracking_factor"]
        gamma_theta = options["filter_gamma_theta"]
        gamma_phi = options["filter_gamma_phi"]
        delta = options["filter_delta"]
        s_theta = options["filter_s_theta"]
        s_phi = options["filter_s_phi"]
        eta_phi = options["filter_eta_phi"]
        max_soc = options["filter_max_soc"]
        kappa_soc = options["filter_kappa_soc"]
        obj_max_inc = 5.0
        Sm = self._mx(S)
        nwd = S[~wd[S]]
        ref_theta = np.zeros(K)
        ref_barr = np.zeros(K)
        ref_dphi = np.zeros(K)
        if nwd.size:
            m = self._mx(nwd)
            th = be.theta(self.vars, self.grad, m)
            dp = be.dphi(self.mu, self.vars, self.res, self.px, self.diag, m)
            ref_theta[nwd], ref_dphi[nwd] = th[nwd], dp[nwd]
            ref_barr[nwd] = phi_cur[nwd]
        W = S[wd[S]]
        ref_theta[W], ref_barr[W], ref_dphi[W] = self.wd_theta[W], self.wd_barr[W], self.wd_dphi[W]
        theta_0 = np.where(np.isnan(self.theta0), ref_theta, self.theta0)
        theta_min = 1e-4 * np.maximum(1.0, theta_0)
        theta_max = 1e4 * np.maximum(1.0, theta_0)
        neg = ref_dphi < 0.0
        nd = np.where(neg, -ref_dphi, 1.0)
        # The powers are read only for the searching problems with
        # ref_dphi < 0 (the single-problem tests short-circuit otherwise).
        SN = S[neg[S]]
        nd_sphi = _pow_at(nd, s_phi, SN)
        rt_stheta = _pow_at(ref_theta, s_theta, SN)
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            am = np.where(neg, np.minimum(gamma_theta, gamma_phi * ref_theta / nd), gamma_theta)
            am = np.where(neg & (ref_theta <= theta_min),
                          np.minimum(am, delta * rt_stheta / nd_sphi), am)
        alpha_min = np.where(wd, alpha_x, am * 0.05)

        def is_ftype(a):
            with np.errstate(invalid="ignore", over="ignore"):
                return (ref_dphi < 0.0) & (a * nd_sphi > delta * rt_stheta)

        def armijo(tb, a):
            return (tb - ref_barr) - eta_phi * a * ref_dphi <= EPS10 * np.abs(ref_barr)

        def acceptable_to_iterate(tb, tt, m):
            """For the problems of mask m (other entries are not used)."""
            ok = np.ones(K, bool)
            # The logarithms only where the objective went up, as the
            # single-problem test takes them.
            with np.errstate(divide="ignore", invalid="ignore"):
                for i in np.nonzero(m & (tb > ref_barr))[0]:
                    rb = abs(float(ref_barr[i]))
                    basval = np.log10(rb) if rb > 10.0 else 1.0
                    if np.log10(max(float(tb[i]) - float(ref_barr[i]), 1e-300)) > obj_max_inc + basval:
                        ok[i] = False
            dec = ((tt - (1.0 - gamma_theta) * ref_theta <= EPS10 * np.abs(ref_theta))
                   | ((tb - ref_barr) - (-gamma_phi * ref_theta) <= EPS10 * np.abs(ref_barr)))
            return ok & dec

        def check(I, tb, tt, a):
            """(accepted, filter-rejected) for the problems I."""
            with np.errstate(all="ignore"):
                return _check(I, tb, tt, a)

        def _check(I, tb, tt, a):
            acc = np.zeros(K, bool)
            frej = np.zeros(K, bool)
            m = self._mx(I)
            tm = m & ~(tt > theta_max)
            fa = tm & (a > 0.0) & is_ftype(a) & (ref_theta <= theta_min)
            pass1 = np.where(fa, armijo(tb, a), acceptable_to_iterate(tb, tt, tm & ~fa)) & tm
            J = np.nonzero(pass1)[0]
            if J.size:
                ok = self.filter.acceptable(J, tb[J], tt[J])
                acc[J[ok]] = True
                frej[J[~ok]] = True
            return acc, frej

        def update_filter(I, tb, a):
            with np.errstate(all="ignore"):
                keep = is_ftype(a) & armijo(tb, a)
            J = I[~keep[I]]
            if J.size:
                self.filter.add(J, ref_barr[J], ref_theta[J])

        if use_soc:
            be.residual(self.mu, self.vars, self.grad, self.res, Sm)
            res_orig = self.res.clone()
            upd_bk = be.create_vec()
            be.copy_vec(upd_bk, self.update)
            px_orig = self.px.clone()
        alpha_p = alpha_x.copy()
        n_steps = np.zeros(K, np.int64)
        last_filt = np.zeros(K, bool)
        done = ~Sm
        out_alpha = np.zeros(K)
        out_steps = np.zeros(K, np.int64)
        out_acc = np.zeros(K, bool)
        out_lf = np.zeros(K, bool)
        finished = np.zeros(K, bool)  # left the loop without acceptance
        trial_theta = np.zeros(K)
        trial_barr = np.zeros(K)
        while True:
            Lm = ~done & ((alpha_p > alpha_min) | (n_steps == 0)) & (n_steps < max_iters)
            ex = ~done & ~Lm
            finished |= ex
            done |= ex
            if not Lm.any():
                break
            be.apply_step(alpha_p, alpha_z, self.vars, self.update, self.temp, Lm)
            self._update_gradient(self.temp.x, Lm)
            th = be.theta(self.temp, self.grad, Lm)
            tb = self._barrier_objective(self.temp, Lm)
            trial_theta[Lm], trial_barr[Lm] = th[Lm], tb[Lm]
            L = np.nonzero(Lm)[0]
            acc, frej = check(L, trial_barr, trial_theta, alpha_p)
            Aok = np.nonzero(acc)[0]
            if Aok.size:
                update_filter(Aok, trial_barr, alpha_p)
                be.copy_vec(self.vars, self.temp, self._mx(Aok))
                out_alpha[Aok] = alpha_p[Aok] / alpha_x[Aok]
                out_steps[Aok] = n_steps[Aok] + 1
                out_acc[Aok] = True
                out_lf[Aok] = last_filt[Aok]
                done[Aok] = True
            rej = Lm & ~acc
            last_filt[rej] = frej[rej]
            soc_ok = np.zeros(K, bool)
            Q = rej & (n_steps == 0) & (trial_theta >= ref_theta) if use_soc else np.zeros(K, bool)
            if Q.any():
                c_soc = res_orig.clone()
                alpha_soc = alpha_p.copy()
                theta_old = np.zeros(K)
                act = Q.copy()
                mult = self.mult_dev
                for soc_count in range(max_soc):
                    if soc_count > 0:
                        act &= ~(trial_theta > kappa_soc * theta_old)
                    if not act.any():
                        break
                    theta_old[act] = trial_theta[act]
                    be.residual(self.mu, self.temp, self.grad, self.res, act)
                    asoc = torch.as_tensor(alpha_soc, device=self.x.device)[be.xprob]
                    sel = mult & torch.as_tensor(act, device=self.x.device)[be.xprob]
                    c_soc = torch.where(sel, self.res + asoc * c_soc, c_soc)
                    be.copy_x(self.res, c_soc, act)
                    self._update_gradient(self.vars.x, act)
                    self.solver.solve(self.res, self._sol, act)
                    be.copy_x(self.px, self._sol, act)
                    be.update(self.mu, self.vars, self.px, self.update, act)
                    sax, _, saz, _ = be.max_step(self.tau, self.vars, self.update, act)
                    be.apply_step(sax, saz, self.vars, self.update, self.temp, act)
                    self._update_gradient(self.temp.x, act)
                    th = be.theta(self.temp, self.grad, act)
                    tb = self._barrier_objective(self.temp, act)
                    trial_theta[act], trial_barr[act] = th[act], tb[act]
                    sacc, _ = check(np.nonzero(act)[0], trial_barr, trial_theta, alpha_p)
                    Ok = np.nonzero(sacc)[0]
                    if Ok.size:
                        update_filter(Ok, trial_barr, alpha_p)
                        be.copy_vec(self.vars, self.temp, self._mx(Ok))
                        soc_ok[Ok] = True
                        act &= ~sacc
                    alpha_soc = np.where(act, sax, alpha_soc)
                Sok = np.nonzero(soc_ok)[0]
                out_alpha[Sok] = 1.0
                out_steps[Sok] = n_steps[Sok] + 1
                out_acc[Sok] = True
                out_lf[Sok] = False
                done[Sok] = True
                back = Q & ~soc_ok
                if back.any():
                    be.copy_vec(self.update, upd_bk, back)
                    be.copy_x(self.px, px_orig, back)
            cont = rej & ~soc_ok
            alpha_p = np.where(cont, alpha_p * alpha_red, alpha_p)
            n_steps = np.where(cont, n_steps + 1, n_steps)
        F = np.nonzero(finished & Sm)[0]
        if F.size:
            self._update_gradient(self.vars.x, self._mx(F))
            out_alpha[F] = alpha_p[F] / alpha_x[F]
            out_steps[F] = n_steps[F]
            out_acc[F] = False
            out_lf[F] = last_filt[F]
        return out_alpha, out_steps, out_acc, out_lf

    def _line_search_with_watchdog(self, S, alpha_x, alpha_z, tau, options, use_soc):
        """filter_line_search.py:395-438 (watchdog), per problem."""
        be, K = self.be, self.K
        Sm = self._mx(S)
        R1 = Sm & self.in_wd & (alpha_x < 1e-16)
        if R1.any():
            self._wd_restore(R1)
            self._update_gradient(self.vars.x, R1)
            self.in_wd[R1] = False
            self.wd_short[R1] = 0
        Wm = Sm & ~self.in_wd & (self.wd_trigger > 0) & (self.wd_short >= self.wd_trigger)
        if Wm.any():
            self._wd_save(Wm, alpha_x)
            th = be.theta(self.vars, self.grad, Wm)
            tb = self._barrier_objective(self.vars, Wm)
            dp = be.dphi(self.mu, self.vars, self.res, self.px, self.diag, Wm)
            self.wd_theta[Wm], self.wd_barr[Wm], self.wd_dphi[Wm] = th[Wm], tb[Wm], dp[Wm]
            self.in_wd[Wm] = True
            self.wd_trial[Wm] = 0
        phi = self._barrier_objective(self.vars, Sm)
        wd = self.in_wd.copy()
        alpha, steps, acc, frej = self._line_search(S, alpha_x, alpha_z, tau, options, use_soc, phi, wd)
        W = Sm & wd
        done_ok = W & acc
        self.in_wd[done_ok] = False
        self.wd_short[done_ok] = 0
        fail = W & ~acc
        self.wd_trial[fail] += 1
        over = fail & (self.wd_trial > self.wd_max_trials)
        take = fail & ~over
        acc = acc | take
        if over.any():
            self._wd_restore(over)
            self._update_gradient(self.vars.x, over)
            self.in_wd[over] = False
            self.wd_short[over] = 0
            ax2, _, az2, _ = be.max_step(tau, self.vars, self.update, over)
            phi2 = self._barrier_objective(self.vars, over)
            O = np.nonzero(over)[0]
            ax_n, az_n = alpha_x.copy(), alpha_z.copy()
            ax_n[O], az_n[O] = ax2[O], az2[O]
            a2, s2, c2, f2 = self._line_search(O, ax_n, az_n, tau, options, use_soc, phi2, np.zeros(K, bool))
            alpha[O], steps[O], acc[O], frej[O] = a2[O], s2[O], c2[O], f2[O]
        return alpha, steps, acc, frej

    def _wd_save(self, m, alpha_x):
        be = self.be
        be.copy_vec(self.wd_iter, self.vars, m, what=1)
        be.copy_vec(self.wd_upd, self.update, m)
        be.copy_x(self.wd_px, self.px, m)
        self.wd_alpha[m] = alpha_x[m]

    def _wd_restore(self, m):
        be = self.be
        be.copy_vec(self.vars, self.wd_iter, m, what=1)
        be.copy_vec(self.update, self.wd_upd, m)
        be.copy_x(self.px, self.wd_px, m)

    # ---- main loop (ipm_driver.py:106-483) ------------------------------------
    def optimize(self, options={}, on_done=None):
        """Solve every problem; returns one result dict per problem (the
        keys of Optimizer.optimize, plus "batched").  on_done(k, result, x)
        is called as soon as problem k has finished, with its unscaled
        primal-dual solution on the host."""
        start = time.perf_counter()
        options = get_default_options(options)
        if options["filter_monotone_mode"] if "filter_monotone_mode" in options else False:
            raise NotImplementedError
        K, be = self.K, self.be
        self._reset()
        data = [{"options": options, "converged": False, "iterations": []} for _ in range(K)]
        max_iters = options["max_iterations"]
        tau_min = options["tau_min"]
        max_rej = options["max_consecutive_rejections"]
        barrier_inc = options["barrier_increase_factor"]
        initial_barrier = options["initial_barrier_param"]
        self.mu = np.full(K, float(initial_barrier))
        self._initialize_iterate(options)
        self.tau = np.ones(K)
        z = lambda: np.zeros(K)
        zi = lambda: np.zeros(K, np.int64)
        line_iters, st_ax, st_az, st_xi, st_zi = zi(), z(), z(), np.full(K, -1), np.full(K, -1)
        self.prev_res = np.full(K, np.inf)
        self.acc, self.floor, self.rej = zi(), zi(), zi()
        filt_rej, filt_resets = zi(), zi()
        self.ic = ic = BatchIC(K, self.mu)
        use_soc = bool(options["second_order_correction"])
        filter_ls = options["filter_line_search"]
        if not filter_ls:
            raise NotImplementedError("the batched driver runs the filter line search")
        self.filter = BatchFilter(K, options["filter_gamma_phi"], options["filter_gamma_theta"])
        self.theta0 = np.full(K, np.nan)
        self.in_wd = np.zeros(K, bool)
        self.wd_short, self.wd_trial = zi(), zi()
        self.wd_trigger = np.full(K, options["watchdog_shortened_iter_trigger"])
        self.wd_max_trials = options["watchdog_trial_iter_max"]
        self.wd_theta, self.wd_barr, self.wd_dphi, self.wd_alpha = z(), z(), z(), z()
        self.wd_iter, self.wd_upd = be.create_vec(), be.create_vec()
        self.wd_px = torch.zeros_like(self.px)
        active = np.ones(K, bool)
        self.fallback[:] = False
        self._alone = {}
        self._alone_futs = {}
        self._alone_pool = None
        record_obj = options.get("record_objective", True)
        self.iterations_run = 0
        self.active_history = []
        for i in range(max_iters):
            tp = time.perf_counter()
            if self.profile_phases:
                torch.cuda.synchronize()
            A = active.copy()
            self.active_history.append(int(A.sum()))
            res_norm = be.residual(self.mu, self.vars, self.grad, self.res, A)
            ic.mu = self.mu.copy()
            need = A & np.isnan(self.theta0)
            if need.any():
                th = be.theta(self.vars, self.grad, need)
                self.theta0[need] = th[need]
            kkt = be.kkt_error_mu(0.0, self.vars, self.grad, A)
            status, err = self._check_convergence(i, options, A, res_norm, kkt)
            obj = self._barrier_objective(self.vars, A) if record_obj else None
            fsz = self.filter.size()
            now = time.perf_counter() - start
            for k in np.nonzero(A)[0]:
                it = {"iteration": i, "time": now, "residual": float(res_norm[k]), "barrier_param": float(self.mu[k]),
                      "line_iters": int(line_iters[k]), "alpha_x": float(st_ax[k]), "x_index": int(st_xi[k]),
                      "alpha_z": float(st_az[k]), "z_index": int(st_zi[k]), "inertia_delta": float(ic.last_dw[k]),
                      "inf_du": float(kkt[0][k]), "inf_pr": float(kkt[1][k]), "compl": float(kkt[2][k]),
                      "nlp_error": float(err[k]), "filter_size": int(fsz[k])}
                if record_obj:
                    it["objective"] = float(obj[k])
                data[k]["iterations"].append(it)
                s = status[k]
                if s == CONVERGED:
                    data[k]["converged"] = True
                elif s in (CONVERGED_ACCEPTABLE, PRECISION_FLOOR):
                    data[k]["converged"] = True
                    data[k]["acceptable"] = True
                    if s == PRECISION_FLOOR:
                        data[k]["precision_floor"] = True
            if on_done is not None:
                for k in np.nonzero(A & (status != CONTINUE))[0]:
                    data[k]["inertia_history"] = list(ic.history[k])
                    data[k]["solver_inertia_history"] = self._solver_history(k)
                    data[k]["num_iterations"] = len(data[k]["iterations"]) - 1
                    data[k]["time"] = time.perf_counter() - start
                    data[k]["batched"] = True
                    on_done(int(k), data[k], self._solution_of(int(k)))
            active &= status == CONTINUE
            A = active.copy()
            if not A.any():
                break
            self.iterations_run = i + 1
            self.prev_res[A] = res_norm[A]
            be.diagonal(self.vars, self.diag, A)
            before = self.mu.copy()
            # _barrier_step: barrier update, then the Newton direction.
            self._barrier_update(i, res_norm, options, A)
            ic.mu = self.mu.copy()
            tp = self._tick("convergence+barrier", tp)
            fok = self._factorize(A, options)
            tp = self._tick("factorize+IC", tp)
            B = A & fok
            if B.any():
                self._solve_with_mu(B)
            tp = self._tick("solve+refinement", tp)
            changed = A & (self.mu != before)
            if changed.any():
                C = np.nonzero(changed)[0]
                self.filter.clear(C)
                th = be.theta(self.vars, self.grad, changed)
                self.theta0[changed] = th[changed]
                self.in_wd[changed] = False
                self.wd_short[changed] = 0
                self.wd_trial[changed] = 0
                filt_rej[changed] = 0
                filt_resets[changed] = 0
            bad = A & ~fok
            if bad.any():
                self.rej[bad] += 1
                self.mu[bad] = before[bad]
                line_iters[bad] = 0
                st_ax[bad] = st_az[bad] = 0.0
                st_xi[bad] = st_zi[bad] = -1
                self._increase_barrier_on_rejections(np.nonzero(bad)[0], max_rej, barrier_inc, initial_barrier)
            if not B.any():
                continue
            be.residual(self.mu, self.vars, self.grad, self.res, B)
            self.tau = np.maximum(tau_min, 1.0 - self.mu) if options["adaptive_tau"] else np.full(
                K, options["fraction_to_boundary"])
            ax, xi_, az, zi_ = be.max_step(self.tau, self.vars, self.update, B)
            if options["equal_primal_dual_step"]:
                ax = az = np.minimum(ax, az)
            Sx = np.nonzero(B)[0]
            alpha, li, accepted, frej = self._line_search_with_watchdog(Sx, ax, az, self.tau, options, use_soc)
            tp = self._tick("line_search", tp)
            line_iters[B] = li[B]
            acc_m = B & accepted
            filt_rej[acc_m & frej] += 1
            filt_rej[acc_m & ~frej] = 0
            reset = acc_m & (filt_rej >= options["filter_reset_trigger"]) & (filt_resets < options["max_filter_resets"])
            if reset.any():
                self.filter.clear(np.nonzero(reset)[0])
                filt_resets[reset] += 1
                filt_rej[reset] = 0
            # Feasibility restoration is not batched: those problems leave
            # the batch and are solved again alone, starting now, beside the
            # batch's remaining iterations.
            rest = B & ~accepted
            if rest.any():
                self.fallback |= rest
                active &= ~rest
                for k in np.nonzero(rest)[0]:
                    self._start_alone(int(k), options, on_done)
            G = acc_m
            if G.any():
                eps = np.finfo(np.float64).eps
                smin = eps * np.minimum(1.0, self.mu)
                smin = np.where(smin == 0.0, np.finfo(np.float64).tiny, smin)
                be.floor_slacks(self.vars, self.mu, smin, 1.81898940354586e-12, G)
                be.reset_bound_multipliers(self.mu, 1e10, self.vars, G)
                self._update_gradient(self.vars.x, G)
                st_ax[G] = alpha[G] * ax[G]
                st_az[G] = az[G]
                st_xi[G], st_zi[G] = xi_[G], zi_[G]
                self.rej[G] = 0
                nw = G & ~self.in_wd
                full = (alpha == 1.0) | (line_iters == 1)
                self.wd_short[nw & full] = 0
                self.wd_short[nw & ~full] += 1
            self._tick("step", tp)
        self.results = data
        for k in range(K):
            data[k]["inertia_history"] = list(ic.history[k])
            data[k]["solver_inertia_history"] = self._solver_history(k)
            data[k]["num_iterations"] = len(data[k]["iterations"]) - 1
            data[k]["time"] = time.perf_counter() - start
            data[k]["batched"] = not self.fallback[k]
        self.batch_time = time.perf_counter() - start
        if on_done is not None:
            # Problems still iterating when the iteration limit ended the loop.
            for k in np.nonzero(active)[0]:
                data[k]["inertia_history"] = list(ic.history[k])
                data[k]["solver_inertia_history"] = self._solver_history(int(k))
                data[k]["num_iterations"] = len(data[k]["iterations"]) - 1
                data[k]["batched"] = True
                on_done(int(k), data[k], self._solution_of(int(k)))
        # The problems that left the batch: their alone solves (started when
        # they left) replace the batch records.
        for k, fut in sorted(self._alone_futs.items()):
            _, o, d = fut.result()
            d["batched"] = False
            data[k] = d
            self._alone[k] = o
        if self._alone_pool is not None:
            self._alone_pool.shutdown(wait=True)
            self._alone_pool = None
        self.total_time = time.perf_counter() - start
        return data

    def _start_alone(self, k, options, on_done):
        """Solve problem k again on its own (a thread with its own stream,
        no graph capture beside the batch's work); on_done, if given, gets
        its result as soon as it finishes."""
        import concurrent.futures as cf

        if self._alone_pool is None:
            self._alone_pool = cf.ThreadPoolExecutor(max_workers=8)
        opts = dict(options)

        def one():
            stream = torch.cuda.Stream(self.bp.device)
            with torch.cuda.stream(stream):
                o = Optimizer(self.problems[k], order=self.order, scaling=self.scaling,
                              precision=self.precision)
                o.solver.ldl.set_graphs(False)
                d = o.optimize(opts)
                stream.synchronize()
                d["solver_inertia_history"] = list(o.solver.inertia_history)
                d["batched"] = False
                if on_done is not None:
                    on_done(k, d, o.get_solution())
            return k, o, d

        self._alone_futs[k] = self._alone_pool.submit(one)

    def _solver_history(self, k):
        """Inertia of every factorization of problem k, as the
        single-problem solver records it (AmigoSolver.inertia_history)."""
        h = [] if getattr(self, "_ls_inertia", None) is None or self._ls_inertia[k] is None \
            else [self._ls_inertia[k]]
        return h + list(self.ic.history[k])

    def _solution_of(self, k):
        """Problem k's unscaled x (host copy) from the batch vectors."""
        be = self.be
        seg = slice(int(self.bp.xseg[k]), int(self.bp.xseg[k + 1]))
        x = self.vars.x[seg].cpu().numpy().copy()
        if be.scaling_active:
            c0, c1 = int(be.cseg[k]), int(be.cseg[k + 1])
            ci = be.cidx_g_h[c0:c1] - int(self.bp.xseg[k])
            x[ci] /= be.constr_scale[c0:c1].cpu().numpy()
        return x

    def get_solution(self, k=None):
        """Primal and multiplier values in the unscaled space (host copies),
        for problem k or a list of all problems."""
        x = self.vars.x.clone()
        self.be.unscale_multipliers(x)
        xh = x.cpu().numpy()
        seg = self.bp.xseg
        sols = []
        for j in range(self.K):
            if getattr(self, "_alone", None) and j in self._alone:
                sols.append(self._alone[j].get_solution())
            else:
                sols.append(xh[seg[j]:seg[j + 1]].copy())
        return sols if k is None else sols[k]
