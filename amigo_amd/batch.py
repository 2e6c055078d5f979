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
* ``BatchOptimizer``: the native interior-point driver (csrc/ipm_driver.cpp)
  over the three: every problem keeps its own barrier parameter, inertia
  correction, filter, watchdog and restoration state on the host, and each
  phase of an iteration is one set of launches over the problems that take
  part in it (a mask), with one read-back of their scalars.  A problem that
  needs feasibility restoration restores inside the batch.

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

    @classmethod
    def wrap(cls, native):
        """The native driver's iterate (its x, zl, zu, sl, su tensors)."""
        v = cls.__new__(cls)
        v.x, v.zl, v.zu, v.sl, v.su = native.x, native.zl, native.zu, native.sl, native.su
        v._st = _lib.IpmState(v.x.data_ptr(), v.zl.data_ptr(), v.zu.data_ptr(), v.sl.data_ptr(),
                              v.su.data_ptr())
        return v

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

    def get_solution(self):
        return self.x

    def get_zl(self):
        return self.zl

    def get_zu(self):
        return self.zu

    def get_sl(self):
        return self.sl

    def get_su(self):
        return self.su


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
# The lockstep optimizer
# ---------------------------------------------------------------------------
class BatchOptimizer:
    """Interior-point solves of K problems in lockstep: one call of the
    native driver (include/amigo_driver.h) steps every problem, each with
    its own decisions, through shared launches."""

    def __init__(self, problems, order="ND", scaling=True, verbose=False, precision="fp64"):
        from .driver import NativeIPM

        self.problems = list(problems)
        self.order = order
        self.scaling = scaling
        self.precision = precision
        self.verbose = verbose
        t = time.perf_counter()
        bp = self.bp = BatchProblem(self.problems)
        self.K = bp.K
        self.solver = BatchSolver(bp, order=order, scaling=scaling, precision=precision)
        self.be = BatchIPM(bp, bp.lower, bp.upper)
        self.native = NativeIPM(bp, self.solver, self.be)
        self.x = self.native.x
        self.vars = BVec.wrap(self.native)
        self.setup_time = time.perf_counter() - t
        self.results = None
        self.profile_phases = False
        self.phase_times = {}
        self.fp64_fallbacks = 0
        self.iterations_run = 0
        self.restorations = np.zeros(self.K, np.int64)

    @property
    def diag(self):
        """The diagonal shift of the last factorization (device copy)."""
        d = torch.zeros(self.bp.n, dtype=torch.float64, device=self.bp.device)
        self.native.copy_vector(2, d)
        return d

    def optimize(self, options={}, on_done=None):
        """Solve every problem; returns one result dict per problem (the
        keys of Optimizer.optimize).  on_done(k, result, x) is called as
        soon as problem k has finished, with its unscaled primal-dual
        solution on the host."""
        start = time.perf_counter()
        options = get_default_options(options)
        cb = None
        if on_done is not None:
            def cb(k):
                on_done(k, self.native.result(k, options), self.native.solution(k))
        data = self.native.solve(options, self.precision, cb, profile=self.profile_phases)
        self.results = data
        self.fp64_fallbacks = int(sum(d["fp64_fallbacks"] for d in data))
        self.restorations = np.array([d["restorations"] for d in data], np.int64)
        self.iterations_run = max((d["num_iterations"] for d in data), default=0)
        for k, v in self.native.phase_times.items():
            self.phase_times[k] = self.phase_times.get(k, 0.0) + v
        for d in data:
            d["batched"] = True
        self.total_time = self.batch_time = time.perf_counter() - start
        return data

    def get_solution(self, k=None):
        """Primal and multiplier values in the unscaled space (host copies),
        for problem k or a list of all problems."""
        return self.native.solution(k)
