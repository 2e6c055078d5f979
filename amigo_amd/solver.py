"""KKT solver adapter: the LinearSolver duck type of the reference
(amigo/algorithm/solvers/linear_solver.py:6-264, amigo_solver.py:13-58) on
the GPU multifrontal LDL^T.

Differences of mechanism (not of result):
- the diagonal Sigma + delta is fused into the factorization
  (amigo_ldl_factor_shifted) instead of being added to the assembled matrix,
  so inertia-correction retries refactor without re-assembling the Hessian;
- iterative refinement runs on the device (SpMV + fused residual kernels).
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .sparse import SparseLDL, current_stream_ptr


def _s():
    return _lib.vp(current_stream_ptr())


class AmigoSolver:
    """Direct KKT solver on the GPU LDL^T with inertia (AmigoSolver,
    amigo/algorithm/solvers/amigo_solver.py:13-58)."""

    supports_inertia = True
    solver_name = "amigo_amd.SparseLDL"

    def __init__(self, problem, ustab=0.01, pivot_tol=1e-14, order="AMD", precision="fp64",
                 delay_growth=2.0, scaling=False):
        self.problem = problem
        self.hess = problem.create_matrix()
        self.nrows = self.hess.nrows
        groups = getattr(problem, "var_groups", None)
        self.ldl = SparseLDL(
            self.hess, "LDL", ustab=ustab, pivot_tol=pivot_tol, delay_growth=delay_growth,
            order=order, precision=precision, groups=groups, scaling=scaling,
        )
        self.shift = torch.zeros(self.nrows, dtype=torch.float64, device=problem.device)
        self._diag_idx = problem.diag_idx_d
        self._tmp = torch.zeros(self.nrows, dtype=torch.float64, device=problem.device)
        self.n_factor = 0
        self.n_solve = 0
        self.inertia_history = []
        self.delay_history = []
        # Optionally keep the first factorized matrix (values, shift).
        self.keep_first = False
        self.first_values = None
        self.fp64_fallbacks = 0
        self.ir_steps_total = 0

    # ---- assembly + factorization ------------------------------------------
    def _factor_current(self):
        if self.keep_first and self.first_values is None:
            self.first_values = (self.hess.data.clone(), self.shift.clone())
        self.ldl.factor(shift=self.shift)
        self.n_factor += 1
        self.inertia_history.append(self.ldl.get_inertia())
        self.delay_history.append(self.ldl.stats())

    def factor(self, alpha, x, diag, post_hessian=None):
        """Assemble the Hessian, add diag and factor (linear_solver.py:49-57)."""
        self.problem.hessian(alpha, x, self.hess)
        if post_hessian is not None:
            post_hessian(self.hess)
        self.shift.copy_(diag)
        self._factor_current()

    def add_diagonal_and_factor(self, diag):
        self.shift.copy_(diag)
        self._factor_current()

    def assemble_hessian(self, alpha, x):
        self.problem.hessian(alpha, x, self.hess)
        return self.hess.data[self._diag_idx].clone()

    def get_hessian_diagonal(self, alpha, x):
        return self.assemble_hessian(alpha, x)

    def get_inertia(self):
        return self.ldl.get_inertia()

    # ---- solves ---------------------------------------------------------------
    def solve(self, bx, px):
        px.copy_(bx)
        self.ldl.solve(px)
        self.n_solve += 1

    def kkt_matvec(self, x, y):
        """y = (H + diag(shift)) x, the matrix that was factored."""
        rowp_d, cols_d = self.hess.pattern_device()
        _lib.check(
            _lib.lib().amigo_csr_spmv(
                self.nrows, _lib.dptr(rowp_d), _lib.dptr(cols_d), _lib.dptr(self.hess.data),
                _lib.dptr(x), _lib.dptr(y), _s(),
            ),
            "ipm",
        )
        y.addcmul_(self.shift, x)

    def iterative_refinement(self, mu, mult_ind, rhs_saved, vars_, grad, lbx, ubx, px, res, corr,
                             backend, max_steps=10):
        """FP64 refinement on the unreduced 8-block system
        (linear_solver.py:109-264).  The stopping rule is the reference's:
        ratio < 1e-10 after at least one step, or a stall, at most 10.

        With the mixed-precision factor (FP32 Schur updates), refinement
        that ends above the 1e-10 ratio falls back to an FP64 factorization
        of the same matrix and refines again from a fresh solve."""
        steps, ratio = self._refine(mu, rhs_saved, vars_, grad, px, res, corr, backend, max_steps)
        if self.ldl.precision == "mixed" and not ratio < 1e-10:
            # The FP64 factor then serves the rest of this iteration's solves.
            self.ldl.factor(shift=self.shift, precision="fp64")
            self.n_factor += 1
            self.fp64_fallbacks += 1
            self.solve(rhs_saved, px)
            s2, ratio = self._refine(mu, rhs_saved, vars_, grad, px, res, corr, backend, max_steps)
            steps += s2
        self.ir_steps_total += steps
        self.last_ir = (steps, ratio)
        return steps, ratio

    def _refine(self, mu, rhs_saved, vars_, grad, px, res, corr, backend, max_steps):
        L = _lib.lib()
        info = backend.info
        st = vars_.state()
        dev = px.device
        np_ = backend.np_
        dzl = torch.empty(np_, dtype=torch.float64, device=dev)
        dzu = torch.empty(np_, dtype=torch.float64, device=dev)
        ezl = torch.empty(np_, dtype=torch.float64, device=dev)
        ezu = torch.empty(np_, dtype=torch.float64, device=dev)
        Kpx = self._tmp
        out = (ctypes.c_double * 4)()
        _lib.check(
            L.amigo_ipm_ir_init(
                info, st, float(mu), _lib.dptr(grad), _lib.dptr(rhs_saved), _lib.dptr(px),
                _lib.dptr(dzl), _lib.dptr(dzu), _lib.dptr(backend.ws), out, _s(),
            ),
            "ipm",
        )
        nrm_rhs = out[0]
        ratio_max, improve, min_steps, max_cond = 1e-10, 0.999999999, 1, 1e6
        ratio_old = 1e30
        steps = 0
        ratio = 0.0
        for step in range(max_steps):
            self.kkt_matvec(px, Kpx)
            _lib.check(
                L.amigo_ipm_ir_residual(
                    info, st, float(mu), _lib.dptr(grad), _lib.dptr(rhs_saved), _lib.dptr(Kpx),
                    _lib.dptr(px), _lib.dptr(dzl), _lib.dptr(dzu), _lib.dptr(res), _lib.dptr(ezl),
                    _lib.dptr(ezu), _lib.dptr(backend.ws), out, _s(),
                ),
                "ipm",
            )
            nrm_resid, nrm_res = out[0], out[1]
            denom = min(nrm_res, max_cond * nrm_rhs) + nrm_rhs
            ratio = nrm_resid / denom if denom > 0 else 0.0
            if ratio < ratio_max and step >= min_steps:
                break
            if step >= min_steps and ratio > improve * ratio_old:
                break
            ratio_old = ratio
            self.solve(res, corr)
            _lib.check(
                L.amigo_ipm_ir_update(
                    info, st, self.nrows, _lib.dptr(corr), _lib.dptr(ezl), _lib.dptr(ezu),
                    _lib.dptr(px), _lib.dptr(dzl), _lib.dptr(dzu), _s(),
                ),
                "ipm",
            )
            steps += 1
        return steps, ratio
