"""Device interior-point backend: OptVector and InteriorPointOptimizer.

Method names, arguments and return tuples follow the reference binding
(amigo/amigo.cpp:747-907 over include/interior_point_optimizer.h:111-656);
every vector lives in HBM and every kernel is a HIP kernel from
csrc/ipm_kernels.hip.  Scalar reductions come back to the host.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .sparse import current_stream_ptr

RED = {
    "compl": 0,
    "kkt_err": 1,
    "logsum": 2,
    "dphi_kkt": 3,
    "theta": 4,
    "kkt_sq": 5,
    "max_comp_dev": 6,
    "compl_sq": 7,
    "dphi_direct": 8,
    "abs_sums": 9,
}


def _s():
    return _lib.vp(current_stream_ptr())


class OptVector:
    """x (primals + multipliers), bound duals zl, zu and bound slacks sl, su
    (include/interior_point_optimizer.h:21-90)."""

    def __init__(self, n, np_, x=None, device="cuda"):
        dev = torch.device(device)
        self.x = x if x is not None else torch.zeros(n, dtype=torch.float64, device=dev)
        self.zl = torch.zeros(np_, dtype=torch.float64, device=dev)
        self.zu = torch.zeros(np_, dtype=torch.float64, device=dev)
        self.sl = torch.ones(np_, dtype=torch.float64, device=dev)
        self.su = torch.ones(np_, dtype=torch.float64, device=dev)
        self._st = None

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

    def copy(self, other):
        self.x.copy_(other.x)
        self.zl.copy_(other.zl)
        self.zu.copy_(other.zu)
        self.sl.copy_(other.sl)
        self.su.copy_(other.su)

    def state(self):
        if self._st is None:
            self._st = _lib.IpmState(
                self.x.data_ptr(), self.zl.data_ptr(), self.zu.data_ptr(), self.sl.data_ptr(), self.su.data_ptr()
            )
        return ctypes.byref(self._st)


class InteriorPointOptimizer:
    """Device counterpart of InteriorPointOptimizer<T, policy>
    (include/interior_point_optimizer.h:111-656)."""

    def __init__(self, problem, lower, upper):
        self.problem = problem
        dev = problem.device
        self.device = dev
        mult = problem.get_multiplier_indicator()
        self.n = problem.n
        lower = np.asarray(lower, np.float64)
        upper = np.asarray(upper, np.float64)
        self.pidx_h = np.nonzero(~mult)[0].astype(np.int32)
        self.cidx_h = np.nonzero(mult)[0].astype(np.int32)
        self.np_ = len(self.pidx_h)
        self.nc_ = len(self.cidx_h)
        T = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
        self.pidx = T(self.pidx_h, torch.int32)
        self.cidx = T(self.cidx_h, torch.int32)
        self.lbx0_h = lower[self.pidx_h].copy()
        self.ubx0_h = upper[self.pidx_h].copy()
        self.lbx0 = T(self.lbx0_h)
        self.ubx0 = T(self.ubx0_h)
        self.lbx = self.lbx0.clone()
        self.ubx = self.ubx0.clone()
        self.lbh_h = lower[self.cidx_h].copy()
        self.lbh = T(self.lbh_h)
        self.lbh0 = self.lbh.clone()
        self.relaxed = False
        self.ws = torch.zeros(_lib.lib().amigo_ipm_workspace_doubles(), dtype=torch.float64, device=dev)
        self._out = (ctypes.c_double * 4)()
        self._set_info()
        self.obj_scale = 1.0
        self.scaling_active = False
        self.constr_scale = None
        self.scale_vec = None

    def _set_info(self):
        self._info = _lib.IpmProblem(
            self.np_, self.nc_, self.pidx.data_ptr(), self.cidx.data_ptr(),
            self.lbx.data_ptr(), self.ubx.data_ptr(), self.lbh.data_ptr(),
        )
        self.info = ctypes.byref(self._info)

    # ---- factories ---------------------------------------------------------
    def create_opt_vector(self, x=None):
        return OptVector(self.n, self.np_, x, device=self.device)

    # ---- index utilities (backend.h:88-118) --------------------------------
    def set_multipliers_value(self, v, x):
        x[self.cidx.long()] = v

    def set_design_vars_value(self, v, x):
        x[self.pidx.long()] = v

    def copy_multipliers(self, d, s):
        c = self.cidx.long()
        d[c] = s[c]

    def copy_design_vars(self, d, s):
        p = self.pidx.long()
        d[p] = s[p]

    # ---- initialization ----------------------------------------------------
    def relax_bounds(self, factor=1e-8, constr_viol_tol=1e-4):
        """ipm::relax_bounds (backend.h:124-145)."""
        if factor <= 0:
            return
        dummy = OptVector(self.n, self.np_, device=self.device)
        _lib.check(
            _lib.lib().amigo_ipm_initialize(
                _lib.dptr(self.lbx0), _lib.dptr(self.ubx0), self.np_, _lib.dptr(self.pidx),
                float(factor), float(constr_viol_tol), 1e-2, 1e-2, _lib.dptr(self.lbx),
                _lib.dptr(self.ubx), dummy.state(), 1, 0, _s(),
            ),
            "ipm",
        )
        self.relaxed = True

    def initialize_multipliers_and_slacks(self, mu, grad, vars_):
        """project_primals_into_interior + initialize_bound_duals +
        initialize_slacks (optimizer.h:186-199)."""
        _lib.check(
            _lib.lib().amigo_ipm_initialize(
                _lib.dptr(self.lbx0), _lib.dptr(self.ubx0), self.np_, _lib.dptr(self.pidx),
                0.0, 0.0, 1e-2, 1e-2, _lib.dptr(self.lbx), _lib.dptr(self.ubx), vars_.state(),
                0, 1, _s(),
            ),
            "ipm",
        )

    def has_slacks(self):
        return False

    # ---- residual / diagonal / update --------------------------------------
    def compute_residual(self, mu, vars_, grad, res):
        _lib.check(
            _lib.lib().amigo_ipm_residual(
                self.info, vars_.state(), float(mu), _lib.dptr(grad), _lib.dptr(res),
                _lib.dptr(self.ws), self._out, _s(),
            ),
            "ipm",
        )
        return float(np.sqrt(self._out[0]))

    def compute_residual_and_infeasibility(self, mu, vars_, grad, res):
        _lib.check(
            _lib.lib().amigo_ipm_residual(
                self.info, vars_.state(), float(mu), _lib.dptr(grad), _lib.dptr(res),
                _lib.dptr(self.ws), self._out, _s(),
            ),
            "ipm",
        )
        return self._out[1], self._out[2]

    def compute_diagonal(self, vars_, diag):
        _lib.check(_lib.lib().amigo_ipm_diagonal(self.info, vars_.state(), _lib.dptr(diag), _s()), "ipm")

    def compute_update(self, mu, vars_, px, upd):
        upd.x.copy_(px)
        _lib.check(
            _lib.lib().amigo_ipm_bound_dual_step(
                self.info, vars_.state(), float(mu), _lib.dptr(px), _lib.dptr(upd.zl), _lib.dptr(upd.zu), _s()
            ),
            "ipm",
        )

    def compute_max_step(self, tau, vars_, upd):
        _lib.check(
            _lib.lib().amigo_ipm_max_step(
                self.info, vars_.state(), float(tau), _lib.dptr(upd.x), _lib.dptr(upd.zl),
                _lib.dptr(upd.zu), _lib.dptr(self.ws), self._out, _s(),
            ),
            "ipm",
        )
        o = self._out
        return o[0], int(o[1]), o[2], int(o[3])

    def apply_step_update(self, ax, az, vars_, upd, tmp):
        _lib.check(
            _lib.lib().amigo_ipm_apply_step(
                self.info, vars_.state(), float(ax), float(az), _lib.dptr(upd.x), _lib.dptr(upd.zl),
                _lib.dptr(upd.zu), self.n, tmp.state(), _s(),
            ),
            "ipm",
        )

    # ---- reductions ---------------------------------------------------------
    def _reduce(self, kind, vars_, mu=0.0, a=None, b=None, c=None):
        _lib.check(
            _lib.lib().amigo_ipm_reduce(
                RED[kind], self.info, vars_.state(), float(mu), _lib.dptr(a), _lib.dptr(b), _lib.dptr(c),
                _lib.dptr(self.ws), self._out, _s(),
            ),
            "ipm",
        )
        return self._out[0], self._out[1], self._out[2], self._out[3]

    def compute_complementarity(self, vars_):
        """(avg, xi) (optimizer.h:287-300)."""
        s, cnt, mn, _ = self._reduce("compl", vars_)
        avg = s / cnt if cnt > 0 else 0.0
        if cnt == 0:
            mn = np.finfo(np.float64).max
        xi = max(0.0, min(1.0, mn / avg)) if avg > 0 else 1.0
        return avg, xi

    def compute_complementarity_sq(self, vars_):
        return self._reduce("compl_sq", vars_, 0.0)[0]

    def compute_max_comp_deviation(self, vars_, mu):
        return self._reduce("max_comp_dev", vars_, mu)[0]

    def compute_kkt_error(self, vars_, grad):
        d, p, c, _ = self._reduce("kkt_sq", vars_, 0.0, grad)
        return d, p, c

    def compute_kkt_error_mu(self, mu, vars_, grad):
        d, p, c, _ = self._reduce("kkt_err", vars_, mu, grad)
        return d, p, c

    def compute_constraint_violation_1norm(self, vars_, grad):
        return self._reduce("theta", vars_, 0.0, grad)[0]

    def compute_barrier_log_sum(self, mu, vars_):
        return self._reduce("logsum", vars_, mu)[0]

    def compute_barrier_dphi(self, mu, vars_, update, res, px, diag):
        return self._reduce("dphi_kkt", vars_, mu, res, px, diag)[0]

    def compute_barrier_dphi_direct(self, mu, vars_, grad, px):
        return self._reduce("dphi_direct", vars_, mu, grad, px)[0]

    def abs_sums(self, vars_):
        """[sum|zl|+|zu|, sum|y|, n_bounds, n_constraints]."""
        return self._reduce("abs_sums", vars_)

    def reset_bound_multipliers(self, mu, kappa, vars_):
        _lib.check(
            _lib.lib().amigo_ipm_reset_bound_multipliers(self.info, vars_.state(), float(mu), float(kappa), _s()),
            "ipm",
        )

    def compute_affine_start_point(self, beta_min, vars_, upd, dst):
        """ipm::compute_affine_start_point (backend.h:480-489)."""
        _lib.check(
            _lib.lib().amigo_ipm_affine_start_point(
                self.info, vars_.state(), float(beta_min), _lib.dptr(upd.zl), _lib.dptr(upd.zu),
                _lib.dptr(dst.zl), _lib.dptr(dst.zu), _s(),
            ),
            "ipm",
        )

    def compute_dual_residual_vector(self, vars_, grad, out):
        _lib.check(
            _lib.lib().amigo_ipm_dual_residual_vector(
                self.info, vars_.state(), _lib.dptr(grad), _lib.dptr(out), self.n, _s()
            ),
            "ipm",
        )

    def get_kkt_element_counts(self):
        nb = int(np.sum(np.isfinite(self.lbx0_h)) + np.sum(np.isfinite(self.ubx0_h)))
        return self.np_, self.nc_, nb

    def floor_slacks(self, vars_, mu, s_min, slack_move):
        _lib.check(
            _lib.lib().amigo_ipm_floor_slacks(
                self.info, vars_.state(), _lib.dptr(self.lbx0), _lib.dptr(self.ubx0), float(mu),
                float(s_min), float(slack_move), _lib.dptr(self.ws), self._out, _s(),
            ),
            "ipm",
        )
        return int(self._out[0])

    # ---- NLP scaling (optimizer.h:502-598) -----------------------------------
    def compute_nlp_scaling(self, x, grad, max_gradient=100.0, min_value=1e-8):
        g_obj = grad[self.pidx.long()]
        obj_max = float(torch.max(torch.abs(g_obj))) if self.np_ else 0.0
        self.obj_scale = max_gradient / obj_max if obj_max > max_gradient else 1.0
        if self.obj_scale < min_value:
            self.obj_scale = min_value
        hess = self.problem.create_matrix()
        self.problem.hessian(1.0, x, hess)
        rowp_d, _ = hess.pattern_device()
        dc = torch.ones(self.nc_, dtype=torch.float64, device=self.device)
        _lib.check(
            _lib.lib().amigo_ipm_constraint_scaling(
                self.nc_, _lib.dptr(self.cidx), _lib.dptr(rowp_d), _lib.dptr(hess.data),
                float(max_gradient), float(min_value), _lib.dptr(dc), _s(),
            ),
            "ipm",
        )
        self.constr_scale = dc
        # lbh *= dc where dc != 1 (same value either way)
        self.lbh.mul_(dc)
        self.lbh_h = self.lbh.cpu().numpy()
        sv = torch.ones(self.n, dtype=torch.float64, device=self.device)
        sv[self.cidx.long()] = dc
        self.scale_vec = sv
        self.scaling_active = True
        self._set_info()

    def apply_gradient_scaling(self, grad):
        if not self.scaling_active:
            return
        _lib.check(
            _lib.lib().amigo_ipm_scale_rows(
                self.nc_, _lib.dptr(self.cidx), _lib.dptr(self.constr_scale), _lib.dptr(grad), 0, _s()
            ),
            "ipm",
        )

    def apply_hessian_scaling(self, hess):
        if not self.scaling_active:
            return
        rowp_d, cols_d = hess.pattern_device()
        _lib.check(
            _lib.lib().amigo_csr_scale_symmetric(
                self.n, _lib.dptr(rowp_d), _lib.dptr(cols_d), _lib.dptr(self.scale_vec), _lib.dptr(hess.data), _s()
            ),
            "ipm",
        )

    def scale_multipliers(self, x):
        if self.scaling_active:
            _lib.check(
                _lib.lib().amigo_ipm_scale_rows(
                    self.nc_, _lib.dptr(self.cidx), _lib.dptr(self.constr_scale), _lib.dptr(x), 0, _s()
                ),
                "ipm",
            )

    def unscale_multipliers(self, x):
        if self.scaling_active:
            _lib.check(
                _lib.lib().amigo_ipm_scale_rows(
                    self.nc_, _lib.dptr(self.cidx), _lib.dptr(self.constr_scale), _lib.dptr(x), 1, _s()
                ),
                "ipm",
            )

    def get_obj_scale(self):
        return self.obj_scale

    def has_scaling(self):
        return self.scaling_active

    # ---- accessors ----------------------------------------------------------
    def get_num_design_variables(self):
        return self.np_

    def get_num_constraints(self):
        return self.nc_

    def get_lbx(self):
        return self.lbx0_h

    def get_ubx(self):
        return self.ubx0_h

    def get_lbx_relaxed(self):
        return self.lbx.cpu().numpy()

    def get_ubx_relaxed(self):
        return self.ubx.cpu().numpy()
