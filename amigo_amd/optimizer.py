"""Primal-dual interior-point optimizer on the device backend.

Restates the reference's Python driver (amigo/algorithm/ipm_driver.py:41-483
and its mixins) step for step: heuristic (LOQO) / monotone barrier update,
inertia correction (Algorithm IC), filter line search with second-order
correction and watchdog, feasibility restoration, the convergence tests, NLP
gradient scaling and least-squares multiplier initialization.  Every vector
stays in HBM; only scalars cross to the host.

    opt = Optimizer(problem)                 # problem: ComponentProblem
    data = opt.optimize({"max_iterations": 100})
"""

import time

import numpy as np
import torch

from .default_options import get_default_options
from .ipm import InteriorPointOptimizer
from .solver import AmigoSolver

CONTINUE, CONVERGED, CONVERGED_ACCEPTABLE, DIVERGED, PRECISION_FLOOR = 0, 1, 2, 3, 4


class Filter:
    """(phi, theta) filter with Eq. 18 margins (filter_acceptance.py:13-51)."""

    def __init__(self, gamma_phi=1e-8, gamma_theta=1e-5):
        self.entries = []
        self.gamma_phi = gamma_phi
        self.gamma_theta = gamma_theta

    def is_acceptable(self, phi, theta):
        for pf, tf in self.entries:
            if phi > pf and theta > tf:
                return False
        return True

    def add(self, phi, theta):
        pe = phi - self.gamma_phi * theta
        te = (1.0 - self.gamma_theta) * theta
        self.entries = [(p, t) for p, t in self.entries if not (p >= pe and t >= te)]
        self.entries.append((pe, te))

    def clear(self):
        self.entries.clear()

    def __len__(self):
        return len(self.entries)


class InertiaCorrector:
    """Algorithm IC of Wachter & Biegler (inertia_correction.py:13-402)."""

    _NOT_YET, _NOT_DEGEN, _DEGENERATE = 0, 1, 2
    _NO_TEST, _TEST_DC0_DX0, _TEST_DC1_DX0, _TEST_DC0_DX1, _TEST_DC1_DX1 = 0, 1, 2, 3, 4

    def __init__(self, mult_ind, mult_ind_dev, barrier_param, options):
        self.mult_ind = mult_ind
        self.mult_dev = mult_ind_dev
        self.primal_dev = ~mult_ind_dev
        self._barrier = barrier_param
        self._verbose = options.get("verbose_barrier", False)
        self.numerical_eps = 1e-12
        self._delta_x_last = 0.0
        self._delta_c_last = 0.0
        self._delta_x_curr = 0.0
        self._delta_c_curr = 0.0
        self.last_delta_w = 0.0
        self.last_delta_c = 0.0
        self._dw_init, self._dw_min, self._dw_max = 1e-4, 1e-20, 1e20
        self._kw_inc, self._kw_first_inc, self._kw_dec = 8.0, 100.0, 1.0 / 3
        self._dc_val, self._dc_exp = 1e-8, 0.25
        self._hess_degen = self._NOT_YET
        self._jac_degen = self._NOT_YET
        self._degen_iters = 0
        self._degen_iters_max = 3
        self._test_status = self._NO_TEST
        self._pivtol, self._pivtolmax = 1e-6, 0.1
        self.history = []  # (n_pos, n_neg) of every factorization

    def update_barrier(self, mu):
        self._barrier = mu

    def _delta_cd(self):
        return self._dc_val * self._barrier**self._dc_exp

    def _get_deltas_for_wrong_inertia(self):
        prev = self._delta_x_curr
        if self._delta_x_curr == 0.0:
            if self._delta_x_last == 0.0:
                self._delta_x_curr = self._dw_init
            else:
                self._delta_x_curr = max(self._dw_min, self._delta_x_last * self._kw_dec)
        else:
            if self._delta_x_last == 0.0 or 1e5 * self._delta_x_last < self._delta_x_curr:
                self._delta_x_curr *= self._kw_first_inc
            else:
                self._delta_x_curr *= self._kw_inc
        if self._delta_x_curr > self._dw_max:
            self._delta_x_curr = prev
            self._delta_x_last = 0.0
            return False
        return True

    def _perturb_for_wrong_inertia(self):
        self._finalize_test()
        if self._get_deltas_for_wrong_inertia():
            return True
        if self._delta_c_curr == 0.0:
            self._delta_c_curr = self._delta_cd()
            self._delta_x_curr = 0.0
            if self._hess_degen == self._DEGENERATE:
                self._hess_degen = self._NOT_YET
            self._test_status = self._NO_TEST
            return self._get_deltas_for_wrong_inertia()
        return False

    def _perturb_for_singularity(self):
        if self._hess_degen == self._NOT_YET or self._jac_degen == self._NOT_YET:
            ts = self._test_status
            if ts == self._TEST_DC0_DX0:
                if self._jac_degen == self._NOT_YET:
                    self._delta_c_curr = self._delta_cd()
                    self._test_status = self._TEST_DC1_DX0
                else:
                    if not self._get_deltas_for_wrong_inertia():
                        return False
                    self._test_status = self._TEST_DC0_DX1
            elif ts == self._TEST_DC1_DX0:
                self._delta_c_curr = 0.0
                if not self._get_deltas_for_wrong_inertia():
                    return False
                self._test_status = self._TEST_DC0_DX1
            elif ts == self._TEST_DC0_DX1:
                self._delta_c_curr = self._delta_cd()
                if not self._get_deltas_for_wrong_inertia():
                    return False
                self._test_status = self._TEST_DC1_DX1
            elif ts == self._TEST_DC1_DX1:
                if not self._get_deltas_for_wrong_inertia():
                    return False
        else:
            if self._delta_c_curr > 0.0:
                if not self._get_deltas_for_wrong_inertia():
                    return False
            else:
                self._delta_c_curr = self._delta_cd()
        return True

    def _finalize_test(self):
        ts = self._test_status
        if ts == self._NO_TEST:
            return
        if ts == self._TEST_DC0_DX0:
            if self._hess_degen == self._NOT_YET:
                self._hess_degen = self._NOT_DEGEN
            if self._jac_degen == self._NOT_YET:
                self._jac_degen = self._NOT_DEGEN
        elif ts == self._TEST_DC1_DX0:
            if self._hess_degen == self._NOT_YET:
                self._hess_degen = self._NOT_DEGEN
            if self._jac_degen == self._NOT_YET:
                self._degen_iters += 1
                if self._degen_iters >= self._degen_iters_max:
                    self._jac_degen = self._DEGENERATE
        elif ts == self._TEST_DC0_DX1:
            if self._jac_degen == self._NOT_YET:
                self._jac_degen = self._NOT_DEGEN
            if self._hess_degen == self._NOT_YET:
                self._degen_iters += 1
                if self._degen_iters >= self._degen_iters_max:
                    self._hess_degen = self._DEGENERATE
        elif ts == self._TEST_DC1_DX1:
            self._degen_iters += 1
            if self._degen_iters >= self._degen_iters_max:
                self._hess_degen = self._DEGENERATE
                self._jac_degen = self._DEGENERATE
        self._test_status = self._NO_TEST

    def _consider_new_system(self):
        self._finalize_test()
        if self._delta_x_curr > 0.0:
            self._delta_x_last = self._delta_x_curr
        if self._delta_c_curr > 0.0:
            self._delta_c_last = self._delta_c_curr
        if self._hess_degen == self._NOT_YET or self._jac_degen == self._NOT_YET:
            self._test_status = self._TEST_DC0_DX0
        else:
            self._test_status = self._NO_TEST
        self._delta_c_curr = self._delta_cd() if self._jac_degen == self._DEGENERATE else 0.0
        self._delta_x_curr = 0.0
        if self._hess_degen == self._DEGENERATE:
            self._get_deltas_for_wrong_inertia()

    def _increase_quality(self, solver):
        if self._pivtol >= self._pivtolmax:
            return False
        self._pivtol = min(self._pivtolmax, self._pivtol**0.5)
        return True

    def factorize(self, solver, x, diag, diag_base, zero_hessian_indices, zero_hessian_eps,
                  comm_rank, max_corrections=40, inertia_tolerance=0, obj_scale=1.0,
                  hessian_scaling_fn=None, log=None):
        """Assemble, regularize, factorize (inertia_correction.py:261-393)."""
        solver.assemble_hessian(obj_scale, x)
        if hessian_scaling_fn is not None:
            hessian_scaling_fn(solver.hess)
        n_primal = int(np.sum(~self.mult_ind))
        n_dual = int(np.sum(self.mult_ind))
        n_total = n_primal + n_dual
        itol = inertia_tolerance

        def _ok(p, q):
            return abs(p - n_primal) <= itol and abs(q - n_dual) <= itol and p + q >= n_total - itol

        diag[self.primal_dev] += self.numerical_eps
        if zero_hessian_indices is not None and zero_hessian_eps is not None:
            zi = zero_hessian_indices
            diag[zi] = torch.clamp(diag[zi], min=zero_hessian_eps)
        reg_diag = diag.clone()
        self._consider_new_system()
        augsys_improved = False

        def _apply_and_factor(first):
            diag.copy_(reg_diag)
            if self._delta_x_curr > 0:
                diag[self.primal_dev] += self._delta_x_curr
            if self._delta_c_curr > 0:
                diag[self.mult_dev] -= self._delta_c_curr
            try:
                if first:
                    solver.add_diagonal_and_factor(diag)
                else:
                    solver.add_diagonal_and_factor(diag)
                p, q = solver.get_inertia()
                self.history.append((p, q))
                return p, q, False
            except Exception:
                return 0, 0, True

        for attempt in range(max_corrections + 1):
            n_pos, n_neg, singular = _apply_and_factor(attempt == 0)
            if not singular and _ok(n_pos, n_neg):
                self.last_delta_w = self._delta_x_curr
                self.last_delta_c = self._delta_c_curr
                if self._delta_x_curr > 0 and log is not None:
                    log(f"  Inertia correction: delta_w={self._delta_x_curr:.2e}, "
                        f"delta_c={self._delta_c_curr:.2e}, attempts={attempt + 1}")
                return True
            if singular and n_dual > 0:
                if not self._perturb_for_singularity():
                    break
            elif not singular and n_neg < n_dual:
                assume_singular = True
                if not augsys_improved:
                    augsys_improved = self._increase_quality(solver)
                    if augsys_improved:
                        assume_singular = False
                if assume_singular:
                    if not self._perturb_for_singularity():
                        break
            else:
                if not self._perturb_for_wrong_inertia():
                    break
        self.last_delta_w = self._delta_x_curr
        self.last_delta_c = self._delta_c_curr
        return False


class WatchdogState:
    """Watchdog bookkeeping (filter_line_search.py:395-438)."""

    def __init__(self, backend):
        self.in_watchdog = False
        self.shortened_iter = 0
        self.trial_iter = 0
        self.iterate = None
        self.update_backup = backend.create_opt_vector()
        self.px_backup = None
        self.alpha_primal_test = 0.0
        self.theta = self.barr = self.dphi = 0.0
        self.trigger = 0
        self.max_trials = 0

    def save_iterate(self, vars_, update, px, alpha_x):
        self.iterate = vars_.get_solution().clone()
        self.update_backup.copy(update)
        self.px_backup = px.clone()
        self.alpha_primal_test = alpha_x

    def restore_iterate(self, vars_, update, px):
        vars_.get_solution().copy_(self.iterate)
        update.copy(self.update_backup)
        px.copy_(self.px_backup)

    def reset(self):
        self.in_watchdog = False
        self.shortened_iter = 0
        self.trial_iter = 0
        self.iterate = None
        self.px_backup = None


class Optimizer:
    """Primal-dual interior-point optimizer (ipm_driver.py:41-483)."""

    def __init__(self, problem, x=None, lower=None, upper=None, solver=None, verbose=False,
                 precision="fp64", order="AMD", scaling=False):
        self.problem = problem
        self.barrier_param = 1.0
        self.verbose = verbose
        dev = problem.device
        self.device = dev
        x0 = problem.x0 if x is None else np.asarray(x, np.float64)
        self.lower = problem.lower if lower is None else np.asarray(lower, np.float64)
        self.upper = problem.upper if upper is None else np.asarray(upper, np.float64)
        self.x = torch.as_tensor(x0.copy(), dtype=torch.float64, device=dev)
        self._x0 = self.x.clone()
        if solver is None or (isinstance(solver, str) and solver.lower() in ("amigo", "gpu")):
            self.solver = AmigoSolver(problem, order=order, precision=precision, scaling=scaling)
        elif isinstance(solver, str):
            raise ValueError(f"Unknown solver string '{solver}'. Expected 'amigo' or 'gpu'.")
        else:
            self.solver = solver
        self.optimizer = InteriorPointOptimizer(problem, self.lower, self.upper)
        self.vars = self.optimizer.create_opt_vector(self.x)
        self.update = self.optimizer.create_opt_vector()
        self.temp = self.optimizer.create_opt_vector()
        self.grad = problem.create_vector()
        self.res = problem.create_vector()
        self.diag = problem.create_vector()
        self.px = problem.create_vector()
        self.ir_corr = problem.create_vector()
        self.mult_ind = problem.get_multiplier_indicator()
        self.mult_dev = torch.as_tensor(self.mult_ind, device=dev)
        self.cidx = self.optimizer.cidx.long()
        self.distribute = False
        self.phase_times = {}
        self.profile_phases = False

    def _timed(self, name, fn, *args, **kw):
        """Run fn, accumulating device-synchronized wall time under name
        when profile_phases is set."""
        if not self.profile_phases:
            return fn(*args, **kw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn(*args, **kw)
        torch.cuda.synchronize()
        self.phase_times[name] = self.phase_times.get(name, 0.0) + time.perf_counter() - t
        return r

    def reset(self):
        """Restore the initial point and the unscaled problem data so that
        optimize() can run again on the resident symbolic factor."""
        be = self.optimizer
        self.x.copy_(self._x0)
        be.lbx.copy_(be.lbx0)
        be.ubx.copy_(be.ubx0)
        be.lbh.copy_(be.lbh0)
        be.scaling_active = False
        be.obj_scale = 1.0
        be._set_info()
        self.vars.zl.zero_()
        self.vars.zu.zero_()
        self.vars.sl.fill_(1.0)
        self.vars.su.fill_(1.0)

    def _log(self, msg):
        if self.verbose:
            print(msg)

    def get_options(self, options={}):
        return get_default_options(options)

    # ---- problem evaluation (newton_direction.py:17-27) ---------------------
    def _update_gradient(self, x):
        self.problem.update(x)
        self.problem.gradient(self._obj_scale, x, self.grad)
        self.optimizer.apply_gradient_scaling(self.grad)

    def _compute_barrier_objective(self, vars_):
        """phi_mu = alpha f(x) - mu sum log(gap) (filter_line_search.py:19-46)."""
        x = vars_.get_solution()
        lam = x[self.cidx].clone()
        x[self.cidx] = 0.0
        f_obj = self.problem.lagrangian(self._obj_scale, x)
        x[self.cidx] = lam
        return f_obj + self.optimizer.compute_barrier_log_sum(self.barrier_param, vars_)

    def _compute_filter_theta(self, vars_=None):
        if vars_ is None:
            vars_ = self.vars
        return self.optimizer.compute_constraint_violation_1norm(vars_, self.grad)

    # ---- initialization (iterate_initialization.py:16-69) -------------------
    def _initialize_iterate(self, options):
        self._obj_scale = 1.0
        self._hessian_scaling_fn = None
        x = self.vars.get_solution()
        self.optimizer.relax_bounds(1e-8, options["constr_viol_tol"])
        self.optimizer.set_multipliers_value(0.0, x)
        self.optimizer.initialize_multipliers_and_slacks(self.barrier_param, self.grad, self.vars)
        self._update_gradient(x)
        nlp_max_grad = options["nlp_scaling_max_gradient"]
        if nlp_max_grad > 0:
            self.optimizer.compute_nlp_scaling(x, self.grad, max_gradient=nlp_max_grad)
            self._obj_scale = self.optimizer.get_obj_scale()
            if self.optimizer.has_scaling():
                self._hessian_scaling_fn = self.optimizer.apply_hessian_scaling
                self._update_gradient(x)
                self._log(f"  NLP scaling: obj_scale={self._obj_scale:.4e}")
        if options["init_affine_step_multipliers"]:
            self._compute_least_squares_multipliers()
            self.barrier_param = self._compute_affine_multipliers(beta_min=self.barrier_param)
        elif options["init_least_squares_multipliers"]:
            self._compute_least_squares_multipliers()
        self._update_gradient(x)

    def _compute_least_squares_multipliers(self, lambda_max=1e3):
        """multiplier_initialization.py:19-58."""
        x = self.vars.get_solution()
        self.optimizer.compute_dual_residual_vector(self.vars, self.grad, self.res)
        self.res.mul_(-1.0)
        self.diag.zero_()
        self.optimizer.set_design_vars_value(1.0, self.diag)
        self.solver.factor(0.0, x, self.diag, post_hessian=self._hessian_scaling_fn)
        self.solver.solve(self.res, self.px)
        lam = self.px[self.cidx]
        if lam.numel() > 0 and float(torch.max(torch.abs(lam))) > lambda_max:
            self.optimizer.set_multipliers_value(0.0, x)
            return
        self.optimizer.copy_multipliers(x, self.px)

    def _compute_affine_multipliers(self, beta_min=1.0):
        """multiplier_initialization.py:60-93."""
        x = self.vars.get_solution()
        self._update_gradient(x)
        mu = 0.0
        self.optimizer.compute_residual(mu, self.vars, self.grad, self.res)
        self.optimizer.compute_diagonal(self.vars, self.diag)
        self.solver.factor(self._obj_scale, x, self.diag, post_hessian=self._hessian_scaling_fn)
        self.solver.solve(self.res, self.px)
        self.optimizer.compute_update(mu, self.vars, self.px, self.update)
        self.optimizer.copy_multipliers(x, self.update.get_solution())
        self.optimizer.compute_affine_start_point(beta_min, self.vars, self.update, self.vars)
        barrier, _ = self.optimizer.compute_complementarity(self.vars)
        return max(barrier, beta_min)

    # ---- scaling / convergence (optimality_scaling.py, convergence_check.py)
    def _compute_optimality_scaling(self):
        s_max = 100.0
        z_asum, y_asum, n_bounds, n_constraints = self.optimizer.abs_sums(self.vars)
        s_c = 1.0 if n_bounds == 0 else max(s_max, z_asum / n_bounds) / s_max
        n_all = n_constraints + n_bounds
        s_d = 1.0 if n_all == 0 else max(s_max, (y_asum + z_asum) / n_all) / s_max
        return s_d, s_c

    def _compute_scaled_barrier_error(self, mu):
        d, p, c = self.optimizer.compute_kkt_error_mu(mu, self.vars, self.grad)
        s_d, s_c = self._compute_optimality_scaling()
        return max(d / s_d, p, c / s_c)

    def _check_convergence(self, i, options, res_norm, prev_res_norm, acc_count, floor_count,
                           kkt=None):
        tol = options["convergence_tolerance"]
        if kkt is None:
            kkt = self.optimizer.compute_kkt_error_mu(0.0, self.vars, self.grad)
        d_inf, p_inf, c_inf = kkt
        s_d, s_c = self._compute_optimality_scaling()
        err = max(d_inf / s_d, p_inf, c_inf / s_c)
        self._last_nlp_error = (err, d_inf, p_inf, c_inf)
        if (err <= tol and d_inf <= options["dual_inf_tol"] and p_inf <= options["constr_viol_tol"]
                and c_inf <= options["compl_inf_tol"]):
            return CONVERGED, err, acc_count, floor_count
        x_max = float(torch.max(torch.abs(self.vars.get_solution())))
        if x_max > options["diverging_iterates_tol"]:
            return DIVERGED, err, acc_count, floor_count
        is_acc = (err <= options["acceptable_tol"] and d_inf <= options["acceptable_dual_inf_tol"]
                  and p_inf <= options["acceptable_constr_viol_tol"]
                  and c_inf <= options["acceptable_compl_inf_tol"])
        if options["acceptable_iter"] > 0 and is_acc:
            acc_count += 1
            if acc_count >= options["acceptable_iter"]:
                return CONVERGED_ACCEPTABLE, err, acc_count, floor_count
        else:
            acc_count = 0
        rel = abs(res_norm - prev_res_norm) / max(res_norm, 1e-30)
        floor_count = floor_count + 1 if (rel < 1e-14 and i > 0) else 0
        if floor_count >= 3 and is_acc:
            return PRECISION_FLOOR, err, acc_count, floor_count
        return CONTINUE, err, acc_count, floor_count

    # ---- Newton direction (newton_direction.py:29-195) ----------------------
    def _factorize_kkt(self, x, diag_base, ic, options):
        self.diag.copy_(diag_base)
        return self._timed("factorize+IC", ic.factorize,
            self.solver, x, self.diag, diag_base, None, None, 0,
            max_corrections=options["max_inertia_corrections"],
            inertia_tolerance=options.get("inertia_tolerance", 0),
            obj_scale=self._obj_scale, hessian_scaling_fn=self._hessian_scaling_fn,
            log=self._log if self.verbose else None,
        )

    def _solve_with_mu(self, mu, ic=None):
        self.optimizer.compute_residual(mu, self.vars, self.grad, self.res)
        rhs_copy = self.res.clone()
        self._timed("solve", self.solver.solve, self.res, self.px)
        if ic is not None and hasattr(self.solver, "iterative_refinement"):
            self._timed("refinement", self.solver.iterative_refinement,
                mu, self.mult_ind, rhs_copy, self.vars, self.grad, None, None, self.px, self.res,
                self.ir_corr, self.optimizer,
            )
        self.optimizer.compute_update(mu, self.vars, self.px, self.update)

    def _find_direction(self, x, diag_base, ic, options):
        if not self._factorize_kkt(x, diag_base, ic, options):
            return False
        self._solve_with_mu(self.barrier_param, ic)
        return True

    # ---- barrier strategy (barrier_strategy/heuristic.py) -------------------
    @staticmethod
    def _loqo_heuristic(xi, comp, gamma, r, tol=1e-12):
        if xi > 1e-10:
            term = (1 - r) * (1 - xi) / xi
            factor = min(term, 2.0) ** 3
        else:
            factor = 2.0**3
        return max(gamma * factor * comp, tol), factor

    def _maybe_reduce_barrier(self, res_norm, tol, compl_inf_tol, options):
        kappa_eps = options["barrier_tol_factor"]
        heuristic = options["barrier_strategy"] == "heuristic"
        if heuristic:
            comp_h, xi_h = self.optimizer.compute_complementarity(self.vars)
        elif options["progress_based_barrier"]:
            if not res_norm <= kappa_eps * self.barrier_param:
                return
        elif not res_norm <= 0.1 * self.barrier_param:
            return
        if heuristic:
            mu_floor = min(tol, compl_inf_tol) / (kappa_eps + 1.0)
            self.barrier_param, _ = self._loqo_heuristic(
                xi_h, comp_h, options["heuristic_barrier_gamma"], options["heuristic_barrier_r"], mu_floor
            )
        else:
            while True:
                mu = self.barrier_param
                if self._compute_scaled_barrier_error(mu) > kappa_eps * mu:
                    break
                new_mu = min(options["mu_linear_decrease_factor"] * mu, mu ** options["mu_superlinear_decrease_power"])
                new_mu = max(new_mu, min(tol, compl_inf_tol) / (kappa_eps + 1.0))
                if new_mu >= mu:
                    break
                self.barrier_param = new_mu

    def _barrier_step(self, i, res_norm, x, diag_base, ic, options, st):
        tol = options["convergence_tolerance"]
        cit = options["compl_inf_tol"]
        if st["filter_monotone_mode"]:
            if res_norm <= options["barrier_progress_tol"] * st["filter_monotone_mu"]:
                st["filter_monotone_mu"] = max(st["filter_monotone_mu"] * options["monotone_barrier_fraction"], tol)
            self.barrier_param = st["filter_monotone_mu"]
        elif i > 0 and self.barrier_param > min(tol, cit) / (options["barrier_tol_factor"] + 1.0):
            self._maybe_reduce_barrier(res_norm, tol, cit, options)
        ic.update_barrier(self.barrier_param)
        return self._find_direction(x, diag_base, ic, options)

    def _increase_barrier_on_rejections(self, rejections, max_rejections, inc, initial):
        if rejections >= max_rejections:
            new = min(self.barrier_param * inc, initial)
            if new > self.barrier_param:
                self._log(f"  Barrier increased: {self.barrier_param:.2e} -> {new:.2e}")
                self.barrier_param = new
            return 0
        return rejections

    @staticmethod
    def _compute_adaptive_tau(mu, tau_min=0.99):
        return max(tau_min, 1.0 - mu)

    def _ensure_positive_slacks(self, mu):
        """bound_safeguards.py:16-87 (fused device kernel)."""
        eps = np.finfo(np.float64).eps
        s_min = eps * min(1.0, mu)
        if s_min == 0.0:
            s_min = np.finfo(np.float64).tiny
        return self.optimizer.floor_slacks(self.vars, mu, s_min, 1.81898940354586e-12)

    # ---- filter line search (filter_line_search.py:48-280) ------------------
    def _filter_line_search(self, alpha_x, alpha_z, inner_filter, options, tau, use_soc,
                            phi_current, watchdog_ref=None):
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
        EPS10 = 10.0 * np.finfo(float).eps
        if watchdog_ref is not None:
            ref_theta, ref_barr, ref_dphi = watchdog_ref
        else:
            ref_theta = self._compute_filter_theta()
            ref_barr = phi_current
            ref_dphi = self.optimizer.compute_barrier_dphi(
                self.barrier_param, self.vars, self.update, self.res, self.px, self.diag
            )
        theta_0 = self._filter_theta_0 if self._filter_theta_0 is not None else ref_theta
        theta_min = 1e-4 * max(1.0, theta_0)
        theta_max = 1e4 * max(1.0, theta_0)
        if watchdog_ref is not None:
            alpha_min = alpha_x
        else:
            alpha_min = gamma_theta
            if ref_dphi < 0.0:
                alpha_min = min(gamma_theta, gamma_phi * ref_theta / (-ref_dphi))
                if ref_theta <= theta_min:
                    alpha_min = min(alpha_min, delta * ref_theta**s_theta / (-ref_dphi) ** s_phi)
            alpha_min *= 0.05

        def is_ftype(a):
            return ref_dphi < 0.0 and a * (-ref_dphi) ** s_phi > delta * ref_theta**s_theta

        def armijo(tb, a):
            return (tb - ref_barr) - eta_phi * a * ref_dphi <= EPS10 * abs(ref_barr)

        def acceptable_to_iterate(tb, tt):
            if tb > ref_barr:
                basval = np.log10(abs(ref_barr)) if abs(ref_barr) > 10.0 else 1.0
                if np.log10(max(tb - ref_barr, 1e-300)) > obj_max_inc + basval:
                    return False
            return (tt - (1.0 - gamma_theta) * ref_theta <= EPS10 * abs(ref_theta)) or (
                (tb - ref_barr) - (-gamma_phi * ref_theta) <= EPS10 * abs(ref_barr)
            )

        def check(tb, tt, a):
            if tt > theta_max:
                return False, False
            if a > 0.0 and is_ftype(a) and ref_theta <= theta_min:
                if not armijo(tb, a):
                    return False, False
            elif not acceptable_to_iterate(tb, tt):
                return False, False
            ok = inner_filter.is_acceptable(tb, tt)
            return ok, not ok

        def update_filter(tb, a):
            if not (is_ftype(a) and armijo(tb, a)):
                inner_filter.add(ref_barr, ref_theta)

        if use_soc:
            self.optimizer.compute_residual(self.barrier_param, self.vars, self.grad, self.res)
            res_orig = self.res.clone()
            update_backup = self.optimizer.create_opt_vector()
            update_backup.copy(self.update)
            px_orig = self.px.clone()

        alpha_primal = alpha_x
        n_steps = 0
        last_filt = False
        while alpha_primal > alpha_min or n_steps == 0:
            if n_steps >= max_iters:
                break
            self.optimizer.apply_step_update(alpha_primal, alpha_z, self.vars, self.update, self.temp)
            self._update_gradient(self.temp.get_solution())
            trial_theta = self._compute_filter_theta(self.temp)
            trial_barr = self._compute_barrier_objective(self.temp)
            a_test = alpha_primal
            acc, frej = check(trial_barr, trial_theta, a_test)
            if acc:
                update_filter(trial_barr, a_test)
                self.vars.copy(self.temp)
                return alpha_primal / alpha_x, n_steps + 1, True, last_filt
            last_filt = frej
            if use_soc and n_steps == 0 and trial_theta >= ref_theta:
                c_soc = res_orig.clone()
                alpha_soc = alpha_primal
                theta_soc_old = 0.0
                soc_ok = False
                for soc_count in range(max_soc):
                    if soc_count > 0 and trial_theta > kappa_soc * theta_soc_old:
                        break
                    theta_soc_old = trial_theta
                    self.optimizer.compute_residual(self.barrier_param, self.temp, self.grad, self.res)
                    c_soc[self.mult_dev] = self.res[self.mult_dev] + alpha_soc * c_soc[self.mult_dev]
                    self.res.copy_(c_soc)
                    self._update_gradient(self.vars.get_solution())
                    try:
                        self.solver.solve(self.res, self.px)
                    except Exception:
                        break
                    self.optimizer.compute_update(self.barrier_param, self.vars, self.px, self.update)
                    soc_ax, _, soc_az, _ = self.optimizer.compute_max_step(tau, self.vars, self.update)
                    self.optimizer.apply_step_update(soc_ax, soc_az, self.vars, self.update, self.temp)
                    self._update_gradient(self.temp.get_solution())
                    trial_theta = self._compute_filter_theta(self.temp)
                    trial_barr = self._compute_barrier_objective(self.temp)
                    sacc, _ = check(trial_barr, trial_theta, a_test)
                    if sacc:
                        update_filter(trial_barr, a_test)
                        self.vars.copy(self.temp)
                        soc_ok = True
                        break
                    alpha_soc = soc_ax
                if soc_ok:
                    return 1.0, n_steps + 1, True, False
                self.update.copy(update_backup)
                self.px.copy_(px_orig)
            alpha_primal *= alpha_red
            n_steps += 1
        self._update_gradient(self.vars.get_solution())
        return alpha_primal / alpha_x, n_steps, False, last_filt

    def _filter_line_search_with_watchdog(self, alpha_x, alpha_z, inner_filter, options, tau,
                                          use_soc, watchdog, factorize_ok):
        if watchdog.in_watchdog and (not factorize_ok or alpha_x < 1e-16):
            watchdog.restore_iterate(self.vars, self.update, self.px)
            self._update_gradient(self.vars.get_solution())
            watchdog.in_watchdog = False
            watchdog.shortened_iter = 0
        if not watchdog.in_watchdog and watchdog.trigger > 0 and watchdog.shortened_iter >= watchdog.trigger:
            watchdog.save_iterate(self.vars, self.update, self.px, alpha_x)
            watchdog.theta = self._compute_filter_theta()
            watchdog.barr = self._compute_barrier_objective(self.vars)
            watchdog.dphi = self.optimizer.compute_barrier_dphi(
                self.barrier_param, self.vars, self.update, self.res, self.px, self.diag
            )
            watchdog.in_watchdog = True
            watchdog.trial_iter = 0
            self._log("  Watchdog started")
        phi_current = self._compute_barrier_objective(self.vars)
        wd_ref = (watchdog.theta, watchdog.barr, watchdog.dphi) if watchdog.in_watchdog else None
        skip_first = False
        while True:
            alpha, line_iters, acc, frej = self._filter_line_search(
                alpha_x, alpha_z, inner_filter, options, tau, use_soc, phi_current,
                watchdog_ref=wd_ref if not skip_first else None,
            )
            if watchdog.in_watchdog and not skip_first:
                if acc:
                    watchdog.in_watchdog = False
                    watchdog.shortened_iter = 0
                    break
                watchdog.trial_iter += 1
                if watchdog.trial_iter > watchdog.max_trials:
                    watchdog.restore_iterate(self.vars, self.update, self.px)
                    self._update_gradient(self.vars.get_solution())
                    watchdog.in_watchdog = False
                    watchdog.shortened_iter = 0
                    alpha_x, _, alpha_z, _ = self.optimizer.compute_max_step(tau, self.vars, self.update)
                    phi_current = self._compute_barrier_objective(self.vars)
                    skip_first = True
                    continue
                acc = True
                break
            break
        return alpha, line_iters, acc, frej

    # ---- restoration (feasibility_restoration.py:16-172) --------------------
    def _restoration_phase(self, ic, inner_filter, options, x):
        backtrack = options["backtracking_factor"]
        max_ls = options["max_line_search_iterations"]
        tau_min = options["tau_min"]
        tol = options["convergence_tolerance"]
        theta_k = self._compute_filter_theta()
        theta_start = theta_k
        theta_min_restore = 1e-4 * max(1.0, self._filter_theta_0 if self._filter_theta_0 is not None else theta_k)
        tau_of = lambda mu: self._compute_adaptive_tau(mu, tau_min) if options["adaptive_tau"] else options["fraction_to_boundary"]
        if theta_k < max(tol, theta_min_restore):
            res_current = self.optimizer.compute_residual(self.barrier_param, self.vars, self.grad, self.res)
            ax, _, az, _ = self.optimizer.compute_max_step(tau_of(self.barrier_param), self.vars, self.update)
            alpha = 1.0
            for _ in range(max_ls):
                self.optimizer.apply_step_update(alpha * ax, az, self.vars, self.update, self.temp)
                self._update_gradient(self.temp.get_solution())
                res_trial = self.optimizer.compute_residual(self.barrier_param, self.temp, self.grad, self.res)
                if res_trial < res_current:
                    phi_t = self._compute_barrier_objective(self.temp)
                    theta_t = self._compute_filter_theta(self.temp)
                    if inner_filter.is_acceptable(phi_t, theta_t):
                        self.vars.copy(self.temp)
                        return True
                    if res_trial < 0.99 * res_current:
                        self.vars.copy(self.temp)
                        inner_filter.clear()
                        self._filter_theta_0 = theta_t
                        return True
                alpha *= backtrack
            self._update_gradient(self.vars.get_solution())
            inner_filter.clear()
            self._filter_theta_0 = theta_k
            return True
        phi_k = self._compute_barrier_objective(self.vars)
        inner_filter.add(phi_k, theta_k)
        target = 0.9 * theta_k
        saved = self.barrier_param
        levels = sorted(set(m for m in [max(saved, 1e-2), 0.1, 1.0] if m >= saved))
        per_level = max(3, options["filter_restoration_max_iter"] // len(levels))
        for rmu in levels:
            self.barrier_param = rmu
            ic.update_barrier(rmu)
            stop = False
            for _ in range(per_level):
                xs = self.vars.get_solution()
                self._update_gradient(xs)
                self.optimizer.compute_diagonal(self.vars, self.diag)
                diag_r = self.diag.clone()
                self._find_direction(xs, diag_r, ic, options)
                ax, _, az, _ = self.optimizer.compute_max_step(tau_of(rmu), self.vars, self.update)
                alpha = 1.0
                taken = False
                for _ in range(max_ls):
                    self.optimizer.apply_step_update(alpha * ax, alpha * az, self.vars, self.update, self.temp)
                    self._update_gradient(self.temp.get_solution())
                    tt = self._compute_filter_theta(self.temp)
                    if tt < theta_k:
                        self.vars.copy(self.temp)
                        theta_k = tt
                        taken = True
                        break
                    alpha *= backtrack
                if not taken or theta_k < target:
                    stop = theta_k < target
                    break
            if theta_k < target or stop:
                break
        self.barrier_param = saved
        ic.update_barrier(saved)
        self._update_gradient(self.vars.get_solution())
        self.optimizer.compute_diagonal(self.vars, self.diag)
        return theta_k < theta_start * 0.99

    # ---- main loop (ipm_driver.py:106-483) ----------------------------------
    def optimize(self, options={}):
        start = time.perf_counter()
        options = self.get_options(options=options)
        self.reset()
        opt_data = {"options": options, "converged": False, "iterations": []}
        max_iters = options["max_iterations"]
        tau_min = options["tau_min"]
        tol = options["convergence_tolerance"]
        max_rej = options["max_consecutive_rejections"]
        barrier_inc = options["barrier_increase_factor"]
        initial_barrier = options["initial_barrier_param"]
        self.barrier_param = options["initial_barrier_param"]
        self._initialize_iterate(options)
        x = self.vars.get_solution()
        st = {"line_iters": 0, "ax": 0.0, "az": 0.0, "xi": -1, "zi": -1, "prev_res": float("inf"),
              "acc": 0, "floor": 0, "rej": 0, "filter_monotone_mode": False, "filter_monotone_mu": None,
              "filt_rej": 0, "filt_resets": 0}
        ic = InertiaCorrector(self.mult_ind, self.mult_dev, self.barrier_param, options)
        self.inertia_corrector = ic
        use_soc = bool(options["second_order_correction"])
        filter_ls = options["filter_line_search"]
        inner_filter = Filter(gamma_phi=options["filter_gamma_phi"], gamma_theta=options["filter_gamma_theta"])
        self._filter_theta_0 = None
        watchdog = WatchdogState(self.optimizer)
        watchdog.trigger = options["watchdog_shortened_iter_trigger"]
        watchdog.max_trials = options["watchdog_trial_iter_max"]
        res_norm = 0.0
        for i in range(max_iters):
            res_norm = self.optimizer.compute_residual(self.barrier_param, self.vars, self.grad, self.res)
            ic.update_barrier(self.barrier_param)
            if filter_ls and self._filter_theta_0 is None:
                self._filter_theta_0 = self._compute_filter_theta()
            kkt = self.optimizer.compute_kkt_error_mu(0.0, self.vars, self.grad)
            status, err, st["acc"], st["floor"] = self._check_convergence(
                i, options, res_norm, st["prev_res"], st["acc"], st["floor"], kkt=kkt
            )
            it = {"iteration": i, "time": time.perf_counter() - start, "residual": res_norm,
                  "barrier_param": self.barrier_param, "line_iters": st["line_iters"],
                  "alpha_x": st["ax"], "x_index": st["xi"], "alpha_z": st["az"], "z_index": st["zi"],
                  "inertia_delta": ic.last_delta_w, "inf_du": kkt[0], "inf_pr": kkt[1], "compl": kkt[2],
                  "nlp_error": err, "filter_size": len(inner_filter)}
            if options.get("record_objective", True):
                it["objective"] = self._compute_barrier_objective(self.vars)
            opt_data["iterations"].append(it)
            if self.verbose:
                self._write_log(i, it)
            if status == CONVERGED:
                opt_data["converged"] = True
                break
            if status in (CONVERGED_ACCEPTABLE, PRECISION_FLOOR):
                opt_data["converged"] = True
                opt_data["acceptable"] = True
                if status == PRECISION_FLOOR:
                    opt_data["precision_floor"] = True
                break
            if status == DIVERGED:
                break
            st["prev_res"] = res_norm
            step_rejected = False
            self.optimizer.compute_diagonal(self.vars, self.diag)
            diag_base = self.diag.clone()
            barrier_before = self.barrier_param
            factorize_ok = self._barrier_step(i, res_norm, x, diag_base, ic, options, st)
            if self.barrier_param != barrier_before:
                ic.update_barrier(self.barrier_param)
                if filter_ls:
                    inner_filter.clear()
                    self._filter_theta_0 = self._compute_filter_theta()
                watchdog.reset()
                st["filt_rej"] = 0
                st["filt_resets"] = 0
            if not factorize_ok:
                st["rej"] += 1
                self.barrier_param = barrier_before
                st["line_iters"] = 0
                st["ax"] = st["az"] = 0.0
                st["xi"] = st["zi"] = -1
                st["rej"] = self._increase_barrier_on_rejections(st["rej"], max_rej, barrier_inc, initial_barrier)
                continue
            # Condensed residual at the new mu (ipm_driver.py:314-316); the line
            # search's barrier derivative reads it.
            self.optimizer.compute_residual(self.barrier_param, self.vars, self.grad, self.res)
            tau = self._compute_adaptive_tau(self.barrier_param, tau_min) if options["adaptive_tau"] else options["fraction_to_boundary"]
            alpha_x, x_index, alpha_z, z_index = self.optimizer.compute_max_step(tau, self.vars, self.update)
            if options["equal_primal_dual_step"]:
                alpha_x = alpha_z = min(alpha_x, alpha_z)
            alpha, st["line_iters"], accepted, frej = self._timed("line_search", self._filter_line_search_with_watchdog,
                alpha_x, alpha_z, inner_filter, options, tau, use_soc, watchdog, factorize_ok
            )
            if accepted:
                st["filt_rej"] = st["filt_rej"] + 1 if frej else 0
                if st["filt_rej"] >= options["filter_reset_trigger"] and st["filt_resets"] < options["max_filter_resets"]:
                    inner_filter.clear()
                    st["filt_resets"] += 1
                    st["filt_rej"] = 0
            if not accepted:
                if self._restoration_phase(ic, inner_filter, options, x):
                    accepted = True
                    st["line_iters"] = 0
                    watchdog.shortened_iter = 0
                else:
                    step_rejected = True
                    st["rej"] += 1
            if accepted:
                self._ensure_positive_slacks(self.barrier_param)
                self.optimizer.reset_bound_multipliers(self.barrier_param, 1e10, self.vars)
                self._update_gradient(self.vars.get_solution())
            if step_rejected:
                st["ax"] = st["az"] = 0.0
                st["xi"] = st["zi"] = -1
                self.barrier_param = barrier_before
                st["rej"] = self._increase_barrier_on_rejections(st["rej"], max_rej, barrier_inc, initial_barrier)
            else:
                st["ax"] = alpha * alpha_x
                st["az"] = alpha_z
                st["xi"], st["zi"] = x_index, z_index
                st["rej"] = 0
                if not watchdog.in_watchdog:
                    if alpha == 1.0 or st["line_iters"] == 1:
                        watchdog.shortened_iter = 0
                    else:
                        watchdog.shortened_iter += 1
        opt_data["inertia_history"] = list(ic.history)
        opt_data["num_iterations"] = len(opt_data["iterations"]) - 1
        opt_data["time"] = time.perf_counter() - start
        return opt_data

    def _write_log(self, i, it):
        if i % 20 == 0:
            print(f"{'iter':>4s}  {'nlp_error':>9s}  {'objective':>12s}  {'inf_pr':>9s}  {'inf_du':>9s}  "
                  f"{'compl':>9s}  {'mu':>9s}  {'delta_w':>8s}  {'alpha_x':>8s}  {'alpha_z':>8s}  {'ls':>2s}")
        print(f"{i:4d}  {it['nlp_error']:9.2e}  {it.get('objective', 0.0):12.5e}  {it['inf_pr']:9.2e}  "
              f"{it['inf_du']:9.2e}  {it['compl']:9.2e}  {it['barrier_param']:9.2e}  {it['inertia_delta']:8.1e}  "
              f"{it['alpha_x']:8.2e}  {it['alpha_z']:8.2e}  {it['line_iters']:2d}")

    # ---- results ------------------------------------------------------------
    def get_solution(self):
        """Primal and multiplier values in the unscaled space (host copy)."""
        x = self.vars.get_solution().clone()
        self.optimizer.unscale_multipliers(x)
        return x.cpu().numpy()
