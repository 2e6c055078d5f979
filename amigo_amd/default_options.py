"""Interior-point options and their defaults.

The keys and default values are the reference's option table
(amigo/algorithm/default_options.py:9-122) so that option dictionaries written
for the reference work unchanged; unknown keys raise.  One key is added:
``record_objective`` (store the barrier objective in every iteration record).
"""

_DEFAULTS = {
    "max_iterations": 100,
    "barrier_strategy": "heuristic",
    "monotone_barrier_fraction": 0.1,
    "convergence_tolerance": 1e-8,
    "dual_inf_tol": 1.0,
    "constr_viol_tol": 1e-4,
    "compl_inf_tol": 1e-4,
    "diverging_iterates_tol": 1e20,
    "fraction_to_boundary": 0.99,
    "initial_barrier_param": 1.0,
    "max_line_search_iterations": 40,
    "check_update_step": False,
    "backtracking_factor": 0.5,
    "record_components": [],
    "equal_primal_dual_step": False,
    "init_least_squares_multipliers": True,
    "init_affine_step_multipliers": False,
    "heuristic_barrier_gamma": 0.1,
    "heuristic_barrier_r": 0.95,
    "verbose_barrier": False,
    "continuation_control": None,
    "regularization_eps_x": 1e-8,
    "regularization_eps_z": 1e-8,
    "adaptive_regularization": True,
    "max_regularization": 1e-2,
    "regularization_increase_factor": 10.0,
    "armijo_constant": 1e-4,
    "use_armijo_line_search": True,
    "second_order_correction": True,
    "filter_line_search": True,
    "filter_gamma_theta": 1e-5,
    "filter_gamma_phi": 1e-8,
    "filter_delta": 1.0,
    "filter_s_theta": 1.1,
    "filter_s_phi": 2.3,
    "filter_eta_phi": 1e-8,
    "filter_max_soc": 4,
    "filter_kappa_soc": 0.99,
    "filter_restoration_max_iter": 20,
    "filter_reset_trigger": 5,
    "max_filter_resets": 5,
    "watchdog_shortened_iter_trigger": 10,
    "watchdog_trial_iter_max": 3,
    "acceptable_tol": 1e-6,
    "acceptable_iter": 15,
    "acceptable_dual_inf_tol": 1e10,
    "acceptable_constr_viol_tol": 1e-2,
    "acceptable_compl_inf_tol": 1e-2,
    "adaptive_tau": True,
    "tau_min": 0.99,
    "progress_based_barrier": True,
    "barrier_progress_tol": 10.0,
    "zero_hessian_variables": [],
    "regularization_eps_x_zero_hessian": 1.0,
    "max_consecutive_rejections": 5,
    "barrier_increase_factor": 5.0,
    "max_inertia_corrections": 40,
    "inertia_tolerance": 0,
    "mu_max_fact": 1e3,
    "mu_min": 1e-11,
    "mu_linear_decrease_factor": 0.2,
    "mu_superlinear_decrease_power": 1.5,
    "barrier_tol_factor": 10.0,
    "adaptive_mu_globalization": "obj-constr-filter",
    "adaptive_mu_kkterror_red_iters": 4,
    "adaptive_mu_kkterror_red_fact": 0.9999,
    "adaptive_mu_monotone_init_factor": 0.8,
    "adaptive_mu_safeguard_factor": 0.0,
    "quality_function_sigma_max": 100.0,
    "quality_function_sigma_min": 1e-6,
    "quality_function_section_sigma_tol": 1e-2,
    "quality_function_section_qf_tol": 0.0,
    "quality_function_golden_iters": 8,
    "quality_function_norm_scaling": True,
    "quality_function_centrality": "none",
    "quality_function_balancing_term": "none",
    "quality_function_predictor_corrector": False,
    "nlp_scaling_max_gradient": 100.0,
    "debug_cycling": False,
    "record_objective": True,
}


def get_default_options(options={}):
    """Merged options; unrecognised names raise ValueError."""
    merged = dict(_DEFAULTS)
    merged["record_components"] = []
    merged["zero_hessian_variables"] = []
    for name, value in options.items():
        if name not in merged:
            raise ValueError(f"Unrecognized option {name}")
        merged[name] = value
    return merged
