"""The interior-point optimizer: ``am.Optimizer`` on the GPU.

    opt = Optimizer(problem)                 # problem: ComponentProblem
    data = opt.optimize({"max_iterations": 100})
    x = opt.get_solution()

Same constructor arguments and result dictionary as the reference's
``am.Optimizer`` (amigo/algorithm/ipm_driver.py:41-483).  The iterations run
in the native driver (csrc/ipm_driver.cpp, include/amigo_driver.h): inertia
correction, Newton direction with iterative refinement, filter line search
with second-order correction and watchdog, restoration, barrier updates and
convergence tests are C++ over the device kernels, one library call per
``optimize``.  A single problem is the K = 1 case of the lockstep batch
(amigo_amd.batch), so a problem solved alone and in a batch takes the same
steps.
"""

import time

from .batch import BatchOptimizer
from .default_options import get_default_options


class SolverRecord:
    """Counters and histories of the KKT solver across optimize() calls (the
    reference's LinearSolver objects keep the factor; these are what callers
    read back): factorization and solve counts, the inertia of every
    factorization, mixed-precision fallbacks and refinement steps; ``ldl``
    is the SparseLDL handle (timing, work, graphs)."""

    def __init__(self, batch):
        self._b = batch
        self.ldl = batch.solver.ldl
        self.n_factor = 0
        self.n_solve = 0
        self.inertia_history = []
        self.fp64_fallbacks = 0
        self.ir_steps_total = 0
        self.keep_first = False
        self.first_values = None

    @property
    def hess(self):
        return self._b.solver.hess


class Optimizer:
    """Primal-dual interior-point optimizer of one problem on the GPU
    (ipm_driver.py:41-483)."""

    def __init__(self, problem, x=None, lower=None, upper=None, solver=None, verbose=False,
                 precision="fp64", order="AMD", scaling=False):
        if solver is not None and not (isinstance(solver, str) and solver.lower() in ("amigo", "gpu")):
            if isinstance(solver, str):
                raise ValueError(f"Unknown solver string '{solver}'. Expected 'amigo' or 'gpu'.")
            raise ValueError("the native driver factors with its own GPU LDL^T; pass solver=None")
        if x is not None or lower is not None or upper is not None:
            import copy

            import numpy as np

            problem = copy.copy(problem)  # the caller's problem keeps its own values
            if x is not None:
                problem.x0 = np.asarray(x, np.float64).copy()
            if lower is not None:
                problem.lower = np.asarray(lower, np.float64).copy()
            if upper is not None:
                problem.upper = np.asarray(upper, np.float64).copy()
        self.problem = problem
        self.verbose = verbose
        self.precision = precision
        self._batch = BatchOptimizer([problem], order=order, scaling=scaling, precision=precision)
        self.solver = SolverRecord(self._batch)
        self.profile_phases = False
        self.phase_times = {}
        self._obj_scale = 1.0
        self._last = None

    @property
    def vars(self):
        """The iterate on the device: x, zl, zu, sl, su."""
        return self._batch.vars

    def get_options(self, options={}):
        return get_default_options(options)

    def reset(self):
        """Nothing to do: every optimize() starts from the initial point."""

    def optimize(self, options={}):
        start = time.perf_counter()
        b = self._batch
        if self.solver.keep_first and self.solver.first_values is None:
            b.native.capture_first_factorization(True)
        b.profile_phases = self.profile_phases
        out = b.optimize(options)[0]
        sr = self.solver
        f, s = b.native.counts()
        sr.n_factor += f
        sr.n_solve += s
        sr.inertia_history.extend(out["solver_inertia_history"])
        sr.fp64_fallbacks += out["fp64_fallbacks"]
        sr.ir_steps_total += out["ir_steps"]
        if sr.keep_first and sr.first_values is None and b.native.capture_hess is not None:
            sr.first_values = (b.native.capture_hess.clone(), b.native.capture_shift.clone())
            b.native.capture_first_factorization(False)
        self._obj_scale = out["obj_scale"]
        if self.profile_phases:
            for k, v in b.native.phase_times.items():
                self.phase_times[k] = self.phase_times.get(k, 0.0) + v
        out["time"] = time.perf_counter() - start
        if self.verbose:
            _write_log(out["iterations"])
        self._last = out
        return out

    def get_solution(self):
        """Primal and multiplier values in the unscaled space (host copy)."""
        return self._batch.get_solution(0)

    def compute_post_opt_derivatives(self, of=[], wrt=[], method="adjoint"):
        """Derivatives of solution entries (or "objective") with respect to
        element data (group, element, slot) at the optimum
        (post_optimization.py:23-77); see amigo_amd.post_optimization."""
        from .post_optimization import PostOptimization

        return PostOptimization(self).compute_post_opt_derivatives(of, wrt, method)


def _write_log(iterations):
    for it in iterations:
        i = it["iteration"]
        if i % 20 == 0:
            print(f"{'iter':>4s}  {'nlp_error':>9s}  {'objective':>12s}  {'inf_pr':>9s}  {'inf_du':>9s}  "
                  f"{'compl':>9s}  {'mu':>9s}  {'delta_w':>8s}  {'alpha_x':>8s}  {'alpha_z':>8s}  {'ls':>2s}")
        print(f"{i:4d}  {it['nlp_error']:9.2e}  {it.get('objective', 0.0):12.5e}  {it['inf_pr']:9.2e}  "
              f"{it['inf_du']:9.2e}  {it['compl']:9.2e}  {it['barrier_param']:9.2e}  {it['inertia_delta']:8.1e}  "
              f"{it['alpha_x']:8.2e}  {it['alpha_z']:8.2e}  {it['line_iters']:2d}")
