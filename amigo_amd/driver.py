"""ctypes binding of the native interior-point driver (include/amigo_driver.h).

``NativeIPM`` runs the driver over a ``BatchProblem`` of K >= 1 problems
(K = 1 is the single solve), its ``BatchSolver`` (the union KKT factor) and
``BatchIPM`` (index maps, bounds and block tables).  Every iteration of
every problem runs in C++ inside one library call; the result dictionaries
carry the keys of the reference's ``optimize()``
(amigo/algorithm/ipm_driver.py:238-483).
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .sparse import current_stream_ptr

_V = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_L = ctypes.c_longlong


class NlpGroup(ctypes.Structure):
    _fields_ = [("etype", _I), ("E", _I), ("ndata", _I), ("idx", _V), ("data", _V), ("blk_prob", _V),
                ("blk_prob_host", _V), ("alpha_blk", _V), ("goff", _L), ("hoff", _L), ("boff", _I)]


class NlpDesc(ctypes.Structure):
    _fields_ = [
        ("K", _I), ("n", _I), ("nnz", _I), ("n_primal", _I), ("n_constraints", _I),
        ("xseg_host", _V), ("pseg_host", _V), ("cseg_host", _V),
        ("x", _V), ("x0", _V), ("lbx0", _V), ("ubx0", _V), ("lbx", _V), ("ubx", _V), ("lbh0", _V), ("lbh", _V),
        ("pidx", _V), ("cidx", _V), ("pidx_global", _V), ("cidx_global", _V), ("cidx_global32", _V),
        ("row_prob", _V), ("is_mult", _V), ("xseg", _V), ("pseg", _V), ("cseg", _V), ("xseg64", _V), ("hseg64", _V),
        ("nblk_total", _I), ("blk_total", _V), ("first_total", _V), ("count_total", _V),
        ("nblk_primal", _I), ("blk_primal", _V), ("first_primal", _V), ("count_primal", _V),
        ("ngroups", _I), ("groups", ctypes.POINTER(NlpGroup)), ("gbuf", _V), ("hbuf", _V), ("lpart", _V),
        ("g_ptr", _V), ("g_src", _V), ("h_ptr", _V), ("h_src", _V), ("lptr", _V), ("lidx", _V),
        ("rowp", _V), ("cols", _V), ("hess", _V), ("ldl", _V),
        ("zl", _V), ("zu", _V), ("sl", _V), ("su", _V), ("dc", _V),
        ("capture_hess", _V), ("capture_shift", _V),
    ]


_OPTION_FIELDS = [
    ("max_iterations", _I), ("barrier_heuristic", _I), ("progress_based_barrier", _I),
    ("convergence_tolerance", _D), ("dual_inf_tol", _D), ("constr_viol_tol", _D), ("compl_inf_tol", _D),
    ("diverging_iterates_tol", _D), ("fraction_to_boundary", _D), ("initial_barrier_param", _D),
    ("max_line_search_iterations", _I), ("backtracking_factor", _D), ("equal_primal_dual_step", _I),
    ("init_least_squares_multipliers", _I), ("init_affine_step_multipliers", _I),
    ("heuristic_barrier_gamma", _D), ("heuristic_barrier_r", _D), ("second_order_correction", _I),
    ("filter_gamma_theta", _D), ("filter_gamma_phi", _D), ("filter_delta", _D), ("filter_s_theta", _D),
    ("filter_s_phi", _D), ("filter_eta_phi", _D), ("filter_max_soc", _I), ("filter_kappa_soc", _D),
    ("filter_restoration_max_iter", _I), ("filter_reset_trigger", _I), ("max_filter_resets", _I),
    ("watchdog_shortened_iter_trigger", _I), ("watchdog_trial_iter_max", _I), ("acceptable_tol", _D),
    ("acceptable_iter", _I), ("acceptable_dual_inf_tol", _D), ("acceptable_constr_viol_tol", _D),
    ("acceptable_compl_inf_tol", _D), ("adaptive_tau", _I), ("tau_min", _D), ("barrier_progress_tol", _D),
    ("max_consecutive_rejections", _I), ("barrier_increase_factor", _D), ("max_inertia_corrections", _I),
    ("inertia_tolerance", _I), ("mu_linear_decrease_factor", _D), ("mu_superlinear_decrease_power", _D),
    ("barrier_tol_factor", _D), ("nlp_scaling_max_gradient", _D), ("record_objective", _I), ("precision", _I),
    ("profile", _I),
]


class IpmOptions(ctypes.Structure):
    _fields_ = _OPTION_FIELDS


class IpmIteration(ctypes.Structure):
    _fields_ = [(k, _I) for k in ("iteration", "line_iters", "x_index", "z_index", "filter_size")] + [
        (k, _D) for k in ("time", "residual", "barrier_param", "alpha_x", "alpha_z", "inertia_delta", "inf_du",
                          "inf_pr", "compl", "nlp_error", "objective")]


class IpmResult(ctypes.Structure):
    _fields_ = [("status", _I), ("num_records", _I), ("num_inertia", _I), ("num_solver_inertia", _I),
                ("fp64_fallbacks", _I), ("ir_steps", _I), ("restorations", _I), ("obj_scale", _D),
                ("time", _D)]


DONE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int)

CONVERGED, ACCEPTABLE, DIVERGED, PRECISION_FLOOR = 1, 2, 3, 4
PHASES = ["initialize", "convergence+barrier", "factorize+IC", "solve+refinement", "line_search",
          "restoration", "step"]

_SIGS = {
    "amigo_ipm_driver_last_error": (ctypes.c_char_p, []),
    "amigo_ipm_driver_create": (_I, [ctypes.POINTER(NlpDesc), ctypes.POINTER(_V)]),
    "amigo_ipm_driver_destroy": (_I, [_V]),
    "amigo_ipm_driver_solve": (_I, [_V, ctypes.POINTER(IpmOptions), DONE_FN, _V, _V]),
    "amigo_ipm_driver_result": (_I, [_V, _I, ctypes.POINTER(IpmResult)]),
    "amigo_ipm_driver_iterations": (_I, [_V, _I, ctypes.POINTER(IpmIteration), _I]),
    "amigo_ipm_driver_inertia": (_I, [_V, _I, _I, ctypes.POINTER(_I), _I]),
    "amigo_ipm_driver_counts": (_I, [_V, ctypes.POINTER(_L), ctypes.POINTER(_L)]),
    "amigo_ipm_driver_copy_vector": (_I, [_V, _I, _V, _V]),
    "amigo_ipm_driver_phase_times": (_I, [_V, ctypes.POINTER(_D)]),
    "amigo_ipm_regularization_trace": (_I, [ctypes.POINTER(_I), _I, ctypes.POINTER(_D), _I,
                                            ctypes.POINTER(_D), _I]),
}


def lib():
    L = _lib.lib()
    if not getattr(L, "_driver_sigs", False):
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L._driver_sigs = True
    return L


def check(rc):
    if rc != 0:
        msg = lib().amigo_ipm_driver_last_error()
        raise _lib.AmigoError(rc, msg.decode() if msg else "")
    return rc


def make_options(options, precision, profile=False):
    """The option dictionary (default_options.py keys) as the C struct."""
    strategy = options["barrier_strategy"]
    if strategy not in ("heuristic", "monotone"):
        raise ValueError(f"barrier_strategy {strategy!r} is not supported (heuristic or monotone)")
    o = IpmOptions()
    for name, ctype in _OPTION_FIELDS:
        if name == "barrier_heuristic":
            v = strategy == "heuristic"
        elif name == "precision":
            v = 1 if precision == "mixed" else 0
        elif name == "profile":
            v = profile
        elif name == "record_objective":
            v = options.get("record_objective", True)
        else:
            v = options[name]
        setattr(o, name, int(bool(v)) if ctype is _I and isinstance(v, bool) else ctype(v).value)
    return o


def regularization_trace(outcomes, mus, has_constraints=True, cap=1000):
    """(delta_w, delta_c) of every factorization the inertia-correction rules
    ask for, given the factorization outcomes (0 correct, 1 singular, 2 too
    few negative pivots, 3 wrong inertia; -1 starts a new matrix whose mu is
    the next entry of mus)."""
    o = np.ascontiguousarray(outcomes, np.int32)
    m = np.ascontiguousarray(mus, np.float64)
    out = np.zeros(2 * cap)
    n = lib().amigo_ipm_regularization_trace(
        o.ctypes.data_as(ctypes.POINTER(_I)), len(o), m.ctypes.data_as(ctypes.POINTER(_D)),
        int(bool(has_constraints)), out.ctypes.data_as(ctypes.POINTER(_D)), cap)
    if n < 0:
        check(n)
    return out[: 2 * n].reshape(n, 2)


class NativeIPM:
    """The native driver over a BatchProblem (bp), its BatchSolver (solver)
    and BatchIPM (be)."""

    def __init__(self, bp, solver, be):
        dev = bp.device
        self.K, self.n = bp.K, bp.n
        T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
        self.xseg_h = np.ascontiguousarray(bp.xseg, np.int32)
        self.pseg_h = np.ascontiguousarray(be.pseg, np.int32)
        self.cseg_h = np.ascontiguousarray(be.cseg, np.int32)
        self.x = T(bp.x0, torch.float64)
        self.x0 = self.x.clone()
        z = lambda: torch.zeros(max(be.np_, 1), dtype=torch.float64, device=dev)
        self.zl, self.zu, self.sl, self.su = z(), z(), z(), z()
        self.dc = torch.ones(max(be.nc_, 1), dtype=torch.float64, device=dev)
        self.row_prob = T(np.repeat(np.arange(self.K, dtype=np.int32), np.diff(bp.xseg)), torch.int32)
        self.is_mult = T(bp.mult_ind.astype(np.int32), torch.int32)
        self.cidx_g32 = be.cidx_g.to(torch.int32)
        rowp_d, cols_d = solver.hess.pattern_device()
        self._keep = [rowp_d, cols_d]
        grps = []
        self._bprob_h = []
        for G in bp.ugroups:
            bh = np.ascontiguousarray(G.bprob, np.int32)
            self._bprob_h.append(bh)
            grps.append(NlpGroup(G.etype, G.E, G.nd, G.idx_d.data_ptr(), G.data_d.data_ptr(),
                                 G.bprob32_d.data_ptr(), bh.ctypes.data, G.alpha_d.data_ptr(), G.goff, G.hoff,
                                 G.boff))
        self._groups = (NlpGroup * max(len(grps), 1))(*grps)
        (nt, bt, ft, ct), (npb, bpp, fp, cp) = be._tabs
        self.capture_hess = None
        self.capture_shift = None
        p = lambda t: t.data_ptr()
        self.desc = NlpDesc(
            self.K, bp.n, bp.nnz, be.np_, be.nc_,
            self.xseg_h.ctypes.data, self.pseg_h.ctypes.data, self.cseg_h.ctypes.data,
            p(self.x), p(self.x0), p(be.lbx0), p(be.ubx0), p(be.lbx), p(be.ubx), p(be.lbh0), p(be.lbh),
            p(be.pidx), p(be.cidx), p(be.pidx_g), p(be.cidx_g), p(self.cidx_g32),
            p(self.row_prob), p(self.is_mult), p(be.xseg_d), p(be.pseg_d), p(be.cseg_d),
            p(bp.xseg64_d), p(bp.hseg64_d),
            nt, p(bt), p(ft), p(ct), npb, p(bpp), p(fp), p(cp),
            len(grps), self._groups, p(bp.gbuf), p(bp.hbuf), p(bp.lpart),
            p(bp.g_ptr_d), p(bp.g_src_d), p(bp.h_ptr_d), p(bp.h_src_d), p(bp.lptr_d), p(bp.lidx_d),
            p(rowp_d), p(cols_d), p(solver.hess.data), solver.ldl._h.value,
            p(self.zl), p(self.zu), p(self.sl), p(self.su), p(self.dc), None, None,
        )
        self.bp, self.solver, self.be = bp, solver, be
        h = ctypes.c_void_p()
        check(lib().amigo_ipm_driver_create(ctypes.byref(self.desc), ctypes.byref(h)))
        self._h = h
        self.phase_times = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().amigo_ipm_driver_destroy(h)
            except Exception:
                pass
            self._h = None

    def capture_first_factorization(self, on=True):
        """Keep the values and shift of the next solve's first Newton-system
        factorization (capture_hess, capture_shift)."""
        if on:
            dev = self.x.device
            self.capture_hess = torch.zeros(self.bp.nnz, dtype=torch.float64, device=dev)
            self.capture_shift = torch.zeros(self.n, dtype=torch.float64, device=dev)
            self.desc.capture_hess = self.capture_hess.data_ptr()
            self.desc.capture_shift = self.capture_shift.data_ptr()
        else:
            self.capture_hess = self.capture_shift = None
            self.desc.capture_hess = self.desc.capture_shift = None
        # the driver keeps its own copy of the description: make a new one
        lib().amigo_ipm_driver_destroy(self._h)
        h = ctypes.c_void_p()
        check(lib().amigo_ipm_driver_create(ctypes.byref(self.desc), ctypes.byref(h)))
        self._h = h

    def solve(self, options, precision="fp64", on_done=None, profile=False):
        """Run optimize(options) on every problem; returns one result dict per
        problem.  on_done(k) is called (from this thread) when problem k
        finishes."""
        o = make_options(options, precision, profile)
        cb = DONE_FN(lambda ctx, k: on_done(int(k))) if on_done is not None else DONE_FN()
        stream = _lib.vp(current_stream_ptr())
        check(lib().amigo_ipm_driver_solve(self._h, ctypes.byref(o), cb, None, stream))
        pt = (_D * 8)()
        lib().amigo_ipm_driver_phase_times(self._h, pt)
        self.phase_times = dict(zip(PHASES, pt[:7]))
        return [self.result(k, options) for k in range(self.K)]

    def counts(self):
        f, s = _L(), _L()
        check(lib().amigo_ipm_driver_counts(self._h, ctypes.byref(f), ctypes.byref(s)))
        return f.value, s.value

    def copy_vector(self, which, dst):
        check(lib().amigo_ipm_driver_copy_vector(self._h, int(which), _lib.vp(dst.data_ptr()),
                                                 _lib.vp(current_stream_ptr())))

    def result(self, k, options):
        r = IpmResult()
        check(lib().amigo_ipm_driver_result(self._h, k, ctypes.byref(r)))
        its = (IpmIteration * max(r.num_records, 1))()
        n = lib().amigo_ipm_driver_iterations(self._h, k, its, r.num_records)
        record_obj = options.get("record_objective", True)
        iters = []
        for i in range(n):
            it = its[i]
            d = {"iteration": it.iteration, "time": it.time, "residual": it.residual,
                 "barrier_param": it.barrier_param, "line_iters": it.line_iters, "alpha_x": it.alpha_x,
                 "x_index": it.x_index, "alpha_z": it.alpha_z, "z_index": it.z_index,
                 "inertia_delta": it.inertia_delta, "inf_du": it.inf_du, "inf_pr": it.inf_pr,
                 "compl": it.compl, "nlp_error": it.nlp_error, "filter_size": it.filter_size}
            if record_obj:
                d["objective"] = it.objective
            iters.append(d)
        out = {"options": options, "converged": r.status in (CONVERGED, ACCEPTABLE, PRECISION_FLOOR),
               "iterations": iters}
        if r.status in (ACCEPTABLE, PRECISION_FLOOR):
            out["acceptable"] = True
        if r.status == PRECISION_FLOOR:
            out["precision_floor"] = True
        out["inertia_history"] = self._inertia(k, 0, r.num_inertia)
        out["solver_inertia_history"] = self._inertia(k, 1, r.num_solver_inertia)
        out["num_iterations"] = len(iters) - 1
        out["time"] = r.time
        out["status"] = r.status
        out["obj_scale"] = r.obj_scale
        out["fp64_fallbacks"] = r.fp64_fallbacks
        out["ir_steps"] = r.ir_steps
        out["restorations"] = r.restorations
        return out

    def _inertia(self, k, which, count):
        buf = (_I * (2 * max(count, 1)))()
        n = lib().amigo_ipm_driver_inertia(self._h, k, which, buf, count)
        return [(buf[2 * i], buf[2 * i + 1]) for i in range(n)]

    def solution(self, k=None):
        """Primal and multiplier values (multipliers unscaled) on the host:
        one array for problem k, or the list of all."""
        x = self.x.clone()
        if self.be.nc_:
            x[self.be.cidx_g] = x[self.be.cidx_g] / self.dc[: self.be.nc_]
        xh = x.cpu().numpy()
        xs = self.xseg_h
        if k is not None:
            return xh[xs[k]:xs[k + 1]].copy()
        return [xh[xs[j]:xs[j + 1]].copy() for j in range(self.K)]
