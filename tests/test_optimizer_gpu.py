"""End-to-end interior-point solves on the GPU.

- Double integrator N=256 (a QP whose bounds stay inactive): the solution
  equals the exact discrete optimum from one dense KKT solve (oracle) to
  1e-7, and every factorization has the inertia (n_primal, n_constraints).
- 2-D elliptic control 32x32 and cart-pole swing-up N=64: converged, and an
  independent CPU check of the returned point with the oracle element
  Lagrangians: constraint residual <= 1e-8, bounds respected, and the
  stationarity residual ||alpha grad f + J^T lambda - zl + zu||_inf small.
- The same problem solved twice gives bit-identical iterates.
"""

import numpy as np
import pytest

from oracle import ipm as O
from tests import assembly_oracle as AO

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _solve(prob, **kw):
    from amigo_amd.optimizer import Optimizer

    opt = Optimizer(prob, **kw)
    out = opt.optimize({"max_iterations": 200})
    return opt, out


def test_double_integrator_matches_exact_discrete_optimum():
    from amigo_amd.problems import double_integrator

    prob = double_integrator.build(N=256)
    opt, out = _solve(prob)
    assert out["converged"]
    q, v = prob.states(opt)
    u = prob.controls(opt)
    qe, ve, ue, _, _ = O.di_discrete_optimum(256)
    assert np.max(np.abs(u - ue)) < 1e-7
    assert np.max(np.abs(q - qe)) < 1e-9 and np.max(np.abs(v - ve)) < 1e-9
    n_p = int(np.sum(~prob.mult_ind))
    n_c = int(np.sum(prob.mult_ind))
    assert (n_p, n_c) == (771, 516)
    assert all(tuple(i) == (n_p, n_c) for i in opt.solver.inertia_history)
    print(f"double integrator: {out['num_iterations']} iterations, "
          f"{len(opt.solver.inertia_history)} factorizations")


def _certify(prob, opt, tol_c=1e-8, tol_d=1e-8):
    """Independent CPU check of the device iterate (the scaled problem the
    optimizer solves: objective weight obj_scale, multipliers as stored):
    constraint residual <= tol_c, bounds respected, and the stationarity
    residual <= tol_d in the scaled form of the convergence test
    (|grad L - zl + zu|_inf / s_d with the optimality scaling s_d of
    amigo/algorithm/optimality_scaling.py:19-48), with 1% room for the
    different summation order of the CPU check."""
    x = opt.vars.get_solution().cpu().numpy()
    alpha = opt._obj_scale
    _, g, _ = AO.assemble(prob, alpha, x) if prob.n < 4000 else (None, _grad_only(prob, alpha, x), None)
    mult = prob.mult_ind
    c = g[mult] - prob.lower[mult]
    assert np.max(np.abs(c)) <= tol_c, np.max(np.abs(c))
    xp = x[~mult]
    lo, hi = prob.lower[~mult], prob.upper[~mult]
    assert np.all(xp >= lo - 1e-10) and np.all(xp <= hi + 1e-10)
    zl = opt.vars.zl.cpu().numpy()
    zu = opt.vars.zu.cpu().numpy()
    stat = g[~mult] - zl + zu
    fl, fu = np.isfinite(lo), np.isfinite(hi)
    nb = int(fl.sum() + fu.sum())
    nc = int(mult.sum())
    zsum = float(np.sum(np.abs(zl)) + np.sum(np.abs(zu)))
    s_d = max(100.0, (np.sum(np.abs(x[mult])) + zsum) / max(nb + nc, 1)) / 100.0
    sd = np.max(np.abs(stat)) / s_d
    assert sd <= tol_d * 1.01, sd
    return np.max(np.abs(c)), sd


def _grad_only(prob, alpha, x):
    n = prob.n
    g = np.zeros(n)
    for grp in prob.groups:
        for e in range(grp.E):
            idx = grp.idx[e]
            ok = idx >= 0
            z = np.where(ok, x[np.maximum(idx, 0)], 0.0)
            ge = AO.cs_grad(AO.element_function(grp.etype, grp.data[e], alpha), z)
            g[idx[ok]] += ge[ok]
    return g


def test_elliptic_2d_converges_and_certifies():
    from amigo_amd.problems import elliptic

    prob = elliptic.EllipticControl(m=32, dim=2, seed=0)
    opt, out = _solve(prob, order="ND", scaling=True)
    assert out["converged"]
    assert out["iterations"][-1]["nlp_error"] <= 1e-8
    ec, ed = _certify(prob, opt)
    n_p, n_c = int(np.sum(~prob.mult_ind)), int(np.sum(prob.mult_ind))
    last = opt.solver.inertia_history[-1]
    assert tuple(last) == (n_p, n_c)
    print(f"elliptic 32^2: {out['num_iterations']} iterations, |c| {ec:.1e}, |stat| {ed:.1e}")


def test_cart_pole_swing_up():
    from amigo_amd.problems import cart_pole

    prob = cart_pole.build(N=64, mc=1.0, mp=0.3, l=0.5)
    opt, out = _solve(prob, order="ND")
    assert out["converged"]
    ec, ed = _certify(prob, opt)
    x = opt.get_solution()
    assert abs(x[prob.ith[-1]] - np.pi) < 1e-8 and abs(x[prob.ith[0]]) < 1e-8
    assert np.max(np.abs(prob.controls(opt))) <= 20.0 + 1e-8
    print(f"cart-pole N=64: {out['num_iterations']} iterations, |c| {ec:.1e}, |stat| {ed:.1e}")


def test_repeat_solve_is_bit_identical():
    from amigo_amd.problems import elliptic

    prob = elliptic.EllipticControl(m=24, dim=2, seed=2)
    from amigo_amd.optimizer import Optimizer

    opt = Optimizer(prob, order="ND", scaling=True)
    out1 = opt.optimize({"max_iterations": 100})
    x1 = opt.get_solution().copy()
    opt.reset()
    out2 = opt.optimize({"max_iterations": 100})
    x2 = opt.get_solution()
    assert out1["num_iterations"] == out2["num_iterations"]
    assert np.array_equal(x1, x2)


def test_elliptic_3d_converges_and_certifies():
    from amigo_amd.problems import elliptic

    prob = elliptic.EllipticControl(m=10, dim=3, seed=0)
    opt, out = _solve(prob, order="ND", scaling=True)
    assert out["converged"]
    assert out["iterations"][-1]["nlp_error"] <= 1e-8
    ec, ed = _certify(prob, opt)
    n_p, n_c = int(np.sum(~prob.mult_ind)), int(np.sum(prob.mult_ind))
    assert tuple(opt.solver.inertia_history[-1]) == (n_p, n_c)
    print(f"elliptic 10^3: {out['num_iterations']} iterations, |c| {ec:.1e}, |stat|/s_d {ed:.1e}")
