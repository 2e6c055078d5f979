"""Pins the numpy restatement of the interior-point kernels (oracle/ipm.py)
with properties the mathematics fixes, not by retyping its formulas:
stationary points give zero residual, the bound-dual step is the
linearised complementarity, the fraction-to-boundary step is the largest
feasible step found by bisection, the barrier directional derivative equals
a finite difference of the barrier function, and the element Lagrangians
match closed forms and an independent mass-matrix solve."""

import numpy as np
import pytest

from oracle import ipm as O


def _problem(rng, n_p=40, n_c=15, frac_bounds=0.7):
    n = n_p + n_c
    mult = np.zeros(n, bool)
    mult[rng.choice(n, n_c, replace=False)] = True
    lower = np.full(n, -np.inf)
    upper = np.full(n, np.inf)
    p = np.nonzero(~mult)[0]
    has_l = rng.random(n_p) < frac_bounds
    has_u = rng.random(n_p) < frac_bounds
    lower[p[has_l]] = rng.uniform(-2, 0, has_l.sum())
    upper[p[has_u]] = rng.uniform(1, 3, has_u.sum())
    c = np.nonzero(mult)[0]
    lower[c] = upper[c] = rng.normal(size=n_c)
    return O.Info(mult, lower, upper)


def _interior_state(rng, P, mu=None):
    x = rng.normal(size=P.n)
    xp = x[P.pidx]
    xp = np.where(np.isfinite(P.lbx) & np.isfinite(P.ubx), 0.5 * (P.lbx + P.ubx), xp)
    xp = np.where(np.isfinite(P.lbx) & ~np.isfinite(P.ubx), P.lbx + 0.5, xp)
    xp = np.where(~np.isfinite(P.lbx) & np.isfinite(P.ubx), P.ubx - 0.5, xp)
    x[P.pidx] = xp
    sl = np.where(P.fl, xp - P.lbx, 1.0)
    su = np.where(P.fu, P.ubx - xp, 1.0)
    if mu is None:
        zl = np.where(P.fl, rng.uniform(0.1, 2, len(xp)), 0.0)
        zu = np.where(P.fu, rng.uniform(0.1, 2, len(xp)), 0.0)
    else:
        zl = np.where(P.fl, mu / sl, 0.0)
        zu = np.where(P.fu, mu / su, 0.0)
    return x, zl, zu, sl, su


def test_residual_vanishes_at_a_central_point_and_has_the_right_signs():
    rng = np.random.default_rng(0)
    P = _problem(rng)
    mu = 0.1
    S = _interior_state(rng, P, mu=mu)
    _, zl, zu, sl, su = S
    grad = np.zeros(P.n)
    grad[P.pidx] = zl - zu              # stationarity of the Lagrangian
    grad[P.cidx] = P.lbh                 # constraints satisfied
    res, nrm2, dual2, primal2 = O.residual(mu, P, S, grad)
    assert np.max(np.abs(res)) < 1e-13 and nrm2 < 1e-25
    # d res / d grad = -I on every row
    e = np.zeros(P.n)
    e[P.pidx[3]] = 1e-3
    e[P.cidx[2]] = 2e-3
    res2, *_ = O.residual(mu, P, S, grad + e)
    assert np.allclose(res2 - res, -e, atol=1e-15)
    # raising mu by dmu moves a lower-bounded-only row by +dmu/sl
    i = np.nonzero(P.fl & ~P.fu)[0][0]
    res3, *_ = O.residual(mu + 0.05, P, S, grad)
    assert np.isclose(res3[P.pidx[i]], 0.05 / sl[i])


def test_bound_dual_step_is_linearised_complementarity():
    rng = np.random.default_rng(1)
    P = _problem(rng)
    S = _interior_state(rng, P)
    _, zl, zu, sl, su = S
    mu = 0.3
    px = rng.normal(size=P.n)
    dzl, dzu = O.bound_dual_step(mu, P, S, px)
    dx = px[P.pidx]
    # (sl + dx)(zl + dzl) = mu and (su - dx)(zu + dzu) = mu to first order
    assert np.allclose((sl * dzl + zl * dx)[P.fl], (mu - sl * zl)[P.fl])
    assert np.allclose((su * dzu - zu * dx)[P.fu], (mu - su * zu)[P.fu])
    assert np.all(dzl[~P.fl] == 0) and np.all(dzu[~P.fu] == 0)


def test_max_step_is_the_largest_fraction_to_boundary_step():
    rng = np.random.default_rng(2)
    P = _problem(rng)
    S = _interior_state(rng, P)
    x, zl, zu, sl, su = S
    px = 3.0 * rng.normal(size=P.n)
    dzl = rng.normal(size=len(P.pidx))
    dzu = rng.normal(size=len(P.pidx))
    tau = 0.99
    ax, xi, az, zi = O.max_step(tau, P, S, px, dzl, dzu)

    def ok_x(a):
        dx = a * px[P.pidx]
        return np.all((sl + dx)[P.fl] >= (1 - tau) * sl[P.fl] - 1e-14) and np.all(
            (su - dx)[P.fu] >= (1 - tau) * su[P.fu] - 1e-14)

    def ok_z(a):
        return np.all((zl + a * dzl)[P.fl] >= (1 - tau) * zl[P.fl] - 1e-14) and np.all(
            (zu + a * dzu)[P.fu] >= (1 - tau) * zu[P.fu] - 1e-14)

    for a, ok in ((ax, ok_x), (az, ok_z)):
        assert 0 < a <= 1 and ok(a)
        if a < 1:
            assert not ok(a * (1 + 1e-9))
    assert xi in P.pidx and zi in P.pidx


def test_apply_step_keeps_slacks_consistent_with_bounds():
    rng = np.random.default_rng(3)
    P = _problem(rng)
    S = _interior_state(rng, P)
    d = rng.normal(size=P.n)
    dzl = rng.normal(size=len(P.pidx))
    dzu = rng.normal(size=len(P.pidx))
    x, zl, zu, sl, su = O.apply_step(0.3, 0.7, P, S, d, dzl, dzu)
    xp = x[P.pidx]
    assert np.allclose((xp - P.lbx)[P.fl], sl[P.fl])
    assert np.allclose((P.ubx - xp)[P.fu], su[P.fu])
    assert np.allclose(zl[P.fl], (S[1] + 0.7 * dzl)[P.fl])


def test_barrier_dphi_equals_directional_derivative_of_the_barrier_function():
    """On a QP with equality constraints J x = b, the step of the condensed
    KKT system gives dphi = d/dt phi(x + t dx) at t = 0 for
    phi = f - mu sum log(sl) - mu sum log(su)."""
    rng = np.random.default_rng(4)
    P = _problem(rng, n_p=30, n_c=10)
    mu = 0.05
    S = _interior_state(rng, P)
    x, zl, zu, sl, su = S
    np_, nc = len(P.pidx), len(P.cidx)
    Q = rng.normal(size=(np_, np_))
    Q = Q @ Q.T / np_ + np.eye(np_)
    c = rng.normal(size=np_)
    J = rng.normal(size=(nc, np_))
    xp, lam = x[P.pidx], x[P.cidx]
    grad = np.zeros(P.n)
    grad[P.pidx] = Q @ xp + c + J.T @ lam
    grad[P.cidx] = J @ xp
    res, *_ = O.residual(mu, P, S, grad)
    diag = O.diagonal(P, S)
    delta_c = 1e-6
    diag[P.cidx] = -delta_c
    K = np.block([[Q + np.diag(diag[P.pidx]), J.T], [J, -delta_c * np.eye(nc)]])
    sol = np.linalg.solve(K, np.concatenate([res[P.pidx], res[P.cidx]]))
    px = np.zeros(P.n)
    px[P.pidx], px[P.cidx] = sol[:np_], sol[np_:]
    dphi = O.barrier_dphi_from_kkt(P, S, res, px, diag)

    def phi(t):
        y = xp + t * px[P.pidx]
        f = 0.5 * y @ Q @ y + c @ y
        lo = (y - P.lbx)[P.fl]
        up = (P.ubx - y)[P.fu]
        return f - mu * np.sum(np.log(lo)) - mu * np.sum(np.log(up))

    h = 1e-6
    fd = (phi(h) - phi(-h)) / (2 * h)
    assert abs(dphi - fd) < 1e-6 * max(1.0, abs(fd))
    # and the log-barrier sum is that phi's barrier part
    b = O.barrier_log_sum(mu, P, S)
    assert np.isclose(b, phi(0.0) - (0.5 * xp @ Q @ xp + c @ xp))


def test_complementarity_kkt_error_and_reset():
    rng = np.random.default_rng(5)
    P = _problem(rng)
    mu = 0.2
    S = _interior_state(rng, P, mu=mu)
    avg, xi = O.complementarity(P, S)
    assert np.isclose(avg, mu) and np.isclose(xi, 1.0)
    grad = np.zeros(P.n)
    grad[P.pidx] = S[1] - S[2]
    grad[P.cidx] = P.lbh
    d, p, c = O.kkt_error(mu, P, S, grad)
    assert max(d, p, c) < 1e-14
    assert O.constraint_violation(P, grad) == 0.0
    # reset: multipliers land in [mu/(k s), k mu/s]; inside values unchanged
    x, zl, zu, sl, su = _interior_state(rng, P)
    kappa = 10.0
    zl2, zu2 = O.reset_bound_multipliers(mu, kappa, P, (x, zl, zu, sl, su))
    lo, hi = mu / (kappa * sl), kappa * mu / sl
    assert np.all((zl2 >= lo - 1e-15)[P.fl]) and np.all((zl2 <= hi + 1e-15)[P.fl])
    inside = P.fl & (zl >= lo) & (zl <= hi)
    assert np.array_equal(zl2[inside], zl[inside])


def test_initial_point_is_strictly_interior():
    rng = np.random.default_rng(6)
    P = _problem(rng)
    lb, ub = O.relax_bounds(P)
    assert np.all((lb <= P.lbx)[P.fl]) and np.all((P.lbx - lb)[P.fl] <= 1e-4 + 1e-15)
    x0 = rng.normal(scale=5.0, size=P.n)
    x, zl, zu, sl, su = O.project_and_init(P, x0)
    xp = x[P.pidx]
    assert np.all((xp > P.lbx)[P.fl]) and np.all((xp < P.ubx)[P.fu])
    assert np.all(sl[P.fl] > 0) and np.all(su[P.fu] > 0)
    assert np.array_equal(x[P.cidx], x0[P.cidx])


def test_fd_tool_on_a_polynomial():
    f = lambda z: z[0] ** 3 + 2 * z[0] * z[1] ** 2 - z[1]
    z = np.array([0.7, -1.3])
    g, H = O.fd_grad_hess(f, z)
    assert np.allclose(g, [3 * z[0] ** 2 + 2 * z[1] ** 2, 4 * z[0] * z[1] - 1], atol=1e-8)
    assert np.allclose(H, [[6 * z[0], 4 * z[1]], [4 * z[1], 4 * z[0]]], atol=1e-4)


def test_di_interval_hessian_closed_form():
    dt, alpha = 0.1, 1.7
    z = np.random.default_rng(7).normal(size=8)
    _, H = O.fd_grad_hess(lambda w: O.di_interval(w, dt, alpha), z)
    E = np.zeros((8, 8))
    E[2, 2] = E[5, 5] = alpha * dt / 2
    h = dt / 2
    # lq * (q1 - q0 - h (v0 + v1)); lv * (v1 - v0 - h (u0 + u1))
    for j, c in ((3, 1.0), (0, -1.0), (1, -h), (4, -h)):
        E[6, j] = E[j, 6] = c
    for j, c in ((4, 1.0), (1, -1.0), (2, -h), (5, -h)):
        E[7, j] = E[j, 7] = c
    assert np.allclose(H, E, atol=1e-5)


@pytest.mark.parametrize("dim", [2, 3])
def test_elliptic_second_derivative_closed_form(dim):
    h, yd, a, alpha = 0.05, 0.3, 1e-4, 0.9
    rng = np.random.default_rng(8)
    z = rng.normal(size=2 * dim + 3)
    _, H = O.fd_grad_hess(lambda w: O.elliptic(w, h, yd, a, alpha, dim), z, eps=1e-4)
    y, lam = z[0], z[-1]
    assert np.isclose(H[0, 0], alpha * h**dim + 6 * lam * h**2 * y, rtol=1e-5, atol=1e-7)
    assert np.isclose(H[-2, -2], alpha * a * h**dim, atol=1e-7)
    assert np.isclose(H[-1, 0], 2 * dim + 3 * h**2 * y**2, rtol=1e-6)
    assert np.isclose(H[-1, -2], -h**2, atol=1e-7)


def test_cartpole_accelerations_solve_the_lagrangian_mass_matrix():
    """Cart x, pole angle th from hanging, point mass mp at (x + l sin th,
    -l cos th); Lagrange's equations give
    [[mc+mp, mp l c], [mp l c, mp l^2]] [xdd, thdd] =
    [u + mp l s w^2, -mp g l s]."""
    rng = np.random.default_rng(9)
    for _ in range(20):
        th, w, u = rng.uniform(-4, 4), rng.normal(), rng.normal(scale=10)
        mc, mp, l = rng.uniform(0.5, 2), rng.uniform(0.1, 0.5), rng.uniform(0.3, 1.0)
        s, c = np.sin(th), np.cos(th)
        Mm = np.array([[mc + mp, mp * l * c], [mp * l * c, mp * l * l]])
        rhs = np.array([u + mp * l * s * w * w, -mp * 9.81 * l * s])
        ref = np.linalg.solve(Mm, rhs)
        assert np.allclose(O.cartpole_accel(th, w, u, mc, mp, l), ref, rtol=1e-12, atol=1e-12)


def test_double_integrator_discrete_optimum():
    """Feasible, and the interval means of u converge to the continuous
    optimum u(t) = 6 - 12 t (first order: the end nodes carry half
    weight in the trapezoidal objective)."""
    errs = []
    for N in (32, 64, 128):
        q, v, u, lq, lv = O.di_discrete_optimum(N)
        dt = 1.0 / N
        assert np.allclose(q[1:] - q[:-1], dt / 2 * (v[1:] + v[:-1]), atol=1e-12)
        assert np.allclose(v[1:] - v[:-1], dt / 2 * (u[1:] + u[:-1]), atol=1e-12)
        assert abs(q[0]) < 1e-12 and abs(v[0]) < 1e-12 and abs(q[-1] - 1) < 1e-12 and abs(v[-1]) < 1e-12
        tm = (np.arange(N) + 0.5) * dt
        errs.append(np.max(np.abs(0.5 * (u[1:] + u[:-1]) - (6 - 12 * tm))))
    assert errs[-1] < 0.03 and errs[0] / errs[-1] > 3.0
