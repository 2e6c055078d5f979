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


# ---------------------------------------------------------------------------
# Pins for the remaining restated functions: each check is chosen so that a
# plausible slip (a wrong product, a sum for a max, a dropped term, a flipped
# sign) fails it, and the planted-mistake tests show that it does.

def _sigma_check(P, S, diag_fn, rng):
    """Sigma is the condensation of the bound-dual step: for any px,
    Sigma dx = (dzl(0) - dzl(px)) - (dzu(0) - dzu(px)) on the primal rows."""
    mu = 0.17
    px = rng.normal(size=P.n)
    dzl0, dzu0 = O.bound_dual_step(mu, P, S, np.zeros(P.n))
    dzl1, dzu1 = O.bound_dual_step(mu, P, S, px)
    lhs = diag_fn(P, S)[P.pidx] * px[P.pidx]
    rhs = (dzl0 - dzl1) - (dzu0 - dzu1)
    return np.allclose(lhs, rhs, rtol=1e-13, atol=1e-14) and np.all(diag_fn(P, S)[P.cidx] == 0.0)


def test_diagonal_is_the_condensed_bound_dual_step():
    rng = np.random.default_rng(20)
    P = _problem(rng)
    S = _interior_state(rng, P)
    assert _sigma_check(P, S, O.diagonal, rng)


def test_planted_sigma_mistake_fails_the_condensation_check():
    rng = np.random.default_rng(21)
    P = _problem(rng)
    S = _interior_state(rng, P)

    def wrong(P, S):  # zl * sl instead of zl / sl
        _, zl, zu, sl, su = S
        d = np.zeros(P.n)
        d[P.pidx] = np.where(P.fl, zl * sl, 0.0) + np.where(P.fu, zu / su, 0.0)
        return d

    assert not _sigma_check(P, S, wrong, rng)


def _planted_state(rng, P, mu):
    """A central state (every product equals mu, stationarity and
    feasibility exact) with known violations planted in each component."""
    S = _interior_state(rng, P, mu=mu)
    x, zl, zu, sl, su = (a.copy() for a in S)
    grad = np.zeros(P.n)
    grad[P.pidx] = zl - zu
    grad[P.cidx] = P.lbh
    ip = np.nonzero(P.fl)[0][:2]
    iu = np.nonzero(P.fu)[0][:1]
    # stationarity: +0.3 and -0.7 on two primal rows
    grad[P.pidx[5]] += 0.3
    grad[P.pidx[9]] -= 0.7
    # feasibility: residuals 0.25, -1.5, 0.5 on three constraint rows
    grad[P.cidx[[1, 4, 7]]] += np.array([0.25, -1.5, 0.5])
    # complementarity: sl*zl = mu + 0.4 and mu - 0.05; su*zu = mu + 0.9
    zl[ip[0]] = (mu + 0.4) / sl[ip[0]]
    zl[ip[1]] = (mu - 0.05) / sl[ip[1]]
    zu[iu[0]] = (mu + 0.9) / su[iu[0]]
    # keep the planted duals out of the stationarity rows
    for i in (ip[0], ip[1], iu[0]):
        grad[P.pidx[i]] = zl[i] - zu[i] + (0.3 if i == 5 else -0.7 if i == 9 else 0.0)
    return (x, zl, zu, sl, su), grad


def test_kkt_error_is_the_infinity_norm_of_planted_violations():
    rng = np.random.default_rng(22)
    P = _problem(rng)
    mu = 0.2
    S, grad = _planted_state(rng, P, mu)
    d, p, c = O.kkt_error(mu, P, S, grad)
    assert np.isclose(d, 0.7, rtol=1e-12) and np.isclose(p, 1.5, rtol=1e-12) and np.isclose(c, 0.9, rtol=1e-12)
    # the max deviation and the squared sums follow from the same plants
    assert np.isclose(O.max_comp_deviation(mu, P, S), 0.9, rtol=1e-12)
    assert np.isclose(O.complementarity_sq(mu, P, S), 0.4**2 + 0.05**2 + 0.9**2, rtol=1e-10)
    assert np.isclose(O.constraint_violation(P, grad), 0.25 + 1.5 + 0.5, rtol=1e-13)
    dsq, psq, csq = O.kkt_error_sq(P, S, grad)
    assert np.isclose(dsq, 0.3**2 + 0.7**2, rtol=1e-10)
    assert np.isclose(psq, 0.25**2 + 1.5**2 + 0.5**2, rtol=1e-12)
    # comp_sq sums the squared products themselves (not their deviation)
    _, zl, zu, sl, su = S
    n_l, n_u = int(P.fl.sum()), int(P.fu.sum())
    exp = (n_l - 2) * mu**2 + (mu + 0.4) ** 2 + (mu - 0.05) ** 2 + (n_u - 1) * mu**2 + (mu + 0.9) ** 2
    assert np.isclose(csq, exp, rtol=1e-10)


def test_planted_norm_mistakes_fail():
    rng = np.random.default_rng(23)
    P = _problem(rng)
    mu = 0.2
    S, grad = _planted_state(rng, P, mu)
    r = grad[P.cidx] - P.lbh
    assert not np.isclose(np.sum(np.abs(r)), np.max(np.abs(r)))   # sum-for-max would fail
    assert not np.isclose(np.sum(r**2), 0.25 + 1.5 + 0.5)          # squared-for-abs would fail


def test_complementarity_uniformity_below_one():
    rng = np.random.default_rng(24)
    P = _problem(rng)
    S = _interior_state(rng, P, mu=1.0)
    x, zl, zu, sl, su = (a.copy() for a in S)
    il = np.nonzero(P.fl)[0]
    # products 1 everywhere except one 0.1 and one 3.1 on lower bounds
    zl[il[0]] = 0.1 / sl[il[0]]
    zl[il[1]] = 3.1 / sl[il[1]]
    n = int(P.fl.sum() + P.fu.sum())
    avg, xi = O.complementarity(P, (x, zl, zu, sl, su))
    assert np.isclose(avg, (n - 2 + 0.1 + 3.1) / n, rtol=1e-13)
    assert np.isclose(xi, 0.1 / avg, rtol=1e-13) and xi < 1.0


def test_barrier_dphi_direct_is_the_barrier_directional_derivative():
    """dphi = d/dt [f(x + t dx) - mu sum log(sl + t dx) - mu sum log(su -
    t dx)] at t = 0, with grad = grad f on the primal rows."""
    rng = np.random.default_rng(25)
    P = _problem(rng)
    S = _interior_state(rng, P)
    x, zl, zu, sl, su = S
    mu = 0.08
    Q = rng.normal(size=(len(P.pidx),) * 2)
    Q = Q @ Q.T
    c = rng.normal(size=len(P.pidx))
    xp = x[P.pidx]
    grad = np.zeros(P.n)
    grad[P.pidx] = Q @ xp + c
    grad[P.cidx] = rng.normal(size=len(P.cidx))  # not read
    px = rng.normal(size=P.n)

    def phi(t):
        y = xp + t * px[P.pidx]
        return 0.5 * y @ Q @ y + c @ y - mu * np.sum(np.log((sl + t * px[P.pidx])[P.fl])) - mu * np.sum(
            np.log((su - t * px[P.pidx])[P.fu]))

    h = 1e-6
    fd = (phi(h) - phi(-h)) / (2 * h)
    assert abs(O.barrier_dphi(mu, P, S, grad, px) - fd) < 1e-6 * max(1.0, abs(fd))


def test_dual_residual_vector_is_the_residual_at_a_central_point():
    rng = np.random.default_rng(26)
    P = _problem(rng)
    mu = 0.3
    S = _interior_state(rng, P, mu=mu)
    grad = rng.normal(size=P.n)
    r = O.dual_residual_vector(P, S, grad)
    res, *_ = O.residual(mu, P, S, grad)
    assert np.allclose(r[P.pidx], -res[P.pidx], rtol=1e-13, atol=1e-14)
    assert np.all(r[P.cidx] == 0.0)


def test_affine_start_point_floors_at_beta_min():
    rng = np.random.default_rng(27)
    P = _problem(rng)
    S = _interior_state(rng, P)
    _, zl, zu, _, _ = S
    dzl = rng.normal(size=len(zl))
    dzu = rng.normal(size=len(zu))
    beta = 0.5
    zl2, zu2 = O.affine_start_point(beta, P, S, dzl, dzu)
    assert np.all(zl2[P.fl] >= beta) and np.all(zu2[P.fu] >= beta)
    big = P.fl & (zl + dzl > beta)
    assert np.array_equal(zl2[big], (zl + dzl)[big])
    assert np.array_equal(zl2[~P.fl], zl[~P.fl]) and np.array_equal(zu2[~P.fu], zu[~P.fu])


def test_optimality_scaling_closed_form():
    rng = np.random.default_rng(28)
    P = _problem(rng)
    S = _interior_state(rng, P)
    x, zl, zu, sl, su = (a.copy() for a in S)
    nb = int(P.fl.sum() + P.fu.sum())
    nc = len(P.cidx)
    # every bound dual 300, every multiplier -500
    zl[P.fl] = 300.0
    zu[P.fu] = 300.0
    zl[~P.fl] = 0.0
    zu[~P.fu] = 0.0
    x[P.cidx] = -500.0
    z, y, nbo, nco = O.abs_sums(P, (x, zl, zu, sl, su))
    assert (z, y, nbo, nco) == (300.0 * nb, 500.0 * nc, nb, nc)
    s_d, s_c = O.optimality_scaling(P, (x, zl, zu, sl, su))
    assert np.isclose(s_c, 3.0) and np.isclose(s_d, (300.0 * nb + 500.0 * nc) / (nb + nc) / 100.0)
    # small duals: both scales are 1
    zl[P.fl] = 0.5
    zu[P.fu] = 0.5
    x[P.cidx] = 0.1
    assert O.optimality_scaling(P, (x, zl, zu, sl, su)) == (1.0, 1.0)


def test_floor_slacks_resets_small_slacks_and_snaps_primals():
    rng = np.random.default_rng(29)
    P = _problem(rng)
    S = _interior_state(rng, P)
    x, zl, zu, sl, su = (a.copy() for a in S)
    mu = 1e-3
    il = np.nonzero(P.fl)[0]
    iu = np.nonzero(P.fu)[0]
    sl[il[0]] = 1e-30         # below s_min: target mu/z, capped by slack_move*max(1,|lb|)
    sl[il[1]] = -1e-20        # negative: g = 0
    su[iu[0]] = 0.0
    zl[il[0]] = 2.0
    (x2, zl2, zu2, sl2, su2), cnt = O.floor_slacks(mu, P, (x, zl, zu, sl, su), P.lbx, P.ubx)
    assert cnt == 3
    move = 1.81898940354586e-12
    assert sl2[il[0]] == min(max(mu / 2.0, np.finfo(float).eps * mu), move * max(1.0, abs(P.lbx[il[0]])) + 1e-30)
    assert sl2[il[1]] == min(max(mu / zl[il[1]], np.finfo(float).eps * mu), move * max(1.0, abs(P.lbx[il[1]])))
    assert su2[iu[0]] > 0
    # untouched slacks keep their values; every bounded primal is snapped
    keep = P.fl.copy()
    keep[il[:2]] = False
    assert np.array_equal(sl2[keep], sl[keep])
    only_l = P.fl & ~P.fu
    assert np.array_equal(x2[P.pidx][only_l], (P.lbx + sl2)[only_l])
    assert np.array_equal(x2[P.pidx][P.fu], (P.ubx - su2)[P.fu])
    # nothing below s_min: nothing changes
    S3, cnt3 = O.floor_slacks(mu, P, S, P.lbx, P.ubx)
    assert cnt3 == 0 and all(np.array_equal(a, b) for a, b in zip(S3, S))


def _kkt_csr(rng, P, scale_rows=()):
    """A symmetric KKT [[W, J^T], [J, 0]] in CSR (full pattern, diagonal
    included) with some constraint rows given large entries."""
    import scipy.sparse as sp

    n = P.n
    A = np.zeros((n, n))
    W = rng.normal(size=(len(P.pidx),) * 2)
    W = W + W.T
    W[np.abs(W) < 1.0] = 0.0
    A[np.ix_(P.pidx, P.pidx)] = W
    J = rng.normal(size=(len(P.cidx), len(P.pidx)))
    J[np.abs(J) < 0.8] = 0.0
    for j, s in scale_rows:
        J[j] *= s
    A[np.ix_(P.cidx, P.pidx)] = J
    A[np.ix_(P.pidx, P.cidx)] = J.T
    M = sp.csr_matrix(A + np.diag(np.full(n, 1e-300)))
    M.sort_indices()
    return M, A


def test_nlp_scaling_is_dense_dhd_and_bounds_row_norms():
    rng = np.random.default_rng(30)
    P = _problem(rng)
    M, A = _kkt_csr(rng, P, scale_rows=((0, 1e3), (3, 5e9), (6, 0.01)))
    grad = rng.normal(size=P.n)
    grad[P.pidx[2]] = 400.0
    obj, dc, lbh, sv = O.nlp_scaling(P, grad, M.indptr, M.data)
    assert np.isclose(obj, 100.0 / 400.0)
    rowmax = np.max(np.abs(A[P.cidx]), axis=1)
    assert np.allclose(dc, np.maximum(np.where(rowmax > 100, 100 / rowmax, 1.0), 1e-8))
    assert np.all(dc[rowmax <= 100.0] == 1.0)
    assert np.array_equal(lbh[dc == 1.0], P.lbh[dc == 1.0])
    scaled = O.hessian_scaling(M.indptr, M.indices, M.data, sv)
    D = np.diag(sv)
    dense = (D @ (A + np.diag(np.full(P.n, 1e-300))) @ D)
    import scipy.sparse as sp

    got = sp.csr_matrix((scaled, M.indices, M.indptr), shape=M.shape).toarray()
    assert np.allclose(got, dense, rtol=1e-15, atol=0)
    # after scaling, each constraint row's inf-norm is at most max_gradient
    # (rows floored at min_value excepted)
    ok = dc > 1e-8
    assert np.all(np.max(np.abs(got[P.cidx]), axis=1)[ok] <= 100.0 * (1 + 1e-14))


def _unreduced_solution(mu, P, S, grad, K):
    """Dense solve of the unreduced Newton system in (dx, dlam, dzl, dzu):
    K[dx;dlam] - dzl + dzu = -(grad_x - zl + zu) (primal rows), J dx -
    delta_c dlam = rhs_c (constraint rows), zl dx + sl dzl = mu - sl zl,
    -zu dx + su dzu = mu - su zu, with K the matrix without Sigma."""
    _, zl, zu, sl, su = S
    n, npr = P.n, len(P.pidx)
    fl, fu = P.fl, P.fu
    N = n + 2 * npr
    M = np.zeros((N, N))
    b = np.zeros(N)
    M[:n, :n] = K
    for i, idx in enumerate(P.pidx):
        M[idx, n + i] = -1.0
        M[idx, n + npr + i] = 1.0
        b[idx] = -(grad[idx] - zl[i] + zu[i])
        r = n + i
        if fl[i]:
            M[r, idx] = zl[i]
            M[r, r] = sl[i]
            b[r] = mu - sl[i] * zl[i]
        else:
            M[r, r] = 1.0
        r = n + npr + i
        if fu[i]:
            M[r, idx] = -zu[i]
            M[r, r] = su[i]
            b[r] = mu - su[i] * zu[i]
        else:
            M[r, r] = 1.0
    b[P.cidx] = -(grad[P.cidx] - P.lbh)
    sol = np.linalg.solve(M, b)
    return sol[:n], sol[n:n + npr], sol[n + npr:]


def _ir_setup(seed):
    rng = np.random.default_rng(seed)
    P = _problem(rng, n_p=30, n_c=10)
    S = _interior_state(rng, P)
    mu = 0.05
    Q = rng.normal(size=(30, 30))
    Q = Q @ Q.T / 30 + np.eye(30)
    J = rng.normal(size=(10, 30))
    dcr = 1e-8
    K = np.zeros((P.n, P.n))
    K[np.ix_(P.pidx, P.pidx)] = Q
    K[np.ix_(P.cidx, P.pidx)] = J
    K[np.ix_(P.pidx, P.cidx)] = J.T
    K[P.cidx, P.cidx] = -dcr
    grad = rng.normal(size=P.n)
    sig = O.diagonal(P, S)
    Kfull = K + np.diag(sig)
    res, *_ = O.residual(mu, P, S, grad)
    return P, S, mu, grad, K, Kfull, res


def test_refinement_reaches_the_unreduced_solution_from_an_inexact_factor():
    P, S, mu, grad, K, Kfull, res = _ir_setup(31)
    rng = np.random.default_rng(32)
    E = rng.normal(size=Kfull.shape) * 1e-3
    Kbad = Kfull + (E + E.T)  # a 1e-3 perturbed "factor"
    solve = lambda r: np.linalg.solve(Kbad, r)
    px0 = solve(res)
    px, dzl, dzu, steps, ratio, why = O.iterative_refinement(mu, P, S, grad, res, px0, lambda v: Kfull @ v, solve)
    dx, dzl_e, dzu_e = _unreduced_solution(mu, P, S, grad, K)
    assert why == "converged" and ratio < 1e-10 and 2 <= steps <= 10
    assert np.allclose(px, dx, rtol=1e-9, atol=1e-11)
    assert np.allclose(dzl[P.fl], dzl_e[P.fl], rtol=1e-9, atol=1e-11)
    assert np.allclose(dzu[P.fu], dzu_e[P.fu], rtol=1e-9, atol=1e-11)


def test_refinement_takes_one_step_with_an_exact_factor_and_stalls_on_a_useless_one():
    P, S, mu, grad, K, Kfull, res = _ir_setup(33)
    solve = lambda r: np.linalg.solve(Kfull, r)
    px, dzl, dzu, steps, ratio, why = O.iterative_refinement(mu, P, S, grad, res, solve(res), lambda v: Kfull @ v, solve)
    assert (why, steps) == ("converged", 1) and ratio < 1e-12
    # a solve that returns zero corrections cannot improve the ratio
    px0 = np.linalg.solve(Kfull + 0.01 * np.eye(P.n), res)
    _, _, _, steps, ratio, why = O.iterative_refinement(mu, P, S, grad, res, px0, lambda v: Kfull @ v,
                                                        lambda r: np.zeros_like(r))
    assert (why, steps) == ("stalled", 1) and ratio > 1e-10
    # the rule itself
    assert O.ir_stop(0, 1e-20, 1e30) is None            # minimum one step
    assert O.ir_stop(1, 9e-11, 1.0) == "converged"
    assert O.ir_stop(1, 0.5, 0.5) == "stalled"
    assert O.ir_stop(1, 0.5, 0.6) is None


def test_planted_refinement_sign_mistake_misses_the_unreduced_solution():
    P, S, mu, grad, K, Kfull, res = _ir_setup(34)
    rng = np.random.default_rng(35)
    E = rng.normal(size=Kfull.shape) * 1e-3
    solve = lambda r: np.linalg.solve(Kfull + E + E.T, r)
    orig = O.ir_blocks

    def wrong(mu, P, S, grad, rhs, Kpx, px, dzl, dzu):  # e_zu with its dx term's sign flipped
        nr, ns, ec, ezl, ezu = orig(mu, P, S, grad, rhs, Kpx, px, dzl, dzu)
        _, zl, zu, sl, su = S
        ezu2 = np.where(P.fu, ezu - 2 * zu * px[P.pidx], 0.0)
        ec2 = ec.copy()
        ec2[P.pidx] -= np.where(P.fu, (ezu2 - ezu) / np.where(P.fu, su, 1.0), 0.0)
        return nr, ns, ec2, ezl, ezu2

    O.ir_blocks = wrong
    try:
        px, *_ = O.iterative_refinement(mu, P, S, grad, res, solve(res), lambda v: Kfull @ v, solve)
    finally:
        O.ir_blocks = orig
    dx, _, _ = _unreduced_solution(mu, P, S, grad, K)
    assert not np.allclose(px, dx, rtol=1e-6, atol=1e-8)
