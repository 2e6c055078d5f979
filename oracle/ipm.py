"""Numpy restatement of the reference's interior-point kernels
(include/interior_point_backend.h) and optimizer wrappers
(include/interior_point_optimizer.h), plus the element Lagrangians of the
workloads.  ORACLE infrastructure: tests and bench.py's cpu_baseline only.

Every function cites the reference lines it restates.  Pinned by
tests/test_oracle_ipm.py against closed forms, brute-force loops written
from the reference's comments, and finite differences.
"""

import numpy as np


class Info:
    """ipm::ProblemInfo (backend.h:55-63) built like the optimizer's
    constructor (optimizer.h:114-166)."""

    def __init__(self, mult_ind, lower, upper):
        mult = np.asarray(mult_ind, bool)
        self.pidx = np.nonzero(~mult)[0]
        self.cidx = np.nonzero(mult)[0]
        self.lbx = np.asarray(lower, float)[self.pidx].copy()
        self.ubx = np.asarray(upper, float)[self.pidx].copy()
        self.lbh = np.asarray(lower, float)[self.cidx].copy()
        self.n = len(mult)

    @property
    def fl(self):
        return np.isfinite(self.lbx)

    @property
    def fu(self):
        return np.isfinite(self.ubx)


def relax_bounds(P, factor=1e-8, cvt=1e-4):
    """backend.h:124-145"""
    lb, ub = P.lbx.copy(), P.ubx.copy()
    fl, fu = np.isfinite(lb), np.isfinite(ub)
    lb[fl] = lb[fl] - np.minimum(cvt, factor * np.maximum(1.0, np.abs(lb[fl])))
    ub[fu] = ub[fu] + np.minimum(cvt, factor * np.maximum(1.0, np.abs(ub[fu])))
    return lb, ub


def project_and_init(P, x, kappa1=1e-2, kappa2=1e-2):
    """project_primals_into_interior + initialize_bound_duals +
    initialize_slacks (backend.h:79-86, :147-200)."""
    x = x.copy()
    lb, ub = P.lbx, P.ubx
    fl, fu = np.isfinite(lb), np.isfinite(ub)
    xp = x[P.pidx]
    both = fl & fu
    rng = ub - lb
    pl = np.minimum(kappa1 * np.maximum(1.0, np.abs(lb)), kappa2 * rng)
    pu = np.minimum(kappa1 * np.maximum(1.0, np.abs(ub)), kappa2 * rng)
    xp = np.where(both, np.maximum(np.minimum(xp, ub - pu), lb + pl), xp)
    onlyl = fl & ~fu
    onlyu = fu & ~fl
    xp = np.where(onlyl, np.maximum(xp, lb + kappa1 * np.maximum(1.0, np.abs(lb))), xp)
    xp = np.where(onlyu, np.minimum(xp, ub - kappa1 * np.maximum(1.0, np.abs(ub))), xp)
    x[P.pidx] = xp
    zl = np.where(fl, 1.0, 0.0)
    zu = np.where(fu, 1.0, 0.0)
    sl = np.where(fl, xp - lb, 1.0)
    su = np.where(fu, ub - xp, 1.0)
    return x, zl, zu, sl, su


def residual(mu, P, S, grad):
    """compute_residual_and_infeasibility (backend.h:227-266) and the
    squared norm of compute_residual (optimizer.h:214-225)."""
    x, zl, zu, sl, su = S
    g = grad[P.pidx]
    stat = g - zl + zu
    aug = stat.copy()
    fl, fu = P.fl, P.fu
    aug[fl] += (sl[fl] * zl[fl] - mu) / sl[fl]
    aug[fu] -= (su[fu] * zu[fu] - mu) / su[fu]
    res = np.zeros(P.n)
    res[P.pidx] = -aug
    rp = grad[P.cidx] - P.lbh
    res[P.cidx] = -rp
    return res, float(np.dot(res, res)), float(np.dot(stat, stat)), float(np.dot(rp, rp))


def diagonal(P, S):
    """compute_diagonal (backend.h:277-288)."""
    _, zl, zu, sl, su = S
    d = np.zeros(P.n)
    sig = np.where(P.fl, zl / np.where(P.fl, sl, 1.0), 0.0) + np.where(P.fu, zu / np.where(P.fu, su, 1.0), 0.0)
    d[P.pidx] = sig
    return d


def bound_dual_step(mu, P, S, px):
    """compute_bound_dual_step (backend.h:302-321)."""
    _, zl, zu, sl, su = S
    dx = px[P.pidx]
    dzl = np.where(P.fl, -((sl * zl - mu) + zl * dx) / np.where(P.fl, sl, 1.0), 0.0)
    dzu = np.where(P.fu, -((su * zu - mu) - zu * dx) / np.where(P.fu, su, 1.0), 0.0)
    return dzl, dzu


def max_step(tau, P, S, px, dzl, dzu):
    """compute_max_step (backend.h:328-369), loop form, with the binding's
    initial values ax = az = 1, xi = zi = -1 (amigo.cpp:768-775)."""
    _, zl, zu, sl, su = S
    ax = az = 1.0
    xi = zi = -1
    for i in range(len(P.pidx)):
        idx = P.pidx[i]
        dx = px[idx]
        if np.isfinite(P.lbx[i]):
            if dx < 0.0:
                a = -tau * sl[i] / dx
                if a < ax:
                    ax, xi = a, idx
            if dzl[i] < 0.0:
                a = -tau * zl[i] / dzl[i]
                if a < az:
                    az, zi = a, idx
        if np.isfinite(P.ubx[i]):
            if dx > 0.0:
                a = tau * su[i] / dx
                if a < ax:
                    ax, xi = a, idx
            if dzu[i] < 0.0:
                a = -tau * zu[i] / dzu[i]
                if a < az:
                    az, zi = a, idx
    return ax, xi, az, zi


def apply_step(ax, az, P, S, dxlam, dzl, dzu):
    """apply_step (backend.h:375-392)."""
    x, zl, zu, sl, su = S
    xn = x + ax * dxlam
    zln = np.where(P.fl, zl + az * dzl, zl)
    zun = np.where(P.fu, zu + az * dzu, zu)
    dx = ax * dxlam[P.pidx]
    sln = np.where(P.fl, sl + dx, sl)
    sun = np.where(P.fu, su - dx, su)
    return xn, zln, zun, sln, sun


def complementarity(P, S):
    """compute_complementarity + the wrapper's avg / xi (backend.h:396-414,
    optimizer.h:287-300)."""
    _, zl, zu, sl, su = S
    c = np.concatenate([(sl * zl)[P.fl], (su * zu)[P.fu]])
    if len(c) == 0:
        return 0.0, 1.0
    avg = c.sum() / len(c)
    xi = max(0.0, min(1.0, c.min() / avg)) if avg > 0 else 1.0
    return avg, xi


def kkt_error(mu, P, S, grad):
    """compute_kkt_error (backend.h:470-489): infinity norms."""
    _, zl, zu, sl, su = S
    dual = np.max(np.abs(grad[P.pidx] - zl + zu), initial=0.0)
    primal = np.max(np.abs(grad[P.cidx] - P.lbh), initial=0.0)
    comp = max(np.max(np.abs((sl * zl)[P.fl] - mu), initial=0.0), np.max(np.abs((su * zu)[P.fu] - mu), initial=0.0))
    return dual, primal, comp


def barrier_log_sum(mu, P, S):
    """compute_barrier_log_sum (backend.h:494-508)."""
    _, _, _, sl, su = S
    b = 0.0
    m = P.fl & (sl > 0)
    b -= mu * np.sum(np.log(sl[m]))
    m = P.fu & (su > 0)
    b -= mu * np.sum(np.log(su[m]))
    return b


def barrier_dphi_from_kkt(P, S, res, px, diag):
    """compute_barrier_dphi_from_kkt (backend.h:592-606)."""
    x = S[0]
    d = -np.sum(res[P.pidx] * px[P.pidx])
    jdx = res[P.cidx] - diag[P.cidx] * px[P.cidx]
    return d - np.sum(x[P.cidx] * jdx)


def constraint_violation(P, grad):
    """compute_constraint_violation_1norm (backend.h:553-562)."""
    return float(np.sum(np.abs(grad[P.cidx] - P.lbh)))


def reset_bound_multipliers(mu, kappa, P, S):
    """reset_bound_multipliers (backend.h:459-476)."""
    _, zl, zu, sl, su = S
    zl = np.where(P.fl, np.maximum(np.minimum(zl, kappa * mu / sl), mu / (kappa * sl)), zl)
    zu = np.where(P.fu, np.maximum(np.minimum(zu, kappa * mu / su), mu / (kappa * su)), zu)
    return zl, zu


# ---------------------------------------------------------------------------
# Element Lagrangians of the workloads: value, gradient, Hessian by formula.

def di_interval(z, dt, alpha):
    """Double-integrator trapezoidal collocation interval."""
    q0, v0, u0, q1, v1, u1, lq, lv = z
    h = 0.5 * dt
    cq = q1 - q0 - h * (v0 + v1)
    cv = v1 - v0 - h * (u0 + u1)
    return alpha * (u0 * u0 + u1 * u1) * dt / 4 + lq * cq + lv * cv


def elliptic(z, h, yd, a, alpha, dim):
    """Semilinear elliptic node: alpha f_e + lambda c_e."""
    nb = 2 * dim
    y = z[0]
    s = np.sum(z[1:1 + nb])
    u, lam = z[nb + 1], z[nb + 2]
    hd = h**dim
    c = (2 * dim * y - s) + h**2 * (y**3 - u)
    return alpha * (0.5 * hd * (y - yd) ** 2 + 0.5 * a * hd * u * u) + lam * c


def cartpole_accel(th, w, u, mc, mp, l, g=9.81):
    """Cart-pole accelerations with th measured from hanging: the residual
    form of examples/trajectory/cart/cart_pole.py:76-86 solved for
    (xdd, thdd)."""
    s, c = np.sin(th), np.cos(th)
    D = mc + mp * s * s
    xdd = (u + mp * s * (l * w * w + g * c)) / D
    thdd = (-u * c - mp * l * w * w * s * c - (mc + mp) * g * s) / (l * D)
    return xdd, thdd


def cartpole_interval(z, dt, mc, mp, l, alpha):
    s0, u0, s1, u1, lam = z[0:4], z[4], z[5:9], z[9], z[10:14]

    def f(s, u):
        xdd, thdd = cartpole_accel(s[1], s[3], u, mc, mp, l)
        return np.array([s[2], s[3], xdd, thdd])

    c = s1 - s0 - 0.5 * dt * (f(s0, u0) + f(s1, u1))
    return alpha * (u0 * u0 + u1 * u1) * dt / 2 + np.dot(lam, c)


def fd_grad_hess(fun, z, eps=1e-5):
    """Central finite-difference gradient and Hessian of a scalar function."""
    n = len(z)
    g = np.zeros(n)
    H = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = eps
        g[i] = (fun(z + e) - fun(z - e)) / (2 * eps)
        for j in range(n):
            f = np.zeros(n)
            f[j] = eps
            H[i, j] = (fun(z + e + f) - fun(z + e - f) - fun(z - e + f) + fun(z - e - f)) / (4 * eps * eps)
    return g, H


def di_discrete_optimum(N, T=1.0):
    """Exact optimum of the double-integrator collocation QP (bounds
    inactive): one dense KKT solve of min sum (u_k^2 + u_{k+1}^2) dt/4
    s.t. the trapezoidal defects and q(0)=v(0)=0, q(T)=1, v(T)=0.
    Returns (q, v, u) at the N+1 nodes and the defect multipliers (lq, lv)."""
    dt = T / N
    M = N + 1
    nz = 3 * M                       # z = [q_0..q_N, v_0..v_N, u_0..u_N]
    iq, iv, iu = np.arange(M), M + np.arange(M), 2 * M + np.arange(M)
    H = np.zeros((nz, nz))
    for k in range(N):
        H[iu[k], iu[k]] += dt / 2
        H[iu[k + 1], iu[k + 1]] += dt / 2
    rows = []
    rhs = []
    for k in range(N):
        r = np.zeros(nz)
        r[iq[k + 1]], r[iq[k]], r[iv[k]], r[iv[k + 1]] = 1.0, -1.0, -dt / 2, -dt / 2
        rows.append(r)
        rhs.append(0.0)
    for k in range(N):
        r = np.zeros(nz)
        r[iv[k + 1]], r[iv[k]], r[iu[k]], r[iu[k + 1]] = 1.0, -1.0, -dt / 2, -dt / 2
        rows.append(r)
        rhs.append(0.0)
    for idx, val in ((iq[0], 0.0), (iv[0], 0.0), (iq[N], 1.0), (iv[N], 0.0)):
        r = np.zeros(nz)
        r[idx] = 1.0
        rows.append(r)
        rhs.append(val)
    A = np.array(rows)
    m = A.shape[0]
    K = np.block([[H, A.T], [A, np.zeros((m, m))]])
    sol = np.linalg.solve(K, np.concatenate([np.zeros(nz), rhs]))
    z, lam = sol[:nz], sol[nz:]
    return z[iq], z[iv], z[iu], lam[:N], lam[N:2 * N]


# ---------------------------------------------------------------------------
# Further reductions and vector kernels of interior_point_backend.h.

def kkt_error_sq(P, S, grad):
    """compute_kkt_error_sq (backend.h:530-549): sums of squares of the
    stationarity rows, the constraint residuals and the complementarity
    products (not their deviation from mu)."""
    _, zl, zu, sl, su = S
    rd = grad[P.pidx] - zl + zu
    rp = grad[P.cidx] - P.lbh
    cl = (sl * zl)[P.fl]
    cu = (su * zu)[P.fu]
    return float(rd @ rd), float(rp @ rp), float(cl @ cl + cu @ cu)


def max_comp_deviation(mu, P, S):
    """compute_max_comp_deviation (backend.h:417-427), starting at 0."""
    _, zl, zu, sl, su = S
    d = np.concatenate([np.abs((sl * zl)[P.fl] - mu), np.abs((su * zu)[P.fu] - mu)])
    return float(np.max(d, initial=0.0))


def complementarity_sq(mu, P, S):
    """compute_complementarity_sq (backend.h:430-444)."""
    _, zl, zu, sl, su = S
    d = np.concatenate([(sl * zl)[P.fl] - mu, (su * zu)[P.fu] - mu])
    return float(d @ d)


def barrier_dphi(mu, P, S, grad, px):
    """compute_barrier_dphi (backend.h:513-524): grad.dx - mu dx/sl +
    mu dx/su over the primals."""
    _, _, _, sl, su = S
    dx = px[P.pidx]
    d = np.sum(grad[P.pidx] * dx)
    d -= mu * np.sum((dx / np.where(P.fl, sl, 1.0))[P.fl])
    d += mu * np.sum((dx / np.where(P.fu, su, 1.0))[P.fu])
    return float(d)


def affine_start_point(beta_min, P, S, dzl, dzu):
    """compute_affine_start_point (backend.h:480-489): z + dz floored at
    beta_min where the bound is finite; other entries unchanged."""
    _, zl, zu, _, _ = S
    zl2 = np.where(P.fl, np.maximum(zl + dzl, beta_min), zl)
    zu2 = np.where(P.fu, np.maximum(zu + dzu, beta_min), zu)
    return zl2, zu2


def dual_residual_vector(P, S, grad):
    """compute_dual_residual_vector (backend.h:566-574)."""
    _, zl, zu, _, _ = S
    out = np.zeros(P.n)
    out[P.pidx] = grad[P.pidx] - zl + zu
    return out


def abs_sums(P, S):
    """The sums behind the optimality scaling of
    amigo/algorithm/optimality_scaling.py:19-48:
    [sum |zl| + |zu|, sum |y|, number of finite bounds, number of
    constraints]."""
    x, zl, zu, _, _ = S
    return (float(np.sum(np.abs(zl)) + np.sum(np.abs(zu))), float(np.sum(np.abs(x[P.cidx]))),
            float(np.sum(P.fl) + np.sum(P.fu)), float(len(P.cidx)))


def optimality_scaling(P, S, s_max=100.0):
    """(s_d, s_c) of optimality_scaling.py:19-48."""
    z, y, nb, nc = abs_sums(P, S)
    s_c = 1.0 if nb == 0 else max(s_max, z / nb) / s_max
    s_d = 1.0 if nb + nc == 0 else max(s_max, (y + z) / (nb + nc)) / s_max
    return s_d, s_c


def floor_slacks(mu, P, S, lbx0, ubx0, slack_move=1.81898940354586e-12):
    """Slack flooring of amigo/algorithm/bound_safeguards.py:16-87.  Slacks
    below s_min = eps min(1, mu) become min(max(mu/z, s_min), slack_move
    max(1, |bound|) + max(s, 0)); when any slack moved, every bounded
    primal is set from its slack (x = lb + sl, then x = ub - su), as the
    reference's consistency loops do.  Returns the new state and the count
    of adjusted slacks."""
    x, zl, zu, sl, su = (a.copy() for a in S)
    s_min = np.finfo(float).eps * min(1.0, mu)
    if s_min == 0.0:
        s_min = np.finfo(float).tiny
    tiny = np.finfo(float).tiny
    fl, fu = np.isfinite(lbx0), np.isfinite(ubx0)
    adjusted = 0
    for s, z, b, f in ((sl, zl, lbx0, fl), (su, zu, ubx0, fu)):
        small = f & (s < s_min)
        if small.any():
            g = np.maximum(s[small], 0.0)
            target = np.maximum(mu / np.maximum(z[small], tiny), s_min)
            cap = slack_move * np.maximum(1.0, np.abs(b[small])) + g
            s[small] = np.minimum(target, cap)
            adjusted += int(small.sum())
    if adjusted:
        xp = x[P.pidx]
        xp = np.where(fl, lbx0 + sl, xp)
        xp = np.where(fu, ubx0 - su, xp)
        x[P.pidx] = xp
    return (x, zl, zu, sl, su), adjusted


# ---------------------------------------------------------------------------
# NLP scaling (include/interior_point_optimizer.h:502-584).

def nlp_scaling(P, grad, rowp, vals, max_gradient=100.0, min_value=1e-8):
    """compute_nlp_scaling (optimizer.h:502-563): the objective scale from
    the inf-norm of the primal gradient, and per constraint row j
    dc_j = min(1, max_gradient / |row_j|_inf) floored at min_value, from
    the KKT matrix assembled at lambda = 0.  Returns (obj_scale, dc, lbh
    scaled by dc, the per-variable scale vector)."""
    g = np.abs(grad[P.pidx])
    gmax = float(np.max(g, initial=0.0))
    obj = max_gradient / gmax if gmax > max_gradient else 1.0
    obj = max(obj, min_value)
    dc = np.ones(len(P.cidx))
    for j, r in enumerate(P.cidx):
        row = np.abs(vals[rowp[r]:rowp[r + 1]])
        m = float(np.max(row, initial=0.0))
        d = max_gradient / m if m > max_gradient else 1.0
        dc[j] = max(d, min_value)
    lbh = np.where(dc != 1.0, P.lbh * dc, P.lbh)
    sv = np.ones(P.n)
    sv[P.cidx] = dc
    return obj, dc, lbh, sv


def hessian_scaling(rowp, cols, vals, sv):
    """apply_hessian_scaling (optimizer.h:573-584): a_ij *= s_i s_j."""
    out = vals.copy()
    for i in range(len(rowp) - 1):
        for p in range(rowp[i], rowp[i + 1]):
            out[p] *= sv[i] * sv[cols[p]]
    return out


# ---------------------------------------------------------------------------
# Iterative refinement on the unreduced system
# (amigo/algorithm/solvers/linear_solver.py:109-264).

IR_RATIO_MAX = 1e-10
IR_IMPROVEMENT = 0.999999999
IR_MIN_STEPS = 1
IR_MAX_COND = 1e6


def ir_blocks(mu, P, S, grad, rhs_saved, Kpx, px, dzl, dzu):
    """Residuals of the 8-block system at (px, dzl, dzu), their inf-norm,
    the inf-norm of the iterate, and the condensed correction right-hand
    side (linear_solver.py:203-245).  Kpx = (K + Sigma) px with the matrix
    that was factored."""
    _, zl, zu, sl, su = S
    fl, fu = P.fl, P.fu
    gl = np.where(fl, sl, 1.0)
    gu = np.where(fu, su, 1.0)
    sigma = np.where(fl, zl / gl, 0.0) + np.where(fu, zu / gu, 0.0)
    dx = px[P.pidx]
    rhs_x = -(grad[P.pidx] - zl + zu)
    e_x = rhs_x - Kpx[P.pidx] + sigma * dx + dzl - dzu
    e_c = rhs_saved[P.cidx] - Kpx[P.cidx]
    e_zl = np.where(fl, (mu - gl * zl) - zl * dx - gl * dzl, 0.0)
    e_zu = np.where(fu, (mu - gu * zu) + zu * dx - gu * dzu, 0.0)
    nrm_resid = max(np.max(np.abs(a), initial=0.0) for a in (e_x, e_c, e_zl, e_zu))
    nrm_sol = max(np.max(np.abs(a), initial=0.0) for a in (px, dzl, dzu))
    econd = np.zeros(P.n)
    econd[P.pidx] = e_x + np.where(fl, e_zl / gl, 0.0) - np.where(fu, e_zu / gu, 0.0)
    econd[P.cidx] = e_c
    return nrm_resid, nrm_sol, econd, e_zl, e_zu


def ir_rhs_norm(mu, P, S, grad, rhs_saved):
    """Inf-norm of the 8-block right-hand side (linear_solver.py:187-198)."""
    _, zl, zu, sl, su = S
    a = np.abs(grad[P.pidx] - zl + zu)
    b = np.abs(rhs_saved[P.cidx])
    c = np.abs(mu - sl * zl)[P.fl]
    d = np.abs(mu - su * zu)[P.fu]
    return max(np.max(x, initial=0.0) for x in (a, b, c, d))


def ir_initial_dz(mu, P, S, px):
    """Bound-dual steps by back-substitution (linear_solver.py:177-179)."""
    _, zl, zu, sl, su = S
    dx = px[P.pidx]
    gl = np.where(P.fl, sl, 1.0)
    gu = np.where(P.fu, su, 1.0)
    dzl = np.where(P.fl, (mu - gl * zl - zl * dx) / gl, 0.0)
    dzu = np.where(P.fu, (mu - gu * zu + zu * dx) / gu, 0.0)
    return dzl, dzu


def ir_stop(step, ratio, ratio_old):
    """The stopping rule of linear_solver.py:226-236: converged (ratio below
    1e-10 after the minimum number of steps) or stalled (ratio above
    0.999999999 times the previous one after the minimum steps)."""
    if step >= IR_MIN_STEPS and ratio < IR_RATIO_MAX:
        return "converged"
    if step >= IR_MIN_STEPS and ratio > IR_IMPROVEMENT * ratio_old:
        return "stalled"
    return None


def iterative_refinement(mu, P, S, grad, rhs_saved, px, kkt_matvec, solve, max_steps=10):
    """The refinement loop (linear_solver.py:109-264) with the caller's
    operator y = (K + Sigma) x and solve.  Returns (px, dzl, dzu, steps,
    final ratio, reason)."""
    px = px.copy()
    _, zl, zu, sl, su = S
    dzl, dzu = ir_initial_dz(mu, P, S, px)
    nrm_rhs = ir_rhs_norm(mu, P, S, grad, rhs_saved)
    ratio_old = 1e30
    ratio = 0.0
    steps = 0
    reason = "max_steps"
    gl = np.where(P.fl, sl, 1.0)
    gu = np.where(P.fu, su, 1.0)
    for step in range(max_steps):
        nrm_resid, nrm_sol, econd, e_zl, e_zu = ir_blocks(mu, P, S, grad, rhs_saved, kkt_matvec(px), px, dzl, dzu)
        denom = min(nrm_sol, IR_MAX_COND * nrm_rhs) + nrm_rhs
        ratio = nrm_resid / denom if denom > 0 else 0.0
        why = ir_stop(step, ratio, ratio_old)
        if why is not None:
            reason = why
            break
        ratio_old = ratio
        corr = solve(econd)
        dc = corr[P.pidx]
        dzl = dzl + np.where(P.fl, (e_zl - zl * dc) / gl, 0.0)
        dzu = dzu + np.where(P.fu, (e_zu + zu * dc) / gu, 0.0)
        px = px + corr
        steps += 1
    return px, dzl, dzu, steps, ratio, reason
