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
