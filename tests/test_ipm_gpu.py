"""GPU interior-point vector kernels (csrc/ipm_kernels.hip through the
device InteriorPointOptimizer) against the oracle restatement
(oracle/ipm.py) on seeded random problems: element-wise results to FP64
rounding, reductions to 1e-13 relative, argmin indices exactly."""

import numpy as np
import pytest

from oracle import ipm as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


class _Prob:
    def __init__(self, mult):
        self.mult = mult
        self.n = len(mult)
        self.device = torch.device("cuda")

    def get_multiplier_indicator(self):
        return self.mult


def _setup(seed, n_p=3000, n_c=1200):
    from amigo_amd.ipm import InteriorPointOptimizer

    rng = np.random.default_rng(seed)
    n = n_p + n_c
    mult = np.zeros(n, bool)
    mult[rng.choice(n, n_c, replace=False)] = True
    lower = np.full(n, -np.inf)
    upper = np.full(n, np.inf)
    p = np.nonzero(~mult)[0]
    hl = rng.random(n_p) < 0.6
    hu = rng.random(n_p) < 0.6
    lower[p[hl]] = rng.uniform(-2, 0, hl.sum())
    upper[p[hu]] = rng.uniform(1, 3, hu.sum())
    c = np.nonzero(mult)[0]
    lower[c] = upper[c] = rng.normal(size=n_c)
    P = O.Info(mult, lower, upper)
    opt = InteriorPointOptimizer(_Prob(mult), lower, upper)
    x = rng.normal(size=n)
    xp = x[P.pidx]
    xp = np.where(P.fl & P.fu, P.lbx + rng.uniform(0.1, 0.9, n_p) * (P.ubx - P.lbx), xp)
    xp = np.where(P.fl & ~P.fu, P.lbx + rng.uniform(0.01, 2, n_p), xp)
    xp = np.where(~P.fl & P.fu, P.ubx - rng.uniform(0.01, 2, n_p), xp)
    x[P.pidx] = xp
    sl = np.where(P.fl, xp - P.lbx, 1.0)
    su = np.where(P.fu, P.ubx - xp, 1.0)
    zl = np.where(P.fl, rng.uniform(0.01, 3, n_p), 0.0)
    zu = np.where(P.fu, rng.uniform(0.01, 3, n_p), 0.0)
    S = (x, zl, zu, sl, su)
    v = opt.create_opt_vector(torch.tensor(x, device="cuda"))
    for name, a in zip(("zl", "zu", "sl", "su"), S[1:]):
        getattr(v, name).copy_(torch.tensor(a))
    return rng, P, S, opt, v


def _d(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _h(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("seed", [0, 1])
def test_residual_diagonal_and_update(seed):
    rng, P, S, opt, v = _setup(seed)
    mu = 0.037
    grad = rng.normal(size=P.n)
    res = torch.zeros(P.n, dtype=torch.float64, device="cuda")
    nrm = opt.compute_residual(mu, v, _d(grad), res)
    r_o, n2, d2, p2 = O.residual(mu, P, S, grad)
    assert np.allclose(_h(res), r_o, rtol=1e-14, atol=1e-14)
    assert np.isclose(nrm, np.sqrt(n2), rtol=1e-13)
    d_sq, p_sq = opt.compute_residual_and_infeasibility(mu, v, _d(grad), res)
    assert np.isclose(d_sq, d2, rtol=1e-13) and np.isclose(p_sq, p2, rtol=1e-13)
    diag = torch.zeros(P.n, dtype=torch.float64, device="cuda")
    opt.compute_diagonal(v, diag)
    assert np.allclose(_h(diag), O.diagonal(P, S), rtol=1e-15, atol=0)
    px = rng.normal(size=P.n)
    upd = opt.create_opt_vector()
    opt.compute_update(mu, v, _d(px), upd)
    dzl, dzu = O.bound_dual_step(mu, P, S, px)
    assert np.allclose(_h(upd.zl), dzl, rtol=1e-14, atol=1e-14)
    assert np.allclose(_h(upd.zu), dzu, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("seed", [2, 3])
def test_max_step_and_apply_step(seed):
    rng, P, S, opt, v = _setup(seed)
    px = 2.0 * rng.normal(size=P.n)
    upd = opt.create_opt_vector(_d(px))
    dzl = rng.normal(size=len(P.pidx))
    dzu = rng.normal(size=len(P.pidx))
    upd.zl.copy_(_d(dzl))
    upd.zu.copy_(_d(dzu))
    ax, xi, az, zi = opt.compute_max_step(0.995, v, upd)
    ax_o, xi_o, az_o, zi_o = O.max_step(0.995, P, S, px, dzl, dzu)
    assert ax == ax_o and az == az_o and xi == xi_o and zi == zi_o
    tmp = opt.create_opt_vector()
    opt.apply_step_update(ax, az, v, upd, tmp)
    got = [_h(t) for t in (tmp.x, tmp.zl, tmp.zu, tmp.sl, tmp.su)]
    exp = O.apply_step(ax, az, P, S, px, dzl, dzu)
    for a, b in zip(got, exp):
        assert np.allclose(a, b, rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("seed", [4])
def test_reductions(seed):
    rng, P, S, opt, v = _setup(seed)
    mu = 0.21
    grad = rng.normal(size=P.n)
    avg, xi = opt.compute_complementarity(v)
    avg_o, xi_o = O.complementarity(P, S)
    assert np.isclose(avg, avg_o, rtol=1e-13) and np.isclose(xi, xi_o, rtol=1e-13)
    d, p, c = opt.compute_kkt_error_mu(mu, v, _d(grad))
    d_o, p_o, c_o = O.kkt_error(mu, P, S, grad)
    assert (d, p, c) == pytest.approx((d_o, p_o, c_o), rel=1e-15)
    assert np.isclose(opt.compute_barrier_log_sum(mu, v), O.barrier_log_sum(mu, P, S), rtol=1e-13)
    assert np.isclose(opt.compute_constraint_violation_1norm(v, _d(grad)), O.constraint_violation(P, grad), rtol=1e-13)
    res, px, diag = rng.normal(size=P.n), rng.normal(size=P.n), rng.normal(size=P.n)
    dphi = opt.compute_barrier_dphi(mu, v, None, _d(res), _d(px), _d(diag))
    assert np.isclose(dphi, O.barrier_dphi_from_kkt(P, S, res, px, diag), rtol=1e-12)


def test_reset_bound_multipliers():
    rng, P, S, opt, v = _setup(5)
    mu, kappa = 0.05, 10.0
    opt.reset_bound_multipliers(mu, kappa, v)
    zl, zu = O.reset_bound_multipliers(mu, kappa, P, S)
    assert np.allclose(_h(v.zl), zl, rtol=1e-15, atol=0) and np.allclose(_h(v.zu), zu, rtol=1e-15, atol=0)
