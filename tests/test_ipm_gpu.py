"""GPU interior-point vector kernels (csrc/ipm_kernels.hip through the
device InteriorPointOptimizer) against the oracle restatement
(oracle/ipm.py) on seeded random problems: element-wise results to FP64
rounding, reductions to 1e-13 relative, argmin indices exactly."""

import ctypes

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


def test_initialize_relaxes_and_projects_like_the_oracle():
    rng, P, S, opt, v = _setup(6)
    x0 = rng.normal(scale=4.0, size=P.n)
    v.x.copy_(_d(x0))
    opt.relax_bounds(1e-8, 1e-4)
    lb, ub = O.relax_bounds(P, 1e-8, 1e-4)
    assert np.array_equal(_h(opt.lbx), lb) and np.array_equal(_h(opt.ubx), ub)
    opt.initialize_multipliers_and_slacks(1.0, None, v)
    Pr = O.Info(~np.isin(np.arange(P.n), P.pidx), np.zeros(P.n), np.zeros(P.n))
    Pr.lbx, Pr.ubx, Pr.lbh = lb, ub, P.lbh
    exp = O.project_and_init(Pr, x0)
    # to FP64 rounding: the device contracts a*b+c into one fused operation
    for a, b in zip((v.x, v.zl, v.zu, v.sl, v.su), exp):
        assert np.allclose(_h(a), b, rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("seed", [7, 8])
def test_remaining_reductions(seed):
    rng, P, S, opt, v = _setup(seed)
    mu = 0.13
    grad = rng.normal(size=P.n)
    px = rng.normal(size=P.n)
    assert opt.compute_kkt_error(v, _d(grad)) == pytest.approx(O.kkt_error_sq(P, S, grad), rel=1e-13)
    assert np.isclose(opt.compute_max_comp_deviation(v, mu), O.max_comp_deviation(mu, P, S), rtol=1e-15)
    assert np.isclose(opt.compute_complementarity_sq(v), O.complementarity_sq(0.0, P, S), rtol=1e-13)
    assert np.isclose(opt.compute_barrier_dphi_direct(mu, v, _d(grad), _d(px)),
                      O.barrier_dphi(mu, P, S, grad, px), rtol=1e-12)
    got = opt.abs_sums(v)
    exp = O.abs_sums(P, S)
    assert np.allclose(got, exp, rtol=1e-13, atol=0)


def test_affine_start_point_and_dual_residual():
    rng, P, S, opt, v = _setup(9)
    upd = opt.create_opt_vector()
    dzl = rng.normal(size=len(P.pidx))
    dzu = rng.normal(size=len(P.pidx))
    upd.zl.copy_(_d(dzl))
    upd.zu.copy_(_d(dzu))
    opt.compute_affine_start_point(0.4, v, upd, v)
    zl, zu = O.affine_start_point(0.4, P, S, dzl, dzu)
    assert np.array_equal(_h(v.zl), zl) and np.array_equal(_h(v.zu), zu)
    grad = rng.normal(size=P.n)
    out = torch.full((P.n,), 7.0, dtype=torch.float64, device="cuda")
    opt.compute_dual_residual_vector(v, _d(grad), out)
    S2 = (S[0], zl, zu, S[3], S[4])
    assert np.array_equal(_h(out), O.dual_residual_vector(P, S2, grad))


def test_floor_slacks_matches_the_oracle():
    rng, P, S, opt, v = _setup(10)
    x, zl, zu, sl, su = (a.copy() for a in S)
    il = np.nonzero(P.fl)[0]
    iu = np.nonzero(P.fu)[0]
    sl[il[:3]] = [1e-30, -1e-20, 0.0]
    su[iu[:2]] = [1e-25, 0.0]
    for name, a in zip(("x", "zl", "zu", "sl", "su"), (x, zl, zu, sl, su)):
        getattr(v, name).copy_(_d(a))
    mu = 1e-3
    s_min = np.finfo(float).eps * min(1.0, mu)
    cnt = opt.floor_slacks(v, mu, s_min, 1.81898940354586e-12)
    exp, cnt_o = O.floor_slacks(mu, P, (x, zl, zu, sl, su), P.lbx, P.ubx)
    assert cnt == cnt_o == 5
    for a, b in zip((v.x, v.zl, v.zu, v.sl, v.su), exp):
        assert np.allclose(_h(a), b, rtol=1e-15, atol=1e-300)


def test_nlp_scaling_kernels_match_the_oracle():
    import scipy.sparse as sp

    from amigo_amd import _lib
    from amigo_amd.ipm import _s

    rng, P, S, opt, v = _setup(11, n_p=600, n_c=250)
    n = P.n
    A = sp.random(n, n, density=0.01, random_state=3, format="csr") * 50.0
    A = A + A.T + sp.eye(n)
    A = A.tocsr()
    A.sort_indices()
    rows = P.cidx[:20]
    D = sp.diags(np.where(np.isin(np.arange(n), rows), 1e4, 1.0))
    A = (D @ A @ D).tocsr()
    A.sort_indices()
    grad = rng.normal(size=n) * 30
    grad[P.pidx[0]] = -700.0
    obj, dc, lbh, sv = O.nlp_scaling(P, grad, A.indptr, A.data)
    rowp = torch.tensor(A.indptr.astype(np.int32), device="cuda")
    cols = torch.tensor(A.indices.astype(np.int32), device="cuda")
    vals = _d(A.data)
    dcd = torch.ones(len(P.cidx), dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().amigo_ipm_constraint_scaling(
        len(P.cidx), _lib.dptr(opt.cidx), _lib.dptr(rowp), _lib.dptr(vals), 100.0, 1e-8,
        _lib.dptr(dcd), _s()), "ipm")
    assert np.array_equal(_h(dcd), dc) and np.any(dc < 1.0)
    svd = _d(sv)
    _lib.check(_lib.lib().amigo_csr_scale_symmetric(
        n, _lib.dptr(rowp), _lib.dptr(cols), _lib.dptr(svd), _lib.dptr(vals), _s()), "ipm")
    assert np.array_equal(_h(vals), O.hessian_scaling(A.indptr, A.indices, A.data, sv))
    g = _d(grad)
    _lib.check(_lib.lib().amigo_ipm_scale_rows(len(P.cidx), _lib.dptr(opt.cidx), _lib.dptr(dcd),
                                               _lib.dptr(g), 0, _s()), "ipm")
    exp = grad.copy()
    exp[P.cidx] *= dc
    assert np.array_equal(_h(g), exp)


def test_refinement_kernels_match_the_oracle():
    from amigo_amd import _lib
    from amigo_amd.ipm import _s

    rng, P, S, opt, v = _setup(12)
    mu = 0.07
    grad = rng.normal(size=P.n)
    rhs = rng.normal(size=P.n)
    px = rng.normal(size=P.n)
    Kpx = rng.normal(size=P.n)
    L = _lib.lib()
    # device copies held for the whole test (a temporary's memory could be
    # reused by the next one before the kernels read it)
    grad_d, rhs_d, px_d, Kpx_d = _d(grad), _d(rhs), _d(px), _d(Kpx)
    npr = len(P.pidx)
    dzl = torch.zeros(npr, dtype=torch.float64, device="cuda")
    dzu = torch.zeros_like(dzl)
    out = (ctypes.c_double * 4)()
    _lib.check(L.amigo_ipm_ir_init(opt.info, v.state(), mu, _lib.dptr(grad_d), _lib.dptr(rhs_d),
                                   _lib.dptr(px_d), _lib.dptr(dzl), _lib.dptr(dzu), _lib.dptr(opt.ws),
                                   out, _s()), "ipm")
    dzl_o, dzu_o = O.ir_initial_dz(mu, P, S, px)
    # fused multiply-adds on the device; cancellation near zero
    assert np.allclose(_h(dzl), dzl_o, rtol=1e-13, atol=1e-14) and np.allclose(_h(dzu), dzu_o, rtol=1e-13, atol=1e-14)
    assert np.isclose(out[0], O.ir_rhs_norm(mu, P, S, grad, rhs), rtol=1e-15)
    econd = torch.zeros(P.n, dtype=torch.float64, device="cuda")
    ezl = torch.zeros_like(dzl)
    ezu = torch.zeros_like(dzl)
    _lib.check(L.amigo_ipm_ir_residual(opt.info, v.state(), mu, _lib.dptr(grad_d), _lib.dptr(rhs_d),
                                       _lib.dptr(Kpx_d), _lib.dptr(px_d), _lib.dptr(dzl), _lib.dptr(dzu),
                                       _lib.dptr(econd), _lib.dptr(ezl), _lib.dptr(ezu), _lib.dptr(opt.ws),
                                       out, _s()), "ipm")
    nr, ns, ec, ezl_o, ezu_o = O.ir_blocks(mu, P, S, grad, rhs, Kpx, px, dzl_o, dzu_o)
    assert np.isclose(out[0], nr, rtol=1e-13) and np.isclose(out[1], ns, rtol=1e-13)
    assert np.allclose(_h(econd), ec, rtol=1e-13, atol=1e-14)
    assert np.allclose(_h(ezl), ezl_o, rtol=1e-13, atol=1e-14)
    corr = rng.normal(size=P.n)
    corr_d = _d(corr)
    pxd = _d(px)
    _lib.check(L.amigo_ipm_ir_update(opt.info, v.state(), P.n, _lib.dptr(corr_d), _lib.dptr(ezl),
                                     _lib.dptr(ezu), _lib.dptr(pxd), _lib.dptr(dzl), _lib.dptr(dzu), _s()),
               "ipm")
    _, zl, zu, sl, su = S
    dc = corr[P.pidx]
    exp_l = dzl_o + np.where(P.fl, (_h(ezl) - zl * dc) / np.where(P.fl, sl, 1.0), 0.0)
    exp_u = dzu_o + np.where(P.fu, (_h(ezu) + zu * dc) / np.where(P.fu, su, 1.0), 0.0)
    assert np.allclose(_h(pxd), px + corr, rtol=1e-15, atol=0)
    assert np.allclose(_h(dzl), exp_l, rtol=1e-13, atol=1e-14)
    assert np.allclose(_h(dzu), exp_u, rtol=1e-13, atol=1e-14)
