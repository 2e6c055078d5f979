"""GPU gradient / Hessian / Lagrangian assembly of the component groups
against the oracle element Lagrangians (complex-step derivatives) on the
CPU, at random points, for every element type the workloads use."""

import numpy as np
import pytest

from tests import assembly_oracle as AO

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _problems():
    from amigo_amd.problems import cart_pole, double_integrator, elliptic

    return {
        "double_integrator": lambda: double_integrator.build(N=8),
        "elliptic2d": lambda: elliptic.EllipticControl(m=5, dim=2, seed=3),
        "elliptic3d": lambda: elliptic.EllipticControl(m=3, dim=3, seed=4),
        "cart_pole": lambda: cart_pole.build(N=6, mc=1.3, mp=0.2, l=0.7),
    }


@pytest.mark.parametrize("name", ["double_integrator", "elliptic2d", "elliptic3d", "cart_pole"])
@pytest.mark.parametrize("alpha", [1.0, 0.37])
def test_assembly_matches_oracle(name, alpha):
    prob = _problems()[name]()
    rng = np.random.default_rng(hash(name) % 1000)
    x = rng.normal(size=prob.n)
    xd = torch.tensor(x, device="cuda")
    g = torch.zeros(prob.n, dtype=torch.float64, device="cuda")
    prob.gradient(alpha, xd, g)
    mat = prob.create_matrix()
    prob.hessian(alpha, xd, mat)
    L = prob.lagrangian(alpha, xd)
    torch.cuda.synchronize()
    Lo, go, Ho = AO.assemble(prob, alpha, x)
    # gradient: complex-step oracle is exact to rounding
    assert np.allclose(g.cpu().numpy(), go, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(go).max()))
    assert abs(L - Lo) <= 1e-12 * max(1.0, abs(Lo))
    # Hessian: central difference of exact gradients, O(1e-8) truncation
    Hd = mat.to_scipy().toarray()
    scale = max(1.0, np.abs(Ho).max())
    assert np.allclose(Hd, Ho, rtol=1e-6, atol=1e-7 * scale)
    # nothing of the oracle Hessian falls outside the pattern
    P = mat.to_scipy().copy()
    P.data[:] = 1.0
    outside = np.abs(Ho) * (1.0 - P.toarray())
    assert outside.max() <= 1e-7 * scale
    # symmetric pattern and values
    assert np.allclose(Hd, Hd.T, atol=0.0)


def test_assembly_is_deterministic():
    from amigo_amd.problems import elliptic

    prob = elliptic.EllipticControl(m=40, dim=2, seed=5)
    x = torch.tensor(np.random.default_rng(0).normal(size=prob.n), device="cuda")
    m1, m2 = prob.create_matrix(), prob.create_matrix()
    prob.hessian(1.0, x, m1)
    prob.hessian(1.0, x, m2)
    g1 = torch.zeros(prob.n, dtype=torch.float64, device="cuda")
    g2 = torch.zeros_like(g1)
    prob.gradient(1.0, x, g1)
    prob.gradient(1.0, x, g2)
    assert torch.equal(m1.data, m2.data) and torch.equal(g1, g2)


@pytest.mark.parametrize("name", ["double_integrator", "elliptic2d", "elliptic3d", "cart_pole"])
def test_gradient_only_evaluation_equals_the_hessian_evaluations_gradient(name):
    """The element kernels' gradient-only variant (no Hessian wanted) gives
    bitwise the gradient entries of the variant that also forms the
    Hessian."""
    from amigo_amd import _lib
    from amigo_amd.problem import current_stream_ptr

    prob = _problems()[name]()
    x = torch.tensor(np.random.default_rng(7).normal(size=prob.n), device="cuda")
    s = _lib.vp(current_stream_ptr())
    for grp in prob.groups:
        g1 = torch.zeros(grp.E * 16, dtype=torch.float64, device="cuda")
        g2 = torch.zeros_like(g1)
        hb = torch.zeros(grp.E * 16 * 16, dtype=torch.float64, device="cuda")
        for gb, h in ((g1, _lib.vp(0)), (g2, _lib.dptr(hb))):
            _lib.check(_lib.lib().amigo_elements_eval(
                grp.etype, grp.E, _lib.dptr(grp.idx_d), _lib.dptr(grp.data_d), grp.nd,
                _lib.dptr(x), 0.37, _lib.dptr(gb), h, _lib.vp(0), s), "asm")
        torch.cuda.synchronize()
        assert torch.equal(g1, g2), grp.etype
