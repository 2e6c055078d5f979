"""Post-optimality derivatives and problem documents on the GPU.

The oracle of a sensitivity is a finite difference of whole re-solves: the
problem's data value is moved by +-h, the IPM solves both problems to tight
tolerance, and the central difference of the solutions (and of the optimal
objective) must match the adjoint and the direct derivatives, which must
also match each other.  A problem read back from its document solves with
the same iterations to the same solution bit for bit, and a serving request
carrying a document gives that solution too.
"""

import numpy as np
import pytest

from amigo_amd.problems import components as C

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

OPTS = {"max_iterations": 200, "convergence_tolerance": 1e-11, "record_objective": False}


def _solve(prob):
    from amigo_amd.optimizer import Optimizer

    opt = Optimizer(prob, order="AMD")
    out = opt.optimize(OPTS)
    assert out["converged"]
    return opt, out


def _objective(prob, x):
    """Optimal objective on the CPU: the element Lagrangians with the
    multipliers set to zero (L = alpha f + lambda^T c)."""
    from amigo_amd.post_optimization import _derivatives_of

    f = 0.0
    for g in prob.groups:
        der = _derivatives_of(g.etype)
        z = np.where(g.idx >= 0, x[np.maximum(g.idx, 0)], 0.0)
        z[:, der.comp.multiplier_slots()] = 0.0
        L, _, _ = der.evaluate(z, g.data[:, :der.ndata] if der.ndata else g.data, alpha=1.0)
        f += float(np.sum(L))
    return f


def _fd(build, param, of, h):
    gi, e, j = param
    vals = []
    for sgn in (1.0, -1.0):
        prob = build()
        prob.groups[gi].data[e, j] += sgn * h
        prob._upload()
        opt, _ = _solve(prob)
        x = opt.get_solution()
        vals.append(np.array([_objective(prob, x) if o == "objective" else x[o] for o in of]))
    return (vals[0] - vals[1]) / (2 * h)


def _check(build, params, of, h, rtol):
    prob = build()
    opt, _ = _solve(prob)
    from amigo_amd.post_optimization import PostOptimization

    po = PostOptimization(opt)
    adj = po.compute_post_opt_derivatives(of, params, method="adjoint")
    dirr = po.compute_post_opt_derivatives(of, params, method="direct")
    scale = max(1.0, float(np.max(np.abs(adj))))
    assert np.max(np.abs(adj - dirr)) <= 1e-10 * scale
    for c, p in enumerate(params):
        fd = _fd(build, p, of, h)
        assert np.max(np.abs(adj[:, c] - fd)) <= rtol * scale, (p, adj[:, c], fd)
    return adj


def test_hanging_chain_sensitivities_match_finite_differences():
    """Node positions and the optimal energy with respect to a segment's
    length and weight (a generated element type)."""
    K = 16
    build = lambda: C.hanging_chain(K=K, seg=0.1, span=1.2)[0]
    of = [2 * (K // 2) + 1, 2 * 3, 2 * 3 + 1, "objective"]  # mid-node y, node 3 x and y, energy
    _check(build, [(0, 2, 0), (0, 5, 1), (0, K // 2, 0)], of, h=1e-5, rtol=2e-5)


def test_elliptic_tracking_target_sensitivities_match_finite_differences():
    """State, control and optimal objective of a small 2-D elliptic control
    problem with respect to target values y_d at two nodes (built-in element
    type; active control bounds stay active under the perturbation)."""
    from amigo_amd.problems import elliptic

    build = lambda: elliptic.EllipticControl(m=6, dim=2, seed=2)
    prob = build()
    g0 = prob.groups[0]
    mid = g0.E // 2
    y_mid = int(g0.idx[mid, 0])
    u_mid = int(g0.idx[mid, g0.idx.shape[1] - 2])
    _check(build, [(0, mid, 1), (0, 3, 1)], [y_mid, u_mid, "objective"], h=1e-4, rtol=1e-5)


def test_problem_document_solves_like_the_original():
    from amigo_amd import wire
    from amigo_amd.serve import SolveServer

    prob, _ = C.hanging_chain(K=20, seg=0.1, span=1.5)
    opt0, out0 = _solve(prob)
    text = wire.dumps(prob)
    opt1, out1 = _solve(wire.loads(text))
    assert out0["num_iterations"] == out1["num_iterations"]
    assert np.array_equal(opt0.get_solution(), opt1.get_solution())
    server = SolveServer(slots=2)
    try:
        res = server.submit({"problem": text, "order": "AMD", "scaling": False,
                             "options": OPTS}).result(timeout=120)
    finally:
        server.close()
    assert res["converged"]
    assert np.array_equal(res["x"], opt0.get_solution())
