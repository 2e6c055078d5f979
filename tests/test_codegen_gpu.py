"""Generated element kernels on the GPU.

- The kernels generated from the workload components equal the hand-written
  kernels of the same element types (gradient, Hessian, Lagrangian partial
  sums) to rounding: the derivatives are the same mathematics in another
  operation order.
- The kernels of the two example components equal the numpy evaluation of
  the same symbolic derivatives (which tests/test_codegen.py pins against
  complex-step derivatives), and their gradient-only variant gives the
  gradient of the Hessian variant bit for bit.
- Whole solves on generated kernels take the iterations of the hand-written
  path (same KKT pattern, same inertia sequence, same optimum to 1e-9).
- A hanging chain, a component with no hand-written kernel, solves to a
  point that satisfies the KKT conditions checked on the CPU.
"""

import ctypes

import numpy as np
import pytest

from amigo_amd import codegen
from amigo_amd.problems import components as C

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _eval(etype, idx, data, x, alpha, want_h):
    from amigo_amd import _lib
    from amigo_amd.sparse import current_stream_ptr

    E, nc = idx.shape
    idx_d = torch.tensor(idx, dtype=torch.int32, device="cuda")
    data_d = torch.tensor(data, dtype=torch.float64, device="cuda")
    x_d = torch.tensor(x, dtype=torch.float64, device="cuda")
    g = torch.full((E * nc,), np.nan, dtype=torch.float64, device="cuda")
    H = torch.full((E * nc * nc,), np.nan, dtype=torch.float64, device="cuda") if want_h else None
    lp = torch.zeros((E + 255) // 256, dtype=torch.float64, device="cuda")
    rc = _lib.lib().amigo_elements_eval(
        etype, E, _lib.dptr(idx_d), _lib.dptr(data_d), data.shape[1], _lib.dptr(x_d), alpha,
        _lib.dptr(g), _lib.dptr(H) if want_h else None, _lib.dptr(lp), _lib.vp(current_stream_ptr()))
    _lib.check(rc, "asm")
    torch.cuda.synchronize()
    return (g.cpu().numpy().reshape(E, nc), H.cpu().numpy().reshape(E, nc, nc) if want_h else None,
            lp.cpu().numpy())


def _inputs(nc, nd, E, seed, n=5000):
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, n, size=(E, nc)).astype(np.int32)
    idx[rng.random((E, nc)) < 0.05] = -1  # absent inputs
    idx[:, 0] = np.abs(idx[:, 0])
    data = np.abs(rng.normal(size=(E, max(nd, 1)))) + 0.2
    x = rng.normal(size=n)
    return idx, data, x


def _close(a, b, rtol):
    scale = max(1.0, float(np.max(np.abs(b))))
    return float(np.max(np.abs(a - b))) <= rtol * scale


@pytest.mark.parametrize("etype", sorted(C.workload_components()))
def test_generated_kernel_matches_the_handwritten_kernel(etype):
    comp = C.workload_components()[etype]
    ge = codegen.element_type(comp, build=False)
    idx, data, x = _inputs(ge.nc, ge.ndata, 777, seed=etype)
    gh, Hh, lh = _eval(etype, idx, data, x, 0.37, True)
    gg, Hg, lg = _eval(ge.etype, idx, data, x, 0.37, True)
    assert _close(gg, gh, 1e-13) and _close(Hg, Hh, 1e-13) and _close(lg, lh, 1e-13)
    # structural zeros of the generated Hessian are exact zeros of the hand-written one
    assert np.all(Hh[:, ~ge.derivatives.structure] == 0.0)
    g2, _, l2 = _eval(ge.etype, idx, data, x, 0.37, False)
    assert np.array_equal(g2, gg) and np.array_equal(l2, lg)


@pytest.mark.parametrize("name", ["ChainSegment", "TranscendentalElement"])
def test_generated_kernel_matches_numpy_derivatives(name):
    comp = {"ChainSegment": C.ChainSegment, "TranscendentalElement": C.TranscendentalElement}[name]()
    ge = codegen.element_type(comp, build=False)
    idx, data, x = _inputs(ge.nc, ge.ndata, 300, seed=11)
    idx[idx < 0] = 0
    if name == "TranscendentalElement":
        x = 0.3 + 0.6 * np.abs(x)  # a ** -2: away from 0
    g, H, lp = _eval(ge.etype, idx, data, x, 0.8, True)
    z = x[idx]
    L, go, Ho = ge.derivatives.evaluate(z, data[:, :ge.ndata] if ge.ndata else data, alpha=0.8)
    assert _close(g, go, 1e-12) and _close(H, Ho, 1e-12)
    nb = (idx.shape[0] + 255) // 256
    lo = np.array([L[b * 256:(b + 1) * 256].sum() for b in range(nb)])
    assert _close(lp, lo, 1e-12)
    g2, _, _ = _eval(ge.etype, idx, data, x, 0.8, False)
    assert np.array_equal(g2, g)


@pytest.mark.parametrize("name", ["double_integrator", "cart_pole"])
def test_solve_on_generated_kernels_takes_the_handwritten_iterations(name):
    from amigo_amd.optimizer import Optimizer
    from amigo_amd.problems import cart_pole, double_integrator

    build = {"double_integrator": lambda: double_integrator.build(N=64),
             "cart_pole": lambda: cart_pole.build(N=40, mc=1.1, mp=0.25, l=0.6)}[name]
    outs, xs, inert = [], [], []
    for gen in (False, True):
        prob = build()
        if gen:
            prob = C.generated_problem(prob)
            assert all(g.etype >= codegen.GENERATED_BASE for g in prob.groups)
        opt = Optimizer(prob, order="ND" if name == "cart_pole" else "AMD")
        out = opt.optimize({"max_iterations": 100})
        outs.append(out)
        xs.append(opt.get_solution())
        inert.append(out["inertia_history"])
    assert outs[0]["converged"] and outs[1]["converged"]
    assert outs[0]["num_iterations"] == outs[1]["num_iterations"]
    assert inert[0] == inert[1]
    assert np.max(np.abs(xs[0] - xs[1])) <= 1e-9 * max(1.0, np.max(np.abs(xs[0])))


def test_hanging_chain_solves_to_a_kkt_point():
    from amigo_amd.optimizer import Optimizer

    K, seg, span = 40, 0.05, 1.5
    prob, ge = C.hanging_chain(K=K, seg=seg, span=span)
    opt = Optimizer(prob, order="AMD")
    out = opt.optimize({"max_iterations": 200})
    assert out["converged"], out["iterations"][-1]
    sol = opt.get_solution()
    nn = K + 1
    xy = sol[:2 * nn].reshape(nn, 2)
    lam = sol[2 * nn:2 * nn + K]
    lam_end = sol[2 * nn + K:]
    # feasibility: every segment has its length, the ends are held
    d = np.diff(xy, axis=0)
    assert np.max(np.abs(np.hypot(d[:, 0], d[:, 1]) - seg)) < 1e-8
    assert np.allclose(xy[[0, -1]].ravel(), [0.0, 0.0, span, 0.0], atol=1e-9)
    # stationarity of the Lagrangian (CPU, numpy derivatives of the component)
    idx = np.stack([np.arange(K) * 2, np.arange(K) * 2 + 1, np.arange(K) * 2 + 2,
                    np.arange(K) * 2 + 3], axis=1)
    z = np.concatenate([sol[idx], lam[:, None]], axis=1)
    _, g, _ = ge.derivatives.evaluate(z, np.tile([seg, seg], (K, 1)), alpha=1.0)
    grad = np.zeros(2 * nn)
    np.add.at(grad, idx.ravel(), g[:, :4].ravel())
    grad[[0, 1, 2 * K, 2 * K + 1]] += lam_end
    assert np.max(np.abs(grad)) < 1e-6 * max(1.0, np.max(np.abs(lam)))
    # the chain hangs symmetrically below its ends
    assert np.allclose(xy[:, 1], xy[::-1, 1], atol=1e-7)
    assert np.all(xy[1:-1, 1] < 0.0)
