"""Mixed-precision form: FP32 MFMA Schur updates of the blocked fronts,
FP64 pivot decisions, inertia, 2x2 blocks, solves and refinement.

- The inertia of the mixed factor equals the FP64 factor's.
- One mixed solve is accurate to FP32 level (1e-2 relative residual on
  these 1/h^2-scaled grid KKTs); FP64 iterative refinement
  on the FP64 matrix recovers 1e-12 relative residual.
- An interior-point solve in mixed precision reaches the same optimum
  (1e-6 relative), with the same inertia at every factorization and the
  iteration count within one of the FP64 run."""

import numpy as np
import pytest

from tests import matrices as M

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _factor(A, precision):
    from amigo_amd.sparse import CSRMat, SparseLDL

    ldl = SparseLDL(CSRMat.from_scipy(A), order="ND", precision=precision)
    ldl.factor()
    return ldl


@pytest.mark.parametrize("side", [40, 56])
def test_mixed_factor_inertia_and_refined_solve(side):
    A = M.grid_kkt(side, seed=5)[0]
    l64 = _factor(A, "fp64")
    l32 = _factor(A, "mixed")
    assert l32.stats()[1] > 128  # the blocked path (and its Schur updates) ran
    assert l32.get_inertia() == l64.get_inertia()
    rng = np.random.default_rng(0)
    xs = rng.normal(size=A.shape[0])
    b = A @ xs
    x = torch.tensor(b, device="cuda")
    l32.solve(x)
    r0 = np.linalg.norm(A @ x.cpu().numpy() - b) / np.linalg.norm(b)
    assert r0 < 1e-2  # FP32 Schur updates on an ill-conditioned (1/h^2) KKT
    xh = x.cpu().numpy()
    for _ in range(10):
        d = torch.tensor(b - A @ xh, device="cuda")
        l32.solve(d)
        xh = xh + d.cpu().numpy()
    r = np.linalg.norm(A @ xh - b) / np.linalg.norm(b)
    assert r < 1e-12, (r0, r)


def test_mixed_interior_point_matches_fp64():
    from amigo_amd.optimizer import Optimizer
    from amigo_amd.problems import elliptic

    prob = elliptic.EllipticControl(m=48, dim=2, seed=1)
    o64 = Optimizer(prob, order="ND", scaling=True, precision="fp64")
    r64 = o64.optimize({"max_iterations": 100})
    x64 = o64.get_solution()
    o32 = Optimizer(prob, order="ND", scaling=True, precision="mixed")
    r32 = o32.optimize({"max_iterations": 100})
    x32 = o32.get_solution()
    assert r64["converged"] and r32["converged"]
    assert abs(r64["num_iterations"] - r32["num_iterations"]) <= 1
    assert r32["iterations"][-1]["nlp_error"] <= 1e-8
    # north_star's bar: the same optimum within 1e-8 relative.
    diff = np.max(np.abs(x32 - x64)) / max(1.0, np.max(np.abs(x64)))
    print(f"mixed vs FP64 solution: {diff:.2e} relative")
    assert diff <= 1e-8
    n = min(len(o64.solver.inertia_history), len(o32.solver.inertia_history))
    assert o64.solver.inertia_history[:n] == o32.solver.inertia_history[:n]
    print(f"mixed: {r32['num_iterations']} iterations, {o32.solver.fp64_fallbacks} FP64 fallbacks, "
          f"{o32.solver.ir_steps_total} IR steps (FP64 run: {o64.solver.ir_steps_total})")
