"""The native driver against the recorded decision sequences of the
driver it replaced (tests/golden/driver_regression.npz, written by
tests/golden/make_driver_fixtures.py): on the double integrator N = 256,
2-D elliptic control 32 x 32 and eight seeded cart-pole swing-ups, the
same iteration count, the same inertia at every factorization, the same
line-search trial counts, barrier parameters and step lengths, and the
same solution to 1e-10 relative (the refinement product sums the shift in
one rounding instead of two, so iterates may differ at rounding level)."""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))


def _fixtures():
    import sys

    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_driver_fixtures as F

    data = np.load(os.path.join(HERE, "golden", "driver_regression.npz"))
    return F, data


def _cases():
    F, _ = _fixtures()
    return [c[0] for c in F.cases()]


@pytest.mark.parametrize("name", _cases())
def test_decisions_match_the_recorded_sequence(name):
    F, data = _fixtures()
    case = {c[0]: c for c in F.cases()}[name]
    _, make, kw, keep_x = case
    rec = F.record(make, kw, keep_x)
    g = lambda k: data[f"{name}/{k}"]
    assert bool(rec["converged"]) == bool(g("converged"))
    assert int(rec["num_iterations"]) == int(g("num_iterations"))
    assert np.array_equal(rec["inertia"], g("inertia"))
    assert np.array_equal(rec["line_iters"], g("line_iters"))
    for k in ("mu", "alpha_x", "alpha_z", "delta_w"):
        assert np.allclose(rec[k], g(k), rtol=1e-9, atol=0), (k, rec[k], g(k))
    if keep_x:
        x, x0 = rec["x"], g("x")
        assert np.max(np.abs(x - x0)) <= 1e-10 * max(1.0, np.max(np.abs(x0)))


def test_restoration_inside_a_batch_equals_alone():
    """A problem that enters feasibility restoration does so inside the
    batch, and its iterates equal its solve alone bit for bit."""
    from amigo_amd.batch import BatchOptimizer
    from amigo_amd.optimizer import Optimizer
    from amigo_amd.problems import cart_pole

    probs = [cart_pole.seeded(s, nmin=16, nmax=120) for s in range(24)]
    bo = BatchOptimizer(probs, order="ND", scaling=True)
    res = bo.optimize({"record_objective": False})
    xs = bo.get_solution()
    restored = [k for k, r in enumerate(res) if r["restorations"] > 0]
    check = (restored + [0, 1])[:4]
    for k in check:
        o = Optimizer(probs[k], order="ND", scaling=True)
        d = o.optimize({"record_objective": False})
        assert d["num_iterations"] == res[k]["num_iterations"], k
        assert d["inertia_history"] == res[k]["inertia_history"], k
        assert np.array_equal(o.get_solution(), xs[k]), k
    print(f"restored in the batch: {restored}")


def test_last_problems_of_a_batch_equal_alone():
    """The last problems still iterating in a batch run on launches that
    list only their blocks and rows (a quarter of the batch or fewer);
    their iterates equal their solves alone bit for bit, and so do those of
    a problem that finished early."""
    from amigo_amd.batch import BatchOptimizer
    from amigo_amd.optimizer import Optimizer
    from amigo_amd.problems import cart_pole

    probs = [cart_pole.seeded(100 + s, nmin=16, nmax=160) for s in range(32)]
    bo = BatchOptimizer(probs, order="ND", scaling=True)
    res = bo.optimize({"record_objective": False})
    xs = bo.get_solution()
    its = np.array([r["num_iterations"] for r in res])
    order = np.argsort(-its, kind="stable")
    # the slowest problems run their last iterations with <= 8 of 32 active
    assert its[order[0]] > np.sort(its)[-9], its
    for k in list(order[:3]) + [int(order[-1])]:
        o = Optimizer(probs[k], order="ND", scaling=True)
        d = o.optimize({"record_objective": False})
        assert d["num_iterations"] == res[k]["num_iterations"], k
        assert d["inertia_history"] == res[k]["inertia_history"], k
        assert np.array_equal(o.get_solution(), xs[k]), k
