"""HIP graph replay of the factorization and solve launch sequences.

Two handles on the same matrix, one replaying captured graphs (the default)
and one launching eagerly, factor a sequence of different values in the same
buffers and solve into the same right-hand-side buffer, with unrelated
device work and host synchronizations between the replays.  Every replay
must give bit-identical inertia, pivot / delay counts and solutions."""

import numpy as np
import pytest

from tests import matrices as M

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("case", ["grid_nd", "delays_amd"])
def test_graph_replays_match_eager(case):
    from amigo_amd.sparse import CSRMat, SparseLDL

    if case == "grid_nd":
        A = M.grid_kkt(40, seed=7)[0]
        kw = dict(order="ND")
    else:
        A = M.kkt(1500, 700, 4)
        kw = dict(order="AMD", ustab=0.1)
    mg, me = CSRMat.from_scipy(A), CSRMat.from_scipy(A)
    lg = SparseLDL(mg, **kw)
    le = SparseLDL(me, **kw)
    le.set_graphs(False)
    base = mg.data.clone()
    n = A.shape[0]
    rng = np.random.default_rng(0)
    xg = torch.zeros(n, dtype=torch.float64, device="cuda")
    xe = torch.zeros_like(xg)
    shift_g = torch.zeros(n, dtype=torch.float64, device="cuda")
    shift_e = torch.zeros_like(shift_g)
    for it in range(6):
        scale = 1.0 + 0.05 * it
        mg.data.copy_(base * scale)
        me.data.copy_(base * scale)
        d = torch.tensor(rng.uniform(0, 0.1, n), device="cuda")
        shift_g.copy_(d)
        shift_e.copy_(d)
        lg.factor(shift=shift_g)
        le.factor(shift=shift_e)
        # unrelated eager work and a synchronization between replays
        junk = torch.randn(1000, 1000, device="cuda") @ torch.randn(1000, 1000, device="cuda")
        torch.cuda.synchronize()
        assert lg.get_inertia3() == le.get_inertia3()
        pg, dg = lg.pivot_counts()
        pe, de = le.pivot_counts()
        assert np.array_equal(pg, pe) and np.array_equal(dg, de)
        for r in range(3):
            b = torch.tensor(rng.normal(size=n), device="cuda")
            xg.copy_(b)
            xe.copy_(b)
            lg.solve(xg)
            le.solve(xe)
            assert torch.equal(xg, xe), (it, r)
        del junk
