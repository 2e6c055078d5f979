"""GPU tests of the batched form: many problems in one block-diagonal KKT.

* Per-block inertia and solves of a union of independent test matrices equal
  each matrix factored alone (bitwise: the union keeps every block's own
  symbolic plan, and each front is computed by the same kernel code).
* A masked factorization refactors only the blocks it names; the others
  keep their factors (solves with them still give their old solutions).
* The lockstep interior-point driver reproduces every problem's
  single-problem solve bit for bit (iterations, inertia sequence, solution),
  in FP64 and in the mixed-precision form.
"""

import numpy as np
import pytest

from tests import matrices as M

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _union(As, order="AMD", scaling=False):
    from scipy.sparse import block_diag

    from amigo_amd.sparse import CSRMat, SparseLDL

    blocks = []
    for A in As:
        A = A.tocsr()
        A.sort_indices()
        blocks.append(SparseLDL(CSRMat.from_scipy(A, device="cpu"), order=order))
    U = block_diag([A.tocsr() for A in As], format="csr")
    U.sort_indices()
    mat = CSRMat.from_scipy(U)
    ldl = SparseLDL.union(mat, blocks, scaling=scaling)
    offs = np.concatenate([[0], np.cumsum([A.shape[0] for A in As])])
    K = [b.sizes()["num_nodes"] for b in blocks]
    ldl.set_batch(np.concatenate([[0], np.cumsum(K)]).astype(np.int32))
    return U, mat, ldl, offs


def _alone(A, order="AMD", scaling=False, shift=None):
    from amigo_amd.sparse import CSRMat, SparseLDL

    mat = CSRMat.from_scipy(A)
    ldl = SparseLDL(mat, order=order, scaling=scaling)
    if shift is None:
        ldl.factor()
    else:
        ldl.factor(shift=torch.full((A.shape[0],), shift, dtype=torch.float64, device="cuda"))
    return ldl


def _mats():
    As = [M.small(i)[0] for i in range(3)]
    As.append(M.grid_kkt(6, seed=3)[0])
    As.append(M.kkt(40, 15, seed=2, indefinite_h=True))
    As.append(M.poisson_2d(10, seed=0))
    # Full diagonal patterns, so that a diagonal shift keeps the pattern.
    return [M.with_full_diagonal(A.tocsr()) for A in As]


@pytest.mark.parametrize("scaling", [False, True])
def test_union_inertia_and_solves_equal_each_block_alone(scaling):
    As = _mats()
    U, mat, ldl, offs = _union(As, scaling=scaling)
    ldl.factor()
    npos, nneg, nzero, nnan = ldl.get_batch_inertia()
    rng = np.random.default_rng(0)
    b = rng.normal(size=U.shape[0])
    xu = torch.tensor(b, device="cuda")
    ldl.solve(xu)
    xu = xu.cpu().numpy()
    for k, A in enumerate(As):
        a = _alone(A, scaling=scaling)
        assert (npos[k], nneg[k]) == a.get_inertia()
        assert nnan[k] == 0
        xa = torch.tensor(b[offs[k]:offs[k + 1]], device="cuda")
        a.solve(xa)
        assert np.array_equal(xa.cpu().numpy(), xu[offs[k]:offs[k + 1]])


def test_union_solves_equal_alone_whatever_the_launch_width():
    """A level of at most 256 fronts with one of at least 256 rows is
    forward-solved with 1,024-thread workgroups, other levels with 256.
    Alone, the KKT below has such levels (fronts of 274, 280 and 311 rows at
    three levels of 1-3 fronts); in the union with 300 chains (one front at
    each of 8 levels) the same levels hold over 300 fronts and take the
    narrow launch.  Its solution must not change by a bit."""
    import scipy.sparse as sp

    big = M.with_full_diagonal(M.kkt(600, 200, seed=4, density_h=0.005, indefinite_h=True).tocsr())
    rng = np.random.default_rng(3)
    chains = []
    for _ in range(300):
        n = 500
        chains.append(sp.diags([np.full(n - 1, -1.0), 4.0 + rng.random(n), np.full(n - 1, -1.0)],
                               [-1, 0, 1], format="csr"))
    As = [big] + chains
    U, mat, ldl, offs = _union(As)
    ldl.factor()
    b = rng.normal(size=U.shape[0])
    xu = torch.tensor(b, device="cuda")
    ldl.solve(xu)
    a = _alone(big)
    xa = torch.tensor(b[:offs[1]], device="cuda")
    a.solve(xa)
    assert np.array_equal(xa.cpu().numpy(), xu.cpu().numpy()[:offs[1]])
    r = big @ xa.cpu().numpy() - b[:offs[1]]
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b[:offs[1]])


def test_masked_factorization_keeps_left_out_factors():
    As = _mats()
    U, mat, ldl, offs = _union(As)
    n = U.shape[0]
    dev = "cuda"
    s0 = torch.zeros(n, dtype=torch.float64, device=dev)
    ldl.factor(shift=s0)
    rng = np.random.default_rng(1)
    b = rng.normal(size=n)
    x0 = torch.tensor(b, device=dev)
    ldl.solve(x0)
    x0 = x0.cpu().numpy()
    # Shift every block, refactor only blocks 0, 2, 4.
    s1 = torch.full((n,), 0.5, dtype=torch.float64, device=dev)
    mask = np.array([k % 2 == 0 for k in range(len(As))])
    ldl.set_batch_mask(mask)
    ldl.factor(shift=s1)
    npos, nneg, _, _ = ldl.get_batch_inertia()
    ldl.set_batch_mask(None)
    x1 = torch.tensor(b, device=dev)
    ldl.solve(x1)
    x1 = x1.cpu().numpy()
    for k, A in enumerate(As):
        seg = slice(offs[k], offs[k + 1])
        if mask[k]:
            a = _alone(A, shift=0.5)
            assert (npos[k], nneg[k]) == a.get_inertia()
            xa = torch.tensor(b[seg], device=dev)
            a.solve(xa)
            assert np.array_equal(xa.cpu().numpy(), x1[seg])
        else:
            # Left out: still the factor of the unshifted matrix.
            assert np.array_equal(x0[seg], x1[seg])


def _problems():
    from amigo_amd.problems import cart_pole, double_integrator, elliptic

    ps = [cart_pole.seeded(200 + s, nmin=16, nmax=80) for s in range(5)]
    ps.append(double_integrator.build(N=24))
    ps.append(elliptic.EllipticControl(m=10, dim=2, seed=4))
    return ps


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_solves_equal_single_problem_solves(precision):
    from amigo_amd.batch import BatchOptimizer
    from amigo_amd.optimizer import Optimizer

    probs = _problems()
    opts = {"record_objective": False}
    bo = BatchOptimizer(probs, order="ND", scaling=True, precision=precision)
    res = bo.optimize(opts)
    xs = bo.get_solution()
    assert all(r["converged"] for r in res)
    for k, p in enumerate(probs):
        o = Optimizer(p, order="ND", scaling=True, precision=precision)
        d = o.optimize(opts)
        assert d["converged"]
        assert d["num_iterations"] == res[k]["num_iterations"], k
        assert d["inertia_history"] == res[k]["inertia_history"], k
        assert np.array_equal(o.get_solution(), xs[k]), (k, np.max(np.abs(o.get_solution() - xs[k])))


def test_batched_solve_repeats_bitwise():
    from amigo_amd.batch import BatchOptimizer

    probs = _problems()[:4]
    bo = BatchOptimizer(probs)
    bo.optimize({"record_objective": False})
    x1 = [x.copy() for x in bo.get_solution()]
    bo.optimize({"record_objective": False})
    x2 = bo.get_solution()
    assert all(np.array_equal(a, b) for a, b in zip(x1, x2))


def test_concurrent_factorizations_on_separate_streams_repeat_bitwise():
    """Three batches factored at once from three threads, each on its own
    stream, so that their kernels share the CUs: every factorization and
    solve equals the one computed with the device to itself.  (Waves of a
    workgroup that fall behind under such sharing must still see the pivot
    entries of the front before any wave interchanges rows.)"""
    import threading

    from amigo_amd.batch import BatchOptimizer
    from amigo_amd.problems import cart_pole, elliptic

    def build():
        ps = [cart_pole.seeded(300 + s, nmin=300, nmax=900) for s in range(6)]
        ps.append(elliptic.EllipticControl(m=60, dim=2, seed=5))
        return ps

    def run(out, tag, reps):
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            bo = BatchOptimizer(build(), order="ND", scaling=True)
            sol = bo.solver
            sol.ldl.set_graphs(False)
            bo.optimize({"max_iterations": 3, "record_objective": False})
            sol.assemble_hessian(True, bo.x)
            bo.be.apply_hessian_scaling(sol.hess)
            diag = bo.diag.clone()
            b = torch.tensor(np.random.default_rng(0).normal(size=diag.numel()), device="cuda")
            xs = []
            for _ in range(reps):
                sol.add_diagonal_and_factor(diag)
                x = b.clone()
                sol.ldl.solve(x)
                xs.append((x.cpu().numpy(), [tuple(v) for v in zip(*sol.ldl.get_batch_inertia())]))
            s.synchronize()
        out[tag] = xs

    alone = {}
    run(alone, "a", 1)
    x0, in0 = alone["a"][0]
    out = {}
    ths = [threading.Thread(target=run, args=(out, t, 12)) for t in range(3)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert len(out) == 3
    for t, xs in out.items():
        for r, (x, inert) in enumerate(xs):
            assert inert == in0, (t, r)
            assert np.array_equal(x, x0), (t, r, float(np.max(np.abs(x - x0))))


def test_segment_absmax_equals_numpy_max_per_segment():
    """The per-problem divergence-test reduction: max |x| of each segment,
    exact, NaN propagating as in np.max, 0 for an empty segment."""
    import ctypes

    from amigo_amd import _lib
    from amigo_amd.batch import _p, _s

    rng = np.random.default_rng(0)
    lens = [0, 1, 5, 256, 1000, 3, 4097, 64]
    seg = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    x = rng.normal(size=seg[-1]) * 10.0 ** rng.integers(-5, 5, size=seg[-1])
    x[seg[4] + 17] = np.nan
    x[seg[6] + 4000] = -1e300
    xd = torch.tensor(x, device="cuda")
    sd = torch.tensor(seg, device="cuda")
    out = torch.zeros(len(lens), dtype=torch.float64, device="cuda")
    h = np.zeros(len(lens))
    _lib.check(_lib.lib().amigo_segment_absmax(_p(xd), _p(sd), len(lens), _p(out),
                                                h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                _s()))
    for k in range(len(lens)):
        s = x[seg[k]:seg[k + 1]]
        exp = np.max(np.abs(s)) if len(s) else 0.0
        if np.isnan(exp):
            assert np.isnan(h[k]), k
        else:
            assert h[k] == exp, k
    assert np.array_equal(out.cpu().numpy(), h, equal_nan=True)


def test_batch_of_one_problem_equals_its_single_problem_solve():
    from amigo_amd.batch import BatchOptimizer
    from amigo_amd.optimizer import Optimizer
    from amigo_amd.problems import cart_pole

    p = cart_pole.seeded(321, nmin=40, nmax=200)
    bo = BatchOptimizer([p], order="ND", scaling=True)
    r = bo.optimize({"record_objective": False})[0]
    o = Optimizer(p, order="ND", scaling=True)
    d = o.optimize({"record_objective": False})
    assert r["converged"] and d["converged"]
    assert r["num_iterations"] == d["num_iterations"]
    assert r["inertia_history"] == d["inertia_history"]
    assert np.array_equal(bo.get_solution()[0], o.get_solution())


def test_listed_launches_equal_masked_launches():
    """Batched IPM kernels launched over lists of the active problems'
    blocks (and final passes over the list of those problems) give the
    results of the full launches with the same mask, bit for bit, and
    leave the other problems' entries alone."""
    from amigo_amd.batch import BatchIPM, BatchProblem, BVec
    from amigo_amd.problems import cart_pole

    probs = [cart_pole.seeded(300 + s, nmin=16, nmax=200) for s in range(12)]
    bp = BatchProblem(probs)
    ipm = BatchIPM(bp, bp.lower, bp.upper)
    K, n, np_ = bp.K, bp.n, ipm.np_
    g = torch.Generator(device="cuda").manual_seed(5)
    R = lambda m, lo=0.1: lo + torch.rand(m, dtype=torch.float64, device="cuda", generator=g)
    v = BVec(n, np_, "cuda", R(n, -0.5))
    for t in (v.zl, v.zu, v.sl, v.su):
        t.copy_(R(np_))
    upd = BVec(n, np_, "cuda", R(n, -0.5))
    for t in (upd.zl, upd.zu):
        t.copy_(R(np_, -0.5))
    grad = R(n, -0.5)
    act = [3, 8, 9]
    mask = np.zeros(K, np.int32)
    mask[act] = 1

    def run():
        out = {}
        out["abs"] = ipm.reduce("abs_sums", v, mask=mask)[act]
        out["kkt"] = ipm.reduce("kkt_err", v, 0.1, grad, mask=mask)[act]
        res = torch.full((n,), 7.0, dtype=torch.float64, device="cuda")
        out["res"] = ipm.residual(0.1, v, grad, res, mask=mask)[act]
        out["res_vec"] = res.cpu().numpy()
        out["step"] = np.stack(ipm.max_step(0.99, v, upd, mask=mask), axis=1)[act]
        tmp = BVec(n, np_, "cuda", torch.full((n,), 3.0, dtype=torch.float64, device="cuda"))
        ipm.apply_step(0.5, 0.25, v, upd, tmp, mask=mask)
        out["tmp"] = np.concatenate([t.cpu().numpy() for t in (tmp.x, tmp.zl, tmp.zu, tmp.sl, tmp.su)])
        torch.cuda.synchronize()
        return out

    full = run()
    # launch lists of the active problems' blocks
    keep = []
    for tab, (nf, cnt) in zip(ipm._tabs, (("nlaunch_total", "launch_total"), ("nlaunch_primal", "launch_primal"))):
        first, count = tab[2].cpu().numpy(), tab[3].cpu().numpy()
        lst = np.concatenate([first[k] + np.arange(count[k]) for k in act]).astype(np.int32)
        d = torch.as_tensor(lst, device="cuda")
        keep.append(d)
        setattr(ipm._batch, nf, len(lst))
        setattr(ipm._batch, cnt, d.data_ptr())
    plist = torch.as_tensor(np.asarray(act, np.int32), device="cuda")
    ipm._batch.nlaunch_problems = len(act)
    ipm._batch.launch_problems = plist.data_ptr()
    try:
        listed = run()
    finally:
        ipm._batch.launch_total = ipm._batch.launch_primal = ipm._batch.launch_problems = None
        ipm._batch.nlaunch_total = ipm._batch.nlaunch_primal = ipm._batch.nlaunch_problems = 0
    for key in full:
        assert np.array_equal(full[key].view(np.int64), listed[key].view(np.int64)), key


def test_listed_gather_sum_equals_masked_gather_sum():
    """amigo_gather_sum_listed writes exactly what the masked gather writes
    for the listed groups and nothing else."""
    import ctypes

    from amigo_amd import _lib

    rng = np.random.default_rng(11)
    seg = np.concatenate([[0], np.cumsum(rng.integers(0, 900, size=9))]).astype(np.int64)
    n = int(seg[-1])
    cnt = rng.integers(0, 5, size=n)
    ptr = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    buf = rng.standard_normal(int(ptr[-1]) + 7)
    src = rng.integers(0, len(buf), size=int(ptr[-1])).astype(np.int64)
    act = np.array([1, 4, 5, 8], np.int32)
    mask = np.zeros(len(seg) - 1, np.int32)
    mask[act] = 1
    T = lambda a: torch.as_tensor(a, device="cuda")
    d = {k: T(a) for k, a in dict(ptr=ptr, src=src, buf=buf, seg=seg, act=act, mask=mask).items()}
    L = _lib.lib()
    outs = []
    for accumulate in (0, 1):
        o1 = T(rng.standard_normal(n))
        o2 = o1.clone()
        _lib.check(L.amigo_gather_sum_masked(n, d["ptr"].data_ptr(), d["src"].data_ptr(), d["buf"].data_ptr(),
                                             o1.data_ptr(), accumulate, d["seg"].data_ptr(), len(seg) - 1,
                                             d["mask"].data_ptr(), None), "asm")
        _lib.check(L.amigo_gather_sum_listed(d["ptr"].data_ptr(), d["src"].data_ptr(), d["buf"].data_ptr(),
                                             o2.data_ptr(), accumulate, d["seg"].data_ptr(), d["act"].data_ptr(),
                                             len(act), int(np.max(np.diff(seg)[act])), None), "asm")
        torch.cuda.synchronize()
        assert torch.equal(o1, o2), accumulate
        outs.append(o1)
