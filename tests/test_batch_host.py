"""Host logic of the batched form (CPU): the union problem's pattern and
gather lists, and the union symbolic plan against each problem's own
analysis."""

import numpy as np
import pytest

from amigo_amd.batch import BatchProblem, problem_ldl
from amigo_amd.problems import cart_pole, double_integrator
from amigo_amd.sparse import CSRMat, SparseLDL


def _problems():
    return ([cart_pole.seeded(s, nmin=16, nmax=64, device="cpu") for s in range(4)]
            + [double_integrator.build(N=20, device="cpu")])


def test_batch_problem_is_block_diagonal_union():
    from scipy.sparse import block_diag, csr_matrix

    probs = _problems()
    bp = BatchProblem(probs)
    blocks = [csr_matrix((np.ones(p.nnz), p.cols, p.rowp), shape=(p.n, p.n)) for p in probs]
    Bd = block_diag(blocks, format="csr")
    Bd.sort_indices()
    assert np.array_equal(Bd.indptr, bp.rowp) and np.array_equal(Bd.indices, bp.cols)
    rng = np.random.default_rng(0)
    for k, p in enumerate(probs):
        # Gather lists: the same sources in the same order, remapped.
        hb = rng.standard_normal(p.hsize)
        ub = np.zeros(bp.hsize)
        ub[bp._remap(k, p, np.arange(p.hsize), 2)] = hb
        off = bp.rowp[bp.xseg[k]]
        for i in range(p.nnz):
            a = hb[p.h_src[p.h_ptr[i]:p.h_ptr[i + 1]]]
            c = ub[bp.h_src[bp.h_ptr[off + i]:bp.h_ptr[off + i + 1]]]
            assert np.array_equal(a, c)
        gb = rng.standard_normal(p.gsize)
        ug = np.zeros(bp.gsize)
        ug[bp._remap(k, p, np.arange(p.gsize), 1)] = gb
        x0 = bp.xseg[k]
        for i in range(p.n):
            assert np.array_equal(gb[p.g_src[p.g_ptr[i]:p.g_ptr[i + 1]]],
                                  ug[bp.g_src[bp.g_ptr[x0 + i]:bp.g_ptr[x0 + i + 1]]])
        assert np.array_equal(np.where(p.diag_idx >= 0, p.diag_idx + off, -1), bp.diag_idx[x0:x0 + p.n])
        # Lagrangian partials: one block per 256 elements of each group.
        assert bp.lptr[k + 1] - bp.lptr[k] == sum((g.E + 255) // 256 for g in p.groups)


@pytest.mark.parametrize("order", ["ND", "AMD"])
def test_union_plan_keeps_each_problems_fronts(order):
    probs = _problems()
    bp = BatchProblem(probs)
    m = CSRMat(bp.n, bp.n, bp.rowp, bp.cols, None, device="cpu")
    U = SparseLDL.union(m, [problem_ldl(p, order) for p in probs]).get_symbolic()
    k0 = 0
    for k, p in enumerate(probs):
        mk = CSRMat(p.n, p.n, p.rowp, p.cols, None, device="cpu")
        S = SparseLDL(mk, "LDL", order=order, groups=p.var_groups).get_symbolic()
        nk = len(S["node_ptr"]) - 1
        x0 = bp.xseg[k]
        assert np.array_equal(U["node_ptr"][k0:k0 + nk + 1] - x0, S["node_ptr"])
        assert np.array_equal(U["node_to_var"][x0:x0 + p.n] - x0, S["node_to_var"])
        cp = U["contrib_ptr"][k0:k0 + nk + 1]
        assert np.array_equal(cp - cp[0], S["contrib_ptr"])
        assert np.array_equal(U["contrib_rows"][cp[0]:cp[-1]] - x0, S["contrib_rows"])
        assert np.array_equal(U["invperm"][x0:x0 + p.n] - x0, S["invperm"])
        # Children of every node, in the same order.
        for j in range(nk):
            cu = U["children"][U["children_ptr"][k0 + j]:U["children_ptr"][k0 + j + 1]] - k0
            ca = S["children"][S["children_ptr"][j]:S["children_ptr"][j + 1]]
            assert np.array_equal(cu, ca)
        k0 += nk
    assert k0 == len(U["node_ptr"]) - 1


def test_union_rejects_mismatched_pattern():
    from amigo_amd._lib import AmigoError

    probs = _problems()[:2]
    bp = BatchProblem(probs)
    m = CSRMat(bp.n, bp.n, bp.rowp, bp.cols, None, device="cpu")
    with pytest.raises(AmigoError):
        SparseLDL.union(m, [problem_ldl(p, "ND") for p in probs[::-1]])


def test_server_takes_one_problem_type_per_batch_in_arrival_order():
    """BatchSolveServer._take_batch: the oldest request's type first, in
    arrival order, at most max_batch; the rest stay queued in order."""
    from types import SimpleNamespace

    from amigo_amd.serve import BatchSolveServer

    kinds = ["cart_pole", "elliptic", "cart_pole", "double_integrator", "cart_pole", "elliptic"]
    pending = [({"type": t, "i": i}, None, float(i)) for i, t in enumerate(kinds)]
    srv = SimpleNamespace(max_batch=2, group_by_type=True)
    take, rest = BatchSolveServer._take_batch(srv, pending)
    assert [it[0]["i"] for it in take] == [0, 2]
    assert [it[0]["i"] for it in rest] == [1, 3, 4, 5]
    take, rest = BatchSolveServer._take_batch(srv, rest)
    assert [it[0]["i"] for it in take] == [1, 5]
    assert [it[0]["i"] for it in rest] == [3, 4]
    srv.group_by_type = False
    take, rest = BatchSolveServer._take_batch(srv, pending)
    assert [it[0]["i"] for it in take] == [0, 1] and len(rest) == 4
