"""Model-level renumbering (reorder_model) against the compiled reference:
the element graph equals OrderingUtils::create_csr_from_element_conn's and
the AMD permutation equals the reference BlockAMD's on it (oracle/_ref)."""

import numpy as np
import pytest

from amigo_amd.reorder import element_graph, reorder_model

ref = pytest.importorskip("oracle.ref")
pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")


def _trajectory(N, nx=4, nu=1, seed=0):
    """Collocation intervals (x_k, u_k, x_k+1, u_k+1, lambda_k) plus bound
    links, numbered stage by stage then shuffled."""
    rng = np.random.default_rng(seed)
    nv = (N + 1) * (nx + nu) + N * nx
    relabel = rng.permutation(nv)
    st = lambda k: list(range(k * (nx + nu), (k + 1) * (nx + nu)))
    lam = lambda k: list(range((N + 1) * (nx + nu) + k * nx, (N + 1) * (nx + nu) + (k + 1) * nx))
    elems = np.array([st(k) + st(k + 1) + lam(k) for k in range(N)])
    links = np.array([[i] for i in range(nx)])
    return [relabel[elems], relabel[links]], [nx, 0]


@pytest.mark.parametrize("seed", range(3))
def test_element_graph_equals_the_reference(seed):
    arrays, nc = _trajectory(12 + 5 * seed, seed=seed)
    rowp, cols = element_graph(arrays, nc)
    n = len(rowp) - 1
    rp, cl = ref.element_csr(n, arrays, nc)
    assert np.array_equal(rowp, rp) and np.array_equal(cols, cl)
    # With no constraint nodes every element is a full clique.
    rowp0, cols0 = element_graph(arrays, [0, 0])
    rp0, cl0 = ref.element_csr(n, arrays, [0, 0])
    assert np.array_equal(rowp0, rp0) and np.array_equal(cols0, cl0)


@pytest.mark.parametrize("seed", range(3))
def test_reorder_model_amd_equals_the_reference(seed):
    arrays, nc = _trajectory(20 + 7 * seed, seed=seed)
    rowp, cols = element_graph(arrays, nc)
    iperm = reorder_model("AMD", arrays, nc)
    perm = ref.block_amd(rowp, cols)
    exp = np.empty_like(perm)
    exp[perm] = np.arange(len(perm))
    assert np.array_equal(iperm, exp)
    # reorder_block: the constraint variables ordered last.
    cons = np.unique(arrays[0][:, -4:])
    iperm_b = reorder_model("AMD", arrays, nc, con_indices=cons)
    perm_b = ref.block_amd(rowp, cols, cons)
    exp_b = np.empty_like(perm_b)
    exp_b[perm_b] = np.arange(len(perm_b))
    assert np.array_equal(iperm_b, exp_b)


def test_reorder_model_natural_and_block_partition():
    arrays, nc = _trajectory(6)
    n = len(element_graph(arrays, nc)[0]) - 1
    assert np.array_equal(reorder_model("NATURAL", arrays, nc), np.arange(n))
    cons = np.array([1, 5, 7])
    iperm = reorder_model("NATURAL", arrays, nc, con_indices=cons)
    perm = np.argsort(iperm)
    assert set(perm[-3:]) == set(cons) and list(perm[:-3]) == [i for i in range(n) if i not in cons]
    with pytest.raises(ValueError):
        element_graph([np.array([[0, -1]])])
