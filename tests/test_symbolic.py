"""Host symbolic analysis of libamigo_amd against the reference.

The committed fixtures (tests/golden/make_golden.py, written by the
reference's own BlockAMD / SparseLDL compiled from its headers) pin the AMD
permutation, elimination tree, postorder, column counts and the supernodal
plan bit for bit; when oracle/_ref is built the same comparison runs live on
larger seeded matrices.  CPU only: the symbolic phase never touches the GPU.
"""

import ctypes

import numpy as np
import pytest

from tests import fixtures as FX
from tests import matrices as M

_lib = pytest.importorskip("amigo_amd._lib")


def _ours_order(A, order):
    L = _lib.lib()
    n = A.shape[0]
    rowp = np.ascontiguousarray(A.indptr, np.int32)
    cols = np.ascontiguousarray(A.indices, np.int32)
    perm = np.zeros(n, np.int32)
    iperm = np.zeros(n, np.int32)
    _lib.check(L.amigo_order(_lib.AMIGO_ORDER[order], n, _lib.ip(rowp), _lib.ip(cols), _lib.ip(perm), _lib.ip(iperm)))
    return perm, iperm


def _ours_etree(A, perm=None, iperm=None):
    L = _lib.lib()
    n = A.shape[0]
    rowp = np.ascontiguousarray(A.indptr, np.int32)
    cols = np.ascontiguousarray(A.indices, np.int32)
    out = [np.zeros(n, np.int32) for _ in range(3)]
    pp = None if perm is None else _lib.ip(np.ascontiguousarray(perm, np.int32))
    ipp = None if iperm is None else _lib.ip(np.ascontiguousarray(iperm, np.int32))
    _lib.check(L.amigo_etree_post_colcount(n, _lib.ip(rowp), _lib.ip(cols), pp, ipp, *[_lib.ip(a) for a in out]))
    return out


def _ours_symbolic(A, order):
    from amigo_amd.sparse import CSRMat, SparseLDL

    ldl = SparseLDL(CSRMat.from_scipy(A, device="cpu"), order=order)
    return ldl.get_symbolic()


@pytest.mark.parametrize("name", FX.ldl_cases())
def test_amd_permutation_matches_reference(name):
    A, _ = FX.matrix(name)
    perm, iperm = _ours_order(A, "AMD")
    assert np.array_equal(perm, FX.field(name, "AMD", "perm"))
    assert np.array_equal(iperm, FX.field(name, "AMD", "iperm"))


@pytest.mark.parametrize("name", FX.ldl_cases())
@pytest.mark.parametrize("order", ["NATURAL", "AMD"])
def test_etree_postorder_colcounts_match_reference(name, order):
    A, _ = FX.matrix(name)
    if order == "AMD":
        parent, post, cc = _ours_etree(A, FX.field(name, order, "perm"), FX.field(name, order, "iperm"))
    else:
        parent, post, cc = _ours_etree(A)
    assert np.array_equal(parent, FX.field(name, order, "parent"))
    assert np.array_equal(post, FX.field(name, order, "post"))
    assert np.array_equal(cc, FX.field(name, order, "colcount"))


@pytest.mark.parametrize("name", FX.ldl_cases())
@pytest.mark.parametrize("order", ["NATURAL", "AMD"])
def test_supernodal_plan_matches_reference(name, order):
    A, _ = FX.matrix(name)
    sym = _ours_symbolic(A, order)
    for key in FX.SYMBOLIC_KEYS:
        assert np.array_equal(sym[key], FX.field(name, order, key)), key


def _live_ref():
    from oracle import ref

    if not ref.available():
        pytest.skip("oracle/_ref not built")
    return ref


@pytest.mark.parametrize(
    "make",
    [
        lambda: M.with_full_diagonal(M.poisson_2d(50)),
        lambda: M.kkt(2000, 800, 7),
        lambda: M.kkt(1500, 900, 8, indefinite_h=True, density_h=0.003),
        lambda: M.grid_kkt(24)[0],
    ],
    ids=["poisson50", "kkt2000", "kkt_indef1500", "grid24"],
)
def test_symbolic_matches_live_reference(make):
    ref = _live_ref()
    A = make()
    perm, iperm = _ours_order(A, "AMD")
    rp, rip = ref.order(A.indptr, A.indices, "AMD")
    assert np.array_equal(perm, rp) and np.array_equal(iperm, rip)
    R = ref.RefLDL(A.indptr, A.indices, order="AMD")
    sym = _ours_symbolic(A, "AMD")
    rs = R.symbolic()
    for key in FX.SYMBOLIC_KEYS:
        assert np.array_equal(sym[key], rs[key]), key


def test_nested_dissection_is_a_permutation_with_groups_kept_together():
    """The build's own ND (the reference's needs METIS, absent here): a valid
    permutation; with groups, each group's variables are contiguous, in
    index order except that group-internal variables come last."""
    A, _, N = M.grid_kkt(16)
    n = A.shape[0]
    perm, iperm = _ours_order(A, "ND")
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert np.array_equal(iperm[perm], np.arange(n))
    from amigo_amd.sparse import CSRMat, SparseLDL

    # Group = grid node: y_i, u_i (internal: coupled to lambda_i only), lambda_i.
    groups = np.concatenate([np.arange(N), np.arange(N), np.arange(N)]).astype(np.int32)
    ldl = SparseLDL(CSRMat.from_scipy(A, device="cpu"), order="ND", groups=groups)
    ip = ldl.get_symbolic()["invperm"]
    p = np.argsort(ip)
    g = groups[p]
    # contiguous groups
    starts = np.nonzero(np.diff(g))[0] + 1
    runs = np.split(g, starts)
    assert len(runs) == N and all(len(r) == 3 for r in runs)
    # order inside a group: y, lambda, then u
    for r0 in range(0, n, 3):
        y, lam, u = p[r0:r0 + 3]
        assert y < N and lam >= 2 * N and N <= u < 2 * N
