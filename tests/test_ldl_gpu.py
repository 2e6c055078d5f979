"""GPU parity of the multifrontal LDL^T against the reference's SparseLDL.

For every matrix: the inertia equals the reference's (and the eigenvalue
inertia for the small ones), and the solve residual ||Ax - b|| / ||b|| is at
FP64 level.  Per-node pivot/delay counts are compared with the reference.
"""

import numpy as np
import pytest

from tests import matrices as M

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _ref():
    from oracle import ref

    if not ref.available():
        pytest.skip("reference oracle library not built")
    return ref


def _gpu_ldl(A, order, ustab=0.01, delay_growth=2.0, precision="fp64"):
    from amigo_amd.sparse import CSRMat, SparseLDL

    mat = CSRMat.from_scipy(A)
    ldl = SparseLDL(mat, ustab=ustab, delay_growth=delay_growth, order=order, precision=precision)
    assert ldl.factor() == 0
    return mat, ldl


def _residual(A, ldl, nrhs=1, seed=0):
    rng = np.random.default_rng(seed)
    n = A.shape[0]
    xs = rng.normal(size=(n, nrhs))
    b = A @ xs
    bt = torch.tensor(np.asfortranarray(b), device="cuda")
    bt = bt.t().contiguous().t() if nrhs > 1 else bt[:, 0].contiguous()
    ldl.solve(bt)
    x = bt.cpu().numpy().reshape(n, nrhs)
    r = np.linalg.norm(A @ x - b) / np.linalg.norm(b)
    return r, x


@pytest.mark.parametrize("index", range(7))
@pytest.mark.parametrize("order", ["NATURAL", "AMD"])
def test_small_matrices(index, order):
    A, inertia = M.small(index)
    _, ldl = _gpu_ldl(A, order, ustab=0.4)
    assert ldl.get_inertia() == inertia
    r, _ = _residual(A, ldl)
    assert r < 1e-13


@pytest.mark.parametrize("n", [10, 20, 50])
@pytest.mark.parametrize("order", ["NATURAL", "AMD"])
def test_poisson(n, order):
    ref = _ref()
    A = M.poisson_2d(n)
    _, ldl = _gpu_ldl(A, order, ustab=0.4, delay_growth=1.0)
    R = ref.RefLDL(A.indptr, A.indices, ustab=0.4, delay_growth=1.0, order=order)
    R.factor(A.data)
    assert ldl.get_inertia() == R.inertia() == (A.shape[0], 0)
    r, x = _residual(A, ldl, nrhs=7)
    assert r < 1e-12


@pytest.mark.parametrize(
    "nx,nc,seed,order",
    [(300, 120, 3, "AMD"), (300, 120, 3, "NATURAL"), (1500, 700, 4, "AMD"), (2000, 900, 5, "NATURAL")],
)
def test_kkt_inertia_and_solve(nx, nc, seed, order):
    ref = _ref()
    A = M.kkt(nx, nc, seed)
    _, ldl = _gpu_ldl(A, order)
    R = ref.RefLDL(A.indptr, A.indices, order=order)
    R.factor(A.data)
    assert ldl.get_inertia() == R.inertia() == (nx, nc)
    r, _ = _residual(A, ldl)
    assert r < 1e-10
    np_g, nd_g = ldl.pivot_counts()
    np_r, nd_r = R.pivot_counts()
    # The same threshold tests run on values that differ only by rounding.
    assert np.sum(nd_g != nd_r) <= max(2, len(nd_r) // 100)


@pytest.mark.parametrize("seed", [0, 1])
def test_indefinite_hessian_kkt(seed):
    ref = _ref()
    A = M.kkt(400, 150, seed, indefinite_h=True)
    _, ldl = _gpu_ldl(A, "AMD")
    R = ref.RefLDL(A.indptr, A.indices, order="AMD")
    R.factor(A.data)
    ev = np.linalg.eigvalsh(A.toarray())
    assert ldl.get_inertia() == R.inertia() == (int(np.sum(ev >= 0)), int(np.sum(ev < 0)))
    r, _ = _residual(A, ldl)
    assert r < 1e-9


@pytest.mark.parametrize("side", [16, 40])
def test_grid_kkt(side):
    ref = _ref()
    A, npr, ncon = M.grid_kkt(side, seed=side)
    _, ldl = _gpu_ldl(A, "AMD")
    R = ref.RefLDL(A.indptr, A.indices, order="AMD")
    R.factor(A.data)
    assert ldl.get_inertia() == R.inertia()
    r, _ = _residual(A, ldl)
    assert r < 1e-10


def test_refactor_same_handle():
    A = M.kkt(500, 200, 7)
    mat, ldl = _gpu_ldl(A, "AMD")
    for s in (1.0, 2.0, 0.5):
        mat.data.copy_(torch.tensor(A.data * s, device="cuda"))
        ldl.factor()
        assert ldl.get_inertia() == (500, 200)
        As = A * s
        r, _ = _residual(As, ldl)
        assert r < 1e-10


def test_shifted_factor():
    A = M.kkt(300, 100, 9, sigma=0.0, indefinite_h=True)
    mat, ldl = _gpu_ldl(A, "AMD")
    n = A.shape[0]
    shift = np.zeros(n)
    shift[:300] = 10.0
    shift[300:] = -1e-3
    ldl.factor(shift=torch.tensor(shift, device="cuda"))
    from scipy.sparse import diags

    B = (A + diags(shift)).tocsr()
    ev = np.linalg.eigvalsh(B.toarray())
    assert ldl.get_inertia() == (int(np.sum(ev >= 0)), int(np.sum(ev < 0)))
    r, _ = _residual(B, ldl)
    assert r < 1e-10


def _dense_block_tail(nb=200, tail=60, seed=1):
    from scipy.sparse import csr_matrix

    rng = np.random.default_rng(seed)
    B = rng.normal(size=(nb, nb))
    S = B @ B.T + nb * np.eye(nb)
    n = nb + tail
    A = np.zeros((n, n))
    A[:nb, :nb] = S
    for i in range(nb, n):
        A[i, i] = 4.0
        A[i, i - 1] = A[i - 1, i] = 1.0
    return csr_matrix(A)


@pytest.mark.parametrize("case", ["root", "nonroot", "nonroot_shift", "indefinite"])
def test_blocked_path_fronts(case):
    """Fronts above the workgroup limit (blocked panel path + grid-wide
    assembly): a dense root front, a dense non-root front, the fused
    diagonal shift, and an indefinite dense front with 2x2 pivots."""
    from amigo_amd.sparse import CSRMat, SparseLDL
    from scipy.sparse import csr_matrix

    if case == "root":
        A = _dense_block_tail(300, 0)
    elif case == "indefinite":
        rng = np.random.default_rng(3)
        B = rng.normal(size=(260, 260))
        A = csr_matrix(B + B.T)
    else:
        A = _dense_block_tail(200, 60)
    A.sort_indices()
    ldl = SparseLDL(CSRMat.from_scipy(A), order="NATURAL")
    Ad = A.toarray()
    if case == "nonroot_shift":
        sh = np.linspace(0.1, 2.0, A.shape[0])
        ldl.factor(shift=torch.tensor(sh, device="cuda"))
        Ad = Ad + np.diag(sh)
    else:
        ldl.factor()
    assert ldl.stats()[1] > 128
    ev = np.linalg.eigvalsh(Ad)
    assert ldl.get_inertia() == (int(np.sum(ev >= 0)), int(np.sum(ev < 0)))
    b = np.random.default_rng(0).normal(size=A.shape[0])
    x = torch.tensor(b, device="cuda")
    ldl.solve(x)
    r = np.linalg.norm(Ad @ x.cpu().numpy() - b) / np.linalg.norm(b)
    assert r < 1e-10


def test_delays_with_wide_fronts_match_oracle():
    """Elliptic-control KKT with many delayed pivots: per-node pivot and
    delay counts equal the oracle restatement's on the same plan."""
    from amigo_amd.sparse import CSRMat, SparseLDL
    from oracle.ldl import OracleLDL

    A = M.grid_kkt(16, seed=2)[0]
    ldl = SparseLDL(CSRMat.from_scipy(A), order="AMD", ustab=0.1)
    ldl.factor()
    O = OracleLDL(A.shape[0], A.indptr, A.indices, ldl.get_symbolic(), ustab=0.1)
    O.factor(A.data)
    npiv, ndel = ldl.pivot_counts()
    assert ldl.get_inertia() == O.get_inertia()
    assert np.array_equal(npiv, O.npiv) and np.array_equal(ndel, O.ndel)
    assert ndel.sum() > 50
    b = np.random.default_rng(1).normal(size=A.shape[0])
    x = torch.tensor(b, device="cuda")
    ldl.solve(x)
    assert np.allclose(x.cpu().numpy(), O.solve(b), rtol=1e-9, atol=1e-9)


def _delays_beyond_static_panels_matrix(nleaf=40, ns=64, tail=60, nroot=0, seed=5):
    """NATURAL order: nleaf one-variable leaves with a zero diagonal, each
    coupled to all ns variables of a dense block (ns = 64 columns, so no
    leaf is merged into it), then a tail of rows coupled to the whole block,
    then nroot rows coupled to the whole tail.  Every leaf delays its pivot,
    so the front that takes them holds 40 more candidates than its static
    panel plan (its own pivots plus 16).  With nroot = 0 the block and the
    tail form one root front; with nroot > 0 the block's front (contribution
    rows: the tail) has the tail and root rows' front as its parent."""
    from scipy.sparse import csr_matrix

    rng = np.random.default_rng(seed)
    n = nleaf + ns + tail + nroot
    A = np.zeros((n, n))
    b0, t0, r0 = nleaf, nleaf + ns, nleaf + ns + tail
    B = rng.normal(size=(ns, ns))
    A[b0:t0, b0:t0] = B + B.T + np.diag(np.full(ns, 2.0 * ns))
    for i in range(nleaf):
        A[i, b0:t0] = A[b0:t0, i] = rng.normal(size=ns)
    for t in range(tail):
        i = t0 + t
        A[i, i] = 4.0 if t % 2 else -3.0
        A[i, b0:t0] = A[b0:t0, i] = 0.1 * rng.normal(size=ns)
        if nroot:
            A[i, r0:] = A[r0:, i] = 0.1 * rng.normal(size=nroot)
    for q in range(nroot):
        A[r0 + q, r0 + q] = 5.0
    M_ = csr_matrix(A + np.diag(np.full(n, 1e-300)))
    M_.data[np.abs(M_.data) <= 1e-299] = 0.0
    M_.sort_indices()
    return M_


@pytest.mark.parametrize("nroot", [0, 8])
def test_delays_beyond_the_static_panel_plan(nroot):
    """A blocked front with more delayed pivots than its planned panel
    launches hold is found unfinished (as a root front, and as a front
    whose parent's level then sees it) and the factorization is redone
    with per-level sizes: inertia equal to the eigenvalue inertia, pivot
    counts showing every leaf delayed, residual of a solve <= 1e-10."""
    from amigo_amd.sparse import CSRMat, SparseLDL

    A = _delays_beyond_static_panels_matrix(nroot=nroot)
    ldl = SparseLDL(CSRMat.from_scipy(A), order="NATURAL")
    for _ in range(2):  # the second factorization reuses the handle's plan
        ldl.factor()
        Ad = A.toarray()
        ev = np.linalg.eigvalsh(Ad)
        assert ldl.get_inertia() == (int(np.sum(ev >= 0)), int(np.sum(ev < 0)))
        npiv, ndel = ldl.pivot_counts()
        assert ndel.sum() >= 40
        b = np.random.default_rng(2).normal(size=A.shape[0])
        x = torch.tensor(b, device="cuda")
        ldl.solve(x)
        r = np.linalg.norm(Ad @ x.cpu().numpy() - b) / np.linalg.norm(b)
        assert r < 1e-10


def _tall_front_matrix(nb=1100, tail=48, seed=11):
    """A dense symmetric indefinite block of nb rows (every tenth diagonal
    entry zero, so 2x2 pivots occur) followed by a tail of
    sparsely coupled rows.  In NATURAL order the dense block is cut into
    64-column supernodes whose first fronts hold more than 1024 rows: the
    global-memory panel path."""
    from scipy.sparse import csr_matrix

    rng = np.random.default_rng(seed)
    B = rng.normal(size=(nb, nb))
    S = B + B.T
    S[np.arange(0, nb, 10), np.arange(0, nb, 10)] = 0.0
    n = nb + tail
    A = np.zeros((n, n))
    A[:nb, :nb] = S
    for t in range(tail):
        i = nb + t
        A[i, i] = 3.0 if t % 2 else -2.0
        for j in rng.choice(nb, 3, replace=False):
            A[i, j] = A[j, i] = rng.normal()
        if t:
            A[i, i - 1] = A[i - 1, i] = 0.5
    M_ = csr_matrix(A + np.diag(np.full(n, 1e-300)))
    M_.data[np.abs(M_.data) <= 1e-299] = 0.0
    M_.sort_indices()
    return M_


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fronts_taller_than_1024_rows(precision):
    """The global-memory panel (fronts of more than 1024 rows): inertia
    equals the eigenvalue inertia and the reference's; in FP64 the per-node
    pivot and delay counts equal the oracle restatement's; solves reach
    FP64 residuals (the mixed form after FP64 iterative refinement)."""
    from amigo_amd.sparse import CSRMat, SparseLDL
    from oracle.ldl import OracleLDL

    ref = _ref()
    A = _tall_front_matrix()
    ldl = SparseLDL(CSRMat.from_scipy(A), order="NATURAL", precision=precision)
    ldl.factor()
    assert ldl.stats()[1] > 1024
    Ad = A.toarray()
    ev = np.linalg.eigvalsh(Ad)
    eig = (int(np.sum(ev >= 0)), int(np.sum(ev < 0)))
    R = ref.RefLDL(A.indptr, A.indices, order="NATURAL")
    R.factor(A.data)
    assert ldl.get_inertia() == R.inertia() == eig
    b = np.random.default_rng(2).normal(size=A.shape[0])
    x = torch.tensor(b, device="cuda")
    ldl.solve(x)
    r = np.linalg.norm(Ad @ x.cpu().numpy() - b) / np.linalg.norm(b)
    if precision == "fp64":
        O = OracleLDL(A.shape[0], A.indptr, A.indices, ldl.get_symbolic())
        O.factor(A.data)
        npiv, ndel = ldl.pivot_counts()
        assert O.get_inertia() == eig
        assert np.array_equal(npiv, O.npiv) and np.array_equal(ndel, O.ndel)
        np_r, _ = R.pivot_counts()
        assert np.array_equal(npiv, np_r)
        assert r < 1e-10
    else:
        # FP32 Schur updates: the plain solve is accurate to the FP32 level
        # only; FP64 refinement on the FP64 matrix recovers FP64 residuals.
        assert r < 1e-1
        xh = x.cpu().numpy()
        for _ in range(15):
            d = torch.tensor(b - Ad @ xh, device="cuda")
            ldl.solve(d)
            xh = xh + d.cpu().numpy()
        r = np.linalg.norm(Ad @ xh - b) / np.linalg.norm(b)
        assert r < 1e-12
