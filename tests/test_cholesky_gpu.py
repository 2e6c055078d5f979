"""The Cholesky solver type of SparseLDL and the SparseCholesky drop-in
against the reference's SparseCholesky (oracle/_ref) and dense solves.

- The three SPD matrices of the reference's tests/unit/test_cholesky.py:
  solution x = 1 recovered (the reference's own check compares a vector
  with itself), for SparseLDL(CHOLESKY) and SparseCholesky.
- 2-D Poisson n = 10, 20, 50 with 7 right-hand sides in an F-ordered NumPy
  array, as tests/unit/test_ldl.py:230-314 (error < 1e-11 against x = 1).
- Quasi-definite [A B^T; B -C] split at sqdef_index: factor() returns 0 as
  the reference's does, the solution equals the reference's; with the split
  left out (plain Cholesky) both report a failure at the first variable of
  the -C block.
- NumPy solves: F-order in place, C-order through a copy, 1-D, and the
  binding's errors.
"""

import numpy as np
import pytest
import scipy.sparse as sp

from tests import matrices as M

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SPD = [
    [[2, 1, 0, 0, 0], [1, 2, 1, 0, 0], [0, 1, 2, 1, 0], [0, 0, 1, 2, 1], [0, 0, 0, 1, 2]],
    [[2, 0, 1, 0, 0], [0, 2, 0, 1, 0], [1, 0, 2, 0, 1], [0, 1, 0, 2, 1], [0, 0, 1, 1, 2]],
    [[2, 1, 1, 0, 0], [1, 2, 0, 1, 0], [1, 0, 2, 0, 1], [0, 1, 0, 2, 1], [0, 0, 1, 1, 2]],
]


def _ref():
    from oracle import ref

    if not ref.available():
        pytest.skip("reference oracle library not built")
    return ref


def _sqd(nA=300, nC=120, seed=0):
    A = sp.random(nA, nA, density=0.02, random_state=seed + 1)
    A = A @ A.T + 2.0 * sp.eye(nA)
    C = sp.random(nC, nC, density=0.05, random_state=seed + 2)
    C = C @ C.T + sp.eye(nC)
    B = sp.random(nC, nA, density=0.03, random_state=seed + 3)
    K = sp.bmat([[A, B.T], [B, -C]]).tocsr()
    K = (K + sp.eye(nA + nC) * 1e-300).tocsr()
    K.sort_indices()
    return K, nA


@pytest.mark.parametrize("index", range(3))
def test_reference_spd_matrices(index):
    from amigo_amd.sparse import CSRMat, SparseCholesky, SparseLDL

    ref = _ref()
    D = np.array(SPD[index], dtype=float)
    A = sp.csr_matrix(D)
    A.sort_indices()
    b = D @ np.ones(5)
    for solver in ("ldl_cholesky", "sparse_cholesky"):
        mat = CSRMat.from_scipy(A)
        if solver == "ldl_cholesky":
            s = SparseLDL(mat, solver_type="CHOLESKY", ustab=0.4)
            assert s.factor() == 0
        else:
            s = SparseCholesky(mat)
            assert s.factor() == 0
        x = b.copy()
        s.solve(x)
        assert np.max(np.abs(x - 1.0)) < 1e-14
    R = ref.RefCholesky(A.indptr, A.indices)
    assert R.factor(A.data) == 0
    assert np.allclose(R.solve(b), x, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("n", [10, 20, 50])
def test_poisson_seven_rhs(n):
    from amigo_amd.sparse import CSRMat, SparseLDL

    A = M.poisson_2d(n)
    N = A.shape[0]
    s = SparseLDL(CSRMat.from_scipy(A), solver_type="CHOLESKY", order="AMD")
    s.factor()
    x = np.asfortranarray(np.tile((A @ np.ones(N))[:, None], (1, 7)))
    s.solve(x)
    assert np.max(np.abs(x - 1.0)) < 1e-11


def test_quasi_definite_split_matches_the_reference():
    from amigo_amd.sparse import CSRMat, SparseCholesky

    ref = _ref()
    K, nA = _sqd()
    N = K.shape[0]
    chol = SparseCholesky(CSRMat.from_scipy(K, sqdef_index=nA))
    assert chol.factor() == 0
    R = ref.RefCholesky(K.indptr, K.indices, sqdef_index=nA)
    assert R.factor(K.data) == 0
    b = np.random.default_rng(4).normal(size=N)
    x = b.copy()
    chol.solve(x)
    xr = R.solve(b)
    assert np.linalg.norm(K @ x - b) / np.linalg.norm(b) < 1e-13
    assert np.allclose(x, xr, rtol=1e-11, atol=1e-12)
    # Without the split the matrix is not positive definite: both fail at
    # the first variable of the -C block (1-based).
    plain = SparseCholesky(CSRMat.from_scipy(K))
    R2 = ref.RefCholesky(K.indptr, K.indices)
    assert plain.factor() == R2.factor(K.data) == nA + 1


def test_not_quasi_definite_fails_like_the_reference():
    from amigo_amd.sparse import CSRMat, SparseCholesky

    ref = _ref()
    K, nA = _sqd(seed=5)
    K = K.tolil()
    K[7, 7] = -50.0  # an indefinite A block
    K = K.tocsr()
    K.sort_indices()
    mine = SparseCholesky(CSRMat.from_scipy(K, sqdef_index=nA)).factor()
    theirs = ref.RefCholesky(K.indptr, K.indices, sqdef_index=nA).factor(K.data)
    assert mine > 0 and theirs > 0


def test_numpy_solve_layouts_and_errors():
    from amigo_amd.sparse import CSRMat, SparseLDL

    A = M.kkt(200, 80, 3)
    N = A.shape[0]
    ldl = SparseLDL(CSRMat.from_scipy(A), order="AMD")
    ldl.factor()
    rng = np.random.default_rng(0)
    X = rng.normal(size=(N, 3))
    B = A @ X
    f = np.asfortranarray(B.copy())
    ptr = f.ctypes.data
    ldl.solve(f)
    assert f.ctypes.data == ptr and np.allclose(f, X, rtol=1e-9, atol=1e-9)
    c = np.ascontiguousarray(B.copy())
    ldl.solve(c)
    assert c.flags["C_CONTIGUOUS"] and np.allclose(c, X, rtol=1e-9, atol=1e-9)
    v = B[:, 1].copy()
    ldl.solve(v)
    assert np.allclose(v, X[:, 1], rtol=1e-9, atol=1e-9)
    with pytest.raises(TypeError):
        ldl.solve(B.astype(np.float32))
    ro = B.copy()
    ro.setflags(write=False)
    with pytest.raises(TypeError):
        ldl.solve(ro)
    with pytest.raises(TypeError):
        ldl.solve(np.zeros((N, 2, 2)))
    with pytest.raises(RuntimeError):
        ldl.solve([0.0] * N)
