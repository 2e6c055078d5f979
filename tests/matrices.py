"""Seeded test matrices for the LDL^T parity tests.

- ``small(index)``: the seven hand matrices of the reference's
  tests/unit/test_ldl.py:20-82 (SPD and indefinite, 2x2..5x5).
- ``poisson_2d(n)``: the 5-point Laplacian on an (n-1)^2 grid, randomly
  symmetric-permuted with np.random.seed(0) as tests/unit/test_ldl.py:85-118.
- ``kkt(...)``: seeded symmetric indefinite KKT matrices
  [[H + sigma, J^T], [J, -delta]] with the full pattern (diagonal included),
  the structure the interior-point solver factors.
"""

import numpy as np
from scipy.sparse import bmat, csr_matrix, diags, eye, kron
from scipy.sparse import random as sprandom

SMALL = [
    [[2, 1, 0, 0, 0], [1, 2, 1, 0, 0], [0, 1, 2, 1, 0], [0, 0, 1, 2, 1], [0, 0, 0, 1, 2]],
    [[0, 1, 0, 0, 0], [1, 0, 1, 0, 0], [0, 1, 2, 1, 0], [0, 0, 1, 2, 1], [0, 0, 0, 1, 2]],
    [[0, 0, 0, 1, 1], [0, 0, 1, 0, 1], [0, 1, 2, 0, 1], [1, 0, 0, 0, 1], [1, 1, 1, 1, 0]],
    [[0, 1, 0, 0, 1], [1, 0, 2, 0, 1], [0, 2, 0, 0, 1], [0, 0, 0, 2, 1], [1, 1, 1, 1, 0]],
    [[0, 1], [1, 0]],
    [[1, 1], [1, 0]],
    [[1, 2, 3], [2, 1, 3], [3, 3, 0]],
]


def eig_inertia(M, tol=1e-12):
    """(npos, nneg) from the eigenvalues of a small dense matrix
    (tests/unit/test_ldl.py:10-17)."""
    ev = np.linalg.eigvalsh(M)
    if np.min(np.abs(ev)) < tol:
        raise ValueError("singular matrix")
    return int(np.sum(ev >= 0)), int(np.sum(ev < 0))


def small(index):
    M = np.array(SMALL[index], dtype=float)
    A = csr_matrix(M)
    A.sort_indices()
    return A, eig_inertia(M)


def poisson_2d(n, seed=0):
    m = n - 1
    T = diags([-np.ones(m - 1), 2.0 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1], format="csr")
    A = ((kron(eye(m), T) + kron(T, eye(m))) * n**2).tocsr()
    np.random.seed(seed)
    p = np.random.permutation(m * m)
    A = A[p, :][:, p].tocsr()
    A.sort_indices()
    return A


def with_full_diagonal(A):
    n = A.shape[0]
    D = csr_matrix((np.full(n, 1e-300), (np.arange(n), np.arange(n))), shape=(n, n))
    B = (A + D).tocsr()
    B.data[np.abs(B.data) <= 1e-299] = 0.0
    B.sort_indices()
    return B


def kkt(nx, nc, seed, density_h=0.01, density_j=None, sigma=1.0, delta=0.0, indefinite_h=False):
    """Seeded KKT matrix with full diagonal pattern."""
    rng = np.random.default_rng(seed)
    if density_j is None:
        density_j = min(1.0, 4.0 / nx)
    H = sprandom(nx, nx, density=density_h, random_state=seed, format="csr")
    H = (H + H.T) * 0.5
    if indefinite_h:
        d = rng.uniform(-1.0, 1.0, nx)
    else:
        d = np.abs(H).sum(axis=1).A.ravel() + sigma
    H = H + diags(d)
    J = sprandom(nc, nx, density=density_j, random_state=seed + 1, format="csr")
    # Guarantee full row rank pieces: a unit entry per constraint row.
    cols = rng.choice(nx, size=nc, replace=nc > nx)
    J = J + csr_matrix((np.ones(nc), (np.arange(nc), cols)), shape=(nc, nx))
    C = diags(np.full(nc, -delta)) if delta > 0 else None
    K = bmat([[H, J.T], [J, C]]).tocsr()
    return with_full_diagonal(K)


def grid_kkt(nx_side, seed=0, alpha=1e-4):
    """KKT of a 2-D elliptic control problem at a random interior point:
    variables (y, u) per grid node, one constraint per node."""
    rng = np.random.default_rng(seed)
    m = nx_side
    N = m * m
    h = 1.0 / (m + 1)
    T = diags([-np.ones(m - 1), 2.0 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
    Lap = ((kron(eye(m), T) + kron(T, eye(m))) / h**2).tocsr()
    y = rng.uniform(0, 1, N)
    lam = rng.normal(0, 1, N)
    Hyy = diags(h * h + 6.0 * lam * y * 1e-2)
    Huu = diags(alpha * h * h + rng.uniform(0, 1, N))
    J = bmat([[Lap + diags(3 * y**2), -eye(N)]]).tocsr()
    H = bmat([[Hyy, None], [None, Huu]]).tocsr()
    K = bmat([[H, J.T], [J, None]]).tocsr()
    return with_full_diagonal(K), 2 * N, N
