"""Writes tests/golden/ldl_fixtures.npz from the reference's own SparseLDL
and BlockAMD, compiled from the reference headers by oracle/Makefile
(oracle/_ref/libamigo_ref.so).  Run in a container that has the reference
tree:

    python tests/golden/make_golden.py

Cases: the seven hand matrices and the seeded Poisson matrices of the
reference's tests/unit/test_ldl.py, plus seeded KKT matrices (with and
without delayed pivots) and a small elliptic-control KKT.  Stored per case
and ordering: the CSR matrix, AMD/NATURAL perm and iperm, etree / postorder
/ column counts, the symbolic plan, inertia, per-node pivot and delay
counts, and the solution of A x = b for b = 1..n.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402
from tests import matrices as M  # noqa: E402


def cases():
    out = []
    for i in range(7):
        A, _ = M.small(i)
        out.append((f"small{i}", A, 0.4))
    for n in (10, 20):
        out.append((f"poisson{n}", M.with_full_diagonal(M.poisson_2d(n)), 0.01))
    out.append(("kkt", M.kkt(200, 80, 1), 0.01))
    out.append(("kkt_indef", M.kkt(150, 60, 2, indefinite_h=True), 0.01))
    out.append(("kkt_delay", M.kkt(150, 60, 3, indefinite_h=True, sigma=0.0), 0.1))
    A, _, _ = M.grid_kkt(8)
    out.append(("grid8", A, 0.01))
    return out


def main():
    ref.build_ref()
    data = {}
    for name, A, ustab in cases():
        A = A.tocsr()
        A.sort_indices()
        n = A.shape[0]
        data[f"{name}/rowp"] = A.indptr.astype(np.int32)
        data[f"{name}/cols"] = A.indices.astype(np.int32)
        data[f"{name}/vals"] = A.data.astype(np.float64)
        data[f"{name}/ustab"] = np.array(ustab)
        for order in ("NATURAL", "AMD"):
            key = f"{name}/{order}"
            perm, iperm = ref.order(A.indptr, A.indices, order)
            R = ref.RefLDL(A.indptr, A.indices, ustab=ustab, order=order)
            if order == "AMD":
                data[f"{key}/perm"] = perm
                data[f"{key}/iperm"] = iperm
                par, post, cc = R.etree(perm, iperm)
            else:
                par, post, cc = R.etree()
            data[f"{key}/parent"] = par
            data[f"{key}/post"] = post
            data[f"{key}/colcount"] = cc
            for k, v in R.symbolic().items():
                data[f"{key}/{k}"] = v
            R.factor(A.data)
            data[f"{key}/inertia"] = np.array(R.inertia(), np.int32)
            npiv, ndel = R.pivot_counts()
            data[f"{key}/npiv"] = npiv
            data[f"{key}/ndel"] = ndel
            data[f"{key}/x"] = R.solve(np.arange(1, n + 1, dtype=np.float64))
    np.savez_compressed(os.path.join(HERE, "ldl_fixtures.npz"), **data)
    print(f"wrote {len(data)} arrays")


if __name__ == "__main__":
    main()
