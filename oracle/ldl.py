"""Numpy restatement of the reference's numeric multifrontal LDL^T
(include/sparse_ldl.h).  ORACLE infrastructure: only tests/, the smoke check
and bench.py's cpu_baseline leg use it; the product never does.

Given the reference's symbolic plan (node_ptr, node_to_var, children,
contribution rows, invperm) and the CSR values, ``factor`` builds every front
in postorder, factors its fully summed block with the reference's pivot
rules and passes delayed pivots and the Schur complement to the parent:

- front variables: supernode variables, then the children's delayed pivots
  in child order, then the static contribution rows (sparse_ldl.h:957-987,
  :610-627);
- assembly: original entries with invperm[i] >= invperm[j] of the supernode
  columns, plus the children's contribution blocks (sparse_ldl.h:1004-1156);
- non-root fronts (fully summed < front size): threshold 1x1 pivot, the
  row-imax 1x1 pivot, the Duff 2x2 test, else delay by a symmetric swap to
  the end of the candidates (sparse_ldl.h:1298-1535).  Restated unblocked:
  the reference's 64-column panels take the same decisions on the same
  (lazily updated) values;
- root fronts (fully summed == front size): LAPACK dsytrf("L"), the
  Bunch-Kaufman partial pivoting of Bunch & Kaufman (Math. Comp. 31, 1977)
  with alpha = (1 + sqrt(17)) / 8 as LAPACK's dsytf2 states it
  (sparse_ldl.h:1208-1230);
- inertia: 1x1 by the sign of d (zero counted positive), 2x2 by det >= 0
  -> both of sign(d11), else one of each (sparse_ldl.h:1917-1969);
- Cholesky solver type: no pivoting, every pivot 1x1 (sparse_ldl.h:1992-2026
  uses potrf/trsm/syrk, which on a definite matrix takes the same pivots).

The factor is kept as an explicit permuted L, D so that ``solve`` is a dense
triangular solve; that restates the mathematics of sparse_ldl.h:1807-1908,
not its node loop.  Pinned by tests/test_oracle_ldl.py against the compiled
reference (oracle/_ref) and the committed fixtures, against eigenvalue
inertia (Sylvester) and against dense solves.
"""

import numpy as np

BK_ALPHA = (1.0 + np.sqrt(17.0)) / 8.0


def _imax(v):
    """First index of the largest |v| (idamax, 0-based)."""
    return int(np.argmax(np.abs(v)))


def _swap(F, vars_, a, b):
    """Symmetric swap of rows/columns a, b of a full symmetric front
    (sparse_ldl.h:1252-1274 on the lower triangle)."""
    if a == b:
        return
    F[[a, b], :] = F[[b, a], :]
    F[:, [a, b]] = F[:, [b, a]]
    vars_[a], vars_[b] = vars_[b], vars_[a]


def _eliminate(F, k, step):
    """Right-looking update of the trailing block after pivot(s) at k."""
    if step == 1:
        d = F[k, k]
        l = F[k + 1:, k] / d
        F[k + 1:, k + 1:] -= np.outer(F[k + 1:, k], l)
        F[k + 1:, k] = l
        F[k, k + 1:] = l
    else:
        D = F[k:k + 2, k:k + 2].copy()
        Dinv = np.array([[D[1, 1], -D[1, 0]], [-D[1, 0], D[0, 0]]]) / (D[0, 0] * D[1, 1] - D[1, 0] ** 2)
        B = F[k + 2:, k:k + 2].copy()
        L = B @ Dinv
        F[k + 2:, k + 2:] -= L @ B.T
        F[k + 2:, k:k + 2] = L
        F[k:k + 2, k + 2:] = L.T


def factor_front(F, fs, ustab, pivot_tol):
    """Pivots of a non-root front (sparse_ldl.h:1298-1535).  F is the full
    symmetric front (modified in place); returns (order, steps, npiv):
    the front permutation, the pivot step sizes and the number of pivots."""
    m = F.shape[0]
    order = list(range(m))
    steps = []
    nc = fs
    k = 0
    while k < nc:
        abswkk = abs(F[k, k])
        abswik, imax = 0.0, 0
        if k + 1 < nc:
            imax = k + 1 + _imax(F[k + 1:nc, k])
            abswik = abs(F[imax, k])
        ismax = nc + _imax(F[nc:m, k])
        colmax = max(abswik, abs(F[ismax, k]))
        step, kpiv = 0, -1
        if abswkk >= pivot_tol and abswkk >= ustab * colmax:
            step, kpiv = 1, k
        elif k + 1 < nc:
            col = F[:, imax]
            rowmax = abs(col[k + _imax(col[k:imax])])
            if imax + 1 < m:
                rowmax = max(rowmax, abs(col[imax + 1 + _imax(col[imax + 1:m])]))
            abswii = abs(F[imax, imax])
            if abswii >= pivot_tol and abswii >= ustab * rowmax:
                step, kpiv = 1, imax
            else:
                det = abs(F[k, k] * F[imax, imax] - F[imax, k] ** 2)
                if det >= pivot_tol:
                    delta = det / max(abswii + abswik, abswkk + abswik)
                    if delta >= ustab * colmax and delta >= ustab * rowmax:
                        step, kpiv = 2, imax
        if step == 0:
            if k < nc - 1:
                _swap(F, order, k, nc - 1)
            nc -= 1
            continue
        ksrc = k + step - 1
        _swap(F, order, ksrc, kpiv)
        _eliminate(F, k, step)
        steps.append(step)
        k += step
    return order, steps, nc


def factor_front_nopivot(F, fs):
    """Cholesky solver type: all fs pivots 1x1 in order."""
    for k in range(fs):
        _eliminate(F, k, 1)
    return list(range(F.shape[0])), [1] * fs, fs


def bunch_kaufman(F):
    """LAPACK dsytf2("L") decisions on the full symmetric root front (in
    place); returns (order, steps).  A zero column is a 1x1 pivot of 0."""
    n = F.shape[0]
    order = list(range(n))
    steps = []
    k = 0
    while k < n:
        absakk = abs(F[k, k])
        if k + 1 < n:
            imax = k + 1 + _imax(F[k + 1:, k])
            colmax = abs(F[imax, k])
        else:
            imax, colmax = k, 0.0
        if max(absakk, colmax) == 0.0:
            steps.append(1)
            k += 1
            continue
        if absakk >= BK_ALPHA * colmax:
            kp, kstep = k, 1
        else:
            row = np.concatenate([F[imax, k:imax], F[imax + 1:, imax]])
            rowmax = float(np.max(np.abs(row))) if row.size else 0.0
            if absakk >= BK_ALPHA * colmax * (colmax / rowmax):
                kp, kstep = k, 1
            elif abs(F[imax, imax]) >= BK_ALPHA * rowmax:
                kp, kstep = imax, 1
            else:
                kp, kstep = imax, 2
        kk = k + kstep - 1
        _swap(F, order, kk, kp)
        _eliminate(F, k, kstep)
        steps.append(kstep)
        k += kstep
    return order, steps


def _block_inertia(F, k, step):
    if step == 1:
        return (1, 0) if F[k, k] >= 0.0 else (0, 1)
    d11, d21, d22 = F[k, k], F[k + 1, k], F[k + 1, k + 1]
    if d11 * d22 - d21 * d21 >= 0.0:
        return (2, 0) if d11 >= 0.0 else (0, 2)
    return (1, 1)


class OracleLDL:
    """Multifrontal LDL^T on a given symbolic plan.

    sym: dict with node_ptr, node_to_var, children_ptr, children,
    contrib_ptr, contrib_rows, invperm (the reference's layout,
    sparse_ldl.h:3252-3268)."""

    def __init__(self, n, rowp, cols, sym, ustab=0.01, pivot_tol=1e-14, cholesky=False):
        self.n = n
        self.rowp = np.asarray(rowp)
        self.cols = np.asarray(cols)
        self.sym = {k: np.asarray(v) for k, v in sym.items()}
        self.ustab = min(max(ustab, 0.0), 0.5)
        self.pivot_tol = pivot_tol
        self.cholesky = cholesky

    def factor(self, vals):
        s = self.sym
        n = self.n
        K = len(s["node_ptr"]) - 1
        inv = s["invperm"]
        vals = np.asarray(vals, float)
        kp, kids = s["children_ptr"], s["children"]
        contrib = {}
        self.npiv = np.zeros(K, int)
        self.ndel = np.zeros(K, int)
        self.inertia = [0, 0]
        # Global factor: pivot sequence, L (n x n, unit lower in pivot
        # order) and the block diagonal D.
        pos = {}
        seq = []
        Lcols = []
        Dblocks = []
        for ks in range(K):
            sv = list(s["node_to_var"][s["node_ptr"][ks]:s["node_ptr"][ks + 1]])
            front = list(sv)
            for c in kids[kp[ks]:kp[ks + 1]]:
                cv, nd, _ = contrib[c]
                front.extend(cv[:nd])
            fs = len(front)
            front.extend(s["contrib_rows"][s["contrib_ptr"][ks]:s["contrib_ptr"][ks + 1]])
            m = len(front)
            idx = {v: i for i, v in enumerate(front)}
            F = np.zeros((m, m))
            for j, var in enumerate(sv):
                for p in range(self.rowp[var], self.rowp[var + 1]):
                    i = self.cols[p]
                    if i in idx and inv[i] >= inv[var]:
                        F[idx[i], j] += vals[p]
            # symmetric fill of the original entries (lower -> upper)
            F = np.tril(F) + np.tril(F, -1).T
            for c in kids[kp[ks]:kp[ks + 1]]:
                cv, nd, C = contrib.pop(c)
                ii = np.array([idx[v] for v in cv])
                F[np.ix_(ii, ii)] += C
            is_root = fs == m and not self.cholesky
            if self.cholesky:
                order, steps, npv = factor_front_nopivot(F, fs)
            elif is_root:
                order, steps = bunch_kaufman(F)
                npv = m
            else:
                order, steps, npv = factor_front(F, fs, self.ustab, self.pivot_tol)
            fv = [front[o] for o in order]
            self.npiv[ks] = npv
            self.ndel[ks] = fs - npv
            k = 0
            for st in steps:
                p, q = _block_inertia(F, k, st)
                self.inertia[0] += p
                self.inertia[1] += q
                Dblocks.append(F[k:k + st, k:k + st].copy())
                for t in range(st):
                    pos[fv[k + t]] = len(seq)
                    seq.append(fv[k + t])
                    Lcols.append((fv[k + st:], F[k + st:, k + t].copy()))
                k += st
            if npv < m:
                contrib[ks] = (fv[npv:], fs - npv, F[npv:, npv:].copy())
        if len(seq) != n:
            raise RuntimeError("oracle factor: not every variable was pivoted")
        L = np.eye(n)
        for j, (rows, col) in enumerate(Lcols):
            for r, v in zip(rows, col):
                L[pos[r], j] = v
        D = np.zeros((n, n))
        k = 0
        for B in Dblocks:
            st = B.shape[0]
            D[k:k + st, k:k + st] = B
            k += st
        self.seq = np.array(seq)
        self.L, self.D = L, D
        return 0

    def get_inertia(self):
        return tuple(self.inertia)

    def solve(self, b):
        """x = A^{-1} b with P^T A P = L D L^T (b: (n,) or (n, nrhs))."""
        b = np.asarray(b, float)
        y = b[self.seq]
        z = np.linalg.solve(self.L, y)
        z = np.linalg.solve(self.D, z)
        w = np.linalg.solve(self.L.T, z)
        x = np.empty_like(w)
        x[self.seq] = w
        return x
