"""Post-optimality sensitivities of a converged interior-point solve.

The reference's PostOptimization (amigo/algorithm/post_optimization.py:
23-77) differentiates model outputs with respect to model data at the
optimum: it factors the KKT matrix of the barrier problem at the final
iterate and solves it once per output (adjoint) or once per data value
(direct).  Here the same two methods work on a ComponentProblem:

- parameters are element data values, named (group, element, slot);
- outputs are entries of the solution vector (primal values or
  multipliers), and "objective" (the optimal objective value);
- the stationarity conditions F(x, p) = [grad_x L + bound terms; c(x)] = 0
  give dx/dp = -K^{-1} dF/dp with K the KKT matrix [H + Sigma, J^T; J, 0]
  at the solution (Sigma = zl / sl + zu / su, the barrier diagonal of the
  final iterate), so for outputs y = x[of]

      direct:  dy/dp = -x[of] rows of K^{-1} dF/dp      (one solve per p)
      adjoint: dy/dp = -(K^{-1} e_of)^T dF/dp           (one solve per y)

  and the optimal objective's derivative is dL/dp at the solution
  (envelope theorem).

K is assembled by the problem's Hessian kernels at the unscaled solution,
factored by the GPU LDL^T (its inertia checked: n_primal positive,
n_multiplier negative pivots) and the solves refined in FP64 against the
assembled matrix.  dF/dp comes from the symbolic derivatives of the element
Lagrangians with respect to their data (amigo_amd.codegen).
"""

import numpy as np
import scipy.sparse as sp
import torch

from .codegen import ElementDerivatives
from .sparse import CSRMat, SparseLDL


def _derivatives_of(etype):
    from . import codegen
    from .problems import components as C

    comps = C.workload_components()
    if etype in comps:
        return ElementDerivatives(comps[etype])
    for ge in codegen.registered():
        if ge.etype == etype:
            return ge.derivatives
    raise KeyError(f"no component known for element type {etype}")


class PostOptimization:
    """Sensitivities at the solution of ``opt`` (an Optimizer after a
    converged optimize())."""

    def __init__(self, opt, order="AMD", refine=2):
        prob = opt.problem
        self.prob = prob
        self.refine = refine
        res = opt._last
        if res is None:
            raise RuntimeError("optimize() has not run")
        s = float(res["obj_scale"])
        mult = np.asarray(prob.mult_ind, bool)
        x = np.asarray(opt.get_solution(), np.float64).copy()
        # multipliers and bound duals of the unscaled problem (the iterate's
        # are those of the objective scaled by obj_scale)
        x[mult] /= s
        self.x = x
        v = opt.vars
        prim = np.nonzero(~mult)[0]
        zl, zu = v.zl.double().cpu().numpy(), v.zu.double().cpu().numpy()
        sl, su = v.sl.double().cpu().numpy(), v.su.double().cpu().numpy()
        with np.errstate(divide="ignore", invalid="ignore"):
            sig = np.where(zl != 0.0, zl / sl, 0.0) + np.where(zu != 0.0, zu / su, 0.0)
        self.sigma = np.zeros(prob.n)
        self.sigma[prim] = sig / s
        self.n_primal, self.n_mult = len(prim), int(mult.sum())
        # K at the solution
        xd = torch.tensor(x, device=prob.device)
        mat = prob.create_matrix()
        prob.hessian(1.0, xd, mat)
        torch.cuda.synchronize()
        self.K = (mat.to_scipy().tocsr() + sp.diags(self.sigma)).tocsr()
        self.K.sort_indices()
        kmat = CSRMat.from_scipy(self.K)
        self.ldl = SparseLDL(kmat, order=order, groups=getattr(prob, "var_groups", None))
        self.ldl.factor()
        self.inertia = self.ldl.get_inertia()
        if self.inertia != (self.n_primal, self.n_mult):
            raise RuntimeError(f"KKT inertia at the solution is {self.inertia}, expected "
                               f"({self.n_primal}, {self.n_mult}): the solution is not a strict local minimum "
                               "with independent constraints")

    # ---- building blocks ------------------------------------------------------
    def _element_terms(self, params):
        """dF/dp (n x P) and dL/dp (P) for parameters (group, element, slot)."""
        prob, x = self.prob, self.x
        P = len(params)
        G = np.zeros((prob.n, P))
        dLdp = np.zeros(P)
        cache = {}
        for c, (gi, e, j) in enumerate(params):
            g = prob.groups[gi]
            if not (0 <= e < g.E and 0 <= j < g.data.shape[1]):
                raise IndexError(f"parameter {(gi, e, j)} out of range")
            der = cache.get(g.etype)
            if der is None:
                der = cache[g.etype] = _derivatives_of(g.etype)
            if j >= der.ndata:
                raise IndexError(f"element type {g.etype} has {der.ndata} data values")
            idx = g.idx[e]
            ok = idx >= 0
            z = np.where(ok, x[np.maximum(idx, 0)], 0.0)
            dL, dg = der.data_derivatives(z[None, :], g.data[e][None, :der.ndata], alpha=1.0)
            G[idx[ok], c] += dg[0, ok, j]
            dLdp[c] = dL[0, j]
        return G, dLdp

    def solve(self, rhs):
        """K^{-1} rhs (n x k), GPU LDL^T solves refined in FP64."""
        rhs = np.asarray(rhs, np.float64)
        one = rhs.ndim == 1
        B = rhs.reshape(self.prob.n, -1)
        X = np.asfortranarray(B.copy())
        self.ldl.solve(X)
        for _ in range(self.refine):
            R = np.asfortranarray(B - self.K @ X)
            self.ldl.solve(R)
            X = X + R
        return X[:, 0] if one else X

    # ---- sensitivities ----------------------------------------------------------
    def solution_derivatives(self, params):
        """dx*/dp (n x P): derivatives of the whole solution vector (primal
        values and multipliers) with respect to the parameters."""
        G, _ = self._element_terms(params)
        return self.solve(-G)

    def compute_post_opt_derivatives(self, of, wrt, method="adjoint"):
        """d of / d wrt (len(of) x len(wrt)).  of: solution indices and/or
        "objective"; wrt: parameters (group, element, slot); method
        "adjoint" (one solve per output) or "direct" (one per parameter)."""
        if method not in ("adjoint", "direct"):
            raise ValueError("method must be 'adjoint' or 'direct'")
        G, dLdp = self._element_terms(wrt)
        out = np.zeros((len(of), len(wrt)))
        rows = [(r, o) for r, o in enumerate(of) if not (isinstance(o, str) and o == "objective")]
        for r, o in enumerate(of):
            if isinstance(o, str):
                if o != "objective":
                    raise ValueError(f"unknown output {o}")
                out[r] = dLdp
        if rows:
            sel = np.array([int(o) for _, o in rows])
            if method == "direct":
                X = self.solve(-G)
                out[[r for r, _ in rows]] = X[sel]
            else:
                E = np.zeros((self.prob.n, len(sel)))
                E[sel, np.arange(len(sel))] = 1.0
                Y = self.solve(E)  # K symmetric
                out[[r for r, _ in rows]] = -(Y.T @ G)
        return out
