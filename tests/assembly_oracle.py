"""Oracle assembly of a ComponentProblem on the CPU: element Lagrangians from
oracle/ipm.py, gradients by complex step (exact to rounding), Hessians by a
central difference of the complex-step gradient."""

import numpy as np

from oracle import ipm as O

LINK, DI, ELL2, ELL3, CART = 0, 1, 2, 3, 4


def element_function(etype, d, alpha):
    if etype == LINK:
        return lambda z: z[1] * z[0]
    if etype == DI:
        return lambda z: O.di_interval(z, d[0], alpha)
    if etype in (ELL2, ELL3):
        dim = 2 if etype == ELL2 else 3
        return lambda z: O.elliptic(z, d[0], d[1], d[2], alpha, dim)
    if etype == CART:
        return lambda z: O.cartpole_interval(z, d[0], d[1], d[2], d[3], alpha)
    raise ValueError(etype)


def cs_grad(f, z, h=1e-30):
    g = np.zeros(len(z))
    for i in range(len(z)):
        w = z.astype(complex)
        w[i] += 1j * h
        g[i] = f(w).imag / h
    return g


def element_terms(f, z, delta=1e-4):
    g = cs_grad(f, z)
    n = len(z)
    H = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = delta
        H[:, j] = (cs_grad(f, z + e) - cs_grad(f, z - e)) / (2 * delta)
    return float(np.real(f(z.astype(complex)))), g, 0.5 * (H + H.T)


def assemble(prob, alpha, x):
    """(L, g, dense H) of alpha f + lambda^T c at x."""
    n = prob.n
    gtot = np.zeros(n)
    H = np.zeros((n, n))
    L = 0.0
    for grp in prob.groups:
        for e in range(grp.E):
            idx = grp.idx[e]
            ok = idx >= 0
            z = np.where(ok, x[np.maximum(idx, 0)], 0.0)
            f = element_function(grp.etype, grp.data[e], alpha)
            le, ge, he = element_terms(f, z)
            L += le
            ii = idx[ok]
            gtot[ii] += ge[ok]
            H[np.ix_(ii, ii)] += he[np.ix_(ok, ok)]
    return L, gtot, H
