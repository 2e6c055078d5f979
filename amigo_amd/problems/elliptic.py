"""Semilinear elliptic optimal control on the unit square / cube.

    min  1/2 h^d ||y - y_d||^2 + alpha/2 h^d ||u||^2
    s.t. -Lap(y) + y^3 = u   (5-point / 7-point finite differences,
                              homogeneous Dirichlet boundary; each node's
                              equation is multiplied by h^2)
         0 <= u <= 50

on an m^d interior grid, h = 1 / (m + 1).  The target y_d is a seeded sum of
Fourier modes prod_i sin(k_i pi x_i) with amplitudes N(0,1) / |k|^2: in 2-D
all 16 modes k in {1..4}^2, in 3-D 24 modes drawn from {1..3}^3.

Variable layout: per grid node (row-major) [y_i, u_i, lambda_i].
"""

import itertools

import numpy as np

from ..problem import ELEM_ELLIPTIC2D, ELEM_ELLIPTIC3D, ComponentProblem, ElementGroup


def target_field(m, dim, seed, n_modes=None):
    rng = np.random.default_rng(seed)
    h = 1.0 / (m + 1)
    grid = np.arange(1, m + 1) * h
    if dim == 2:
        modes = list(itertools.product(range(1, 5), repeat=2))
    else:
        cand = list(itertools.product(range(1, 4), repeat=3))
        pick = rng.choice(len(cand), size=n_modes or 24, replace=False)
        modes = [cand[p] for p in sorted(pick)]
    amps = rng.normal(0.0, 1.0, len(modes))
    coords = np.meshgrid(*([grid] * dim), indexing="ij")
    yd = np.zeros([m] * dim)
    for a, k in zip(amps, modes):
        term = np.ones([m] * dim)
        for ax in range(dim):
            term = term * np.sin(k[ax] * np.pi * coords[ax])
        yd += a / float(np.dot(k, k)) * term
    return yd.reshape(-1)


class EllipticControl(ComponentProblem):
    def __init__(self, m=256, dim=2, alpha=1e-4, umax=50.0, seed=0, n_modes=None, device="cuda"):
        self.m, self.dim, self.alpha = m, dim, alpha
        N = m**dim
        h = 1.0 / (m + 1)
        node = np.arange(N)
        self.iy = 3 * node
        self.iu = 3 * node + 1
        self.il = 3 * node + 2
        n = 3 * N
        sub = np.array(np.unravel_index(node, [m] * dim)).T  # (N, dim)
        nb_cols = []
        for ax in range(dim):
            for step in (-1, 1):
                s = sub.copy()
                s[:, ax] += step
                ok = (s[:, ax] >= 0) & (s[:, ax] < m)
                nbr = np.ravel_multi_index(np.clip(s, 0, m - 1).T, [m] * dim)
                nb_cols.append(np.where(ok, self.iy[nbr], -1))
        idx = np.column_stack([self.iy] + nb_cols + [self.iu, self.il])
        self.yd = target_field(m, dim, seed, n_modes)
        data = np.column_stack([np.full(N, h), self.yd, np.full(N, alpha)])
        etype = ELEM_ELLIPTIC2D if dim == 2 else ELEM_ELLIPTIC3D
        g = ElementGroup(etype, idx, data)
        mult = np.zeros(n, bool)
        mult[self.il] = True
        lower = np.full(n, -np.inf)
        upper = np.full(n, np.inf)
        lower[self.iu] = 0.0
        upper[self.iu] = umax
        lower[self.il] = upper[self.il] = 0.0
        x0 = np.zeros(n)
        super().__init__(n, [g], mult, lower, upper, x0, device=device,
                         name=f"elliptic{dim}d_m{m}", var_groups=np.repeat(node, 3))

    def fields(self, opt):
        x = opt.get_solution()
        return x[self.iy], x[self.iu], x[self.il]


def build_2d(m=256, seed=0, device="cuda"):
    return EllipticControl(m=m, dim=2, seed=seed, device=device)


def build_3d(m=48, seed=0, device="cuda"):
    return EllipticControl(m=m, dim=3, seed=seed, device=device)
