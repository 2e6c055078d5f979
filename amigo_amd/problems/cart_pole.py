"""Cart-pole swing-up, trapezoidal collocation.

N intervals on [0, T]; node k has states (x, th, xd, w) and force u,
|u| <= umax.  The pole angle th is measured from hanging (th = pi upright)
as in the reference's examples/trajectory/cart/cart_pole.py:53-86.
Constraints: the trapezoidal defects of (x, th, xd, w) on every interval and
the boundary conditions: hanging at rest at x = 0 at t = 0, upright at rest
at x = 0 at t = T.  Objective sum (u_k^2 + u_{k+1}^2) dt / 2
(cart_pole.py:347-349).  Initial guess: th linear 0 -> pi, everything else 0.

Variable layout (node-interleaved): [x, th, xd, w, u, l1, l2, l3, l4] for
node k < N (l = defect multipliers of interval k), then [x, th, xd, w, u]
for node N, then the eight boundary multipliers.
"""

import numpy as np

from ..problem import ELEM_CARTPOLE, ELEM_LINK, ComponentProblem, ElementGroup


class CartPole(ComponentProblem):
    def __init__(self, N=128, T=2.0, mc=1.0, mp=0.3, l=0.5, umax=20.0, device="cuda"):
        self.N, self.T = N, T
        dt = T / N
        base = lambda k: 9 * k
        nodes = np.array([base(k) for k in range(N + 1)])
        self.ix, self.ith, self.ixd, self.iw, self.iu = (nodes + j for j in range(5))
        ilam = 9 * np.arange(N)[:, None] + 5 + np.arange(4)[None, :]
        nbase = 9 * N + 5
        ibc = nbase + np.arange(8)
        n = nbase + 8
        k = np.arange(N)
        st = lambda kk: np.stack([self.ix[kk], self.ith[kk], self.ixd[kk], self.iw[kk], self.iu[kk]], axis=1)
        idx = np.concatenate([st(k), st(k + 1), ilam], axis=1)
        data = np.tile([dt, mc, mp, l], (N, 1))
        g1 = ElementGroup(ELEM_CARTPOLE, idx, data)
        bvars = np.array([self.ix[0], self.ith[0], self.ixd[0], self.iw[0],
                          self.ix[N], self.ith[N], self.ixd[N], self.iw[N]])
        g2 = ElementGroup(ELEM_LINK, np.stack([bvars, ibc], axis=1), np.zeros((8, 1)))
        mult = np.zeros(n, bool)
        mult[ilam.ravel()] = True
        mult[ibc] = True
        lower = np.full(n, -np.inf)
        upper = np.full(n, np.inf)
        lower[self.iu] = -umax
        upper[self.iu] = umax
        lower[mult] = upper[mult] = 0.0
        lower[ibc] = upper[ibc] = [0.0, 0.0, 0.0, 0.0, 0.0, np.pi, 0.0, 0.0]
        x0 = np.zeros(n)
        x0[self.ith] = np.linspace(0.0, np.pi, N + 1)
        groups = np.zeros(n, np.int64)
        groups[: 9 * N] = np.repeat(np.arange(N), 9)
        groups[9 * N: 9 * N + 5] = N
        groups[ibc] = [0, 0, 0, 0, N, N, N, N]
        super().__init__(n, [g1, g2], mult, lower, upper, x0, device=device,
                         name=f"cart_pole_N{N}", var_groups=groups)

    def controls(self, opt):
        return opt.get_solution()[self.iu]


def build(N=128, T=2.0, mc=1.0, mp=0.3, l=0.5, device="cuda"):
    return CartPole(N=N, T=T, mc=mc, mp=mp, l=l, device=device)


def seeded(seed, N=None, nmin=16, nmax=512, T=2.0, device="cuda"):
    """One problem of the seeded batch: N ~ U{nmin..nmax} unless given,
    mc ~ U(0.5, 2), mp ~ U(0.1, 0.5), l ~ U(0.3, 1.0)."""
    rng = np.random.default_rng(seed)
    NN = int(rng.integers(nmin, nmax + 1))
    mc, mp, l = rng.uniform(0.5, 2.0), rng.uniform(0.1, 0.5), rng.uniform(0.3, 1.0)
    return CartPole(N=N or NN, T=T, mc=mc, mp=mp, l=l, device=device)
