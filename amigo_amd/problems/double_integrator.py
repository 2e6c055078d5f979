"""Double-integrator minimum-effort trajectory, trapezoidal collocation.

N intervals on [0, T]; every node k has states q_k, v_k and control u_k.
Constraints: defects q' = v, v' = u on every interval, boundary conditions
q(0) = 0, v(0) = 0, q(T) = 1, v(T) = 0.  Objective sum (u_k^2 + u_{k+1}^2)
dt / 4 with |u| <= 10.  Initial guess: q linear 0 -> 1, v = u = 0.
For T = 1 the continuous optimum is u(t) = 6 - 12 t.

Variable layout (node-interleaved): [q_k, v_k, u_k, lq_k, lv_k] for k < N,
then [q_N, v_N, u_N], then the four boundary multipliers.
"""

import numpy as np

from ..problem import ELEM_DI_INTERVAL, ELEM_LINK, ComponentProblem, ElementGroup


class DoubleIntegrator(ComponentProblem):
    def __init__(self, N=256, T=1.0, umax=10.0, device="cuda"):
        self.N, self.T = N, T
        dt = T / N
        n_nodes = N + 1
        q = lambda k: 5 * k if k < N else 5 * N
        self.iq = np.array([q(k) for k in range(n_nodes)])
        self.iv = self.iq + 1
        self.iu = self.iq + 2
        ilq = 5 * np.arange(N) + 3
        ilv = 5 * np.arange(N) + 4
        nbase = 5 * N + 3
        ibc = nbase + np.arange(4)
        n = nbase + 4
        k = np.arange(N)
        idx = np.stack([self.iq[k], self.iv[k], self.iu[k], self.iq[k + 1], self.iv[k + 1], self.iu[k + 1],
                        ilq, ilv], axis=1)
        g1 = ElementGroup(ELEM_DI_INTERVAL, idx, np.full((N, 1), dt))
        bc_vars = np.array([self.iq[0], self.iv[0], self.iq[N], self.iv[N]])
        g2 = ElementGroup(ELEM_LINK, np.stack([bc_vars, ibc], axis=1), np.zeros((4, 1)))
        mult = np.zeros(n, bool)
        mult[ilq] = mult[ilv] = True
        mult[ibc] = True
        lower = np.full(n, -np.inf)
        upper = np.full(n, np.inf)
        lower[self.iu] = -umax
        upper[self.iu] = umax
        lower[mult] = upper[mult] = 0.0
        lower[ibc] = upper[ibc] = [0.0, 0.0, 1.0, 0.0]
        x0 = np.zeros(n)
        x0[self.iq] = np.linspace(0.0, 1.0, n_nodes)
        groups = np.zeros(n, np.int64)
        groups[: 5 * N] = np.repeat(np.arange(N), 5)
        groups[5 * N: 5 * N + 3] = N
        groups[ibc] = [0, 0, N, N]
        super().__init__(n, [g1, g2], mult, lower, upper, x0, device=device,
                         name=f"double_integrator_N{N}", var_groups=groups)

    def controls(self, opt):
        return opt.get_solution()[self.iu]

    def states(self, opt):
        x = opt.get_solution()
        return x[self.iq], x[self.iv]


def build(N=256, T=1.0, device="cuda"):
    return DoubleIntegrator(N=N, T=T, device=device)
