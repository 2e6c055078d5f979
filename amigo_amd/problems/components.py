"""The workloads' element types as components.

Each class writes one element of a built-in element type of
amigo_amd.problem (the hand-written kernels of csrc/assembly.hip) with the
component API, in the same input order, so that a problem can run on either
the hand-written or the generated kernels:

- ``Link``: (x, lambda), c = x.
- ``DoubleIntegratorInterval``: trapezoidal interval of q' = v, v' = u with
  the control effort (u0^2 + u1^2) dt / 4 (the reference's double-integrator
  example).
- ``EllipticNode(dim)``: one grid node of -Lap(y) + y^3 = u (constraint
  scaled by h^2) with the tracking objective.
- ``CartPoleInterval``: trapezoidal interval of the cart-pole dynamics of the
  reference's examples/trajectory/cart/cart_pole.py:53-86 (angle from
  hanging) with the objective (u0^2 + u1^2) dt / 2.

Two components have no built-in kernel: ``ChainSegment`` (a hanging chain:
segment length constraint, potential energy objective) and
``TranscendentalElement`` (every function of the expression module in one
element, for kernel checks).
"""

from .. import expressions as am
from ..component import Component


class Link(Component):
    def __init__(self):
        super().__init__()
        self.add_input("x")
        self.add_constraint("c")

    def compute(self):
        self.constraints["c"] = self.inputs["x"]


class DoubleIntegratorInterval(Component):
    def __init__(self):
        super().__init__()
        self.add_data("dt")
        for s in ("q0", "v0", "u0", "q1", "v1", "u1"):
            self.add_input(s)
        self.add_objective("effort")
        self.add_constraint("cq")
        self.add_constraint("cv")

    def compute(self):
        dt = self.data["dt"]
        h = 0.5 * dt
        q0, v0, u0 = self.inputs["q0"], self.inputs["v0"], self.inputs["u0"]
        q1, v1, u1 = self.inputs["q1"], self.inputs["v1"], self.inputs["u1"]
        self.objective["effort"] = (u0 * u0 + u1 * u1) * dt * 0.25
        self.constraints["cq"] = q1 - q0 - h * (v0 + v1)
        self.constraints["cv"] = v1 - v0 - h * (u0 + u1)


class EllipticNode(Component):
    def __init__(self, dim=2):
        super().__init__(name=f"EllipticNode{dim}d")
        self.dim = dim
        self.add_data("h")
        self.add_data("yd")
        self.add_data("a")
        self.add_input("y")
        self.add_input("y_nb", shape=2 * dim)
        self.add_input("u")
        self.add_objective("tracking")
        self.add_constraint("pde")

    def compute(self):
        h, yd, a = self.data["h"], self.data["yd"], self.data["a"]
        y, nb, u = self.inputs["y"], self.inputs["y_nb"], self.inputs["u"]
        hd = h * h if self.dim == 2 else h * h * h
        s = nb[0]
        for q in range(1, 2 * self.dim):
            s = s + nb[q]
        self.objective["tracking"] = 0.5 * hd * (y - yd) * (y - yd) + 0.5 * a * hd * u * u
        self.constraints["pde"] = (2.0 * self.dim * y - s) + h * h * (y * y * y - u)


class CartPoleInterval(Component):
    def __init__(self):
        super().__init__()
        self.add_constant("g", 9.81)
        for name in ("dt", "mc", "mp", "l"):
            self.add_data(name)
        self.add_input("s0", shape=4)  # x, th, xd, w at the start
        self.add_input("u0")
        self.add_input("s1", shape=4)
        self.add_input("u1")
        self.add_objective("effort")
        self.add_constraint("defect", shape=4)

    def _rates(self, s, u):
        g = self.constants["g"]
        mc, mp, l = self.data["mc"], self.data["mp"], self.data["l"]
        th, w = s[1], s[3]
        sn, cs = am.sin(th), am.cos(th)
        D = mc + mp * (sn * sn)
        xdd = (u + mp * (sn * (l * (w * w) + g * cs))) / D
        thdd = (-(u * cs) - (mp * l) * ((w * w) * (sn * cs)) - ((mc + mp) * g) * sn) / (l * D)
        return [s[2], s[3], xdd, thdd]

    def compute(self):
        dt = self.data["dt"]
        h = 0.5 * dt
        s0, u0, s1, u1 = self.inputs["s0"], self.inputs["u0"], self.inputs["s1"], self.inputs["u1"]
        f0, f1 = self._rates(s0, u0), self._rates(s1, u1)
        self.objective["effort"] = (u0 * u0 + u1 * u1) * dt * 0.5
        self.constraints["defect"] = [s1[r] - s0[r] - h * (f0[r] + f1[r]) for r in range(4)]


class ChainSegment(Component):
    """Segment of a hanging chain between nodes (x0, y0) and (x1, y1): fixed
    length, potential energy of its mass at its midpoint."""

    def __init__(self):
        super().__init__()
        self.add_data("length")
        self.add_data("weight")
        self.add_input("p0", shape=2)
        self.add_input("p1", shape=2)
        self.add_objective("energy")
        self.add_constraint("length_res")

    def compute(self):
        p0, p1 = self.inputs["p0"], self.inputs["p1"]
        ln, w = self.data["length"], self.data["weight"]
        dx, dy = p1[0] - p0[0], p1[1] - p0[1]
        self.objective["energy"] = 0.5 * w * (p0[1] + p1[1])
        self.constraints["length_res"] = dx * dx + dy * dy - ln * ln


class TranscendentalElement(Component):
    """Every function of the expression module in one element (kernel checks
    against the numpy evaluation of the same derivatives)."""

    def __init__(self):
        super().__init__()
        self.add_constant("k", 0.3)
        self.add_data("c", shape=2)
        self.add_input("v", shape=4)
        self.add_objective("f")
        self.add_constraint("r", shape=2)

    def compute(self):
        v, c, k = self.inputs["v"], self.data["c"], self.constants["k"]
        a, b, p, q = v[0], v[1], v[2], v[3]
        f = am.exp(k * a) * am.cosh(0.5 * b) + am.log(1.0 + p * p) + am.atan(q) * am.sin(a)
        f = f + am.sqrt(2.0 + am.cos(b) * am.cos(b)) + am.tanh(p - q) + am.sinh(0.2 * q) * c[0]
        f = f + am.atan2(a, 2.0 + q * q) + am.asinh(b) + am.log10(3.0 + a * a) + (2.0 + p * p) ** 1.5
        f = f + am.asin(0.5 * am.tanh(a)) + am.acos(0.4 * am.sin(q)) + am.atanh(0.3 * am.tanh(b))
        f = f + am.acosh(2.0 + b * b) + am.tan(0.3 * am.sin(p)) + a ** -2 * c[1] + am.fabs(b) * p
        f = f + am.min2(a, p) * q + am.max2(b, q) * a
        self.objective["f"] = f
        self.constraints["r"] = [a * b - am.exp(p) / (1.0 + q * q), p ** 3 - c[1] * am.cos(a + q)]


def workload_components():
    """The components of the built-in element types, by type number."""
    from ..problem import ELEM_CARTPOLE, ELEM_DI_INTERVAL, ELEM_ELLIPTIC2D, ELEM_ELLIPTIC3D, ELEM_LINK

    return {
        ELEM_LINK: Link(),
        ELEM_DI_INTERVAL: DoubleIntegratorInterval(),
        ELEM_ELLIPTIC2D: EllipticNode(2),
        ELEM_ELLIPTIC3D: EllipticNode(3),
        ELEM_CARTPOLE: CartPoleInterval(),
    }


def all_components():
    """The workload components and the two example components."""
    return list(workload_components().values()) + [ChainSegment(), TranscendentalElement()]


def generated_problem(prob, builtin_structure=True):
    """The same problem on generated kernels: every element group of a
    built-in type is retyped to the generated kernel of its workload
    component.  With builtin_structure the generated types take the built-in
    kernels' Hessian pattern (a superset of the symbolic one), so both
    problems have the same KKT pattern and ordering."""
    import copy

    from .. import codegen
    from ..problem import STRUCTURE, ElementGroup

    comps = workload_components()
    out = copy.copy(prob)
    groups = []
    for g in prob.groups:
        ge = codegen.element_type(comps[g.etype], build=False,
                                  structure=STRUCTURE[g.etype] if builtin_structure else None)
        groups.append(ElementGroup(ge.etype, g.idx, g.data))
    out.groups = groups
    out._build_pattern()
    out._upload()
    return out


def hanging_chain(K=40, seg=0.05, span=1.5, device="cuda"):
    """A chain of K segments of length seg hanging between (0, 0) and
    (span, 0): nodes (x_i, y_i), one length constraint per segment, the ends
    held by link constraints.  Returns (ComponentProblem, element kernel)."""
    import numpy as np

    from .. import codegen
    from ..problem import ELEM_LINK, ComponentProblem, ElementGroup

    ge = codegen.element_type(ChainSegment(), build=False)
    nn = K + 1
    nseg = 2 * nn  # first segment multiplier
    nlink = nseg + K  # first end-link multiplier
    n = nlink + 4
    s = np.arange(K)
    idx = np.stack([2 * s, 2 * s + 1, 2 * s + 2, 2 * s + 3, nseg + s], axis=1)
    data = np.tile([seg, seg], (K, 1))
    ends = np.array([0, 1, 2 * K, 2 * K + 1])
    links = np.stack([ends, nlink + np.arange(4)], axis=1)
    mult = np.zeros(n, bool)
    mult[nseg:] = True
    lower = np.full(n, -np.inf)
    upper = np.full(n, np.inf)
    lower[nseg:nlink] = upper[nseg:nlink] = 0.0
    lower[nlink:] = upper[nlink:] = [0.0, 0.0, span, 0.0]
    x0 = np.zeros(n)
    t = np.linspace(0.0, 1.0, nn)
    x0[0:2 * nn:2] = span * t
    x0[1:2 * nn:2] = -0.4 * span * np.sin(np.pi * t)
    prob = ComponentProblem(n, [ElementGroup(ge.etype, idx, data),
                                ElementGroup(ELEM_LINK, links, np.zeros((4, 1)))],
                            mult, lower, upper, x0, device=device, name=f"hanging_chain_{K}")
    return prob, ge
