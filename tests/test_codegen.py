"""Component expressions and generated element kernels (host side).

The symbolic derivatives that the generated kernels compute are checked on
the CPU: against complex-step derivatives of the same expressions (exact to
rounding for analytic functions), against finite differences where the
functions have kinks (fabs, min2, max2), and, for the workload components,
against the assembly oracle's complex-step derivatives of the independently
written element Lagrangians (tests/assembly_oracle.py).  A planted mistake
in a derivative rule fails the complex-step check.  The generated source is
deterministic and compiles with hipcc; the C registration entry point
rejects bad arguments without touching a GPU.
"""

import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from amigo_amd import codegen
from amigo_amd import expressions as ex
from amigo_amd.component import Component
from amigo_amd.problems import components as C
from tests import assembly_oracle as AO

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _cs_grad(node_fn, z, d, alpha=1.0, h=1e-30):
    """Complex-step gradient of a scalar node function at z (one element)."""
    n = len(z)
    g = np.zeros(n)
    for i in range(n):
        w = z.astype(complex)
        w[i] += 1j * h
        g[i] = ex.evaluate([node_fn], w, d, dtype=complex, alpha=alpha)[0].imag / h
    return g


def _smooth_component():
    class Smooth(Component):
        def __init__(self):
            super().__init__()
            self.add_data("c")
            self.add_input("v", shape=3)
            self.add_objective("f")
            self.add_constraint("r", shape=2)

        def compute(self):
            a, b, p = self.inputs["v"][0], self.inputs["v"][1], self.inputs["v"][2]
            c = self.data["c"]
            f = ex.exp(0.3 * a) * ex.cosh(b) + ex.log(2.0 + p * p) + ex.atan(a * b) * ex.sin(p)
            f = f + ex.sqrt(3.0 + ex.cos(b)) + ex.tanh(a - p) + ex.sinh(0.5 * c * b)
            f = f + ex.atan2(a, 2.0 + p * p) + ex.asinh(b) + ex.log10(4.0 + a * a)
            f = f + (1.5 + p * p) ** 1.5 + ex.asin(0.4 * ex.tanh(a)) + ex.acos(0.3 * ex.sin(p))
            f = f + ex.atanh(0.2 * ex.tanh(b)) + ex.acosh(2.0 + b * b) + ex.tan(0.2 * p)
            f = f + a ** -2 + (2.0 + a * a) ** b
            self.objective["f"] = f
            self.constraints["r"] = [a * b - ex.exp(p) / (1.0 + a * a), p ** 3 - c * ex.cos(a + b)]

    return Smooth()


def test_gradient_and_hessian_equal_complex_step():
    """g = complex-step gradient of L, H = complex-step Jacobian of g, for an
    element using every analytic function (and a general power)."""
    comp = _smooth_component()
    der = codegen.ElementDerivatives(comp)
    rng = np.random.default_rng(0)
    for _ in range(5):
        z = rng.uniform(0.3, 1.2, der.nc) * rng.choice([-1.0, 1.0], der.nc)
        z[0] = abs(z[0]) + 0.2  # a ** -2 and a ** b: away from 0
        d = np.array([rng.uniform(0.5, 1.5)])
        L, g, H = der.evaluate(z, d, alpha=0.7)
        gcs = _cs_grad(der.L, z, d, alpha=0.7)
        assert np.allclose(g[0], gcs, rtol=1e-13, atol=1e-13)
        for i in range(der.nc):
            hcs = _cs_grad(der.g[i], z, d, alpha=0.7)
            assert np.allclose(H[0, i], hcs, rtol=1e-12, atol=1e-12), i
        assert np.array_equal(H[0], H[0].T)


def test_kinked_functions_against_central_differences():
    class Kinks(Component):
        def __init__(self):
            super().__init__()
            self.add_input("v", shape=3)
            self.add_objective("f")

        def compute(self):
            a, b, c = (self.inputs["v"][q] for q in range(3))
            self.objective["f"] = ex.fabs(a) * b + ex.min2(a, c) * b * b + ex.max2(b, c) * a * c

    der = codegen.ElementDerivatives(Kinks())
    z = np.array([-0.7, 0.4, 0.9])
    d = np.zeros(1)
    L, g, H = der.evaluate(z, d)
    eps = 1e-6
    fd = np.zeros(3)
    for i in range(3):
        e = np.zeros(3)
        e[i] = eps
        fd[i] = (ex.evaluate([der.L], z + e, d)[0] - ex.evaluate([der.L], z - e, d)[0]) / (2 * eps)
    assert np.allclose(g[0], fd, rtol=1e-8, atol=1e-8)
    # min2(a, c) = a here (a < c): d/da = b^2, max2(b, c) = c: d/dc = a c + ...
    assert np.isclose(g[0][0], -z[1] + z[1] ** 2 + z[2] * z[2])


def test_planted_derivative_mistake_fails(monkeypatch):
    """A sign mistake in the derivative of sin is caught by the complex-step
    check above."""
    orig = ex._unary_slope

    def wrong(op, a, f):
        if op == "sin":
            return ex.unary("neg", ex.unary("cos", a))
        return orig(op, a, f)

    monkeypatch.setattr(ex, "_unary_slope", wrong)

    class S(Component):
        def __init__(self):
            super().__init__()
            self.add_input("v", shape=2)
            self.add_objective("f")

        def compute(self):
            self.objective["f"] = ex.sin(self.inputs["v"][0]) * self.inputs["v"][1]

    der = codegen.ElementDerivatives(S())
    z, d = np.array([0.4, 1.3]), np.zeros(1)
    _, g, _ = der.evaluate(z, d)
    assert not np.allclose(g[0], _cs_grad(der.L, z, d), rtol=1e-10)


@pytest.mark.parametrize("etype", sorted(C.workload_components()))
def test_workload_components_match_the_assembly_oracle(etype):
    """The workload components' derivatives equal the complex-step
    derivatives of the oracle's element Lagrangians (written independently of
    the components); their Hessian structure is contained in the built-in
    kernels' pattern."""
    from amigo_amd.problem import STRUCTURE

    comp = C.workload_components()[etype]
    der = codegen.ElementDerivatives(comp)
    rng = np.random.default_rng(etype)
    for _ in range(3):
        z = rng.normal(size=der.nc)
        d = np.abs(rng.normal(size=max(der.ndata, 1))) + 0.2
        L, g, H = der.evaluate(z, d, alpha=0.7)
        f = AO.element_function(etype, d, 0.7)
        le, ge, he = AO.element_terms(f, z)
        assert abs(L[0] - le) <= 1e-13 * max(1.0, abs(le))
        assert np.allclose(g[0], ge, rtol=1e-12, atol=1e-12)
        # the oracle's Hessian is a central difference of complex-step gradients
        assert np.allclose(H[0], he, rtol=1e-6, atol=1e-6)
    assert not np.any(der.structure & ~STRUCTURE[etype])


def test_generated_source_is_independent_of_build_order():
    """The source text (and so the kernel library's name) does not depend on
    what else the process built before."""
    code = (
        "import sys; sys.path.insert(0, {root!r});"
        "from amigo_amd import codegen; from amigo_amd.problems import components as C;"
        "{pre}"
        "import hashlib; print(hashlib.sha1(codegen.generate_source(C.CartPoleInterval())[0].encode()).hexdigest())"
    )
    outs = []
    for pre in ("", "[codegen.generate_source(c) for c in C.all_components()];"):
        r = subprocess.run([sys.executable, "-c", code.format(root=ROOT, pre=pre)],
                           capture_output=True, text=True, check=True)
        outs.append(r.stdout.strip())
    assert outs[0] == outs[1]


def test_generated_source_compiles():
    with tempfile.TemporaryDirectory() as tmp:
        src, der = codegen.generate_source(C.ChainSegment())
        out = codegen.compile_source(src, os.path.join(tmp, "chain.so"))
        assert os.path.getsize(out) > 0
        assert "amigo_gen_eval" in src and f"constexpr int NC = {der.nc};" in src


def test_register_rejects_bad_arguments():
    import ctypes

    from amigo_amd import _lib

    L = _lib.lib()
    assert L.amigo_elem_register(99, 4, ctypes.c_void_p(1)) == -1
    assert L.amigo_elem_register(100 + 256, 4, ctypes.c_void_p(1)) == -1
    assert L.amigo_elem_register(100, 4, None) == -1
    assert L.amigo_elem_register(100, 0, ctypes.c_void_p(1)) == -1


def test_component_declaration_errors():
    class Bad(Component):
        def __init__(self):
            super().__init__()
            self.add_input("x", shape=2)
            self.add_constraint("c", shape=2)

        def compute(self):
            self.constraints["c"] = [self.inputs["x"][0]]  # one expression for two

    with pytest.raises(ValueError):
        Bad().lagrangian()

    class Missing(Component):
        def __init__(self):
            super().__init__()
            self.add_input("x")
            self.add_constraint("c")

    with pytest.raises(ValueError):
        Missing().lagrangian()
    comp = C.CartPoleInterval()
    assert comp.num_inputs() == 14 and comp.multiplier_slots() == [10, 11, 12, 13]
    assert comp.input_slot("s1", 2) == 7 and comp.input_slot("defect", 3) == 13
