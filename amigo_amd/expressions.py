"""Symbolic expressions of component inputs, data and constants.

The component modelling API (``amigo_amd.component``) builds the objective
and constraint expressions of one element from its inputs with ordinary
Python arithmetic, as the reference's ``Expr`` does (amigo/expressions.py:
1-386, amigo/unary_operations.py).  Here the expressions are a hash-consed
DAG: structurally equal subexpressions are one node, so derivatives built
from them share their common parts, and constant operands fold at
construction.  ``amigo_amd.codegen`` differentiates the element Lagrangian
symbolically (first and second derivatives) and writes the DAG out as
straight-line HIP code.

Leaves: ``const`` (a float), ``in`` (element input q), ``data`` (element
data value q) and ``alpha`` (the objective factor of the Lagrangian).
Operations: unary functions, + - * /, powers, atan2, min2 / max2 and a
comparison select used by the derivatives of min2 / max2.
"""

import math

import numpy as np

# ---- nodes -------------------------------------------------------------------

_TABLE = {}


class Node:
    """One interned DAG node: op, argument nodes, and a value for leaves."""

    __slots__ = ("op", "args", "val", "id")

    def __init__(self, op, args, val, nid):
        self.op = op
        self.args = args
        self.val = val
        self.id = nid

    def __repr__(self):
        return f"Node({self.op}, {[a.id for a in self.args]}, {self.val})"


def _node(op, args=(), val=None):
    key = (op, tuple(a.id for a in args), val)
    n = _TABLE.get(key)
    if n is None:
        n = Node(op, tuple(args), val, len(_TABLE))
        _TABLE[key] = n
    return n


def _const_node(v):
    v = float(v)
    if v == 0.0:
        v = 0.0  # one zero node (-0.0 folds to 0.0)
    return _node("const", (), v)


UNARY = ("neg", "sin", "cos", "tan", "asin", "acos", "atan", "sinh", "cosh", "tanh",
         "asinh", "acosh", "atanh", "exp", "log", "log10", "sqrt", "fabs", "sign")
BINARY = ("add", "sub", "mul", "div", "pow", "atan2", "min2", "max2")

_NP = {
    "sin": np.sin, "cos": np.cos, "tan": np.tan, "asin": np.arcsin, "acos": np.arccos,
    "atan": np.arctan, "sinh": np.sinh, "cosh": np.cosh, "tanh": np.tanh,
    "asinh": np.arcsinh, "acosh": np.arccosh, "atanh": np.arctanh, "exp": np.exp,
    "log": np.log, "log10": np.log10, "sqrt": np.sqrt,
}


def _is_const(n, v=None):
    return n.op == "const" and (v is None or n.val == v)


def _fold(op, vals):
    a = vals[0]
    if op == "neg":
        return -a
    if op in _NP:
        return float(_NP[op](a))
    if op == "fabs":
        return math.fabs(a)
    if op == "sign":
        return 0.0 if a == 0.0 else math.copysign(1.0, a)
    b = vals[1]
    if op == "add":
        return a + b
    if op == "sub":
        return a - b
    if op == "mul":
        return a * b
    if op == "div":
        return a / b
    if op == "pow":
        return a ** b
    if op == "atan2":
        return math.atan2(a, b)
    if op == "min2":
        return min(a, b)
    if op == "max2":
        return max(a, b)
    raise ValueError(op)


def unary(op, a):
    if op not in UNARY:
        raise ValueError(f"unknown function {op}")
    if a.op == "const":
        return _const_node(_fold(op, [a.val]))
    if op == "neg" and a.op == "neg":
        return a.args[0]
    return _node(op, (a,))


def binary(op, a, b):
    if op not in BINARY:
        raise ValueError(f"unknown operation {op}")
    if a.op == "const" and b.op == "const":
        return _const_node(_fold(op, [a.val, b.val]))
    if op == "add":
        if _is_const(a, 0.0):
            return b
        if _is_const(b, 0.0):
            return a
    elif op == "sub":
        if _is_const(b, 0.0):
            return a
        if _is_const(a, 0.0):
            return unary("neg", b)
        if a is b:
            return _const_node(0.0)
    elif op == "mul":
        if _is_const(a, 0.0) or _is_const(b, 0.0):
            return _const_node(0.0)
        if _is_const(a, 1.0):
            return b
        if _is_const(b, 1.0):
            return a
        if _is_const(a, -1.0):
            return unary("neg", b)
        if _is_const(b, -1.0):
            return unary("neg", a)
    elif op == "div":
        if _is_const(a, 0.0):
            return _const_node(0.0)
        if _is_const(b, 1.0):
            return a
    elif op == "pow":
        if _is_const(b, 1.0):
            return a
        if _is_const(b, 0.0):
            return _const_node(1.0)
    return _node(op, (a, b))


def select_lt(a, b, x, y):
    """a < b ? x : y (the branch rule of min2 / max2 derivatives)."""
    if x is y:
        return x
    return _node("lt", (a, b, x, y))


# ---- user-facing expression ---------------------------------------------------

def _as_node(v):
    if isinstance(v, Expr):
        return v.node
    if isinstance(v, Node):
        return v
    if isinstance(v, (int, float, np.floating, np.integer)):
        return _const_node(float(v))
    raise TypeError(f"unsupported operand {type(v).__name__}")


class Expr:
    """An expression of element inputs, data and constants."""

    __slots__ = ("node",)

    def __init__(self, node):
        self.node = node

    def __add__(self, o):
        return Expr(binary("add", self.node, _as_node(o)))

    def __radd__(self, o):
        return Expr(binary("add", _as_node(o), self.node))

    def __sub__(self, o):
        return Expr(binary("sub", self.node, _as_node(o)))

    def __rsub__(self, o):
        return Expr(binary("sub", _as_node(o), self.node))

    def __mul__(self, o):
        return Expr(binary("mul", self.node, _as_node(o)))

    def __rmul__(self, o):
        return Expr(binary("mul", _as_node(o), self.node))

    def __truediv__(self, o):
        return Expr(binary("div", self.node, _as_node(o)))

    def __rtruediv__(self, o):
        return Expr(binary("div", _as_node(o), self.node))

    def __neg__(self):
        return Expr(unary("neg", self.node))

    def __pos__(self):
        return self

    def __pow__(self, o):
        # Integer powers expand to products (the reference's Expr.__pow__,
        # amigo/expressions.py:247-262); other exponents stay pow().
        if isinstance(o, (int, np.integer)) and not isinstance(o, bool):
            o = int(o)
            if o == 0:
                return Expr(_const_node(1.0))
            base = self.node
            acc = base
            for _ in range((-o if o < 0 else o) - 1):
                acc = binary("mul", acc, base)
            return Expr(acc) if o > 0 else Expr(binary("div", _const_node(1.0), acc))
        return Expr(binary("pow", self.node, _as_node(o)))

    def __rpow__(self, o):
        return Expr(binary("pow", _as_node(o), self.node))

    def is_zero(self):
        return _is_const(self.node, 0.0)

    def __repr__(self):
        return f"Expr(#{self.node.id} {self.node.op})"


def _fn(op):
    def f(x):
        return Expr(unary(op, _as_node(x)))

    f.__name__ = op
    f.__doc__ = f"{op}(x) of an expression."
    return f


sin, cos, tan = _fn("sin"), _fn("cos"), _fn("tan")
asin, acos, atan = _fn("asin"), _fn("acos"), _fn("atan")
sinh, cosh, tanh = _fn("sinh"), _fn("cosh"), _fn("tanh")
asinh, acosh, atanh = _fn("asinh"), _fn("acosh"), _fn("atanh")
exp, log, log10, sqrt = _fn("exp"), _fn("log"), _fn("log10"), _fn("sqrt")
fabs = _fn("fabs")
abs = fabs  # noqa: A001  (the reference's am.abs)


def _bin(op):
    def f(a, b):
        if not isinstance(a, Expr) and not isinstance(b, Expr):
            raise ValueError("Neither argument is active")
        return Expr(binary(op, _as_node(a), _as_node(b)))

    f.__name__ = op
    return f


atan2, min2, max2 = _bin("atan2"), _bin("min2"), _bin("max2")


def input_expr(q):
    return Expr(_node("in", (), int(q)))


def data_expr(q):
    return Expr(_node("data", (), int(q)))


def alpha_expr():
    return Expr(_node("alpha", (), None))


# ---- derivatives ---------------------------------------------------------------

def _c(v):
    return _const_node(v)


def derivative(node, q, memo):
    """d node / d input q as a node (memo: dict keyed by (node id, q)); a
    negative q = -(j + 1) differentiates with respect to data value j."""
    key = (node.id, q)
    r = memo.get(key)
    if r is not None:
        return r
    op = node.op
    if op in ("const", "alpha"):
        r = _c(0.0)
    elif op == "data":
        r = _c(1.0 if q < 0 and node.val == -q - 1 else 0.0)
    elif op == "in":
        r = _c(1.0 if node.val == q else 0.0)
    elif op == "lt":
        a, b, x, y = node.args
        r = select_lt(a, b, derivative(x, q, memo), derivative(y, q, memo))
    elif op in UNARY:
        a = node.args[0]
        da = derivative(a, q, memo)
        if _is_const(da, 0.0):
            r = _c(0.0)
        elif op == "neg":
            r = unary("neg", da)
        else:
            r = binary("mul", _unary_slope(op, a, node), da)
    else:
        a, b = node.args
        da, db = derivative(a, q, memo), derivative(b, q, memo)
        if op == "add":
            r = binary("add", da, db)
        elif op == "sub":
            r = binary("sub", da, db)
        elif op == "mul":
            r = binary("add", binary("mul", da, b), binary("mul", a, db))
        elif op == "div":
            # (da - (a / b) db) / b
            r = binary("div", binary("sub", da, binary("mul", node, db)), b)
        elif op == "pow":
            if b.op == "const":
                r = binary("mul", binary("mul", b, binary("pow", a, _c(b.val - 1.0))), da)
            else:
                # a^b (db log a + b da / a)
                t = binary("add", binary("mul", db, unary("log", a)),
                           binary("div", binary("mul", b, da), a))
                r = binary("mul", node, t)
        elif op == "atan2":
            den = binary("add", binary("mul", a, a), binary("mul", b, b))
            r = binary("div", binary("sub", binary("mul", b, da), binary("mul", a, db)), den)
        elif op == "min2":
            r = select_lt(a, b, da, db)
        elif op == "max2":
            r = select_lt(a, b, db, da)
        else:
            raise ValueError(op)
    memo[key] = r
    return r


def _unary_slope(op, a, f):
    """d op(a) / da as a node (f is op(a) itself)."""
    one = _c(1.0)
    if op == "sin":
        return unary("cos", a)
    if op == "cos":
        return unary("neg", unary("sin", a))
    if op == "tan":
        return binary("add", one, binary("mul", f, f))
    if op == "asin":
        return binary("div", one, unary("sqrt", binary("sub", one, binary("mul", a, a))))
    if op == "acos":
        return unary("neg", binary("div", one, unary("sqrt", binary("sub", one, binary("mul", a, a)))))
    if op == "atan":
        return binary("div", one, binary("add", one, binary("mul", a, a)))
    if op == "sinh":
        return unary("cosh", a)
    if op == "cosh":
        return unary("sinh", a)
    if op == "tanh":
        return binary("sub", one, binary("mul", f, f))
    if op == "asinh":
        return binary("div", one, unary("sqrt", binary("add", binary("mul", a, a), one)))
    if op == "acosh":
        return binary("div", one, unary("sqrt", binary("sub", binary("mul", a, a), one)))
    if op == "atanh":
        return binary("div", one, binary("sub", one, binary("mul", a, a)))
    if op == "exp":
        return f
    if op == "log":
        return binary("div", one, a)
    if op == "log10":
        return binary("div", one, binary("mul", a, _c(math.log(10.0))))
    if op == "sqrt":
        return binary("div", _c(0.5), f)
    if op == "fabs":
        return unary("sign", a)
    if op == "sign":
        return _c(0.0)
    raise ValueError(op)


# ---- numerical evaluation (checks and host use) --------------------------------

_NP_COMPLEX = dict(_NP)


def evaluate(nodes, z, d, dtype=np.float64, alpha=1.0):
    """Values of the given nodes for inputs z and data d (numpy arrays of
    shape (E, nin) / (E, ndata), or 1-D for one element).  dtype complex128
    evaluates in complex arithmetic (complex-step derivative checks of the
    analytic functions)."""
    z = np.asarray(z, dtype=dtype)
    d = np.asarray(d, dtype=np.float64)
    single = z.ndim == 1
    if single:
        z, d = z[None, :], d[None, :]
    order = sorted(_reachable(nodes), key=lambda n: n.id)
    val = {}
    for n in order:
        op = n.op
        if op == "const":
            v = np.full(z.shape[0], n.val, dtype=dtype)
        elif op == "in":
            v = z[:, n.val]
        elif op == "data":
            v = d[:, n.val].astype(dtype)
        elif op == "alpha":
            v = np.full(z.shape[0], alpha, dtype=dtype)
        elif op == "lt":
            a, b, x, y = (val[c.id] for c in n.args)
            v = np.where(np.real(a) < np.real(b), x, y)
        elif op in UNARY:
            a = val[n.args[0].id]
            if op == "neg":
                v = -a
            elif op == "fabs":
                v = np.where(np.real(a) < 0, -a, a)
            elif op == "sign":
                v = np.sign(np.real(a)).astype(dtype)
            else:
                v = _NP_COMPLEX[op](a)
        else:
            a, b = val[n.args[0].id], val[n.args[1].id]
            if op == "add":
                v = a + b
            elif op == "sub":
                v = a - b
            elif op == "mul":
                v = a * b
            elif op == "div":
                v = a / b
            elif op == "pow":
                v = a ** b
            elif op == "atan2":
                if np.iscomplexobj(a) or np.iscomplexobj(b):
                    # atan(a / b) with the quadrant of the real parts
                    ra, rb = np.real(a), np.real(b)
                    v = np.arctan(a / b) + np.where(rb < 0, np.where(ra >= 0, np.pi, -np.pi), 0.0)
                else:
                    v = np.arctan2(a, b)
            elif op == "min2":
                v = np.where(np.real(a) < np.real(b), a, b)
            elif op == "max2":
                v = np.where(np.real(a) < np.real(b), b, a)
            else:
                raise ValueError(op)
        val[n.id] = v
    out = [val[n.id] for n in nodes]
    return [o[0] for o in out] if single else out


def _reachable(nodes):
    seen = {}
    stack = list(nodes)
    while stack:
        n = stack.pop()
        if n.id in seen:
            continue
        seen[n.id] = n
        stack.extend(n.args)
    return seen.values()


def reachable(nodes):
    """Every node the given nodes depend on, in dependency order."""
    return sorted(_reachable(nodes), key=lambda n: n.id)
