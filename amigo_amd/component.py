"""Components: the element-level modelling API (``am.Component``).

A component declares its constants, inputs, data, constraints and objective
and writes them as expressions of the inputs in ``compute()``, with the same
method names as the reference (amigo/component.py:521-745):

    class Interval(Component):
        def __init__(self):
            super().__init__()
            self.add_data("dt")
            self.add_input("q", shape=2)
            self.add_constraint("res")

        def compute(self):
            dt = self.data["dt"]
            q = self.inputs["q"]
            self.constraints["res"] = q[1] - q[0] - dt

One element of the component has the inputs in declaration order (arrays
flattened row-major) followed by one multiplier per scalar constraint, and
the data values in declaration order; its Lagrangian contribution is

    L_e = alpha * (sum of the objectives) + sum_i lambda_i c_i

(the reference's _compute_lagrangian, component.py:775-800, which keeps
only the last objective when there are several; here they are summed).
``amigo_amd.codegen`` turns the component into a HIP element kernel.
"""

import numpy as np

from . import expressions as ex
from .expressions import Expr


def _shape(shape):
    if shape is None:
        return ()
    if isinstance(shape, (int, np.integer)):
        return (int(shape),)
    shape = tuple(int(s) for s in shape)
    if len(shape) > 2:
        raise ValueError("at most two dimensions")
    return shape


def _size(shape):
    n = 1
    for s in shape:
        n *= s
    return n


class _Array:
    """A 1-D or 2-D array of expressions, indexed like the reference's
    expression arrays (a[i], a[i, j])."""

    def __init__(self, items, shape):
        self.items = list(items)
        self.shape = shape

    def __getitem__(self, i):
        if isinstance(i, tuple):
            if len(self.shape) != 2:
                raise IndexError("two indices into a one-dimensional array")
            r, c = int(i[0]), int(i[1])
            if not (0 <= r < self.shape[0] and 0 <= c < self.shape[1]):
                raise IndexError(f"index {i} out of range {self.shape}")
            return self.items[r * self.shape[1] + c]
        i = int(i)
        if len(self.shape) == 2:
            raise IndexError("two indices needed")
        if not 0 <= i < self.shape[0]:
            raise IndexError(f"index {i} out of range {self.shape}")
        return self.items[i]

    def __len__(self):
        return self.shape[0]


class _Declared:
    """Named entries in declaration order, each with a shape and a slot
    offset into the element's flat vector."""

    def __init__(self):
        self.names = []
        self.shapes = {}
        self.meta = {}
        self.offset = {}
        self.size = 0

    def _add(self, name, shape, meta):
        if name in self.shapes:
            raise ValueError(f"{name} declared twice")
        self.names.append(name)
        self.shapes[name] = _shape(shape)
        self.meta[name] = meta
        self.offset[name] = self.size
        self.size += _size(self.shapes[name])

    def __iter__(self):
        return iter(self.names)

    def __len__(self):
        return len(self.names)

    def get_shape(self, name):
        s = self.shapes[name]
        return None if s == () else s

    def get_meta(self, name):
        return self.meta[name]


class _Leaves(_Declared):
    """Inputs or data: indexing returns their leaf expressions."""

    def __init__(self, leaf):
        super().__init__()
        self._leaf = leaf
        self._base = 0

    def add(self, name, shape=None, **kwargs):
        self._add(name, shape, kwargs)

    def __getitem__(self, name):
        shape = self.shapes[name]
        o = self._base + self.offset[name]
        if shape == ():
            return self._leaf(o)
        return _Array([self._leaf(o + q) for q in range(_size(shape))], shape)


class _Constants(_Declared):
    def __init__(self):
        super().__init__()
        self.values = {}

    def add(self, name, value, type=float, **kwargs):
        self._add(name, None, kwargs)
        self.values[name] = type(value)

    def __getitem__(self, name):
        return self.values[name]


class _Results(_Declared):
    """Constraints or objectives: compute() assigns their expressions."""

    def __init__(self, scalar_only=False):
        super().__init__()
        self.exprs = {}
        self.scalar_only = scalar_only

    def add(self, name, shape=None, **kwargs):
        if self.scalar_only and shape is not None:
            raise ValueError("objectives are scalars")
        self._add(name, shape, kwargs)

    def __setitem__(self, name, value):
        if name not in self.shapes:
            raise KeyError(f"{name} was not declared")
        n = _size(self.shapes[name])
        if self.shapes[name] == ():
            vals = [value]
        else:
            vals = list(np.asarray(value, dtype=object).ravel()) if not isinstance(value, _Array) else value.items
        if len(vals) != n:
            raise ValueError(f"{name} needs {n} expressions, got {len(vals)}")
        self.exprs[name] = [v if isinstance(v, Expr) else Expr(ex._as_node(v)) for v in vals]

    def __getitem__(self, name):
        e = self.exprs[name]
        shape = self.shapes[name]
        return e[0] if shape == () else _Array(e, shape)


class Component:
    """One element type of an optimization problem (amigo/component.py:
    521-745): constants, inputs, data, constraints, objective and compute()."""

    def __init__(self, name=None):
        self.name = name if name is not None else type(self).__name__
        self.constants = _Constants()
        self.inputs = _Leaves(ex.input_expr)
        self.data = _Leaves(ex.data_expr)
        self.constraints = _Results()
        self.objective = _Results(scalar_only=True)
        self.alpha = ex.alpha_expr()
        self.args = [{}]
        self._lagrangian = None

    # ---- declarations (component.py:638-684) --------------------------------
    def add_constant(self, name, value, **kwargs):
        self.constants.add(name, value, **kwargs)

    def add_input(self, name, shape=None, **kwargs):
        self.inputs.add(name, shape=shape, **kwargs)

    def add_data(self, name, shape=None, **kwargs):
        self.data.add(name, shape=shape, **kwargs)

    def add_constraint(self, name, shape=None, **kwargs):
        self.constraints.add(name, shape=shape, **kwargs)

    def add_objective(self, name, **kwargs):
        self.objective.add(name, **kwargs)

    def compute(self, **kwargs):
        pass

    # ---- queries (component.py:686-745) --------------------------------------
    def get_input_names(self):
        return list(self.inputs)

    def get_constraint_names(self):
        return list(self.constraints)

    def get_data_names(self):
        return list(self.data)

    def get_num_constraints(self):
        return self.constraints.size

    def num_inputs(self):
        """Element inputs: the inputs, then one multiplier per constraint."""
        return self.inputs.size + self.constraints.size

    def num_data(self):
        return self.data.size

    def input_slot(self, name, index=0):
        """Position of an input (or of a constraint's multiplier) in the
        element's input vector."""
        if name in self.inputs.shapes:
            return self.inputs.offset[name] + int(index)
        if name in self.constraints.shapes:
            return self.inputs.size + self.constraints.offset[name] + int(index)
        raise KeyError(name)

    def multiplier_slots(self):
        n0 = self.inputs.size
        return list(range(n0, n0 + self.constraints.size))

    # ---- the element Lagrangian ------------------------------------------------
    def lagrangian(self):
        """alpha * (objectives) + sum_i lambda_i c_i as an expression."""
        if self._lagrangian is None:
            args = self.args[0] if self.args else {}
            for name in self.constraints.names:
                self.constraints.exprs.pop(name, None)
            for name in self.objective.names:
                self.objective.exprs.pop(name, None)
            self.compute(**args)
            for name in self.constraints.names:
                if name not in self.constraints.exprs:
                    raise ValueError(f"compute() did not set constraint {name}")
            for name in self.objective.names:
                if name not in self.objective.exprs:
                    raise ValueError(f"compute() did not set objective {name}")
            L = Expr(ex._const_node(0.0))
            f = Expr(ex._const_node(0.0))
            for name in self.objective.names:
                f = f + self.objective.exprs[name][0]
            if not f.is_zero():
                L = self.alpha * f
            base = self.inputs.size
            for name in self.constraints.names:
                o = base + self.constraints.offset[name]
                for q, c in enumerate(self.constraints.exprs[name]):
                    L = L + c * ex.input_expr(o + q)
            self._lagrangian = L
        return self._lagrangian
