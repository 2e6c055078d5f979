"""Request / on-disk format of problems: JSON documents.

The reference serializes a model as JSON: its components (each with its
expressions), sizes, links and data (amigo/model.py:1624-1720,
amigo/component.py:552-598, Expr.serialize / deserialize in
amigo/expressions.py).  Here the unit is a ComponentProblem, the form a
solve request reaches the GPU in:

    {"format": "amigo_amd.problem", "version": 1,
     "name": ..., "n": ...,
     "components": [component, ...],     # generated element types
     "groups": [{"etype": int} or {"component": i}, "idx": array, "data": array],
     "mult": array, "lower": array, "upper": array, "x0": array,
     "var_groups": array or null}

A component is stored by its declarations and its element Lagrangian as an
expression list (post-order; each entry [op, [argument positions], value]),
which is all the code generator needs; loading it regenerates the same
kernel source, so a compiled kernel is found again by the hash of that
source.  Arrays are stored bit-exact: {"dtype", "shape", "b64"} with the
little-endian bytes in base64 (doubles round-trip exactly, and a large
element array stays compact).

    doc = wire.dumps(prob, components)      # JSON text
    prob = wire.loads(doc)                   # a ComponentProblem on the GPU
"""

import base64
import json

import numpy as np

from . import expressions as ex
from .component import Component

FORMAT = "amigo_amd.problem"
VERSION = 1


def pack_array(a):
    a = np.ascontiguousarray(a)
    dt = a.dtype.newbyteorder("<")
    return {"dtype": dt.str, "shape": list(a.shape),
            "b64": base64.b64encode(a.astype(dt, copy=False).tobytes()).decode("ascii")}


def unpack_array(d):
    raw = base64.b64decode(d["b64"])
    return np.frombuffer(raw, dtype=np.dtype(d["dtype"])).reshape(d["shape"]).copy()


# ---- expressions and components ------------------------------------------------

def _nodes_to_list(roots):
    from .codegen import _postorder

    order = _postorder(roots)
    pos = {n.id: i for i, n in enumerate(order)}
    items = [[n.op, [pos[a.id] for a in n.args], n.val] for n in order]
    return items, [pos[r.id] for r in roots]


def _list_to_nodes(items):
    built = []
    for op, args, val in items:
        a = [built[i] for i in args]
        if op == "const":
            n = ex._const_node(val)
        elif op == "in":
            n = ex.input_expr(val).node
        elif op == "data":
            n = ex.data_expr(val).node
        elif op == "alpha":
            n = ex.alpha_expr().node
        elif op == "lt":
            n = ex.select_lt(*a)
        elif op in ex.UNARY:
            n = ex.unary(op, a[0])
        elif op in ex.BINARY:
            n = ex.binary(op, a[0], a[1])
        else:
            raise ValueError(f"unknown expression op {op!r}")
        built.append(n)
    return built


class StoredComponent(Component):
    """A component read back from its stored form: declarations and the
    element Lagrangian (compute() is not needed any more)."""

    def __init__(self, name, lagrangian):
        super().__init__(name=name)
        self._stored = lagrangian

    def lagrangian(self):
        return self._stored


def serialize_component(comp):
    L = comp.lagrangian()
    items, roots = _nodes_to_list([L.node])

    def decl(s):
        return [[n, list(s.shapes[n])] for n in s.names]

    return {
        "name": comp.name,
        "constants": {n: comp.constants.values[n] for n in comp.constants.names},
        "inputs": decl(comp.inputs),
        "data": decl(comp.data),
        "constraints": decl(comp.constraints),
        "objective": list(comp.objective.names),
        "lagrangian": {"nodes": items, "root": roots[0]},
    }


def deserialize_component(d):
    nodes = _list_to_nodes(d["lagrangian"]["nodes"])
    comp = StoredComponent(d["name"], ex.Expr(nodes[d["lagrangian"]["root"]]))
    for n, v in d["constants"].items():
        comp.add_constant(n, v)
    for n, s in d["inputs"]:
        comp.add_input(n, shape=tuple(s) or None)
    for n, s in d["data"]:
        comp.add_data(n, shape=tuple(s) or None)
    for n, s in d["constraints"]:
        comp.add_constraint(n, shape=tuple(s) or None)
    for n in d["objective"]:
        comp.add_objective(n)
    return comp


# ---- problems -------------------------------------------------------------------

def problem_to_dict(prob, components=None):
    """The document of a ComponentProblem.  components: {etype: Component}
    for its generated element types (built-in types are stored by number)."""
    from . import codegen

    components = dict(components or {})
    for ge in codegen.registered():
        components.setdefault(ge.etype, ge.comp)
    comp_list, comp_pos = [], {}
    groups = []
    for g in prob.groups:
        entry = {"idx": pack_array(g.idx.astype(np.int32)), "data": pack_array(g.data.astype(np.float64))}
        if g.etype >= codegen.GENERATED_BASE:
            if g.etype not in components:
                raise KeyError(f"no component given for generated element type {g.etype}")
            if g.etype not in comp_pos:
                comp_pos[g.etype] = len(comp_list)
                comp_list.append(serialize_component(components[g.etype]))
            entry["component"] = comp_pos[g.etype]
        else:
            entry["etype"] = int(g.etype)
        groups.append(entry)
    vg = getattr(prob, "var_groups", None)
    return {
        "format": FORMAT, "version": VERSION, "name": prob.name, "n": int(prob.n),
        "components": comp_list, "groups": groups,
        "mult": pack_array(np.asarray(prob.mult_ind, np.uint8)),
        "lower": pack_array(prob.lower), "upper": pack_array(prob.upper),
        "x0": pack_array(prob.x0),
        "var_groups": None if vg is None else pack_array(np.asarray(vg, np.int32)),
    }


def problem_from_dict(doc, device="cuda", build=True):
    """A ComponentProblem from its document; generated element types are
    generated (and compiled if no cached kernel matches) and registered."""
    from . import codegen
    from .problem import ComponentProblem, ElementGroup

    if doc.get("format") != FORMAT or int(doc.get("version", 0)) != VERSION:
        raise ValueError("not an amigo_amd problem document of a known version")
    etypes = [codegen.element_type(deserialize_component(c), build=build).etype
              for c in doc["components"]]
    groups = []
    for g in doc["groups"]:
        et = etypes[g["component"]] if "component" in g else int(g["etype"])
        groups.append(ElementGroup(et, unpack_array(g["idx"]), unpack_array(g["data"])))
    vg = doc.get("var_groups")
    return ComponentProblem(int(doc["n"]), groups, unpack_array(doc["mult"]).astype(bool),
                            unpack_array(doc["lower"]), unpack_array(doc["upper"]),
                            unpack_array(doc["x0"]), device=device, name=doc.get("name", "problem"),
                            var_groups=None if vg is None else unpack_array(vg))


def dumps(prob, components=None):
    return json.dumps(problem_to_dict(prob, components))


def loads(text, device="cuda", build=True):
    return problem_from_dict(json.loads(text), device=device, build=build)
