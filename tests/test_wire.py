"""Problem documents (amigo_amd.wire), host side: arrays round-trip bit for
bit, a stored component regenerates the same kernel source, and a problem
read back has the same element groups, bounds and KKT pattern."""

import json

import numpy as np
import pytest

from amigo_amd import codegen, wire
from amigo_amd.problems import components as C


def test_arrays_round_trip_bit_exact():
    a = np.array([0.0, -0.0, 1.0 / 3.0, np.pi, np.inf, -np.inf, 1e-310, np.nan])
    b = wire.unpack_array(json.loads(json.dumps(wire.pack_array(a))))
    assert b.dtype == a.dtype and np.array_equal(a.view(np.int64), b.view(np.int64))
    i = np.arange(-5, 7, dtype=np.int32).reshape(3, 4)
    assert np.array_equal(wire.unpack_array(wire.pack_array(i)), i)


@pytest.mark.parametrize("comp", C.all_components(), ids=lambda c: c.name)
def test_stored_component_regenerates_the_same_kernel(comp):
    d = json.loads(json.dumps(wire.serialize_component(comp)))
    back = wire.deserialize_component(d)
    src0, der0 = codegen.generate_source(comp)
    src1, der1 = codegen.generate_source(back)
    assert src0 == src1
    assert back.num_inputs() == comp.num_inputs() and back.num_data() == comp.num_data()
    assert np.array_equal(der0.structure, der1.structure)


def test_problem_document_round_trip():
    prob, ge = C.hanging_chain(K=12, device="cpu")
    text = wire.dumps(prob)
    doc = json.loads(text)
    assert doc["format"] == "amigo_amd.problem" and len(doc["components"]) == 1
    back = wire.loads(text, device="cpu", build=False)
    assert back.n == prob.n and len(back.groups) == len(prob.groups)
    for g0, g1 in zip(prob.groups, back.groups):
        assert g0.etype == g1.etype
        assert np.array_equal(g0.idx, g1.idx) and np.array_equal(g0.data, g1.data)
    for f in ("lower", "upper", "x0"):
        assert np.array_equal(getattr(prob, f), getattr(back, f))
    assert np.array_equal(prob.mult_ind, back.mult_ind)
    assert np.array_equal(prob.rowp, back.rowp) and np.array_equal(prob.cols, back.cols)


def test_builtin_problem_document_round_trip():
    from amigo_amd.problems import cart_pole

    prob = cart_pole.seeded(5, nmin=16, nmax=40, device="cpu")
    back = wire.loads(wire.dumps(prob), device="cpu", build=False)
    assert [g.etype for g in back.groups] == [g.etype for g in prob.groups]
    assert np.array_equal(prob.var_groups, back.var_groups)
    assert np.array_equal(prob.rowp, back.rowp) and np.array_equal(prob.cols, back.cols)


def test_bad_documents_are_rejected():
    with pytest.raises(ValueError):
        wire.problem_from_dict({"format": "something else", "version": 1})
