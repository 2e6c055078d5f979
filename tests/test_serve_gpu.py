"""Serving form: requests of different sizes solved as they arrive, several
in flight on their own HIP streams, admitted by predicted front memory.

- Each result is bit-identical alone (one slot) and in the mix (several
  slots, overlapping on the device): every kernel is deterministic.
- Admission respects the memory budget: with a budget below two requests'
  predicted memory, no two run together.
- A Poisson stream completes with every solve converged and reports
  p50 / p99 latency."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SPECS = [
    {"type": "cart_pole", "N": 48, "seed": 11},
    {"type": "elliptic", "side": 40, "seed": 3},
    {"type": "double_integrator", "N": 96},
    {"type": "cart_pole", "N": 96, "seed": 12},
    {"type": "elliptic", "side": 56, "seed": 4},
    {"type": "cart_pole", "N": 64, "seed": 13},
]


def _run(slots, specs, **kw):
    from amigo_amd.serve import SolveServer

    with SolveServer(slots=slots, **kw) as srv:
        futs = [srv.submit(s) for s in specs]
        res = [f.result() for f in futs]
    return res, srv


def test_results_are_identical_alone_and_in_the_mix():
    alone, _ = _run(1, SPECS)
    mixed, srv = _run(4, SPECS)
    assert srv.peak_inflight > 1
    for a, b in zip(alone, mixed):
        assert a["converged"] and b["converged"]
        assert a["iterations"] == b["iterations"]
        assert a["inertia"] == b["inertia"]
        assert np.array_equal(a["x"], b["x"])


def test_admission_respects_the_memory_budget():
    res, srv = _run(4, SPECS[:4], memory_budget=1)
    assert all(r["converged"] for r in res)
    assert srv.peak_inflight == 1
    big = max(r["predicted_bytes"] for r in res)
    res2, srv2 = _run(4, SPECS[:4], memory_budget=2 * big)
    assert srv2.peak_bytes <= 2 * big
    for a, b in zip(res, res2):
        assert np.array_equal(a["x"], b["x"])


def test_poisson_stream_reports_latency():
    from amigo_amd.serve import SolveServer, request_mix, run_poisson

    specs = request_mix(8, seed=1, cart_n=(32, 128), side=(32, 64))
    with SolveServer(slots=4) as srv:
        st = run_poisson(srv, specs, rate=50.0, seed=2)
    assert st["converged"] == len(specs)
    assert 0 < st["p50_ms"] <= st["p99_ms"]
    assert st["solves_per_s"] > 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lockstep_batches_equal_requests_alone(precision):
    """Requests gathered into lockstep batches (fused level-kernel launches)
    give bitwise the result each gets alone."""
    from amigo_amd.serve import BatchSolveServer

    alone, _ = _run(1, SPECS, precision=precision)
    with BatchSolveServer(max_batch=4, max_batches=2, gather_s=0.05, precision=precision) as srv:
        futs = [srv.submit(s) for s in SPECS]
        res = [f.result() for f in futs]
    sizes = [b for b, _ in srv.batches]
    assert sum(sizes) == len(SPECS) and max(sizes) > 1
    for a, b in zip(alone, res):
        assert b["converged"]
        assert a["iterations"] == b["iterations"]
        assert a["inertia"] == b["inertia"]
        assert np.array_equal(a["x"], b["x"])


def test_lockstep_poisson_stream():
    from amigo_amd.serve import BatchSolveServer, request_mix, run_poisson

    specs = request_mix(12, seed=5, cart_n=(32, 128), side=(32, 48))
    with BatchSolveServer(max_batch=8, max_batches=2) as srv:
        st = run_poisson(srv, specs, rate=200.0, seed=6)
    assert st["converged"] == len(specs)
    assert 0 < st["p50_ms"] <= st["p99_ms"]


def test_lockstep_batches_hold_one_problem_type():
    from amigo_amd.serve import BatchSolveServer

    with BatchSolveServer(max_batch=8, max_batches=1, gather_s=0.05) as srv:
        futs = [srv.submit(s) for s in SPECS]
        res = [f.result() for f in futs]
    assert all(r["converged"] for r in res)
    assert sum(len(b) for b in srv.batch_specs) == len(SPECS)
    for specs in srv.batch_specs:
        assert len({s["type"] for s in specs}) == 1
