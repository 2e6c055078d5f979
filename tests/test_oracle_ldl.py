"""Pins the numpy restatement of the numeric multifrontal LDL^T
(oracle/ldl.py): on the reference's own symbolic plans it reproduces the
reference's inertia, per-node pivot and delayed-pivot counts and solutions
(committed fixtures written by the compiled reference, and live when
oracle/_ref is built), the eigenvalue inertia (Sylvester) and dense solves."""

import numpy as np
import pytest

from oracle.ldl import OracleLDL
from tests import fixtures as FX
from tests import matrices as M


def _oracle_from_fixture(name, order):
    A, ustab = FX.matrix(name)
    sym = {k: FX.field(name, order, k) for k in FX.SYMBOLIC_KEYS}
    O = OracleLDL(A.shape[0], A.indptr, A.indices, sym, ustab=ustab)
    O.factor(A.data)
    return A, O


@pytest.mark.parametrize("name", FX.ldl_cases())
@pytest.mark.parametrize("order", ["NATURAL", "AMD"])
def test_oracle_matches_reference_fixtures(name, order):
    A, O = _oracle_from_fixture(name, order)
    assert O.get_inertia() == tuple(FX.field(name, order, "inertia"))
    assert np.array_equal(O.npiv, FX.field(name, order, "npiv"))
    assert np.array_equal(O.ndel, FX.field(name, order, "ndel"))
    b = np.arange(1, A.shape[0] + 1, dtype=float)
    x = O.solve(b)
    xr = FX.field(name, order, "x")
    assert np.max(np.abs(x - xr)) <= 1e-10 * np.max(np.abs(xr))


def test_fixtures_exercise_delays_and_two_by_two_pivots():
    total = sum(int(FX.field(n, o, "ndel").sum()) for n in FX.ldl_cases() for o in ("NATURAL", "AMD"))
    assert total > 50
    _, O = _oracle_from_fixture("kkt_delay", "AMD")
    assert np.any(np.abs(np.diag(O.D, 1)) > 0)   # at least one 2x2 block


@pytest.mark.parametrize("name", FX.ldl_cases())
def test_oracle_inertia_is_sylvester_and_solve_is_dense(name):
    A, O = _oracle_from_fixture(name, "AMD")
    ev = np.linalg.eigvalsh(A.toarray())
    assert O.get_inertia() == (int(np.sum(ev >= 0)), int(np.sum(ev < 0)))
    b = np.random.default_rng(0).normal(size=(A.shape[0], 3))
    x = O.solve(b)
    assert np.allclose(x, np.linalg.solve(A.toarray(), b), rtol=1e-9, atol=1e-9)


def test_planted_threshold_mistake_is_caught():
    """A restatement that skipped the threshold test (ustab = 0) takes no
    delays; the reference takes some on this matrix."""
    A, ustab = FX.matrix("kkt_delay")
    sym = {k: FX.field("kkt_delay", "AMD", k) for k in FX.SYMBOLIC_KEYS}
    O = OracleLDL(A.shape[0], A.indptr, A.indices, sym, ustab=0.0)
    O.factor(A.data)
    assert not np.array_equal(O.ndel, FX.field("kkt_delay", "AMD", "ndel"))


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("ustab", [0.01, 0.1])
def test_oracle_matches_live_reference(seed, ustab):
    """Elliptic-control KKT matrices at random interior points: 80-170
    delayed pivots each.  (Matrices whose trailing blocks cancel to rounding
    level are not used: there every pivot decision compares noise.)"""
    from oracle import ref

    if not ref.available():
        pytest.skip("oracle/_ref not built")
    A = M.grid_kkt(12, seed=seed)[0]
    R = ref.RefLDL(A.indptr, A.indices, order="AMD", ustab=ustab)
    R.factor(A.data)
    O = OracleLDL(A.shape[0], A.indptr, A.indices, R.symbolic(), ustab=ustab)
    O.factor(A.data)
    assert O.get_inertia() == R.inertia()
    rp, rd = R.pivot_counts()
    assert np.array_equal(O.npiv, rp) and np.array_equal(O.ndel, rd)
