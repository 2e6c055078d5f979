"""The native driver's inertia-correction rules (amigo_ipm_regularization_
trace, host only) against the oracle restatement of the reference's rules
(oracle/driver.py), and both against sequences derived by hand from the
constants of Waechter & Biegler (2006), Table 3: delta_w starts at 1e-4 or
a third of the last value, grows by 100 on its first increase and by 8
after, caps at 1e20; delta_c = 1e-8 mu^(1/4); a singular matrix is probed
with delta_c alone, then delta_w alone, then both; three systems in a row
that needed a perturbation make it structural."""

import numpy as np
import pytest

from amigo_amd.driver import regularization_trace
from oracle import driver as OD


def _both(outcomes, mus, has_c=True):
    native = [tuple(r) for r in regularization_trace(outcomes, mus, has_c)]
    oracle = OD.ic_trace(outcomes, mus, has_c)
    return native, oracle


def test_wrong_inertia_ladder():
    native, oracle = _both([-1, 3, 3, 3, 3, 0], [0.1])
    exp = [(0.0, 0.0), (1e-4, 0.0), (1e-2, 0.0), (1.0, 0.0), (100.0, 0.0)]
    assert native == oracle == exp
    # the next matrix starts from zero, then a third of the last value, then x8
    native, oracle = _both([-1, 3, 0, -1, 3, 3, 0], [0.1, 0.1])
    last = 1e-4
    exp = [(0.0, 0.0), (1e-4, 0.0), (0.0, 0.0), (max(1e-20, last / 3), 0.0), (max(1e-20, last / 3) * 8, 0.0)]
    assert native == exp and oracle == exp


def test_delta_w_cap_then_delta_c_then_give_up():
    # 0, 1e-4 * 100^k up to <= 1e20: 1e-4, 1e-2, ..., 1e20 (13 values), then delta_c on and again
    native, oracle = _both([-1] + [3] * 40, [1.0])
    assert native == oracle
    dws = [t[0] for t in native]
    assert dws[:14] == [0.0] + [1e-4 * 100.0**k for k in range(13)]
    # after the cap: delta_c = 1e-8 mu^0.25 switched on and delta_w restarts
    assert native[14] == (1e-4, 1e-8) and len(native) == 27


def test_singular_probe_order_and_structural_jacobian():
    mu = 0.0625
    dc = 1e-8 * mu**0.25
    # probe: bare, delta_c only, delta_w only, both
    native, oracle = _both([-1, 1, 1, 1, 0], [mu])
    assert native == oracle == [(0.0, 0.0), (0.0, dc), (1e-4, 0.0), (1e-2, dc)]
    # three matrices repaired by delta_c alone: the Jacobian becomes structurally
    # degenerate and the fourth matrix starts with delta_c on
    seq = [-1, 1, 0] * 3 + [-1, 0]
    native, oracle = _both(seq, [mu] * 4)
    assert native == oracle
    assert native[-1] == (0.0, dc)


def test_too_few_negative_pivots_retries_the_same_matrix_three_times_per_solve():
    # each of the first three such matrices gets one identical retry (pivot
    # quality 1e-6 -> 1e-3 -> 10^-1.5 -> 0.1), after which it is treated as singular
    seq = []
    for _ in range(4):
        seq += [-1, 2, 2, 0]
    native, oracle = _both(seq, [0.5] * 4)
    assert native == oracle
    per = [native[i:i + 3] for i in range(0, 9, 3)]
    for p in per:
        assert p[0] == p[1]           # the retry factors the same matrix
    assert native[9] != native[10]    # the fourth: straight to the singular probe


def test_no_constraints_singular_is_wrong_inertia():
    native, oracle = _both([-1, 1, 1, 0], [0.3], has_c=False)
    assert native == oracle == [(0.0, 0.0), (1e-4, 0.0), (1e-2, 0.0)]


@pytest.mark.parametrize("seed", range(8))
def test_random_outcome_sequences_match_the_oracle(seed):
    rng = np.random.default_rng(seed)
    seq, mus = [], []
    for _ in range(60):
        seq.append(-1)
        mus.append(float(10.0 ** rng.uniform(-10, 0)))
        for _ in range(int(rng.integers(1, 8))):
            seq.append(int(rng.choice([1, 2, 3], p=[0.4, 0.2, 0.4])))
        seq.append(0)
    for has_c in (True, False):
        native, oracle = _both(seq, mus, has_c)
        assert native == oracle


def test_planted_rule_change_fails():
    """A ladder that grows by 10 instead of 100 on the first increase is
    caught by the hand-derived sequence."""
    orig = OD._grow

    def wrong(st):
        prev = st["dw"]
        if st["dw"] == 0.0:
            st["dw"] = 1e-4
        else:
            st["dw"] *= 10.0
        return st["dw"] <= 1e20 or (st.update(dw=prev) and False)

    OD._grow = wrong
    try:
        assert OD.ic_trace([-1, 3, 3, 0], [0.1]) != [(0.0, 0.0), (1e-4, 0.0), (1e-2, 0.0)]
    finally:
        OD._grow = orig
