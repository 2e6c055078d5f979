"""Restatement of the reference's inertia-correction rules
(amigo/algorithm/inertia_correction.py:41-393), ORACLE infrastructure for
tests: the native driver's regularization state machine
(csrc/ipm_driver.h Regularization) is checked against it on scripted
factorization outcomes.  Pinned by tests/test_driver_host.py against
sequences derived by hand from the published constants of Waechter &
Biegler (2006), Table 3.

The state is a dict; ``ic_trace`` replays outcomes (0 inertia correct,
1 singular, 2 too few negative pivots, 3 wrong inertia, -1 a new KKT matrix
whose barrier parameter is the next mu) and returns the (delta_w, delta_c)
of every factorization tried.
"""

UNKNOWN, NOT_DEGENERATE, DEGENERATE = 0, 1, 2
# degeneracy probe of the current matrix, by the perturbations it has on
NONE, BARE, DC_ONLY, DW_ONLY, BOTH = 0, 1, 2, 3, 4


def new_state():
    return {"dw": 0.0, "dc": 0.0, "dw_last": 0.0, "dc_last": 0.0, "hess": UNKNOWN, "jac": UNKNOWN,
            "count": 0, "probe": NONE, "pivtol": 1e-6, "mu": 1.0}


def _dc(st):
    """delta_c = 1e-8 mu^0.25 (inertia_correction.py:85-87)."""
    return 1e-8 * st["mu"] ** 0.25


def _grow(st):
    """inertia_correction.py:89-113: first 1e-4, or a third of the last
    value (at least 1e-20); then x100 while the last is zero or far below,
    else x8; above 1e20 the step is undone and False returned."""
    prev = st["dw"]
    if st["dw"] == 0.0:
        st["dw"] = 1e-4 if st["dw_last"] == 0.0 else max(1e-20, st["dw_last"] * (1.0 / 3))
    elif st["dw_last"] == 0.0 or 1e5 * st["dw_last"] < st["dw"]:
        st["dw"] *= 100.0
    else:
        st["dw"] *= 8.0
    if st["dw"] > 1e20:
        st["dw"] = prev
        st["dw_last"] = 0.0
        return False
    return True


def _finish_probe(st):
    """inertia_correction.py:200-230."""
    p = st["probe"]
    if p == NONE:
        return
    if p == BARE:
        for key in ("hess", "jac"):
            if st[key] == UNKNOWN:
                st[key] = NOT_DEGENERATE
    elif p == DC_ONLY:
        if st["hess"] == UNKNOWN:
            st["hess"] = NOT_DEGENERATE
        if st["jac"] == UNKNOWN:
            st["count"] += 1
            if st["count"] >= 3:
                st["jac"] = DEGENERATE
    elif p == DW_ONLY:
        if st["jac"] == UNKNOWN:
            st["jac"] = NOT_DEGENERATE
        if st["hess"] == UNKNOWN:
            st["count"] += 1
            if st["count"] >= 3:
                st["hess"] = DEGENERATE
    else:
        st["count"] += 1
        if st["count"] >= 3:
            st["hess"] = st["jac"] = DEGENERATE
    st["probe"] = NONE


def new_matrix(st, mu):
    """inertia_correction.py:232-258 (consider a new system)."""
    st["mu"] = mu
    _finish_probe(st)
    if st["dw"] > 0.0:
        st["dw_last"] = st["dw"]
    if st["dc"] > 0.0:
        st["dc_last"] = st["dc"]
    st["probe"] = BARE if (st["hess"] == UNKNOWN or st["jac"] == UNKNOWN) else NONE
    st["dc"] = _dc(st) if st["jac"] == DEGENERATE else 0.0
    st["dw"] = 0.0
    st["raised"] = False
    if st["hess"] == DEGENERATE:
        _grow(st)


def _wrong_inertia(st):
    """inertia_correction.py:115-134."""
    _finish_probe(st)
    if _grow(st):
        return True
    if st["dc"] == 0.0:
        st["dc"] = _dc(st)
        st["dw"] = 0.0
        if st["hess"] == DEGENERATE:
            st["hess"] = UNKNOWN
        st["probe"] = NONE
        return _grow(st)
    return False


def _singular(st):
    """inertia_correction.py:136-198."""
    if st["hess"] == UNKNOWN or st["jac"] == UNKNOWN:
        p = st["probe"]
        if p == BARE:
            if st["jac"] == UNKNOWN:
                st["dc"] = _dc(st)
                st["probe"] = DC_ONLY
            else:
                if not _grow(st):
                    return False
                st["probe"] = DW_ONLY
        elif p == DC_ONLY:
            st["dc"] = 0.0
            if not _grow(st):
                return False
            st["probe"] = DW_ONLY
        elif p == DW_ONLY:
            st["dc"] = _dc(st)
            if not _grow(st):
                return False
            st["probe"] = BOTH
        elif p == BOTH:
            if not _grow(st):
                return False
    else:
        if st["dc"] > 0.0:
            if not _grow(st):
                return False
        else:
            st["dc"] = _dc(st)
    return True


def after_failure(st, outcome, has_constraints):
    """The branch of inertia_correction.py:342-388 for one failed
    factorization; False when the rules give up on this matrix."""
    if outcome == 1 and has_constraints:
        return _singular(st)
    if outcome == 2:
        if not st["raised"]:
            # inertia_correction.py:240-246: a better pivot tolerance first
            if st["pivtol"] < 0.1:
                st["pivtol"] = min(0.1, st["pivtol"] ** 0.5)
                st["raised"] = True
                return True
        return _singular(st)
    return _wrong_inertia(st)


def ic_trace(outcomes, mus, has_constraints=True):
    st = new_state()
    out = []
    live = False
    m = iter(mus)
    for o in outcomes:
        if o == -1:
            new_matrix(st, next(m))
            live = True
            continue
        if not live:
            continue
        out.append((st["dw"], st["dc"]))
        if o == 0:
            live = False
        elif not after_failure(st, o, has_constraints):
            live = False
    return out
