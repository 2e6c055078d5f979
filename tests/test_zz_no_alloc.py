"""amigo_ipm_reduce_noalloc (the reduction entry point that allocates
nothing, its workspace a static device array) against amigo_ipm_reduce
(caller's workspace) through the C ABI: every reduction kind, bit for bit,
on two calls in a row with the second problem smaller, so that the static
array holds the first call's partials when the second starts."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _problem(seed, n_p, n_c):
    from amigo_amd import _lib

    rng = np.random.default_rng(seed)
    n = n_p + n_c
    mult = np.zeros(n, bool)
    mult[rng.choice(n, n_c, replace=False)] = True
    pidx = np.nonzero(~mult)[0].astype(np.int32)
    cidx = np.nonzero(mult)[0].astype(np.int32)
    lbx = np.where(rng.random(n_p) < 0.6, rng.uniform(-2, 0, n_p), -np.inf)
    ubx = np.where(rng.random(n_p) < 0.6, rng.uniform(1, 3, n_p), np.inf)
    lbh = rng.normal(size=n_c)
    D = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    I = lambda a: torch.tensor(a, dtype=torch.int32, device="cuda")
    keep = {
        "pidx": I(pidx), "cidx": I(cidx), "lbx": D(lbx), "ubx": D(ubx), "lbh": D(lbh),
        "x": D(rng.normal(size=n)), "zl": D(np.where(np.isfinite(lbx), rng.uniform(0.1, 3, n_p), 0.0)),
        "zu": D(np.where(np.isfinite(ubx), rng.uniform(0.1, 3, n_p), 0.0)),
        "sl": D(rng.uniform(0.01, 2, n_p)), "su": D(rng.uniform(0.01, 2, n_p)),
        "a": D(rng.normal(size=n)), "b": D(rng.normal(size=n)), "c": D(rng.normal(size=n)),
    }
    P = _lib.IpmProblem(n_p, n_c, keep["pidx"].data_ptr(), keep["cidx"].data_ptr(),
                        keep["lbx"].data_ptr(), keep["ubx"].data_ptr(), keep["lbh"].data_ptr())
    S = _lib.IpmState(keep["x"].data_ptr(), keep["zl"].data_ptr(), keep["zu"].data_ptr(),
                      keep["sl"].data_ptr(), keep["su"].data_ptr())
    return P, S, keep


def _both(kind, P, S, keep, mu=0.07):
    from amigo_amd import _lib
    from amigo_amd.sparse import current_stream_ptr

    L = _lib.lib()
    s = _lib.vp(current_stream_ptr())
    ws = torch.zeros(L.amigo_ipm_workspace_doubles(), dtype=torch.float64, device="cuda")
    a, b, c = (_lib.dptr(keep[k]) for k in "abc")
    ref = (ctypes.c_double * 4)()
    _lib.check(L.amigo_ipm_reduce(kind, ctypes.byref(P), ctypes.byref(S), mu, a, b, c, _lib.dptr(ws),
                                  ref, s), "ipm")
    got = (ctypes.c_double * 4)(*([np.nan] * 4))
    _lib.check(L.amigo_ipm_reduce_noalloc(kind, ctypes.byref(P), ctypes.byref(S), mu, a, b, c, got, s),
               "ipm")
    return np.array(ref[:]), np.array(got[:])


def test_noalloc_reduce_equals_workspace_reduce_on_two_calls():
    big = _problem(0, 200000, 90000)      # 1024 partial blocks: the whole array
    small = _problem(1, 700, 300)         # 4 partial blocks: leftovers beyond them
    for kind in range(10):
        for P, S, keep in (big, small):
            ref, got = _both(kind, P, S, keep)
            # kinds fill 1 to 4 result slots; the others are zero in both
            assert np.array_equal(ref, got, equal_nan=True), (kind, ref, got)


def test_noalloc_reduce_rejects_bad_arguments():
    from amigo_amd import _lib
    from amigo_amd.sparse import current_stream_ptr

    L = _lib.lib()
    P, S, keep = _problem(2, 100, 40)
    out = (ctypes.c_double * 4)()
    s = _lib.vp(current_stream_ptr())
    assert L.amigo_ipm_reduce_noalloc(42, ctypes.byref(P), ctypes.byref(S), 0.0, None, None, None, out, s) == -1
    assert b"unknown reduction kind" in L.amigo_ipm_last_error()
    bad = _lib.IpmProblem(-5, 0, 0, 0, 0, 0, 0)
    assert L.amigo_ipm_reduce_noalloc(0, ctypes.byref(bad), ctypes.byref(S), 0.0, None, None, None, out, s) == -1
    assert b"n_primal + n_constraints" in L.amigo_ipm_last_error()
