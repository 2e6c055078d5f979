"""Model-level variable renumbering: ``reorder_model`` of the reference
(amigo/amigo.cpp:162-281), which orders a model's variables from the graph
of its elements before the KKT is built.

The element graph follows OrderingUtils::create_csr_from_element_conn
(include/ordering_utils.h:374-470) as reorder_model calls it: no diagonal,
sorted rows, no linear elements, and in an element with ncon constraint
nodes last, a constraint node is joined only to the element's other
(non-constraint) nodes.  The ordering is OrderingUtils::reorder
(ordering_utils.h:38-56) on that graph, or reorder_block (:75-104) with the
constraint indices ordered last.  AMD is the raw BlockAMD run on the graph
(bit-identical to the reference's); NATURAL is the identity; nested
dissection is this build's own (the reference's needs METIS).  Host-side
integer work only.
"""

import numpy as np

from . import _lib
from .sparse import _order_code

_ND, _AMD, _MC, _NAT = 0, 1, 2, 3


def element_graph(arrays, num_cons=None, nrows=None):
    """(rowp, cols) int32 of the element graph of element arrays (each
    E x npe node indices >= 0; num_cons[g] constraint nodes last)."""
    arrays = [np.asarray(a, np.int64) for a in arrays]
    if num_cons is None:
        num_cons = [0] * len(arrays)
    if len(num_cons) != len(arrays):
        raise ValueError("The constraints and input arrays must be the same size")
    for a in arrays:
        if a.ndim != 2:
            raise ValueError("Each input array must be 2D")
        if a.size and a.min() < 0:
            raise ValueError("element node indices must be >= 0")
    n = nrows if nrows is not None else 1 + max([int(a.max()) for a in arrays if a.size] + [-1])
    rows, cols = [], []
    for a, nc in zip(arrays, num_cons):
        E, npe = a.shape
        if E == 0:
            continue
        nc = int(nc)
        # First position of each entry's node in its element.
        same = a[:, :, None] == a[:, None, :]
        first = np.argmax(same, axis=2)  # (E, npe)
        for p in range(npe):
            ii = first[:, p]
            for q in range(npe):
                # A constraint row node (first position among the last nc)
                # is joined to non-constraint columns only.
                keep = np.ones(E, bool) if nc == 0 else ((ii < npe - nc) | (q < npe - nc))
                rows.append(a[keep, p])
                cols.append(a[keep, q])
    if rows:
        r = np.concatenate(rows)
        c = np.concatenate(cols)
        off = r != c
        key = np.unique(r[off] * n + c[off])
        r, c = key // n, key % n
    else:
        r = c = np.zeros(0, np.int64)
    rowp = np.zeros(n + 1, np.int64)
    np.add.at(rowp, r + 1, 1)
    return np.cumsum(rowp).astype(np.int32), c.astype(np.int32)


def reorder_model(order_type, arrays, num_cons=None, con_indices=None):
    """iperm of the model's variables (new index of each variable), as the
    reference's reorder_model returns it."""
    rowp, cols = element_graph(arrays, num_cons)
    n = len(rowp) - 1
    code = _order_code(order_type)
    perm = np.zeros(n, np.int32)
    if code == _AMD:
        mult = np.ascontiguousarray([] if con_indices is None else con_indices, np.int32)
        _lib.check(_lib.lib().amigo_block_amd(n, _lib.ip(rowp), _lib.ip(cols), int(mult.size),
                                              _lib.ip(mult if mult.size else np.zeros(1, np.int32)),
                                              _lib.ip(perm)))
    elif code == _MC:
        raise NotImplementedError("multi-color ordering is not part of this build")
    else:
        iperm0 = np.zeros(n, np.int32)
        if code == _NAT:
            perm[:] = np.arange(n, dtype=np.int32)
        else:
            _lib.check(_lib.lib().amigo_order(code, n, _lib.ip(rowp), _lib.ip(cols), _lib.ip(perm),
                                              _lib.ip(iperm0)))
        if con_indices is not None:
            # reorder_block: a stable partition with the constraints last.
            is_c = np.zeros(n, bool)
            ci = np.asarray(con_indices, np.int64)
            is_c[ci[(ci >= 0) & (ci < n)]] = True
            perm = np.concatenate([perm[~is_c[perm]], perm[is_c[perm]]]).astype(np.int32)
    iperm = np.empty(n, np.int32)
    iperm[perm] = np.arange(n, dtype=np.int32)
    return iperm
