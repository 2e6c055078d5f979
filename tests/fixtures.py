"""Loader for the committed reference fixtures (tests/golden/*.npz)."""

import os

import numpy as np
from scipy.sparse import csr_matrix

HERE = os.path.dirname(os.path.abspath(__file__))
_cache = {}


def ldl():
    if "ldl" not in _cache:
        with np.load(os.path.join(HERE, "golden", "ldl_fixtures.npz")) as z:
            _cache["ldl"] = {k: z[k] for k in z.files}
    return _cache["ldl"]


def ldl_cases():
    return sorted({k.split("/")[0] for k in ldl()})


def matrix(name):
    d = ldl()
    rowp, cols, vals = d[f"{name}/rowp"], d[f"{name}/cols"], d[f"{name}/vals"]
    n = len(rowp) - 1
    return csr_matrix((vals, cols, rowp), shape=(n, n)), float(d[f"{name}/ustab"])


def field(name, order, key):
    return ldl()[f"{name}/{order}/{key}"]


SYMBOLIC_KEYS = ["node_ptr", "node_to_var", "children_ptr", "children", "contrib_ptr", "contrib_rows", "invperm"]
