"""Writes tests/golden/driver_regression.npz: the interior-point driver's
decision sequence on fixed workloads, recorded on the GPU so that a
rewrite of the driver can be checked to take the same steps.

    python tests/golden/make_driver_fixtures.py      # needs cuda:0

Per case: the iteration count, the inertia (n_pos, n_neg) of every
factorization, and per iteration the barrier parameter, step lengths,
line-search trials and inertia-correction delta; the solution for the two
smaller cases.  Cases: double integrator N = 256 (default options, AMD),
2-D elliptic control 32 x 32 (ND, equilibration) and eight seeded
cart-pole swing-ups (N ~ U{16..512}, ND, equilibration), all in FP64.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

CART_SEEDS = list(range(1000, 1008))


def cases():
    from amigo_amd.problems import cart_pole, double_integrator, elliptic

    yield "double_integrator_256", lambda: double_integrator.build(N=256), {}, True
    yield "elliptic2d_32", lambda: elliptic.EllipticControl(m=32, dim=2, seed=0), {"order": "ND", "scaling": True}, True
    for s in CART_SEEDS:
        yield f"cart_pole_{s}", (lambda s=s: cart_pole.seeded(s, nmin=16, nmax=512)), {"order": "ND", "scaling": True}, False


def record(make, kw, keep_x):
    from amigo_amd.optimizer import Optimizer

    prob = make()
    opt = Optimizer(prob, **kw)
    out = opt.optimize({"max_iterations": 200})
    its = out["iterations"]
    rec = {
        "converged": np.array(bool(out["converged"])),
        "num_iterations": np.array(out["num_iterations"]),
        "inertia": np.array(out["inertia_history"], dtype=np.int64).reshape(-1, 2),
        "mu": np.array([it["barrier_param"] for it in its]),
        "alpha_x": np.array([it["alpha_x"] for it in its]),
        "alpha_z": np.array([it["alpha_z"] for it in its]),
        "line_iters": np.array([it["line_iters"] for it in its], dtype=np.int64),
        "delta_w": np.array([it["inertia_delta"] for it in its]),
        "nlp_error": np.array([it["nlp_error"] for it in its]),
    }
    if keep_x:
        rec["x"] = opt.get_solution()
    return rec


def main():
    data = {}
    for name, make, kw, keep_x in cases():
        rec = record(make, kw, keep_x)
        print(f"{name}: {int(rec['num_iterations'])} iterations, {len(rec['inertia'])} factorizations, "
              f"converged {bool(rec['converged'])}", flush=True)
        for k, v in rec.items():
            data[f"{name}/{k}"] = v
    out = os.path.join(HERE, "driver_regression.npz")
    np.savez_compressed(out, **data)
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
