"""Serving loop: NLP solves of different sizes, taken as they arrive.

    server = SolveServer(slots=8, precision="fp64")
    fut = server.submit({"type": "cart_pole", "N": 128, "seed": 3})
    res = fut.result()          # {"x", "converged", "iterations", "latency_s", ...}
    server.close()

    stats = run_poisson(server, specs, rate=200.0, seed=0)   # p50 / p99 / solves/s

Each slot is a worker thread with its own HIP stream.  A request is built
(problem, symbolic analysis of its KKT) on the host, then admitted once the
predicted device memory of its factorization (SparseLDL.memory_bytes) fits
the budget next to the solves already in flight; it keeps its iterate,
symbolic factor and fronts resident in HBM until it completes.  Requests
that arrive while others run are admitted as slots and memory free up, and
their kernels overlap on the device.  Every kernel is deterministic (no
floating-point atomics), so a request's result does not depend on what else
runs beside it.

The reference has no serving layer; its `am.Optimizer.optimize`
(amigo/algorithm/ipm_driver.py:62-483) is what each slot runs.
"""

import concurrent.futures as cf
import queue
import threading
import time

import numpy as np
import torch

from .optimizer import Optimizer


def build_problem(spec, device="cuda"):
    """Problem of a request spec: {"type": "double_integrator" | "cart_pole" |
    "elliptic", sizes, seed}."""
    kind = spec["type"]
    if kind == "double_integrator":
        from .problems import double_integrator

        return double_integrator.build(N=int(spec.get("N", 256)), device=device)
    if kind == "cart_pole":
        from .problems import cart_pole

        return cart_pole.seeded(int(spec.get("seed", 0)), N=int(spec["N"]) if "N" in spec else None,
                                device=device)
    if kind == "elliptic":
        from .problems import elliptic

        return elliptic.EllipticControl(m=int(spec.get("side", 64)), dim=int(spec.get("dim", 2)),
                                        seed=int(spec.get("seed", 0)), device=device)
    raise ValueError(f"unknown problem type {kind!r}")


def _order_for(spec):
    return spec.get("order", "ND")


class SolveServer:
    """Continuous-batching solve server on one GPU (see the module doc)."""

    def __init__(self, slots=8, memory_budget=None, precision="fp64", options=None,
                 device="cuda", scaling=True):
        self.device = torch.device(device)
        self.slots = int(slots)
        if memory_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            memory_budget = int(0.8 * free)
        self.budget = int(memory_budget)
        self.precision = precision
        self.options = dict(options or {})
        self.scaling = scaling
        self._q = queue.Queue()
        self._cv = threading.Condition()
        self._inflight = 0
        self._inflight_bytes = 0
        self.peak_inflight = 0
        self.peak_bytes = 0
        self._closed = False
        self._threads = [threading.Thread(target=self._worker, daemon=True, name=f"amigo-slot-{i}")
                         for i in range(self.slots)]
        for t in self._threads:
            t.start()

    # ---- public -----------------------------------------------------------
    def submit(self, spec, arrival=None):
        fut = cf.Future()
        self._q.put((spec, fut, time.perf_counter() if arrival is None else arrival))
        return fut

    def close(self):
        if self._closed:
            return
        self._closed = True
        for _ in self._threads:
            self._q.put(None)
        for t in self._threads:
            t.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- admission ----------------------------------------------------------
    def _admit(self, need):
        with self._cv:
            # A request larger than the whole budget runs alone.
            while self._inflight > 0 and self._inflight_bytes + need > self.budget:
                self._cv.wait()
            self._inflight += 1
            self._inflight_bytes += need
            self.peak_inflight = max(self.peak_inflight, self._inflight)
            self.peak_bytes = max(self.peak_bytes, self._inflight_bytes)

    def _release(self, need):
        with self._cv:
            self._inflight -= 1
            self._inflight_bytes -= need
            self._cv.notify_all()

    # ---- worker -------------------------------------------------------------
    def _worker(self):
        stream = torch.cuda.Stream(self.device)
        with torch.cuda.device(self.device), torch.cuda.stream(stream):
            while True:
                item = self._q.get()
                if item is None:
                    return
                spec, fut, t_arr = item
                t_start = time.perf_counter()
                try:
                    prob = build_problem(spec, self.device)
                    opt = Optimizer(prob, order=_order_for(spec), precision=self.precision,
                                    scaling=spec.get("scaling", self.scaling))
                    # No HIP-graph capture beside other slots: a stream capture
                    # makes the other threads' legacy-stream copies fail.
                    opt.solver.ldl.set_graphs(False)
                    need = opt.solver.ldl.memory_bytes() + 8 * (prob.nnz + 24 * prob.n)
                    self._admit(need)
                    t_admit = time.perf_counter()
                    try:
                        out = opt.optimize(dict(self.options, **spec.get("options", {})))
                        x = opt.get_solution()
                        stream.synchronize()
                    finally:
                        self._release(need)
                    t_done = time.perf_counter()
                    fut.set_result({
                        "spec": spec,
                        "x": x,
                        "converged": bool(out["converged"]),
                        "iterations": int(out["num_iterations"]),
                        "nlp_error": float(out["iterations"][-1]["nlp_error"]),
                        "inertia": list(opt.solver.inertia_history),
                        "predicted_bytes": int(need),
                        "latency_s": t_done - t_arr,
                        "queue_s": t_admit - t_arr,
                        "build_s": t_admit - t_start,
                        "solve_s": t_done - t_admit,
                    })
                except Exception as e:  # the error goes to the caller
                    fut.set_exception(e)


def request_mix(n, seed=0, cart_pole_frac=0.7, cart_n=(32, 1024), side=(32, 192)):
    """Seeded request specs: cart-pole swing-ups with N ~ U{cart_n} and 2-D
    elliptic control problems with side ~ U{side}."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        if rng.random() < cart_pole_frac:
            specs.append({"type": "cart_pole", "N": int(rng.integers(cart_n[0], cart_n[1] + 1)),
                          "seed": int(rng.integers(0, 2**31 - 1))})
        else:
            specs.append({"type": "elliptic", "side": int(rng.integers(side[0], side[1] + 1)),
                          "seed": int(rng.integers(0, 2**31 - 1))})
    return specs


def run_poisson(server, specs, rate, seed=0):
    """Open-loop Poisson stream: request i arrives at the i-th event of a
    rate-`rate` Poisson process (from the start of the call) and is queued
    then.  Returns per-request results and the p50 / p99 latency and solves/s
    of the stream."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(specs)))
    t0 = time.perf_counter()
    futs = []
    for spec, a in zip(specs, arrivals):
        now = time.perf_counter() - t0
        if a > now:
            time.sleep(a - now)
        futs.append(server.submit(spec, arrival=t0 + a))
    results = [f.result() for f in futs]
    wall = time.perf_counter() - t0
    lat = np.array([r["latency_s"] for r in results])
    return {
        "results": results,
        "solves_per_s": len(results) / wall,
        "p50_ms": 1e3 * float(np.percentile(lat, 50)),
        "p99_ms": 1e3 * float(np.percentile(lat, 99)),
        "wall_s": wall,
        "converged": sum(r["converged"] for r in results),
        "peak_inflight": server.peak_inflight,
    }
