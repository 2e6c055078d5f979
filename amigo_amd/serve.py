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
    "elliptic", sizes, seed}, or {"problem": document} with a problem
    document of amigo_amd.wire (a dict or its JSON text)."""
    if "problem" in spec:
        from . import wire

        doc = spec["problem"]
        return wire.loads(doc, device=device) if isinstance(doc, str) else wire.problem_from_dict(doc, device=device)
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


class BatchSolveServer:
    """Continuous batching with fused launches: requests queued while earlier
    batches run are gathered into the next lockstep batch
    (batch.BatchOptimizer), whose problems share every level-kernel launch
    of the KKT factorization and every IPM kernel launch.  Up to
    `max_batches` batches run at once, each on its own HIP stream, so a new
    batch starts while an older one finishes its slowest problems.  A batch
    takes at most `max_batch` requests, by default of one problem type
    (group_by_type), and is admitted by predicted device
    memory (the sum of its problems' SparseLDL.memory_bytes and vectors)
    against the budget.  Each request's result is returned when its own
    problem finishes, not when its batch does, and equals the request
    solved alone bit for bit.

        with BatchSolveServer(max_batch=512) as srv:
            fut = srv.submit({"type": "cart_pole", "N": 256, "seed": 7})
            res = fut.result()
    """

    def __init__(self, max_batch=512, max_batches=2, gather_s=0.005, memory_budget=None,
                 precision="mixed", options=None, device="cuda", scaling=True, order="ND",
                 group_by_type=True):
        self.device = torch.device(device)
        if memory_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            memory_budget = int(0.8 * free)
        self.budget = int(memory_budget)
        self.max_batch = int(max_batch)
        self.gather_s = float(gather_s)
        self.precision = precision
        self.options = dict(options or {})
        self.scaling = scaling
        self.order = order
        self.group_by_type = bool(group_by_type)
        self._q = queue.Queue()
        self._cv = threading.Condition()
        self._inflight = 0
        self._inflight_bytes = 0
        self._free_slots = int(max_batches)
        self.peak_inflight = 0
        self.peak_bytes = 0
        self.batches = []  # (size, wall seconds) of every batch
        self.batch_specs = []
        self._closed = False
        self._workers = []
        self._disp = threading.Thread(target=self._dispatch, daemon=True, name="amigo-batcher")
        self._disp.start()

    def submit(self, spec, arrival=None):
        fut = cf.Future()
        self._q.put((spec, fut, time.perf_counter() if arrival is None else arrival))
        return fut

    def close(self):
        if self._closed:
            return
        self._closed = True
        self._q.put(None)
        self._disp.join()
        for t in list(self._workers):
            t.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- dispatcher: gather queued requests into batches --------------------
    def _dispatch(self):
        pending = []
        stop = False
        while not stop or pending:
            if not pending:
                item = self._q.get()
                if item is None:
                    stop = True
                    continue
                pending.append(item)
            # Gather what arrives within gather_s (or until the batch is full).
            t_end = time.perf_counter() + self.gather_s
            while len(pending) < self.max_batch:
                try:
                    item = self._q.get(timeout=max(0.0, t_end - time.perf_counter()))
                except queue.Empty:
                    break
                if item is None:
                    stop = True
                    break
                pending.append(item)
            with self._cv:
                while self._free_slots == 0:
                    self._cv.wait()
                self._free_slots -= 1
            # Requests that arrived while waiting for a slot join this batch.
            while len(pending) < self.max_batch:
                try:
                    item = self._q.get_nowait()
                except queue.Empty:
                    break
                if item is None:
                    stop = True
                    break
                pending.append(item)
            batch, pending = self._take_batch(pending)
            t = threading.Thread(target=self._run_batch, args=(batch,), daemon=True)
            self._workers = [w for w in self._workers if w.is_alive()] + [t]
            t.start()

    def _take_batch(self, pending):
        """The next batch from the queued requests, in arrival order.  With
        group_by_type, only requests of the oldest request's problem type:
        a lockstep factorization takes as many level launches as its
        deepest elimination tree, so a trajectory (a few levels) batched
        with a grid problem (tens of levels) would pay for the grid's
        depth at every iteration."""
        if not self.group_by_type:
            return pending[: self.max_batch], pending[self.max_batch:]
        kind = pending[0][0].get("type")
        take, rest = [], []
        for item in pending:
            if len(take) < self.max_batch and item[0].get("type") == kind:
                take.append(item)
            else:
                rest.append(item)
        return take, rest

    # ---- one batch -----------------------------------------------------------
    def _run_batch(self, batch):
        from .batch import BatchOptimizer, problem_ldl

        stream = torch.cuda.Stream(self.device)
        need = 0
        admitted = False
        try:
            with torch.cuda.device(self.device), torch.cuda.stream(stream):
                t_start = time.perf_counter()
                probs = [build_problem(spec, self.device) for spec, _, _ in batch]
                need = sum(problem_ldl(p, self.order).memory_bytes() + 8 * (p.nnz + 24 * p.n)
                           for p in probs)
                self._admit(len(batch), need)
                admitted = True
                t_admit = time.perf_counter()
                bo = BatchOptimizer(probs, order=self.order, scaling=self.scaling,
                                    precision=self.precision)
                # No HIP-graph capture beside other batches' streams.
                bo.solver.ldl.set_graphs(False)

                def done(k, out, x):
                    spec, fut, t_arr = batch[k]
                    t_done = time.perf_counter()
                    fut.set_result({
                        "spec": spec,
                        "x": x,
                        "converged": bool(out["converged"]),
                        "iterations": int(out["num_iterations"]),
                        "nlp_error": float(out["iterations"][-1]["nlp_error"]),
                        "inertia": list(out.get("solver_inertia_history", out["inertia_history"])),
                        "batched": bool(out.get("batched", True)),
                        "batch_size": len(batch),
                        "latency_s": t_done - t_arr,
                        "queue_s": t_admit - t_arr,
                        "build_s": t_admit - t_start,
                        "solve_s": t_done - t_admit,
                    })

                bo.optimize(dict(self.options), on_done=done)
                stream.synchronize()
                self.batches.append((len(batch), time.perf_counter() - t_start))
                self.batch_specs.append([spec for spec, _, _ in batch])
        except Exception as e:  # every unfinished request gets the error
            for _, fut, _ in batch:
                if not fut.done():
                    fut.set_exception(e)
        finally:
            with self._cv:
                if admitted:
                    self._inflight -= len(batch)
                    self._inflight_bytes -= need
                self._free_slots += 1
                self._cv.notify_all()

    def _admit(self, count, need):
        with self._cv:
            while self._inflight > 0 and self._inflight_bytes + need > self.budget:
                self._cv.wait()
            self._inflight += count
            self._inflight_bytes += need
            self.peak_inflight = max(self.peak_inflight, self._inflight)
            self.peak_bytes = max(self.peak_bytes, self._inflight_bytes)


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
