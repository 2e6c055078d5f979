"""Benchmark: interior-point solves/s on one MI355X.

Workload (default): 2-D semilinear elliptic optimal control on a 256 x 256
interior grid (131,072 primal variables, 65,536 equality constraints, KKT of
196,608 rows), alpha = 1e-4, 0 <= u <= 50, 16-mode seeded target (seed 0),
nested-dissection ordering, solved to the reference's default tolerances
(NLP error <= 1e-8).  One step = one full interior-point solve from the
initial point, with the problem, the symbolic factor and every buffer
resident in HBM.

    python bench.py [--steps K] [--warmup W] [--m 256] [--gpus 1]
                    [--precision fp64|mixed] [--no-cpu-baseline]

Prints one JSON line (see the repository's bench contract).  `value` is
solves/s over the K timed steps.  The roofline entry describes the kernel
class with the largest accumulated device time, timed with HIP events
recorded around every launch on the launch stream, over --timing-steps
further solves of the same workload (per-launch events cost host time, so
they are kept out of the solves/s steps).  The CPU baseline times the
reference's own SparseLDL (oracle/_ref, OpenMP + MKL) on a bounded sample of
the same KKT matrices.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--m", type=int, default=256, help="grid side of the elliptic workload")
    ap.add_argument("--dim", type=int, default=2, choices=[2, 3],
                    help="2-D (default; the headline workload) or 3-D elliptic control")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--order", default="ND")
    ap.add_argument("--no-scaling", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the other-precision and serving measurements")
    ap.add_argument("--forms", default="precision,serving,batching_slots,batching,serving_lockstep",
                    help="comma-separated forms to measure")
    ap.add_argument("--forms-only", action="store_true",
                    help="measure only the --forms other than precision (no headline solve; "
                         "for kernel profiles of one form)")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="MiB of one device copy run first (counter calibration)")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--batch-problems", type=int, default=48)
    ap.add_argument("--batch-slots", type=int, default=16)
    ap.add_argument("--lockstep-problems", type=int, default=2048,
                    help="cart-pole problems solved in one lockstep batch")
    ap.add_argument("--lockstep-sample", type=int, default=16,
                    help="of those, solved again one at a time (rate and identity check)")
    ap.add_argument("--serve-lockstep-requests", type=int, default=96,
                    help="requests per arrival rate of the lockstep-batching server")
    ap.add_argument("--serve-lockstep-rates", default="100,1000,5000")
    ap.add_argument("--serve-lockstep-seconds", type=float, default=0.0,
                    help="arrivals per rate: rate x seconds requests (0: --serve-lockstep-requests each)")
    ap.add_argument("--serve-lockstep-batches", type=int, default=4,
                    help="lockstep batches the server runs at once")
    ap.add_argument("--timing-steps", type=int, default=1,
                    help="solves timed per kernel class with HIP events (roofline)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: run with --gpus 1 "
                          "(multi-GPU sharding is a design note only)", "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    from amigo_amd.optimizer import Optimizer
    from amigo_amd.problems import elliptic

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.forms_only:
        args.forms = ",".join(f for f in args.forms.split(",") if f and f != "precision")
        print(json.dumps({"forms": forms(args, None, None, None, None)}))
        return
    prob = (elliptic.build_2d if args.dim == 2 else elliptic.build_3d)(m=args.m, seed=args.seed)
    opt = Optimizer(prob, order=args.order, precision=args.precision, scaling=not args.no_scaling)
    ldl = opt.solver.ldl
    opt.solver.keep_first = True

    def one_solve():
        out = opt.optimize({"record_objective": False})
        return out

    if args.copy_calibration > 0:
        # One plain device-to-device copy of known size, so that hardware
        # counters (FETCH_SIZE / WRITE_SIZE) read in the same run can be
        # scaled to bytes.
        nb = int(args.copy_calibration * 2**20) // 8
        src = torch.ones(nb, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy calibration: {8 * nb} bytes read and written", file=sys.stderr, flush=True)
        del src, dst
    for w in range(args.warmup):
        t = time.perf_counter()
        out = one_solve()
        torch.cuda.synchronize()
        print(f"[bench] warmup {w}: {time.perf_counter() - t:.2f} s, "
              f"{out['num_iterations']} iterations, converged={out['converged']}", file=sys.stderr, flush=True)

    lat = []
    outs = []
    n_fact0, n_solve0 = opt.solver.n_factor, opt.solver.n_solve
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        out = one_solve()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
        outs.append(out)
        print(f"[bench] step {len(lat) - 1}: {lat[-1]:.2f} s, {out['num_iterations']} iterations",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    nfact = (opt.solver.n_factor - n_fact0) / args.steps
    nsolve = (opt.solver.n_solve - n_solve0) / args.steps

    # Per-kernel-class device time: HIP events around every launch, on the
    # launch stream, over further solves of the same workload (the events
    # cost host time per launch, so they stay out of the solves/s steps).
    ldl.set_timing(True, reset=True)
    n_fact1 = opt.solver.n_factor
    for _ in range(args.timing_steps):
        one_solve()
    torch.cuda.synchronize()
    ldl.set_timing(False)
    nfact_t = (opt.solver.n_factor - n_fact1) / max(args.timing_steps, 1)

    times, (fl_small, fl_big) = ldl.timing()
    flops_total, solve_bytes = ldl.work()
    converged = all(o["converged"] for o in outs)
    last = outs[-1]["iterations"][-1]
    iters = outs[-1]["num_iterations"]

    # Dominant kernel class and its algorithmic work (the blocked-front
    # assembly is counted with the level kernels, which assemble the rest).
    times = dict(times)
    asm_ms, asm_n = times.pop("assemble_big")
    times["level"] = (times["level"][0] + asm_ms, times["level"][1])
    cls = max(times, key=lambda k: times[k][0])
    ms, launches = times[cls]
    ms = max(ms, 1e-9)
    if cls == "level":
        work_per_factor = fl_small
        per = work_per_factor * nfact_t * args.timing_steps
        achieved = per / (ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s"}
    elif cls in ("panel", "schur_update"):
        per = fl_big * nfact_t * args.timing_steps
        tot_ms = times["panel"][0] + times["schur_update"][0]
        achieved = per / (tot_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s"}
    else:
        nsolves = nsolve * args.timing_steps
        per = 0.5 * solve_bytes * nsolves
        achieved = per / (ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = _measured_traffic(args, cls)
    if roof["traffic"] is not None:
        roof["traffic_source"] = "profiles/elliptic2d_m256_fp64_traffic.json (copy-calibrated FETCH_SIZE + WRITE_SIZE)"
    roof["kernel"] = cls
    roof["kernel_ms_per_launch"] = ms / max(launches, 1)
    roof["launches"] = launches
    roof["measured_over"] = f"{args.timing_steps} further solve(s) with HIP events per launch"

    value = args.steps / total
    result = {
        "metric": "IPM solves/s + p50/p99 latency, mixed-prec vs FP64 on GPU at KKT resid <=1e-8",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "p50_ms": 1e3 * float(np.percentile(lat, 50)),
        "p99_ms": 1e3 * float(np.percentile(lat, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "precision_form": args.precision,
        "data": "synthetic (seeded 16-mode target field)",
        "config": {
            "workload": (f"{args.dim}-D semilinear elliptic control, "
                         f"{'x'.join([str(args.m)] * args.dim)} grid, alpha=1e-4, 0<=u<=50"),
            "kkt_rows": int(prob.n),
            "kkt_nnz": int(prob.nnz),
            "ordering": args.order,
            "equilibration": not args.no_scaling,
            "seed": args.seed,
            "parallelism": "single GPU",
        },
        "converged": converged,
        "ipm_iterations": iters,
        "final_nlp_error": last["nlp_error"],
        "factorizations_per_solve": nfact,
        "linear_solves_per_solve": nsolve,
        "kernel_ms": dict({k: v[0] for k, v in times.items()}, assemble_big=asm_ms),
        "roofline": roof,
    }
    if not args.no_forms:
        result["forms"] = forms(args, prob, opt, outs, lat)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(prob, opt, nfact, nsolve, args.cpu_threads)
    print(json.dumps(result))


def _measured_traffic(args, cls):
    """HBM bytes per launch of the dominant class, from the copy-calibrated
    FETCH_SIZE / WRITE_SIZE passes recorded for this workload in profiles/
    (None for a workload or class without a recorded measurement)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "elliptic2d_m256_fp64_traffic.json")
    if not (args.dim == 2 and args.m == 256 and args.precision == "fp64" and args.order == "ND"
            and not args.no_scaling and os.path.exists(path)):
        return None
    with open(path) as f:
        rec = json.load(f)
    return rec["bytes_per_launch"] if rec.get("kernel_class") == cls else None


def _inertia_of_last_solve(opt, before):
    return [tuple(i) for i in opt.solver.inertia_history[before:]]


def forms(args, prob, opt, outs, lat):
    """The other precision of the same workload, and the served and batched
    forms on seeded request mixes, each against the same requests solved one
    at a time in FP64."""
    res = {}
    want = set(f for f in args.forms.split(",") if f)
    if "precision" in want:
        res["precision"] = precision_form(args, prob, opt)
    if "serving" in want:
        res["serving"] = serving_slots(args)
    if "batching_slots" in want:
        res["batching_slots"] = batching_slots(args)
    if "batching" in want and args.lockstep_problems > 0:
        res["batching"] = lockstep_batching(args)
    if "serving_lockstep" in want and args.serve_lockstep_requests > 0:
        res["serving_lockstep"] = lockstep_serving(args)
    return res


def precision_form(args, prob, opt):
    """The other precision of the single large solve."""
    import numpy as np
    import torch

    from amigo_amd.optimizer import Optimizer

    other = "mixed" if args.precision == "fp64" else "fp64"
    h0 = len(opt.solver.inertia_history)
    opt.optimize({"record_objective": False})
    ref_inertia = _inertia_of_last_solve(opt, h0)
    x_ref = opt.get_solution()
    o2 = Optimizer(prob, order=args.order, precision=other, scaling=not args.no_scaling)
    o2.optimize({"record_objective": False})
    torch.cuda.synchronize()
    lat2 = []
    for _ in range(args.steps):
        h1 = len(o2.solver.inertia_history)
        t = time.perf_counter()
        out2 = o2.optimize({"record_objective": False})
        torch.cuda.synchronize()
        lat2.append(time.perf_counter() - t)
    x2 = o2.get_solution()
    out = {
        "form": other,
        "value": len(lat2) / sum(lat2),
        "p50_ms": 1e3 * float(np.percentile(lat2, 50)),
        "p99_ms": 1e3 * float(np.percentile(lat2, 99)),
        "ipm_iterations": out2["num_iterations"],
        "final_nlp_error": out2["iterations"][-1]["nlp_error"],
        "inertia_sequence_equal": _inertia_of_last_solve(o2, h1) == ref_inertia,
        "max_abs_diff_vs_" + args.precision: float(np.max(np.abs(x2 - x_ref))),
        "fp64_fallbacks": o2.solver.fp64_fallbacks,
    }
    print(f"[bench] {other}: {out['value']:.3f} solves/s", file=sys.stderr, flush=True)
    del o2
    return out


def serving_slots(args):
    """Slot serving: an open-loop Poisson stream of cart-pole and 2-D
    elliptic requests, several in flight on their own streams, against the
    same requests solved one at a time."""
    import numpy as np

    from amigo_amd.serve import SolveServer, request_mix, run_poisson

    specs = request_mix(args.serve_requests, seed=args.seed + 1)
    with SolveServer(slots=1) as srv:
        t = time.perf_counter()
        seq = [f.result() for f in [srv.submit(sp) for sp in specs]]
        seq_rate = len(seq) / (time.perf_counter() - t)
    with SolveServer(slots=args.serve_slots) as srv:
        st = run_poisson(srv, specs, rate=args.serve_rate, seed=args.seed + 2)
    same = all(np.array_equal(a["x"], b["x"]) for a, b in zip(seq, st["results"]))
    out = {
        "requests": len(specs),
        "mix": "70% cart-pole N~U{32..1024}, 30% 2-D elliptic side~U{32..192}",
        "arrival_rate_per_s": args.serve_rate,
        "slots": args.serve_slots,
        "value": st["solves_per_s"],
        "unit": "solves/s",
        "p50_ms": st["p50_ms"],
        "p99_ms": st["p99_ms"],
        "converged": st["converged"],
        "sequential_fp64_solves_per_s": seq_rate,
        "speedup_vs_sequential": st["solves_per_s"] / seq_rate,
        "results_identical_to_sequential": bool(same),
        "peak_inflight": st["peak_inflight"],
    }
    print(f"[bench] serving: {st['solves_per_s']:.2f} solves/s (sequential {seq_rate:.2f})",
          file=sys.stderr, flush=True)
    return out


def batching_slots(args):
    """Independent cart-pole swing-ups with N ~ U{16..512} and seeded masses
    and length, all queued at once, several in flight on slot streams."""
    import numpy as np

    from amigo_amd.serve import SolveServer, request_mix

    cp = request_mix(args.batch_problems, seed=args.seed + 3, cart_pole_frac=1.0, cart_n=(16, 512))
    with SolveServer(slots=1) as srv:
        t = time.perf_counter()
        seq = [f.result() for f in [srv.submit(sp) for sp in cp]]
        seq_rate = len(seq) / (time.perf_counter() - t)
    with SolveServer(slots=args.batch_slots) as srv:
        t = time.perf_counter()
        bat = [f.result() for f in [srv.submit(sp) for sp in cp]]
        bat_rate = len(bat) / (time.perf_counter() - t)
    out = {
        "problems": len(cp),
        "workload": "cart-pole swing-up, N~U{16..512}, T=2 s, mc~U(0.5,2), mp~U(0.1,0.5), l~U(0.3,1)",
        "slots": args.batch_slots,
        "value": bat_rate,
        "unit": "solves/s",
        "converged": sum(r["converged"] for r in bat),
        "sequential_fp64_solves_per_s": seq_rate,
        "speedup_vs_sequential": bat_rate / seq_rate,
        "results_identical_to_sequential": bool(all(np.array_equal(a["x"], b["x"]) for a, b in zip(seq, bat))),
    }
    print(f"[bench] batching (slots): {bat_rate:.2f} solves/s (sequential {seq_rate:.2f})", file=sys.stderr,
          flush=True)
    return out


def lockstep_serving(args):
    """The served form with fused launches: open-loop Poisson streams of the
    request mix (70% cart-pole swing-ups N ~ U{32..1024}, 30% 2-D elliptic
    problems side ~ U{32..192}) at each rate, gathered into lockstep batches
    as they arrive (BatchSolveServer, mixed precision), against the first
    requests of the same mix solved one at a time in plain FP64."""
    import numpy as np

    from amigo_amd.serve import BatchSolveServer, SolveServer, request_mix, run_poisson

    rates = [float(r) for r in args.serve_lockstep_rates.split(",") if r]
    per_rate = {r: (int(np.ceil(args.serve_lockstep_seconds * r)) if args.serve_lockstep_seconds > 0
                    else args.serve_lockstep_requests) for r in rates}
    n = max([args.serve_lockstep_requests] + list(per_rate.values()))
    specs = request_mix(n, seed=args.seed + 11)
    ns = min(16, n)
    with SolveServer(slots=1, precision="fp64") as srv:
        t = time.perf_counter()
        seq = [f.result() for f in [srv.submit(sp) for sp in specs[:ns]]]
        seq_rate = ns / (time.perf_counter() - t)
    # The same requests one at a time in the served precision: the bitwise
    # reference for the batched results.
    with SolveServer(slots=1, precision="mixed") as srv:
        seq_m = [f.result() for f in [srv.submit(sp) for sp in specs[:ns]]]
    out = {"requests_generated": n,
           "mix": "70% cart-pole N~U{32..1024}, 30% 2-D elliptic side~U{32..192}",
           "form": "lockstep batches of the requests queued while earlier batches run (one problem type per batch), "
                   f"up to {args.serve_lockstep_batches} batches at once on their own streams, mixed precision; "
                   "each result returned when its own problem finishes",
           "sequential_fp64_sample": ns, "sequential_fp64_solves_per_s": seq_rate, "rates": []}
    same = True
    for rate in rates:
        nr = per_rate[rate]
        print(f"[bench] lockstep serving: {nr} requests at {rate:.0f} req/s", file=sys.stderr, flush=True)
        with BatchSolveServer(max_batch=512, max_batches=args.serve_lockstep_batches) as srv:
            st = run_poisson(srv, specs[:nr], rate=rate, seed=args.seed + 12)
        sizes = [b for b, _ in srv.batches]
        mism = [(i, specs[i]["type"], bool(st["results"][i]["batched"]),
                 float(np.max(np.abs(a["x"] - b["x"]))), a["iterations"], b["iterations"])
                for i, (a, b) in enumerate(zip(seq_m, st["results"][:ns])) if not np.array_equal(a["x"], b["x"])]
        same &= not mism
        dmax = max(float(np.max(np.abs(a["x"] - b["x"]))) for a, b in zip(seq, st["results"][:ns]))
        out["rates"].append({
            "arrival_rate_per_s": rate,
            "requests": nr,
            "mean_arrival_span_s": nr / rate,
            "wall_s": float(st["wall_s"]),
            "value": st["solves_per_s"],
            "unit": "solves/s",
            "p50_ms": st["p50_ms"],
            "p99_ms": st["p99_ms"],
            "converged": st["converged"],
            "batches": len(sizes),
            "mean_batch_size": float(np.mean(sizes)) if sizes else 0.0,
            "speedup_vs_sequential_fp64": st["solves_per_s"] / seq_rate,
            "sample_max_abs_diff_vs_fp64": dmax,
            "sample_mismatches_vs_same_precision": mism,
            # Median parts of a request's latency: waiting for its batch to
            # start, building the batch (problems, union plan, buffers), and
            # solving until its own problem finished.
            "p50_queue_ms": 1e3 * float(np.median([r["queue_s"] for r in st["results"]])),
            "p50_build_ms": 1e3 * float(np.median([r["build_s"] for r in st["results"]])),
            "p50_solve_ms": 1e3 * float(np.median([r["solve_s"] for r in st["results"]])),
        })
        print(f"[bench] lockstep serving at {rate:.0f} req/s: {st['solves_per_s']:.1f} solves/s, "
              f"p50 {st['p50_ms']:.0f} ms, p99 {st['p99_ms']:.0f} ms, {len(sizes)} batches",
              file=sys.stderr, flush=True)
    out["sample_identical_to_sequential_same_precision"] = bool(same)
    return out


def lockstep_batching(args):
    """The batching form: K seeded cart-pole swing-ups (N ~ U{16..512},
    T = 2 s, seeded masses and length) solved in lockstep as one
    block-diagonal NLP, every elimination-tree level of every problem in the
    same launches, in the mixed-precision form; a sample of the same
    problems is then solved one at a time in plain FP64 for the rate and
    a bitwise check of iterations, inertia sequence and solution."""
    import numpy as np
    import torch

    from amigo_amd.batch import BatchOptimizer
    from amigo_amd.optimizer import Optimizer
    from amigo_amd.problems import cart_pole

    K = args.lockstep_problems
    t = time.perf_counter()
    probs = [cart_pole.seeded(args.seed + 1000 + s, nmin=16, nmax=512) for s in range(K)]
    t_build = time.perf_counter() - t
    t = time.perf_counter()
    bo = BatchOptimizer(probs, order="ND", scaling=True, precision="mixed")
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t
    opts = {"record_objective": False}
    t = time.perf_counter()
    out = bo.optimize(opts)
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t
    xs = bo.get_solution()
    its = np.array([o["num_iterations"] for o in out])
    print(f"[bench] lockstep batch of {K}: {t_solve:.2f} s ({K / t_solve:.1f} solves/s), setup {t_setup:.1f} s",
          file=sys.stderr, flush=True)
    ns = min(args.lockstep_sample, K)
    same = True
    # The same problems one at a time in plain FP64: construction and solve
    # timed apart, so that both sides are compared with and without it.
    t = time.perf_counter()
    singles = [Optimizer(probs[k], order="ND", scaling=True, precision="fp64") for k in range(ns)]
    torch.cuda.synchronize()
    t_seq_setup = time.perf_counter() - t
    t = time.perf_counter()
    for k in range(ns):
        d = singles[k].optimize(opts)
        same &= (d["num_iterations"] == out[k]["num_iterations"]
                 and d["inertia_history"] == out[k]["inertia_history"]
                 and bool(np.array_equal(singles[k].get_solution(), xs[k])))
    torch.cuda.synchronize()
    t_seq_solve = time.perf_counter() - t
    del singles
    rate = K / t_solve
    seq_rate = ns / t_seq_solve
    rate_incl = K / (t_setup + t_solve)
    seq_rate_incl = ns / (t_seq_setup + t_seq_solve)
    restored = [int(k) for k in np.nonzero(bo.restorations > 0)[0]]
    return {
        "problems": K,
        "workload": "cart-pole swing-up, N~U{16..512}, T=2 s, mc~U(0.5,2), mp~U(0.1,0.5), l~U(0.3,1)",
        "form": "lockstep batch: one block-diagonal KKT, one level-kernel launch per elimination-tree level "
                "for all problems, per-problem inertia correction / line search / refinement / restoration "
                "in the native driver, mixed precision",
        "value": rate,
        "unit": "solves/s",
        "value_with_setup": rate_incl,
        "wall_s": t_solve,
        "setup_s": t_setup,
        "problem_build_s": t_build,
        "converged": int(sum(o["converged"] for o in out)),
        "iterations_mean": float(its.mean()),
        "iterations_max": int(its.max()),
        "not_converged": [int(k) for k, o in enumerate(out) if not o["converged"]],
        "lockstep_iterations": int(bo.iterations_run),
        "restored_in_batch": restored,
        "batch_factorizations": int(bo.native.counts()[0]),
        "kkt_rows_total": int(bo.bp.n),
        "fp64_fallbacks": int(bo.fp64_fallbacks),
        "host_phase_s": {k: round(v, 4) for k, v in bo.phase_times.items()},
        "sequential_fp64_sample": ns,
        "sequential_fp64_solves_per_s": seq_rate,
        "sequential_fp64_solves_per_s_with_setup": seq_rate_incl,
        "speedup_vs_sequential_fp64": rate / seq_rate,
        "speedup_vs_sequential_fp64_with_setup": rate_incl / seq_rate_incl,
        "sample_identical_to_sequential_fp64": bool(same),
    }


def _ruiz(rowp, cols, vals, passes=5):
    """Symmetric Ruiz equilibration of a CSR matrix: passes of
    s_i <- s_i / sqrt(max_j |s_i a_ij s_j|) (the GPU path's equilibration).
    Returns the scaled values."""
    import numpy as np

    n = len(rowp) - 1
    rows = np.repeat(np.arange(n), np.diff(rowp))
    s = np.ones(n)
    for _ in range(passes):
        a = np.abs(vals) * s[rows] * s[cols]
        rmax = np.zeros(n)
        np.maximum.at(rmax, rows, a)
        rmax[rmax == 0.0] = 1.0
        s = s / np.sqrt(rmax)
    return vals * s[rows] * s[cols]


def cpu_baseline(prob, opt, nfact, nsolve, threads):
    """The reference's own SparseLDL (oracle/_ref: OpenMP + MKL) timed on a
    bounded sample: factorizations and solves of the KKT matrix of the first
    IPM iterate, scaled by the GPU run's factorizations and solves per IPM
    solve.  Two configurations of the reference are timed:
    - its default ordering (AMD, what am.Optimizer uses without METIS);
    - the GPU path's own ordering and equilibration: the matrix permuted by
      the GPU plan's nested-dissection permutation and Ruiz-scaled, factored
      by the reference in NATURAL order, so that it forms the same fronts.
    The faster of the two is the reported value.  (Later iterates are not
    sampled: there the AMD configuration's delayed-pivot cascade grows a
    front past 46,340 rows, where the reference's int front_size *
    front_size workspace size overflows.)"""
    import numpy as np
    import torch
    from scipy.sparse import csr_matrix

    try:
        from oracle import ref
    except Exception as e:  # pragma: no cover
        return {"error": f"oracle unavailable: {e}"}
    if not ref.available():
        return {"error": "oracle/_ref not built"}
    ncpu = min(threads, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(ncpu)
    os.environ["MKL_NUM_THREADS"] = str(ncpu)
    lib = ref.lib()
    lib.ref_set_num_threads(ncpu)
    vals, shift = (t.clone() for t in opt.solver.first_values)  # the first Newton system
    diag = torch.as_tensor(prob.diag_idx, device=vals.device)
    vals[diag] += shift
    A = vals.cpu().numpy()
    b = np.ones(prob.n)

    def timed(rowp, cols, data, order):
        R = ref.RefLDL(rowp, cols, order=order)
        t = time.perf_counter()
        R.factor(data)
        tf = time.perf_counter() - t
        t = time.perf_counter()
        R.solve(b)
        R.solve(b)
        ts = (time.perf_counter() - t) / 2
        return tf, ts, R.inertia()

    tf_amd, ts_amd, in_amd = timed(prob.rowp, prob.cols, A, "AMD")
    # The GPU plan's ordering and equilibration.
    perm = opt.solver.ldl.get_perm()
    M = csr_matrix((_ruiz(np.asarray(prob.rowp), np.asarray(prob.cols), A),
                    np.asarray(prob.cols), np.asarray(prob.rowp)), shape=(prob.n, prob.n))
    M = M[perm][:, perm].tocsr()
    M.sort_indices()
    tf_nd, ts_nd, in_nd = timed(M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data,
                                "NATURAL")
    rate_amd = 1.0 / (nfact * tf_amd + nsolve * ts_amd)
    rate_nd = 1.0 / (nfact * tf_nd + nsolve * ts_nd)
    return {
        "value": max(rate_amd, rate_nd),
        "unit": "solves/s",
        "cores": ncpu,
        "kind": "reference",
        "sample": (f"reference SparseLDL (ustab=0.01, OpenMP+MKL, {ncpu} threads) on the KKT of the first "
                   f"IPM iterate, scaled by {nfact:.1f} factorizations + {nsolve:.1f} solves per IPM solve, "
                   "IPM vector work not counted; the faster of: its default AMD ordering "
                   f"(factorization {tf_amd*1e3:.1f} ms, solve {ts_amd*1e3:.2f} ms, {rate_amd:.3f} solves/s) "
                   "and the GPU plan's ND permutation with Ruiz equilibration in NATURAL order "
                   f"(factorization {tf_nd*1e3:.1f} ms, solve {ts_nd*1e3:.2f} ms, {rate_nd:.3f} solves/s)"),
        "configurations": {
            "amd": {"factor_ms": tf_amd * 1e3, "solve_ms": ts_amd * 1e3, "solves_per_s": rate_amd,
                    "inertia": list(in_amd)},
            "gpu_ordering_scaled": {"factor_ms": tf_nd * 1e3, "solve_ms": ts_nd * 1e3,
                                    "solves_per_s": rate_nd, "inertia": list(in_nd)},
        },
    }

if __name__ == "__main__":
    main()
