#!/usr/bin/env python3
"""Benchmark: time to FP64 relative residual 1e-8 of MAS-PCG on one MI355X.

Workload (default): 64^3 Q1-hex linear elasticity, per-element E = 10^U(3,6) (seed 0), nu = 0.3,
bottom face fixed, gravity RHS — 823,875 DOF, 7,081,535 3x3 blocks — and the 200-step sequence
of such systems whose moduli drift by a 5% multiplicative random walk (step 0 is the system
above).  Right-hand sides and solutions stay in HBM (``psmas_solve_device``); factorisations are
done before the timed solves and reported separately.

Paths:
  * fp64    — the reference algorithm: FP64 MAS-PCG, true residual every 4th iteration,
              convergence checked every 10th, relative tolerance 1e-8;
  * mixed   — FP32 inner PCG with FP16 bank inverses and FP64 reliable updates; converged only
              when the FP64 true residual ||b - A x|| <= 1e-8 ||b||;
  * learned — the mixed path plus the learned lattice coarse correction (flexible PCG, plain MAS
              as fallback), its network trained on pairs from the solves of steps 1..50.

A "step" is one solve of one held-out system of the sequence: ``--steps K`` systems spread over
steps 51..199, each refactorised (values only), solved ``--warmup`` times untimed (first system)
or once (the others), then solved once timed between device synchronisations by each of the
three paths.  The headline ``value`` is the learned path's mean ms over the K held-out systems;
the mixed and FP64 means on the same systems are reported beside it, with iterations, fallbacks
and true residuals per step, and ablations of the network on the first held-out steps (the exact
lattice coarse solve it approximates, and a network with only its linear skip layer).

The plain paths are also timed on step 0 itself, where the kernels are timed with HIP events
(the ``roofline`` object) and the solutions are compared with an FP64 direct solve (CPU sparse
Cholesky, the CHOLMOD stand-in) in ``accuracy``.  ``cpu_baseline`` is the reference's CPU
iterative solver (Eigen's Jacobi-preconditioned CG, restated with OpenMP) run to 1e-8 on this
machine's host cores; the direct solver's times are beside it.  ``multi_rhs`` reports
``--multi-rhs`` (default 4) right-hand sides of the step-0 matrix solved together.

One JSON line is printed (driver contract).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_info() -> str:
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def make_problem(config: int):
    from polysolve_amd import problems

    if config == 1:
        return problems.config1()
    if config == 2:
        return problems.config2(seed=0)
    if config == 3:
        return problems.config3(seed=0)
    raise SystemExit(f"config {config} is not a single-system bench workload")


def run_path(solver, b_dev, x_dev, n, steps, warmup, sync):
    for _ in range(warmup):
        solver.solve_device(b_dev.ptr, x_dev.ptr, n)
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.solve_device(b_dev.ptr, x_dev.ptr, n)
    sync()
    return (time.perf_counter() - t0) * 1e3 / max(steps, 1)


# HBM bytes per launch measured with the FETCH_SIZE and WRITE_SIZE counters (rocprofv3 --pmc, one
# counter per run) on the 64^3 workload, each scaled by the reading for a 256 MiB device copy
# (FETCH_SIZE reads 0.500x the bytes such a copy reads, WRITE_SIZE 1.000x).  Keyed by
# (config, kernel); re-measure when a kernel's memory access changes.
MEASURED_TRAFFIC = {
    (2, "spmv_fp32"): 215.0e6 + 3.3e6,  # the symmetric form (k_spmv_sym<3, 4, float, float>)
}


def measured_read_gbs(mib: int = 1024, reps: int = 5) -> float:
    """Streaming-read ceiling of this GPU (psmas_read_bandwidth): `mib` MiB (beyond the 256 MB
    Infinity Cache) read with 16-byte nontemporal loads, best of `reps` — the measured peak
    beside the specification's 8 TB/s."""
    import ctypes
    from polysolve_amd import _native

    g = ctypes.c_double(0.0)
    rc = _native.lib().psmas_read_bandwidth(ctypes.c_longlong(mib << 20), reps, ctypes.byref(g))
    return g.value if rc == 0 else float("nan")


def json_number(v: float, digits: int):
    """v rounded, or None where it is not finite (JSON has no NaN)."""
    return round(v, digits) if np.isfinite(v) else None


def bench_batch(args) -> int:
    """64 uneven Q1-hex systems (edges U{8..32}, seeded moduli) solved as one batch: one launch
    sequence per PCG phase for all systems, graph-captured, tuned by the autotuner; compared with
    solving each system alone with the same kernels, and with the FP64 path."""
    from polysolve_amd import MASSolver, problems
    from polysolve_amd.linear import DeviceArray, synchronize

    t = time.perf_counter()
    probs = problems.config4(seed=0)
    ntot = sum(p.n for p in probs)
    log(f"assembled 64 systems, sum n = {ntot} ({time.perf_counter() - t:.1f}s)")
    base = {"block_dim": 3, "relative_tolerance": 1e-8, "absolute_tolerance": 0.0, "max_iter": args.max_iter,
            "precision": "mixed", "coarse_solver": args.coarse, "spmv_symmetric": args.spmv == "symmetric"}
    bdev = [DeviceArray(p.n, p.b) for p in probs]
    xdev = [DeviceArray(p.n, np.zeros(p.n)) for p in probs]
    bp, xp = [b.ptr for b in bdev], [x.ptr for x in xdev]

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        synchronize()
        return (time.perf_counter() - t0) * 1e3 / max(args.steps, 1)

    s = MASSolver()
    s.set_parameters({"MAS": base})
    t = time.perf_counter()
    s.factorize_batch([p.A for p in probs], coords=[p.coords for p in probs])
    setup_ms = (time.perf_counter() - t) * 1e3
    s.solve_batch_device(bp, xp)
    t = time.perf_counter()
    tune = s.autotune(reps=20)
    tune_s = time.perf_counter() - t
    log(f"autotune ({tune_s:.1f}s): {tune}")
    batch_ms = timed(lambda: s.solve_batch_device(bp, xp))
    infos = [s.batch_info(i) for i in range(len(probs))]
    rels = [float(np.linalg.norm(p.b - p.A @ x.download()) / np.linalg.norm(p.b)) for p, x in zip(probs, xdev)]
    iters = [i["solver_iter"] for i in infos]
    log(f"batched mixed: {batch_ms:.2f} ms for 64 systems, iterations {min(iters)}..{max(iters)}, "
        f"max rel res {max(rels):.2e}")
    s.set_parameters({"MAS": {"precision": "fp64"}})
    batch64_ms = timed(lambda: s.solve_batch_device(bp, xp))
    log(f"batched fp64: {batch64_ms:.2f} ms")
    del s

    # Each system alone, same kernels, the tuned parameters of the batch.
    singles = []
    for p in probs:
        o = MASSolver()
        o.set_parameters({"MAS": dict(base, spmv_unroll=tune["spmv_unroll"], check_interval=tune["check_interval"],
                                      coarse_solver=tune["coarse_solver"])})
        o.set_coordinates(p.coords)
        o.factorize(p.A)
        singles.append(o)

    def solve_each():
        for o, p, b, x in zip(singles, probs, bdev, xdev):
            o.solve_device(b.ptr, x.ptr, p.n)

    single_ms = timed(solve_each)
    log(f"one by one: {single_ms:.2f} ms")
    line = {
        "metric": "ms to 1e-8 rel. residual + PCG iters: learned/mixed MAS-PCG vs plain FP64",
        "value": round(batch_ms, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(batch_ms, 3), "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "fp32+fp16 inner / fp64 refinement", "data": "synthetic (seeded Q1-hex elasticity assembled in-repo)",
        "config": {"workload": f"batch of 64 Q1-hex elasticity systems, edges U{{8..32}} seed 0, sum n = {ntot}",
                   "path": "batched mixed MAS-PCG, autotuned"},
        "iters_min": min(iters), "iters_max": max(iters), "max_true_rel_residual": max(rels),
        "one_by_one_ms": round(single_ms, 3), "batched_fp64_ms": round(batch64_ms, 3),
        "speedup_vs_one_by_one": round(single_ms / batch_ms, 3), "setup_ms": round(setup_ms, 1),
        "autotune": tune, "autotune_s": round(tune_s, 1),
    }
    print(json.dumps(line))
    return 0


def timed_solve(solver, b_dev, x_dev, n, warmup, sync):
    """warmup untimed solves, then one timed solve bracketed by device synchronisations (ms)."""
    for _ in range(warmup):
        solver.solve_device(b_dev.ptr, x_dev.ptr, n)
    sync()
    t0 = time.perf_counter()
    solver.solve_device(b_dev.ptr, x_dev.ptr, n)
    sync()
    return (time.perf_counter() - t0) * 1e3


def kernel_table(kt):
    out = {}
    for name, rec in kt.items():
        if rec["launches"]:
            avg = rec["total_ms"] / rec["launches"]
            out[name] = {"launches": rec["launches"], "avg_us": avg * 1e3, "bytes": rec["bytes"],
                         "GBps": rec["bytes"] / (avg * 1e-3) / 1e9}
    return out


def learned_sequence(args, prob, base, sync):
    """The drifting-modulus protocol on the 64^3 system: the network is trained on pairs from the
    solves of steps 1..T of the sequence, then every path is timed on held-out later steps
    (a spread sample of 51..199).  Returns the per-step records, the summary and the ablations."""
    from polysolve_amd import MASSolver, problems as P
    from polysolve_amd.linear import DeviceArray

    T = args.train_solves
    K = max(args.steps, 1)
    eval_steps = sorted({int(v) for v in np.linspace(T + 1, 199, K).round()})
    while len(eval_steps) < K:  # K > 149: repeat the sample
        eval_steps += eval_steps[:K - len(eval_steps)]
    need = set(range(1, T + 1)) | set(eval_steps)
    moduli = {}
    for t, E in enumerate(P.config5_moduli(prob.ne, steps=max(need) + 1, seed=0)):
        if t in need:
            moduli[t] = E
    n = prob.n
    trainer = MASSolver()
    trainer.set_parameters({"MAS": dict(base, precision="mixed", coarse_solver=args.coarse, learned_collect=True,
                                        lattice_points=args.lattice, learned_hidden=args.hidden,
                                        learned_random_pairs=args.random_pairs, lazy_partitioning=True)})
    trainer.set_coordinates(prob.coords)
    log(f"learned path: pairs from steps 1..{T} of the drifting sequence, held-out steps {eval_steps}")
    t0 = time.perf_counter()
    train_iters = []
    for t in range(1, T + 1):
        pt = P.hex_elasticity(prob.ne, E=moduli[t])
        trainer.factorize(pt.A)
        xt = np.zeros(pt.n)
        trainer.solve(pt.b, xt)
        train_iters.append(trainer.get_info()["solver_iter"])
        if t % 10 == 0:
            log(f"  collected steps 1..{t} ({time.perf_counter() - t0:.0f}s)")
    collect_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = trainer.learned_train(epochs=args.epochs, batch=64, lr=3e-3, seed=0)
    train_s = time.perf_counter() - t0
    pairs = trainer.get_info()["learned"]["pairs"]
    log(f"  {pairs} pairs in {collect_s:.1f}s, trained in {train_s:.1f}s, loss {loss:.3g}")
    trainer.set_parameters({"MAS": {"learned_collect": False}})

    b_dev, x_dev = DeviceArray(n), DeviceArray(n)
    modes = {"learned": {"precision": "mixed", "learned_correction": True},
             "mixed": {"precision": "mixed", "learned_correction": False},
             "fp64": {"precision": "fp64", "learned_correction": False}}

    def run(pt, mode, warmup):
        trainer.set_parameters({"MAS": modes[mode]})
        ms = timed_solve(trainer, b_dev, x_dev, n, warmup, sync)
        info = trainer.get_info()
        x = x_dev.download()
        rel = float(np.linalg.norm(pt.b - pt.A @ x) / np.linalg.norm(pt.b))
        return {"ms": ms, "iters": info["solver_iter"], "rel": rel, "status": info["solver_status"],
                "fallback": bool(info["learned"]["fallback"]), "used": bool(info["learned"]["used"]),
                "reliable_updates": info["reliable_updates"]}, x

    records = []
    systems = {}
    for i, t in enumerate(eval_steps):
        pt = P.hex_elasticity(prob.ne, E=moduli[t])
        trainer.factorize(pt.A)
        b_dev.upload(pt.b)
        rec = {"step": t, "factorize_mode": trainer.get_info()["factorize_mode"]}
        xs = {}
        for mode in ("learned", "mixed", "fp64"):
            rec[mode], xs[mode] = run(pt, mode, max(args.warmup if i == 0 else 1, 1))
        rec["learned_vs_fp64_rel_diff"] = float(np.linalg.norm(xs["learned"] - xs["fp64"]) /
                                                np.linalg.norm(xs["fp64"]))
        records.append(rec)
        if len(systems) < args.ablation_steps:
            systems[t] = pt
        log(f"  step {t}: learned {rec['learned']['ms']:.2f} ms / {rec['learned']['iters']} it, mixed "
            f"{rec['mixed']['ms']:.2f} / {rec['mixed']['iters']}, fp64 {rec['fp64']['ms']:.2f} / "
            f"{rec['fp64']['iters']}, res {rec['learned']['rel']:.2e}")

    # The learned path on step 0 (the sequence's origin, not trained on) for the accuracy check
    # against the direct solve.
    trainer.factorize(prob.A)
    b_dev.upload(prob.b)
    step0, x0_learned = run(prob, "learned", 1)

    # Ablations on the first held-out steps: the network against the linear two-level method
    # it approximates (the exact lattice coarse solve) and against a skip-layer-only network
    # trained on the same pairs.
    abl = {"steps": list(systems), "mlp": [], "linear": [], "exact": []}
    for t in systems:
        r = next(rr for rr in records if rr["step"] == t)
        abl["mlp"].append({"ms": r["learned"]["ms"], "iters": r["learned"]["iters"],
                           "fallback": r["learned"]["fallback"]})
    trainer.set_parameters({"MAS": {"learned_network": "linear"}})
    t1 = time.perf_counter()
    lin_loss = trainer.learned_train(epochs=args.epochs, batch=64, lr=3e-3, seed=0)
    abl["linear_train_s"] = round(time.perf_counter() - t1, 1)
    abl["linear_train_loss"] = lin_loss
    for t, pt in systems.items():
        trainer.factorize(pt.A)
        b_dev.upload(pt.b)
        r, _ = run(pt, "learned", 1)
        abl["linear"].append({"ms": r["ms"], "iters": r["iters"], "fallback": r["fallback"], "rel": r["rel"]})
    exact_setup = []
    for t, pt in systems.items():
        trainer.factorize(pt.A)
        exact_setup.append(trainer.learned_exact())
        b_dev.upload(pt.b)
        r, _ = run(pt, "learned", 1)
        abl["exact"].append({"ms": r["ms"], "iters": r["iters"], "fallback": r["fallback"], "rel": r["rel"]})
    abl["exact_setup_ms"] = [round(v, 1) for v in exact_setup]
    for key in ("mlp", "linear", "exact"):
        abl[key + "_mean_iters"] = float(np.mean([r["iters"] for r in abl[key]]))
        abl[key + "_mean_ms"] = float(np.mean([r["ms"] for r in abl[key]]))
    log(f"  ablation on steps {abl['steps']}: mlp {abl['mlp_mean_iters']:.0f} it / {abl['mlp_mean_ms']:.2f} ms, "
        f"linear {abl['linear_mean_iters']:.0f} / {abl['linear_mean_ms']:.2f}, exact lattice solve "
        f"{abl['exact_mean_iters']:.0f} / {abl['exact_mean_ms']:.2f}")
    b_dev.free()
    x_dev.free()
    del trainer

    def stat(mode, key):
        v = [r[mode][key] for r in records]
        return float(np.mean(v)), float(np.max(v))

    summary = {"train_steps": [1, T], "eval_steps": eval_steps, "pairs": pairs, "collect_s": round(collect_s, 1),
               "train_s": round(train_s, 1), "train_loss": loss, "train_iters_mean": float(np.mean(train_iters)),
               "lattice_points": args.lattice, "hidden": args.hidden}
    for mode in ("learned", "mixed", "fp64"):
        ms_mean, ms_max = stat(mode, "ms")
        it_mean, it_max = stat(mode, "iters")
        summary[mode] = {"ms_mean": round(ms_mean, 3), "ms_max": round(ms_max, 3), "iters_mean": it_mean,
                         "iters_max": it_max, "max_true_rel_residual": max(r[mode]["rel"] for r in records),
                         "converged": sum(r[mode]["status"] == "Reach relative tolerance" for r in records)}
    summary["learned"]["fallbacks"] = sum(r["learned"]["fallback"] for r in records)
    summary["learned"]["max_rel_diff_vs_fp64"] = max(r["learned_vs_fp64_rel_diff"] for r in records)
    return records, summary, abl, step0, x0_learned


def cpu_baselines(args, prob, x_refs):
    """The reference's CPU paths on this machine's host cores: Eigen's Jacobi-CG semantics (full
    run to 1e-8 with OpenMP, plus a bounded one-thread sample as Eigen runs a column-major
    product) and the direct solver (CHOLMOD stand-in: nested-dissection multifrontal Cholesky),
    whose solution is also the FP64 accuracy reference of the GPU results."""
    import oracle
    from oracle.sparse_cholesky import SparseCholesky

    threads = args.cpu_threads or oracle.omp_max_threads()
    out = {"nproc": os.cpu_count(), "threads": threads, "cpu": cpu_info()}
    log(f"CPU: Jacobi-CG (Eigen semantics) to 1e-8, {threads} threads")
    t = time.perf_counter()
    xj, itj, errj = oracle.jacobi_cg_omp(prob.A, prob.b, tol=1e-8, max_iter=args.cpu_max_iter, threads=threads)
    jac_ms = (time.perf_counter() - t) * 1e3
    rel_j = float(np.linalg.norm(prob.b - prob.A @ xj) / np.linalg.norm(prob.b))
    out["jacobi_cg"] = {"ms": round(jac_ms, 1), "iters": itj, "tol_error": errj, "true_rel_residual": rel_j,
                        "converged": errj < 1e-8, "threads": threads}
    log(f"  {jac_ms:.0f} ms, {itj} iterations, rel res {rel_j:.2e}")
    t = time.perf_counter()
    _, it1, _ = oracle.jacobi_cg_omp(prob.A, prob.b, tol=0.0, max_iter=args.cpu_sample_iters, threads=1)
    per_it = (time.perf_counter() - t) / max(it1, 1)
    out["jacobi_cg_1thread_estimate_ms"] = round(per_it * itj * 1e3, 1)
    out["jacobi_cg_1thread_sample"] = f"{it1} iterations on 1 thread ({per_it * 1e3:.1f} ms each) x {itj} iterations"
    if args.no_direct:
        return out, None
    log("CPU: sparse Cholesky (nested dissection, multifrontal) — CHOLMOD stand-in")
    t = time.perf_counter()
    S = SparseCholesky(prob.A, 3, prob.coords)
    S.factorize(progress=lambda f, s: log(f"  factorize {100 * f:.0f}% of the flops after {s:.0f}s"))
    xd = S.solve(prob.b)
    total_ms = (time.perf_counter() - t) * 1e3
    rel_d = float(np.linalg.norm(prob.b - prob.A @ xd) / np.linalg.norm(prob.b))
    out["cholmod_standin"] = {"total_ms": round(total_ms, 1), "analyze_ms": round(S.times["analyze_s"] * 1e3, 1),
                              "factorize_ms": round(S.times["factorize_s"] * 1e3, 1),
                              "solve_ms": round(S.times["solve_s"] * 1e3, 1), "nnz_L": S.nnz_L,
                              "gflop": round(S.flops / 1e9, 1), "true_rel_residual": rel_d}
    log(f"  {total_ms:.0f} ms (factorize {S.times['factorize_s']:.1f}s), rel res {rel_d:.2e}")
    errs = {k: float(np.linalg.norm(x - xd) / np.linalg.norm(xd)) for k, x in x_refs.items()}
    errs["jacobi_cg"] = float(np.linalg.norm(xj - xd) / np.linalg.norm(xd))
    return out, {"reference": "FP64 sparse Cholesky (CHOLMOD stand-in)", "reference_rel_residual": rel_d,
                 "rel_err": errs}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed solves: one per held-out system of the drifting sequence (64^3), else of the system")
    ap.add_argument("--warmup", type=int, default=1, help="untimed solves before the first timed one")
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-direct", action="store_true", help="skip the CPU sparse Cholesky (accuracy reference)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="threads of the CPU Jacobi-CG (0: all)")
    ap.add_argument("--cpu-max-iter", type=int, default=100000)
    ap.add_argument("--cpu-sample-iters", type=int, default=40, help="one-thread Jacobi-CG sample")
    ap.add_argument("--aggregation", default="morton", choices=["morton", "graph"])
    ap.add_argument("--coarse", default="cholesky", choices=["cholesky", "bank"],
                    help="coarsest-level solver of the mixed path")
    ap.add_argument("--profile-only", action="store_true", help="time steps only (for rocprof runs)")
    ap.add_argument("--no-graphs", action="store_true", help="launch kernels one by one (no HIP graph replay)")
    ap.add_argument("--pcg", default="classic", choices=["single", "classic"],
                    help="mixed-path PCG iteration: the classic two-reduction one (default) or the single-reduction one")
    ap.add_argument("--spmv", default="symmetric", choices=["symmetric", "full"],
                    help="form of the matrix the PCG loop's y = A p reads (spmv_symmetric)")
    ap.add_argument("--max-iter", type=int, default=20000, help="iteration cap (short counter-collection runs)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-correction path")
    ap.add_argument("--profile-learned", action="store_true",
                    help="profile mode: also train briefly (2 solves, 20 epochs) and time learned solves")
    ap.add_argument("--train-solves", type=int, default=50,
                    help="learned path: steps 1..T of the drifting sequence provide the training pairs")
    ap.add_argument("--ablation-steps", type=int, default=4, help="held-out steps of the ablations")
    ap.add_argument("--lattice", type=int, default=8, help="learned path: lattice points per axis")
    ap.add_argument("--hidden", type=int, default=128, help="learned path: hidden width of the network")
    ap.add_argument("--random-pairs", type=int, default=100, help="learned path: Gaussian pairs per training solve")
    ap.add_argument("--epochs", type=int, default=150, help="learned path: training epochs")
    ap.add_argument("--multi-rhs", type=int, default=4,
                    help="also solve this many right-hand sides together (0: skip)")
    ap.add_argument("--profile-multi", type=int, default=0,
                    help="profile mode: also solve this many right-hand sides together (kernel counters)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="profile mode: first run one device-to-device copy of this size (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is a design "
                                   "(see DESIGN.md), not a bench mode", "n_gpus": args.gpus}))
        return 2

    from polysolve_amd import MASSolver
    from polysolve_amd.linear import DeviceArray, synchronize

    if args.config == 4:
        return bench_batch(args)

    log(f"assembling config {args.config}")
    t = time.perf_counter()
    prob = make_problem(args.config)
    log(f"  n={prob.n} nnz={prob.A.nnz} ({time.perf_counter() - t:.1f}s)")

    tol = 1e-8
    base = {"block_dim": 3, "relative_tolerance": tol, "absolute_tolerance": 0.0, "max_iter": args.max_iter,
            "aggregation": args.aggregation, "use_graphs": not args.no_graphs,
            "single_reduction": args.pcg == "single", "spmv_symmetric": args.spmv == "symmetric"}
    solver = MASSolver()
    solver.set_parameters({"MAS": dict(base, precision="fp64")})
    solver.set_coordinates(prob.coords)
    t = time.perf_counter()
    solver.factorize(prob.A)
    setup_ms = (time.perf_counter() - t) * 1e3
    levels = solver.level_info()
    log(f"  factorize {setup_ms:.0f} ms, levels {levels}")

    b_dev = DeviceArray(prob.n, prob.b)
    x_dev = DeviceArray(prob.n, np.zeros(prob.n))
    n = prob.n
    bnorm = float(np.linalg.norm(prob.b))
    learned_seq = args.config == 2 and not args.no_learned and not args.profile_only

    def accuracy():
        x = x_dev.download()
        rel = float(np.linalg.norm(prob.b - prob.A @ x) / bnorm)
        return x, rel

    if args.profile_only and args.copy_calibration_mib > 0:
        # A plain device copy of known size: its FETCH_SIZE / WRITE_SIZE counter readings give
        # the factor that turns those counters into bytes for the solver's kernels.
        import ctypes
        from polysolve_amd import _native
        cnt = args.copy_calibration_mib * (1 << 20) // 8
        src, dst = DeviceArray(cnt), DeviceArray(cnt)
        _native.lib().psmas_memcpy(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), cnt * 8, 3)
        synchronize()
        log(f"  calibration copy: {args.copy_calibration_mib} MiB device to device")
        src.free()
        dst.free()

    # The plain paths on this system (the held-out steps of the sequence are timed below).
    path_steps = min(args.steps, 5) if learned_seq else args.steps
    log("fp64 path")
    fp64_ms = run_path(solver, b_dev, x_dev, n, path_steps, args.warmup, synchronize)
    info64 = solver.get_info()
    x64, rel64 = accuracy()
    log(f"  fp64 {fp64_ms:.2f} ms/solve, {info64['solver_iter']} it, true rel res {rel64:.2e}")
    if args.profile_only:
        solver.set_parameters({"MAS": {"precision": "mixed", "coarse_solver": args.coarse}})
        mixed_ms = run_path(solver, b_dev, x_dev, n, args.steps, args.warmup, synchronize)
        log(f"  mixed {mixed_ms:.2f} ms/solve")
        if args.profile_learned and args.config == 2:
            # Kernel durations do not depend on how well the network is trained.
            from polysolve_amd import problems as P

            # The training solves run to convergence (pairs come from their solutions) whatever
            # --max-iter caps the profiled solves at.
            solver.set_parameters({"MAS": {"learned_collect": True, "lattice_points": args.lattice,
                                           "learned_random_pairs": 32, "max_iter": 3000}})
            seq = P.config5_moduli(prob.ne, steps=3, seed=0)
            next(seq)
            for E in seq:
                pt = P.hex_elasticity(prob.ne, E=E)
                solver.factorize(pt.A)
                xt = np.zeros(pt.n)
                solver.solve(pt.b, xt)
            solver.learned_train(epochs=20, batch=64, lr=3e-3, seed=0)
            solver.set_parameters({"MAS": {"learned_collect": False, "learned_correction": True,
                                           "max_iter": args.max_iter}})
            solver.factorize(prob.A)
            learned_ms = run_path(solver, b_dev, x_dev, n, args.steps, args.warmup, synchronize)
            log(f"  learned {learned_ms:.2f} ms/solve, {solver.get_info()['solver_iter']} it")
        if args.profile_multi > 1:
            k = args.profile_multi
            B = np.random.default_rng(7).standard_normal((n, k))
            ms_solver = MASSolver()
            ms_solver.set_parameters({"MAS": dict(base, precision="mixed", coarse_solver=args.coarse)})
            ms_solver.set_coordinates(prob.coords)
            ms_solver.factorize_multi(prob.A, k)
            dB, dX = DeviceArray(n * k, B.ravel(order="F")), DeviceArray(n * k)
            t = time.perf_counter()
            for _ in range(max(args.steps, 1)):
                ms_solver.solve_multi_device(dB.ptr, dX.ptr, n, k)
            synchronize()
            log(f"  {k} right-hand sides together {(time.perf_counter() - t) * 1e3 / max(args.steps, 1):.1f} ms/solve")
            dB.free()
            dX.free()
        return 0

    solver.set_parameters({"MAS": {"precision": "mixed", "coarse_solver": args.coarse}})
    log("mixed path")
    mixed_ms = run_path(solver, b_dev, x_dev, n, path_steps, args.warmup, synchronize)
    infom = solver.get_info()
    xm, relm = accuracy()
    log(f"  mixed {mixed_ms:.2f} ms/solve, {infom['solver_iter']} it, {infom['reliable_updates']} updates, "
        f"true rel res {relm:.2e}")

    # Roofline: dominant kernel of the mixed path, timed with HIP events on the solver stream.
    solver.set_kernel_timing(True)
    solver.solve_device(b_dev.ptr, x_dev.ptr, n)
    synchronize()
    kt = solver.kernel_timing()
    solver.set_kernel_timing(False)
    solver.set_parameters({"MAS": {"precision": "fp64", "coarse_solver": "bank"}})
    solver.set_kernel_timing(True)
    solver.solve_device(b_dev.ptr, x_dev.ptr, n)
    synchronize()
    kt64 = solver.kernel_timing()
    solver.set_kernel_timing(False)
    kernels = kernel_table(kt)
    kernels.update({k + "@fp64path": v for k, v in kernel_table(kt64).items()})
    dom_name = max(kt.items(), key=lambda kv: kv[1]["total_ms"])[0]
    dom = kernels[dom_name]
    read_gbs = measured_read_gbs()
    log(f"  measured streaming read: {read_gbs:.0f} GB/s")
    roofline = {"bound": "hbm", "achieved": round(dom["GBps"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dom["GBps"] / HBM_PEAK_GBS, 4), "traffic": MEASURED_TRAFFIC.get((args.config, dom_name)),
                "kernel": dom_name, "avg_us": round(dom["avg_us"], 2), "bytes_per_launch": dom["bytes"],
                "peak_measured": json_number(read_gbs, 1), "frac_of_measured": json_number(dom["GBps"] / read_gbs, 4),
                "peak_measured_probe": "1 GiB streaming read, 16-byte nontemporal loads, best of 5 (psmas_read_bandwidth)"}

    # Values-only refactorisation (a Newton step's setup): same pattern, new values.
    solver.set_parameters({"MAS": {"lazy_partitioning": True, "precision": "mixed", "coarse_solver": args.coarse}})
    solver.factorize(prob.A)  # records the pattern
    t = time.perf_counter()
    solver.factorize(prob.A)
    refactor_ms = (time.perf_counter() - t) * 1e3
    refactor_info = solver.get_info()
    log(f"  values-only refactorisation {refactor_ms:.1f} ms ({refactor_info['factorize_mode']}), "
        f"phases {refactor_info['setup_ms']}")

    # Several right-hand sides of this matrix solved together (mixed path, factorize_multi): one
    # factorisation, whose SpMV and bank applies read each matrix once for all of them.  Compared
    # with solving the same right-hand sides one by one with the same parameters.
    multi = None
    if args.multi_rhs > 1:
        k = args.multi_rhs
        rng = np.random.default_rng(7)
        B = rng.standard_normal((n, k))
        B[:, 0] = prob.b
        mprm = {"MAS": dict(base, precision="mixed", coarse_solver=args.coarse, learned_correction=False,
                            learned_collect=False)}
        solver.set_parameters(mprm)
        bc = DeviceArray(n)
        solver.solve_device(b_dev.ptr, x_dev.ptr, n)
        synchronize()
        t = time.perf_counter()
        for c in range(k):
            bc.upload(B[:, c])
            solver.solve_device(bc.ptr, x_dev.ptr, n)
        synchronize()
        singles_ms = (time.perf_counter() - t) * 1e3
        bc.free()
        ms_solver = MASSolver()
        ms_solver.set_parameters(mprm)
        ms_solver.set_coordinates(prob.coords)
        t = time.perf_counter()
        ms_solver.factorize_multi(prob.A, k)
        multi_setup = (time.perf_counter() - t) * 1e3
        dB, dX = DeviceArray(n * k, B.ravel(order="F")), DeviceArray(n * k)
        ms_solver.solve_multi_device(dB.ptr, dX.ptr, n, k)
        synchronize()
        reps = max(path_steps // 2, 1)
        t = time.perf_counter()
        for _ in range(reps):
            ms_solver.solve_multi_device(dB.ptr, dX.ptr, n, k)
        synchronize()
        multi_ms = (time.perf_counter() - t) * 1e3 / reps
        X = dX.download().reshape((n, k), order="F")
        worst = max(float(np.linalg.norm(B[:, c] - prob.A @ X[:, c]) / np.linalg.norm(B[:, c])) for c in range(k))
        multi = {"rhs": k, "ms": round(multi_ms, 3), "ms_per_rhs": round(multi_ms / k, 3),
                 "one_by_one_ms": round(singles_ms, 3), "throughput_vs_one_by_one": round(singles_ms / multi_ms, 3),
                 "iters": [ms_solver.batch_info(c)["solver_iter"] for c in range(k)],
                 "max_true_rel_residual": worst, "setup_ms": round(multi_setup, 1)}
        log(f"  multi-RHS: {k} right-hand sides in {multi_ms:.1f} ms against {singles_ms:.1f} ms one by one")
        dB.free()
        dX.free()
        del ms_solver
    b_dev.free()
    x_dev.free()
    del solver

    learned = None
    x_refs = {"fp64": x64, "mixed": xm}
    if learned_seq:
        records, summary, abl, step0, x0l = learned_sequence(args, prob, base, synchronize)
        x_refs["learned"] = x0l
        learned = dict(summary, per_step=records, ablation=abl, step0=step0)

    cpu, acc = (None, None)
    if not args.no_cpu_baseline:
        cpu, acc = cpu_baselines(args, prob, x_refs)

    if learned is not None:
        head_ms = learned["learned"]["ms_mean"]
        head_iters = learned["learned"]["iters_mean"]
        head_path = ("learned-correction mixed MAS-PCG (flexible, reliable updates) on held-out steps of the "
                     f"drifting sequence; network trained on steps 1-{args.train_solves}")
        speed = {"learned_vs_fp64": round(learned["fp64"]["ms_mean"] / head_ms, 3),
                 "mixed_vs_fp64": round(learned["fp64"]["ms_mean"] / learned["mixed"]["ms_mean"], 3),
                 "learned_vs_mixed": round(learned["mixed"]["ms_mean"] / head_ms, 3)}
        workload = (f"Q1-hex elasticity {prob.ne}^3 (n={n}, nnzb={levels['nnzb']}): {len(learned['eval_steps'])} "
                    f"held-out steps of the 200-step 5% drifting-modulus sequence (seed 0)")
        max_rel = learned["learned"]["max_true_rel_residual"]
    else:
        head_ms, head_iters = mixed_ms, infom["solver_iter"]
        head_path = "mixed MAS-PCG (reliable updates)"
        speed = {"mixed_vs_fp64": round(fp64_ms / mixed_ms, 3)}
        workload = f"{prob.name}: Q1-hex elasticity {prob.ne}^3, n={n}, nnzb={levels['nnzb']}"
        max_rel = relm
    cpu_line = None
    if cpu is not None:
        cpu_line = {"value": cpu["jacobi_cg"]["ms"], "unit": "ms", "cores": cpu["threads"], "kind": "port",
                    "sample": (f"full run of the Eigen ConjugateGradient<Lower|Upper, DiagonalPreconditioner> "
                               f"restatement (OpenMP) on the 64^3 step-0 system to rel. residual 1e-8: "
                               f"{cpu['jacobi_cg']['iters']} iterations" if args.config == 2 else
                               f"full Jacobi-CG run to 1e-8: {cpu['jacobi_cg']['iters']} iterations"),
                    **{k: v for k, v in cpu.items() if k != "threads"}}
    line = {
        "metric": "ms to 1e-8 rel. residual + PCG iters: learned/mixed MAS-PCG vs plain FP64",
        "value": round(head_ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head_ms, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "fp32+fp16 inner / fp64 refinement",
        "data": "synthetic (seeded Q1-hex elasticity assembled in-repo)",
        "config": {"workload": workload, "path": head_path, "aggregation": args.aggregation,
                   "parts": levels["parts"], "levels": levels["level_num"]},
        "iters": head_iters,
        "max_true_rel_residual": max_rel,
        "speedup": speed,
        "learned": learned,
        "step0": {"mixed": {"ms": round(mixed_ms, 3), "iters": infom["solver_iter"],
                            "reliable_updates": infom["reliable_updates"], "true_rel_residual": relm},
                  "fp64": {"ms": round(fp64_ms, 3), "iters": info64["solver_iter"], "true_rel_residual": rel64},
                  "learned": learned["step0"] if learned is not None else None},
        "accuracy": acc,
        "setup_ms": round(setup_ms, 1),
        "refactorize_values_ms": round(refactor_ms, 1),
        "refactorize_phases_ms": refactor_info["setup_ms"],
        "roofline": roofline,
        "multi_rhs": multi,
        "kernels": kernels,
        "cpu_baseline": cpu_line,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
