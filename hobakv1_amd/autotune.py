"""Offline autotuner for the PCG kernels' launch shapes.

Sweeps the block-matrix slice height (rows per SELL slice, which sets lanes per row in the SpMV),
the PCG update kernel's workgroup cap (the partial sums the next SpMV re-reduces) and row groups
per wave, how many PCG iterations are queued per group, and (reduced-precision form) the first
inner tolerance and the local aggregates (HOBAK_TUNE_AGG), on a
given workload, and records per configuration the average step time and the sampled SpMV launch
time.  The best configuration per precision is written to
tuning.json next to this file; lib.Context applies it to every instance it creates.

    python -m hobakv1_amd.autotune --n 40 --steps 4 [--precision fp64|fp32] [--write]
    python -m hobakv1_amd.autotune --workload batch --instances 256 [--precision ...] [--write]
    python -m hobakv1_amd.autotune --workload serving --precision fp32 [--write]

The batch workload (the 256-instance C4 set stepped together by hobak_step_batch) is written
as the "<precision>_batch" entry, which contexts created for batched instances apply; the serving
workload (serve.BENCH_STREAM, the serving benchmark's stream, measured by its throughput, three
runs per shape ranked by the median) as the "<precision>_serving" entry.
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")


def load_table():
    if not os.path.exists(TABLE):
        return {}
    with open(TABLE) as f:
        return json.load(f)


def measure(ctx, scene, precond, precision, steps, warmup, inner_tol=1e-4):
    from . import lib

    inst = lib.Instance(ctx, scene, precond=precond, precision=precision,
                        refine_inner_tol=inner_tol)
    for _ in range(warmup):
        inst.step()
    inst.set_state(scene.initial_u(), scene.initial_v())
    t0 = time.perf_counter()
    iters = 0
    for _ in range(steps):
        st = inst.step()
        iters += st["total_pcg_iters"]
    dt = time.perf_counter() - t0
    # the SpMV is timed in a separate step: sampled launches carry events and run uncaptured
    inst.set_state(scene.initial_u(), scene.initial_v())
    inst.set_kernel_sampling(2)
    inst.step()
    ks = inst.kernel_stats()
    inst.close()
    return {"ms_per_step": 1e3 * dt / steps, "pcg_iters": iters / steps,
            "ms_per_pcg_iter": 1e3 * dt / max(1, iters),
            "spmv_ms": ks["spmv_avg_ms"],
            "spmv_GBps": (ks["spmv_bytes"] / (ks["spmv_avg_ms"] * 1e-3) / 1e9
                          if ks["spmv_avg_ms"] > 0 else None)}


def measure_batch(ctx, scs, precision, steps, warmup):
    from . import lib

    insts = [lib.Instance(ctx, sc, precision=precision) for sc in scs]

    def reset():
        for inst, sc in zip(insts, scs):
            inst.set_state(sc.initial_u(), sc.initial_v())

    for _ in range(warmup):
        lib.step_batch(ctx, insts, allow_unconverged=True)
    reset()
    t0 = time.perf_counter()
    iters = 0
    for _ in range(steps):
        st = lib.step_batch(ctx, insts, allow_unconverged=True)
        iters += sum(s["total_pcg_iters"] for s in st)
    dt = time.perf_counter() - t0
    for inst in insts:
        inst.set_kernel_sampling(2)
    reset()
    lib.step_batch(ctx, insts, allow_unconverged=True)
    ks = insts[0].kernel_stats()
    for inst in insts:
        inst.close()
    return {"ms_per_step": 1e3 * dt / steps, "pcg_iters": iters / steps,
            "spmv_ms": ks["spmv_avg_ms"],
            "spmv_GBps": (ks["spmv_bytes"] / (ks["spmv_avg_ms"] * 1e-3) / 1e9
                          if ks["spmv_avg_ms"] > 0 else None)}


def measure_serving(ctx, precision):
    from . import lib, scenes, serve

    b = serve.BENCH_STREAM
    stream = serve.request_stream(b["sessions"], b["rate"], b["steps"], seed=b["seed"])
    cache = {}

    def scene_of(k):
        if k not in cache:
            cache[k] = scenes.serving_instance(k, seed=b["seed"])
        return cache[k]

    r = serve.run(ctx, stream, scene_of, int(b["budget_gib"] * (1 << 30)), precision, mode="batched",
                  keep_states=False)
    # ranked by ms_per_step like the other workloads: the stream's makespan over its requests
    return {"ms_per_step": 1e3 * r["makespan_s"] / max(1, r["requests"]), "tet_steps_per_s": r["value"],
            "p50_step_ms": r["p50_step_ms"], "p99_step_ms": r["p99_step_ms"], "spmv_ms": 0.0,
            "spmv_GBps": None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=40)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--slice-rows", default="8,16,32,64")
    ap.add_argument("--poll", default="8,32")
    ap.add_argument("--update-wgs", default="512")
    ap.add_argument("--update-groups", default="2", help="21-row groups per update wave: 1, 2, 4")
    ap.add_argument("--spmv-pad", default="0,1")
    ap.add_argument("--row-sort", default="0", help="vertex relabelling by row length: 0, 1 or 0,1")
    ap.add_argument("--inner-tol", default="1e-4", help="FP32 inner PCG tolerances to sweep")
    ap.add_argument("--agg", default="0", help="local aggregates (reduced-precision form): 0, 1 or 0,1")
    ap.add_argument("--workload", default="single", choices=["single", "batch", "serving"])
    ap.add_argument("--instances", type=int, default=256, help="batch workload size")
    ap.add_argument("--repeats", type=int, default=None,
                    help="runs per shape, ranked by the median (default: 3 for the serving "
                         "workload, whose run-to-run spread is as large as the differences between "
                         "shapes; 1 otherwise)")
    ap.add_argument("--write", action="store_true")
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    from hobakv1_amd import lib, scenes, serve

    precision = lib.PRECISION_FP64 if args.precision == "fp64" else lib.PRECISION_FP32_REFINE
    sc = scenes.lattice_scene((args.n,) * 3, seed=2, stretch_velocity=True)
    scs = scenes.batch_instances(args.instances) if args.workload == "batch" else None
    results = []
    tols = [float(x) for x in args.inner_tol.split(",")] if args.precision == "fp32" else [0.0]
    for H, poll, uw, pad, rs, ug, tol, ag in itertools.product([int(x) for x in args.slice_rows.split(",")],
                                                           [int(x) for x in args.poll.split(",")],
                                                           [int(x) for x in args.update_wgs.split(",")],
                                                           [int(x) for x in args.spmv_pad.split(",")],
                                                           [int(x) for x in args.row_sort.split(",")],
                                                           [int(x) for x in args.update_groups.split(",")],
                                                           tols, [int(x) for x in args.agg.split(",")]):
        repeats = args.repeats or (3 if args.workload == "serving" else 1)
        runs = []
        for _ in range(repeats):
            ctx = lib.Context(0, apply_table=False)
            ctx.set_tuning(slice_rows=H, poll_chunk=poll, update_wgs=uw, spmv_pad=pad, row_sort=rs,
                           update_groups=ug, agg=ag)
            if args.workload == "serving":
                ctx.refine_inner_tol = tol or 1e-4
                runs.append(measure_serving(ctx, precision))
            elif scs is not None:
                ctx.refine_inner_tol = tol or 1e-4
                runs.append(measure_batch(ctx, scs, precision, args.steps, args.warmup))
            else:
                runs.append(measure(ctx, sc, scenes.PRECOND_BLOCK_JACOBI, precision, args.steps,
                                    args.warmup, inner_tol=tol or 1e-4))
            ctx.close()
        # the run of median ms_per_step stands for the shape; every run's value is kept
        runs.sort(key=lambda q: q["ms_per_step"])
        r = dict(runs[len(runs) // 2])
        if repeats > 1:
            r["ms_per_step_runs"] = [q["ms_per_step"] for q in runs]
        r.update({"slice_rows": H, "poll_chunk": poll, "update_wgs": uw, "spmv_pad": pad,
                  "row_sort": rs, "update_groups": ug, "agg": ag})
        if args.precision == "fp32":
            r["refine_inner_tol"] = tol
        results.append(r)
        print(json.dumps(r), flush=True)
    best = min(results, key=lambda r: r["ms_per_step"])
    print("best", json.dumps(best), flush=True)
    if args.write:
        table = load_table()
        key = args.precision + {"single": "", "batch": "_batch", "serving": "_serving"}[args.workload]
        bs = serve.BENCH_STREAM
        what = {"single": f"{args.n}^3 lattice", "batch": f"{args.instances}-instance batch",
                "serving": f"serving stream: {bs['sessions']} sessions, Poisson {bs['rate']:g}/s, "
                           f"{bs['steps']} steps each, {bs['budget_gib']:g} GiB"}[args.workload]
        table[key] = {"slice_rows": best["slice_rows"], "poll_chunk": best["poll_chunk"],
                      "update_wgs": best["update_wgs"], "spmv_pad": best["spmv_pad"],
                      "row_sort": best["row_sort"], "update_groups": best["update_groups"],
                      "agg": best["agg"], "measured": {"workload": what, **best}}
        if "refine_inner_tol" in best:
            table[key]["refine_inner_tol"] = best["refine_inner_tol"]
        with open(TABLE, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
