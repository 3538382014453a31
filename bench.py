#!/usr/bin/env python3
"""Benchmark of the backward-Euler Newton step on one MI355X.

A "step" is one timestep of the workload: one Newton iteration over the whole mesh (element
kernel, assembly, right-hand side, PCG to tol 1e-8, state update).  The default workload is the
40^3 Kuhn-lattice cube (384,000 tets, 68,921 vertices), Stable Neo-Hookean E=1e5 nu=0.45 rho=1000,
x=0 face pinned, seeded [0.7, 1.3] stretch applied as a velocity field, dt = 1/60, FP64,
block-Jacobi PCG.  Inputs are resident in HBM before timing starts; the warmup steps are followed
by a reset to the initial state, so the timed steps are steps 1..K of the workload.

Output: one JSON line (see the field list in main()).
  value      tet-Newton-steps/s = nt * Newton iterations * K / timed seconds
  roofline   the PCG's BSR SpMV (the dominant kernel): algorithmic bytes per launch
             nnzb*(9*8+4) + 4*(nv+1) + 2*3*8*nv divided by its average launch time, sampled live
             during the timed steps with HIP start/stop events that the SpMV's own dispatch
             carries (hipExtLaunchKernelGGL) on the library's stream; `traffic` is the HBM bytes
             per launch from the committed rocprofv3 --pmc summary of this workload
  cpu_baseline  the literal reference (oracle/_ref, compiled from its sources) timed on a bounded
             sample of the same workload on this host's cores
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "tet-Newton-steps/s + p50/p99 step latency vs FP64 GPU path at 1e-6 error"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=40, help="lattice cells per axis")
    ap.add_argument("--precond", default="block", choices=["block", "jacobi"])
    ap.add_argument("--cpu-baseline", default="auto", choices=["auto", "reference", "port", "none"])
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--sample-every", type=int, default=4)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run the calibration copy kernel (for counter scaling)")
    ap.add_argument("--precision-form", default="on", choices=["on", "off"],
                    help="also time the FP32 + FP64-refinement path on the same workload")
    return ap.parse_args()


def measured_traffic(n, precision="fp64"):
    """HBM bytes per SpMV launch from the committed rocprofv3 --pmc summary, when it was taken on
    this workload (FETCH_SIZE and WRITE_SIZE scaled by a calibration copy kernel)."""
    path = os.path.join(ROOT, "profiles", f"cg_spmv_{n}cube_{precision}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith((" ", "#", "|")):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    if vals.get("workload_n") != str(n) or vals.get("precision") != precision:
        return None
    return {"bytes": float(vals["traffic_bytes_per_launch"]), "kernel": vals.get("kernel"),
            "source": os.path.relpath(path, ROOT)}


def cpu_baseline(scene, steps, kind):
    """Time the CPU checker on the first `steps` steps of the same scene."""
    import oracle  # test/baseline infrastructure only

    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if kind in ("auto", "reference") and oracle.ref_available():
        sim = oracle.RefSim(scene)
        k = "reference"
    elif kind in ("auto", "port"):
        sim = oracle.Oracle(scene, precond=0)
        k = "port"
    else:
        return None
    secs = []
    iters = []
    for _ in range(steps):
        it, _, sec = sim.step()
        secs.append(sec)
        iters.append(int(it))
    sim.close()
    total = float(np.sum(secs))
    return {
        "value": scene.nt * steps / total,
        "unit": "tet-Newton-steps/s",
        "cores": cores,
        "kind": k,
        "sample": f"first {steps} steps of the same {scene.dims[0]}^3 scene, Jacobi PCG tol 1e-8 "
                  f"(the reference's preconditioner); {total:.2f} s; PCG iters {iters}",
        "ms_per_step": 1e3 * total / steps,
    }


def precision_form(lib, sc, precond, args, fp64_total, u64):
    """The same K steps with FP32 element/assembly/PCG and FP64 refinement; compared with the
    FP64 GPU path's positions after the same steps."""
    ctx = lib.Context(0, precision="fp32")
    inst = lib.Instance(ctx, sc, precond=precond, precision=lib.PRECISION_FP32_REFINE)
    for _ in range(args.warmup):
        inst.step()
    inst.set_state(sc.initial_u(), sc.initial_v())
    step_s, iters, sweeps = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        st = inst.step()
        step_s.append(time.perf_counter() - t1)
        iters.append(st["pcg_iters"][0])
        sweeps.append(st["refine_iters"][0])
    total = time.perf_counter() - t0
    u, _ = inst.state()
    X = sc.rest.ravel()
    err = float(np.linalg.norm(u - u64) / np.linalg.norm(X + u64))
    inst.close()
    ctx.close()
    return {
        "value": sc.nt * args.steps / total,
        "unit": "tet-Newton-steps/s",
        "ms_per_step": 1e3 * total / args.steps,
        "p50_step_ms": 1e3 * float(np.percentile(step_s, 50)),
        "p99_step_ms": 1e3 * float(np.percentile(step_s, 99)),
        "speedup_vs_fp64": fp64_total / total,
        "rel_position_error_vs_fp64_after_steps": err,
        "inner_pcg_iters_per_step": iters,
        "refine_sweeps_per_step": sweeps,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards by instance, see DESIGN.md)", file=sys.stderr)
        return 2
    from hobakv1_amd import lib, scenes

    sc = scenes.lattice_scene((args.n, args.n, args.n), seed=2, stretch_velocity=True,
                              name=f"cube{args.n}_snh_stretch")
    precond = scenes.PRECOND_BLOCK_JACOBI if args.precond == "block" else scenes.PRECOND_JACOBI
    ctx = lib.Context(0)
    t0 = time.time()
    inst = lib.Instance(ctx, sc, precond=precond)
    setup_s = time.time() - t0
    for _ in range(args.warmup):
        inst.step()
    inst.set_state(sc.initial_u(), sc.initial_v())
    inst.set_kernel_sampling(args.sample_every)
    step_s = []
    iters = []
    phase = np.zeros(5)
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        st = inst.step()
        step_s.append(time.perf_counter() - t1)
        iters.append(st["pcg_iters"][0])
        ph = inst.timing_ms()
        phase += np.array([ph["element"], ph["assemble"], ph["rhs"], ph["pcg"], ph["update"]])
    total = time.perf_counter() - t_start
    ks = inst.kernel_stats()
    u, _ = inst.state()
    finite = bool(np.all(np.isfinite(u)))

    spmv_gbs = ks["spmv_bytes"] / (ks["spmv_avg_ms"] * 1e-3) / 1e9 if ks["spmv_avg_ms"] > 0 else 0.0
    tr = measured_traffic(args.n)
    roofline = {
        "bound": "hbm",
        "kernel": "cg_spmv_kernel<double>",
        "achieved": round(spmv_gbs, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(spmv_gbs / HBM_PEAK_GBS, 4),
        "traffic": tr["bytes"] if tr else None,
        "traffic_source": tr["source"] if tr else None,
        "bytes_per_launch": ks["spmv_bytes"],
        "avg_launch_ms": ks["spmv_avg_ms"],
        "launches_sampled": ks["spmv_count"],
    }
    out = {
        "metric": METRIC,
        "value": sc.nt * args.steps / total,
        "unit": "tet-Newton-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "p50_step_ms": 1e3 * float(np.percentile(step_s, 50)),
        "p99_step_ms": 1e3 * float(np.percentile(step_s, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.n}^3 Kuhn lattice cube ({sc.nt} tets, {sc.nv} vertices), SNH E=1e5 "
                        f"nu=0.45 rho=1000, x=0 pinned, seeded [0.7,1.3] stretch velocity, "
                        f"dt=1/60, 1 Newton iteration/step, FP64, "
                        f"{'block-Jacobi' if precond else 'Jacobi'} PCG tol 1e-8",
            "tets": sc.nt, "vertices": sc.nv, "precision": "fp64",
            "precond": args.precond, "parallelism": "single GPU",
        },
        "pcg_iters_per_step": iters,
        "phase_ms_per_step": dict(zip(["element", "assemble", "rhs", "pcg", "update"],
                                      (phase / args.steps).round(4).tolist())),
        "setup_s": round(setup_s, 3),
        "finite": finite,
        "roofline": roofline,
    }
    if args.precision_form == "on":
        out["fp32_refine"] = precision_form(lib, sc, precond, args, total, u)
    if args.copy_probe:
        nb = 512 << 20
        ms = ctx.copy_probe(nb, 20)
        out["copy_probe"] = {"bytes": nb, "ms": ms, "GBps": 2 * nb / (ms * 1e-3) / 1e9}
    inst.close()
    ctx.close()
    if args.cpu_baseline != "none":
        cb = cpu_baseline(sc, args.cpu_steps, args.cpu_baseline)
        if cb:
            out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
