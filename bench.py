#!/usr/bin/env python3
"""Benchmark of the backward-Euler Newton step on one MI355X.

A "step" is one timestep of the workload: one Newton iteration over the whole mesh (element
kernel, assembly, right-hand side, PCG to tol 1e-8, state update).  The headline workload is the
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
  fp32_refine   the same K steps in the precision form (FP32 element/assembly/PCG, FP64
             refinement), its speed-up and its position error against the FP64 run
  batch      256 independent instances (dims U{4..24}^3, SNH/ARAP, seeded E and nu) advanced
             together by hobak_step_batch, FP64 and FP32+refinement, against stepping each alone
  serving    a Poisson stream of sessions (tets log-uniform in [1e3, 4e5], penalty ground,
             20 step requests each) served by the continuous-batching scheduler in the precision
             form, against the plain FP64 path serving the same stream one step at a time
  beyond_mall   the SpMV's roofline on one 88^3 lattice instance (4.09 M tets, 0.83 GB FP64 /
             0.44 GB FP32 per launch), FP64 and FP32, where no cache holds the matrix
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
FP64_VALU_PEAK_TFLOPS = 78.6
# FP64 flops the element kernel executes per tet (forces and clamped stiffness), counted with
# rocprofv3 (64 x SQ_INSTS_VALU_FLOPS_FP64 incl. transcendental ops / tets), and the kernel's
# VALUBusy from the same runs: profiles/element_kernel_40cube_valu.md
ELEMENT_FP64_FLOPS_PER_TET = 2158.0
ELEMENT_VALU_BUSY = 0.121
GRAPHS = [1]  # --graphs, applied to every context the benchmark creates


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=40, help="lattice cells per axis")
    ap.add_argument("--precond", default="block", choices=["block", "jacobi"])
    ap.add_argument("--cpu-baseline", default="auto", choices=["auto", "reference", "port", "none"])
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--sample-every", type=int, default=256,
                    help="time every n-th SpMV launch of the timed steps with HIP events; a PCG "
                         "group holding one runs uncaptured, the others replay their graph")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run the calibration copy kernel (for counter scaling)")
    ap.add_argument("--precision-form", default="on", choices=["on", "off"],
                    help="also time the FP32 + FP64-refinement path on the same workload")
    ap.add_argument("--batch", default="on", choices=["on", "off"],
                    help="also time 256 instances stepped together (hobak_step_batch)")
    ap.add_argument("--batch-instances", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--batch-alone", default="on", choices=["on", "off"],
                    help="also step each batch instance alone (speedup and bitwise comparison)")
    ap.add_argument("--serving", default="on", choices=["on", "off"],
                    help="also run the serving stream against the plain FP64 path")
    ap.add_argument("--sessions", type=int, default=48)
    ap.add_argument("--rate", type=float, default=400.0, help="session arrivals per second")
    ap.add_argument("--session-steps", type=int, default=20)
    ap.add_argument("--budget-gib", type=float, default=8.0)
    ap.add_argument("--beyond-mall", default="on", choices=["on", "off"],
                    help="the SpMV on one large instance whose matrix exceeds the Infinity Cache")
    ap.add_argument("--large-n", type=int, default=88, help="lattice cells per axis of that instance")
    ap.add_argument("--graphs", default="on", choices=["on", "off"],
                    help="off: launch every kernel directly (for profilers that do not trace "
                         "graph launches)")
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
    # the reference prints progress to stdout; keep stdout for the one JSON line
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        return _cpu_baseline(oracle, scene, steps, kind, cores)
    finally:
        sys.stdout.flush()
        os.dup2(saved, 1)
        os.close(saved)


def _cpu_baseline(oracle, scene, steps, kind, cores):
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
    ctx.set_tuning(graphs=GRAPHS[0])
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


def c3_errors(lib, scenes, steps):
    """C3: the 40^3 stretch scene in FP32 + FP64 refinement against the FP64 path, SNH and ARAP
    (mu = 3.4e4), after every step: the position error, and at the FP64 state the errors of the
    FP32 forces and of the FP32 stiffness applied to a fixed random vector (normwise relative)."""
    out = {}
    rng = np.random.default_rng(12)
    for mat, name in ((scenes.MAT_SNH, "snh"), (scenes.MAT_ARAP, "arap")):
        sc = scenes.config3(mat)
        c64, c32 = lib.Context(0), lib.Context(0, precision="fp32")
        a = lib.Instance(c64, sc, precond=scenes.PRECOND_BLOCK_JACOBI)
        b = lib.Instance(c32, sc, precond=scenes.PRECOND_BLOCK_JACOBI, precision=lib.PRECISION_FP32_REFINE)
        rp, ci = a.bsr_pattern()
        w = rng.uniform(-1, 1, sc.dofs)
        X = sc.rest.ravel()

        rows = np.repeat(np.arange(sc.nv), np.diff(rp))

        def kw(K):  # the block matrix applied to w
            y = np.zeros((sc.nv, 3))
            np.add.at(y, rows, np.einsum("kab,kb->ka", K, w.reshape(-1, 3)[ci]))
            return y.ravel()

        pos, frc, kv = [], [], []
        for _ in range(steps):
            a.step()
            b.step()
            ua, _ = a.state()
            ub, _ = b.state()
            pos.append(float(np.linalg.norm(ub - ua) / np.linalg.norm(X + ua)))
        Ra, Ka = a.forces_hessian(ua)
        Rb, Kb = b.forces_hessian(ua)
        frc = float(np.linalg.norm(Rb - Ra) / np.linalg.norm(Ra))
        ya, yb = kw(Ka), kw(Kb)
        kv = float(np.linalg.norm(yb - ya) / np.linalg.norm(ya))
        out[name] = {"rel_position_error_per_step": pos, "max_rel_position_error": max(pos),
                     "rel_force_error_at_last_state": frc, "rel_stiffness_vector_error_at_last_state": kv}
        for x in (a, b):
            x.close()
        c64.close()
        c32.close()
    return out


def batch_form(lib, scenes, args):
    """hobak_step_batch over many independent instances against stepping each one alone."""
    scs = scenes.batch_instances(args.batch_instances)
    nt = sum(sc.nt for sc in scs)
    out = {"instances": len(scs), "tets": nt, "steps": args.batch_steps,
           "workload": f"{len(scs)} instances, dims U{{4..24}}^3 Kuhn lattices, SNH/ARAP 50/50, "
                       f"E log-U[1e4,1e6], nu U[0.3,0.48], seeded stretch velocity, block-Jacobi"}
    final = {}
    for prec, name in ((lib.PRECISION_FP64, "fp64"), (lib.PRECISION_FP32_REFINE, "fp32_refine")):
        ctx = lib.Context(0, precision="fp64" if prec == lib.PRECISION_FP64 else "fp32",
                          workload="batch")
        ctx.set_tuning(graphs=GRAPHS[0])
        insts = [lib.Instance(ctx, sc, precision=prec) for sc in scs]

        def reset():
            for inst, sc in zip(insts, scs):
                inst.set_state(sc.initial_u(), sc.initial_v())

        lib.step_batch(ctx, insts, allow_unconverged=True)
        reset()
        step_s = []
        t0 = time.perf_counter()
        for _ in range(args.batch_steps):
            t1 = time.perf_counter()
            lib.step_batch(ctx, insts, allow_unconverged=True)
            step_s.append(time.perf_counter() - t1)
        tb = time.perf_counter() - t0
        ub = [inst.state()[0] for inst in insts]
        final[name] = ub
        res = {"value": nt * args.batch_steps / tb, "unit": "tet-Newton-steps/s",
               "ms_per_batch_step": 1e3 * tb / args.batch_steps,
               "p50_batch_step_ms": 1e3 * float(np.percentile(step_s, 50)),
               "p99_batch_step_ms": 1e3 * float(np.percentile(step_s, 99))}
        if args.batch_alone == "on":
            reset()
            t0 = time.perf_counter()
            for _ in range(args.batch_steps):
                for inst in insts:
                    inst.step(allow_unconverged=True)
            ta = time.perf_counter() - t0
            res["alone_value"] = nt * args.batch_steps / ta
            res["speedup_vs_alone"] = ta / tb
            res["bitwise_equal_to_alone"] = all(np.array_equal(inst.state()[0], u)
                                                for inst, u in zip(insts, ub))
        for inst in insts:
            inst.set_kernel_sampling(2)
        reset()
        lib.step_batch(ctx, insts, allow_unconverged=True)
        ks = insts[0].kernel_stats()
        nv = sum(sc.nv for sc in scs)
        rf = spmv_roofline(ks, prec == lib.PRECISION_FP64, nv)
        if rf:
            rf["note"] = "launches during which every instance was solving (its first inner solve)"
            res["roofline"] = rf
        if name == "fp32_refine":
            res["max_rel_position_error_vs_fp64"] = max(
                float(np.linalg.norm(u32 - u64) / np.linalg.norm(sc.rest.ravel() + u64))
                for u32, u64, sc in zip(final["fp32_refine"], final["fp64"], scs))
        out[name] = res
        for inst in insts:
            inst.close()
        ctx.close()
    return out


def spmv_roofline(ks, fp64, nv):
    """Roofline entry of the sampled SpMV launches: the algorithmic bytes per launch
    nnzb*(9s+4) + 4*(nv+1) + 2*3s*nv over the average launch time, and the same with the
    kernel's two vector writes (t and p_new, 2*3s*nv) added."""
    if ks["spmv_avg_ms"] <= 0:
        return None
    s = 8 if fp64 else 4
    sec = ks["spmv_avg_ms"] * 1e-3
    gbs = ks["spmv_bytes"] / sec / 1e9
    gbs_w = (ks["spmv_bytes"] + 2 * 3 * s * nv) / sec / 1e9
    return {"bound": "hbm", "kernel": f"cg_spmv_kernel<{'double' if fp64 else 'float'}>",
            "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "achieved_incl_vector_writes": round(gbs_w, 1),
            "frac_incl_vector_writes": round(gbs_w / HBM_PEAK_GBS, 4),
            "bytes_per_launch": ks["spmv_bytes"], "avg_launch_ms": ks["spmv_avg_ms"],
            "launches_sampled": ks["spmv_count"]}


def beyond_mall_form(lib, scenes, args):
    """The SpMV on a working set well beyond the 256 MB Infinity Cache: one n^3 lattice instance
    of about the 256-instance batch's total size, one step per precision with every SpMV launch
    of the first (inner) solve sampled.  Launch shapes: the batch entry's, except slice height 64
    with the padded-width loop in both precisions (on this one large instance they measured
    faster than the FP64 batch entry's slice 32 row-length loop: 178 against 180 us)."""
    n = args.large_n
    sc = scenes.lattice_scene((n, n, n), seed=2, stretch_velocity=True)
    out = {"workload": f"{n}^3 Kuhn lattice cube ({sc.nt} tets, {sc.nv} vertices), stretch scene, "
                       f"block-Jacobi, one step", "tets": sc.nt, "vertices": sc.nv}
    for prec, name in ((lib.PRECISION_FP64, "fp64"), (lib.PRECISION_FP32_REFINE, "fp32_refine")):
        ctx = lib.Context(0, precision="fp64" if prec == lib.PRECISION_FP64 else "fp32",
                          workload="batch")
        ctx.set_tuning(graphs=GRAPHS[0], slice_rows=64, spmv_pad=1)
        inst = lib.Instance(ctx, sc, precision=prec)
        inst.set_kernel_sampling(1)
        t0 = time.perf_counter()
        st = inst.step(allow_unconverged=True)
        dt = time.perf_counter() - t0
        rf = spmv_roofline(inst.kernel_stats(), prec == lib.PRECISION_FP64, sc.nv)
        out[name] = {"ms_step_sampled": 1e3 * dt, "pcg_iters": st["total_pcg_iters"],
                     "roofline": rf}
        u_after = inst.state()[0]
        if prec == lib.PRECISION_FP64:
            u64 = u_after
        else:
            out[name]["rel_position_error_vs_fp64"] = float(
                np.linalg.norm(u_after - u64) / np.linalg.norm(sc.rest.ravel() + u64))
        inst.close()
        ctx.close()
    return out


def serving_form(lib, scenes, args):
    from hobakv1_amd import serve

    stream = serve.request_stream(args.sessions, args.rate, args.session_steps, seed=5)
    cache = {}

    def scene_of(k):
        if k not in cache:
            cache[k] = scenes.serving_instance(k, seed=5)
        return cache[k]

    for s_ in stream:
        scene_of(s_.k)
    ca = lib.Context(0, precision="fp32", workload="serving")
    cb = lib.Context(0)
    ca.set_tuning(graphs=GRAPHS[0])
    cb.set_tuning(graphs=GRAPHS[0])
    out = serve.compare(ca, cb, stream, scene_of, int(args.budget_gib * (1 << 30)),
                        sample_every=args.sample_every)
    ks = out.pop("spmv", None)
    if ks and ks["spmv_avg_ms"] > 0:
        gbs = ks["spmv_bytes"] / (ks["spmv_avg_ms"] * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": "cg_spmv_kernel<float>", "achieved": round(gbs, 1),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                           "bytes_per_launch": ks["spmv_bytes"], "avg_launch_ms": ks["spmv_avg_ms"],
                           "launches_sampled": ks["spmv_count"],
                           "note": "sampled launches of the serving engine; bytes are those of the "
                                   "instances solving in the launch's tick"}
    ca.close()
    cb.close()
    out["rate_sessions_per_s"] = args.rate
    out["mean_tets"] = float(np.mean([scene_of(s_.k).nt for s_ in stream]))
    out["workload"] = (f"{args.sessions} sessions, Poisson arrivals {args.rate:g}/s, "
                       f"{args.session_steps} step requests each, tets log-U[1e3,4e5], aspect "
                       f"[0.5,2], SNH/ARAP, penalty ground y=-0.5 k=1e6, HBM budget "
                       f"{args.budget_gib:g} GiB; instance set-up excluded from both clocks")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards by instance: run one process per GPU, each with its own instances)", file=sys.stderr)
        return 2
    from hobakv1_amd import lib, scenes

    sc = scenes.lattice_scene((args.n, args.n, args.n), seed=2, stretch_velocity=True,
                              name=f"cube{args.n}_snh_stretch")
    precond = scenes.PRECOND_BLOCK_JACOBI if args.precond == "block" else scenes.PRECOND_JACOBI
    GRAPHS[0] = 1 if args.graphs == "on" else 0
    ctx = lib.Context(0)
    ctx.set_tuning(graphs=GRAPHS[0])
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
    # one more step outside the timed region, with direct launches, so that the element kernel
    # has events of its own (a captured step times its prologue as one piece)
    ctx.set_tuning(graphs=0)
    inst.step()
    el_ms = inst.timing_ms()["element"]
    ctx.set_tuning(graphs=GRAPHS[0])
    el_tflops = ELEMENT_FP64_FLOPS_PER_TET * sc.nt / (el_ms * 1e-3) / 1e12 if el_ms > 0 else 0.0
    element = {"kernel": "element_kernel<double>", "bound": "latency", "ms": el_ms,
               "achieved": round(el_tflops, 3), "peak": FP64_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
               "frac": round(el_tflops / FP64_VALU_PEAK_TFLOPS, 4), "valu_busy": ELEMENT_VALU_BUSY,
               "flops_per_tet": ELEMENT_FP64_FLOPS_PER_TET,
               "source": "profiles/element_kernel_40cube_valu.md"}

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
        "element_kernel": element,
    }
    if args.precision_form == "on":
        out["fp32_refine"] = precision_form(lib, sc, precond, args, total, u)
        out["fp32_refine"]["c3_errors_vs_fp64"] = c3_errors(lib, scenes, args.steps)
    if args.batch == "on":
        out["batch"] = batch_form(lib, scenes, args)
    if args.serving == "on":
        out["serving"] = serving_form(lib, scenes, args)
    if args.beyond_mall == "on":
        out["beyond_mall"] = beyond_mall_form(lib, scenes, args)
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
