#!/usr/bin/env python3
"""Benchmark of the backward-Euler Newton step on one MI355X.

The headline is the metric's own comparison: the serving request stream (tets log-uniform in
[1e3, 4e5], aspect [0.5, 2], mixed SNH/ARAP, a penalty ground at y = -0.5 with k = 1e6, instances
dropped from y in [0, 1]), served by the continuous-batching scheduler in the precision form (FP32
element/assembly/PCG with FP64 refinement) under an HBM budget, against the plain FP64 GPU path
serving the same stream one step at a time.  A "step" is one step request: one Newton iteration
of one instance (element kernel, assembly, right-hand side, PCG to tol 1e-8, state update).  Each
session issues --steps requests one after another; --warmup sessions run through the scheduler
before the timed stream.  Instance set-up (host preprocessing and uploads, before a session's
first request) is outside the clock; the end-to-end rate that charges it is reported beside it.

Output: one JSON line (see the field list in main()).
  value      serving tet-Newton-steps/s = sum over requests of nt / the stream's makespan
  roofline   the dominant kernel of the serving stream, the PCG's BSR SpMV (FP32): algorithmic
             bytes nnzb*(9*4+4) + 4*(nv+1) + 2*3*4*nv of the instances solving in a launch's tick,
             over its duration, sampled live with HIP start/stop events that the SpMV's own
             dispatch carries (hipExtLaunchKernelGGL) on the library's stream; `traffic` is the
             HBM bytes per launch from the committed rocprofv3 --pmc summary of this stream
  cpu_baseline  the oracle's FP64 restatement (OpenMP) timed on a bounded sample of the same
             stream's sessions on this host's cores (the literal reference has no penalty
             ground); cpu_baseline_c2 times the literal reference (oracle/_ref) on the 40^3 scene
  plain_fp64_c2  the 40^3 Kuhn-lattice cube (384,000 tets), SNH E=1e5 nu=0.45 rho=1000, x=0
             pinned, seeded stretch velocity, FP64 block-Jacobi: K timed steps, its SpMV roofline,
             the element kernel's FLOP rate and the per-phase times
  fp32_refine   the same 40^3 steps in the precision form, and C3's per-step errors (SNH, ARAP)
  batch      256 independent instances (dims U{4..24}^3) advanced together by hobak_step_batch
  beyond_mall   the SpMV's roofline on one 88^3 lattice instance (4.09 M tets), FP64 and FP32
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
ELEMENT_VALU_BUSY = 0.264
GRAPHS = [1]  # --graphs, applied to every context the benchmark creates


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="step requests per serving session (and timed steps of the 40^3 section)")
    ap.add_argument("--warmup", type=int, default=2,
                    help="untimed sessions served before the timed stream (and untimed steps of "
                         "the 40^3 section)")
    ap.add_argument("--n", type=int, default=40, help="lattice cells per axis of the C2 section")
    ap.add_argument("--precond", default="block", choices=["block", "jacobi"])
    ap.add_argument("--cpu-baseline", default="auto", choices=["auto", "reference", "port", "none"])
    ap.add_argument("--cpu-steps", type=int, default=20, help="steps of the 40^3 reference timing")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="bound on the CPU time of the serving-stream sample")
    ap.add_argument("--sample-every", type=int, default=256,
                    help="time every n-th SpMV launch with HIP events; a PCG group holding one "
                         "runs uncaptured, the others replay their graph")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run the calibration copy kernel (for counter scaling)")
    ap.add_argument("--plain-c2", default="on", choices=["on", "off"],
                    help="the 40^3 plain FP64 section")
    ap.add_argument("--precision-form", default="on", choices=["on", "off"],
                    help="also time the FP32 + FP64-refinement path on the 40^3 workload")
    ap.add_argument("--batch", default="on", choices=["on", "off"],
                    help="also time 256 instances stepped together (hobak_step_batch)")
    ap.add_argument("--batch-instances", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--batch-alone", default="on", choices=["on", "off"],
                    help="also step each batch instance alone (speedup and bitwise comparison)")
    ap.add_argument("--serving", default="on", choices=["on", "off"],
                    help="the serving stream (the headline)")
    ap.add_argument("--plain-serving", default="on", choices=["on", "off"],
                    help="the plain FP64 path on the same stream (the comparison)")
    ap.add_argument("--sessions", type=int, default=None)
    ap.add_argument("--rate", type=float, default=None, help="session arrivals per second")
    ap.add_argument("--budget-gib", type=float, default=None)
    ap.add_argument("--beyond-mall", default="on", choices=["on", "off"],
                    help="the SpMV on one large instance whose matrix exceeds the Infinity Cache")
    ap.add_argument("--large-n", type=int, default=88, help="lattice cells per axis of that instance")
    ap.add_argument("--graphs", default="on", choices=["on", "off"],
                    help="off: launch every kernel directly (for profilers that do not trace "
                         "graph launches)")
    args = ap.parse_args()
    from hobakv1_amd import serve

    for key, attr in (("sessions", "sessions"), ("rate", "rate"), ("budget_gib", "budget_gib")):
        if getattr(args, attr) is None:
            setattr(args, attr, serve.BENCH_STREAM[key])
    return args


def measured_traffic(name, precision):
    """HBM bytes per SpMV launch from the committed rocprofv3 --pmc summary of a workload
    (profiles/cg_spmv_<name>_<precision>_hbm_traffic.md: FETCH_SIZE and WRITE_SIZE scaled by a
    calibration copy kernel), with the algorithmic bytes per launch the same run counted."""
    path = os.path.join(ROOT, "profiles", f"cg_spmv_{name}_{precision}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith((" ", "#", "|")):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    if vals.get("workload") != name or vals.get("precision") != precision:
        return None
    out = {"bytes": float(vals["traffic_bytes_per_launch"]), "kernel": vals.get("kernel"),
           "source": os.path.relpath(path, ROOT)}
    if "algorithmic_bytes_per_launch" in vals:
        out["algorithmic"] = float(vals["algorithmic_bytes_per_launch"])
    if "traffic_to_algorithmic" in vals:
        out["ratio"] = float(vals["traffic_to_algorithmic"])
    return out


def _quiet_stdout(fn, *a):
    """Run fn with stdout sent to stderr (the reference prints progress; stdout carries only the
    one JSON line)."""
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        return fn(*a)
    finally:
        sys.stdout.flush()
        os.dup2(saved, 1)
        os.close(saved)


def _cores():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_baseline_serving(stream, scene_of, steps, seconds):
    """The oracle's FP64 restatement (OpenMP, block-Jacobi PCG, the penalty ground) on the serving
    stream's sessions in arrival order, each for its `steps` step requests, until `seconds` of CPU
    step time have run: tet-Newton-steps/s and the step latency percentiles of that sample."""
    import oracle  # test/baseline infrastructure only

    def run():
        from hobakv1_amd import scenes

        secs, tets, done = [], 0, []
        for s_ in stream:
            sc = scene_of(s_.k)
            o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
            for _ in range(steps):
                _, _, sec = o.step()
                secs.append(sec)
                tets += sc.nt
            o.close()
            done.append(sc.nt)
            if sum(secs) >= seconds:
                break
        total = float(np.sum(secs))
        return {
            "value": tets / total, "unit": "tet-Newton-steps/s", "cores": _cores(), "kind": "port",
            "sample": f"the first {len(done)} sessions of the serving stream (nt {done}), "
                      f"{steps} steps each ({len(secs)} steps, {total:.1f} s), the oracle's FP64 "
                      f"restatement with block-Jacobi PCG tol 1e-8 and the penalty ground (the "
                      f"literal reference has no ground)",
            "p50_step_ms": 1e3 * float(np.percentile(secs, 50)),
            "p99_step_ms": 1e3 * float(np.percentile(secs, 99)),
        }

    return _quiet_stdout(run)


def cpu_baseline(scene, steps, kind):
    """Time the CPU checker on the first `steps` steps of the same scene."""
    import oracle  # test/baseline infrastructure only

    return _quiet_stdout(_cpu_baseline, oracle, scene, steps, kind, _cores())


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
        "p50_step_ms": 1e3 * float(np.percentile(secs, 50)),
        "p99_step_ms": 1e3 * float(np.percentile(secs, 99)),
    }


# The FP32 inner PCG's preconditioner for pinned instances under the fp32 and fp32_batch tuning
# entries (HOBAK_TUNE_AGG 1): the plain FP64 path keeps the reference's block Jacobi.
INNER_PRECOND_PINNED = ("block Jacobi + local aggregates: each chunk of 42 consecutive vertices of a "
                        "Hilbert-curve vertex order solves its own translation (HOBAK_TUNE_AGG)")


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
        "inner_preconditioner": INNER_PRECOND_PINNED,
    }


def c3_errors(lib, scenes, steps):
    """C3: the 40^3 stretch scene in FP32 + FP64 refinement against the plain FP64 GPU path (itself
    within 1e-9 of the FP64 oracle in positions and 1e-12 in forces and stiffness, by the parity
    tests), SNH and ARAP (mu = 3.4e4): after every step the position error, and the error of the
    FP32 forces evaluated at the FP64 path's state; at steps 1, 10 and the last, the error of the
    FP32 stiffness applied to a fixed random vector at that state (normwise relative)."""
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

        pos, frc, kv = [], [], {}
        for k in range(1, steps + 1):
            a.step()
            b.step()
            ua, _ = a.state()
            ub, _ = b.state()
            pos.append(float(np.linalg.norm(ub - ua) / np.linalg.norm(X + ua)))
            with_k = k in (1, 10, steps)
            Ra, Ka = a.forces_hessian(ua, with_k=with_k)
            Rb, Kb = b.forces_hessian(ua, with_k=with_k)
            frc.append(float(np.linalg.norm(Rb - Ra) / np.linalg.norm(Ra)))
            if with_k:
                ya, yb = kw(Ka), kw(Kb)
                kv[str(k)] = float(np.linalg.norm(yb - ya) / np.linalg.norm(ya))
        out[name] = {"rel_position_error_per_step": pos, "max_rel_position_error": max(pos),
                     "rel_force_error_per_step": frc, "max_rel_force_error": max(frc),
                     "rel_stiffness_vector_error_at_steps": kv}
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
               **({"inner_preconditioner": INNER_PRECOND_PINNED} if name == "fp32_refine" else {}),
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
    """The headline: the serving stream in the precision form against the plain FP64 path."""
    from hobakv1_amd import serve

    stream = serve.request_stream(args.sessions, args.rate, args.steps, seed=5)
    cache = {}

    def scene_of(k):
        if k not in cache:
            cache[k] = scenes.serving_instance(k, seed=5)
        return cache[k]

    for s_ in stream:  # the scenes' host-side generation runs before any clock
        scene_of(s_.k)
    budget = int(args.budget_gib * (1 << 30))
    ca = lib.Context(0, precision="fp32", workload="serving")
    ca.set_tuning(graphs=GRAPHS[0])
    cb = None
    if args.plain_serving == "on":
        cb = lib.Context(0)
        cb.set_tuning(graphs=GRAPHS[0])
    if args.warmup > 0:  # untimed sessions of another seed: module loading, clocks, arenas
        warm = serve.request_stream(args.warmup, args.rate, args.steps, seed=6)
        wsc = {w.k: scenes.serving_instance(w.k, seed=6) for w in warm}
        serve.run(ca, warm, lambda k: wsc[k], budget, lib.PRECISION_FP32_REFINE, keep_states=False)
        if cb is not None:
            serve.run(cb, warm, lambda k: wsc[k], budget, lib.PRECISION_FP64, mode="plain",
                      keep_states=False)
    out = serve.compare(ca, cb, stream, scene_of, budget, sample_every=args.sample_every)
    ks = out.pop("spmv", None)
    eng = out.get("engine") or {}
    if ks and ks["spmv_avg_ms"] > 0:
        gbs = ks["spmv_bytes"] / (ks["spmv_avg_ms"] * 1e-3) / 1e9
        tr = measured_traffic("serving_stream", "fp32")
        out["roofline"] = {"bound": "hbm", "kernel": "cg_spmv_kernel<float>", "achieved": round(gbs, 1),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                           # launches differ in size: the profiled run's traffic/algorithmic
                           # ratio, applied to this run's bytes per launch
                           "traffic": round(tr["ratio"] * ks["spmv_bytes"]) if tr and "ratio" in tr
                           else None,
                           "traffic_to_algorithmic": tr.get("ratio") if tr else None,
                           "traffic_source": tr["source"] if tr else None,
                           "bytes_per_launch": ks["spmv_bytes"], "avg_launch_ms": ks["spmv_avg_ms"],
                           "launches_sampled": ks["spmv_count"],
                           "stream_avg_bytes_per_launch": eng.get("spmv_bytes_per_launch"),
                           "note": "average time of every n-th SpMV launch of the serving engine; "
                                   "algorithmic bytes per launch averaged over all its launches, "
                                   "each instance's matrix and vectors counted in the launches "
                                   "that streamed them (counted on the device)"}
    ca.close()
    if cb is not None:
        cb.close()
    out["rate_sessions_per_s"] = args.rate
    out["mean_tets"] = float(np.mean([scene_of(s_.k).nt for s_ in stream]))
    arr = [s_.arrival for s_ in stream]
    out["arrival_span_s"] = float(arr[-1] - arr[0]) if arr else 0.0
    out["workload"] = (f"serving request stream: {args.sessions} sessions, Poisson arrivals "
                       f"{args.rate:g}/s, {args.steps} step requests each, tets log-U[1e3,4e5], "
                       f"aspect [0.5,2], SNH/ARAP, penalty ground y=-0.5 k=1e6, dropped from "
                       f"y in [0,1], HBM budget {args.budget_gib:g} GiB, FP32 + FP64 refinement, "
                       f"block-Jacobi PCG tol 1e-8 (the FP32 inner PCG adds a rigid-translation "
                       f"coarse correction; the plain FP64 path does not); instance set-up outside "
                       f"the clock")
    return out, stream, scene_of


def plain_c2_form(lib, scenes, args):
    """The 40^3 scene on the plain FP64 path: K timed steps after W warmup steps and a reset."""
    sc = scenes.lattice_scene((args.n, args.n, args.n), seed=2, stretch_velocity=True,
                              name=f"cube{args.n}_snh_stretch")
    precond = scenes.PRECOND_BLOCK_JACOBI if args.precond == "block" else scenes.PRECOND_JACOBI
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
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        st = inst.step()
        step_s.append(time.perf_counter() - t1)
        iters.append(st["pcg_iters"][0])
    total = time.perf_counter() - t_start
    ks = inst.kernel_stats()
    u, _ = inst.state()
    finite = bool(np.all(np.isfinite(u)))
    # one more step outside the timed region, with direct launches, so that every phase has
    # events of its own (a captured step times its prologue as one piece)
    ctx.set_tuning(graphs=0)
    inst.step()
    ph = inst.timing_ms()
    el_ms = ph["element"]
    ctx.set_tuning(graphs=GRAPHS[0])
    el_tflops = ELEMENT_FP64_FLOPS_PER_TET * sc.nt / (el_ms * 1e-3) / 1e12 if el_ms > 0 else 0.0
    element = {"kernel": "element_kernel<double>", "bound": "latency", "ms": el_ms,
               "achieved": round(el_tflops, 3), "peak": FP64_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
               "frac": round(el_tflops / FP64_VALU_PEAK_TFLOPS, 4), "valu_busy": ELEMENT_VALU_BUSY,
               "flops_per_tet": ELEMENT_FP64_FLOPS_PER_TET,
               "source": "profiles/element_kernel_40cube_valu.md"}
    spmv_gbs = ks["spmv_bytes"] / (ks["spmv_avg_ms"] * 1e-3) / 1e9 if ks["spmv_avg_ms"] > 0 else 0.0
    tr = measured_traffic(f"{args.n}cube", "fp64")
    roofline = {
        "bound": "hbm", "kernel": "cg_spmv_kernel<double>", "achieved": round(spmv_gbs, 1),
        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(spmv_gbs / HBM_PEAK_GBS, 4),
        "traffic": tr["bytes"] if tr else None, "traffic_source": tr["source"] if tr else None,
        "bytes_per_launch": ks["spmv_bytes"], "avg_launch_ms": ks["spmv_avg_ms"],
        "launches_sampled": ks["spmv_count"],
    }
    out = {
        "value": sc.nt * args.steps / total, "unit": "tet-Newton-steps/s", "steps": args.steps,
        "ms_per_step": 1e3 * total / args.steps,
        "p50_step_ms": 1e3 * float(np.percentile(step_s, 50)),
        "p99_step_ms": 1e3 * float(np.percentile(step_s, 99)),
        "workload": f"{args.n}^3 Kuhn lattice cube ({sc.nt} tets, {sc.nv} vertices), SNH E=1e5 "
                    f"nu=0.45 rho=1000, x=0 pinned, seeded [0.7,1.3] stretch velocity, dt=1/60, "
                    f"1 Newton iteration/step, FP64, "
                    f"{'block-Jacobi' if precond else 'Jacobi'} PCG tol 1e-8",
        "pcg_iters_per_step": iters,
        "phase_ms_direct_launch_step": {k: round(v, 4) for k, v in ph.items()},
        "setup_s": round(setup_s, 3), "finite": finite, "roofline": roofline,
        "element_kernel": element,
    }
    inst.close()
    ctx.close()
    return out, sc, precond, total, u


def _progress(msg):
    """A progress line on stderr (stdout carries only the JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards by instance: run one process per GPU, each with its own instances)",
              file=sys.stderr)
        return 2
    from hobakv1_amd import lib, scenes

    GRAPHS[0] = 1 if args.graphs == "on" else 0
    out = {"metric": METRIC}
    stream = scene_of = None
    if args.serving == "on":
        _progress(f"serving stream: {args.sessions} sessions")
        sv, stream, scene_of = serving_form(lib, scenes, args)
        _progress(f"serving stream done: {sv['value'] / 1e6:.1f} M tet-steps/s")
        pl = sv.get("plain_fp64") or {}
        out.update({
            "value": sv["value"],
            "unit": "tet-Newton-steps/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            # the stream's makespan per step request (requests overlap on the device)
            "ms_per_step": 1e3 * sv["makespan_s"] / max(1, sv["requests"]),
            "p50_step_ms": sv["p50_step_ms"],
            "p99_step_ms": sv["p99_step_ms"],
            "vs_fp64_gpu": sv.get("speedup_vs_plain_fp64"),
            "fp64_gpu": {"value": pl.get("value"), "p50_step_ms": pl.get("p50_step_ms"),
                         "p99_step_ms": pl.get("p99_step_ms")},
            "max_rel_position_error_vs_fp64_gpu": sv.get("max_rel_position_error_vs_plain_fp64"),
            "value_incl_setup": sv.get("value_incl_setup"),
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic",
            "config": {
                "workload": sv["workload"], "sessions": args.sessions, "requests": sv["requests"],
                "rate_sessions_per_s": args.rate, "hbm_budget_gib": args.budget_gib,
                "precision": "fp32 element/assembly/PCG with fp64 refinement (fp64 SVD, residual "
                             "and Newton test)",
                "parallelism": "single GPU",
            },
        })
        if "roofline" in sv:
            out["roofline"] = sv.pop("roofline")
        out["serving"] = sv
    if args.plain_c2 == "on":
        _progress("40^3 plain FP64")
        c2, sc, precond, total, u = plain_c2_form(lib, scenes, args)
        out["plain_fp64_c2"] = c2
        if args.precision_form == "on":
            _progress("40^3 precision form")
            out["fp32_refine"] = precision_form(lib, sc, precond, args, total, u)
            _progress("40^3 precision-form errors")
            out["fp32_refine"]["c3_errors_vs_fp64"] = c3_errors(lib, scenes, args.steps)
        if args.serving != "on":  # the 40^3 scene is the headline when the stream is off
            out.update({"value": c2["value"], "unit": "tet-Newton-steps/s", "n_gpus": 1,
                        "steps": args.steps, "warmup": args.warmup, "ms_per_step": c2["ms_per_step"],
                        "p50_step_ms": c2["p50_step_ms"], "p99_step_ms": c2["p99_step_ms"],
                        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                        "dtype": "f64", "data": "synthetic",
                        "config": {"workload": c2["workload"], "parallelism": "single GPU"},
                        "roofline": c2["roofline"]})
    if args.batch == "on":
        _progress("batch")
        out["batch"] = batch_form(lib, scenes, args)
    if args.beyond_mall == "on":
        _progress("88^3 instance")
        out["beyond_mall"] = beyond_mall_form(lib, scenes, args)
    if args.copy_probe:
        ctx = lib.Context(0)
        nb = 512 << 20
        ms = ctx.copy_probe(nb, 20)
        out["copy_probe"] = {"bytes": nb, "ms": ms, "GBps": 2 * nb / (ms * 1e-3) / 1e9}
        ctx.close()
    if args.cpu_baseline != "none":
        _progress("CPU baselines")
        if stream is not None and args.cpu_baseline in ("auto", "port"):
            out["cpu_baseline"] = cpu_baseline_serving(stream, scene_of, args.steps, args.cpu_seconds)
        if args.plain_c2 == "on":
            cb = cpu_baseline(sc, args.cpu_steps, args.cpu_baseline)
            if cb:
                out["cpu_baseline_c2"] = cb
                if "cpu_baseline" not in out:
                    out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
