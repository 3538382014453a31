"""Serving loop: many HBM-resident simulations of uneven size, each request advancing one instance
by one timestep, continuously batched on one GPU.

The reference has no serving layer; its caller steps one scene per SIMULATION_SCENE::stepSimulation
(src/Scenes/SIMULATION_SCENE.h:120-126).  Here a *session* is one simulated object: it arrives
(Poisson arrivals at a fixed rate), is admitted when its instance fits the HBM budget, then issues
`steps` step requests one after another (the next when the previous completes), and leaves,
freeing its memory.

    batched  hobak_sched: every tick finishes the solves that ended, starts the requests that
             arrived since (one segmented launch per prologue kernel over just those), runs
             groups of PCG iterations over the instances still solving (one captured graph
             over every resident instance while all of them solve), and polls
    plain    the same stream served one hobak_step at a time, round-robin over admitted
             sessions (no batching): the plain FP64 GPU path

Clock: instance set-up (host mesh preprocessing, uploads, the rest-state stiffness) and tear-down
run while the GPU is idle between ticks; both loops exclude them from their clocks and report
their total separately, so the throughput figures compare the stepping paths.  The end-to-end
rate, which charges that set-up time as well, is reported beside it (`value_incl_setup`).

Failure isolation: a step whose solve produced a NaN or an infinity (HOBAK_E_NONFINITE) ends its
session at once.  The session is quarantined: its remaining requests are dropped, its instance
is destroyed (freeing its budget), and it is listed under `failed`.  Every other session runs
on unaffected, because no kernel lets one instance's values reach another's.  A step that hits
the PCG iteration limit (HOBAK_E_NOT_CONVERGED) has still advanced the state, as the reference's
does, so its session continues and the step is counted in `nonzero_status`.

    python -m hobakv1_amd.serve --sessions 32 --rate 100 --steps 20
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from collections import deque
from contextlib import contextmanager
from dataclasses import dataclass

import numpy as np

from . import lib, scenes

NONFINITE = -5  # HOBAK_E_NONFINITE

# The stream the serving benchmark and the serving tuning workload measure: enough sessions that
# the budget binds and the run is mostly steady state, arriving faster than the GPU serves them.
BENCH_STREAM = {"sessions": 192, "rate": 160.0, "steps": 20, "budget_gib": 6.0, "seed": 5}


@dataclass
class Session:
    k: int
    arrival: float  # seconds after the stream starts
    steps: int


def request_stream(n_sessions=32, rate=100.0, steps=20, seed=5):
    """Seeded Poisson arrivals (exponential gaps from the SplitMix64 uniform stream); the first
    session arrives at t = 0."""
    u = scenes.uniform(seed, 30, n_sessions, 0.0, 1.0)
    gaps = -np.log1p(-np.minimum(u, 1.0 - 1e-12)) / rate
    t = np.concatenate([[0.0], np.cumsum(gaps[1:])])
    return [Session(k, float(t[k]), steps) for k in range(n_sessions)]


class Clock:
    """Wall clock that can leave out intervals (instance set-up and tear-down)."""

    def __init__(self):
        self.t0 = time.perf_counter()
        self.excluded = 0.0

    def now(self):
        return time.perf_counter() - self.t0 - self.excluded

    @contextmanager
    def paused(self):
        a = time.perf_counter()
        try:
            yield
        finally:
            self.excluded += time.perf_counter() - a


def _pctl(x, q):
    return float(np.percentile(np.asarray(x), q)) if len(x) else 0.0


def _admit(ctx, waiting, live, budget, precision, scene_of, clock, make_inst):
    """Admit arrived sessions in order while their instances fit the budget (a session that does
    not fit waits; the head of the queue is never skipped)."""
    waits = 0
    while waiting:
        s = waiting[0]
        sc = scene_of(s.k)
        need = lib.inst_bytes_estimate(sc.nv, sc.nt, precision)
        if budget and live and ctx.bytes_in_use() + need > budget:
            waits += 1
            break
        waiting.popleft()
        with clock.paused():
            inst = make_inst(sc)
        live[s.k] = {"inst": inst, "left": s.steps, "steps": s.steps, "scene": sc,
                     "submit": clock.now()}
    return waits


def run(ctx, stream, scene_of, budget=0, precision=lib.PRECISION_FP32_REFINE,
        precond=scenes.PRECOND_BLOCK_JACOBI, mode="batched", keep_states=True, sample_every=0):
    """Serve `stream`; returns throughput, latency percentiles and (optionally) every session's
    final displacement.  With sample_every (batched mode), every sample_every-th SpMV launch is
    timed and the result carries the sampled average (`spmv`)."""
    clock = Clock()
    pending = deque(sorted(stream, key=lambda s: s.arrival))
    waiting = deque()
    live = {}
    lat = []
    tet_steps = 0
    admission_waits = 0
    max_live = 0
    final = {}
    iters = []
    statuses = []
    failed = []

    def make_inst(sc):
        return lib.Instance(ctx, sc, precond=precond, precision=precision)

    sch = lib.Scheduler(ctx) if mode == "batched" else None
    if sch is not None and sample_every > 0:
        sch.set_kernel_sampling(sample_every)
    rr = deque()  # plain mode: round-robin queue of session ids with a pending request

    def finish(k, quarantine=False):
        ent = live.pop(k)
        with clock.paused():
            if keep_states and not quarantine:
                final[k] = ent["inst"].state()
            if sch is not None:
                sch.detach(ent["inst"])
            ent["inst"].close()
        if quarantine:
            failed.append({"session": k, "steps_done": ent["steps"] - ent["left"]})

    while pending or waiting or live:
        now = clock.now()
        while pending and pending[0].arrival <= now:
            waiting.append(pending.popleft())
        before = set(live)
        admission_waits += _admit(ctx, waiting, live, budget, precision, scene_of, clock, make_inst)
        for k in live:
            if k not in before:
                live[k]["submit"] = clock.now()
                if sch is not None:
                    sch.submit(live[k]["inst"], k)
                else:
                    rr.append(k)
        max_live = max(max_live, len(live))
        if not live:
            if pending:
                time.sleep(max(0.0, pending[0].arrival - clock.now()))
            continue
        if sch is not None:
            comps, _ = sch.tick()
            now = clock.now()
            for c in comps:
                k = int(c["tag"])
                ent = live[k]
                statuses.append(c["stats"]["status"])
                ent["left"] -= 1
                if c["stats"]["status"] == NONFINITE:
                    finish(k, quarantine=True)
                    continue
                lat.append(now - ent["submit"])
                tet_steps += ent["scene"].nt * c["stats"]["newton_iters"]
                iters.append(c["stats"]["total_pcg_iters"])
                if ent["left"] > 0:
                    ent["submit"] = now
                    sch.submit(ent["inst"], k)
                else:
                    finish(k)
        else:
            k = rr.popleft()
            ent = live[k]
            try:
                st = ent["inst"].step(allow_unconverged=True)
            except lib.HobakError as e:
                if e.code != NONFINITE:
                    raise
                statuses.append(NONFINITE)
                ent["left"] -= 1
                finish(k, quarantine=True)
                continue
            now = clock.now()
            lat.append(now - ent["submit"])
            tet_steps += ent["scene"].nt * st["newton_iters"]
            iters.append(st["total_pcg_iters"])
            statuses.append(st["status"])
            ent["left"] -= 1
            if ent["left"] > 0:
                ent["submit"] = now
                rr.append(k)
            else:
                finish(k)
    makespan = clock.now()
    engine = None
    spmv = None
    if sch is not None:
        engine = sch.stats()
        if sample_every > 0:
            spmv = sch.kernel_stats()
        sch.close()
    return {
        "mode": mode,
        "precision": "fp32_refine" if precision == lib.PRECISION_FP32_REFINE else "fp64",
        "value": tet_steps / makespan if makespan > 0 else 0.0,
        "unit": "tet-Newton-steps/s",
        "value_incl_setup": (tet_steps / (makespan + clock.excluded)
                             if makespan + clock.excluded > 0 else 0.0),
        "makespan_s": makespan,
        "setup_s": clock.excluded,
        "requests": len(lat),
        "tet_newton_steps": tet_steps,
        "p50_step_ms": 1e3 * _pctl(lat, 50),
        "p99_step_ms": 1e3 * _pctl(lat, 99),
        "mean_pcg_iters": float(np.mean(iters)) if iters else 0.0,
        "nonzero_status": int(sum(1 for s in statuses if s != 0)),
        "max_concurrent_sessions": max_live,
        "admission_waits": admission_waits,
        "failed": failed,
        "engine": engine,
        "spmv": spmv,
        "final": final,
    }


def compare(ctx_batched, ctx_plain, stream, scene_of, budget, sample_every=0):
    """Serving (FP32 + refinement, batched) and the plain FP64 path on the same stream; the
    position error of every session after its last step, against the plain FP64 result.  With
    ctx_plain None only the serving side runs."""
    sv = run(ctx_batched, stream, scene_of, budget, lib.PRECISION_FP32_REFINE, mode="batched",
             sample_every=sample_every)
    if ctx_plain is None:
        sv.pop("final")
        return {**sv, "sessions": len(stream), "steps_per_session": stream[0].steps if stream else 0,
                "hbm_budget_bytes": budget}
    pl = run(ctx_plain, stream, scene_of, budget, lib.PRECISION_FP64, mode="plain")
    errs = []
    for k, (u, _) in sv["final"].items():
        if k not in pl["final"]:
            continue
        X = scene_of(k).rest.ravel()
        u64 = pl["final"][k][0]
        errs.append(float(np.linalg.norm(u - u64) / np.linalg.norm(X + u64)))
    sv.pop("final")
    pl.pop("final")
    pl.pop("spmv", None)
    return {
        **sv,
        "plain_fp64": pl,
        "speedup_vs_plain_fp64": sv["value"] / pl["value"] if pl["value"] > 0 else None,
        "max_rel_position_error_vs_plain_fp64": max(errs) if errs else None,
        "sessions": len(stream),
        "steps_per_session": stream[0].steps if stream else 0,
        "hbm_budget_bytes": budget,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sessions", type=int, default=32)
    ap.add_argument("--rate", type=float, default=100.0, help="session arrivals per second")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--budget-gib", type=float, default=3.0)
    ap.add_argument("--seed", type=int, default=5)
    args = ap.parse_args()
    stream = request_stream(args.sessions, args.rate, args.steps, args.seed)
    cache = {}

    def scene_of(k):
        if k not in cache:
            cache[k] = scenes.serving_instance(k, seed=args.seed)
        return cache[k]

    for s in stream:
        scene_of(s.k)
    ca = lib.Context(0, precision="fp32", workload="serving")
    cb = lib.Context(0)
    out = compare(ca, cb, stream, scene_of, int(args.budget_gib * (1 << 30)))
    ca.close()
    cb.close()
    out["rate_sessions_per_s"] = args.rate
    out["mean_tets"] = float(np.mean([scene_of(s.k).nt for s in stream]))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
