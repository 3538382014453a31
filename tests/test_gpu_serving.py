"""The serving form (hobakv1_amd/serve.py over hobak_sched_*) on a seeded request stream, with the
tuning the serving benchmark uses (Context(precision="fp32", workload="serving")): FP32 element,
assembly and PCG with FP64 refinement, mixed SNH/ARAP instances of uneven size, the penalty
ground plane in contact, HBM-budget admission that binds, and failure isolation.

Tolerances:
  alone vs in the mix   bitwise (positions, velocities): an instance's arithmetic does not depend
                        on what else is batched with it
  vs the FP64 oracle    |x - x64| / |x64| <= 1e-6 after the session's last step, x = X + u
                        (block-Jacobi PCG, tol 1e-8, the penalty ground on)
"""
import numpy as np
import pytest

import oracle
from helpers import rel
from hobakv1_amd import lib, scenes, serve

pytestmark = pytest.mark.gpu

SEED = 5
SESSIONS = (0, 1, 5, 6, 8, 11)  # serving_instance ids: 4.7k-29k tets, SNH and ARAP
STEPS = 5


def _say(msg):
    print(msg, flush=True)


def _scene(k):
    """serving_instance k with the ground raised to 0.02 above its lowest rest vertex, so the
    bottom layer starts in penalty contact within the test's few steps."""
    sc = scenes.serving_instance(k, seed=SEED)
    sc.ground_y = float(sc.rest[:, 1].min()) + 0.02
    return sc


@pytest.fixture(scope="module")
def serving_ctx():
    ctx = lib.Context(0, precision="fp32", workload="serving")
    yield ctx
    ctx.close()


@pytest.fixture(scope="module")
def alone(serving_ctx):
    """Each session's scene stepped STEPS times on its own, in the same context."""
    out = {}
    for i, k in enumerate(SESSIONS):
        sc = _scene(k)
        inst = lib.Instance(serving_ctx, sc, precision=lib.PRECISION_FP32_REFINE)
        for _ in range(STEPS):
            st = inst.step()
            assert st["status"] == 0
        out[i] = inst.state()
        inst.close()
    return out


def _stream(rate=200.0):
    return serve.request_stream(n_sessions=len(SESSIONS), rate=rate, steps=STEPS, seed=SEED)


def test_serving_stream_budget_binds_and_results_match(serving_ctx, alone):
    """Six sessions arriving within ~30 ms under a budget of 1.5x the largest instance: some
    sessions wait for admission, others share the GPU, and every session ends bitwise where it
    ends alone and within 1e-6 of the FP64 oracle."""
    est = [lib.inst_bytes_estimate(_scene(k).nv, _scene(k).nt, lib.PRECISION_FP32_REFINE)
           for k in SESSIONS]
    budget = int(1.5 * max(est))
    scs = {i: _scene(k) for i, k in enumerate(SESSIONS)}
    res = serve.run(serving_ctx, _stream(), lambda i: scs[i], budget=budget,
                    precision=lib.PRECISION_FP32_REFINE)
    _say(f"serving: {res['requests']} requests, waits {res['admission_waits']}, "
         f"max concurrent {res['max_concurrent_sessions']}, p50 {res['p50_step_ms']:.2f} ms, "
         f"p99 {res['p99_step_ms']:.2f} ms, failed {res['failed']}")
    assert res["requests"] == len(SESSIONS) * STEPS
    assert res["nonzero_status"] == 0 and not res["failed"]
    assert res["admission_waits"] > 0
    assert res["max_concurrent_sessions"] >= 2
    assert serving_ctx.bytes_in_use() == 0
    worst = 0.0
    for i, k in enumerate(SESSIONS):
        u, v = res["final"][i]
        ua, va = alone[i]
        assert np.array_equal(u, ua) and np.array_equal(v, va), i
        sc = scs[i]
        o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
        for _ in range(STEPS):
            o.step()
        uo, _ = o.state()
        X = sc.rest.ravel()
        e = rel(X + u, X + uo)
        worst = max(worst, e)
        _say(f"  session {i} (serving_instance {k}, nt {sc.nt}, mat {sc.material}): "
             f"rel x vs FP64 oracle {e:.2e}")
        assert e <= 1e-6, (i, e)
        o.close()
    _say(f"  worst rel x vs FP64 oracle {worst:.2e}")


def test_nonfinite_instance_is_isolated_in_a_batch(serving_ctx, alone):
    """A NaN in one instance's velocity: its step reports HOBAK_E_NONFINITE, and the instances
    batched with it end bitwise where they end alone."""
    idx = [0, 2, 3]
    insts = [lib.Instance(serving_ctx, _scene(SESSIONS[i]), precision=lib.PRECISION_FP32_REFINE)
             for i in idx]
    bad = 1
    u, v = insts[bad].state()
    v[7] = np.nan
    insts[bad].set_state(u, v)
    sch = lib.Scheduler(serving_ctx)
    left = [STEPS] * len(insts)
    status = {}
    for j, inst in enumerate(insts):
        sch.submit(inst, j)
    ticks = 0
    while True:
        comps, inflight = sch.tick()
        for c in comps:
            j = int(c["tag"])
            left[j] -= 1
            status.setdefault(j, []).append(c["stats"]["status"])
            if j == bad:
                continue  # quarantined: no further requests
            if left[j] > 0:
                sch.submit(insts[j], j)
        ticks += 1
        assert ticks < 100000
        if inflight == 0 and not comps:
            break
    _say(f"  statuses per instance: {status}")
    assert status[bad] == [-5]  # HOBAK_E_NONFINITE
    for j, i in enumerate(idx):
        if j == bad:
            continue
        assert status[j] == [0] * STEPS
        uj, vj = insts[j].state()
        ua, va = alone[i]
        assert np.array_equal(uj, ua) and np.array_equal(vj, va), j
    sch.close()
    for inst in insts:
        inst.close()


def test_serve_quarantines_a_failed_session(serving_ctx, alone):
    """serve.run with one session whose initial velocity holds a NaN: that session is quarantined
    after its first step (its instance destroyed, its later requests dropped); every other
    session completes bitwise as it does alone."""
    scs = {i: _scene(k) for i, k in enumerate(SESSIONS[:4])}
    bad = 2
    v0 = np.zeros(scs[bad].dofs)
    v0[5] = np.nan
    scs[bad].v0 = v0
    stream = serve.request_stream(n_sessions=4, rate=200.0, steps=STEPS, seed=SEED)
    res = serve.run(serving_ctx, stream, lambda i: scs[i], precision=lib.PRECISION_FP32_REFINE)
    _say(f"  failed {res['failed']}, requests {res['requests']}")
    assert res["failed"] == [{"session": bad, "steps_done": 1}]
    assert res["requests"] == 3 * STEPS
    assert bad not in res["final"]
    assert serving_ctx.bytes_in_use() == 0
    for i in range(4):
        if i == bad:
            continue
        u, v = res["final"][i]
        ua, va = alone[i]
        assert np.array_equal(u, ua) and np.array_equal(v, va), i


def test_plain_path_quarantines_a_failed_session(gpu_ctx):
    """The plain FP64 serving loop (one hobak_step at a time) applies the same rule."""
    scs = {i: _scene(k) for i, k in enumerate((6, 11))}
    v0 = np.zeros(scs[0].dofs)
    v0[0] = np.inf
    scs[0].v0 = v0
    stream = serve.request_stream(n_sessions=2, rate=200.0, steps=2, seed=SEED)
    res = serve.run(gpu_ctx, stream, lambda i: scs[i], precision=lib.PRECISION_FP64, mode="plain")
    assert res["failed"] == [{"session": 0, "steps_done": 1}]
    assert list(res["final"]) == [1]


def test_admission_over_budget_is_refused(gpu_ctx):
    """hobak_inst_create returns HOBAK_E_NOMEM when the instance would exceed the context's HBM
    budget, leaves nothing allocated, and succeeds once budget has been freed."""
    sc = scenes.config1()
    need = lib.inst_bytes_estimate(sc.nv, sc.nt, lib.PRECISION_FP64)
    small = lib.Context(0, budget_bytes=need // 4)
    with pytest.raises(lib.HobakError) as e:
        lib.Instance(small, sc)
    assert e.value.code == -3
    assert small.bytes_in_use() == 0
    small.close()
    ctx = lib.Context(0, budget_bytes=int(1.5 * need))
    first = lib.Instance(ctx, sc)
    assert 0 < ctx.bytes_in_use() <= int(1.5 * need)
    with pytest.raises(lib.HobakError) as e:
        lib.Instance(ctx, sc)
    assert e.value.code == -3
    first.close()
    assert ctx.bytes_in_use() == 0
    again = lib.Instance(ctx, sc)
    again.step()
    again.close()
    ctx.close()
