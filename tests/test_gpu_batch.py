"""Batched and continuously batched stepping (hobak_step_batch, hobak_sched_*) against stepping
each instance alone, and against the oracle.

Every kernel of a batch is one segmented launch over all instances; each instance's arithmetic
(its reductions' partial counts and order included) does not depend on what else is in the
batch.  So the bar for "same result alone and in the mix" is bitwise equality of positions,
velocities and PCG iteration counts, stronger than the 1e-6 position tolerance stated for the
serving form.  Against the oracle, batched FP64 instances meet the plain form's tolerances
(iterations +-1, positions 1e-9 relative).
"""
import numpy as np
import pytest

import oracle
from helpers import rel
from hobakv1_amd import lib, scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _mixed_scenes(k=5, seed=21):
    """Small instances of different sizes, materials and stiffness (C4-style, scaled down)."""
    return scenes.batch_instances(count=k, seed=seed, lo=3, hi=9)


def _run_alone(ctx, scs, steps, precision, **kw):
    out = []
    for sc in scs:
        inst = lib.Instance(ctx, sc, precision=precision, **kw)
        its = [inst.step()["pcg_iters"] for _ in range(steps)]
        out.append((inst.state(), its))
        inst.close()
    return out


def _run_batched(ctx, scs, steps, precision, **kw):
    insts = [lib.Instance(ctx, sc, precision=precision, **kw) for sc in scs]
    its = [[] for _ in scs]
    for _ in range(steps):
        st = lib.step_batch(ctx, insts)
        for i, s in enumerate(st):
            its[i].append(s["pcg_iters"])
    out = [(inst.state(), its[i]) for i, inst in enumerate(insts)]
    for inst in insts:
        inst.close()
    return out


@pytest.mark.parametrize("precision", [lib.PRECISION_FP64, lib.PRECISION_FP32_REFINE])
def test_batch_bitwise_equals_alone(gpu_ctx, precision):
    scs = _mixed_scenes()
    _say(f"batch of {len(scs)} (nt {[s.nt for s in scs]}) precision {precision}")
    alone = _run_alone(gpu_ctx, scs, 4, precision)
    batched = _run_batched(gpu_ctx, scs, 4, precision)
    for i, ((ua, va), ia) in enumerate(alone):
        (ub, vb), ib = batched[i]
        assert ia == ib, (i, ia, ib)
        assert np.array_equal(ua, ub) and np.array_equal(va, vb), i


def test_batch_graphs_match_direct_launches():
    scs = _mixed_scenes(k=4, seed=22)
    res = []
    for graphs in (1, 0):
        ctx = lib.Context(0)
        ctx.set_tuning(graphs=graphs)
        res.append(_run_batched(ctx, scs, 3, lib.PRECISION_FP64))
        ctx.close()
    for (sa, ia), (sb, ib) in zip(res[0], res[1]):
        assert ia == ib
        assert np.array_equal(sa[0], sb[0]) and np.array_equal(sa[1], sb[1])


def test_batch_persistent_grid_matches():
    """A capped launch grid (workgroups striding over virtual workgroups) changes nothing."""
    scs = _mixed_scenes(k=4, seed=23)
    res = []
    for cap in (0, 7):
        ctx = lib.Context(0)
        ctx.set_tuning(grid_cap=cap)
        res.append(_run_batched(ctx, scs, 2, lib.PRECISION_FP64))
        ctx.close()
    for (sa, ia), (sb, ib) in zip(res[0], res[1]):
        assert ia == ib
        assert np.array_equal(sa[0], sb[0]) and np.array_equal(sa[1], sb[1])


@pytest.mark.parametrize("precision", [lib.PRECISION_FP64, lib.PRECISION_FP32_REFINE])
def test_streamed_matrix_loads_match(precision):
    """Non-temporal matrix loads (chosen for batches larger than the Infinity Cache; forced here
    with a zero threshold) read the same values: bitwise the same trajectory."""
    scs = _mixed_scenes(k=3, seed=24)
    res = []
    for mb in (192, 0):
        ctx = lib.Context(0, precision="fp64" if precision == lib.PRECISION_FP64 else "fp32")
        ctx.set_tuning(spmv_stream_mb=mb)
        _say(f"stream threshold {mb} MiB, precision {precision}")
        res.append(_run_batched(ctx, scs, 2, precision))
        ctx.close()
    for (sa, ia), (sb, ib) in zip(res[0], res[1]):
        assert ia == ib
        assert np.array_equal(sa[0], sb[0]) and np.array_equal(sa[1], sb[1])


@pytest.mark.parametrize("precision", [lib.PRECISION_FP64, lib.PRECISION_FP32_REFINE])
def test_compacted_groups_match(precision):
    """PCG groups over only the still-solving instances (HOBAK_TUNE_COMPACT_PCT) against groups
    over the whole batch: bitwise the same trajectory, for never, sometimes and always compacting.
    The instances differ in size, so they finish their solves at different ticks."""
    scs = _mixed_scenes(k=6, seed=26)
    res = []
    for pct in (0, 40, 100):
        ctx = lib.Context(0, precision="fp64" if precision == lib.PRECISION_FP64 else "fp32")
        ctx.set_tuning(compact_pct=pct, poll_chunk=4)
        _say(f"compact at {pct}%, precision {precision}")
        res.append(_run_batched(ctx, scs, 3, precision))
        ctx.close()
    for r in res[1:]:
        for (sa, ia), (sb, ib) in zip(res[0], r):
            assert ia == ib
            assert np.array_equal(sa[0], sb[0]) and np.array_equal(sa[1], sb[1])


@pytest.mark.parametrize("groups", [1, 4])
def test_update_groups_match_oracle(groups):
    """The PCG update kernel with 1 and 4 groups of 21 rows per wave (the default is 2): C1 for
    20 steps against the oracle, iterations +-1 and positions within 1e-9; and a batch of
    instances in that layout equals each instance alone, bitwise."""
    ctx = lib.Context(0)
    ctx.set_tuning(update_groups=groups)
    sc = scenes.config1()
    _say(f"update groups {groups}: C1, 20 steps against the oracle")
    inst = lib.Instance(ctx, sc)
    o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    for k in range(20):
        st = inst.step()
        it, _, _ = o.step()
        assert abs(st["pcg_iters"][0] - it) <= 1, k
    u, v = inst.state()
    uo, vo = o.state()
    assert rel(u, uo) <= 1e-9 and rel(v, vo) <= 1e-9
    with pytest.raises(lib.HobakError):  # fixed while the context holds instances
        ctx.set_tuning(update_groups=2)
    inst.close()
    o.close()
    scs = _mixed_scenes(k=3, seed=27)
    alone = _run_alone(ctx, scs, 2, lib.PRECISION_FP64)
    batched = _run_batched(ctx, scs, 2, lib.PRECISION_FP64)
    for (sa, ia), (sb, ib) in zip(alone, batched):
        assert ia == ib
        assert np.array_equal(sa[0], sb[0]) and np.array_equal(sa[1], sb[1])
    with pytest.raises(lib.HobakError):
        ctx.set_tuning(update_groups=3)
    ctx.close()


def test_batch_fp64_matches_oracle(gpu_ctx):
    scs = _mixed_scenes(k=4, seed=24)
    insts = [lib.Instance(gpu_ctx, sc) for sc in scs]
    orcs = [oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI) for sc in scs]
    for k in range(5):
        st = lib.step_batch(gpu_ctx, insts)
        for i, o in enumerate(orcs):
            it, _, _ = o.step()
            assert abs(st[i]["pcg_iters"][0] - it) <= 1, (k, i, st[i]["pcg_iters"], it)
    for i, (inst, o) in enumerate(zip(insts, orcs)):
        u, _ = inst.state()
        uo, _ = o.state()
        X = scs[i].rest.ravel()
        assert rel(X + u, X + uo) <= 1e-9, i
        inst.close()


def test_batch_bdf_newton_equals_alone(gpu_ctx):
    scs = _mixed_scenes(k=3, seed=25)
    alone = _run_alone(gpu_ctx, scs, 2, lib.PRECISION_FP64, newton_iters=3)
    batched = _run_batched(gpu_ctx, scs, 2, lib.PRECISION_FP64, newton_iters=3)
    for i, ((ua, va), ia) in enumerate(alone):
        (ub, vb), ib = batched[i]
        _say(f"  bdf inst {i}: newton iters alone {ia} batched {ib}")
        assert ia == ib
        assert np.array_equal(ua, ub) and np.array_equal(va, vb)


@pytest.mark.parametrize("sample_every", [0, 4])
def test_scheduler_continuous_batching_equals_alone(gpu_ctx, sample_every):
    """Requests arrive while others run; each result equals the instance stepped alone (also with
    the scheduler's SpMV sampling on, whose sampled groups run uncaptured)."""
    scs = [scenes.serving_instance(k, seed=7) for k in range(4)]
    scs = [s for s in scs if s.nt <= 60000][:3] or scs[:2]
    steps = 3
    alone = _run_alone(gpu_ctx, scs, steps, lib.PRECISION_FP32_REFINE)
    insts = [lib.Instance(gpu_ctx, sc, precision=lib.PRECISION_FP32_REFINE) for sc in scs]
    sch = lib.Scheduler(gpu_ctx)
    if sample_every:
        sch.set_kernel_sampling(sample_every)
    remaining = [steps] * len(insts)
    # staggered arrivals: instance i's first request arrives after i ticks
    submitted_first = [False] * len(insts)
    overlapped = False
    running = set()
    tick = 0
    done_at = {}
    while any(remaining) or running:
        for i in range(len(insts)):
            if not submitted_first[i] and tick >= i:
                sch.submit(insts[i], i)
                submitted_first[i] = True
                running.add(i)
        if len(running) > 1:
            overlapped = True
        comps, _ = sch.tick()
        for c in comps:
            i = int(c["tag"])
            assert c["stats"]["status"] == 0
            remaining[i] -= 1
            running.discard(i)
            if remaining[i] > 0:
                sch.submit(insts[i], i)
                running.add(i)
            else:
                done_at[i] = tick
        tick += 1
        assert tick < 100000
    assert overlapped
    for i, inst in enumerate(insts):
        u, v = inst.state()
        (ua, va), _ = alone[i]
        assert np.array_equal(u, ua) and np.array_equal(v, va), i
    ks = sch.kernel_stats()
    if sample_every:
        # bytes per launch over the window: each instance's matrix and vectors count in the
        # launches that streamed them, so the average is positive and at most all of them
        total = sum(_spmv_bytes(inst, 4) for inst in insts)
        assert ks["spmv_count"] > 0 and ks["spmv_avg_ms"] > 0
        assert 0 < ks["spmv_bytes"] <= total
    else:
        assert ks["spmv_count"] == 0
    sch.close()
    for inst in insts:
        inst.close()


def _spmv_bytes(inst, s):
    """Algorithmic bytes of one SpMV over the instance: nnzb (9 s + 4) + 4 (nv + 1) + 2 * 3 s nv."""
    row_ptr, _ = inst.bsr_pattern()
    nnzb = int(row_ptr[-1])
    return nnzb * (9 * s + 4) + 4 * (inst.nv + 1) + 2 * 3 * s * inst.nv


@pytest.mark.parametrize("precision", [lib.PRECISION_FP64, lib.PRECISION_FP32_REFINE])
def test_scheduler_counts_spmv_bytes_of_working_launches(gpu_ctx, precision):
    """The engine's algorithmic SpMV bytes count an instance once per launch that streamed its
    matrix: per (inner) solve, its iterations plus the first launch (p = z), not the launches
    queued after it converged.  Two instances of different sizes finish at different launches of
    the same queued groups."""
    scs = [scenes.lattice_scene((6, 6, 6), seed=3, stretch_velocity=True),
           scenes.lattice_scene((9, 8, 7), seed=4, stretch_velocity=True)]
    ctx = gpu_ctx if precision == lib.PRECISION_FP64 else lib.Context(0, precision="fp32")
    insts = [lib.Instance(ctx, sc, precision=precision) for sc in scs]
    s = 8 if precision == lib.PRECISION_FP64 else 4
    sch = lib.Scheduler(ctx)
    expect = 0
    for _ in range(2):
        for i, inst in enumerate(insts):
            sch.submit(inst, i)
        got = 0
        while got < len(insts):
            comps, _ = sch.tick()
            for c in comps:
                st = c["stats"]
                assert st["status"] == 0
                inst = insts[int(c["tag"])]
                # FP64: pcg_iters per Newton iteration; refinement: inner iterations summed over
                # the sweeps, one first launch per sweep
                n = st["newton_iters"]
                launches = (sum(st["pcg_iters"][:n]) + n if precision == lib.PRECISION_FP64 else
                            st["total_pcg_iters"] + sum(st["refine_iters"][:n]))
                expect += launches * _spmv_bytes(inst, s)
                got += 1
    eng = sch.stats()
    _say(f"  engine SpMV bytes {eng['spmv_bytes']:.0f}, expected {expect}, launches "
         f"{eng['spmv_launches']:.0f}")
    assert eng["spmv_bytes"] == expect
    assert eng["spmv_bytes"] < eng["spmv_launches"] * sum(_spmv_bytes(i, s) for i in insts)
    sch.close()
    for inst in insts:
        inst.close()
    if ctx is not gpu_ctx:
        ctx.close()


@pytest.mark.parametrize("precision", [lib.PRECISION_FP64, lib.PRECISION_FP32_REFINE])
def test_full_c4_batch_step(precision):
    """The C4 batch generator (dims U{4..24}^3, SNH/ARAP, seeded materials) at 300 instances,
    about 4.7 M tets, with the batch tuning entry, one step: every instance bitwise equal to
    stepping it alone, and, in FP64, six instances of assorted size and material against the
    oracle (PCG counts +-1, positions 1e-9).  Beyond 256 instances the segmented instance lookup
    (kernels.h seg_find) takes a second round of prefix loads, which this batch runs."""
    scs = scenes.batch_instances(count=300)
    name = "fp64" if precision == lib.PRECISION_FP64 else "fp32"
    ctx = lib.Context(0, precision=name, workload="batch")
    _say(f"full C4 batch {name}: {len(scs)} instances, {sum(s.nt for s in scs)} tets")
    insts = [lib.Instance(ctx, sc, precision=precision) for sc in scs]
    st = lib.step_batch(ctx, insts)
    batched = [inst.state() for inst in insts]
    for inst, sc in zip(insts, scs):  # back to the initial state, then one step each alone
        inst.set_state(sc.initial_u(), sc.initial_v())
    for i, inst in enumerate(insts):
        sa = inst.step()
        ua, va = inst.state()
        assert sa["pcg_iters"] == st[i]["pcg_iters"], i
        assert np.array_equal(ua, batched[i][0]) and np.array_equal(va, batched[i][1]), i
    _say(f"  bitwise alone == batched for all {len(scs)}")
    if precision == lib.PRECISION_FP64:
        order = np.argsort([sc.nt for sc in scs])
        pick = [int(order[0]), int(order[len(order) // 3]), int(order[len(order) // 2]),
                int(order[2 * len(order) // 3]), int(order[-1])]
        pick += [i for i in range(len(scs)) if scs[i].material != scs[pick[0]].material][:1]
        for i in pick:
            o = oracle.Oracle(scs[i], precond=scenes.PRECOND_BLOCK_JACOBI)
            it, _, _ = o.step()
            uo, _ = o.state()
            X = scs[i].rest.ravel()
            e = rel(X + batched[i][0], X + uo)
            _say(f"  instance {i} nt {scs[i].nt} mat {scs[i].material}: iters {st[i]['pcg_iters'][0]} "
                 f"oracle {it} rel x {e:.2e}")
            assert abs(st[i]["pcg_iters"][0] - it) <= 1
            assert e <= 1e-9
            o.close()
    for inst in insts:
        inst.close()
    ctx.close()
