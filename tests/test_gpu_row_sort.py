"""Vertex relabelling by row length (HOBAK_TUNE_ROW_SORT): the device holds the mesh relabelled,
every entry point keeps the caller's numbering.

What must hold, against an instance of the same scene without relabelling:
- forces, stiffness blocks and the last projected system are bitwise equal: the element kernel
  and the ascending-tet gathers see the same tets in the same order, and each row keeps its
  columns in the caller's order, so only the positions of rows and slots move;
- the pattern reported by hobak_get_bsr_pattern is the caller's;
- states after K steps agree to the plain form's bar (PCG iterations +-1, positions 1e-9): the
  row sums are the same, but the PCG dot products are summed in another order;
- pinned vertices, plane constraints and the damping hooks use the caller's ids;
- a relabelled instance batched with others equals itself alone, bitwise.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from helpers import rel
from hobakv1_amd import lib, scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def sorted_ctx():
    ctx = lib.Context(0)
    ctx.set_tuning(row_sort=1)
    yield ctx
    ctx.close()


def test_relabelled_forces_system_and_pattern_bitwise(gpu_ctx, sorted_ctx):
    sc = scenes.lattice_scene((7, 5, 6), seed=3, stretch_velocity=True)
    a = lib.Instance(gpu_ctx, sc)
    b = lib.Instance(sorted_ctx, sc)
    rpa, cia = a.bsr_pattern()
    rpb, cib = b.bsr_pattern()
    assert np.array_equal(rpa, rpb) and np.array_equal(cia, cib)
    u = 0.02 * np.random.default_rng(5).uniform(-1, 1, sc.dofs)
    Ra, Ka = a.forces_hessian(u)
    Rb, Kb = b.forces_hessian(u)
    assert np.array_equal(Ra, Rb) and np.array_equal(Ka, Kb)
    a.step()
    b.step()
    ra, la = a.last_system()
    rb, lb = b.last_system()
    assert np.array_equal(ra, rb) and np.array_equal(la, lb)
    a.close()
    b.close()


def test_relabelled_steps_match_oracle(sorted_ctx):
    """C1 (pinned x = 0 face) in a relabelling context: iterations +-1 and positions 1e-9 of the
    oracle after 20 steps, block Jacobi."""
    sc = scenes.config1()
    inst = lib.Instance(sorted_ctx, sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    for k in range(20):
        st = inst.step()
        it, _, _ = o.step()
        assert abs(st["pcg_iters"][0] - it) <= 1, k
    u, v = inst.state()
    uo, vo = o.state()
    assert rel(u, uo) <= 1e-9 and rel(v, vo) <= 1e-9
    inst.close()
    o.close()


def test_relabelled_contact_and_damping_use_caller_ids(gpu_ctx, sorted_ctx):
    """The drop scene with energy damping: plane constraint ids in the caller's numbering every
    step, states within 1e-9 of the unrelabelled instance, damping hooks bitwise equal."""
    sc = scenes.drop_scene(n=(5, 5, 5), seed=13)
    sc = dataclasses.replace(sc, green_mu=0.05 * sc.mu, fdot_v=sc.v0 + 0.3 * np.sin(np.arange(sc.dofs)))
    a = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    b = lib.Instance(sorted_ctx, sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    u = 0.01 * np.random.default_rng(6).uniform(-1, 1, sc.dofs)
    Ra, Ca = a.damping_forces_hessian(u)
    Rb, Cb = b.damping_forces_hessian(u)
    assert np.array_equal(Ra, Rb) and np.array_equal(Ca, Cb)
    seen = 0
    for k in range(25):
        sa = a.step()
        sb = b.step()
        assert abs(sa["pcg_iters"][0] - sb["pcg_iters"][0]) <= 1, k
        assert np.array_equal(a.plane_constraints(), b.plane_constraints()), k
        seen = max(seen, a.plane_constraints().size)
    assert seen > 0
    assert rel(b.state()[0], a.state()[0]) <= 1e-9
    a.close()
    b.close()


def test_relabelled_batch_equals_alone(sorted_ctx):
    scs = scenes.batch_instances(count=4, seed=41, lo=3, hi=7)
    _say("4 relabelled instances, batched vs alone, 6 steps")
    alone = []
    for sc in scs:
        inst = lib.Instance(sorted_ctx, sc)
        its = [inst.step()["pcg_iters"][0] for _ in range(6)]
        alone.append((inst.state(), its))
        inst.close()
    insts = [lib.Instance(sorted_ctx, sc) for sc in scs]
    its = [[] for _ in insts]
    for _ in range(6):
        for k, st in enumerate(lib.step_batch(sorted_ctx, insts)):
            its[k].append(st["pcg_iters"][0])
    for k, inst in enumerate(insts):
        u, v = inst.state()
        assert its[k] == alone[k][1], k
        assert np.array_equal(u, alone[k][0][0]) and np.array_equal(v, alone[k][0][1]), k
        inst.close()
