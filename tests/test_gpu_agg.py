"""The reduced-precision form's local aggregates (HOBAK_TUNE_AGG): instances held in a spatial
(Hilbert-curve) vertex order, and for those where the rigid-translation correction does not
apply (pinned vertices) the inner FP32 PCG's preconditioner
  z = M^-1 r + sum_I P_I (P_I^T A P_I)^-1 P_I^T r,
I = every chunk of 21 * update-groups consecutive vertices, P_I the translations of its free
vertices.

Tolerances:
  vs the FP64 oracle       |x - x64| / |x64| <= 1e-6 after the last step, x = X + u (block-Jacobi
                           PCG, tol 1e-8), with the aggregates on and off
  on vs off                fewer inner PCG iterations over the steps on a pinned lattice
  batched vs alone         bitwise, with pinned (aggregates) and free (coarse correction)
                           instances in one batch
"""
import numpy as np
import pytest

import oracle
from helpers import rel
from hobakv1_amd import lib, scenes

pytestmark = pytest.mark.gpu

STEPS = 4


def _say(msg):
    print(msg, flush=True)


def _ctx(agg, **tuning):
    ctx = lib.Context(0, precision="fp32")
    ctx.set_tuning(agg=agg, **tuning)
    return ctx


@pytest.fixture(scope="module")
def ctx_agg():
    ctx = _ctx(1)
    yield ctx
    ctx.close()


@pytest.fixture(scope="module")
def ctx_plain():
    ctx = _ctx(0)
    yield ctx
    ctx.close()


def _lattice(n=10, seed=3):
    return scenes.lattice_scene((n, n, n), seed=seed, stretch_velocity=True)


def _run(ctx, sc, steps=STEPS, **kw):
    inst = lib.Instance(ctx, sc, precision=lib.PRECISION_FP32_REFINE, **kw)
    iters = 0
    for _ in range(steps):
        st = inst.step()
        assert st["status"] == 0
        iters += st["total_pcg_iters"]
    u, v = inst.state()
    inst.close()
    return u, v, iters


def _oracle(sc, steps=STEPS, **kw):
    o = oracle.Oracle(sc, **kw)
    for _ in range(steps):
        o.step()
    u, _ = o.state()
    o.close()
    return u


@pytest.mark.parametrize("n", [10, 14])
def test_agg_matches_oracle(ctx_agg, ctx_plain, n):
    sc = _lattice(n)
    _say(f"lattice {n}^3 (nv {sc.nv}, pinned {len(sc.pinned())})")
    u1, _, i1 = _run(ctx_agg, sc)
    u0, _, i0 = _run(ctx_plain, sc)
    uo = _oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    X = sc.rest.ravel()
    e1, e0 = rel(X + u1, X + uo), rel(X + u0, X + uo)
    _say(f"  inner iterations over {STEPS} steps: aggregates {i1}, plain {i0}; rel x vs FP64 oracle: "
         f"{e1:.2e}, {e0:.2e}")
    assert e1 <= 1e-6 and e0 <= 1e-6
    assert i1 < 0.9 * i0, (i1, i0)


@pytest.mark.parametrize("groups,row_sort", [(1, 0), (4, 0), (2, 1)])
def test_agg_chunk_sizes_and_row_sort(ctx_plain, groups, row_sort):
    """Chunks of 21, 84 and 42 vertices (the last ordered by row length within each chunk)."""
    sc = _lattice(12, seed=4)
    ctx = _ctx(1, update_groups=groups, row_sort=row_sort)
    try:
        u1, _, i1 = _run(ctx, sc)
    finally:
        ctx.close()
    u0, _, i0 = _run(ctx_plain, sc)
    uo = _oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    X = sc.rest.ravel()
    e1 = rel(X + u1, X + uo)
    _say(f"  groups {groups}, row sort {row_sort}: iterations {i1} (plain {i0}); rel x vs oracle {e1:.2e}")
    assert e1 <= 1e-6
    assert i1 < i0


def test_agg_with_point_jacobi(ctx_agg, ctx_plain):
    sc = _lattice(12, seed=5)
    u1, _, i1 = _run(ctx_agg, sc, precond=scenes.PRECOND_JACOBI)
    u0, _, i0 = _run(ctx_plain, sc, precond=scenes.PRECOND_JACOBI)
    uo = _oracle(sc, precond=scenes.PRECOND_JACOBI)
    X = sc.rest.ravel()
    e1, e0 = rel(X + u1, X + uo), rel(X + u0, X + uo)
    _say(f"  Jacobi: iterations aggregates {i1}, plain {i0}; rel x vs oracle {e1:.2e}, {e0:.2e}")
    assert e1 <= 1e-6 and e0 <= 1e-6
    assert i1 < i0


def test_agg_bdf_newton(ctx_agg):
    sc = _lattice(10, seed=6)
    u1, _, _ = _run(ctx_agg, sc, steps=3, newton_iters=3)
    uo = _oracle(sc, steps=3, precond=scenes.PRECOND_BLOCK_JACOBI, newton_iters=3)
    X = sc.rest.ravel()
    e1 = rel(X + u1, X + uo)
    _say(f"  BDF_1: rel x vs oracle {e1:.2e}")
    assert e1 <= 1e-6


def test_agg_batch_equals_alone(ctx_agg):
    """Pinned lattices (aggregates) and a free instance (coarse correction) in one batch: each ends
    bitwise where it ends stepped alone."""
    free = scenes.serving_instance(6, seed=5)
    free.ground_y = float(free.rest[:, 1].min()) - 0.5
    scs = [_lattice(10, seed=3), free, _lattice(12, seed=4)]
    alone = [_run(ctx_agg, sc, steps=3)[:2] for sc in scs]
    insts = [lib.Instance(ctx_agg, sc, precision=lib.PRECISION_FP32_REFINE) for sc in scs]
    for _ in range(3):
        for st in lib.step_batch(ctx_agg, insts):
            assert st["status"] == 0
    for i, inst in enumerate(insts):
        u, v = inst.state()
        same = np.array_equal(u, alone[i][0]) and np.array_equal(v, alone[i][1])
        _say(f"  instance {i}: bitwise {same}")
        assert same, i
        inst.close()


def test_pinning_switches_to_aggregates(ctx_agg, ctx_plain):
    """A free instance of the spatial order runs the coarse correction; pinning vertices switches it
    to the aggregates (fewer iterations than without them, the same positions within 1e-6)."""
    sc = scenes.serving_instance(6, seed=5)
    sc.ground_y = float(sc.rest[:, 1].min()) - 0.5
    ids = np.nonzero(sc.rest[:, 1] > sc.rest[:, 1].max() - 1e-9)[0].astype(np.int32)
    a = lib.Instance(ctx_agg, sc, precision=lib.PRECISION_FP32_REFINE)
    b = lib.Instance(ctx_plain, sc, precision=lib.PRECISION_FP32_REFINE)
    for inst in (a, b):
        inst.set_pinned(ids)
    ia = sum(a.step()["total_pcg_iters"] for _ in range(3))
    ib = sum(b.step()["total_pcg_iters"] for _ in range(3))
    ua, _ = a.state()
    ub, _ = b.state()
    X = sc.rest.ravel()
    d = rel(X + ua, X + ub)
    _say(f"  {len(ids)} pinned: iterations aggregates {ia}, plain {ib}; positions {d:.2e} apart")
    assert d <= 1e-6
    assert ia < ib
    a.close()
    b.close()


def test_agg_smaller_than_one_chunk(ctx_agg):
    """A 2x2x2-cell lattice (27 vertices, one partial 42-vertex chunk) with its x = 0 face pinned."""
    sc = _lattice(2, seed=7)
    u1, _, _ = _run(ctx_agg, sc, steps=3)
    uo = _oracle(sc, steps=3, precond=scenes.PRECOND_BLOCK_JACOBI)
    X = sc.rest.ravel()
    e1 = rel(X + u1, X + uo)
    _say(f"  27 vertices: rel x vs oracle {e1:.2e}")
    assert e1 <= 1e-6


def test_agg_every_vertex_pinned(ctx_agg):
    """Every vertex pinned where it rests (zero displacement targets): every chunk's matrix is empty
    (its inverse is taken as zero), and the displacements stay where they are."""
    sc = _lattice(4, seed=8)
    inst = lib.Instance(ctx_agg, sc, precision=lib.PRECISION_FP32_REFINE)
    inst.set_pinned(np.arange(sc.nv, dtype=np.int32))
    u0, _ = inst.state()
    for _ in range(2):
        assert inst.step()["status"] == 0
    u, v = inst.state()
    inst.close()
    _say(f"  all pinned: |u - u0| {np.abs(u - u0).max():.2e}, |v| {np.abs(v).max():.2e}")
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(v))
    assert np.abs(u - u0).max() <= 1e-12


@pytest.fixture(scope="module")
def ctx_both():
    ctx = _ctx(2)
    yield ctx
    ctx.close()


@pytest.mark.parametrize("lift", [0.02, -0.5])
def test_agg_with_the_coarse_correction(ctx_agg, ctx_both, lift):
    """HOBAK_TUNE_AGG 2: a free instance adds the aggregates to the rigid-translation correction.
    Resting on the penalty ground (lift 0.02) that takes fewer inner iterations than the
    translation alone; in free fall (lift -0.5) only the oracle bound is asserted."""
    sc = scenes.serving_instance(6, seed=5)
    sc.ground_y = float(sc.rest[:, 1].min()) + lift
    u2, _, i2 = _run(ctx_both, sc)
    u1, _, i1 = _run(ctx_agg, sc)
    uo = _oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    X = sc.rest.ravel()
    e2, e1 = rel(X + u2, X + uo), rel(X + u1, X + uo)
    _say(f"  lift {lift}: iterations both {i2}, translation alone {i1}; rel x vs oracle {e2:.2e}, {e1:.2e}")
    assert e2 <= 1e-6 and e1 <= 1e-6
    if lift > 0:
        assert i2 < i1


def test_agg_with_the_coarse_correction_batch_equals_alone(ctx_both):
    free = scenes.serving_instance(6, seed=5)
    free.ground_y = float(free.rest[:, 1].min()) + 0.02
    scs = [free, _lattice(10, seed=3)]
    alone = [_run(ctx_both, sc, steps=3)[:2] for sc in scs]
    insts = [lib.Instance(ctx_both, sc, precision=lib.PRECISION_FP32_REFINE) for sc in scs]
    for _ in range(3):
        for st in lib.step_batch(ctx_both, insts):
            assert st["status"] == 0
    for i, inst in enumerate(insts):
        u, v = inst.state()
        same = np.array_equal(u, alone[i][0]) and np.array_equal(v, alone[i][1])
        _say(f"  instance {i}: bitwise {same}")
        assert same, i
        inst.close()
