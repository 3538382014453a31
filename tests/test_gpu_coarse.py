"""The reduced-precision form's coarse correction over the rigid translations (HOBAK_TUNE_COARSE:
z = M^-1 r + P (P^T A P)^-1 P^T r in the inner FP32 PCG, for instances with no pinned vertex).

Tolerances:
  vs the FP64 oracle       |x - x64| / |x64| <= 1e-6 after the last step, x = X + u, both with the
                           correction on and off (block-Jacobi PCG, tol 1e-8, penalty ground on)
  on vs off                fewer inner PCG iterations over the steps, positions within 1e-6
  batched vs alone         bitwise, with a pinned (correction off) and a free (on) instance in one
                           batch
  pinned after creation    bitwise equal to the same instance in a context with the correction off
"""
import numpy as np
import pytest

import oracle
from helpers import rel
from hobakv1_amd import lib, scenes

pytestmark = pytest.mark.gpu

SEED = 5
STEPS = 4


def _say(msg):
    print(msg, flush=True)


def _free(k, lift=-0.5):
    """serving_instance k with the ground `lift` above its lowest rest vertex (free fall below 0)."""
    sc = scenes.serving_instance(k, seed=SEED)
    sc.ground_y = float(sc.rest[:, 1].min()) + lift
    return sc


@pytest.fixture(scope="module")
def ctx_on():
    ctx = lib.Context(0, precision="fp32", workload="serving")
    yield ctx
    ctx.close()


@pytest.fixture(scope="module")
def ctx_off():
    ctx = lib.Context(0, precision="fp32", workload="serving")
    ctx.set_tuning(coarse=0)
    yield ctx
    ctx.close()


def _run(ctx, sc, steps=STEPS):
    inst = lib.Instance(ctx, sc, precision=lib.PRECISION_FP32_REFINE)
    iters = 0
    for _ in range(steps):
        st = inst.step()
        assert st["status"] == 0
        iters += st["total_pcg_iters"]
    u, v = inst.state()
    inst.close()
    return u, v, iters


@pytest.mark.parametrize("k,lift", [(6, -0.5), (17, -0.5), (0, 0.02)])
def test_coarse_matches_oracle(ctx_on, ctx_off, k, lift):
    sc = _free(k, lift)
    _say(f"serving_instance {k} (nv {sc.nv}, nt {sc.nt}), ground lift {lift}")
    u1, _, i1 = _run(ctx_on, sc)
    u0, _, i0 = _run(ctx_off, sc)
    o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    for _ in range(STEPS):
        o.step()
    uo, _ = o.state()
    o.close()
    X = sc.rest.ravel()
    e1, e0, d = rel(X + u1, X + uo), rel(X + u0, X + uo), rel(X + u1, X + u0)
    _say(f"  inner iterations over {STEPS} steps: on {i1}, off {i0}; rel x vs FP64 oracle: on "
         f"{e1:.2e}, off {e0:.2e}; on vs off {d:.2e}")
    assert e1 <= 1e-6 and e0 <= 1e-6
    assert d <= 1e-6
    if lift < 0:  # free fall: the translation is block Jacobi's slowest mode
        assert i1 < 0.9 * i0, (i1, i0)


def test_coarse_batch_equals_alone(ctx_on):
    """A free instance (correction on) and a pinned lattice (off) in one batch: each ends bitwise
    where it ends stepped alone."""
    scs = [_free(6), scenes.lattice_scene((6, 6, 6), seed=3, stretch_velocity=True), _free(1)]
    alone = []
    for sc in scs:
        u, v, _ = _run(ctx_on, sc, steps=3)
        alone.append((u, v))
    insts = [lib.Instance(ctx_on, sc, precision=lib.PRECISION_FP32_REFINE) for sc in scs]
    for _ in range(3):
        for st in lib.step_batch(ctx_on, insts):
            assert st["status"] == 0
    for i, inst in enumerate(insts):
        u, v = inst.state()
        _say(f"  instance {i}: bitwise {np.array_equal(u, alone[i][0]) and np.array_equal(v, alone[i][1])}")
        assert np.array_equal(u, alone[i][0]) and np.array_equal(v, alone[i][1]), i
        inst.close()


def test_pinning_turns_the_correction_off(ctx_on, ctx_off):
    """Pinning vertices of an instance created free switches the correction off for it: it then
    steps bitwise as in a context where the correction is off; unpinning switches it back on."""
    sc = _free(6)
    ids = np.arange(0, sc.nv, 97, dtype=np.int32)
    inst_on = lib.Instance(ctx_on, sc, precision=lib.PRECISION_FP32_REFINE)
    inst_off = lib.Instance(ctx_off, sc, precision=lib.PRECISION_FP32_REFINE)
    for inst in (inst_on, inst_off):
        inst.set_pinned(ids)
    a = [inst_on.step()["total_pcg_iters"] for _ in range(2)]
    b = [inst_off.step()["total_pcg_iters"] for _ in range(2)]
    ua, va = inst_on.state()
    ub, vb = inst_off.state()
    _say(f"  pinned: iterations in the on-context {a}, in the off-context {b}")
    assert a == b and np.array_equal(ua, ub) and np.array_equal(va, vb)
    for inst in (inst_on, inst_off):
        inst.set_pinned(np.zeros(0, dtype=np.int32))
        inst.set_state(sc.initial_u(), sc.initial_v())
    a = inst_on.step()["total_pcg_iters"]
    b = inst_off.step()["total_pcg_iters"]
    _say(f"  unpinned, free fall: iterations on {a}, off {b}")
    assert a < b
    inst_on.close()
    inst_off.close()


def test_coarse_bdf_newton_matches_oracle(ctx_on, ctx_off):
    """Three Newton iterations per step (BDF_1): every Newton solve of a free instance takes the
    correction; the state after the steps agrees with the FP64 oracle's BDF_1 within 1e-6."""
    sc = _free(6)
    out = []
    for ctx in (ctx_on, ctx_off):
        inst = lib.Instance(ctx, sc, precision=lib.PRECISION_FP32_REFINE, newton_iters=3)
        iters = 0
        for _ in range(3):
            st = inst.step()
            assert st["status"] == 0 and st["newton_iters"] >= 1
            iters += st["total_pcg_iters"]
        out.append((inst.state()[0], iters))
        inst.close()
    o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI, newton_iters=3)
    for _ in range(3):
        o.step()
    uo, _ = o.state()
    o.close()
    X = sc.rest.ravel()
    (u1, i1), (u0, i0) = out
    e1, e0 = rel(X + u1, X + uo), rel(X + u0, X + uo)
    _say(f"  BDF_1: inner iterations on {i1}, off {i0}; rel x vs FP64 oracle on {e1:.2e}, off {e0:.2e}")
    assert e1 <= 1e-6 and e0 <= 1e-6
    assert i1 < i0


def test_coarse_with_point_jacobi(ctx_on, ctx_off):
    """The correction adds to point Jacobi as to block Jacobi: a free instance with the diagonal
    preconditioner, against the FP64 oracle's Jacobi trajectory."""
    sc = _free(6)
    out = []
    for ctx in (ctx_on, ctx_off):
        inst = lib.Instance(ctx, sc, precond=scenes.PRECOND_JACOBI, precision=lib.PRECISION_FP32_REFINE)
        iters = sum(inst.step()["total_pcg_iters"] for _ in range(STEPS))
        out.append((inst.state()[0], iters))
        inst.close()
    o = oracle.Oracle(sc, precond=scenes.PRECOND_JACOBI)
    for _ in range(STEPS):
        o.step()
    uo, _ = o.state()
    o.close()
    X = sc.rest.ravel()
    (u1, i1), (u0, i0) = out
    e1, e0 = rel(X + u1, X + uo), rel(X + u0, X + uo)
    _say(f"  Jacobi: inner iterations on {i1}, off {i0}; rel x vs FP64 oracle on {e1:.2e}, off {e0:.2e}")
    assert e1 <= 1e-6 and e0 <= 1e-6
    assert i1 < i0
