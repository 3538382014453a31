"""The penalty ground plane (psi = k/2 (y_g - y)^2 below y_g; the serving stream's contact, which
has no reference counterpart) on the GPU.

Its parity is pinned by mathematics instead of a reference fixture:
- the closed-form step of helpers.penalty_ground_case (ARAP with mu = 0, so each vertex solves
  its own equation), to 1e-12, which the oracle meets as well (tests/test_oracle_math.py);
- a lattice dropped onto the plane, 20 steps against the oracle at the plain form's bar
  (PCG iterations +-1 every step, positions and velocities within 1e-9);
- the same drop in the FP32 + refinement form against the oracle, within the serving bar 1e-6.
"""
import numpy as np
import pytest

import oracle
from helpers import penalty_ground_case, rel
from hobakv1_amd import lib, scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _drop(seed=9):
    sc = scenes.lattice_scene((6, 5, 4), seed=seed, E=2e5, nu=0.4, pin=False)
    sc.ground_y = -0.02
    sc.ground_k = 1e6
    sc.v0 = np.tile([0.0, -1.5, 0.0], sc.nv)
    return sc


def test_penalty_ground_closed_form(gpu_ctx):
    sc, du = penalty_ground_case()
    _say(f"penalty ground closed form, nv {sc.nv}")
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    inst.step()
    u, v = inst.state()
    assert rel(u - sc.u0, du) < 1e-12
    assert rel(v, du / sc.dt) < 1e-12
    inst.close()


def test_penalty_ground_drop_matches_oracle(gpu_ctx):
    sc = _drop()
    _say(f"drop onto the penalty ground, nv {sc.nv}, 20 steps against the oracle")
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    touched = False
    for k in range(20):
        st = inst.step()
        it, _, _ = o.step()
        assert abs(st["pcg_iters"][0] - it) <= 1, k
        u, _ = inst.state()
        touched |= bool(np.any(sc.rest[:, 1] + u.reshape(-1, 3)[:, 1] < sc.ground_y))
    assert touched
    u, v = inst.state()
    uo, vo = o.state()
    assert rel(u, uo) <= 1e-9 and rel(v, vo) <= 1e-9
    inst.close()
    o.close()


def test_penalty_ground_drop_fp32_refine(gpu_ctx):
    sc = _drop(seed=10)
    inst = lib.Instance(gpu_ctx, sc, precision=lib.PRECISION_FP32_REFINE)
    o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    for _ in range(20):
        assert inst.step()["pcg_error"][0] < 1e-8
        o.step()
    u, _ = inst.state()
    uo, _ = o.state()
    X = sc.rest.ravel()
    e = rel(X + u, X + uo)
    _say(f"  fp32+refine drop after 20 steps: rel x {e:.3e}")
    assert e <= 1e-6
    inst.close()
    o.close()
