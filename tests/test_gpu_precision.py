"""The precision form: FP32 element kernel, assembly and inner PCG with FP64 iterative refinement
of the linear solve (residual r = b - A x in FP64) to the same 1e-8 tolerance, against the FP64
oracle.

Tolerances, as stated for this form:
  vertex positions x = X + u after 20 steps      |x - x64| / |x64| <= 1e-6
  forces, stiffness-vector products at a state   normwise relative <= 1e-5 (FP32 element math)
"""
import numpy as np
import pytest

import oracle
from helpers import rel
from hobakv1_amd import lib, scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_forces_and_stiffness_fp32_close_to_oracle(gpu_ctx):
    sc = scenes.lattice_scene((6, 6, 6), seed=9, pin=False)
    rng = np.random.default_rng(10)
    u = 0.05 * rng.uniform(-1, 1, sc.dofs)
    inst = lib.Instance(gpu_ctx, sc, precision=lib.PRECISION_FP32_REFINE)
    R, K = inst.forces_hessian(u)
    o = oracle.Oracle(sc)
    Ro, Ko = o.forces_hessian(u)
    rp, ci = inst.bsr_pattern()
    w = rng.uniform(-1, 1, sc.dofs)
    Kw = oracle.bsr_to_dense(rp, ci, K, sc.nv) @ w
    Kow = oracle.bsr_to_dense(rp, ci, Ko, sc.nv) @ w
    _say(f"  fp32 forces rel {rel(R, Ro):.3e}  K rel {rel(K, Ko):.3e}  Kw rel {rel(Kw, Kow):.3e}")
    assert rel(R, Ro) <= 1e-5 and rel(K, Ko) <= 1e-5 and rel(Kw, Kow) <= 1e-5
    inst.close()


def test_c1_fp32_refine_positions_match_fp64_oracle(gpu_ctx):
    sc = scenes.config1()
    inst = lib.Instance(gpu_ctx, sc, precision=lib.PRECISION_FP32_REFINE)
    o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    X = sc.rest.ravel()
    for k in range(20):
        st = inst.step()
        o.step()
        assert st["pcg_error"][0] < 1e-8
    u, _ = inst.state()
    uo, _ = o.state()
    e = rel(X + u, X + uo)
    _say(f"  c1 fp32+refine after 20 steps: rel x {e:.3e} rel u {rel(u, uo):.3e} "
         f"sweeps {st['refine_iters']} inner iters {st['pcg_iters']}")
    assert e <= 1e-6


@pytest.mark.parametrize("material", [scenes.MAT_SNH, scenes.MAT_ARAP])
def test_c3_fp32_refine_against_fp64_oracle(gpu_ctx, material):
    """40^3 stretch scene, 20 steps, SNH or ARAP (mu = 3.4e4): per-step errors of positions,
    forces and stiffness-vector products against the FP64 oracle."""
    sc = scenes.config3(material)
    inst = lib.Instance(gpu_ctx, sc, precision=lib.PRECISION_FP32_REFINE)
    o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    X = sc.rest.ravel()
    rng = np.random.default_rng(0)
    w = rng.uniform(-1, 1, sc.dofs)
    errs = []
    for k in range(20):
        st = inst.step()
        o.step()
        u, _ = inst.state()
        uo, _ = o.state()
        errs.append(rel(X + u, X + uo))
        if k in (0, 9, 19):
            R, K = inst.forces_hessian(uo)
            Ro, Ko = o.forces_hessian(uo)
            rp, ci = inst.bsr_pattern()
            Kw = oracle.bsr_to_dense(rp, ci, K, sc.nv) @ w if sc.nv < 5000 else None
            _say(f"  mat {material} step {k + 1}: rel x {errs[-1]:.3e} rel u {rel(u, uo):.3e} "
                 f"forces {rel(R, Ro):.3e} K {rel(K, Ko):.3e} sweeps {st['refine_iters']} "
                 f"inner {st['pcg_iters']}")
            assert rel(R, Ro) <= 1e-5 and rel(K, Ko) <= 1e-5
    _say(f"  per-step rel x: {[f'{e:.1e}' for e in errs]}")
    assert max(errs) <= 1e-6
    inst.close()
