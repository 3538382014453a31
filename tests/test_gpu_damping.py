"""Energy damping on the GPU (hobak_inst_set_damping with HOBAK_DAMPING_GREEN,
hobak_inst_compute_fdots, hobak_eval_damping): GREEN_DAMPING through
BACKWARD_EULER_POSITION::solveEnergyDamped.

Parity bar, as for the plain FP64 form: forces and the damping stiffness within 1e-12 (entrywise,
relative to the largest entry), PCG iterations within +-1 and states within 1e-9 relative after
20 steps, against the literal reference (golden fixture green_damping.npz, written by
tests/golden/make_golden.py from oracle/_ref) and against the oracle.  An energy-damped instance
in a batch computes bitwise what it computes alone.  The FP32 refinement form is held to its own
bar (positions within 1e-6 after 20 steps).
"""
import dataclasses

import numpy as np
import pytest

import oracle
from conftest import golden
from helpers import rel
from hobakv1_amd import lib, scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _maxrel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def _mesh_scene(g):
    m = g["mesh_mat"]
    return scenes.Scene(name="green_mesh", rest=g["mesh_rest"], tets=g["mesh_tets"],
                        material=int(m[0]), mu=float(m[1]), lam=float(m[2]), pin_cube=None,
                        green_mu=float(g["mesh_green_mu"]), fdot_v=g["mesh_fdot_v"])


def test_damping_forces_and_stiffness_match_reference(gpu_ctx):
    """computeInternalForces(hyperelastic, damping) and computeDampingHessian at a seeded state and
    velocity gradient; the elastic-force hook leaves the damping force out."""
    g = golden("green_damping.npz")
    sc = _mesh_scene(g)
    _say(f"damped mesh nv {sc.nv} nt {sc.nt}")
    inst = lib.Instance(gpu_ctx, sc)
    rp, ci = inst.bsr_pattern()
    assert np.array_equal(rp, g["mesh_row_ptr"]) and np.array_equal(ci, g["mesh_col_idx"])
    R, C = inst.damping_forces_hessian(g["mesh_u"])
    _say(f"  R maxrel {_maxrel(R, g['mesh_R']):.2e}  C maxrel {_maxrel(C, g['mesh_C']):.2e}")
    assert _maxrel(R, g["mesh_R"]) <= 1e-12
    assert _maxrel(C, g["mesh_C"]) <= 1e-12
    Re, _ = inst.forces_hessian(g["mesh_u"], with_k=False)
    o = oracle.Oracle(dataclasses.replace(sc, green_mu=None, fdot_v=None))
    Ro, _ = o.forces_hessian(g["mesh_u"], with_k=False)
    assert _maxrel(Re, Ro) <= 1e-12
    assert rel(R, Re) > 1e-3  # the damping force is really there
    inst.close()
    o.close()


def test_energy_damped_steps_match_reference(gpu_ctx):
    """20 solveEnergyDamped steps of scenes.green_scene with the reference's Jacobi PCG."""
    g = golden("green_damping.npz")
    sc = scenes.green_scene()
    _say(f"green scene nv {sc.nv} nt {sc.nt}, 20 steps against the reference")
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_JACOBI)
    for k in range(1, 21):
        st = inst.step()
        assert abs(st["pcg_iters"][0] - int(g["iters"][k - 1])) <= 1, (k, st["pcg_iters"][0])
        if f"u{k}" in g.files:
            u, v = inst.state()
            assert rel(u, g[f"u{k}"]) <= 1e-9, k
            assert rel(v, g[f"v{k}"]) <= 1e-9, k
    inst.close()


@pytest.mark.parametrize("material", [scenes.MAT_SNH, scenes.MAT_ARAP])
def test_energy_damped_block_jacobi_matches_oracle(gpu_ctx, material):
    sc = scenes.green_scene(n=(7, 5, 6), seed=29, material=material, green_frac=0.03)
    _say(f"green scene {sc.dims} material {material}, block Jacobi, 20 steps against the oracle")
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    for k in range(1, 21):
        st = inst.step()
        it, _, _ = o.step()
        assert abs(st["pcg_iters"][0] - it) <= 1, k
    u, v = inst.state()
    uo, vo = o.state()
    assert rel(u, uo) <= 1e-9 and rel(v, vo) <= 1e-9
    inst.close()
    o.close()


def test_fdots_from_current_velocity(gpu_ctx):
    """compute_fdots(None) takes the instance's velocity: the same gradients, bitwise, as passing
    that velocity explicitly."""
    sc = scenes.green_scene(n=(4, 4, 4))
    a = lib.Instance(gpu_ctx, sc)
    b = lib.Instance(gpu_ctx, dataclasses.replace(sc, fdot_v=None))
    b.set_state(v=sc.fdot_v)
    b.compute_fdots()
    u = 0.01 * np.random.default_rng(4).uniform(-1, 1, sc.dofs)
    Ra, Ca = a.damping_forces_hessian(u)
    Rb, Cb = b.damping_forces_hessian(u)
    assert np.array_equal(Ra, Rb) and np.array_equal(Ca, Cb)
    a.close()
    b.close()


def test_energy_damped_batched_equals_alone(gpu_ctx):
    """An energy-damped instance batched with Rayleigh-damped ones: bitwise the same as alone."""
    scs = [scenes.green_scene(n=(6, 5, 4))] + scenes.batch_instances(count=2, seed=37, lo=3, hi=6)
    _say("energy-damped instance + 2 Rayleigh instances, batched vs alone, 10 steps")
    alone = []
    for sc in scs:
        inst = lib.Instance(gpu_ctx, sc)
        its = [inst.step()["pcg_iters"][0] for _ in range(10)]
        alone.append((inst.state(), its))
        inst.close()
    insts = [lib.Instance(gpu_ctx, sc) for sc in scs]
    its = [[] for _ in insts]
    for _ in range(10):
        for k, st in enumerate(lib.step_batch(gpu_ctx, insts)):
            its[k].append(st["pcg_iters"][0])
    for k, inst in enumerate(insts):
        u, v = inst.state()
        assert its[k] == alone[k][1], k
        assert np.array_equal(u, alone[k][0][0]) and np.array_equal(v, alone[k][0][1]), k
        inst.close()


def test_switching_back_to_rayleigh(gpu_ctx):
    """set_damping(RAYLEIGH) after GREEN restores the Rayleigh step bitwise."""
    sc = scenes.lattice_scene((5, 4, 4), seed=3, stretch_velocity=True)
    a = lib.Instance(gpu_ctx, sc)
    a.set_damping(lib.DAMPING_GREEN, 0.05 * sc.mu)
    a.compute_fdots(sc.initial_v())
    a.set_damping(lib.DAMPING_RAYLEIGH)
    b = lib.Instance(gpu_ctx, sc)
    for _ in range(5):
        assert a.step()["pcg_iters"][0] == b.step()["pcg_iters"][0]
    assert np.array_equal(a.state()[0], b.state()[0])
    a.close()
    b.close()


def test_energy_damping_refused_where_the_reference_has_none(gpu_ctx):
    """BDF_1 has no energy-damped solve in the reference (BDF_1.cpp:244-255); Fdot without energy
    damping and the damped-force hook without it are errors too."""
    sc = scenes.lattice_scene((3, 3, 3), seed=2)
    inst = lib.Instance(gpu_ctx, sc, newton_iters=3)
    with pytest.raises(lib.HobakError):
        inst.set_damping(lib.DAMPING_GREEN, 10.0)
    inst.close()
    inst = lib.Instance(gpu_ctx, sc)
    with pytest.raises(lib.HobakError):
        inst.compute_fdots(np.zeros(sc.dofs))
    with pytest.raises(lib.HobakError):
        inst.damping_forces_hessian(np.zeros(sc.dofs))
    with pytest.raises(lib.HobakError):
        inst.set_damping(7, 1.0)
    inst.close()


def test_energy_damping_with_plane_constraints(gpu_ctx):
    """The drop scene with energy damping (Fdot from the falling velocity) against the oracle:
    the same constrained vertices every step, states within 1e-9 after 30 steps."""
    sc = scenes.drop_scene(n=(5, 5, 5), seed=13)
    sc = dataclasses.replace(sc, green_mu=0.05 * sc.mu, fdot_v=sc.v0 + 0.3 * np.sin(np.arange(sc.dofs)))
    _say(f"damped drop scene nv {sc.nv}, 30 steps against the oracle")
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    seen = 0
    for k in range(1, 31):
        st = inst.step()
        it, _, _ = o.step()
        assert abs(st["pcg_iters"][0] - it) <= 1, k
        pc = inst.plane_constraints()
        assert np.array_equal(pc, np.sort(o.plane_constraints())), k
        seen = max(seen, pc.size)
    assert seen > 0
    u, v = inst.state()
    uo, vo = o.state()
    assert rel(u, uo) <= 1e-9 and rel(v, vo) <= 1e-9
    inst.close()
    o.close()


def test_energy_damped_fp32_refine_close_to_oracle(gpu_ctx):
    sc = scenes.green_scene()
    inst = lib.Instance(gpu_ctx, sc, precision=lib.PRECISION_FP32_REFINE)
    o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    X = sc.rest.ravel()
    for _ in range(20):
        st = inst.step()
        o.step()
        assert st["pcg_error"][0] < 1e-8
    u, _ = inst.state()
    uo, _ = o.state()
    e = rel(X + u, X + uo)
    _say(f"  green fp32+refine after 20 steps: rel x {e:.3e} rel u {rel(u, uo):.3e}")
    assert e <= 1e-6
    inst.close()
    o.close()
