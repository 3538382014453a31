"""The velocity-form integrator, BACKWARD_EULER_VELOCITY::solveRayleighDamped
(src/Timestepper/BACKWARD_EULER_VELOCITY.cpp:107-235), on the GPU through the C ABI.

It solves (M - dt^2 K) dv = dt (R + dt K v + f_ext) for the velocity change, then v += dv,
u += dt v, v = S v; the reference empties its Rayleigh matrix C before forming the system (:153,
TIMESTEPPER.cpp:657), so the step carries no Rayleigh damping.

Tolerances (the plain FP64 form's): PCG iterations within +-1 every step; positions x = X + u
within 1e-9 relative; plane-constraint sets identical every step.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, golden
from helpers import rel
from hobakv1_amd import lib, scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_c1_velocity_matches_reference(gpu_ctx):
    """C1 on the reference's own velocity integrator (c1_velocity.npz), Jacobi PCG tol 1e-8."""
    g = golden("c1_velocity.npz")
    sc = scenes.config1_velocity()
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_JACOBI)
    X = sc.rest.ravel()
    diffs = []
    for k in range(1, 21):
        st = inst.step()
        diffs.append(int(st["pcg_iters"][0] - g["iters"][k - 1]))
        if f"u{k}" in g.files:
            u, v = inst.state()
            e = rel(X + u, X + g[f"u{k}"])
            _say(f"  step {k}: iters {st['pcg_iters'][0]} ref {g['iters'][k - 1]} rel x {e:.2e} "
                 f"rel v {rel(v, g[f'v{k}']):.2e}")
            assert e <= 1e-9, k
            assert rel(v, g[f"v{k}"]) <= 1e-8, k
    _say(f"  iteration differences vs reference: {diffs}")
    assert max(abs(d) for d in diffs) <= 1
    inst.close()


def test_c1_velocity_block_jacobi_matches_oracle(gpu_ctx):
    sc = scenes.config1_velocity()
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    X = sc.rest.ravel()
    for k in range(20):
        st = inst.step()
        it, _, _ = o.step()
        u, _ = inst.state()
        uo, _ = o.state()
        assert abs(st["pcg_iters"][0] - it) <= 1, (k, st["pcg_iters"], it)
        assert rel(X + u, X + uo) <= 1e-9, k
    inst.close()
    o.close()


def test_separation_test_matches_reference(gpu_ctx):
    """The reference's collision separation test (projects/unitTests/SCENE_TESTS.h:367-453):
    cube_6.tobj on BACKWARD_EULER_VELOCITY dropped onto the floor CUBE, 42 steps at PCG tol 1e-8,
    against the reference's own run (separation_test.npz): the same constraint set every step,
    iterations +-1, positions 1e-9; at the end every surface vertex is above y = 1e-4."""
    g = golden("separation_test.npz")
    sc = scenes.separation_test(os.path.join(GOLDEN, "cube_6.tobj"))
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_JACOBI)
    X = sc.rest.ravel()
    worst = 0.0
    for k in range(1, 43):
        st = inst.step()
        assert abs(st["pcg_iters"][0] - int(g["iters"][k - 1])) <= 1, (k, st["pcg_iters"])
        want = g["plane_ids"][g["plane_offs"][k - 1]:g["plane_offs"][k]]
        assert np.array_equal(inst.plane_constraints(), want), k
        if f"u{k}" in g.files:
            u, _ = inst.state()
            e = rel(X + u, X + g[f"u{k}"])
            worst = max(worst, e)
            assert e <= 1e-9, k
    u, _ = inst.state()
    o = oracle.Oracle(sc, precond=scenes.PRECOND_JACOBI)
    surf = o.surface_vertices()
    ymin = (X + u).reshape(-1, 3)[surf, 1].min()
    _say(f"  separation test: worst rel x {worst:.2e}, lowest surface y {ymin:.4f}")
    assert ymin > 1e-4
    inst.close()
    o.close()


def test_separation_test_at_eigen_default_tolerance(gpu_ctx):
    """The test as the reference writes it, with Eigen's default CG tolerance (machine epsilon):
    its own criterion, every surface vertex above y = 1e-4 after 42 steps."""
    sc = scenes.separation_test(os.path.join(GOLDEN, "cube_6.tobj"), pcg_tol=0.0)
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_JACOBI)
    for _ in range(42):
        inst.step(allow_unconverged=True)
    u, _ = inst.state()
    o = oracle.Oracle(sc, precond=scenes.PRECOND_JACOBI)
    X = sc.rest.ravel()
    ymin = (X + u).reshape(-1, 3)[o.surface_vertices(), 1].min()
    _say(f"  lowest surface y after 42 steps {ymin:.4f}")
    assert ymin > 1e-4
    inst.close()
    o.close()


def test_velocity_batched_with_position_instances_equals_alone(gpu_ctx):
    """A velocity-form instance batched with position-form ones: every instance bitwise equal
    to stepping it alone."""
    scs = [scenes.config1_velocity(), scenes.lattice_scene((5, 6, 4), seed=3, stretch_velocity=True),
           scenes.separation_test(os.path.join(GOLDEN, "cube_6.tobj"))]
    alone = []
    for sc in scs:
        inst = lib.Instance(gpu_ctx, sc)
        for _ in range(3):
            inst.step()
        alone.append(inst.state())
        inst.close()
    insts = [lib.Instance(gpu_ctx, sc) for sc in scs]
    for _ in range(3):
        lib.step_batch(gpu_ctx, insts)
    for i, inst in enumerate(insts):
        u, v = inst.state()
        assert np.array_equal(u, alone[i][0]) and np.array_equal(v, alone[i][1]), i
        inst.close()


def test_velocity_facade_equals_c_abi(gpu_ctx):
    """lib.BackwardEulerVelocity (the reference-shaped façade) steps bitwise as the C ABI does."""
    sc = scenes.config1_velocity()
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_JACOBI)
    mesh = lib.TetMesh(sc.rest, sc.tets)
    be = lib.BackwardEulerVelocity(mesh, lib.SNH(sc.mu, sc.lam), ctx=gpu_ctx,
                                   precond=scenes.PRECOND_JACOBI)
    be.scaleMass(sc.mass_scale)
    be.addGravity(sc.gravity)
    be.setTolerance(sc.pcg_tol)
    be.attachKinematicConstraints(scenes.Cube(sc.pin_cube[:3], sc.pin_cube[3]))
    for _ in range(4):
        inst.step()
        be.solve()
    u, v = inst.state()
    assert np.array_equal(be.position(), u) and np.array_equal(be.velocity(), v)
    inst.close()


@pytest.mark.parametrize("what", ["fp32", "bdf1", "ground", "green"])
def test_velocity_form_refusals(gpu_ctx, what):
    sc = scenes.config1_velocity()
    kw = {}
    if what == "fp32":
        kw["precision"] = lib.PRECISION_FP32_REFINE
    elif what == "bdf1":
        kw["newton_iters"] = 3
    elif what == "ground":
        sc.ground_y = 0.1
        sc.ground_k = 1e3
    if what == "green":
        inst = lib.Instance(gpu_ctx, sc)
        with pytest.raises(lib.HobakError) as e:
            inst.set_damping(lib.DAMPING_GREEN, 1.0)
        inst.close()
    else:
        with pytest.raises(lib.HobakError) as e:
            lib.Instance(gpu_ctx, sc, **kw)
    assert e.value.code == -1
