"""Plane constraints against kinematic collision cubes on the GPU (hobak_inst_add_collision_cube).

Parity bar, as for the plain FP64 form: PCG iterations within +-1, positions within 1e-9
relative, against the reference's own plane constraints (golden fixture drop_traj.npz, made by
tests/golden/make_golden.py from the literal reference) and against the oracle's restatement; the
set of constrained vertices is discrete and must be identical every step.  A contact instance in
a batch computes bitwise what it computes alone.
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


def test_drop_matches_reference(gpu_ctx):
    """A lattice dropped onto a ground CUBE (after src/Scenes/DROP_TEST.h), Jacobi PCG as the
    reference, 40 steps of contact, sliding and release."""
    g = golden("drop_traj.npz")
    sc = scenes.drop_scene()
    _say(f"drop scene nv {sc.nv} nt {sc.nt}, 40 steps against the reference")
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_JACOBI)
    X = sc.rest.ravel()
    for k in range(1, 41):
        st = inst.step()
        assert abs(st["pcg_iters"][0] - int(g["iters"][k - 1])) <= 1, (k, st["pcg_iters"][0])
        want = g["plane_ids"][g["plane_offs"][k - 1]:g["plane_offs"][k]]
        assert np.array_equal(inst.plane_constraints(), want), k
        if f"u{k}" in g.files:
            u, v = inst.state()
            assert rel(X + u, X + g[f"u{k}"]) <= 1e-9, k
            assert rel(u, g[f"u{k}"]) <= 1e-9, k
            assert rel(v, g[f"v{k}"]) <= 1e-9, k
    inst.close()


def test_drop_block_jacobi_matches_oracle(gpu_ctx):
    """The same scene with the block-Jacobi preconditioner against the oracle, 40 steps."""
    sc = scenes.drop_scene(n=(5, 7, 6), seed=9, speed=2.0)
    _say(f"drop scene {sc.dims}, block Jacobi, 40 steps against the oracle")
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    nmax = 0
    for k in range(1, 41):
        st = inst.step()
        it, _, _ = o.step()
        assert abs(st["pcg_iters"][0] - it) <= 1, k
        pc = inst.plane_constraints()
        assert np.array_equal(pc, np.sort(o.plane_constraints())), k
        nmax = max(nmax, pc.size)
        u, v = inst.state()
        uo, vo = o.state()
        assert rel(u, uo) <= 1e-9, k
        assert rel(v, vo) <= 1e-9, k
    assert nmax > 0  # the scene does make contact
    inst.close()
    o.close()


def test_contact_batched_equals_alone(gpu_ctx):
    """A contact instance batched with instances without collision objects: bitwise the same as
    each alone."""
    scs = [scenes.drop_scene()] + scenes.batch_instances(count=2, seed=31, lo=3, hi=6)
    _say("contact instance + 2 plain instances, batched vs alone, 12 steps")
    alone = []
    for sc in scs:
        inst = lib.Instance(gpu_ctx, sc)
        its = [inst.step()["pcg_iters"][0] for _ in range(12)]
        alone.append((inst.state(), its, inst.plane_constraints()))
        inst.close()
    insts = [lib.Instance(gpu_ctx, sc) for sc in scs]
    its = [[] for _ in scs]
    for _ in range(12):
        for i, s in enumerate(lib.step_batch(gpu_ctx, insts)):
            its[i].append(s["pcg_iters"][0])
    for i, inst in enumerate(insts):
        (ua, va), ia, pa = alone[i]
        ub, vb = inst.state()
        assert ia == its[i], i
        assert np.array_equal(ua, ub) and np.array_equal(va, vb), i
        assert np.array_equal(pa, inst.plane_constraints()), i
        inst.close()


def test_collision_cube_needs_fp64_single_newton():
    sc = scenes.drop_scene()
    ctx = lib.Context(0, precision="fp32")
    with pytest.raises(lib.HobakError):
        lib.Instance(ctx, sc, precision=lib.PRECISION_FP32_REFINE)
    ctx.close()
    ctx = lib.Context(0)
    with pytest.raises(lib.HobakError):
        lib.Instance(ctx, sc, newton_iters=3)
    ctx.close()


def test_drop_test_cylinder_matches_reference(gpu_ctx):
    """src/Scenes/DROP_TEST.h on the reference's cylinder_10.tobj (read by
    hobakv1_amd.formats.read_tobj), Jacobi PCG, 60 steps against the reference: the same plane
    constraint set and PCG iterations +-1 every step, positions within 1e-9 at steps 1, 5, 10,
    20, 40 and 60."""
    g = golden("drop_test_cylinder.npz")
    sc = scenes.drop_test(os.path.join(GOLDEN, "cylinder_10.tobj"))
    _say(f"DROP_TEST cylinder nv {sc.nv} nt {sc.nt}, 60 steps")
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_JACOBI)
    X = sc.rest.ravel()
    for k in range(1, 61):
        st = inst.step()
        assert abs(st["pcg_iters"][0] - int(g["iters"][k - 1])) <= 1, (k, st["pcg_iters"][0])
        want = g["plane_ids"][g["plane_offs"][k - 1]:g["plane_offs"][k]]
        assert np.array_equal(inst.plane_constraints(), want), k
        if f"u{k}" in g.files:
            u, v = inst.state()
            assert rel(X + u, X + g[f"u{k}"]) <= 1e-9, k
            assert rel(u, g[f"u{k}"]) <= 1e-9, k
    inst.close()


@pytest.mark.parametrize("tag,kind,E", [("vf", "vertex_face", 3.0), ("ee", "edge_edge", 1.0)])
def test_two_tet_drop_matches_reference_recordings(gpu_ctx, tag, kind, E):
    """The reference's own regression recordings (data/regression/two_tet_drop_{vf,ee}.json,
    src/Scenes/TWO_TET_DROP_*.h): every frame before the first self-collision, with the ground
    CUBE pinning the bottom tet and acting as a collision object, Jacobi PCG at Eigen's default
    tolerance.  Bar: positions and velocities within 1e-9 relative every frame (the reference's
    own regression test uses 1e-8)."""
    g = golden("two_tet_drop.npz")
    sc = scenes.two_tet_drop(os.path.join(GOLDEN, f"two_tets_{kind}.tobj"), E=E)
    pos, vel = g[f"{tag}_positions"], g[f"{tag}_velocities"]
    _say(f"two_tet_drop_{tag}: {len(pos)} frames against the reference's recording")
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_JACOBI)
    for k in range(len(pos)):
        inst.step(allow_unconverged=True)
        u, v = inst.state()
        assert rel(u, pos[k]) <= 1e-9, k
        assert rel(v, vel[k]) <= 1e-9, k
    inst.close()
