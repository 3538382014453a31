"""The reference-shaped host façades (C++ hobakv1_amd/host/hobak.hpp through its driver, and the
Python mirror in hobakv1_amd/lib.py) against the C ABI and the oracle.

The façades add no arithmetic: they upload position()/velocity(), call hobak_step and download.
So against a direct C-ABI instance of the same scene the bar is bitwise equality; against the
oracle it is the plain form's (PCG iterations +-1 per step, positions <= 1e-9 relative).
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle
from helpers import rel
from hobakv1_amd import build as hb
from hobakv1_amd import lib, scenes

pytestmark = pytest.mark.gpu


def _write_scene(path, sc, steps):
    cube = sc.pin_cube if sc.pin_cube is not None else (0.0, 0.0, 0.0, 1.0)
    ground = sc.collision_cubes[0] if sc.collision_cubes else (0.0, 0.0, 0.0, 1.0)
    flags = ((1 if sc.pin_cube is not None else 0) | (2 if sc.collision_cubes else 0)
             | (4 if sc.green_mu is not None else 0)
             | (8 if sc.integrator == scenes.INTEGRATOR_VELOCITY else 0))
    with open(path, "wb") as f:
        f.write(struct.pack("<5i", sc.nv, sc.nt, sc.material, steps, flags))
        f.write(struct.pack("<18d", sc.mu, sc.lam, sc.dt, sc.rayleigh_alpha, sc.rayleigh_beta,
                            sc.mass_scale, *sc.gravity, sc.pcg_tol, *cube, *ground))
        f.write(np.ascontiguousarray(sc.rest, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(sc.tets, dtype="<i4").tobytes())
        f.write(np.ascontiguousarray(sc.initial_u(), dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(sc.initial_v(), dtype="<f8").tobytes())
        if sc.green_mu is not None:
            fv = sc.fdot_v if sc.fdot_v is not None else np.zeros(sc.dofs)
            f.write(struct.pack("<d", sc.green_mu))
            f.write(np.ascontiguousarray(fv, dtype="<f8").tobytes())


def _read_out(path, dofs, planes=False):
    with open(path, "rb") as f:
        (n,) = struct.unpack("<i", f.read(4))
        iters = np.frombuffer(f.read(4 * n), dtype="<i4").tolist()
        counts = np.frombuffer(f.read(4 * n), dtype="<i4").tolist()
        u = np.frombuffer(f.read(8 * dofs), dtype="<f8").copy()
        v = np.frombuffer(f.read(8 * dofs), dtype="<f8").copy()
    return (iters, u, v, counts) if planes else (iters, u, v)


def test_cpp_facade_matches_c_abi_and_oracle(tmp_path):
    assert os.path.exists(hb.DRIVER), "build the façade driver (python -m hobakv1_amd.build)"
    sc = scenes.config1()
    steps = 5
    scene_bin = tmp_path / "scene.bin"
    out_bin = tmp_path / "out.bin"
    _write_scene(scene_bin, sc, steps)
    print(f"C++ façade: C1 ({sc.nt} tets), {steps} steps", flush=True)
    subprocess.run([hb.DRIVER, str(scene_bin), str(out_bin)], check=True, timeout=120)
    iters, u, v = _read_out(out_bin, sc.dofs)

    # the façade's default context uses the library's default launch shapes (no tuning table);
    # the same shapes give the same summation order, hence bitwise-equal states
    ctx = lib.Context(0, apply_table=False)
    inst = lib.Instance(ctx, sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    direct = [inst.step()["pcg_iters"][0] for _ in range(steps)]
    ud, vd = inst.state()
    inst.close()
    ctx.close()
    assert iters == direct
    assert np.array_equal(u, ud) and np.array_equal(v, vd)

    o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    X = sc.rest.ravel()
    for k in range(steps):
        it, _, _ = o.step()
        assert abs(iters[k] - it) <= 1, (k, iters[k], it)
    uo, _ = o.state()
    assert rel(X + u, X + uo) <= 1e-9


def test_python_facade_matches_c_abi(gpu_ctx):
    sc = scenes.config1()
    mesh = lib.TetMesh(sc.rest, sc.tets)
    mat = lib.SNH(sc.mu, sc.lam)
    be = lib.BackwardEulerPosition(mesh, mat, ctx=gpu_ctx)
    be.setDt(sc.dt)
    be.setRayleigh(sc.rayleigh_alpha, sc.rayleigh_beta)
    be.scaleMass(sc.mass_scale)
    be.addGravity(sc.gravity)
    be.setTolerance(sc.pcg_tol)
    be.attachKinematicConstraints(scenes.Cube(scenes.PIN_X0_CUBE[:3], scenes.PIN_X0_CUBE[3]))
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    for _ in range(3):
        be.solve(False)
        inst.step()
    ud, vd = inst.state()
    inst.close()
    assert np.array_equal(be.position(), ud) and np.array_equal(be.velocity(), vd)


def test_cpp_facade_collision_cube_matches_reference(tmp_path):
    """addKinematicCollisionObject through the C++ façade: the drop scene's constraint counts
    and PCG iterations (Jacobi) follow the reference's, and the state after 20 steps is within
    1e-9 of it."""
    from conftest import golden

    g = golden("drop_traj.npz")
    sc = scenes.drop_scene()
    steps = 20
    scene_bin = tmp_path / "scene.bin"
    out_bin = tmp_path / "out.bin"
    _write_scene(scene_bin, sc, steps)
    env = dict(os.environ, HOBAK_FACADE_PRECOND="0")
    print(f"C++ façade: drop scene ({sc.nt} tets), {steps} steps", flush=True)
    subprocess.run([hb.DRIVER, str(scene_bin), str(out_bin)], check=True, timeout=120, env=env)
    iters, u, v, counts = _read_out(out_bin, sc.dofs, planes=True)
    assert counts == g["counts"][:steps].tolist()
    assert max(abs(a - b) for a, b in zip(iters, g["iters"][:steps])) <= 1
    X = sc.rest.ravel()
    assert rel(X + u, X + g["u20"]) <= 1e-9
    assert rel(v, g["v20"]) <= 1e-9


def test_cpp_facade_energy_damping_matches_reference(tmp_path):
    """setDamping(GREEN_DAMPING) and TET_MESH::computeFdots through the C++ façade: the energy-damped
    steps of scenes.green_scene follow the reference's (Jacobi PCG) to +-1 iteration and 1e-9."""
    from conftest import golden

    g = golden("green_damping.npz")
    sc = scenes.green_scene()
    steps = 20
    scene_bin = tmp_path / "scene.bin"
    out_bin = tmp_path / "out.bin"
    _write_scene(scene_bin, sc, steps)
    env = dict(os.environ, HOBAK_FACADE_PRECOND="0")
    print(f"C++ façade: energy-damped scene ({sc.nt} tets), {steps} steps", flush=True)
    subprocess.run([hb.DRIVER, str(scene_bin), str(out_bin)], check=True, timeout=120, env=env)
    iters, u, v = _read_out(out_bin, sc.dofs)
    assert max(abs(a - b) for a, b in zip(iters, g["iters"][:steps])) <= 1
    assert rel(u, g["u20"]) <= 1e-9
    assert rel(v, g["v20"]) <= 1e-9


def test_python_facade_energy_damping_matches_c_abi(gpu_ctx):
    sc = scenes.green_scene(n=(5, 5, 5))
    mesh = lib.TetMesh(sc.rest, sc.tets)
    be = lib.BackwardEulerPosition(mesh, lib.SNH(sc.mu, sc.lam), ctx=gpu_ctx)
    be.setDt(sc.dt)
    be.scaleMass(sc.mass_scale)
    be.addGravity(sc.gravity)
    be.setTolerance(sc.pcg_tol)
    be.attachKinematicConstraints(scenes.Cube(scenes.PIN_X0_CUBE[:3], scenes.PIN_X0_CUBE[3]))
    be.setDamping(lib.GREEN_DAMPING(sc.green_mu))
    mesh.computeFdots(sc.fdot_v)
    be.velocity()[:] = sc.initial_v()
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    for _ in range(4):
        be.solve(False)
        inst.step()
    ud, vd = inst.state()
    inst.close()
    assert np.array_equal(be.position(), ud) and np.array_equal(be.velocity(), vd)


def test_cpp_facade_velocity_integrator_matches_reference(tmp_path):
    """BACKWARD_EULER_VELOCITY through the C++ façade on the reference's collision separation
    test (SCENE_TESTS.h:367-453, Jacobi PCG tol 1e-8): the same plane-constraint counts every
    step and PCG iterations +-1 as the reference's own run, the state within 1e-9 of it after 20
    steps, and bitwise the state of the C ABI."""
    from conftest import GOLDEN, golden

    g = golden("separation_test.npz")
    sc = scenes.separation_test(os.path.join(GOLDEN, "cube_6.tobj"))
    steps = 20
    scene_bin = tmp_path / "scene.bin"
    out_bin = tmp_path / "out.bin"
    _write_scene(scene_bin, sc, steps)
    env = dict(os.environ, HOBAK_FACADE_PRECOND="0")
    print(f"C++ façade: velocity integrator, separation test ({sc.nt} tets), {steps} steps", flush=True)
    subprocess.run([hb.DRIVER, str(scene_bin), str(out_bin)], check=True, timeout=120, env=env)
    iters, u, v, counts = _read_out(out_bin, sc.dofs, planes=True)
    assert counts == g["counts"][:steps].tolist()
    assert max(abs(a - b) for a, b in zip(iters, g["iters"][:steps])) <= 1
    X = sc.rest.ravel()
    assert rel(X + u, X + g["u20"]) <= 1e-9
    ctx = lib.Context(0, apply_table=False)
    inst = lib.Instance(ctx, sc, precond=scenes.PRECOND_JACOBI)
    for _ in range(steps):
        inst.step()
    ud, vd = inst.state()
    inst.close()
    ctx.close()
    assert np.array_equal(u, ud) and np.array_equal(v, vd)
