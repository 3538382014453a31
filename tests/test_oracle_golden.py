"""The CPU restatement (oracle/) against the literal reference's golden vectors.

The fixtures were written by tests/golden/make_golden.py from oracle/_ref, the reference compiled
from its own sources.  These tests are what makes the oracle trustworthy as the checker of the
GPU path.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, golden
from helpers import bsr_entry_rel, rel, rel_finite
from hobakv1_amd import scenes


@pytest.fixture(scope="module")
def orc():
    return oracle.OracleElement()


def test_svd_rv_matches_reference(orc):
    g = golden("element_fixtures.npz")
    for F, Ur, Sr, Vr in zip(g["F"], g["U"], g["S"], g["V"]):
        U, S, V = orc.svd_rv(F)
        scale = np.abs(F).max()
        assert np.max(np.abs(S - Sr)) <= 1e-14 * scale
        # singular vectors are unique only up to signs within degenerate subspaces; compare the
        # products that the energies use
        assert np.max(np.abs(U @ np.diag(S) @ V.T - Ur @ np.diag(Sr) @ Vr.T)) <= 1e-14 * scale
        assert np.max(np.abs(U @ V.T - Ur @ Vr.T)) <= 1e-12


@pytest.mark.parametrize("key", ["snh", "snh_stiff", "arap"])
def test_material_matches_reference(orc, key):
    g = golden("element_fixtures.npz")
    mat, mu, lam = g[f"{key}_params"]
    mat = int(mat)
    for i, F in enumerate(g["F"]):
        P = orc.pk1(mat, mu, lam, F)
        H = orc.clamped_hessian(mat, mu, lam, F)
        Hr = orc.hessian(mat, mu, lam, F)
        psi = orc.psi(mat, mu, lam, F)
        assert abs(psi - g[f"{key}_psi"][i]) <= 1e-12 * max(1.0, abs(g[f"{key}_psi"][i]))
        assert rel(P, g[f"{key}_P"][i]) <= 1e-12 or np.abs(P - g[f"{key}_P"][i]).max() <= 1e-12 * mu
        assert rel_finite(H, g[f"{key}_H"][i]) <= 1e-11, (key, i)
        assert rel_finite(Hr, g[f"{key}_Hraw"][i]) <= 1e-11, (key, i)


@pytest.mark.parametrize("tag", ["snh", "arap"])
def test_mesh_forces_and_stiffness_match_reference(tag):
    g = golden("mesh_small.npz")
    mat, mu, lam = g[f"{tag}_mat"]
    sc = scenes.Scene(name=tag, rest=g[f"{tag}_rest"], tets=g[f"{tag}_tets"], material=int(mat),
                      mu=mu, lam=lam)
    o = oracle.Oracle(sc)
    rp, ci = o.bsr_pattern()
    assert np.array_equal(rp, g[f"{tag}_row_ptr"]) and np.array_equal(ci, g[f"{tag}_col_idx"])
    R, K = o.forces_hessian(g[f"{tag}_u"])
    assert rel(R, g[f"{tag}_R"]) <= 1e-13
    assert rel(K, g[f"{tag}_K"]) <= 1e-13
    assert bsr_entry_rel(K, g[f"{tag}_K"], rp) <= 1e-12


def _run_oracle(sc, steps, precond):
    o = oracle.Oracle(sc, precond=precond)
    out = []
    for _ in range(steps):
        it, err, _ = o.step()
        u, v = o.state()
        out.append((it, u, v))
    return o, out


def test_c1_trajectory_matches_reference():
    """20 steps of the 10^3 pinned SNH cube with the reference's Jacobi PCG (tol 1e-8)."""
    g = golden("c1_traj.npz")
    sc = scenes.config1()
    _, out = _run_oracle(sc, 20, scenes.PRECOND_JACOBI)
    iters = np.array([o[0] for o in out])
    assert np.max(np.abs(iters - g["iters"])) <= 1
    X = sc.rest.ravel()
    for k in g["keep"]:
        u = out[k - 1][1]
        assert rel(u, g[f"u{k}"]) <= 1e-10, k
        assert rel(X + u, X + g[f"u{k}"]) <= 1e-12, k
        assert rel(out[k - 1][2], g[f"v{k}"]) <= 1e-10, k


def test_c1_first_rhs_and_mass_match_reference():
    g = golden("c1_traj.npz")
    sc = scenes.config1()
    o = oracle.Oracle(sc, precond=scenes.PRECOND_JACOBI)
    o.step()
    rhs, _ = o.last_system()
    b = g["b0"].copy()
    b[np.repeat(sc.pinned(), 3) * 3 + np.tile(np.arange(3), sc.pinned().size)] = 0.0
    assert rel(rhs, b) <= 1e-13
    _, one = o.rest_volumes()
    assert rel(np.repeat(one, 3) * sc.mass_scale, g["mass"]) <= 1e-15


def test_c1_arap_trajectory_matches_reference():
    g = golden("c1_arap.npz")
    sc = scenes.lattice_scene((10, 10, 10), seed=1, material=scenes.MAT_ARAP, mu=3.4e4)
    _, out = _run_oracle(sc, 10, scenes.PRECOND_JACOBI)
    iters = np.array([o[0] for o in out])
    assert np.max(np.abs(iters - g["iters"])) <= 1
    assert rel(out[0][1], g["u1"]) <= 1e-10
    assert rel(out[9][1], g["u10"]) <= 1e-9


def test_c2_first_steps_match_reference():
    """40^3 stretch scene: the first 2 steps (the GPU parity tests check all 20)."""
    g = golden("c2_traj.npz")
    sc = scenes.config2()
    _, out = _run_oracle(sc, 2, scenes.PRECOND_JACOBI)
    s = int(g["stride"])
    for k in (1,):
        assert abs(out[k - 1][0] - g["iters"][k - 1]) <= 1
        assert rel(out[k - 1][1][::s], g[f"u{k}"]) <= 1e-9
    assert abs(out[1][0] - g["iters"][1]) <= 1


def test_c1_bdf1_trajectory_matches_reference():
    """The BDF_1 Newton loop (3 Newton iterations, residual tolerance 1e-7) on C1, 10 steps,
    against the reference's own BDF_1 (src/Timestepper/BDF_1.cpp:78-239)."""
    g = golden("c1_bdf1.npz")
    sc = scenes.config1()
    o = oracle.Oracle(sc, precond=scenes.PRECOND_JACOBI, newton_iters=int(g["newton_iters"]),
                      newton_tol=float(g["newton_tol"]))
    X = sc.rest.ravel()
    for k in range(1, 11):
        o.step()
        assert abs(int(o.last_iters_per_newton[-1]) - int(g["iters_last_newton"][k - 1])) <= 1, k
        if k in (1, 5, 10):
            u, v = o.state()
            assert rel(X + u, X + g[f"u{k}"]) <= 1e-12, k
            assert rel(u, g[f"u{k}"]) <= 1e-10, k
            assert rel(v, g[f"v{k}"]) <= 1e-10, k


def test_drop_plane_constraints_match_reference():
    """A lattice dropped onto a kinematic ground CUBE: the restatement's Baraff-Witkin plane
    constraints (TIMESTEPPER.cpp:252-329, 446-586; BACKWARD_EULER_POSITION.cpp:85-110, 163-314)
    against the reference's own, 40 steps: the same constraint set every step, PCG iterations
    +-1, positions and velocities within 1e-10."""
    g = golden("drop_traj.npz")
    sc = scenes.drop_scene()
    o = oracle.Oracle(sc, precond=scenes.PRECOND_JACOBI)
    for k in range(1, 41):
        it, _, _ = o.step()
        assert abs(it - int(g["iters"][k - 1])) <= 1, k
        pc = np.sort(o.plane_constraints())
        want = g["plane_ids"][g["plane_offs"][k - 1]:g["plane_offs"][k]]
        assert np.array_equal(pc, want), k
        if f"u{k}" in g.files:
            u, v = o.state()
            assert rel(u, g[f"u{k}"]) <= 1e-10, k
            assert rel(v, g[f"v{k}"]) <= 1e-10, k


def test_drop_test_cylinder_matches_reference():
    """src/Scenes/DROP_TEST.h on the reference's own cylinder_10.tobj, 60 steps: the same plane
    constraint set and PCG iterations every step, positions within 1e-9 (contact events amplify
    round-off slowly: ~1e-10 by step 60)."""
    g = golden("drop_test_cylinder.npz")
    sc = scenes.drop_test(os.path.join(GOLDEN, "cylinder_10.tobj"))
    o = oracle.Oracle(sc, precond=scenes.PRECOND_JACOBI)
    for k in range(1, 61):
        it, _, _ = o.step()
        assert abs(it - int(g["iters"][k - 1])) <= 1, k
        want = g["plane_ids"][g["plane_offs"][k - 1]:g["plane_offs"][k]]
        assert np.array_equal(np.sort(o.plane_constraints()), want), k
        if f"u{k}" in g.files:
            u, v = o.state()
            assert rel(u, g[f"u{k}"]) <= 1e-9, k


@pytest.mark.parametrize("tag,kind,E", [("vf", "vertex_face", 3.0), ("ee", "edge_edge", 1.0)])
def test_two_tet_drop_matches_reference_recordings(tag, kind, E):
    """The reference's own regression recordings data/regression/two_tet_drop_{vf,ee}.json
    (src/Scenes/TWO_TET_DROP_*.h, checked by projects/regressionTests at 1e-8): every frame before
    the first self-collision is the plain step with the ground CUBE pinning and colliding.  The
    oracle follows them to 1e-12 (the literal reference built here reproduces them bit for bit)."""
    g = golden("two_tet_drop.npz")
    sc = scenes.two_tet_drop(os.path.join(GOLDEN, f"two_tets_{kind}.tobj"), E=E)
    o = oracle.Oracle(sc, precond=scenes.PRECOND_JACOBI)
    pos, vel = g[f"{tag}_positions"], g[f"{tag}_velocities"]
    assert len(pos) == int(g[f"{tag}_first_collision_frame"]) > 20
    for k in range(len(pos)):
        o.step()
        u, v = o.state()
        assert np.linalg.norm(u - pos[k]) <= 1e-12 * np.linalg.norm(pos[k]), k
        assert np.linalg.norm(v - vel[k]) <= 1e-12 * max(np.linalg.norm(vel[k]), 1e-300), k
    o.close()
