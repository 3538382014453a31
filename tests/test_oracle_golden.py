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
from helpers import bsr_entry_rel, rel, rel_finite, unit_tet_expected
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


def _per_tet(R, K, rp, t):
    return R[12 * t:12 * t + 12], K[rp[4 * t]:rp[4 * t + 4]]


@pytest.mark.parametrize("key", ["snh", "snh_stiff", "arap"])
def test_unit_tet_mesh_matches_reference(key):
    """The element fixtures' deformation gradients as a mesh of disjoint unit tets
    (unit_tets.npz, the reference's own mesh path at that displacement): the oracle's forces and
    stiffness per tet within 1e-12, and, for every tet whose edge matrix is the fixture's F
    bitwise, the fixture's PK1 and clamped dP/dF mapped through dF/dx (helpers.unit_tet_expected)
    equal the reference's mesh output, which pins that mapping for the GPU test that uses it."""
    g = golden("unit_tets.npz")
    e = golden("element_fixtures.npz")
    mat, mu, lam = g[f"{key}_params"]
    sc = scenes.Scene(name=key, rest=g["rest"], tets=g["tets"], material=int(mat), mu=mu, lam=lam)
    o = oracle.Oracle(sc)
    rp, ci = o.bsr_pattern()
    assert np.array_equal(rp, g["row_ptr"]) and np.array_equal(ci, g["col_idx"])
    R, K = o.forces_hessian(g["u"])
    x = (g["rest"].ravel() + g["u"]).reshape(-1, 3)
    exact = 0
    for t in range(len(g["F"])):
        f, k = _per_tet(R, K, rp, t)
        fr, kr = _per_tet(g[f"{key}_R"], g[f"{key}_K"], rp, t)
        assert np.abs(f - fr).max() <= 1e-12 * max(np.abs(fr).max(), mu), (key, t)
        assert rel(k, kr) <= 1e-12, (key, t)
        Ds = np.stack([x[4 * t + a] - x[4 * t] for a in (1, 2, 3)], axis=1)
        if np.array_equal(Ds, e["F"][t]):
            exact += 1
            fm, km = unit_tet_expected(e[f"{key}_P"][t], e[f"{key}_H"][t])
            blocks = np.zeros((4, 4, 3, 3))
            for a in range(4):
                for q in range(rp[4 * t + a], rp[4 * t + a + 1]):
                    blocks[a, ci[q] - 4 * t] = kr[q - rp[4 * t]]
            assert np.abs(fm.ravel() - fr).max() <= 1e-12 * max(np.abs(fr).max(), mu), (key, t)
            assert rel(blocks, km) <= 1e-12, (key, t)
    assert exact >= 20
    o.close()


def test_c1_velocity_form_matches_reference():
    """BACKWARD_EULER_VELOCITY (BACKWARD_EULER_VELOCITY.cpp:107-235) on the C1 scene, Jacobi PCG at
    tol 1e-8, 20 steps (c1_velocity.npz, the reference's own run): iterations equal, states to
    1e-12.  The reference empties its Rayleigh C before forming A (:153, TIMESTEPPER.cpp:657), so
    the trajectory carries no damping; a restatement that kept C would differ by the next check."""
    g = golden("c1_velocity.npz")
    sc = scenes.config1_velocity()
    o = oracle.Oracle(sc, precond=scenes.PRECOND_JACOBI)
    X = sc.rest.ravel()
    for k in range(1, 21):
        it, _, _ = o.step()
        assert it == int(g["iters"][k - 1]), k
        if f"u{k}" in g.files:
            u, v = o.state()
            assert rel(X + u, X + g[f"u{k}"]) <= 1e-12, k
            assert rel(v, g[f"v{k}"]) <= 1e-12, k
    o.close()


def test_velocity_form_drops_rayleigh_damping():
    """The emptied C is what the reference computes: the same scene with the Rayleigh term kept
    (A = M - dt C - dt^2 K, b unchanged) ends measurably elsewhere.  Restated here in numpy from
    the oracle's own system at step 1, so a restatement that kept C fails the fixture check above
    by more than its tolerance."""
    g = golden("c1_velocity.npz")
    sc = scenes.config1_velocity()
    o = oracle.Oracle(sc, precond=scenes.PRECOND_JACOBI)
    o.step()
    rhs, lhs = o.last_system()
    rp, ci = o.bsr_pattern()
    A = oracle.bsr_to_dense(rp, ci, lhs, sc.nv)
    # C = alpha M + beta K_rest, projected like A (pinned rows and columns zero)
    ref = oracle.Oracle(sc, precond=scenes.PRECOND_JACOBI)
    _, Krest = ref.forces_hessian(np.zeros(sc.dofs))
    m = np.zeros(sc.dofs)
    X = sc.rest
    T = sc.tets
    d = np.einsum("ij,ij->i", X[T[:, 3]] - X[T[:, 0]], np.cross(X[T[:, 1]] - X[T[:, 0]], X[T[:, 2]] - X[T[:, 0]]))
    for c in range(4):
        np.add.at(m, 3 * T[:, c], d / 24.0)
    m = sc.mass_scale * np.repeat(m[0::3], 3)
    C = sc.rayleigh_alpha * np.diag(m) + sc.rayleigh_beta * oracle.bsr_to_dense(rp, ci, Krest, sc.nv)
    free = np.ones(sc.dofs, bool)
    free[np.repeat(3 * sc.pinned(), 3) + np.tile([0, 1, 2], sc.pinned().size)] = False
    S = np.diag(free.astype(float))
    Adamped = A - sc.dt * (S @ C @ S)
    dv = np.linalg.solve(A, rhs)
    dv_damped = np.linalg.solve(Adamped, rhs)
    u1 = sc.dt * dv
    assert rel(u1, g["u1"]) <= 1e-7
    assert rel(sc.dt * dv_damped, g["u1"]) > 1e-5
    o.close()
    ref.close()


def test_separation_test_matches_reference():
    """The reference's collision separation test (SCENE_TESTS.h:367-453) on the velocity
    integrator, 42 steps at PCG tolerance 1e-8 (separation_test.npz): the same plane-constraint
    set and iteration count every step, states within 1e-9; at the end every surface vertex is
    above y = 1e-4, the test's own criterion."""
    g = golden("separation_test.npz")
    sc = scenes.separation_test(os.path.join(GOLDEN, "cube_6.tobj"))
    o = oracle.Oracle(sc, precond=scenes.PRECOND_JACOBI)
    X = sc.rest.ravel()
    for k in range(1, 43):
        it, _, _ = o.step()
        assert abs(it - int(g["iters"][k - 1])) <= 1, k
        want = g["plane_ids"][g["plane_offs"][k - 1]:g["plane_offs"][k]]
        assert np.array_equal(np.sort(o.plane_constraints()), want), k
        if f"u{k}" in g.files:
            u, v = o.state()
            assert rel(X + u, X + g[f"u{k}"]) <= 1e-9, k
    u, _ = o.state()
    y = (X + u).reshape(-1, 3)[o.surface_vertices(), 1]
    assert y.min() > 1e-4
    y_ref = (X + g["u42"]).reshape(-1, 3)[o.surface_vertices(), 1]
    assert y_ref.min() > 1e-4
    o.close()
