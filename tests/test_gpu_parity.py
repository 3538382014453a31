"""Parity of the HIP path (through the C ABI) with the oracle and the reference's golden vectors.

Tolerances (FP64 path):
  forces            normwise relative <= 1e-12
  stiffness K       normwise relative <= 1e-12; entrywise <= 1e-12 relative to the largest |entry|
                    of the entry's scalar row; and entrywise <= 1e-12 relative to the entry itself
                    for every entry above 1e-3 x its row's largest (tests/helpers.py:bsr_entry_rel).
                    Entries are sums of ~16 element terms of row size, so round-off of a few ulps
                    of the row scale is what two correct FP64 orders of summation differ by; an
                    entry a thousand times smaller than its row carries that round-off at ~1e-13
                    of itself.
  PCG iterations    within +-1 of the checker, every step
  positions         |x - x_ref| / |x_ref| <= 1e-9 after 20 steps, x = X + u
"""
import numpy as np
import pytest

import oracle
from conftest import golden
from helpers import bsr_entry_rel, rel
from hobakv1_amd import lib, scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("tag", ["snh", "arap"])
def test_forces_and_stiffness_match_reference(gpu_ctx, tag):
    g = golden("mesh_small.npz")
    mat, mu, lam = g[f"{tag}_mat"]
    sc = scenes.Scene(name=tag, rest=g[f"{tag}_rest"], tets=g[f"{tag}_tets"], material=int(mat),
                      mu=mu, lam=lam)
    _say(f"forces/stiffness {tag} nt={sc.nt}")
    inst = lib.Instance(gpu_ctx, sc)
    rp, ci = inst.bsr_pattern()
    assert np.array_equal(rp, g[f"{tag}_row_ptr"]) and np.array_equal(ci, g[f"{tag}_col_idx"])
    R, K = inst.forces_hessian(g[f"{tag}_u"])
    assert rel(R, g[f"{tag}_R"]) <= 1e-12
    assert rel(K, g[f"{tag}_K"]) <= 1e-12
    e_row = bsr_entry_rel(K, g[f"{tag}_K"], rp, floor=1.0)
    e_own = bsr_entry_rel(K, g[f"{tag}_K"], rp)
    _say(f"  {tag}: forces {rel(R, g[f'{tag}_R']):.2e} K {rel(K, g[f'{tag}_K']):.2e} "
         f"entrywise (row scale) {e_row:.2e} (own scale) {e_own:.2e}")
    assert e_row <= 1e-12
    assert e_own <= 1e-12
    inst.close()


@pytest.mark.parametrize("mat", [scenes.MAT_SNH, scenes.MAT_ARAP])
def test_forces_and_stiffness_match_oracle_large_deformation(gpu_ctx, mat):
    """Stretched / compressed / inverted elements on a 6^3 lattice."""
    sc = scenes.lattice_scene((6, 6, 6), seed=9, material=mat, pin=False)
    rng = np.random.default_rng(10)
    u = 0.08 * rng.uniform(-1, 1, sc.dofs)  # ~0.5 cell size: some tets invert
    _say(f"large deformation mat={mat} nt={sc.nt}")
    inst = lib.Instance(gpu_ctx, sc)
    o = oracle.Oracle(sc)
    R, K = inst.forces_hessian(u)
    Ro, Ko = o.forces_hessian(u)
    rp, _ = inst.bsr_pattern()
    assert rel(R, Ro) <= 1e-12
    assert rel(K, Ko) <= 1e-12
    e_own = bsr_entry_rel(K, Ko, rp)
    _say(f"  forces {rel(R, Ro):.2e} K {rel(K, Ko):.2e} entrywise (row scale) "
         f"{bsr_entry_rel(K, Ko, rp, floor=1.0):.2e} (own scale) {e_own:.2e}")
    assert bsr_entry_rel(K, Ko, rp, floor=1.0) <= 1e-12
    assert e_own <= 1e-12
    inst.close()


def _trajectory(gpu_ctx, sc, steps, precond):
    inst = lib.Instance(gpu_ctx, sc, precond=precond)
    out = []
    for k in range(steps):
        st = inst.step()
        u, v = inst.state()
        out.append((st["pcg_iters"][0], u, v, st))
    return inst, out


def test_c1_block_jacobi_matches_oracle(gpu_ctx):
    """10^3 pinned SNH cube, 20 steps, block-Jacobi PCG tol 1e-8: GPU vs the CPU restatement."""
    sc = scenes.config1()
    _say("c1 block-jacobi 20 steps")
    inst, out = _trajectory(gpu_ctx, sc, 20, scenes.PRECOND_BLOCK_JACOBI)
    o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    X = sc.rest.ravel()
    for k in range(20):
        it, _, _ = o.step()
        uo, vo = o.state()
        assert abs(out[k][0] - it) <= 1, (k, out[k][0], it)
        assert rel(X + out[k][1], X + uo) <= 1e-9, k
    _say(f"  final rel u {rel(out[-1][1], uo):.3e} rel x {rel(X + out[-1][1], X + uo):.3e}")
    inst.close()


def test_c1_jacobi_matches_reference(gpu_ctx):
    """Same scene with the reference's own Jacobi PCG against its golden trajectory."""
    g = golden("c1_traj.npz")
    sc = scenes.config1()
    _say("c1 jacobi 20 steps vs reference")
    inst, out = _trajectory(gpu_ctx, sc, 20, scenes.PRECOND_JACOBI)
    iters = np.array([o[0] for o in out])
    assert np.max(np.abs(iters - g["iters"])) <= 1, (iters, g["iters"])
    X = sc.rest.ravel()
    for k in g["keep"]:
        assert rel(X + out[k - 1][1], X + g[f"u{k}"]) <= 1e-9, k
        assert rel(out[k - 1][1], g[f"u{k}"]) <= 1e-8, k
    inst.close()


def test_c1_arap_jacobi_matches_reference(gpu_ctx):
    g = golden("c1_arap.npz")
    sc = scenes.lattice_scene((10, 10, 10), seed=1, material=scenes.MAT_ARAP, mu=3.4e4)
    _say("c1 arap jacobi 10 steps vs reference")
    inst, out = _trajectory(gpu_ctx, sc, 10, scenes.PRECOND_JACOBI)
    iters = np.array([o[0] for o in out])
    assert np.max(np.abs(iters - g["iters"])) <= 1
    X = sc.rest.ravel()
    assert rel(X + out[9][1], X + g["u10"]) <= 1e-9
    inst.close()


def test_last_system_matches_oracle(gpu_ctx):
    sc = scenes.config1()
    inst = lib.Instance(gpu_ctx, sc)
    inst.step()
    rhs, lhs = inst.last_system()
    o = oracle.Oracle(sc)
    o.step()
    rhs_o, lhs_o = o.last_system()
    rp, _ = inst.bsr_pattern()
    assert rel(rhs, rhs_o) <= 1e-12
    assert rel(lhs, lhs_o) <= 1e-12
    assert bsr_entry_rel(lhs, lhs_o, rp, floor=1.0) <= 1e-12
    assert bsr_entry_rel(lhs, lhs_o, rp) <= 1e-12
    inst.close()


def test_bitwise_repeatable(gpu_ctx):
    """No floating-point atomics: two runs of the same steps give identical bits."""
    sc = scenes.config1()
    res = []
    for _ in range(2):
        inst = lib.Instance(gpu_ctx, sc)
        for _ in range(3):
            inst.step()
        res.append(inst.state())
        inst.close()
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_zero_rhs_takes_no_iterations(gpu_ctx):
    sc = scenes.lattice_scene((1, 1, 1), seed=4, E=2.0, nu=0.0, g=0.0, pin=False)
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_JACOBI)
    st = inst.step()
    assert st["pcg_iters"] == [0] and st["pcg_error"] == [0.0]
    u, v = inst.state()
    assert np.all(u == 0) and np.all(v == 0)
    inst.close()


def test_iteration_cap_reports_not_converged(gpu_ctx):
    sc = scenes.config1()
    sc.pcg_max_iters = 3
    inst = lib.Instance(gpu_ctx, sc)
    with pytest.raises(lib.HobakError) as e:
        inst.step()
    assert e.value.code == -4
    st = inst.step(allow_unconverged=True)
    assert st["pcg_iters"] == [3]
    inst.close()


@pytest.mark.parametrize("damped", [False, True])
def test_assembled_stiffness_is_exactly_symmetric(gpu_ctx, damped):
    """Only a tet's 10 upper blocks are staged and every lower slot is written as its mirror's
    transpose, so the off-diagonal blocks of K (and of the damping stiffness C) are symmetric
    bitwise: block (j, i) equals block (i, j) transposed, entry for entry.  A diagonal block is
    its own mirror, formed in one piece, and symmetric to round-off."""
    sc = scenes.lattice_scene((5, 4, 6), seed=12, material=scenes.MAT_ARAP, stretch_velocity=True)
    if damped:
        sc = scenes.green_scene(n=(5, 4, 6))
    inst = lib.Instance(gpu_ctx, sc)
    u = 0.03 * np.random.default_rng(8).uniform(-1, 1, sc.dofs)
    _, K = inst.damping_forces_hessian(u) if damped else inst.forces_hessian(u)
    rp, ci = inst.bsr_pattern()
    where = {}
    for i in range(sc.nv):
        for k in range(rp[i], rp[i + 1]):
            where[(i, int(ci[k]))] = k
    scale = np.abs(K).max()
    for (i, j), k in where.items():
        if i != j:
            assert np.array_equal(K[k], K[where[(j, i)]].T), (i, j)
        else:
            assert np.abs(K[k] - K[k].T).max() <= 1e-14 * scale, i
    inst.close()


def test_tet_repeating_a_vertex_is_rejected(gpu_ctx):
    sc = scenes.lattice_scene((2, 2, 2), seed=1)
    tets = sc.tets.copy()
    tets[3, 2] = tets[3, 0]
    with pytest.raises(lib.HobakError) as e:
        lib.Instance(gpu_ctx, scenes.Scene(name="bad", rest=sc.rest, tets=tets, material=sc.material,
                                           mu=sc.mu, lam=sc.lam))
    assert e.value.code == -1


def test_c2_jacobi_matches_reference(gpu_ctx):
    """40^3 stretch scene, 20 steps, Jacobi PCG, against the reference's golden trajectory:
    positions within 1e-9 at steps 1, 10 and 20, iteration counts within +-1 every step (the
    differences are printed; test_c2_iterations_match_oracle_on_same_inputs also compares the
    counts on identical inputs)."""
    g = golden("c2_traj.npz")
    sc = scenes.config2()
    s = int(g["stride"])
    _say("c2 jacobi 20 steps vs reference")
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_JACOBI)
    X = sc.rest.ravel()[::s]
    diffs = []
    for k in range(1, 21):
        st = inst.step()
        diffs.append(int(st["pcg_iters"][0] - g["iters"][k - 1]))
        if k in (1, 10, 20):
            u, v = inst.state()
            e = rel(X + u[::s], X + g[f"u{k}"])
            _say(f"  step {k}: iters {st['pcg_iters'][0]} ref {g['iters'][k - 1]} rel x {e:.3e} "
                 f"rel u {rel(u[::s], g[f'u{k}']):.3e}")
            assert e <= 1e-9
    _say(f"  iteration differences vs reference: {diffs}")
    assert max(abs(d) for d in diffs) <= 1
    inst.close()


@pytest.mark.parametrize("precond", [scenes.PRECOND_JACOBI, scenes.PRECOND_BLOCK_JACOBI])
def test_c2_iterations_match_oracle_on_same_inputs(gpu_ctx, precond):
    """Every step of the 40^3 trajectory, the oracle is started from the GPU's own state, so both
    solve the same system: PCG counts within +-1 and the step's displacement within 1e-9."""
    sc = scenes.config2()
    _say(f"c2 same-input iteration check precond={precond}")
    inst = lib.Instance(gpu_ctx, sc, precond=precond)
    o = oracle.Oracle(sc, precond=precond)
    X = sc.rest.ravel()
    diffs = []
    for k in range(20):
        u0, v0 = inst.state()
        o.set_state(u0, v0)
        st = inst.step()
        it, _, _ = o.step()
        u, _ = inst.state()
        uo, _ = o.state()
        diffs.append(int(st["pcg_iters"][0] - it))
        assert rel(X + u, X + uo) <= 1e-9, k
        if k in (0, 19):
            _say(f"  step {k + 1}: rel du {rel(u - u0, uo - u0):.3e}")
    _say(f"  iteration differences vs oracle: {diffs}")
    assert max(abs(d) for d in diffs) <= 1
    inst.close()


def test_c2_block_jacobi_matches_oracle(gpu_ctx):
    """The benchmarked configuration itself: the 40^3 stretch scene with block-Jacobi PCG (tol
    1e-8), 20 free-running steps of the GPU FP64 path and of the oracle side by side (neither is
    restarted from the other): PCG counts within +-1 every step, positions within 1e-9 every step."""
    sc = scenes.config2()
    _say("c2 block-jacobi 20 free-running steps vs oracle")
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    o = oracle.Oracle(sc, precond=scenes.PRECOND_BLOCK_JACOBI)
    X = sc.rest.ravel()
    diffs, errs = [], []
    for k in range(20):
        st = inst.step()
        it, _, _ = o.step()
        u, _ = inst.state()
        uo, _ = o.state()
        diffs.append(int(st["pcg_iters"][0] - it))
        errs.append(rel(X + u, X + uo))
        if k in (0, 9, 19):
            _say(f"  step {k + 1}: iters {st['pcg_iters'][0]} oracle {it} rel x {errs[-1]:.3e} "
                 f"rel u {rel(u, uo):.3e}")
    _say(f"  iteration differences vs oracle: {diffs}")
    _say(f"  max rel x over the steps: {max(errs):.3e}")
    assert max(abs(d) for d in diffs) <= 1
    assert max(errs) <= 1e-9
    inst.close()


def test_c1_bdf1_matches_reference(gpu_ctx):
    """The BDF_1 Newton loop (3 Newton iterations, residual test |S b|^2 > 1e-7, Jacobi PCG) on C1
    against the reference's BDF_1: the last Newton iteration's PCG count within +-1 every step,
    positions within 1e-9 after 1, 5 and 10 steps."""
    g = golden("c1_bdf1.npz")
    sc = scenes.config1()
    _say("C1 BDF_1, 10 steps")
    inst = lib.Instance(gpu_ctx, sc, precond=scenes.PRECOND_JACOBI,
                        newton_iters=int(g["newton_iters"]), newton_tol=float(g["newton_tol"]))
    X = sc.rest.ravel()
    for k in range(1, 11):
        st = inst.step()
        assert abs(st["pcg_iters"][-1] - int(g["iters_last_newton"][k - 1])) <= 1, (k, st)
        if k in (1, 5, 10):
            u, _ = inst.state()
            assert rel(X + u, X + g[f"u{k}"]) <= 1e-9, k
    inst.close()
