"""The GPU element path at the reference's hard deformation gradients.

tests/golden/unit_tets.npz holds the 48 deformation gradients of element_fixtures.npz as one mesh
of disjoint unit tets (rest corners 0, e_x, e_y, e_z, so Dm = I and F is the deformed edge
matrix), with the reference's own mesh forces and clamped stiffness at that displacement.  The set
covers repeated singular values, sigma = 1e-6 and 1e-3, reflections (sigma = (1, 1, -1)),
improper U and V, near-rank-deficient F, rotations and the Sigma sets of the reference's
hyperelastic unit tests (projects/unitTests/HYPERELASTIC_TESTS.h:181-312), so the device's own
svd_rv (src/util/MATRIX_UTIL.cpp:76-95) and its singular-value-frame stiffness run where the
reference's branches differ.

Tolerance (FP64 path, per tet): forces within 1e-12 of the tet's largest force entry (or of mu,
where the force vanishes, e.g. SNH at F = I); the tet's 12x12 stiffness within 1e-12 normwise.
The ARAP reflections have sigma_j + sigma_k = 0 up to round-off, where the reference's clamped
eigenvalue mu (2 - 4 / (sigma_j + sigma_k)) is of order 1e15: normwise agreement per tet still
holds there.
"""
import numpy as np
import pytest

from conftest import golden
from helpers import rel
from hobakv1_amd import lib, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("key", ["snh", "snh_stiff", "arap"])
def test_unit_tets_match_reference(gpu_ctx, key):
    g = golden("unit_tets.npz")
    mat, mu, lam = g[f"{key}_params"]
    sc = scenes.Scene(name=key, rest=g["rest"], tets=g["tets"], material=int(mat), mu=mu, lam=lam)
    print(f"unit tets {key}: {sc.nt} tets", flush=True)
    inst = lib.Instance(gpu_ctx, sc)
    rp, ci = inst.bsr_pattern()
    assert np.array_equal(rp, g["row_ptr"]) and np.array_equal(ci, g["col_idx"])
    R, K = inst.forces_hessian(g["u"])
    Rr, Kr = g[f"{key}_R"], g[f"{key}_K"]
    worst_f = worst_k = 0.0
    bad = []
    for t in range(sc.nt):
        f, fr = R[12 * t:12 * t + 12], Rr[12 * t:12 * t + 12]
        k, kr = K[rp[4 * t]:rp[4 * t + 4]], Kr[rp[4 * t]:rp[4 * t + 4]]
        ef = np.abs(f - fr).max() / max(np.abs(fr).max(), mu)
        ek = rel(k, kr)
        worst_f, worst_k = max(worst_f, ef), max(worst_k, ek)
        if ef > 1e-12 or ek > 1e-12:
            bad.append((t, float(ef), float(ek)))
    print(f"  worst per-tet force {worst_f:.2e}, stiffness {worst_k:.2e}; over 1e-12: {bad}",
          flush=True)
    assert not bad
    inst.close()
