"""The C-ABI library loads and exports every symbol include/hobak_c.h declares (no GPU needed,
no compute calls)."""
import ctypes
import os
import re

import pytest

from hobakv1_amd import lib, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "hobak_c.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hobak_[a-z_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("hobak_ctx_create", "hobak_inst_create", "hobak_step", "hobak_step_batch",
              "hobak_eval_forces_hessian", "hobak_get_bsr_pattern", "hobak_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    so = ctypes.CDLL(lib.LIB_PATH)
    for s in declared_symbols():
        assert hasattr(so, s), s


def test_python_binding_covers_the_header():
    assert sorted(lib.SIGNATURES) == declared_symbols()


def test_default_params_match_reference_defaults():
    """BACKWARD_EULER_POSITION::initialize (BACKWARD_EULER_POSITION.cpp:50-79): dt = 1/60,
    Rayleigh alpha = beta = 0.01; Eigen CG defaults (tol = eps, max = 2 DOFs) encoded as 0."""
    p = lib.default_params()
    assert p.dt == 1.0 / 60.0
    assert p.rayleigh_alpha == 0.01 and p.rayleigh_beta == 0.01
    assert p.pcg_tol == 0.0 and p.pcg_max_iters == 0
    assert p.precond == scenes.PRECOND_JACOBI
    assert p.newton_iters == 1 and p.newton_tol == 1e-7


def test_struct_layout_matches_header():
    # offsets are checked against the C compiler's layout by the C++ facade test build; here the
    # sizes are pinned so a field added on one side only is caught
    assert ctypes.sizeof(lib.hobak_material) == 24
    assert ctypes.sizeof(lib.hobak_step_stats) == 4 + 4 + 32 + 32 + 4 + 4 + 64 + 64 + 8


def test_cpp_facade_compiles_against_the_header(tmp_path):
    """The C++ façade (hobakv1_amd/host/hobak.hpp) and its driver compile as host C++17."""
    import shutil
    import subprocess

    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    src = os.path.join(ROOT, "hobakv1_amd", "host", "facade_driver.cpp")
    subprocess.run([cxx, "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I",
                    os.path.join(ROOT, "include"), src], check=True)


def test_offset_range_check():
    """hobak_check_offsets (host-only) accepts the largest workloads and refuses a mesh whose
    matrix planes or padded vectors would overflow the PCG kernels' 32-bit byte offsets; the
    boundaries follow from 2^32 bytes per plane (FP64: 536,870,912 slots) and per vector
    (FP64: 178,956,970 vertices, rounded down to whole 64-row slices)."""
    so = lib.load()
    ok, bad = 0, -1
    # C2 and a 300-instance-sized mesh
    assert so.hobak_check_offsets(68921, 1179648, 0) == ok
    assert so.hobak_check_offsets(4_000_000, 70_000_000, 1) == ok
    # value planes: nslots * 8 < 2^32 (FP64), nslots * 4 < 2^32 (FP32)
    assert so.hobak_check_offsets(1000, (1 << 29) - 1, 0) == ok
    assert so.hobak_check_offsets(1000, 1 << 29, 0) == bad
    assert so.hobak_check_offsets(1000, (1 << 30) - 1, 1) == ok
    assert so.hobak_check_offsets(1000, 1 << 30, 1) == bad
    # padded vectors: 3 * ceil64(nv) * s < 2^32
    nv64 = ((1 << 32) // 24) // 64 * 64
    assert so.hobak_check_offsets(nv64 - 64, 1000, 0) == ok
    assert so.hobak_check_offsets(nv64 + 1, 1000, 0) == bad
    assert so.hobak_check_offsets(0, 1000, 0) == bad
