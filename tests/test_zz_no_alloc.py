"""hobak_eval_forces_noalloc, the entry point that allocates nothing, against
hobak_eval_forces_hessian (the same element forces through the library's own buffers).

The test allocates the device input and output itself (hipMalloc through ctypes, as a caller
that owns its memory would) and calls the entry point through the C ABI twice in a row, the
second time on a smaller instance, so the static scratch array is reused with stale contents
from the first call.  Both element-force paths run the same kernels in the same order, so the
bar is bitwise equality; the forces are also held to the FP64 path's 1e-12 against the oracle.
"""
import ctypes

import numpy as np
import pytest

import oracle
from helpers import rel
from hobakv1_amd import lib, scenes

pytestmark = pytest.mark.gpu

HIP_H2D, HIP_D2H = 1, 2


def _hip():
    lib.load()  # libamdhip64 is loaded with the library
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMalloc.restype = ctypes.c_int
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipFree.restype = ctypes.c_int
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemcpy.restype = ctypes.c_int
    h.hipSetDevice.argtypes = [ctypes.c_int]
    h.hipSetDevice.restype = ctypes.c_int
    return h


def _noalloc_forces(hip, inst, u):
    n = u.size
    du, df = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipSetDevice(0) == 0
    assert hip.hipMalloc(ctypes.byref(du), 8 * n) == 0
    assert hip.hipMalloc(ctypes.byref(df), 8 * n) == 0
    try:
        uu = np.ascontiguousarray(u, dtype=np.float64)
        assert hip.hipMemcpy(du, uu.ctypes.data_as(ctypes.c_void_p), 8 * n, HIP_H2D) == 0
        rc = inst.lib.hobak_eval_forces_noalloc(inst.h, du, df)
        out = np.zeros(n)
        if rc == 0:
            assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), df, 8 * n, HIP_D2H) == 0
        return rc, out
    finally:
        hip.hipFree(du)
        hip.hipFree(df)


def test_noalloc_forces_equal_eval_forces_twice():
    hip = _hip()
    ctx = lib.Context(0)
    for k, sc in enumerate((scenes.config2(), scenes.lattice_scene((6, 5, 4), seed=8, material=scenes.MAT_ARAP,
                                                                    pin=False))):
        inst = lib.Instance(ctx, sc)
        u = 0.02 * np.random.default_rng(30 + k).uniform(-1, 1, sc.dofs)
        R_ref, _ = inst.forces_hessian(u, with_k=False)
        rc, R = _noalloc_forces(hip, inst, u)
        print(f"  call {k + 1}: nt {sc.nt} rc {rc} max |diff| {np.abs(R - R_ref).max():.3e}", flush=True)
        assert rc == 0, ctx.error()
        assert np.array_equal(R, R_ref), k
        o = oracle.Oracle(sc)
        Ro, _ = o.forces_hessian(u, with_k=False)
        assert rel(R, Ro) <= 1e-12
        o.close()
        inst.close()
    ctx.close()


def test_noalloc_refuses_beyond_its_scratch():
    """48^3 x 6 = 663,552 tets in FP64 need 63.7 MB of force staging, more than the 48 MiB array."""
    hip = _hip()
    ctx = lib.Context(0)
    sc = scenes.lattice_scene((48, 48, 48), seed=2, pin=False)
    inst = lib.Instance(ctx, sc)
    rc, _ = _noalloc_forces(hip, inst, np.zeros(sc.dofs))
    assert rc == -1
    assert "at most 524288 tets" in ctx.error()
    inst.close()
    ctx.close()
