"""ctypes binding of libhobak_hip.so (include/hobak_c.h) and a Python mirror of the reference's
surface for this path:

    reference (relative to the reference root)                 here
    TET_MESH(restVertices, tets)  src/Geometry/TET_MESH.h:37  TetMesh(rest, tets)
    VOLUME::SNH / VOLUME::ARAP    src/Hyperelastic/Volume/    SNH(mu, lam) / ARAP(mu, lam)
    BACKWARD_EULER_POSITION       src/Timestepper/BACKWARD_EULER_POSITION.h:36
                                                               BackwardEulerPosition(mesh, material)
    BACKWARD_EULER_VELOCITY       src/Timestepper/BACKWARD_EULER_VELOCITY.h:35
                                                               BackwardEulerVelocity(mesh, material)
      .solve(verbose) / position() / velocity() / addGravity / attachKinematicConstraints /
      setDt / setRayleigh / scaleMass                          same names

The GPU library is the only compute path: when libhobak_hip.so is missing or no gfx950 device is
present, every call raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from . import scenes as _scenes

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhobak_hip.so")

HOBAK_OK = 0
ERRORS = {
    -1: "HOBAK_E_INVALID", -2: "HOBAK_E_HIP", -3: "HOBAK_E_NOMEM", -4: "HOBAK_E_NOT_CONVERGED",
    -5: "HOBAK_E_NONFINITE", -6: "HOBAK_E_NODEVICE",
}
MAX_NEWTON = 8

PRECISION_FP64 = 0
PRECISION_FP32_REFINE = 1

DAMPING_RAYLEIGH = 0
DAMPING_GREEN = 1


class hobak_material(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("mu", ctypes.c_double), ("lam", ctypes.c_double)]


class hobak_step_params(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double), ("rayleigh_alpha", ctypes.c_double),
        ("rayleigh_beta", ctypes.c_double), ("mass_scale", ctypes.c_double),
        ("gravity", ctypes.c_double * 3), ("pcg_tol", ctypes.c_double),
        ("pcg_max_iters", ctypes.c_int32), ("precond", ctypes.c_int32),
        ("newton_iters", ctypes.c_int32), ("newton_tol", ctypes.c_double),
        ("precision", ctypes.c_int32), ("refine_inner_tol", ctypes.c_double),
        ("ground_on", ctypes.c_int32), ("ground_y", ctypes.c_double), ("ground_k", ctypes.c_double),
        ("integrator", ctypes.c_int32),
    ]


class hobak_step_stats(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32), ("newton_iters", ctypes.c_int32),
        ("pcg_iters", ctypes.c_int32 * MAX_NEWTON), ("refine_iters", ctypes.c_int32 * MAX_NEWTON),
        ("total_pcg_iters", ctypes.c_int32), ("pcg_error", ctypes.c_double * MAX_NEWTON),
        ("newton_residual", ctypes.c_double * MAX_NEWTON), ("seconds", ctypes.c_double),
    ]

    def as_dict(self):
        n = self.newton_iters
        return {
            "status": self.status, "newton_iters": n,
            "pcg_iters": list(self.pcg_iters[:n]), "refine_iters": list(self.refine_iters[:n]),
            "total_pcg_iters": self.total_pcg_iters, "pcg_error": list(self.pcg_error[:n]),
            "newton_residual": list(self.newton_residual[:n]), "seconds": self.seconds,
        }


class hobak_completion(ctypes.Structure):
    _fields_ = [
        ("inst", ctypes.c_void_p), ("tag", ctypes.c_uint64), ("stats", hobak_step_stats),
        ("submit_s", ctypes.c_double), ("start_s", ctypes.c_double), ("end_s", ctypes.c_double),
    ]


_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)

# name -> (restype, argtypes); every symbol include/hobak_c.h declares
SIGNATURES = {
    "hobak_default_params": (None, [ctypes.POINTER(hobak_step_params)]),
    "hobak_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "hobak_ctx_destroy": (None, [_vp]),
    "hobak_last_error": (ctypes.c_char_p, [_vp]),
    "hobak_ctx_bytes_in_use": (ctypes.c_size_t, [_vp]),
    "hobak_inst_create": (ctypes.c_int, [_vp, _dp, ctypes.c_int32, _ip, ctypes.c_int32,
                                         ctypes.POINTER(hobak_material),
                                         ctypes.POINTER(hobak_step_params), ctypes.POINTER(_vp)]),
    "hobak_inst_bytes_estimate": (ctypes.c_size_t, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "hobak_check_offsets": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32]),
    "hobak_inst_destroy": (None, [_vp]),
    "hobak_inst_set_pinned": (ctypes.c_int, [_vp, _ip, ctypes.c_int32, _dp]),
    "hobak_inst_add_collision_cube": (ctypes.c_int, [_vp, _dp, ctypes.c_double]),
    "hobak_inst_plane_constraints": (ctypes.c_int, [_vp, _ip, _ip, ctypes.c_int32]),
    "hobak_inst_set_damping": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_double]),
    "hobak_inst_compute_fdots": (ctypes.c_int, [_vp, _dp]),
    "hobak_eval_damping": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "hobak_inst_set_state": (ctypes.c_int, [_vp, _dp, _dp]),
    "hobak_inst_get_state": (ctypes.c_int, [_vp, _dp, _dp]),
    "hobak_step": (ctypes.c_int, [_vp, ctypes.POINTER(hobak_step_stats)]),
    "hobak_step_batch": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_int32,
                                        ctypes.POINTER(hobak_step_stats)]),
    "hobak_get_bsr_pattern": (ctypes.c_int, [_vp, _ip, _ip, _ip]),
    "hobak_eval_forces_hessian": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "hobak_eval_forces_noalloc": (ctypes.c_int, [_vp, ctypes.c_void_p, ctypes.c_void_p]),
    "hobak_get_last_system": (ctypes.c_int, [_vp, _dp, _dp]),
    "hobak_last_step_timing": (ctypes.c_int, [_vp, _dp]),
    "hobak_set_kernel_sampling": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "hobak_kernel_stats": (ctypes.c_int, [_vp, _dp, ctypes.POINTER(ctypes.c_int64), _dp]),
    "hobak_copy_probe": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.c_int32, _dp]),
    "hobak_ctx_set_tuning": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32]),
    "hobak_sched_create": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "hobak_sched_destroy": (None, [_vp]),
    "hobak_sched_submit": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64]),
    "hobak_sched_detach": (ctypes.c_int, [_vp, _vp]),
    "hobak_sched_tick": (ctypes.c_int, [_vp, ctypes.POINTER(hobak_completion), ctypes.c_int32,
                                        _ip, _ip]),
    "hobak_sched_clock": (ctypes.c_double, [_vp]),
    "hobak_sched_stats": (ctypes.c_int, [_vp, _dp]),
    "hobak_sched_set_kernel_sampling": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "hobak_sched_kernel_stats": (ctypes.c_int, [_vp, _dp, ctypes.POINTER(ctypes.c_int64), _dp]),
}

TUNE_SLICE_ROWS = 0
TUNE_POLL_CHUNK = 2
TUNE_MAX_CHUNKS = 3
TUNE_GRAPHS = 4
TUNE_GRID_CAP = 5
TUNE_UPDATE_WGS = 6
TUNE_SPMV_PAD = 7
TUNE_SPMV_STREAM_MB = 8
TUNE_ROW_SORT = 9
TUNE_COMPACT_PCT = 10
TUNE_UPDATE_GROUPS = 11
TUNE_P_FORM = 12
TUNE_COARSE = 13
TUNE_AGG = 14

_lib = None


class HobakError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def load(path: str = LIB_PATH):
    """Load libhobak_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with `python -m hobakv1_amd.build`")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def inst_bytes_estimate(nv: int, nt: int, precision: int) -> int:
    """Device bytes an instance of this size takes (hobak_inst_bytes_estimate)."""
    return int(load().hobak_inst_bytes_estimate(int(nv), int(nt), int(precision)))


def default_params() -> hobak_step_params:
    p = hobak_step_params()
    load().hobak_default_params(ctypes.byref(p))
    return p


class Context:
    """One device + stream + HBM budget (0 = unlimited).

    apply_table: apply the autotuned launch shapes of tuning.json (hobakv1_amd/autotune.py) for
    `precision` and `workload` to the instances this context creates.  workload "single" is one
    instance stepped at a time (tuned on the 40^3 scene), "batch" many instances stepped together
    (tuned on the 256-instance batch), "serving" the continuously batched request stream (tuned on
    the serving stream); a missing serving entry falls back to the batch one, a missing batch
    entry to the single one.
    """

    def __init__(self, device: int = 0, budget_bytes: int = 0, apply_table: bool = True,
                 precision: str = "fp64", workload: str = "single"):
        self.lib = load()
        h = _vp()
        rc = self.lib.hobak_ctx_create(device, budget_bytes, ctypes.byref(h))
        if rc != HOBAK_OK:
            raise HobakError(rc, "hobak_ctx_create failed (no gfx950 device?)")
        self.h = h
        self.refine_inner_tol = None
        if apply_table:
            self.apply_tuning_table(precision, workload)

    def apply_tuning_table(self, precision="fp64", workload="single"):
        path = os.path.join(HERE, "tuning.json")
        if not os.path.exists(path):
            return
        import json

        with open(path) as f:
            table = json.load(f)
        entry = None
        if workload == "serving":
            entry = table.get(f"{precision}_serving") or table.get(f"{precision}_batch")
        elif workload == "batch":
            entry = table.get(f"{precision}_batch")
        entry = entry or table.get(precision)
        if entry:
            self.set_tuning(slice_rows=entry.get("slice_rows"), poll_chunk=entry.get("poll_chunk"),
                            max_chunks=entry.get("max_chunks"), graphs=entry.get("graphs"),
                            grid_cap=entry.get("grid_cap"), update_wgs=entry.get("update_wgs"),
                            spmv_pad=entry.get("spmv_pad"), spmv_stream_mb=entry.get("spmv_stream_mb"),
                            row_sort=entry.get("row_sort"), compact_pct=entry.get("compact_pct"),
                            update_groups=entry.get("update_groups"), coarse=entry.get("coarse"),
                            agg=entry.get("agg"))
            self.refine_inner_tol = entry.get("refine_inner_tol")

    def error(self) -> str:
        return self.lib.hobak_last_error(self.h).decode()

    def check(self, rc, what=""):
        if rc != HOBAK_OK:
            raise HobakError(rc, f"{what}: {self.error()}")
        return rc

    def set_tuning(self, slice_rows=None, poll_chunk=None, max_chunks=None, graphs=None,
                   grid_cap=None, update_wgs=None, spmv_pad=None, spmv_stream_mb=None,
                   row_sort=None, compact_pct=None, update_groups=None, p_form=None, coarse=None,
                   agg=None):
        """Kernel and engine tuning (see hobak_ctx_set_tuning)."""
        for key, val in ((TUNE_SLICE_ROWS, slice_rows), (TUNE_POLL_CHUNK, poll_chunk),
                         (TUNE_MAX_CHUNKS, max_chunks), (TUNE_GRAPHS, graphs),
                         (TUNE_GRID_CAP, grid_cap), (TUNE_UPDATE_WGS, update_wgs),
                         (TUNE_SPMV_PAD, spmv_pad), (TUNE_SPMV_STREAM_MB, spmv_stream_mb),
                         (TUNE_ROW_SORT, row_sort), (TUNE_COMPACT_PCT, compact_pct),
                         (TUNE_UPDATE_GROUPS, update_groups), (TUNE_P_FORM, p_form),
                         (TUNE_COARSE, coarse), (TUNE_AGG, agg)):
            if val is not None:
                self.check(self.lib.hobak_ctx_set_tuning(self.h, key, int(val)), "set_tuning")

    def copy_probe(self, nbytes, reps=20):
        ms = ctypes.c_double(0)
        self.check(self.lib.hobak_copy_probe(self.h, int(nbytes), int(reps), ctypes.byref(ms)),
                   "copy_probe")
        return ms.value

    def bytes_in_use(self) -> int:
        return self.lib.hobak_ctx_bytes_in_use(self.h)

    def close(self):
        if self.h:
            self.lib.hobak_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def make_params(scene, precond=_scenes.PRECOND_BLOCK_JACOBI, precision=PRECISION_FP64,
                newton_iters=None, newton_tol=1e-7, pcg_tol=None, refine_inner_tol=1e-4):
    p = default_params()
    p.dt = scene.dt
    p.rayleigh_alpha = scene.rayleigh_alpha
    p.rayleigh_beta = scene.rayleigh_beta
    p.mass_scale = scene.mass_scale
    for c in range(3):
        p.gravity[c] = scene.gravity[c]
    p.pcg_tol = scene.pcg_tol if pcg_tol is None else pcg_tol
    p.pcg_max_iters = scene.pcg_max_iters
    p.precond = precond
    p.newton_iters = scene.newton_iters if newton_iters is None else newton_iters
    p.newton_tol = newton_tol
    p.precision = precision
    p.refine_inner_tol = refine_inner_tol
    if scene.ground_y is not None:
        p.ground_on = 1
        p.ground_y = scene.ground_y
        p.ground_k = scene.ground_k
    p.integrator = getattr(scene, "integrator", _scenes.INTEGRATOR_POSITION)
    return p


class Instance:
    """One HBM-resident simulation instance of a Scene."""

    def __init__(self, ctx: Context, scene, precond=_scenes.PRECOND_BLOCK_JACOBI,
                 precision=PRECISION_FP64, newton_iters=None, newton_tol=1e-7, pcg_tol=None,
                 refine_inner_tol=None, params: Optional[hobak_step_params] = None):
        if refine_inner_tol is None:
            refine_inner_tol = getattr(ctx, "refine_inner_tol", None) or 1e-4
        self.ctx = ctx
        self.lib = ctx.lib
        self.scene = scene
        self.dofs = scene.dofs
        self.nv = scene.nv
        self.params = params or make_params(scene, precond, precision, newton_iters, newton_tol,
                                            pcg_tol, refine_inner_tol)
        mat = hobak_material(scene.material, scene.mu, scene.lam)
        rest = np.ascontiguousarray(scene.rest, dtype=np.float64)
        tets = np.ascontiguousarray(scene.tets, dtype=np.int32)
        h = _vp()
        ctx.check(self.lib.hobak_inst_create(ctx.h, _d(rest), scene.nv, _i(tets), scene.nt,
                                             ctypes.byref(mat), ctypes.byref(self.params),
                                             ctypes.byref(h)), "hobak_inst_create")
        self.h = h
        pins = np.ascontiguousarray(scene.pinned(), dtype=np.int32)
        if pins.size:
            tg = np.ascontiguousarray(scene.pinned_targets(), dtype=np.float64)
            ctx.check(self.lib.hobak_inst_set_pinned(self.h, _i(pins), pins.size, _d(tg)),
                      "set_pinned")
        for cube in getattr(scene, "collision_cubes", ()):
            self.add_collision_cube(cube[:3], cube[3])
        if getattr(scene, "green_mu", None) is not None:
            self.set_damping(DAMPING_GREEN, scene.green_mu)
            if getattr(scene, "fdot_v", None) is not None:
                self.compute_fdots(scene.fdot_v)
        self.set_state(scene.initial_u(), scene.initial_v())

    def set_pinned(self, vids, targets=None):
        """Pin vertices `vids` (caller numbering) at displacements `targets` (n x 3, default 0)."""
        vids = np.ascontiguousarray(vids, dtype=np.int32)
        tg = None if targets is None else np.ascontiguousarray(targets, dtype=np.float64)
        self.ctx.check(self.lib.hobak_inst_set_pinned(self.h, _i(vids), vids.size,
                                                      _d(tg) if tg is not None else None),
                       "set_pinned")

    def set_damping(self, model, mu=0.0):
        """DAMPING_RAYLEIGH, or DAMPING_GREEN: GREEN_DAMPING(mu) and the energy-damped step."""
        self.ctx.check(self.lib.hobak_inst_set_damping(self.h, int(model), float(mu)), "set_damping")

    def compute_fdots(self, velocity=None):
        """TET_MESH::computeFdots: the held velocity gradients, from `velocity` or the current v."""
        vv = None if velocity is None else np.ascontiguousarray(velocity, dtype=np.float64)
        self.ctx.check(self.lib.hobak_inst_compute_fdots(self.h, _d(vv) if vv is not None else None),
                       "compute_fdots")

    def damping_forces_hessian(self, u, with_c=True):
        """computeInternalForces(hyperelastic, damping) and computeDampingHessian at u."""
        rp, ci = self.bsr_pattern()
        R = np.zeros(self.dofs)
        C = np.zeros(9 * ci.size) if with_c else None
        uu = np.ascontiguousarray(u, dtype=np.float64)
        self.ctx.check(self.lib.hobak_eval_damping(self.h, _d(uu), _d(R), _d(C) if with_c else None),
                       "eval_damping")
        return R, (C.reshape(-1, 3, 3) if with_c else None)

    def add_collision_cube(self, center, scale):
        """addKinematicCollisionObject(CUBE(center, scale)): plane constraints on contact."""
        c = np.ascontiguousarray(center, dtype=np.float64)
        self.ctx.check(self.lib.hobak_inst_add_collision_cube(self.h, _d(c), float(scale)),
                       "add_collision_cube")

    def plane_constraints(self):
        """Vertex ids under a plane constraint after the last step, ascending."""
        n = ctypes.c_int32(0)
        self.ctx.check(self.lib.hobak_inst_plane_constraints(self.h, ctypes.byref(n), None, 0))
        ids = np.zeros(max(n.value, 1), np.int32)
        self.ctx.check(self.lib.hobak_inst_plane_constraints(self.h, ctypes.byref(n), _i(ids),
                                                             ids.size))
        return ids[:n.value]

    def close(self):
        if self.h:
            self.lib.hobak_inst_destroy(self.h)
            self.h = None

    def set_state(self, u=None, v=None):
        uu = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        self.ctx.check(self.lib.hobak_inst_set_state(self.h, _d(uu) if uu is not None else None,
                                                     _d(vv) if vv is not None else None),
                       "set_state")

    def state(self):
        u = np.zeros(self.dofs)
        v = np.zeros(self.dofs)
        self.ctx.check(self.lib.hobak_inst_get_state(self.h, _d(u), _d(v)), "get_state")
        return u, v

    def step(self, allow_unconverged=False):
        st = hobak_step_stats()
        rc = self.lib.hobak_step(self.h, ctypes.byref(st))
        if rc != HOBAK_OK and not (allow_unconverged and rc == -4):
            raise HobakError(rc, self.ctx.error())
        return st.as_dict()

    def timing_ms(self):
        out = np.zeros(5)
        self.ctx.check(self.lib.hobak_last_step_timing(self.h, _d(out)), "timing")
        return dict(zip(["element", "assemble", "rhs", "pcg", "update"], out.tolist()))

    def set_kernel_sampling(self, every):
        self.ctx.check(self.lib.hobak_set_kernel_sampling(self.h, int(every)), "sampling")

    def kernel_stats(self):
        ms = ctypes.c_double(0)
        n = ctypes.c_int64(0)
        b = ctypes.c_double(0)
        self.ctx.check(self.lib.hobak_kernel_stats(self.h, ctypes.byref(ms), ctypes.byref(n),
                                                   ctypes.byref(b)), "kernel_stats")
        return {"spmv_avg_ms": ms.value, "spmv_count": n.value, "spmv_bytes": b.value}

    def bsr_pattern(self):
        n = ctypes.c_int32(0)
        self.ctx.check(self.lib.hobak_get_bsr_pattern(self.h, ctypes.byref(n), None, None))
        rp = np.zeros(self.nv + 1, np.int32)
        ci = np.zeros(n.value, np.int32)
        self.ctx.check(self.lib.hobak_get_bsr_pattern(self.h, None, _i(rp), _i(ci)))
        return rp, ci

    def forces_hessian(self, u, with_k=True):
        rp, ci = self.bsr_pattern()
        R = np.zeros(self.dofs)
        K = np.zeros(9 * ci.size) if with_k else None
        uu = np.ascontiguousarray(u, dtype=np.float64)
        self.ctx.check(self.lib.hobak_eval_forces_hessian(self.h, _d(uu), _d(R),
                                                          _d(K) if with_k else None),
                       "eval_forces_hessian")
        return R, (K.reshape(-1, 3, 3) if with_k else None)

    def last_system(self):
        rp, ci = self.bsr_pattern()
        rhs = np.zeros(self.dofs)
        lhs = np.zeros(9 * ci.size)
        self.ctx.check(self.lib.hobak_get_last_system(self.h, _d(rhs), _d(lhs)), "last_system")
        return rhs, lhs.reshape(-1, 3, 3)


def step_batch(ctx: Context, insts, allow_unconverged=False):
    n = len(insts)
    arr = (_vp * n)(*[i.h for i in insts])
    stats = (hobak_step_stats * n)()
    rc = ctx.lib.hobak_step_batch(ctx.h, arr, n, stats)
    if rc != HOBAK_OK and not (allow_unconverged and rc == -4):
        raise HobakError(rc, ctx.error())
    return [s.as_dict() for s in stats]


class Scheduler:
    """Continuous-batching scheduler (hobak_sched_*): submit step requests for instances, call
    tick() repeatedly; each tick returns the completed requests."""

    def __init__(self, ctx: Context, max_out: int = 4096):
        self.ctx = ctx
        self.lib = ctx.lib
        h = _vp()
        ctx.check(self.lib.hobak_sched_create(ctx.h, ctypes.byref(h)), "sched_create")
        self.h = h
        self._buf = (hobak_completion * max_out)()
        self._by_handle = {}

    def submit(self, inst: "Instance", tag: int):
        self._by_handle[inst.h.value] = inst
        self.ctx.check(self.lib.hobak_sched_submit(self.h, inst.h, int(tag)), "sched_submit")

    def detach(self, inst: "Instance"):
        self.ctx.check(self.lib.hobak_sched_detach(self.h, inst.h), "sched_detach")
        self._by_handle.pop(inst.h.value, None)

    def clock(self) -> float:
        return self.lib.hobak_sched_clock(self.h)

    def set_kernel_sampling(self, every):
        self.ctx.check(self.lib.hobak_sched_set_kernel_sampling(self.h, int(every)), "sched_sampling")

    def kernel_stats(self):
        ms = ctypes.c_double(0)
        n = ctypes.c_int64(0)
        b = ctypes.c_double(0)
        self.ctx.check(self.lib.hobak_sched_kernel_stats(self.h, ctypes.byref(ms), ctypes.byref(n),
                                                         ctypes.byref(b)), "sched_kernel_stats")
        return {"spmv_avg_ms": ms.value, "spmv_count": n.value, "spmv_bytes": b.value}

    def stats(self):
        out = np.zeros(11)
        self.ctx.check(self.lib.hobak_sched_stats(self.h, _d(out)), "sched_stats")
        keys = ["ticks", "pcg_groups", "prologue_gpu_ms", "pcg_gpu_ms", "epilogue_gpu_ms",
                "tick_host_ms", "solves", "spmv_launches", "spmv_bytes", "pcg_groups_graph",
                "pcg_groups_direct"]
        d = dict(zip(keys, out.tolist()))
        d["spmv_bytes_per_launch"] = d["spmv_bytes"] / d["spmv_launches"] if d["spmv_launches"] else 0.0
        return d

    def tick(self):
        """Runs one tick; returns (completions, requests still in flight)."""
        n = ctypes.c_int32(0)
        inflight = ctypes.c_int32(0)
        self.ctx.check(self.lib.hobak_sched_tick(self.h, self._buf, len(self._buf), ctypes.byref(n),
                                                 ctypes.byref(inflight)), "sched_tick")
        out = []
        for k in range(n.value):
            c = self._buf[k]
            out.append({"inst": self._by_handle.get(c.inst), "tag": c.tag, "stats": c.stats.as_dict(),
                        "submit_s": c.submit_s, "start_s": c.start_s, "end_s": c.end_s})
        return out, inflight.value

    def close(self):
        if self.h:
            self.lib.hobak_sched_destroy(self.h)
            self.h = None


# ---------------------------------------------------------------------------------------------
# Reference-shaped facade
# ---------------------------------------------------------------------------------------------
class TetMesh:
    """TET_MESH(restVertices, tets) -- src/Geometry/TET_MESH.h:37-38."""

    def __init__(self, rest_vertices, tets):
        self._rest = np.ascontiguousarray(rest_vertices, dtype=np.float64).reshape(-1, 3)
        self._tets = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
        self._fdot_velocity = None  # the velocity of the last computeFdots call
        self._fdot_serial = 0

    def computeFdots(self, velocity):
        """TET_MESH::computeFdots (TET_MESH.cpp:660-684): the velocity gradients an energy-damped
        step reads.  The device instance picks them up at the next solve()."""
        self._fdot_velocity = np.array(velocity, dtype=np.float64).reshape(-1)
        self._fdot_serial += 1

    def restVertices(self):
        return self._rest

    def tets(self):
        return self._tets

    def DOFs(self):
        return 3 * self._rest.shape[0]

    @staticmethod
    def readTobjFile(path):
        """TET_MESH::readTobjFile (TET_MESH.cpp:1173-1220): (vertices, tets)."""
        from .formats import read_tobj

        return read_tobj(path)

    @staticmethod
    def writeTobjFile(path, mesh):
        """TET_MESH::writeTobjFile (TET_MESH.cpp:1225-1259) of the rest vertices."""
        from .formats import write_tobj

        write_tobj(path, mesh.restVertices(), mesh.tets())


class GREEN_DAMPING:
    """VOLUME::GREEN_DAMPING(mu) (src/Damping/Volume/GREEN_DAMPING.cpp:25-28)."""

    def __init__(self, mu):
        self.mu = float(mu)

    def name(self):
        return "GREEN_DAMPING"


class HYPERELASTIC:
    kind = -1

    def __init__(self, mu, lam):
        self.mu = float(mu)
        self.lam = float(lam)

    @staticmethod
    def computeMu(E, nu):
        return _scenes.lame(E, nu)[0]

    @staticmethod
    def computeLambda(E, nu):
        return _scenes.lame(E, nu)[1]


class SNH(HYPERELASTIC):
    kind = _scenes.MAT_SNH

    def name(self):
        return "SNH"


class ARAP(HYPERELASTIC):
    kind = _scenes.MAT_ARAP

    def name(self):
        return "ARAP"


class BackwardEulerPosition:
    """BACKWARD_EULER_POSITION(tetMesh, hyperelastic) with the TIMESTEPPER setters.

    The device instance is created at the first solve(), after all setters have run, the way the
    reference reads its members inside solveRayleighDamped.
    """

    def __init__(self, mesh: TetMesh, material: HYPERELASTIC, ctx: Optional[Context] = None,
                 precond=_scenes.PRECOND_BLOCK_JACOBI, precision=PRECISION_FP64):
        self.mesh = mesh
        self.material = material
        self.ctx = ctx or Context()
        self._owns_ctx = ctx is None
        n = mesh.DOFs()
        self._position = np.zeros(n)
        self._velocity = np.zeros(n)
        self._gravity = np.zeros(3)
        self._dt = 1.0 / 60.0
        self._alpha = 0.01
        self._beta = 0.01
        self._mass_scale = 1.0
        self._pins = np.zeros(0, np.int32)
        self._targets = np.zeros((0, 3))
        self._collision = []
        self._damping = None
        self._fdot_applied = 0
        self._precond = precond
        self._precision = precision
        self._tol = 0.0
        self._inst = None
        self.last_stats = None

    # TIMESTEPPER.h:44-58
    def position(self):
        return self._position

    def velocity(self):
        return self._velocity

    def setDt(self, dt):
        self._dt = float(dt)

    def setRayleigh(self, alpha, beta):
        self._alpha, self._beta = float(alpha), float(beta)

    def scaleMass(self, s):
        self._mass_scale *= float(s)

    def addGravity(self, g):
        self._gravity = np.asarray(g, dtype=np.float64)

    def setTolerance(self, tol):
        self._tol = float(tol)

    def attachKinematicConstraints(self, cube: _scenes.Cube):
        X = self.mesh.restVertices()
        ids = np.nonzero(cube.inside(X))[0].astype(np.int32)
        self._pins = np.union1d(self._pins, ids).astype(np.int32)
        self._targets = cube.to_world(cube.to_local(X[self._pins])) - X[self._pins]

    def addKinematicCollisionObject(self, cube: _scenes.Cube):
        """TIMESTEPPER::addKinematicCollisionObject (TIMESTEPPER.cpp:433-441) for a CUBE."""
        if any(c is cube for c in self._collision):
            return
        self._collision.append(cube)
        if self._inst is not None:
            self._inst.add_collision_cube(cube.center, cube.scale)

    def setDamping(self, damping: Optional[GREEN_DAMPING]):
        """The integrator's damping object (TIMESTEPPER.h:155): a GREEN_DAMPING makes solve() take
        solveEnergyDamped (BACKWARD_EULER_POSITION.cpp:125-131); None restores Rayleigh."""
        self._damping = damping
        if self._inst is not None:
            if damping is None:
                self._inst.set_damping(DAMPING_RAYLEIGH)
            else:
                self._inst.set_damping(DAMPING_GREEN, damping.mu)

    def planeConstraints(self):
        """Vertex ids of the live plane constraints (TIMESTEPPER.h:46), ascending."""
        return np.zeros(0, np.int32) if self._inst is None else self._inst.plane_constraints()

    _integrator = _scenes.INTEGRATOR_POSITION

    def _ensure(self):
        if self._inst is not None:
            return
        sc = _scenes.Scene(name="facade", rest=self.mesh.restVertices(), tets=self.mesh.tets(),
                           integrator=self._integrator,
                           material=self.material.kind, mu=self.material.mu,
                           lam=self.material.lam, dt=self._dt, rayleigh_alpha=self._alpha,
                           rayleigh_beta=self._beta, mass_scale=self._mass_scale,
                           gravity=tuple(self._gravity), pcg_tol=self._tol,
                           collision_cubes=tuple((*c.center, c.scale) for c in self._collision),
                           green_mu=None if self._damping is None else self._damping.mu)
        self._inst = Instance(self.ctx, sc, precond=self._precond, precision=self._precision)
        if self._pins.size:
            self.ctx.check(self.ctx.lib.hobak_inst_set_pinned(
                self._inst.h, _i(np.ascontiguousarray(self._pins)), self._pins.size,
                _d(np.ascontiguousarray(self._targets, dtype=np.float64))))

    def solve(self, verbose=False):
        self._ensure()
        self._inst.set_state(self._position, self._velocity)
        if self._damping is not None and self.mesh._fdot_serial != self._fdot_applied:
            self._inst.compute_fdots(self.mesh._fdot_velocity)
            self._fdot_applied = self.mesh._fdot_serial
        self.last_stats = self._inst.step(allow_unconverged=True)
        if verbose:
            print(f"  PCG iters: {self.last_stats['pcg_iters'][0]:3d} "
                  f"err: {self.last_stats['pcg_error'][0]:6.4e}")
        u, v = self._inst.state()
        self._position[:] = u
        self._velocity[:] = v
        return True


class BackwardEulerVelocity(BackwardEulerPosition):
    """BACKWARD_EULER_VELOCITY(tetMesh, hyperelastic) (src/Timestepper/BACKWARD_EULER_VELOCITY.h:35,
    .cpp:31-46): the same setters and solve(), solving for the velocity change
    (BACKWARD_EULER_VELOCITY.cpp:107-235)."""

    _integrator = _scenes.INTEGRATOR_VELOCITY
