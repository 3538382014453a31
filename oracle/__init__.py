"""TEST INFRASTRUCTURE ONLY.

ctypes drivers for the two CPU checkers of the backward-Euler Newton step:

* ``Oracle``  -- liboracle.so, this repository's FP64 CPU restatement of the reference path
  (oracle/hobak_oracle.cpp).  It is the checker on machines without the reference tree.
* ``RefSim``  -- _ref/libhobak_ref.so, the literal reference compiled from its own sources
  (oracle/Makefile, oracle/ref_harness.cpp).

Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py may import this
package.  The product library (hobakv1_amd) never does.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libhobak_ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


_ref_lib = None


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def load_ref():
    global _ref_lib
    if _ref_lib is None:
        lib = ctypes.CDLL(REF_SO)
        lib.ref_create.restype = ctypes.c_void_p
        lib.ref_create.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp,
                                   ctypes.c_double, ctypes.c_int]
        lib.ref_create_bdf1.restype = ctypes.c_void_p
        lib.ref_create_bdf1.argtypes = lib.ref_create.argtypes + [ctypes.c_int, ctypes.c_double]
        lib.ref_create_velocity.restype = ctypes.c_void_p
        lib.ref_create_velocity.argtypes = lib.ref_create.argtypes
        lib.ref_destroy.argtypes = [ctypes.c_void_p]
        lib.ref_num_pinned.argtypes = [ctypes.c_void_p]
        lib.ref_pinned_nodes.argtypes = [ctypes.c_void_p, _ip]
        lib.ref_set_state.argtypes = [ctypes.c_void_p, _dp, _dp]
        lib.ref_get_state.argtypes = [ctypes.c_void_p, _dp, _dp]
        lib.ref_step.argtypes = [ctypes.c_void_p, _dp, _dp]
        lib.ref_step.restype = ctypes.c_int
        lib.ref_last_rhs_b.argtypes = [ctypes.c_void_p, _dp]
        lib.ref_add_collision_cube.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double]
        lib.ref_read_tobj.argtypes = [ctypes.c_char_p, _ip, _dp, _ip]
        lib.ref_read_tobj.restype = ctypes.c_int
        lib.ref_write_tobj.argtypes = [ctypes.c_char_p, _dp, ctypes.c_int, _ip, ctypes.c_int]
        lib.ref_write_tobj.restype = ctypes.c_int
        lib.ref_read_scene_json.argtypes = [ctypes.c_char_p, _ip, _dp, _dp, _dp, _ip]
        lib.ref_read_scene_json.restype = ctypes.c_int
        lib.ref_plane_constraints.argtypes = [ctypes.c_void_p, _ip]
        lib.ref_plane_constraints.restype = ctypes.c_int
        lib.ref_mass_diag.argtypes = [ctypes.c_void_p, _dp]
        lib.ref_mesh_nnz.argtypes = [ctypes.c_void_p]
        lib.ref_mesh_nnz.restype = ctypes.c_int
        lib.ref_mesh_forces_hessian.argtypes = [ctypes.c_void_p, _dp, _dp, _ip, _ip, _dp]
        lib.ref_svd_rv.argtypes = [_dp, _dp, _dp, _dp]
        lib.ref_psi.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
        lib.ref_psi.restype = ctypes.c_double
        for f in (lib.ref_pk1, lib.ref_hessian, lib.ref_clamped_hessian):
            f.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
        lib.ref_set_green_damping.argtypes = [ctypes.c_void_p, ctypes.c_double]
        lib.ref_set_green_damping.restype = ctypes.c_int
        lib.ref_compute_fdots.argtypes = [ctypes.c_void_p, _dp]
        lib.ref_mesh_damped_forces_hessian.argtypes = [ctypes.c_void_p, _dp, _dp, _ip, _ip, _dp]
        lib.ref_green_psi.argtypes = [ctypes.c_double, _dp, _dp]
        lib.ref_green_psi.restype = ctypes.c_double
        lib.ref_green_pk1.argtypes = [ctypes.c_double, _dp, _dp, _dp]
        lib.ref_green_hessian.argtypes = [ctypes.c_double, _dp, _dp, ctypes.c_int, _dp]
        _ref_lib = lib
    return _ref_lib


def _colmajor(F):
    return _c64(np.asarray(F, dtype=np.float64).T.ravel())


def _from_colmajor(a, n):
    return a.reshape(n, n).T.copy()


class RefElement:
    """Element-level calls into the literal reference (row-major numpy in and out)."""

    def __init__(self):
        self.lib = load_ref()

    def svd_rv(self, F):
        U = np.zeros(9)
        V = np.zeros(9)
        S = np.zeros(3)
        self.lib.ref_svd_rv(_d(_colmajor(F)), _d(U), _d(S), _d(V))
        return _from_colmajor(U, 3), S, _from_colmajor(V, 3)

    def psi(self, mat, mu, lam, F):
        return self.lib.ref_psi(mat, mu, lam, _d(_colmajor(F)))

    def pk1(self, mat, mu, lam, F):
        P = np.zeros(9)
        self.lib.ref_pk1(mat, mu, lam, _d(_colmajor(F)), _d(P))
        return _from_colmajor(P, 3)

    def hessian(self, mat, mu, lam, F):
        H = np.zeros(81)
        self.lib.ref_hessian(mat, mu, lam, _d(_colmajor(F)), _d(H))
        return _from_colmajor(H, 9)

    def clamped_hessian(self, mat, mu, lam, F):
        H = np.zeros(81)
        self.lib.ref_clamped_hessian(mat, mu, lam, _d(_colmajor(F)), _d(H))
        return _from_colmajor(H, 9)

    def green_psi(self, mu, F, Fdot):
        return self.lib.ref_green_psi(mu, _d(_colmajor(F)), _d(_colmajor(Fdot)))

    def green_pk1(self, mu, F, Fdot):
        P = np.zeros(9)
        self.lib.ref_green_pk1(mu, _d(_colmajor(F)), _d(_colmajor(Fdot)), _d(P))
        return _from_colmajor(P, 3)

    def green_hessian(self, mu, F, Fdot, clamped=True):
        H = np.zeros(81)
        self.lib.ref_green_hessian(mu, _d(_colmajor(F)), _d(_colmajor(Fdot)), int(clamped), _d(H))
        return _from_colmajor(H, 9)


def ref_read_tobj(path):
    """TET_MESH::readTobjFile of the literal reference: (vertices (nv, 3), tets (nt, 4))."""
    lib = load_ref()
    nt = ctypes.c_int(0)
    nv = lib.ref_read_tobj(str(path).encode(), ctypes.byref(nt), None, None)
    if nv < 0:
        raise OSError(f"reference reader failed on {path}")
    v = np.zeros(3 * nv)
    t = np.zeros(4 * nt.value, np.int32)
    lib.ref_read_tobj(str(path).encode(), ctypes.byref(nt), _d(v), _i(t))
    return v.reshape(-1, 3), t.reshape(-1, 4)


def ref_write_tobj(path, vertices, tets):
    """TET_MESH::writeTobjFile of the literal reference (rest vertices)."""
    lib = load_ref()
    v = _c64(np.asarray(vertices).ravel())
    t = _c32(np.asarray(tets).ravel())
    if lib.ref_write_tobj(str(path).encode(), _d(v), v.size // 3, _i(t), t.size // 4) != 0:
        raise OSError(f"reference writer failed on {path}")


def ref_read_scene_json(path):
    """readSceneJSON of the literal reference (an uncompressed record): positions and velocities
    per frame, gravity and the number of kinematic shapes."""
    lib = load_ref()
    dofs = ctypes.c_int(0)
    nf = lib.ref_read_scene_json(str(path).encode(), ctypes.byref(dofs), None, None, None, None)
    if nf < 0:
        raise OSError(f"reference reader failed on {path}")
    pos = np.zeros(nf * dofs.value)
    vel = np.zeros(nf * dofs.value)
    g = np.zeros(3)
    ns = ctypes.c_int(0)
    lib.ref_read_scene_json(str(path).encode(), ctypes.byref(dofs), _d(pos), _d(vel), _d(g),
                            ctypes.byref(ns))
    return {"positions": pos.reshape(nf, -1), "velocities": vel.reshape(nf, -1), "gravity": g,
            "nshapes": ns.value}


class RefSim:
    """One reference scene: TET_MESH_FASTER + SNH/ARAP + BACKWARD_EULER_POSITION, or BDF_1
    (src/Timestepper/BDF_1.cpp:78-239) when newton_iters > 1, or BACKWARD_EULER_VELOCITY
    (src/Timestepper/BACKWARD_EULER_VELOCITY.cpp:107-235) for a scene with the velocity
    integrator."""

    def __init__(self, scene, pcg_tol=None, pcg_max_iters=None, newton_iters=None,
                 newton_tol=1e-7):
        self.lib = load_ref()
        self.scene = scene
        self._rest = _c64(scene.rest.ravel())
        self._tets = _c32(scene.tets.ravel())
        g = _c64(np.asarray(scene.gravity))
        pin = None if scene.pin_cube is None else _c64(np.asarray(scene.pin_cube))
        tol = scene.pcg_tol if pcg_tol is None else pcg_tol
        mi = scene.pcg_max_iters if pcg_max_iters is None else pcg_max_iters
        ni = scene.newton_iters if newton_iters is None else newton_iters
        args = (_d(self._rest), scene.nv, _i(self._tets), scene.nt, scene.material, scene.mu,
                scene.lam, scene.dt, scene.rayleigh_alpha, scene.rayleigh_beta, scene.mass_scale,
                _d(g), None if pin is None else _d(pin), tol, mi)
        if getattr(scene, "integrator", 0) == 1:
            self.h = self.lib.ref_create_velocity(*args)
        elif ni > 1:
            self.h = self.lib.ref_create_bdf1(*args, ni, newton_tol)
        else:
            self.h = self.lib.ref_create(*args)
        for cube in getattr(scene, "collision_cubes", ()):
            c = _c64(np.asarray(cube[:3]))
            self.lib.ref_add_collision_cube(self.h, _d(c), float(cube[3]))
        if getattr(scene, "green_mu", None) is not None:
            if self.lib.ref_set_green_damping(self.h, float(scene.green_mu)) != 0:
                raise ValueError("the reference's BDF_1 has no energy-damped solve")
            if getattr(scene, "fdot_v", None) is not None:
                self.compute_fdots(scene.fdot_v)
        self.dofs = scene.dofs
        self.set_state(scene.initial_u(), scene.initial_v())

    def close(self):
        if self.h:
            self.lib.ref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pinned(self):
        n = self.lib.ref_num_pinned(self.h)
        out = np.zeros(n, np.int32)
        self.lib.ref_pinned_nodes(self.h, _i(out))
        return out

    def set_state(self, u, v):
        self.lib.ref_set_state(self.h, _d(_c64(u)), _d(_c64(v)))

    def state(self):
        u = np.zeros(self.dofs)
        v = np.zeros(self.dofs)
        self.lib.ref_get_state(self.h, _d(u), _d(v))
        return u, v

    def step(self):
        err = ctypes.c_double(0)
        sec = ctypes.c_double(0)
        it = self.lib.ref_step(self.h, ctypes.byref(err), ctypes.byref(sec))
        return it, err.value, sec.value

    def last_b(self):
        b = np.zeros(self.dofs)
        self.lib.ref_last_rhs_b(self.h, _d(b))
        return b

    def plane_constraints(self):
        """Vertex ids of the live plane constraints, in the integrator's order."""
        n = self.lib.ref_plane_constraints(self.h, None)
        out = np.zeros(max(n, 1), np.int32)
        self.lib.ref_plane_constraints(self.h, _i(out))
        return out[:n]

    def mass_diag(self):
        m = np.zeros(self.dofs)
        self.lib.ref_mass_diag(self.h, _d(m))
        return m

    def compute_fdots(self, vel):
        """TET_MESH::computeFdots(vel): the velocity gradients the energy-damped step reads."""
        self.lib.ref_compute_fdots(self.h, _d(_c64(vel)))

    def damped_forces_hessian(self, u):
        """computeInternalForces(hyperelastic, damping) and the damping stiffness C (CSC) at u."""
        nnz = self.lib.ref_mesh_nnz(self.h)
        R = np.zeros(self.dofs)
        colptr = np.zeros(self.dofs + 1, np.int32)
        rowidx = np.zeros(nnz, np.int32)
        vals = np.zeros(nnz)
        self.lib.ref_mesh_damped_forces_hessian(self.h, _d(_c64(u)), _d(R), _i(colptr),
                                                _i(rowidx), _d(vals))
        return R, (colptr, rowidx, vals)

    def forces_hessian(self, u):
        """R and the clamped K (scipy-free CSC arrays) of the mesh at displacement u."""
        nnz = self.lib.ref_mesh_nnz(self.h)
        R = np.zeros(self.dofs)
        colptr = np.zeros(self.dofs + 1, np.int32)
        rowidx = np.zeros(nnz, np.int32)
        vals = np.zeros(nnz)
        self.lib.ref_mesh_forces_hessian(self.h, _d(_c64(u)), _d(R), _i(colptr), _i(rowidx),
                                         _d(vals))
        return R, (colptr, rowidx, vals)


_orc_lib = None


def oracle_available() -> bool:
    return os.path.exists(ORACLE_SO)


def load_oracle():
    global _orc_lib
    if _orc_lib is None:
        lib = ctypes.CDLL(ORACLE_SO)
        lib.orc_svd_rv.argtypes = [_dp, _dp, _dp, _dp]
        lib.orc_psi.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
        lib.orc_psi.restype = ctypes.c_double
        for f in (lib.orc_pk1, lib.orc_hessian, lib.orc_clamped_hessian):
            f.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
        lib.orc_clamped_hessian_usv.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                _dp, _dp, _dp, _dp]
        lib.orc_tet_force_hessian.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                              _dp, _dp, ctypes.c_double, _dp, _dp]
        lib.orc_create.restype = ctypes.c_void_p
        lib.orc_create.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp,
                                   _ip, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_double]
        lib.orc_destroy.argtypes = [ctypes.c_void_p]
        lib.orc_set_state.argtypes = [ctypes.c_void_p, _dp, _dp]
        lib.orc_get_state.argtypes = [ctypes.c_void_p, _dp, _dp]
        lib.orc_set_ground.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        lib.orc_add_collision_cube.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double]
        lib.orc_plane_constraints.argtypes = [ctypes.c_void_p, _ip]
        lib.orc_plane_constraints.restype = ctypes.c_int
        lib.orc_surface_vertices.argtypes = [ctypes.c_void_p, _ip]
        lib.orc_surface_vertices.restype = ctypes.c_int
        lib.orc_step.argtypes = [ctypes.c_void_p, _dp, _dp, _ip, _ip]
        lib.orc_step.restype = ctypes.c_int
        lib.orc_nnzb.argtypes = [ctypes.c_void_p]
        lib.orc_nnzb.restype = ctypes.c_int
        lib.orc_bsr_pattern.argtypes = [ctypes.c_void_p, _ip, _ip]
        lib.orc_forces_hessian.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        lib.orc_last_system.argtypes = [ctypes.c_void_p, _dp, _dp]
        lib.orc_rest_volumes.argtypes = [ctypes.c_void_p, _dp, _dp]
        lib.orc_set_threads.argtypes = [ctypes.c_int]
        lib.orc_green_psi.argtypes = [ctypes.c_double, _dp, _dp]
        lib.orc_green_psi.restype = ctypes.c_double
        lib.orc_green_pk1.argtypes = [ctypes.c_double, _dp, _dp, _dp]
        lib.orc_green_hessian.argtypes = [ctypes.c_double, _dp, ctypes.c_int, _dp]
        lib.orc_set_green_damping.argtypes = [ctypes.c_void_p, ctypes.c_double]
        lib.orc_set_green_damping.restype = ctypes.c_int
        lib.orc_compute_fdots.argtypes = [ctypes.c_void_p, _dp]
        lib.orc_damped_forces_hessian.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        lib.orc_set_velocity_form.argtypes = [ctypes.c_void_p]
        lib.orc_set_velocity_form.restype = ctypes.c_int
        _orc_lib = lib
    return _orc_lib


class OracleElement:
    """Element-level calls into the restatement (row-major numpy in and out)."""

    def __init__(self):
        self.lib = load_oracle()

    def svd_rv(self, F):
        U = np.zeros(9)
        V = np.zeros(9)
        S = np.zeros(3)
        self.lib.orc_svd_rv(_d(_colmajor(F)), _d(U), _d(S), _d(V))
        return _from_colmajor(U, 3), S, _from_colmajor(V, 3)

    def psi(self, mat, mu, lam, F):
        return self.lib.orc_psi(mat, mu, lam, _d(_colmajor(F)))

    def pk1(self, mat, mu, lam, F):
        P = np.zeros(9)
        self.lib.orc_pk1(mat, mu, lam, _d(_colmajor(F)), _d(P))
        return _from_colmajor(P, 3)

    def hessian(self, mat, mu, lam, F):
        H = np.zeros(81)
        self.lib.orc_hessian(mat, mu, lam, _d(_colmajor(F)), _d(H))
        return _from_colmajor(H, 9)

    def clamped_hessian(self, mat, mu, lam, F):
        H = np.zeros(81)
        self.lib.orc_clamped_hessian(mat, mu, lam, _d(_colmajor(F)), _d(H))
        return _from_colmajor(H, 9)

    def clamped_hessian_usv(self, mat, mu, lam, U, S, V):
        H = np.zeros(81)
        self.lib.orc_clamped_hessian_usv(mat, mu, lam, _d(_colmajor(U)), _d(_c64(S)),
                                         _d(_colmajor(V)), _d(H))
        return _from_colmajor(H, 9)

    def green_psi(self, mu, F, Fdot):
        return self.lib.orc_green_psi(mu, _d(_colmajor(F)), _d(_colmajor(Fdot)))

    def green_pk1(self, mu, F, Fdot):
        P = np.zeros(9)
        self.lib.orc_green_pk1(mu, _d(_colmajor(F)), _d(_colmajor(Fdot)), _d(P))
        return _from_colmajor(P, 3)

    def green_hessian(self, mu, F, clamped=True):
        H = np.zeros(81)
        self.lib.orc_green_hessian(mu, _d(_colmajor(F)), int(clamped), _d(H))
        return _from_colmajor(H, 9)

    def tet_force_hessian(self, mat, mu, lam, x4, DmInv, vol):
        f = np.zeros(12)
        H = np.zeros(144)
        self.lib.orc_tet_force_hessian(mat, mu, lam, _d(_c64(np.asarray(x4).ravel())),
                                       _d(_colmajor(DmInv)), vol, _d(f), _d(H))
        return f, _from_colmajor(H, 12)


class Oracle:
    """One scene on the CPU restatement.  precond 0 = Jacobi, 1 = block Jacobi."""

    def __init__(self, scene, precond=1, pcg_tol=None, newton_iters=None, newton_tol=1e-7,
                 threads=None):
        self.lib = load_oracle()
        if threads:
            self.lib.orc_set_threads(int(threads))
        self.scene = scene
        self._rest = _c64(scene.rest.ravel())
        self._tets = _c32(scene.tets.ravel())
        g = _c64(np.asarray(scene.gravity))
        pins = _c32(scene.pinned())
        targets = _c64(scene.pinned_targets().ravel()) if pins.size else np.zeros(1)
        tol = scene.pcg_tol if pcg_tol is None else pcg_tol
        ni = scene.newton_iters if newton_iters is None else newton_iters
        self.newton_iters = ni
        self.h = self.lib.orc_create(_d(self._rest), scene.nv, _i(self._tets), scene.nt,
                                     scene.material, scene.mu, scene.lam, scene.dt,
                                     scene.rayleigh_alpha, scene.rayleigh_beta, scene.mass_scale,
                                     _d(g), _i(pins if pins.size else np.zeros(1, np.int32)),
                                     int(pins.size), _d(targets), tol, scene.pcg_max_iters,
                                     precond, ni, newton_tol)
        if scene.ground_y is not None:
            self.lib.orc_set_ground(self.h, scene.ground_y, scene.ground_k)
        if getattr(scene, "integrator", 0) == 1:
            if self.lib.orc_set_velocity_form(self.h) != 0:
                raise ValueError("the velocity form needs the single-iteration step, no ground")
        for cube in getattr(scene, "collision_cubes", ()):
            c = _c64(np.asarray(cube[:3]))
            self.lib.orc_add_collision_cube(self.h, _d(c), float(cube[3]))
        if getattr(scene, "green_mu", None) is not None:
            if self.lib.orc_set_green_damping(self.h, float(scene.green_mu)) != 0:
                raise ValueError("energy damping needs the single-iteration step")
            if getattr(scene, "fdot_v", None) is not None:
                self.compute_fdots(scene.fdot_v)
        self.dofs = scene.dofs
        self.set_state(scene.initial_u(), scene.initial_v())

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, u, v):
        self.lib.orc_set_state(self.h, _d(_c64(u)), _d(_c64(v)))

    def state(self):
        u = np.zeros(self.dofs)
        v = np.zeros(self.dofs)
        self.lib.orc_get_state(self.h, _d(u), _d(v))
        return u, v

    def step(self):
        err = ctypes.c_double(0)
        sec = ctypes.c_double(0)
        per = np.zeros(max(1, self.newton_iters), np.int32)
        nd = ctypes.c_int(0)
        it = self.lib.orc_step(self.h, ctypes.byref(err), ctypes.byref(sec), _i(per),
                               ctypes.byref(nd))
        self.last_iters_per_newton = per[:nd.value].copy()
        return it, err.value, sec.value

    def plane_constraints(self):
        """Vertex ids of the live plane constraints, in creation order."""
        n = self.lib.orc_plane_constraints(self.h, None)
        out = np.zeros(max(n, 1), np.int32)
        self.lib.orc_plane_constraints(self.h, _i(out))
        return out[:n]

    def surface_vertices(self):
        n = self.lib.orc_surface_vertices(self.h, None)
        out = np.zeros(max(n, 1), np.int32)
        self.lib.orc_surface_vertices(self.h, _i(out))
        return out[:n]

    def bsr_pattern(self):
        nnzb = self.lib.orc_nnzb(self.h)
        rp = np.zeros(self.scene.nv + 1, np.int32)
        ci = np.zeros(nnzb, np.int32)
        self.lib.orc_bsr_pattern(self.h, _i(rp), _i(ci))
        return rp, ci

    def forces_hessian(self, u, with_k=True):
        nnzb = self.lib.orc_nnzb(self.h)
        R = np.zeros(self.dofs)
        K = np.zeros(9 * nnzb) if with_k else None
        self.lib.orc_forces_hessian(self.h, _d(_c64(u)), _d(R), _d(K) if with_k else None)
        return R, (K.reshape(nnzb, 3, 3) if with_k else None)

    def compute_fdots(self, vel):
        self.lib.orc_compute_fdots(self.h, _d(_c64(vel)))

    def damped_forces_hessian(self, u):
        """Elastic + damping forces and the assembled damping stiffness C at u."""
        nnzb = self.lib.orc_nnzb(self.h)
        R = np.zeros(self.dofs)
        C = np.zeros(9 * nnzb)
        self.lib.orc_damped_forces_hessian(self.h, _d(_c64(u)), _d(R), _d(C))
        return R, C.reshape(nnzb, 3, 3)

    def last_system(self):
        nnzb = self.lib.orc_nnzb(self.h)
        rhs = np.zeros(self.dofs)
        lhs = np.zeros(9 * nnzb)
        self.lib.orc_last_system(self.h, _d(rhs), _d(lhs))
        return rhs, lhs.reshape(nnzb, 3, 3)

    def rest_volumes(self):
        vol = np.zeros(self.scene.nt)
        one = np.zeros(self.scene.nv)
        self.lib.orc_rest_volumes(self.h, _d(vol), _d(one))
        return vol, one


def bsr_to_dense(row_ptr, col_idx, vals, nv):
    """Dense (3nv x 3nv) matrix from BSR arrays (row-major 3x3 blocks)."""
    A = np.zeros((3 * nv, 3 * nv))
    for i in range(nv):
        for k in range(row_ptr[i], row_ptr[i + 1]):
            j = col_idx[k]
            A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = vals[k]
    return A


def csc_to_bsr_vals(csc, row_ptr, col_idx):
    """Reference CSC (colptr, rowidx, vals) values rearranged into the given BSR pattern."""
    colptr, rowidx, vals = csc
    nv = len(row_ptr) - 1
    out = np.zeros((len(col_idx), 3, 3))
    lookup = {}
    for i in range(nv):
        for k in range(row_ptr[i], row_ptr[i + 1]):
            lookup[(i, int(col_idx[k]))] = k
    ncol = len(colptr) - 1
    for c in range(ncol):
        for q in range(colptr[c], colptr[c + 1]):
            r = int(rowidx[q])
            k = lookup[(r // 3, c // 3)]
            out[k, r % 3, c % 3] = vals[q]
    return out
