"""Synthetic tetrahedral scenes for the backward-Euler Newton step.

The reference builds its meshes from .tobj files (src/Geometry/TET_MESH.cpp:1174-1226) and
hard-coded scene classes (e.g. src/Scenes/DROP_TEST.h:34-95).  The workloads used here are
regular cube lattices instead, so this module defines them deterministically:

* Kuhn (Freudenthal) split: every lattice cell is cut into the 6 tets that share the cell's
  (0,0,0)-(1,1,1) diagonal.  Vertex order inside each tet is chosen so that every rest volume is
  positive, which the reference asserts (src/Geometry/TET_MESH.cpp:132-137).
* Vertex numbering is x-fastest: v(i, j, k) = i + (nx+1) * (j + (ny+1) * k).
* Tet numbering is cell-major (x-fastest cells), 6 tets per cell in a fixed permutation order.
* Random numbers come from a counter-based SplitMix64 stream, so the same (seed, stream, index)
  gives the same double everywhere.
* Density enters as the reference would take it: the lumped mass (one-ring volumes,
  src/Timestepper/TIMESTEPPER.cpp:176-208) is scaled by rho through scaleMass
  (src/Timestepper/TIMESTEPPER.h:100) and gravity is passed as the body force rho * g
  (src/Timestepper/TIMESTEPPER.cpp:334-345).
* "x = 0 face pinned" is the reference's kinematic CUBE with centre (-5, 0.5, 0.5) and scale 10
  (src/Geometry/CUBE.cpp:27-55): it contains exactly the vertices with X <= 0.  Pinned
  displacements are what TIMESTEPPER::applyKinematicConstraints computes
  (src/Timestepper/TIMESTEPPER.cpp:105-122): world(local(X)) - X.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

MAT_SNH = 0
MAT_ARAP = 1

PRECOND_JACOBI = 0
PRECOND_BLOCK_JACOBI = 1

# TIMESTEPPER::BACKWARD_EULER_POSITION (src/Timestepper/BACKWARD_EULER_POSITION.cpp) or
# TIMESTEPPER::BACKWARD_EULER_VELOCITY (src/Timestepper/BACKWARD_EULER_VELOCITY.cpp)
INTEGRATOR_POSITION = 0
INTEGRATOR_VELOCITY = 1

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform01(seed: int, stream: int, n: int) -> np.ndarray:
    """n doubles in [0, 1) from the SplitMix64 counter stream (seed, stream)."""
    with np.errstate(over="ignore"):
        key = _mix(np.array([(seed * 0x100000001B3 + stream * 0x51_7CC1_B727_220A) & 0xFFFFFFFFFFFFFFFF],
                            dtype=np.uint64) + _GOLDEN)[0]
        ctr = key + (np.arange(1, n + 1, dtype=np.uint64) * _GOLDEN)
    return (_mix(ctr) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform(seed: int, stream: int, n: int, lo: float, hi: float) -> np.ndarray:
    return lo + (hi - lo) * uniform01(seed, stream, n)


# The 6 axis orders of the Kuhn split; odd permutations get vertices 1 and 2 swapped so that
# det[x1-x0, x2-x0, x3-x0] > 0.
_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
_PERM_PARITY = (0, 1, 1, 0, 0, 1)


def kuhn_lattice(nx: int, ny: int, nz: int, h: float = None, origin=(0.0, 0.0, 0.0)):
    """Rest vertices (nv, 3) float64 and tets (nt, 4) int32 of an nx*ny*nz Kuhn lattice."""
    if h is None:
        h = 1.0 / max(nx, ny, nz)
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    # x-fastest numbering
    i = i.transpose(2, 1, 0).ravel()
    j = j.transpose(2, 1, 0).ravel()
    k = k.transpose(2, 1, 0).ravel()
    verts = np.stack([origin[0] + i * h, origin[1] + j * h, origin[2] + k * h], axis=1).astype(np.float64)

    def vid(ci, cj, ck):
        return ci + (nx + 1) * (cj + (ny + 1) * ck)

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci = ci.transpose(2, 1, 0).ravel()
    cj = cj.transpose(2, 1, 0).ravel()
    ck = ck.transpose(2, 1, 0).ravel()
    ncell = ci.size
    tets = np.empty((ncell, 6, 4), dtype=np.int64)
    for p, perm in enumerate(_PERMS):
        c = [np.zeros(ncell, np.int64)] * 3
        corner = [ci.copy(), cj.copy(), ck.copy()]
        chain = [vid(*corner)]
        for ax in perm:
            corner[ax] = corner[ax] + 1
            chain.append(vid(*corner))
        if _PERM_PARITY[p]:
            chain[1], chain[2] = chain[2], chain[1]
        tets[:, p, :] = np.stack(chain, axis=1)
        del c
    return verts, tets.reshape(-1, 4).astype(np.int32)


def interior_mask(nx, ny, nz):
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    i = i.transpose(2, 1, 0).ravel()
    j = j.transpose(2, 1, 0).ravel()
    k = k.transpose(2, 1, 0).ravel()
    return (i > 0) & (i < nx) & (j > 0) & (j < ny) & (k > 0) & (k < nz)


def jitter(verts, nx, ny, nz, h, seed, frac=0.1):
    """Uniform +-frac*h jitter on every vertex that lies on no face of the box."""
    out = verts.copy()
    mask = interior_mask(nx, ny, nz)
    n = verts.shape[0]
    d = uniform(seed, 1, 3 * n, -frac * h, frac * h).reshape(n, 3)
    out[mask] += d[mask]
    return out


def tet_volumes(verts, tets):
    x0 = verts[tets[:, 0]]
    d1 = verts[tets[:, 1]] - x0
    d2 = verts[tets[:, 2]] - x0
    d3 = verts[tets[:, 3]] - x0
    return np.einsum("ij,ij->i", d3, np.cross(d1, d2)) / 6.0


def lame(E, nu):
    """HYPERELASTIC::computeMu / computeLambda (src/Hyperelastic/Volume/HYPERELASTIC.cpp:111-122)."""
    mu = E / (2.0 * (1.0 + nu))
    lam = (E * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return mu, lam


class Cube:
    """Axis-aligned kinematic cube: local = S^-1 R^T (world - t), world = R S local + t with R = I.

    Same transforms as src/Geometry/KINEMATIC_SHAPE.h:55-73 and inside test of
    src/Geometry/CUBE.cpp:44-55.
    """

    def __init__(self, center, scale):
        self.center = np.asarray(center, dtype=np.float64)
        self.scale = float(scale)
        s = self.scale
        self.inv = (s * s) * (1.0 / ((s * s) * s))  # Eigen's cofactor inverse (CUBE.cpp:28)

    def to_local(self, w):
        return self.inv * (w - self.center)

    def to_world(self, l):
        return self.scale * l + self.center

    def inside(self, w):
        l = self.to_local(w)
        return np.all((l <= 0.5) & (l >= -0.5), axis=-1)


PIN_X0_CUBE = (-5.0, 0.5, 0.5, 10.0)


@dataclasses.dataclass
class Scene:
    """Everything one instance of the backward-Euler step needs."""

    name: str
    rest: np.ndarray          # (nv, 3) float64
    tets: np.ndarray          # (nt, 4) int32
    material: int             # MAT_SNH / MAT_ARAP
    mu: float
    lam: float
    dt: float = 1.0 / 60.0
    rayleigh_alpha: float = 0.01
    rayleigh_beta: float = 0.01
    mass_scale: float = 1.0   # density rho, applied like scaleMass(rho)
    gravity: tuple = (0.0, 0.0, 0.0)   # body force per unit mass-scaled volume (rho * g)
    pin_cube: Optional[tuple] = None   # (cx, cy, cz, scale) of the pinning CUBE
    u0: Optional[np.ndarray] = None    # (3 nv) initial displacement
    v0: Optional[np.ndarray] = None    # (3 nv) initial velocity
    pcg_tol: float = 1e-8
    pcg_max_iters: int = 0             # 0 -> 2 * DOFs, the reference default
    newton_iters: int = 1
    ground_y: Optional[float] = None   # penalty ground plane (no reference counterpart)
    ground_k: float = 0.0
    collision_cubes: tuple = ()        # ((cx, cy, cz, scale), ...) kinematic collision CUBEs
    green_mu: Optional[float] = None   # GREEN_DAMPING(mu) energy damping in place of Rayleigh
    fdot_v: Optional[np.ndarray] = None  # (3 nv) velocity given to computeFdots before stepping
    integrator: int = 0                # INTEGRATOR_POSITION or INTEGRATOR_VELOCITY
    dims: tuple = ()

    @property
    def nv(self):
        return self.rest.shape[0]

    @property
    def nt(self):
        return self.tets.shape[0]

    @property
    def dofs(self):
        return 3 * self.nv

    def pinned(self):
        """Vertex ids inside the pin cube, ascending (TIMESTEPPER::attachKinematicConstraints)."""
        if self.pin_cube is None:
            return np.zeros(0, np.int32)
        c = Cube(self.pin_cube[:3], self.pin_cube[3])
        return np.nonzero(c.inside(self.rest))[0].astype(np.int32)

    def pinned_targets(self):
        """Displacements that applyKinematicConstraints writes for the pinned vertices."""
        ids = self.pinned()
        if ids.size == 0:
            return np.zeros((0, 3))
        c = Cube(self.pin_cube[:3], self.pin_cube[3])
        X = self.rest[ids]
        return c.to_world(c.to_local(X)) - X

    def initial_u(self):
        return np.zeros(self.dofs) if self.u0 is None else self.u0.copy()

    def initial_v(self):
        return np.zeros(self.dofs) if self.v0 is None else self.v0.copy()


def lattice_scene(n=(10, 10, 10), seed=1, E=1e5, nu=0.45, material=MAT_SNH, rho=1000.0,
                  g=-9.8, h=None, stretch_velocity=False, name=None, pin=True, mu=None):
    """The pinned-cube workload family.

    n=(10,10,10): 6,000 tets / 1,331 vertices; n=(40,40,40): 384,000 tets / 68,921 vertices.
    stretch_velocity: initial velocity v = (s - 1) * (X - c) / dt with one seeded per-axis
    stretch s ~ U[0.7, 1.3]^3 and c the box centre; pinned vertices start at rest.
    """
    nx, ny, nz = n
    if h is None:
        h = 1.0 / max(n)
    verts, tets = kuhn_lattice(nx, ny, nz, h)
    verts = jitter(verts, nx, ny, nz, h, seed)
    m, l = lame(E, nu)
    if mu is not None:
        m = mu
    sc = Scene(
        name=name or f"lattice{nx}x{ny}x{nz}",
        rest=verts, tets=tets, material=material, mu=m, lam=l,
        mass_scale=rho, gravity=(0.0, rho * g, 0.0),
        pin_cube=PIN_X0_CUBE if pin else None, dims=(nx, ny, nz),
    )
    if stretch_velocity:
        s = uniform(seed, 2, 3, 0.7, 1.3)
        c = np.array([0.5 * nx * h, 0.5 * ny * h, 0.5 * nz * h])
        v = ((s - 1.0)[None, :] * (verts - c[None, :]) / sc.dt)
        ids = sc.pinned()
        v[ids] = 0.0
        sc.v0 = v.ravel().copy()
    return sc


def drop_scene(n=(6, 6, 6), seed=7, E=3.0, nu=0.48, material=MAT_SNH, height=0.02,
               speed=3.0, ground=(0.0, -5.0, 0.0, 10.0), name=None):
    """A lattice dropped onto a kinematic ground CUBE, as src/Scenes/DROP_TEST.h does with its
    cylinder: SNH E=3 nu=0.48, density 1, gravity (0, -1, 0), every vertex moving down at 3, the
    ground CUBE(center (0, -5, 0), scale 10) whose top face is y = 0.  The unit lattice starts
    `height` above the ground; its surface vertices become plane constraints on contact
    (TIMESTEPPER.cpp:446-526).  No vertex is pinned."""
    nx, ny, nz = n
    h = 1.0 / max(n)
    verts, tets = kuhn_lattice(nx, ny, nz, h)
    verts = jitter(verts, nx, ny, nz, h, seed)
    verts = verts + np.array([-0.5 * nx * h, height, -0.5 * nz * h])[None, :]
    m, l = lame(E, nu)
    sc = Scene(name=name or f"drop{nx}x{ny}x{nz}", rest=verts, tets=tets, material=material,
               mu=m, lam=l, mass_scale=1.0, gravity=(0.0, -1.0, 0.0), pin_cube=None,
               collision_cubes=(tuple(float(x) for x in ground),), dims=(nx, ny, nz))
    v = np.zeros((verts.shape[0], 3))
    v[:, 1] = -speed
    sc.v0 = v.ravel().copy()
    return sc


def drop_test(tobj_path, pcg_tol=1e-8):
    """src/Scenes/DROP_TEST.h on a .tobj mesh (the reference uses data/cylinder_10.tobj): SNH
    E=3 nu=0.48, density 1, gravity (0, -1, 0), every vertex moving down at 3, the ground
    CUBE(center (0, -5, 0), scale 10), self-collisions off.  The PCG tolerance is the workloads'
    1e-8 (the scene itself keeps Eigen's default)."""
    from .formats import read_tobj

    verts, tets = read_tobj(tobj_path)
    m, l = lame(3.0, 0.48)
    sc = Scene(name="drop_test", rest=verts, tets=tets, material=MAT_SNH, mu=m, lam=l,
               mass_scale=1.0, gravity=(0.0, -1.0, 0.0), pin_cube=None, pcg_tol=pcg_tol,
               collision_cubes=((0.0, -5.0, 0.0, 10.0),))
    v = np.zeros((verts.shape[0], 3))
    v[:, 1] = -3.0
    sc.v0 = v.ravel().copy()
    return sc


def separation_test(tobj_path, pcg_tol=1e-8):
    """The reference's collision separation test (projects/unitTests/SCENE_TESTS.h:367-453) on
    the velocity integrator: data/cube_6.tobj, SNH(mu = 1, lambda = 10), BACKWARD_EULER_VELOCITY
    with self-collisions off, the floor CUBE(center (0, -5, 0), scale 10) as a kinematic collision
    object, every vertex moving down at 3, gravity (0, -1, 0), 42 steps; afterwards every surface
    vertex must be above y = 1e-4 (the cube bounced off the floor).  The test keeps Eigen's
    default PCG tolerance (machine epsilon); the workloads' 1e-8 is the default here."""
    from .formats import read_tobj

    verts, tets = read_tobj(tobj_path)
    sc = Scene(name="separation_test", rest=verts, tets=tets, material=MAT_SNH, mu=1.0, lam=10.0,
               mass_scale=1.0, gravity=(0.0, -1.0, 0.0), pin_cube=None, pcg_tol=pcg_tol,
               collision_cubes=((0.0, -5.0, 0.0, 10.0),), integrator=INTEGRATOR_VELOCITY)
    v = np.zeros((verts.shape[0], 3))
    v[:, 1] = -3.0
    sc.v0 = v.ravel().copy()
    return sc


def green_scene(n=(8, 8, 8), seed=11, E=1e5, nu=0.45, material=MAT_SNH, green_frac=0.02,
                rate=1.0, name=None):
    """The pinned lattice with GREEN_DAMPING(green_frac * mu) in place of Rayleigh damping
    (BACKWARD_EULER_POSITION::solveEnergyDamped).  The initial velocity is the affine field
    v = B (X - c), B = a seeded diagonal stretch rate plus a seeded spin, entries ~ `rate` per
    second (pinned vertices at rest), and the same field is handed to computeFdots, so the
    energy-damped steps carry the damping force of that velocity gradient."""
    sc = lattice_scene(n, seed=seed, E=E, nu=nu, material=material, name=name or f"green{n[0]}x{n[1]}x{n[2]}")
    sc.green_mu = green_frac * sc.mu
    d = uniform(seed, 20, 3, -rate, rate)
    w = uniform(seed, 21, 3, -rate, rate)
    B = np.diag(d) + np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])
    c = sc.rest.mean(axis=0)
    v = (sc.rest - c[None, :]) @ B.T
    v[sc.pinned()] = 0.0
    sc.v0 = v.ravel().copy()
    sc.fdot_v = v.ravel().copy()
    return sc


def two_tet_drop(tobj_path, E=3.0, name="two_tet_drop"):
    """src/Scenes/TWO_TET_DROP_VF.h (E = 3) / TWO_TET_DROP_EE.h (E = 1) without their
    self-collision response: SNH nu=0.45, density 1, gravity (0, -0.3, 0), dt = 1/30, Rayleigh
    0.01/0.01, the ground
    CUBE(center (0, -4.9, 0), scale 10) both pinning the surface vertices inside it
    (attachKinematicSurfaceConstraints; the two meshes are all surface) and acting as a kinematic
    collision object, Eigen's default PCG tolerance.  Until the first frame that detects a
    self-collision (frame 27 of two_tet_drop_vf.json, 36 of two_tet_drop_ee.json) the reference's
    step is exactly this one."""
    from .formats import read_tobj

    verts, tets = read_tobj(tobj_path)
    m, l = lame(E, 0.45)
    return Scene(name=name, rest=verts, tets=tets, material=MAT_SNH, mu=m, lam=l, dt=1.0 / 30.0,
                 mass_scale=1.0, gravity=(0.0, -0.3, 0.0), pin_cube=(0.0, -4.9, 0.0, 10.0),
                 pcg_tol=0.0, collision_cubes=((0.0, -4.9, 0.0, 10.0),))


def config1_velocity(seed=1):
    """The config-1 scene (10^3 pinned SNH cube) on BACKWARD_EULER_VELOCITY."""
    sc = config1(seed)
    sc.integrator = INTEGRATOR_VELOCITY
    sc.name = "config1_velocity"
    return sc


def config1(seed=1):
    """10^3 lattice, SNH E=1e5 nu=0.45 rho=1000, x=0 pinned, dt=1/60, 1 Newton iteration."""
    return lattice_scene((10, 10, 10), seed=seed, name="cube10_snh")


def config2(seed=2):
    """40^3 lattice with the seeded [0.7, 1.3] stretch applied as a velocity field."""
    return lattice_scene((40, 40, 40), seed=seed, stretch_velocity=True, name="cube40_snh_stretch")


def config3(material=MAT_SNH, seed=2):
    """The 40^3 stretch scene, SNH (E=1e5, nu=0.45) or ARAP (mu=3.4e4)."""
    if material == MAT_ARAP:
        return lattice_scene((40, 40, 40), seed=seed, stretch_velocity=True, material=MAT_ARAP,
                             mu=3.4e4, name="cube40_arap_stretch")
    return config2(seed)


def batch_instances(count=256, seed=4, lo=4, hi=24):
    """Independent instances with per-axis lattice dims ~ U{lo..hi}, SNH/ARAP 50/50,
    E log-uniform in [1e4, 1e6], nu ~ U[0.3, 0.48]."""
    dims = np.floor(uniform(seed, 10, 3 * count, lo, hi + 1)).astype(int).reshape(count, 3)
    dims = np.clip(dims, lo, hi)
    mats = (uniform(seed, 11, count, 0, 1) < 0.5).astype(int)
    Es = np.exp(uniform(seed, 12, count, math.log(1e4), math.log(1e6)))
    nus = uniform(seed, 13, count, 0.3, 0.48)
    out = []
    for i in range(count):
        sc = lattice_scene(tuple(int(d) for d in dims[i]), seed=seed * 100003 + i, E=float(Es[i]),
                           nu=float(nus[i]), material=int(mats[i]), stretch_velocity=True,
                           name=f"inst{i}")
        out.append(sc)
    return out


def serving_instance(k, seed=5, ground_y=-0.5, ground_k=1e6):
    """A request-stream instance: total tets log-uniform in [1e3, 4e5] with aspect ratio in
    [0.5, 2], mixed SNH/ARAP, dropped from a height in [0, 1] onto a penalty ground plane."""
    r = uniform(seed * 7919 + k, 20, 8, 0.0, 1.0)
    tets = math.exp(math.log(1e3) + r[0] * (math.log(4e5) - math.log(1e3)))
    cells = tets / 6.0
    ax = 0.5 + 1.5 * r[1]
    az = 0.5 + 1.5 * r[2]
    base = (cells / (ax * az)) ** (1.0 / 3.0)
    nx = max(2, int(round(base * ax)))
    ny = max(2, int(round(base)))
    nz = max(2, int(round(base * az)))
    mat = MAT_SNH if r[3] < 0.5 else MAT_ARAP
    E = math.exp(math.log(1e4) + r[4] * (math.log(1e6) - math.log(1e4)))
    nu = 0.3 + 0.18 * r[5]
    drop = r[6]
    h = 1.0 / max(nx, ny, nz)
    sc = lattice_scene((nx, ny, nz), seed=seed * 1000003 + k, E=E, nu=nu, material=mat,
                       pin=False, name=f"req{k}", h=h)
    sc.rest = sc.rest + np.array([0.0, drop, 0.0])
    sc.ground_y = ground_y
    sc.ground_k = ground_k
    return sc
