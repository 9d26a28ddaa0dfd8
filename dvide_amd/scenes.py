"""Seeded scene generators (ctypes binding of libdvide_scenes.so).

The generators play the role of the reference's scenario drivers
(models/test_compaction.cu:272-320, models/test_oneBall.cu:241-247): they
produce body arrays that are then added to a System.  All randomness comes
from std::mt19937_64 with u = (x >> 11) * 2^-53, so a (kind, seed, args)
triple names one scene on any host.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

BALL_ON_BOX, LATTICE_BOX, COMPACTION, PILE, CLOUD, FREE_FALL = range(6)

# slots of the parameter block written by the generator (scenes.cpp)
P_H, P_MU, P_TOL, P_MAXIT, P_GX, P_GY, P_GZ, P_ENVELOPE, P_BINS, P_STEPS, P_WARM = range(11)


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libdvide_scenes.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        lib.dvs_generate.restype = ctypes.c_void_p
        lib.dvs_generate.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
        lib.dvs_free.argtypes = [ctypes.c_void_p]
        for name in ("dvs_num_bodies", "dvs_num_bilaterals"):
            getattr(lib, name).argtypes = [ctypes.c_void_p]
            getattr(lib, name).restype = ctypes.c_int
        for name in ("dvs_pos", "dvs_vel", "dvs_mass", "dvs_geom", "dvs_bil_info", "dvs_params"):
            getattr(lib, name).argtypes = [ctypes.c_void_p]
            getattr(lib, name).restype = ctypes.POINTER(ctypes.c_double)
        for name in ("dvs_fixed", "dvs_family", "dvs_bil_dofs"):
            getattr(lib, name).argtypes = [ctypes.c_void_p]
            getattr(lib, name).restype = ctypes.POINTER(ctypes.c_int)
        _LIB = lib
    return _LIB


@dataclass
class Scene:
    """Body arrays of one scene plus its solver parameters."""

    pos: np.ndarray            # (n,3) float64
    vel: np.ndarray            # (n,3) float64
    mass: np.ndarray           # (n,)  float64
    geom: np.ndarray           # (n,3) float64: (r,0,0) sphere, (hx,hy,hz) box
    fixed: np.ndarray          # (n,)  int32
    family: np.ndarray         # (n,)  int32
    bil_dofs: np.ndarray       # (m,2) int32  (dofA, dofB or -1)
    bil_info: np.ndarray       # (m,2) float64 (velocity, start time)
    params: dict = field(default_factory=dict)
    shell: "ShellMesh | None" = None  # a fixed ANCF shell mesh, its DOFs after the bodies'

    @property
    def n(self) -> int:
        return int(self.mass.shape[0])


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, (n,)).astype(dtype, copy=True)


def generate(kind: int, seed: int, *args: float) -> Scene:
    """Build a scene; see dvs_generate in scenes.cpp for the per-kind args."""
    lib = _lib()
    a = (ctypes.c_double * 8)(*(list(args) + [0.0] * (8 - len(args))))
    h = lib.dvs_generate(int(kind), ctypes.c_uint64(seed), a)
    if not h:
        raise ValueError(f"unknown scene kind {kind}")
    try:
        n = lib.dvs_num_bodies(h)
        m = lib.dvs_num_bilaterals(h)
        prm = _arr(lib.dvs_params(h), 16, np.float64)
        sc = Scene(
            pos=_arr(lib.dvs_pos(h), 3 * n, np.float64).reshape(n, 3),
            vel=_arr(lib.dvs_vel(h), 3 * n, np.float64).reshape(n, 3),
            mass=_arr(lib.dvs_mass(h), n, np.float64),
            geom=_arr(lib.dvs_geom(h), 3 * n, np.float64).reshape(n, 3),
            fixed=_arr(lib.dvs_fixed(h), n, np.int32),
            family=_arr(lib.dvs_family(h), n, np.int32),
            bil_dofs=_arr(lib.dvs_bil_dofs(h), 2 * m, np.int32).reshape(m, 2),
            bil_info=_arr(lib.dvs_bil_info(h), 2 * m, np.float64).reshape(m, 2),
        )
        sc.params = dict(
            h=prm[P_H], mu=prm[P_MU], tol=prm[P_TOL], max_iterations=int(prm[P_MAXIT]),
            gravity=(prm[P_GX], prm[P_GY], prm[P_GZ]), envelope=prm[P_ENVELOPE],
            bins=int(prm[P_BINS]), steps=int(prm[P_STEPS]), warm_start=bool(prm[P_WARM]),
        )
        return sc
    finally:
        lib.dvs_free(h)


def ball_on_box(gx: float = 0.0, h: float = 1e-2) -> Scene:
    return generate(BALL_ON_BOX, 0, gx, h)


def lattice_box(seed: int = 1001, per_side: int = 10) -> Scene:
    """1,000 spheres r=0.05 rho=2500 mu=0.25 on a jittered lattice in a 1.2 m box."""
    return generate(LATTICE_BOX, seed, per_side)


LID_FREE, LID_BILATERAL, LID_NONE = 0, 1, 2


def compaction(seed: int = 2002, n: int = 112000, width: float = 5.0, lid: int = LID_FREE,
               lid_mass: float = 0.0, lid_speed: float = 0.0) -> Scene:
    """Settled granular pile under a rigid lid moving down at 0.1 m/s (~250k contacts at 112k).

    LID_FREE: a free lid as heavy as the spheres, released at -0.1 m/s, warm
    starting on (the reference's compaction driver); LID_BILATERAL: the lid's
    DOFs driven by three bilateral rows (y at -0.1 m/s).
    """
    return generate(COMPACTION, seed, n, width, 0, lid, lid_mass, lid_speed)


def pile(seed: int, target_contacts: int, rmin=0.03, rmax=0.06, mu=0.3) -> Scene:
    return generate(PILE, seed, target_contacts, rmin, rmax, mu)


def cloud(seed: int, n: int = 500, edge: float = 1.0, rmin=0.02, rmax=0.06) -> Scene:
    return generate(CLOUD, seed, n, edge, rmin, rmax)


def free_fall(seed: int, n: int = 10) -> Scene:
    return generate(FREE_FALL, seed, n)


def layer(seed: int, per_side: int = 16, rmin: float = 0.03, rmax: float = 0.06, mu: float = 0.3,
          tol: float = 1e-6, max_iterations: int = 100000) -> Scene:
    """One layer of spheres resting on a fixed floor box, far enough apart
    that each touches only the floor: every contact is the only one of its
    sphere, so N = D M^-1 D^T is block diagonal with blocks I / m and the
    contact impulses gamma are unique.  Half of the spheres slide (lateral
    speeds U[0.05, 0.5] m/s, above mu g h), the others stick (below 1e-3 m/s).
    Radii U[rmin, rmax], density 2500, y up, h = 1e-3; numpy's PCG64 seeded
    with `seed`."""
    rng = np.random.default_rng(seed)
    n = per_side * per_side
    pitch = 2.0 * rmax + 0.05
    r = rmin + (rmax - rmin) * rng.random(n)
    ix, iz = np.divmod(np.arange(n), per_side)
    pos = np.zeros((n + 1, 3))
    pos[1:, 0] = (ix - 0.5 * (per_side - 1)) * pitch
    pos[1:, 1] = r  # touching the floor's top face at y = 0
    pos[1:, 2] = (iz - 0.5 * (per_side - 1)) * pitch
    half = 0.5 * per_side * pitch + 0.1
    pos[0] = (0.0, -0.05, 0.0)
    geom = np.zeros((n + 1, 3))
    geom[0] = (half, 0.05, half)
    geom[1:, 0] = r
    ang = 2.0 * np.pi * rng.random(n)
    speed = np.where(np.arange(n) % 2 == 0, 0.05 + 0.45 * rng.random(n), 1e-3 * rng.random(n))
    vel = np.zeros((n + 1, 3))
    vel[1:, 0] = speed * np.cos(ang)
    vel[1:, 2] = speed * np.sin(ang)
    mass = np.ones(n + 1)
    mass[1:] = 2500.0 * 4.0 / 3.0 * np.pi * r ** 3
    fixed = np.zeros(n + 1, np.int32)
    fixed[0] = 1
    sc = Scene(pos=pos, vel=vel, mass=mass, geom=geom, fixed=fixed, family=-np.ones(n + 1, np.int32),
               bil_dofs=np.zeros((0, 2), np.int32), bil_info=np.zeros((0, 2)))
    sc.params = dict(h=1e-3, mu=mu, tol=tol, max_iterations=max_iterations, gravity=(0.0, -9.81, 0.0),
                     envelope=1e-3, bins=20, steps=1, warm_start=False)
    return sc


@dataclass
class ShellMesh:
    """An ANCF shell mesh as System::importMesh reads it (System.cu:1932-2060).

    nodes (n, 9): per node its position, x-gradient and y-gradient; conn
    (m, 4): 0-based nodes at the corners (0,0), (1,0), (1,1), (0,1) of each
    element's (xi, eta) square; geom (m, 4): side a (along xi), side b (along
    eta), thickness, and a fourth field importMesh overwrites with the
    contact points per side; material (m, 4) and shell_map (36 m,) feed the
    elastic forces only; fext (9 n,); mass and inv_mass: (I, J, V) triplets
    of the shell DOFs, 0-based.  A mesh without inverse-mass entries is fixed.
    """

    nodes: np.ndarray
    conn: np.ndarray
    geom: np.ndarray
    material: np.ndarray
    shell_map: np.ndarray
    fext: np.ndarray
    mass: tuple
    inv_mass: tuple
    points: int = 6  # contact points per element side (importMesh's numContactPointsPerElement)

    @property
    def n_nodes(self) -> int:
        return int(self.nodes.shape[0])

    @property
    def n_elements(self) -> int:
        return int(self.conn.shape[0])


def _no_triplets():
    return (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0))


def flat_shell(nx: int, ny: int, a: float, b: float, thickness: float, origin=(0.0, 0.0, 0.0),
               points: int = 6) -> ShellMesh:
    """A flat, fixed nx x ny mesh of a x b elements in the plane y = origin[1]
    (y up), element (i, j) spanning x in origin[0] + [i a, (i+1) a] and z in
    origin[2] + [j b, (j+1) b]: the element's xi runs along x and its eta
    along z, so every node's x-gradient is (1,0,0) and its y-gradient
    (0,0,1).  Node (i, j) is number j (nx+1) + i."""
    x0, y0, z0 = origin
    ii, jj = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1))
    ii, jj = ii.ravel(), jj.ravel()
    nodes = np.zeros((ii.size, 9))
    nodes[:, 0] = x0 + a * ii
    nodes[:, 1] = y0
    nodes[:, 2] = z0 + b * jj
    nodes[:, 3] = 1.0
    nodes[:, 8] = 1.0
    ei, ej = np.meshgrid(np.arange(nx), np.arange(ny))
    ei, ej = ei.ravel(), ej.ravel()
    n0 = ej * (nx + 1) + ei
    conn = np.stack([n0, n0 + 1, n0 + nx + 2, n0 + nx + 1], axis=1).astype(np.int32)
    m = conn.shape[0]
    geom = np.tile([a, b, thickness, float(points)], (m, 1))
    material = np.tile([7200.0, 2e8, 0.3, 0.0], (m, 1))  # density, E, nu, (unused): elastic forces only
    shell_map = (9 * np.repeat(conn, 9, axis=1) + np.tile(np.arange(9), 4)).ravel().astype(np.int32)
    return ShellMesh(nodes=nodes, conn=conn, geom=geom, material=material, shell_map=shell_map,
                     fext=np.zeros(9 * ii.size), mass=_no_triplets(), inv_mass=_no_triplets(), points=points)


def write_mesh(mesh: ShellMesh, path: str) -> None:
    """The reference's mesh file (read by System::importMesh,
    System.cu:1932-2060): comma-separated records, one per line, triplets
    1-based, each node as three lines (position, x-gradient, y-gradient)."""
    def row(v):
        return ",".join(repr(float(x)) for x in v)

    mI, mJ, mV = mesh.mass
    iI, iJ, iV = mesh.inv_mass
    with open(path, "w") as f:
        f.write(f"{mesh.n_nodes},{mesh.n_elements},{len(mV)},{len(iV)}\n")
        for nd in mesh.nodes:
            for k in range(3):
                f.write(row(nd[3 * k:3 * k + 3]) + "\n")
        for c in mesh.conn:
            f.write(",".join(str(int(x)) for x in c) + "\n")
        for m in mesh.material:
            f.write(row(m) + "\n")
        for g in mesh.geom:
            f.write(row(g) + "\n")
        for m in mesh.shell_map:
            f.write(f"{int(m)}\n")
        for x in mesh.fext:
            f.write(repr(float(x)) + "\n")
        for I, J, V in ((mI, mJ, mV), (iI, iJ, iV)):
            for i, j, v in zip(I, J, V):
                f.write(f"{int(i) + 1},{int(j) + 1},{float(v)!r}\n")


def read_mesh(path: str, points: int) -> ShellMesh:
    """Parse a mesh file as System::importMesh does (commas or blanks between
    fields; the geometry's fourth field replaced by `points`)."""
    with open(path) as f:
        lines = [ln.replace(",", " ").split() for ln in f]
    it = iter(lines)
    nn, ne, nm, ninv = (int(x) for x in next(it)[:4])
    rows = [[float(x) for x in next(it)[:3]] for _ in range(3 * nn)]
    nodes = np.array(rows, dtype=np.float64).reshape(nn, 9)
    conn = np.array([[int(x) for x in next(it)[:4]] for _ in range(ne)], np.int32)
    material = np.array([[float(x) for x in next(it)[:4]] for _ in range(ne)])
    geom = np.array([[float(x) for x in next(it)[:4]] for _ in range(ne)])
    geom[:, 3] = points
    shell_map = np.array([int(next(it)[0]) for _ in range(36 * ne)], np.int32)
    fext = np.array([float(next(it)[0]) for _ in range(9 * nn)])

    def triplets(k):
        t = [next(it)[:3] for _ in range(k)]
        return (np.array([int(x[0]) - 1 for x in t], np.int32), np.array([int(x[1]) - 1 for x in t], np.int32),
                np.array([float(x[2]) for x in t]))

    mass = triplets(nm)
    inv = triplets(ninv)
    return ShellMesh(nodes=nodes, conn=conn, geom=geom, material=material, shell_map=shell_map, fext=fext,
                     mass=mass, inv_mass=inv, points=points)


def shell_bed(seed: int, nx: int = 8, ny: int = 16, a: float = 0.1, b: float = 0.1, thickness: float = 0.01,
              per_side: tuple = (6, 12), layers: int = 2, rmin: float = 0.02, rmax: float = 0.03,
              mu: float = 0.4, h: float = 5e-4, points: int = 8, tol: float = 1e-5,
              max_iterations: int = 100000) -> Scene:
    """Spheres dropped onto a fixed flat ANCF shell (the sphere-on-shell
    workload with its shell held fixed): an nx x ny mesh of a x b elements in
    the plane y = 0 with `points` contact points per element side, and a
    lattice of per_side[0] x per_side[1] x `layers` spheres of radius U[rmin,
    rmax], density 2500, jittered, above the mesh's middle, the lowest layer
    touching the highest contact point below each sphere.  With 8 points per
    0.1 m side the widest gap between points (at element corners) is narrower
    than a sphere of radius 0.016 plus a point, so no sphere falls through.  y up; numpy's PCG64 seeded with
    `seed`."""
    rng = np.random.default_rng(seed)
    mesh = flat_shell(nx, ny, a, b, thickness, points=points)
    px, pz = per_side
    n = px * pz * layers
    pitch = 2.0 * rmax + 0.01
    ix = np.arange(n) % px
    iz = (np.arange(n) // px) % pz
    iy = np.arange(n) // (px * pz)
    r = rmin + (rmax - rmin) * rng.random(n)
    pos = np.zeros((n, 3))
    pos[:, 0] = 0.5 * nx * a + (ix - 0.5 * (px - 1)) * pitch + 0.002 * (rng.random(n) - 0.5)
    pos[:, 2] = 0.5 * ny * b + (iz - 0.5 * (pz - 1)) * pitch + 0.002 * (rng.random(n) - 0.5)
    # the lowest layer rests on the highest contact point below each sphere
    # (the points of a flat mesh form the grid xs x zs)
    w = points - 1
    xs = np.unique(np.round((np.arange(nx)[:, None] + np.arange(1, w)[None, :] / w).ravel() * a, 12))
    zs = np.unique(np.round((np.arange(ny)[:, None] + np.arange(1, w)[None, :] / w).ravel() * b, 12))
    rho = 0.5 * thickness
    for k in range(n):
        reach = r[k] + rho
        dx = xs[np.abs(xs - pos[k, 0]) < reach] - pos[k, 0]
        dz = zs[np.abs(zs - pos[k, 2]) < reach] - pos[k, 2]
        d2 = dx[:, None] ** 2 + dz[None, :] ** 2
        top = np.sqrt(np.maximum(reach * reach - d2, 0.0)).max() if d2.size else 0.0
        pos[k, 1] = top + iy[k] * pitch
    mass = 2500.0 * 4.0 / 3.0 * np.pi * r ** 3
    geom = np.zeros((n, 3))
    geom[:, 0] = r
    sc = Scene(pos=pos, vel=np.zeros((n, 3)), mass=mass, geom=geom, fixed=np.zeros(n, np.int32),
               family=-np.ones(n, np.int32), bil_dofs=np.zeros((0, 2), np.int32), bil_info=np.zeros((0, 2)))
    sc.params = dict(h=h, mu=mu, tol=tol, max_iterations=max_iterations, gravity=(0.0, -9.81, 0.0),
                     envelope=1e-3, bins=20, steps=1, warm_start=True)
    sc.shell = mesh
    return sc
