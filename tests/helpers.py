"""Shared helpers of the test-suite (scene construction for both sides)."""
import numpy as np

import oracle
from dvide_amd import scenes, system


def both(scene, solver=system.BUILD_PLAIN, geom_mode=1, **kw):
    """The same scene as an oracle System and a GPU System."""
    o = oracle.OracleSystem(scene, geom_mode=geom_mode, **kw)
    g = system.from_scene(scene, solver, geom_mode=geom_mode, **kw)
    return o, g


def pair_keys(a, b):
    return (a.astype(np.uint64) << np.uint64(32)) | b.astype(np.uint64)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(a - b)
    n = max(np.linalg.norm(b), 1e-300)
    return d / n


def tiny_scene(pos, geom, mass=None, fixed=None, family=None, vel=None, gravity=(0.0, -9.81, 0.0),
               h=1e-3, mu=0.25, tol=1e-8, max_iterations=20000, envelope=1e-3, bilaterals=()):
    """A hand-built scene; bilaterals are (dofA, dofB or -1, velocity, start time)."""
    n = len(pos)
    bil = list(bilaterals)
    sc = scenes.Scene(
        pos=np.asarray(pos, np.float64).reshape(n, 3),
        vel=np.zeros((n, 3)) if vel is None else np.asarray(vel, np.float64).reshape(n, 3),
        mass=np.ones(n) if mass is None else np.asarray(mass, np.float64),
        geom=np.asarray(geom, np.float64).reshape(n, 3),
        fixed=np.zeros(n, np.int32) if fixed is None else np.asarray(fixed, np.int32),
        family=-np.ones(n, np.int32) if family is None else np.asarray(family, np.int32),
        bil_dofs=np.array([[a, b] for a, b, _, _ in bil], np.int32).reshape(-1, 2),
        bil_info=np.array([[v, t] for _, _, v, t in bil], np.float64).reshape(-1, 2))
    sc.params = dict(h=h, mu=mu, tol=tol, max_iterations=max_iterations, gravity=gravity,
                     envelope=envelope, bins=20, steps=1, warm_start=False)
    return sc


def shell_point_grid(mesh):
    """Positions of a flat mesh's contact points, in the order they follow
    the bodies (element by element, i then k), and each one's element and
    (xi, eta)."""
    w = mesh.points - 1
    pts, els, xe = [], [], []
    for e, c in enumerate(mesh.conn):
        a, b = mesh.geom[e, 0], mesh.geom[e, 1]
        x0, y0, z0 = mesh.nodes[c[0], 0:3]
        for i in range(1, w):
            for k in range(1, w):
                pts.append((x0 + a * i / w, y0, z0 + b * k / w))
                els.append(e)
                xe.append((i / w, k / w))
    return np.array(pts), np.array(els), np.array(xe)


def spheres_on_points(every=7, nx=3, ny=2, r=0.025, h=5e-4, mu=0.4, tol=1e-9):
    """A fixed flat shell (nx x ny elements of 0.1 m, 8 contact points per
    side) with a sphere resting exactly on top of every `every`-th contact
    point: each sphere touches that point only (its neighbours lie 14 mm to
    the side, 3 mm beyond the envelope), along the vertical."""
    mesh = scenes.flat_shell(nx, ny, 0.1, 0.1, 0.01, points=8)
    pts, _, _ = shell_point_grid(mesh)
    chosen = np.arange(0, len(pts), every)
    pos = pts[chosen].copy()
    pos[:, 1] += 0.005 + r
    # keep spheres that do not touch each other
    keep = []
    for k in range(len(pos)):
        if all(np.linalg.norm(pos[k] - pos[j]) > 2 * r + 0.005 for j in keep):
            keep.append(k)
    pos = pos[keep]
    n = len(pos)
    sc = tiny_scene(pos, [[r, 0, 0]] * n, mass=np.full(n, 2500.0 * 4.0 / 3.0 * np.pi * r ** 3), h=h, mu=mu,
                    tol=tol)
    sc.shell = mesh
    return sc, chosen[keep]


def bits_equal_or_nan(x, y):
    """Bitwise equality of float arrays, NaN matching NaN."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    if x.shape != y.shape:
        return False
    nx, ny = np.isnan(x), np.isnan(y)
    return np.array_equal(nx, ny) and np.array_equal(x[~nx].view(np.uint64), y[~ny].view(np.uint64))
