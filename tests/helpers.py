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
