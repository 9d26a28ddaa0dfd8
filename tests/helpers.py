"""Shared helpers of the test-suite (scene construction for both sides)."""
import numpy as np

import oracle
from dvide_amd import scenes, system


def both(scene, solver=system.SOLVER_PLAIN, geom_mode=1, **kw):
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
