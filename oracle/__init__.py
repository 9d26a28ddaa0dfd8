"""ORACLE -- test infrastructure only (see oracle.cpp's header).

ctypes binding of liboracle.so, the CPU FP64 restatement of the reference's
contact time step.  Imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

P, V, GAMMA, F_CONTACT, R, B, K, FRICTION = range(8)
# solvers, numbered as System(int solverType) numbers them (System.cu:77-101)
APGD, PGJ, PGS = 1, 5, 6

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint32)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run `make -C oracle`")
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        sig = {
            "or_create": ([], vp),
            "or_destroy": ([vp], None),
            "or_set_geom_mode": ([ctypes.c_int], None),
            "or_get_geom_mode": ([], ctypes.c_int),
            "or_set_threads": ([ctypes.c_int], ctypes.c_int),
            "or_set_params": ([vp, _dp], None),
            "or_set_solver": ([vp, ctypes.c_int], ctypes.c_int),
            "or_set_storage": ([vp, ctypes.c_int], ctypes.c_int),
            "or_set_bins": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_int], None),
            "or_add_bodies": ([vp, ctypes.c_int, _dp, _dp, _dp, _dp, _ip, _ip], ctypes.c_int),
            "or_add_bilateral": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double], ctypes.c_int),
            "or_initialize": ([vp], ctypes.c_int),
            "or_step": ([vp], ctypes.c_int),
            "or_collide": ([vp], ctypes.c_int),
            "or_assemble": ([vp], ctypes.c_int),
            "or_num_contacts": ([vp], ctypes.c_int),
            "or_num_possible": ([vp], ctypes.c_long),
            "or_num_rows": ([vp], ctypes.c_int),
            "or_iterations": ([vp], ctypes.c_int),
            "or_backtracks": ([vp], ctypes.c_int),
            "or_residual": ([vp], ctypes.c_double),
            "or_objective": ([vp], ctypes.c_double),
            "or_time": ([vp], ctypes.c_double),
            "or_get_contacts": ([vp, _up, _up, _dp, _dp], None),
            "or_get_vec": ([vp, ctypes.c_int, _dp, ctypes.c_long], ctypes.c_int),
            "or_set_vec": ([vp, ctypes.c_int, _dp, ctypes.c_long], ctypes.c_int),
            "or_get_D": ([vp, _ip, _ip, _dp, _ip, _ip], ctypes.c_long),
            "or_schur": ([vp, _dp, _dp], ctypes.c_int),
            "or_project3": ([_dp, ctypes.c_double], None),
            "or_residual_of": ([vp, _dp], ctypes.c_double),
            "or_residual_vector": ([vp, _dp, _dp], None),
            "or_frame": ([_dp, _dp, _dp], None),
            "or_narrow_pair": ([vp, ctypes.c_int, ctypes.c_int, _dp, _dp], ctypes.c_int),
            "or_ccp_violation": ([vp, _dp], None),
            "or_add_shell": ([vp, ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, _dp, ctypes.c_long, _ip, _ip, _dp,
                              ctypes.c_long, _ip, _ip, _dp], ctypes.c_int),
            "or_num_geometries": ([vp], ctypes.c_int),
            "or_shell_point": ([vp, ctypes.c_int, _dp, _dp], None),
            "or_shell_shape": ([ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


class OracleSystem:
    """CPU restatement of System (System.cu) driven like the reference's drivers."""

    def __init__(self, scene, *, tol=None, max_iterations=None, h=None, mu=None,
                 warm_start=None, anti_relax=False, bins=None, geom_mode=1, solver=APGD, storage=0):
        L = lib()
        L.or_set_geom_mode(int(geom_mode))
        self._L = L
        self._h = L.or_create()
        prm = dict(scene.params)
        for k_, v_ in (("tol", tol), ("max_iterations", max_iterations), ("h", h), ("mu", mu),
                       ("warm_start", warm_start), ("bins", bins)):
            if v_ is not None:
                prm[k_] = v_
        self.params = prm
        g = prm["gravity"]
        a = np.array([prm["h"], prm["mu"], prm["tol"], prm["max_iterations"], g[0], g[1], g[2],
                      prm["envelope"], prm["bins"], float(bool(prm["warm_start"])), float(bool(anti_relax))],
                     dtype=np.float64)
        L.or_set_params(self._h, _d(a))
        if L.or_set_solver(self._h, int(solver)) != 0:
            raise ValueError("solver is APGD, PGJ or PGS")
        if L.or_set_storage(self._h, int(storage)) != 0:
            raise ValueError("storage is 0 (FP64), 1 (binary32 products) or 2 (binary32 products and iterates)")
        pos = np.ascontiguousarray(scene.pos, np.float64)
        vel = np.ascontiguousarray(scene.vel, np.float64)
        mass = np.ascontiguousarray(scene.mass, np.float64)
        geom = np.ascontiguousarray(scene.geom, np.float64)
        fixed = np.ascontiguousarray(scene.fixed, np.int32)
        fam = np.ascontiguousarray(scene.family, np.int32)
        L.or_add_bodies(self._h, scene.n, pos.ctypes.data_as(_dp), vel.ctypes.data_as(_dp),
                        mass.ctypes.data_as(_dp), geom.ctypes.data_as(_dp),
                        fixed.ctypes.data_as(_ip), fam.ctypes.data_as(_ip))
        mesh = getattr(scene, "shell", None)
        if mesh is not None:
            self._keep = [np.ascontiguousarray(x, dt) for x, dt in (
                (mesh.nodes, np.float64), (mesh.conn, np.int32), (mesh.geom, np.float64), (mesh.fext, np.float64),
                (mesh.mass[0], np.int32), (mesh.mass[1], np.int32), (mesh.mass[2], np.float64),
                (mesh.inv_mass[0], np.int32), (mesh.inv_mass[1], np.int32), (mesh.inv_mass[2], np.float64))]
            geom = self._keep[2].copy()
            geom[:, 3] = mesh.points  # importMesh's numContactPointsPerElement
            self._keep[2] = geom
            k = self._keep
            L.or_add_shell(self._h, mesh.n_nodes, k[0].ctypes.data_as(_dp), mesh.n_elements,
                           k[1].ctypes.data_as(_ip), k[2].ctypes.data_as(_dp), k[3].ctypes.data_as(_dp),
                           len(k[6]), k[4].ctypes.data_as(_ip), k[5].ctypes.data_as(_ip), k[6].ctypes.data_as(_dp),
                           len(k[9]), k[7].ctypes.data_as(_ip), k[8].ctypes.data_as(_ip), k[9].ctypes.data_as(_dp))
        for (a_, b_), (vv, t0) in zip(scene.bil_dofs, scene.bil_info):
            L.or_add_bilateral(self._h, int(a_), int(b_), float(vv), float(t0))
        L.or_initialize(self._h)

    def __del__(self):
        try:
            self._L.or_destroy(self._h)
        except Exception:
            pass

    def step(self):
        self._L.or_step(self._h)

    def collide(self) -> int:
        return self._L.or_collide(self._h)

    def assemble(self) -> int:
        return self._L.or_assemble(self._h)

    @property
    def num_contacts(self):
        return self._L.or_num_contacts(self._h)

    @property
    def iterations(self):
        return self._L.or_iterations(self._h)

    @property
    def residual(self):
        return self._L.or_residual(self._h)

    @property
    def objective(self):
        return self._L.or_objective(self._h)

    @property
    def backtracks(self):
        return self._L.or_backtracks(self._h)

    def vec(self, which):
        n = self._L.or_get_vec(self._h, which, None, 0)
        out = np.zeros(n, np.float64)
        self._L.or_get_vec(self._h, which, out.ctypes.data_as(_dp), n)
        return out

    def set_vec(self, which, x):
        x = np.ascontiguousarray(x, np.float64)
        self._L.or_set_vec(self._h, which, x.ctypes.data_as(_dp), x.size)

    def contacts(self):
        nc = self.num_contacts
        a = np.zeros(nc, np.uint32)
        b = np.zeros(nc, np.uint32)
        nap = np.zeros((nc, 4), np.float64)
        loc = np.zeros((nc, 3), np.float64)
        self._L.or_get_contacts(self._h, a.ctypes.data_as(_up), b.ctypes.data_as(_up),
                                nap.ctypes.data_as(_dp), loc.ctypes.data_as(_dp))
        return a, b, nap, loc

    def D(self):
        rows = ctypes.c_int()
        cols = ctypes.c_int()
        nnz = self._L.or_get_D(self._h, None, None, None, ctypes.byref(rows), ctypes.byref(cols))
        I = np.zeros(nnz, np.int32)
        J = np.zeros(nnz, np.int32)
        Vv = np.zeros(nnz, np.float64)
        self._L.or_get_D(self._h, I.ctypes.data_as(_ip), J.ctypes.data_as(_ip), Vv.ctypes.data_as(_dp),
                         ctypes.byref(rows), ctypes.byref(cols))
        return I, J, Vv, rows.value, cols.value

    def schur(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        self._L.or_schur(self._h, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp))
        return y

    def residual_of(self, gamma):
        """getResidual (APGD.cu:90-102) of `gamma` on the assembled system."""
        g = np.ascontiguousarray(gamma, np.float64)
        return self._L.or_residual_of(self._h, g.ctypes.data_as(_dp))

    def residual_vector(self, gamma):
        g = np.ascontiguousarray(gamma, np.float64)
        out = np.zeros_like(g)
        self._L.or_residual_vector(self._h, g.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
        return out

    def narrow_pair(self, a, b):
        n = np.zeros(3)
        gap = ctypes.c_double()
        kept = self._L.or_narrow_pair(self._h, a, b, n.ctypes.data_as(_dp), ctypes.byref(gap))
        return bool(kept), n, gap.value

    def ccp_violation(self):
        out = np.zeros(4)
        self._L.or_ccp_violation(self._h, out.ctypes.data_as(_dp))
        return out

    @property
    def num_geometries(self):
        """Collision geometries: one per body, then the shells' contact points."""
        return self._L.or_num_geometries(self._h)

    def shell_point(self, j):
        """Position and shape-function weights of shell contact point j."""
        pos, S = np.zeros(3), np.zeros(12)
        self._L.or_shell_point(self._h, int(j), pos.ctypes.data_as(_dp), S.ctypes.data_as(_dp))
        return pos, S


def shell_shape(xi, eta, a, b):
    """The 12 shape-function weights of the ANCF shell element at (xi, eta):
    per corner node, its position, x-gradient and y-gradient weights."""
    S = np.zeros(12)
    lib().or_shell_shape(float(xi), float(eta), float(a), float(b), S.ctypes.data_as(_dp))
    return S


def project3(g, mu):
    g = np.array(g, dtype=np.float64)
    lib().or_project3(g.ctypes.data_as(_dp), float(mu))
    return g


def frame(n):
    n = np.ascontiguousarray(n, np.float64)
    u = np.zeros(3)
    w = np.zeros(3)
    lib().or_frame(n.ctypes.data_as(_dp), u.ctypes.data_as(_dp), w.ctypes.data_as(_dp))
    return u, w


def set_threads(n: int) -> int:
    return lib().or_set_threads(int(n))
