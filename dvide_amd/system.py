"""Python mirror of the reference's System / Body / Solver / APGD surface.

The reference drives a scene through C++ classes (System.cuh:279-326,
Body.cuh:28-99, Solver.cuh:17-46, APGD.cuh:62-67).  This module exposes the
same names over the C ABI of libdvide_hip.so (include/dvide.h) so that tests
and drivers read like the reference's drivers (models/test_oneBall.cu,
models/test_compaction.cu).  All compute happens in the HIP library; if the
library is missing or no GPU is present the calls fail with an error.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdvide_hip.so")

# include/dvide.h
DV_OK = 0
DV_ERR_ARG, DV_ERR_STATE, DV_ERR_HIP, DV_ERR_NOMEM, DV_ERR_UNSUPPORTED, DV_ERR_DIVERGED = -1, -2, -3, -4, -5, -6
SOLVER_APGD = 1  # System(int solverType): 2..7 (PDIP .. PJKIP) are rejected
BUILD_FP64, BUILD_MIXED, BUILD_PLAIN = 0, 1, 2  # DV_APGD_*
(PARAM_H, PARAM_MU, PARAM_TOL, PARAM_MAX_ITER, PARAM_GX, PARAM_GY, PARAM_GZ, PARAM_ENVELOPE,
 PARAM_WARM_START, PARAM_APGD_BUILD, PARAM_GEOM_MODE, PARAM_TIME, PARAM_VERBOSE, PARAM_GRID,
 PARAM_PROFILE, PARAM_ROWS_PER_THREAD, PARAM_ANTI_RELAX) = range(17)
PROFILE_OFF, PROFILE_P1, PROFILE_P2, PROFILE_DECIDE = 0, 1, 2, 4  # DV_PARAM_PROFILE values
VEC_P, VEC_V, VEC_GAMMA, VEC_F_CONTACT, VEC_R, VEC_K, VEC_B = range(7)

EXPORTED = (
    "dv_scene_create", "dv_scene_destroy", "dv_last_error", "dv_add_bodies", "dv_add_bilateral_dof",
    "dv_import_mesh", "dv_add_shell_mesh", "dv_num_shell_nodes", "dv_set_bins_per_axis", "dv_set_param", "dv_get_param", "dv_initialize", "dv_step",
    "dv_num_bodies", "dv_get_bodies", "dv_num_contacts", "dv_get_stats", "dv_get_vector", "dv_set_vector",
    "dv_get_contacts", "dv_ccp_violation", "dv_assemble", "dv_schur_product", "dv_schur_product_noalloc",
    "dv_pass_product", "dv_project_cone", "dv_solve", "dv_get_jacobian", "dv_contact_forces", "dv_energies",
    "dv_batch_create", "dv_batch_destroy", "dv_batch_last_error", "dv_batch_add", "dv_batch_step",
    "dv_batch_get_stats",
    "dv_server_create", "dv_server_destroy", "dv_server_last_error", "dv_server_submit", "dv_server_poll",
    "dv_server_pending", "dv_server_failed",
)


class DvStats(ctypes.Structure):
    _fields_ = [
        ("num_contacts", ctypes.c_int), ("num_bilaterals", ctypes.c_int),
        ("iterations", ctypes.c_int), ("backtracks", ctypes.c_int),
        ("residual", ctypes.c_double), ("objective", ctypes.c_double),
        ("ms_total", ctypes.c_double), ("ms_collision", ctypes.c_double),
        ("ms_assembly", ctypes.c_double), ("ms_solve", ctypes.c_double),
        ("ms_update", ctypes.c_double), ("time", ctypes.c_double),
        ("prof_kernel_ms", ctypes.c_double), ("prof_kernel_count", ctypes.c_int),
        ("passes", ctypes.c_int),
        ("rows_per_thread", ctypes.c_int), ("xcd_order", ctypes.c_int),
        ("wg_per_cu_contacts", ctypes.c_int), ("wg_per_cu_bodies", ctypes.c_int),
        ("schedule_tuning", ctypes.c_int),
    ]


def gpu_visible() -> bool:
    """True when the HIP runtime reports at least one device."""
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return False
    n = ctypes.c_int(0)
    return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0


class DvBatchStats(ctypes.Structure):
    _fields_ = [
        ("num_scenes", ctypes.c_int), ("max_iterations", ctypes.c_int),
        ("total_contacts", ctypes.c_long), ("total_iterations", ctypes.c_long),
        ("ms_total", ctypes.c_double), ("ms_assembly", ctypes.c_double),
        ("ms_solve", ctypes.c_double), ("ms_update", ctypes.c_double),
        ("rows_per_thread", ctypes.c_int), ("xcd_order", ctypes.c_int),
        ("wg_per_cu_contacts", ctypes.c_int), ("wg_per_cu_bodies", ctypes.c_int),
        ("schedule_tuning", ctypes.c_int),
    ]


class DvideError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dvide error {code}: {msg}")
        self.code = code


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint32)


def load_library(path: str = LIB_PATH):
    """Load libdvide_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    sig = {
        "dv_scene_create": ([ctypes.POINTER(vp), ctypes.c_int], ctypes.c_int),
        "dv_scene_destroy": ([vp], ctypes.c_int),
        "dv_last_error": ([vp], ctypes.c_char_p),
        "dv_add_bodies": ([vp, ctypes.c_int, _dp, _dp, _dp, _dp, _ip, _ip], ctypes.c_int),
        "dv_add_bilateral_dof": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double], ctypes.c_int),
        "dv_import_mesh": ([vp, ctypes.c_char_p, ctypes.c_double, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "dv_add_shell_mesh": ([vp, ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, ctypes.c_int, _dp], ctypes.c_int),
        "dv_num_shell_nodes": ([vp, _ip], ctypes.c_int),
        "dv_set_bins_per_axis": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "dv_set_param": ([vp, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "dv_get_param": ([vp, ctypes.c_int, _dp], ctypes.c_int),
        "dv_initialize": ([vp], ctypes.c_int),
        "dv_step": ([vp], ctypes.c_int),
        "dv_num_bodies": ([vp, _ip], ctypes.c_int),
        "dv_get_bodies": ([vp, _dp, _dp, _ip, _ip, ctypes.c_size_t], ctypes.c_int),
        "dv_num_contacts": ([vp, _ip], ctypes.c_int),
        "dv_get_stats": ([vp, ctypes.POINTER(DvStats)], ctypes.c_int),
        "dv_get_vector": ([vp, ctypes.c_int, _dp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
        "dv_set_vector": ([vp, ctypes.c_int, _dp, ctypes.c_size_t], ctypes.c_int),
        "dv_get_contacts": ([vp, _up, _up, _dp, _dp, ctypes.c_size_t], ctypes.c_int),
        "dv_ccp_violation": ([vp, _dp], ctypes.c_int),
        "dv_assemble": ([vp, _ip], ctypes.c_int),
        "dv_schur_product": ([vp, _dp, _dp, ctypes.c_size_t], ctypes.c_int),
        "dv_schur_product_noalloc": ([vp, _dp, _dp, ctypes.c_size_t], ctypes.c_int),
        "dv_pass_product": ([vp, ctypes.c_int, _dp, _dp, ctypes.c_size_t], ctypes.c_int),
        "dv_project_cone": ([vp, _dp, ctypes.c_size_t, ctypes.c_double], ctypes.c_int),
        "dv_solve": ([vp], ctypes.c_int),
        "dv_get_jacobian": ([vp, _ip, _ip, _dp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), _ip, _ip],
                            ctypes.c_int),
        "dv_contact_forces": ([vp, _dp, _dp, ctypes.c_size_t], ctypes.c_int),
        "dv_energies": ([vp, _dp], ctypes.c_int),
        "dv_batch_create": ([ctypes.POINTER(vp), ctypes.c_int], ctypes.c_int),
        "dv_batch_destroy": ([vp], ctypes.c_int),
        "dv_batch_last_error": ([vp], ctypes.c_char_p),
        "dv_batch_add": ([vp, vp], ctypes.c_int),
        "dv_batch_step": ([vp], ctypes.c_int),
        "dv_batch_get_stats": ([vp, ctypes.POINTER(DvBatchStats)], ctypes.c_int),
        "dv_server_create": ([ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "dv_server_destroy": ([vp], ctypes.c_int),
        "dv_server_last_error": ([vp], ctypes.c_char_p),
        "dv_server_submit": ([vp, vp, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "dv_server_poll": ([vp, ctypes.POINTER(ctypes.c_long), ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
                           ctypes.c_int),
        "dv_server_pending": ([vp, _ip], ctypes.c_int),
        "dv_server_failed": ([vp, ctypes.POINTER(ctypes.c_long), ctypes.c_int, _ip], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Body:
    """A 3-DOF point mass with sphere or box geometry (Body.cuh:13-99)."""

    def __init__(self, position=(0.0, 0.0, 0.0)):
        self.pos = np.array(position, dtype=np.float64)
        self.vel = np.zeros(3)
        self.mass = 1.0
        self.fixed = False
        self.geometry = np.array([1.0, 0.0, 0.0])
        self.family = -1

    def setMass(self, mass):
        self.mass = float(mass)

    def setBodyFixed(self, fixed):
        self.fixed = bool(fixed)

    def setGeometry(self, geometry):
        self.geometry = np.array(geometry, dtype=np.float64)

    def setVelocity(self, velocity):
        self.vel = np.array(velocity, dtype=np.float64)

    def setCollisionFamily(self, family):
        self.family = int(family)


class Solver:
    """Solver fields of the reference (Solver.cuh:23-26), bound to a System."""

    def __init__(self, system: "System"):
        self._sys = system

    @property
    def tolerance(self):
        return self._sys._get(PARAM_TOL)

    @tolerance.setter
    def tolerance(self, v):
        self._sys._set(PARAM_TOL, v)

    @property
    def maxIterations(self):
        return int(self._sys._get(PARAM_MAX_ITER))

    @maxIterations.setter
    def maxIterations(self, v):
        self._sys._set(PARAM_MAX_ITER, v)

    @property
    def iterations(self):
        return self._sys.stats().iterations

    @property
    def verbose(self):
        return bool(self._sys._get(PARAM_VERBOSE))

    @verbose.setter
    def verbose(self, v):
        self._sys._set(PARAM_VERBOSE, 1.0 if v else 0.0)

    def setWarmStarting(self, flag):  # APGD::setWarmStarting (APGD.cuh:62)
        self._sys._set(PARAM_WARM_START, 1.0 if flag else 0.0)

    def setAntiRelaxation(self, flag):  # APGD::setAntiRelaxation (APGD.cuh:63)
        self._sys._set(PARAM_ANTI_RELAX, 1.0 if flag else 0.0)


class System:
    """Drop-in mirror of the reference's System (System.cuh:279-326)."""

    def __init__(self, solverType: int = SOLVER_APGD, build: int = BUILD_FP64):
        L = load_library()
        self._L = L
        h = ctypes.c_void_p()
        self._h = None
        self._check(L.dv_scene_create(ctypes.byref(h), int(solverType)), None)
        self._h = h
        self.bodies: List[Body] = []
        self.solver = Solver(self)
        self._initialized = False
        if build != BUILD_FP64:
            self.setBuild(build)

    # -- plumbing
    def _check(self, rc, h="self"):
        if rc != DV_OK:
            msg = self._L.dv_last_error(self._h if h == "self" else None)
            raise DvideError(rc, msg.decode() if msg else "")

    def _set(self, p, v):
        self._check(self._L.dv_set_param(self._h, p, float(v)))

    def _get(self, p):
        out = ctypes.c_double()
        self._check(self._L.dv_get_param(self._h, p, ctypes.byref(out)))
        return out.value

    def close(self):
        if getattr(self, "_h", None):
            self._L.dv_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference surface
    @property
    def h(self):
        return self._get(PARAM_H)

    @property
    def time(self):
        return self._get(PARAM_TIME)

    def setTimeStep(self, step_size):
        self._set(PARAM_H, step_size)

    def setFrictionCoefficient(self, mu):
        self._set(PARAM_MU, mu)

    def setGravity(self, g):
        for i, v in enumerate(g):
            self._set(PARAM_GX + i, v)

    def setEnvelope(self, env):  # CollisionDetector::setEnvelope
        self._set(PARAM_ENVELOPE, env)

    def setGeometryMode(self, mode):
        self._set(PARAM_GEOM_MODE, mode)

    def setBuild(self, build):
        """APGD build: BUILD_FP64, BUILD_MIXED or BUILD_PLAIN."""
        self._set(PARAM_APGD_BUILD, build)

    def setGrid(self, g):
        self._set(PARAM_GRID, g)

    def setRowsPerThread(self, r):
        """Contacts per thread of the solver's contact kernels (0 tuned, 1 or 2)."""
        self._set(PARAM_ROWS_PER_THREAD, r)

    def setProfileKernel(self, k):
        """Time one solver kernel per pass with HIP events (PROFILE_*)."""
        self._set(PARAM_PROFILE, k)

    def add(self, body: Body) -> int:
        self.bodies.append(body)
        return len(self.bodies)

    def add_bodies(self, pos, vel, mass, geom, fixed, family):
        """Bulk form of add(Body*) for generated scenes."""
        pos, vel, mass, geom = _f64(pos), _f64(vel), _f64(mass), _f64(geom)
        fixed = np.ascontiguousarray(fixed, np.int32)
        family = np.ascontiguousarray(family, np.int32)
        self._flush_pending()
        self._check(self._L.dv_add_bodies(
            self._h, int(mass.shape[0]), pos.ctypes.data_as(_dp), vel.ctypes.data_as(_dp),
            mass.ctypes.data_as(_dp), geom.ctypes.data_as(_dp), fixed.ctypes.data_as(_ip),
            family.ctypes.data_as(_ip)))
        self._bulk = getattr(self, "_bulk", 0) + int(mass.shape[0])

    def _flush_pending(self):
        pending = self.bodies[getattr(self, "_flushed", 0):]
        if not pending:
            return
        n = len(pending)
        pos = _f64([b.pos for b in pending])
        vel = _f64([b.vel for b in pending])
        mass = _f64([b.mass for b in pending])
        geom = _f64([b.geometry for b in pending])
        fixed = np.array([1 if b.fixed else 0 for b in pending], np.int32)
        fam = np.array([b.family for b in pending], np.int32)
        self._check(self._L.dv_add_bodies(self._h, n, pos.ctypes.data_as(_dp), vel.ctypes.data_as(_dp),
                                          mass.ctypes.data_as(_dp), geom.ctypes.data_as(_dp),
                                          fixed.ctypes.data_as(_ip), fam.ctypes.data_as(_ip)))
        self._flushed = len(self.bodies)

    def addBilateralConstraintDOF(self, dofA, dofB, velocity=0.0, startTime=0.0):
        self._check(self._L.dv_add_bilateral_dof(self._h, int(dofA), int(dofB), float(velocity),
                                                 float(startTime)))

    def importMesh(self, filename, stiffness, numContactPointsPerElement, pressure=0.0):
        """System::importMesh (System.cu:1932-2060): a fixed ANCF shell mesh
        from the reference's text format (a mesh with inverse mass raises
        DvideError with DV_ERR_UNSUPPORTED)."""
        self._flush_pending()
        self._check(self._L.dv_import_mesh(self._h, os.fsencode(filename), float(stiffness),
                                           int(numContactPointsPerElement), float(pressure)))

    def add_shell_mesh(self, mesh):
        """A fixed shell mesh from a scenes.ShellMesh (the arrays importMesh
        fills), without a file."""
        if len(mesh.inv_mass[2]) and np.any(np.asarray(mesh.inv_mass[2]) != 0):
            raise DvideError(DV_ERR_UNSUPPORTED, "only fixed shells (no inverse mass) are supported")
        self._flush_pending()
        nodes = _f64(mesh.nodes)
        conn = np.ascontiguousarray(mesh.conn, np.int32)
        geom = _f64(mesh.geom)
        fext = _f64(mesh.fext)
        self._check(self._L.dv_add_shell_mesh(self._h, mesh.n_nodes, nodes.ctypes.data_as(_dp), mesh.n_elements,
                                              conn.ctypes.data_as(_ip), geom.ctypes.data_as(_dp),
                                              int(mesh.points), fext.ctypes.data_as(_dp)))

    @property
    def numShellNodes(self):
        n = ctypes.c_int()
        self._check(self._L.dv_num_shell_nodes(self._h, ctypes.byref(n)))
        return n.value

    def initializeSystem(self):
        self._flush_pending()
        self._check(self._L.dv_initialize(self._h))
        self._initialized = True

    def DoTimeStep(self):
        self._check(self._L.dv_step(self._h))

    step = DoTimeStep

    def getCCPViolation(self):
        out = np.zeros(4)
        self._check(self._L.dv_ccp_violation(self._h, out.ctypes.data_as(_dp)))
        return out

    # -- data access
    def stats(self) -> DvStats:
        st = DvStats()
        self._check(self._L.dv_get_stats(self._h, ctypes.byref(st)))
        return st

    @property
    def numCollisions(self):
        n = ctypes.c_int()
        self._check(self._L.dv_num_contacts(self._h, ctypes.byref(n)))
        return n.value

    @property
    def numBodies(self):
        n = ctypes.c_int()
        self._check(self._L.dv_num_bodies(self._h, ctypes.byref(n)))
        return n.value

    def vector(self, which):
        ln = ctypes.c_size_t()
        self._check(self._L.dv_get_vector(self._h, which, None, 0, ctypes.byref(ln)))
        out = np.zeros(ln.value, np.float64)
        if ln.value:
            self._check(self._L.dv_get_vector(self._h, which, out.ctypes.data_as(_dp), ln.value, None))
        return out

    def set_vector(self, which, x):
        x = _f64(x)
        self._check(self._L.dv_set_vector(self._h, which, x.ctypes.data_as(_dp), x.size))

    @property
    def p(self):
        return self.vector(VEC_P)

    @property
    def v(self):
        return self.vector(VEC_V)

    @property
    def gamma(self):
        return self.vector(VEC_GAMMA)

    @property
    def f_contact(self):
        return self.vector(VEC_F_CONTACT)

    def contacts(self):
        nc = self.numCollisions
        a = np.zeros(nc, np.uint32)
        b = np.zeros(nc, np.uint32)
        nap = np.zeros((nc, 4), np.float64)
        loc = np.zeros((nc, 3), np.float64)
        self._check(self._L.dv_get_contacts(self._h, a.ctypes.data_as(_up), b.ctypes.data_as(_up),
                                            nap.ctypes.data_as(_dp), loc.ctypes.data_as(_dp), nc))
        return a, b, nap, loc

    def _bodies(self):
        self._flush_pending()
        n = self.numBodies
        mass = np.zeros(n)
        geom = np.zeros((n, 3))
        fixed = np.zeros(n, np.int32)
        family = np.zeros(n, np.int32)
        self._check(self._L.dv_get_bodies(self._h, mass.ctypes.data_as(_dp), geom.ctypes.data_as(_dp),
                                          fixed.ctypes.data_as(_ip), family.ctypes.data_as(_ip), n))
        return mass, geom, fixed, family

    def masses(self):
        return self._bodies()[0]

    def geometry(self):
        return self._bodies()[1]

    def fixed(self):
        return self._bodies()[2]

    def families(self):
        return self._bodies()[3]

    def inverse_mass(self):
        """Diagonal of System::mass (1/m, 0 for a fixed body), one per body."""
        mass, _, fixed, _ = self._bodies()
        return np.where(fixed != 0, 0.0, 1.0 / np.where(mass != 0, mass, 1.0))

    def jacobian(self):
        """System::D as row-ordered COO (I, J, V, rows, cols)."""
        nnz = ctypes.c_size_t()
        rows = ctypes.c_int()
        cols = ctypes.c_int()
        self._check(self._L.dv_get_jacobian(self._h, None, None, None, 0, ctypes.byref(nnz), ctypes.byref(rows),
                                            ctypes.byref(cols)))
        n = nnz.value
        I = np.zeros(n, np.int32)
        J = np.zeros(n, np.int32)
        V = np.zeros(n, np.float64)
        if n:
            self._check(self._L.dv_get_jacobian(self._h, I.ctypes.data_as(_ip), J.ctypes.data_as(_ip),
                                                V.ctypes.data_as(_dp), n, None, None, None))
        return I, J, V, rows.value, cols.value

    def contactForcePerCollision(self):
        """System::calculateContactForcePerCollision: (normal, friction), each (nc, 3)."""
        nc = self.numCollisions
        nrm = np.zeros((nc, 3))
        fr = np.zeros((nc, 3))
        self._check(self._L.dv_contact_forces(self._h, nrm.ctypes.data_as(_dp), fr.ctypes.data_as(_dp), nc))
        return nrm, fr

    def _energies(self):
        out = np.zeros(3)
        self._check(self._L.dv_energies(self._h, out.ctypes.data_as(_dp)))
        return out

    def getPotentialEnergy(self):
        return float(self._energies()[0])

    def getKineticEnergy(self):
        return float(self._energies()[1])

    def getStrainEnergy(self):
        return float(self._energies()[2])

    def getTotalEnergy(self):
        return float(self._energies().sum())

    def assemble(self) -> int:
        nc = ctypes.c_int()
        self._check(self._L.dv_assemble(self._h, ctypes.byref(nc)))
        return nc.value

    def solve(self):
        self._check(self._L.dv_solve(self._h))

    def schur_product(self, x):
        x = _f64(x)
        y = np.zeros_like(x)
        self._check(self._L.dv_schur_product(self._h, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp), x.size))
        return y

    def pass_product(self, x, build):
        """N x by the graph-driven solver's pass kernels (BUILD_FP64 or BUILD_MIXED)."""
        x = _f64(x)
        y = np.zeros_like(x)
        self._check(self._L.dv_pass_product(self._h, int(build), x.ctypes.data_as(_dp), y.ctypes.data_as(_dp),
                                            x.size))
        return y

    def schur_product_noalloc(self, x):
        """N x through the library's fixed device scratch (no allocation;
        at most 1,048,576 rows)."""
        x = _f64(x)
        y = np.zeros_like(x)
        self._check(self._L.dv_schur_product_noalloc(self._h, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp),
                                                     x.size))
        return y


def project_cone(g, mu):
    """Project n 3-vectors (n,3) onto the Coulomb cone on the GPU (APGD.cu:47-78)."""
    L = load_library()
    g = np.ascontiguousarray(g, np.float64).copy()
    h = ctypes.c_void_p()
    rc = L.dv_scene_create(ctypes.byref(h), SOLVER_APGD)
    if rc != DV_OK:
        raise DvideError(rc, "scene create")
    try:
        rc = L.dv_project_cone(h, g.ctypes.data_as(_dp), g.size // 3, float(mu))
        if rc != DV_OK:
            raise DvideError(rc, L.dv_last_error(h).decode())
    finally:
        L.dv_scene_destroy(h)
    return g


def from_scene(scene, build: int = BUILD_FP64, *, tol=None, max_iterations=None, h=None,
               mu=None, warm_start=None, anti_relax=False, geom_mode: int = 1,
               grid: Optional[int] = None) -> System:
    """Build and initialise an APGD System in the given build from a generated
    scene (dvide_amd.scenes)."""
    prm = dict(scene.params)
    for k, v in (("tol", tol), ("max_iterations", max_iterations), ("h", h), ("mu", mu),
                 ("warm_start", warm_start)):
        if v is not None:
            prm[k] = v
    s = System(SOLVER_APGD, build)
    s.setTimeStep(prm["h"])
    s.setFrictionCoefficient(prm["mu"])
    s.setGravity(prm["gravity"])
    s.setEnvelope(prm["envelope"])
    s.setGeometryMode(geom_mode)
    s.solver.tolerance = prm["tol"]
    s.solver.maxIterations = prm["max_iterations"]
    s.solver.setWarmStarting(prm["warm_start"])
    if anti_relax:
        s.solver.setAntiRelaxation(True)
    if grid:
        s.setGrid(grid)
    s.add_bodies(scene.pos, scene.vel, scene.mass, scene.geom, scene.fixed, scene.family)
    if getattr(scene, "shell", None) is not None:
        s.add_shell_mesh(scene.shell)
    for (a, b), (vel, t0) in zip(scene.bil_dofs, scene.bil_info):
        s.addBilateralConstraintDOF(int(a), int(b), float(vel), float(t0))
    s.initializeSystem()
    return s


class Batch:
    """Scenes stepped together: one APGD solve of all of them in the same
    graph launches (dv_batch_*).  Each member System keeps its own state and
    statistics; its result equals stepping it alone with the same solver."""

    def __init__(self, build: int = BUILD_MIXED):
        L = load_library()
        self._L = L
        h = ctypes.c_void_p()
        rc = L.dv_batch_create(ctypes.byref(h), int(build))
        if rc != DV_OK:
            raise DvideError(rc, "dv_batch_create")
        self._h = h
        self.systems: List[System] = []

    def _check(self, rc):
        if rc != DV_OK:
            msg = self._L.dv_batch_last_error(self._h)
            raise DvideError(rc, msg.decode() if msg else "")

    def add(self, system: "System"):
        self._check(self._L.dv_batch_add(self._h, system._h))
        self.systems.append(system)

    def step(self):
        self._check(self._L.dv_batch_step(self._h))

    def stats(self) -> DvBatchStats:
        st = DvBatchStats()
        self._check(self._L.dv_batch_get_stats(self._h, ctypes.byref(st)))
        return st

    def close(self):
        if getattr(self, "_h", None):
            self._L.dv_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous batching of time-step requests over resident scenes
    (dv_server_*): submit() queues one DoTimeStep of a System, poll() admits
    waiting requests by free slot and HBM budget, runs one chunk of solver
    passes and returns the tickets of completed requests."""

    def __init__(self, build: int = BUILD_MIXED, max_slots: int = 64, hbm_budget: float = 0.0):
        L = load_library()
        self._L = L
        h = ctypes.c_void_p()
        rc = L.dv_server_create(ctypes.byref(h), int(build), int(max_slots), float(hbm_budget))
        if rc != DV_OK:
            msg = L.dv_server_last_error(None)
            raise DvideError(rc, msg.decode() if msg else "")
        self._h = h
        self._buf = (ctypes.c_long * 1024)()

    def _check(self, rc):
        if rc != DV_OK:
            msg = self._L.dv_server_last_error(self._h)
            raise DvideError(rc, msg.decode() if msg else "")

    def submit(self, system: "System") -> int:
        t = ctypes.c_long()
        self._check(self._L.dv_server_submit(self._h, system._h, ctypes.byref(t)))
        return t.value

    def poll(self) -> List[int]:
        """Tickets completed by this poll.  A diverged request raises
        DvideError (DV_ERR_DIVERGED) after its slot is freed; the tickets the
        poll completed are then on the error (.done) and the diverged ones
        are returned by failed()."""
        n = ctypes.c_int()
        rc = self._L.dv_server_poll(self._h, self._buf, len(self._buf), ctypes.byref(n))
        done = [self._buf[i] for i in range(n.value)]
        if rc != DV_OK:
            msg = self._L.dv_server_last_error(self._h)
            err = DvideError(rc, msg.decode() if msg else "")
            err.done = done
            raise err
        return done

    def failed(self) -> List[int]:
        """Tickets of diverged requests not yet reported."""
        n = ctypes.c_int()
        self._check(self._L.dv_server_failed(self._h, self._buf, len(self._buf), ctypes.byref(n)))
        return [self._buf[i] for i in range(n.value)]

    def pending(self) -> int:
        n = ctypes.c_int()
        self._check(self._L.dv_server_pending(self._h, ctypes.byref(n)))
        return n.value

    def drain(self) -> List[int]:
        done = []
        while self.pending():
            done += self.poll()
        return done

    def close(self):
        if getattr(self, "_h", None):
            self._L.dv_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
