"""The C ABI: libdvide_hip.so loads and exports every entry point that
include/dvide.h declares; the Python mirror binds exactly that set; without
a GPU the library refuses to create a scene (it has no CPU fallback)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from dvide_amd import scenes, system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dvide.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dv_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_surface():
    names = declared()
    for must in ("dv_scene_create", "dv_step", "dv_initialize", "dv_add_bodies", "dv_get_vector",
                 "dv_solve", "dv_schur_product", "dv_assemble", "dv_get_contacts"):
        assert must in names


def test_library_exports_every_declared_symbol():
    path = os.path.join(ROOT, "dvide_amd", "libdvide_hip.so")
    assert os.path.exists(path), "libdvide_hip.so is not built"
    lib = ctypes.CDLL(path)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (dv_[a-z0-9_]+)$", out, flags=re.M))
    assert set(declared()) <= exported


def test_python_mirror_binds_the_header_set():
    assert sorted(system.EXPORTED) == declared()


@pytest.mark.skipif(system.gpu_visible(), reason="checks the behaviour on a host without a GPU")
def test_no_gpu_means_an_error_not_a_fallback():
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from dvide_amd import system\n"
        "try:\n"
        "    system.System()\n"
        "except system.DvideError as e:\n"
        "    print('ERR', e.code); sys.exit(0)\n"
        "sys.exit(3)\n" % ROOT
    )
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ERR -3" in r.stdout


def test_factory_rejects_solvers_off_this_path():
    """System(int solverType), System.cu:77-101: 2..7 are PDIP, TPAS, JKIP,
    PGJ, PGS and PJKIP; the library refuses them before it looks for a
    device, so the answer is the same with or without a GPU."""
    lib = system.load_library()
    for t in range(2, 8):
        h = ctypes.c_void_p()
        assert lib.dv_scene_create(ctypes.byref(h), t) == system.DV_ERR_UNSUPPORTED
        assert not h.value
        msg = lib.dv_last_error(None).decode()
        assert ["PDIP", "TPAS", "JKIP", "PGJ", "PGS", "PJKIP"][t - 2] in msg
    if not system.gpu_visible():
        for t in (0, 1, 8, -1):  # APGD (1) and the reference's default case
            h = ctypes.c_void_p()
            assert lib.dv_scene_create(ctypes.byref(h), t) == -3  # DV_ERR_HIP: no device, not refused


def test_scene_generators_are_deterministic():
    a = scenes.pile(3, 2000)
    b = scenes.pile(3, 2000)
    c = scenes.pile(4, 2000)
    assert np.array_equal(a.pos, b.pos) and not np.array_equal(a.pos, c.pos)


def test_deposited_spheres_do_not_overlap():
    sc = scenes.compaction(2002, n=3000, width=1.2)
    sph = sc.geom[:, 1] == 0
    p, r = sc.pos[sph], sc.geom[sph, 0]
    assert np.all(r >= 0.04) and np.all(r <= 0.06)
    # all-pairs check on a cell grid
    cell = np.floor(p / 0.12).astype(np.int64)
    key = (cell[:, 0] * 4096 + cell[:, 1]) * 4096 + cell[:, 2]
    order = np.argsort(key)
    worst = 0.0
    for i in range(len(p)):
        d = np.linalg.norm(p - p[i], axis=1) - (r + r[i])
        d[i] = np.inf
        worst = min(worst, d.min())
    assert worst > -1e-9
    assert order.size == len(p)


def test_compaction_lid_modes():
    free = scenes.compaction(2002, n=2000, width=1.0)
    bil = scenes.compaction(2002, n=2000, width=1.0, lid=scenes.LID_BILATERAL)
    sph = free.geom[:, 1] == 0
    assert len(free.bil_dofs) == 0 and free.params["warm_start"]
    lid = 5  # after the ground and four walls
    assert free.mass[lid] == pytest.approx(free.mass[sph].sum())
    assert free.vel[lid, 1] == -0.1
    assert len(bil.bil_dofs) == 3 and not bil.params["warm_start"]
    assert list(bil.bil_dofs[:, 0]) == [3 * lid, 3 * lid + 1, 3 * lid + 2]
    assert list(bil.bil_info[:, 0]) == [0.0, -0.1, 0.0]


DRIVER = os.path.join(ROOT, "tests", "cpp", "bin", "one_ball")


def test_cpp_driver_is_built_against_the_facade():
    assert os.path.exists(DRIVER), "the C++ driver is built by make -C dvide_amd"
    out = subprocess.run(["ldd", DRIVER], capture_output=True, text=True).stdout
    assert "libdvide_hip.so" in out


@pytest.mark.skipif(system.gpu_visible(), reason="checks the behaviour on a host without a GPU")
def test_cpp_driver_fails_loudly_without_a_gpu():
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "no HIP device" in r.stderr
