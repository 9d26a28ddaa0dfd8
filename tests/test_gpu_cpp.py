"""The C++ drop-in surface (include/dvide/System.hpp) driven like the
reference's models/test_oneBall.cu, in every APGD build: the step's known
answers (gamma_n = m g h, f_contact = m g), getCCPViolation of the converged
contact, System::exportMatrices against the contact's closed-form Jacobian,
System::importSystem's masses, the factory's refusal of PDIP, and a ball
resting on a contact point of a fixed shell read by System::importMesh."""
import os
import re
import subprocess

import numpy as np
import pytest

from dvide_amd import system

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "cpp", "bin",
                      "one_ball")


def read_mtx(path):
    with open(path) as fh:
        head = fh.readline()
        dims = [int(x) for x in fh.readline().split()]
        rows = [line.split() for line in fh if line.strip()]
    return head, dims, rows


@pytest.mark.parametrize("build", [system.BUILD_FP64, system.BUILD_MIXED, system.BUILD_PLAIN])
def test_cpp_one_ball(gpu, build, tmp_path):
    r = subprocess.run([DRIVER, str(build), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    vals = dict(re.findall(r"(\w+) ([-0-9.e+]+)", r.stdout))
    assert int(vals["collisions"]) == 1
    tol = 1e-6 if build != system.BUILD_MIXED else 1e-5
    assert abs(float(vals["gamma"]) - 9.81e-2) < tol * 9.81e-2 * 10
    assert abs(float(vals["f_contact_y"]) - 9.81) < tol * 9.81 * 10
    assert abs(float(vals["time"]) - 0.03) < 1e-12
    # getCCPViolation (System.cu:1605-1643) of a converged resting contact:
    # gamma_n > 0 and inside the cone, the normal velocity ~ 0
    assert float(vals["violation_x"]) == 0.0
    assert -1e-5 <= float(vals["violation_y"]) <= 0.0
    assert abs(float(vals["violation_z"])) < 1e-5
    assert float(vals["violation_w"]) == 0.0

    # exportMatrices: D of the ball (B) on the box (A), normal from B to A =
    # (0,-1,0), u = (1,0,0), v = n x u = (0,0,1); A's columns +, B's -
    head, dims, rows = read_mtx(tmp_path / "D.mtx")
    assert head.startswith("%%MatrixMarket matrix coordinate real general")
    assert dims == [3, 6, 18]
    D = np.zeros((3, 6))
    for i, j, v in rows:
        D[int(i) - 1, int(j) - 1] = float(v)
    blk = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], float)
    # the solved steps leave the ball a few 1e-12 m off the box's axis, which
    # tilts the normal by as much
    np.testing.assert_allclose(D, np.hstack([blk, -blk]), rtol=0, atol=1e-9)
    _, dims, rows = read_mtx(tmp_path / "Minv.mtx")
    assert dims == [6, 6, 6]
    np.testing.assert_array_equal([float(v) for _, _, v in rows], [0, 0, 0, 1, 1, 1])
    for name, n in (("r", 3), ("b", 3), ("k", 6)):
        h, dims, rows = read_mtx(tmp_path / f"{name}.mtx")
        assert h.startswith("%%MatrixMarket matrix array real general") and dims == [n, 1] and len(rows) == n
    # k = M v + h f for the ball at rest: f = m g, v ~ 0
    k = [float(x[0]) for x in read_mtx(tmp_path / "k.mtx")[2]]
    assert k[:3] == [0.0, 0.0, 0.0] and abs(k[3]) < 1e-9 and abs(k[5]) < 1e-9
    assert k[4] == pytest.approx(-1e-2 * 9.81, abs=1e-5)

    # importSystem: the ball's mass is 2600 * 4 * 3.14159 * r^3 / 3
    m = 2600 * 4.0 * 3.14159 * 0.5 ** 3.0 / 3.0
    assert int(vals["imported_bodies"]) == 2
    assert float(vals["imported_mass"]) == pytest.approx(m, rel=1e-15)
    assert int(vals["imported_collisions"]) == 1
    assert float(vals["imported_gamma"]) == pytest.approx(m * 9.81 * 1e-2, rel=10 * tol)

    # System(2) is PDIP in the reference: not on this path
    assert "factory_pdip -5" in r.stdout


def test_cpp_ball_on_an_imported_shell(gpu, tmp_path):
    """System::importMesh on the C++ surface: a ball of 1 kg resting on top of
    one contact point of a fixed flat mesh carries gamma_n = m g h, and the
    DOF vectors hold the ball's 3 DOFs and the mesh's 9 per node."""
    from dvide_amd import scenes
    from tests.helpers import shell_point_grid

    mesh = scenes.flat_shell(2, 2, 0.1, 0.1, 0.01, points=8)
    path = tmp_path / "mesh.txt"
    scenes.write_mesh(mesh, str(path))
    pts, _, _ = shell_point_grid(mesh)
    x, y, z = (float(c) for c in pts[17])
    r = 0.025
    args = [str(path), repr(x), repr(y + 0.005 + r), repr(z), repr(r)]
    for build in (system.BUILD_FP64, system.BUILD_PLAIN):
        out = subprocess.run([DRIVER, str(build), str(tmp_path)] + args, capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        vals = dict(re.findall(r"(\w+) ([-0-9.e+]+)", out.stdout))
        assert int(vals["shell_collisions"]) == 1
        assert int(vals["shell_dofs"]) == 3 + 9 * mesh.n_nodes
        # within the binary32 rounding of the touching distance (a few 1e-10 m / h)
        assert float(vals["shell_gamma"]) == pytest.approx(1.0 * 9.81 * 1e-3, rel=1e-3)
