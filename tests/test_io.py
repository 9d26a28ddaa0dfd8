"""On-disk formats (dvide_amd/io.py): the reference's exportSystem /
importSystem CSV, Matrix Market matrices, contact dumps.  The writers and
readers are host code and run here; the System-level round trips need the
GPU and are in test_gpu_io.py."""
import numpy as np
import scipy.io

from dvide_amd import io


def sample(n=7, seed=0):
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(n, 3))
    vel = rng.normal(size=(n, 3)) * 1e-3
    geom = np.zeros((n, 3))
    geom[:, 0] = rng.uniform(0.01, 0.1, n)
    geom[0] = [2.0, 0.5, 1.5]  # a box
    fixed = np.zeros(n, np.int32)
    fixed[0] = 1
    return pos, vel, geom, fixed


def test_system_csv_layout(tmp_path):
    pos, vel, geom, fixed = sample()
    f = tmp_path / "sys.csv"
    io.write_system(f, pos, vel, geom, fixed)
    lines = f.read_text().splitlines()
    assert lines[0] == "7, 0, 0, 0, "
    # box line: index, fixed, p, "1, 0, 0, 0", v, shape 2, half extents
    tok = [t.strip() for t in lines[1].split(",")]
    assert tok[0] == "0" and tok[1] == "1" and tok[5:9] == ["1", "0", "0", "0"]
    assert tok[12] == "2" and tok[13:16] == ["2", "0.5", "1.5"]
    # sphere line: shape 0 then the radius, C++ default formatting (%g)
    tok = [t.strip() for t in lines[2].split(",")]
    assert tok[12] == "0" and tok[13] == "%g" % geom[1, 0]
    assert tok[2] == "%g" % pos[1, 0]


def test_system_csv_round_trip_exact_at_17_digits(tmp_path):
    pos, vel, geom, fixed = sample(50, 3)
    f = tmp_path / "sys.csv"
    io.write_system(f, pos, vel, geom, fixed, precision=17)
    p2, v2, g2, fx2, m2 = io.read_system(f)
    assert np.array_equal(p2, pos) and np.array_equal(v2, vel) and np.array_equal(g2, geom)
    assert np.array_equal(fx2, fixed)
    # importSystem's masses: 2600 * 4 * 3.14159 * r^3 / 3 for spheres, 1 for boxes
    assert m2[0] == 1.0
    assert m2[1] == 2600 * 4.0 * 3.14159 * geom[1, 0] ** 3.0 / 3.0


def test_matrix_market_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    I = np.array([0, 0, 1, 2, 2, 2])
    J = np.array([0, 3, 1, 0, 2, 3])
    V = rng.normal(size=6)
    io.write_matrix_market_coo(tmp_path / "D.mtx", I, J, V, 3, 5)
    D = scipy.io.mmread(str(tmp_path / "D.mtx")).toarray()
    ref = np.zeros((3, 5))
    ref[I, J] = V
    assert np.array_equal(D, ref)
    x = rng.normal(size=9)
    io.write_matrix_market_array(tmp_path / "r.mtx", x)
    assert np.array_equal(scipy.io.mmread(str(tmp_path / "r.mtx")).ravel(), x)


def test_collision_and_force_dumps(tmp_path):
    pos, vel, geom, fixed = sample()
    a = np.array([0, 1], np.uint32)
    b = np.array([1, 2], np.uint32)
    nap = np.array([[0, 1, 0, -1e-3], [0.6, 0.8, 0, 2e-4]])
    io.write_collisions(tmp_path / "c.csv", a, b, nap, pos, vel, geom)
    rows = [[t.strip() for t in ln.split(",")][:-1] for ln in (tmp_path / "c.csv").read_text().splitlines()]
    assert len(rows) == 2 and len(rows[0]) == 7 + 2 * 9
    assert rows[1][:7] == ["1", "1", "2", "0.6", "0.8", "0", "0.0002"]
    loc = np.zeros((2, 3))
    io.write_contact_forces(tmp_path / "f.csv", a, b, loc, np.ones((2, 3)), np.zeros((2, 3)), 7)
    lines = (tmp_path / "f.csv").read_text().splitlines()
    assert lines[0] == "2, 7, 0, 0, 0, "
    assert lines[2].startswith("1, 1, 2, 0, 0, 0, 1, 1, 1, 0, 0, 0")
