"""Fixed ANCF shells on the GPU path (dv_import_mesh / dv_add_shell_mesh)
against the oracle, through the C ABI.

The shell's contact points are fixed spheres whose positions the device
computes from the nodal coordinates, so the plain FP64 build reproduces the
oracle bit for bit as it does for bodies: contacts, gamma, iterations, v and
p.  The contact force on the shell DOFs (the shell columns of D^T gamma / h)
is gathered point by point, a different order of the same sums: 1e-12.  D's
3x36 shell blocks are exported exactly as the oracle builds them.  The
graph-driven builds are held to the oracle's solution as on other scenes.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from dvide_amd import scenes, system
from tests.helpers import bits_equal_or_nan, both, pair_keys, rel, shell_point_grid, spheres_on_points

pytestmark = pytest.mark.gpu


def bed(seed=4):
    return scenes.shell_bed(seed, nx=4, ny=4, per_side=(4, 4), layers=2)


def _check_step(o, g, sc, exact):
    nb = sc.n
    a0, b0, n0, l0 = o.contacts()
    a1, b1, n1, l1 = g.contacts()
    assert np.array_equal(pair_keys(a0, b0), pair_keys(a1, b1))
    assert (b0 >= nb).any(), "no sphere/point contact"
    fo, fg = o.vec(oracle.F_CONTACT), g.f_contact
    assert fo.shape == fg.shape == (3 * nb + 9 * sc.shell.n_nodes,)
    if exact:
        assert bits_equal_or_nan(n0, n1) and bits_equal_or_nan(l0, l1)
        assert g.stats().iterations == o.iterations
        for which in (oracle.GAMMA, oracle.V, oracle.P):
            assert bits_equal_or_nan(o.vec(which), g.vector(which)), which
        assert bits_equal_or_nan(fo[:3 * nb], fg[:3 * nb])
        scale = np.abs(fo[3 * nb:]).max()
        assert np.abs(fo[3 * nb:] - fg[3 * nb:]).max() <= 1e-12 * scale
    else:
        assert np.array_equal(g.p[3 * nb:], sc.shell.nodes.ravel())
        assert not g.v[3 * nb:].any()


def test_shell_bed_plain_bit_exact(gpu):
    sc = bed()
    o, g = both(sc, system.BUILD_PLAIN, tol=1e-8)
    assert g.numShellNodes == sc.shell.n_nodes and g.numBodies == sc.n
    for k in range(3):
        o.step()
        g.step()
        _check_step(o, g, sc, exact=True)
    g.close()


@pytest.mark.parametrize("build", [system.BUILD_FP64, system.BUILD_MIXED])
def test_shell_bed_graph_builds_match_oracle(gpu, build):
    sc = bed(6)
    tol = 1e-6 if build == system.BUILD_FP64 else 1e-4
    g = system.from_scene(sc, build, tol=tol)
    g.step()
    o = oracle.OracleSystem(sc, tol=tol)
    o.step()
    _check_step(o, g, sc, exact=False)
    exact = oracle.OracleSystem(sc, tol=1e-9, max_iterations=100000)
    exact.step()
    nb = sc.n
    # as close to the exact solution as the oracle's at the same tolerance
    # (the mixed build, whose products are binary32: or within 2e-2)
    slack = 2e-2 if build == system.BUILD_MIXED else 0.0
    star = exact.vec(oracle.V)[:3 * nb]
    err = rel(g.v[:3 * nb], star)
    assert err < max(3 * rel(o.vec(oracle.V)[:3 * nb], star), 1e-6, slack), err
    # the shell's nodal forces follow gamma
    fs = exact.vec(oracle.F_CONTACT)[3 * nb:]
    err = rel(g.f_contact[3 * nb:], fs)
    assert err < max(3 * rel(o.vec(oracle.F_CONTACT)[3 * nb:], fs), 1e-6, slack), err
    g.close()


def test_jacobian_export_matches_oracle(gpu):
    sc = bed(7)
    o, g = both(sc, system.BUILD_PLAIN)
    o.assemble()
    g.assemble()
    I0, J0, V0, r0, c0 = o.D()
    I1, J1, V1, r1, c1 = g.jacobian()
    assert (r0, c0) == (r1, c1) and c0 == 3 * sc.n + 9 * sc.shell.n_nodes
    assert np.array_equal(I0, I1) and np.array_equal(J0, J1)
    assert bits_equal_or_nan(V0, V1)
    assert (np.bincount(I0) == 39).any()  # sphere/point rows: 3 + 36 entries
    g.close()


def test_sphere_on_a_point_carries_its_weight(gpu):
    sc, on = spheres_on_points()
    g = system.from_scene(sc, system.BUILD_FP64)
    g.step()
    a, b, nap, _ = g.contacts()
    nb = sc.n
    assert np.array_equal(a, np.arange(nb)) and np.array_equal(b - nb, on)
    gam = g.gamma.reshape(-1, 3)
    gap, nu, m, h = nap[:, 3], nap[:, 1], sc.mass, sc.params["h"]
    assert np.allclose(gam[:, 0], m * (9.81 * h - gap / (h * nu)) / nu, rtol=1e-6, atol=0)
    assert np.abs(gam[:, 1:]).max() < 1e-12
    fc = g.f_contact[3 * nb:]
    want = np.zeros_like(fc)
    pts, els, xe = shell_point_grid(sc.shell)
    for k, j in enumerate(on):
        S = oracle.shell_shape(xe[j][0], xe[j][1], 0.1, 0.1)
        for q, node in enumerate(sc.shell.conn[els[j]]):
            for s_ in range(3):
                want[9 * node + 3 * s_ + 1] += S[3 * q + s_] * (-nu[k] * gam[k, 0] / h)
    assert np.allclose(fc, want, rtol=1e-12, atol=1e-13)
    g.close()


def test_mesh_file_import_equals_arrays(gpu, tmp_path):
    sc = bed(8)
    path = tmp_path / "mesh.txt"
    scenes.write_mesh(sc.shell, str(path))
    a = system.from_scene(sc, system.BUILD_MIXED)
    sc_nomesh = scenes.Scene(pos=sc.pos, vel=sc.vel, mass=sc.mass, geom=sc.geom, fixed=sc.fixed, family=sc.family,
                             bil_dofs=sc.bil_dofs, bil_info=sc.bil_info, params=sc.params)
    b = system.System(system.SOLVER_APGD, system.BUILD_MIXED)
    prm = sc.params
    b.setTimeStep(prm["h"])
    b.setFrictionCoefficient(prm["mu"])
    b.setGravity(prm["gravity"])
    b.setEnvelope(prm["envelope"])
    b.solver.tolerance = prm["tol"]
    b.solver.maxIterations = prm["max_iterations"]
    b.solver.setWarmStarting(prm["warm_start"])
    b.add_bodies(sc_nomesh.pos, sc_nomesh.vel, sc_nomesh.mass, sc_nomesh.geom, sc_nomesh.fixed, sc_nomesh.family)
    b.importMesh(str(path), 2e8, sc.shell.points, 0.0)
    b.initializeSystem()
    for _ in range(2):
        a.step()
        b.step()
        for which in (system.VEC_GAMMA, system.VEC_V, system.VEC_P, system.VEC_F_CONTACT):
            assert bits_equal_or_nan(a.vector(which), b.vector(which)), which
    a.close()
    b.close()


def test_free_shell_meshes_and_shell_dof_rows_are_refused(gpu, tmp_path):
    sc = bed(9)
    mesh = sc.shell
    mesh.inv_mass = (np.array([0], np.int32), np.array([0], np.int32), np.array([0.5]))
    path = tmp_path / "free.txt"
    scenes.write_mesh(mesh, str(path))
    s = system.System(system.SOLVER_APGD, system.BUILD_FP64)
    s.add_bodies(sc.pos, sc.vel, sc.mass, sc.geom, sc.fixed, sc.family)
    with pytest.raises(system.DvideError) as e:
        s.importMesh(str(path), 2e8, 8, 0.0)
    assert e.value.code == system.DV_ERR_UNSUPPORTED
    s.close()
    # malformed files: a truncated one, and a header naming 2^30 nodes
    scenes.write_mesh(sc.shell, str(path))
    lines = path.read_text().splitlines()
    for bad in (lines[:len(lines) // 2], ["1073741824,1,0,0"] + lines[1:]):
        path.write_text("\n".join(bad) + "\n")
        s = system.System(system.SOLVER_APGD, system.BUILD_FP64)
        with pytest.raises(system.DvideError) as e:
            s.importMesh(str(path), 2e8, 8, 0.0)
        assert e.value.code == system.DV_ERR_ARG
        s.close()
    mesh.inv_mass = (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0))
    # more than 2^26 contact points (8,194 per side: 8,192^2 interior points per element)
    s = system.System(system.SOLVER_APGD, system.BUILD_FP64)
    wide = dataclasses.replace(mesh, points=8194)
    with pytest.raises(system.DvideError) as e:
        s.add_shell_mesh(wide)
    assert e.value.code == system.DV_ERR_ARG
    s.close()
    s = system.System(system.SOLVER_APGD, system.BUILD_FP64)
    s.add_bodies(sc.pos, sc.vel, sc.mass, sc.geom, sc.fixed, sc.family)
    s.add_shell_mesh(mesh)
    s.addBilateralConstraintDOF(3 * sc.n + 4, -1)  # a shell DOF
    with pytest.raises(system.DvideError) as e:
        s.initializeSystem()
    assert e.value.code == system.DV_ERR_UNSUPPORTED
    s.close()
    g = system.from_scene(sc, system.BUILD_FP64)
    p = g.p
    p[3 * sc.n] += 1e-3  # moving a fixed shell node
    with pytest.raises(system.DvideError) as e:
        g.set_vector(system.VEC_P, p)
    assert e.value.code == system.DV_ERR_UNSUPPORTED
    p[3 * sc.n] -= 1e-3
    p[0] += 1e-3  # a body may move
    g.set_vector(system.VEC_P, p)
    assert g.p[0] == p[0]
    with pytest.raises(system.DvideError):
        g.vector(system.VEC_K)
    g.close()


def test_batch_with_a_shell_scene_equals_alone(gpu):
    def members():
        return [system.from_scene(bed(11), system.BUILD_MIXED, tol=1e-5),
                system.from_scene(scenes.pile(12, 800, 0.03, 0.06, 0.2), system.BUILD_MIXED, tol=1e-5)]

    alone, together = members(), members()
    batch = system.Batch(system.BUILD_MIXED)
    for g in together:
        batch.add(g)
    for _ in range(2):
        for g in alone:
            g.step()
        batch.step()
        for x, y in zip(alone, together):
            for which in (system.VEC_GAMMA, system.VEC_V, system.VEC_P, system.VEC_F_CONTACT):
                assert bits_equal_or_nan(x.vector(which), y.vector(which)), which
    batch.close()


def test_full_size_shell_bed_plain_bit_exact(gpu):
    """The bench's shell workload at full size (19,200 spheres, 57,600 contact
    points): the landed pile (400 mixed steps on the GPU) stepped once by the
    plain build and by the oracle, contacts, gamma, v and p bit for bit."""
    import copy

    sc = scenes.shell_bed(3003, nx=20, ny=80, per_side=(24, 100), layers=8, tol=1e-4)
    g = system.from_scene(sc, system.BUILD_MIXED, tol=1e-4)
    for _ in range(400):
        g.step()
    nb = sc.n
    landed = copy.copy(sc)
    landed.pos = g.p[:3 * nb].reshape(-1, 3).copy()
    landed.vel = g.v[:3 * nb].reshape(-1, 3).copy()
    g.close()
    o, g = both(landed, system.BUILD_PLAIN, tol=1e-4)
    o.step()
    g.step()
    assert o.num_contacts > 15000
    _check_step(o, g, landed, exact=True)
    g.close()


def test_rigid_body_formats_refuse_a_shell_scene(gpu, tmp_path):
    """export_system and export_collisions write rigid bodies only, and
    export_matrices needs k, which a shell scene does not give."""
    from dvide_amd import io

    g = system.from_scene(bed(12), system.BUILD_MIXED)
    g.step()
    with pytest.raises(NotImplementedError):
        io.export_system(g, str(tmp_path / "system.csv"))
    with pytest.raises(NotImplementedError):
        io.export_collisions(g, str(tmp_path / "collisions.csv"))
    with pytest.raises(system.DvideError) as e:
        io.export_matrices(g, str(tmp_path / "mtx"))
    assert e.value.code == system.DV_ERR_UNSUPPORTED
    # the per-collision force output has no per-body layout and still works
    io.output_contact_force_per_collision(g, str(tmp_path / "forces.csv"))
    assert (tmp_path / "forces.csv").read_text().count("\n") == 1 + g.numCollisions  # a header line, then one per contact
    g.close()


@pytest.mark.parametrize("points", [2, 3])
def test_meshes_with_no_point_or_one_per_element(gpu, points):
    """importMesh's smallest sampling: 2 points per side puts no contact point
    on an element (the spheres fall freely), 3 puts one at its centre."""
    sc = scenes.shell_bed(11, nx=4, ny=4, per_side=(4, 4), layers=2, points=points)
    o, g = both(sc, system.BUILD_PLAIN, tol=1e-8)
    nb = sc.n
    for _ in range(2):
        o.step()
        g.step()
        a0, b0, _, _ = o.contacts()
        a1, b1, _, _ = g.contacts()
        assert np.array_equal(pair_keys(a0, b0), pair_keys(a1, b1))
        assert int((b1 >= nb).sum()) == (0 if points == 2 else 4)
        assert g.stats().iterations == o.iterations
        for which in (oracle.GAMMA, oracle.V, oracle.P):
            assert bits_equal_or_nan(o.vec(which), g.vector(which)), which
        fo, fg = o.vec(oracle.F_CONTACT), g.f_contact
        assert fg.shape == (3 * nb + 9 * sc.shell.n_nodes,)
        assert bits_equal_or_nan(fo[:3 * nb], fg[:3 * nb])
        scale = max(np.abs(fo[3 * nb:]).max(), 1e-300)
        assert np.abs(fo[3 * nb:] - fg[3 * nb:]).max() <= 1e-12 * scale
        if points == 2:
            assert not fg[3 * nb:].any()
    g.close()


def test_uneven_mesh_plain_bit_exact(gpu):
    """Nodes raised or lowered by up to 1 mm and gradients tilted, so that
    every term of a point's shape-function sum is nonzero: the device places
    the points exactly where the oracle does (contacts, gamma, v and p bit for
    bit over three steps)."""
    sc = bed(12)
    rng = np.random.default_rng(12)
    nodes = sc.shell.nodes.copy()
    nodes[:, 1] += 2e-3 * (rng.random(len(nodes)) - 0.5)
    nodes[:, 3:9] += 0.05 * (rng.random((len(nodes), 6)) - 0.5)
    sc.shell = dataclasses.replace(sc.shell, nodes=nodes)
    o, g = both(sc, system.BUILD_PLAIN, tol=1e-7)
    for _ in range(3):
        o.step()
        g.step()
        _check_step(o, g, sc, exact=True)
    g.close()
