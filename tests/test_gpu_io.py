"""On-disk formats and diagnostics through the C ABI on the GPU: exported
matrices and per-contact forces against the oracle, export/import round
trips, energies in closed form."""
import numpy as np
import pytest
import scipy.io

import oracle
from dvide_amd import io, scenes, system
from tests.helpers import both

pytestmark = pytest.mark.gpu


def test_exported_jacobian_and_vectors_equal_oracle(gpu, tmp_path):
    sc = scenes.cloud(31, n=400, edge=0.6)
    o, g = both(sc, solver=system.BUILD_PLAIN)
    o.assemble()
    # a sphere centre inside a box has no normal (the reference normalises a
    # zero vector the same way), so D has NaN rows; the solve then ends with
    # a non-finite residual, which the step reports instead of updating, and
    # the system it assembled is exported as it stands
    with pytest.raises(system.DvideError) as err:
        g.step()
    assert err.value.code == -6  # DV_ERR_DIVERGED
    io.export_matrices(g, str(tmp_path))
    D = scipy.io.mmread(str(tmp_path / "D.mtx")).tocsr()
    I, J, V, rows, cols = o.D()
    Do = np.zeros((rows, cols))
    np.add.at(Do, (I, J), V)
    assert D.shape == (rows, cols)
    # NaN rows: a sphere centre inside a box has no normal (the reference
    # normalises a zero vector the same way)
    assert np.array_equal(D.toarray(), Do, equal_nan=True)
    Gi, Gj, Gv, _, _ = g.jacobian()
    assert np.array_equal(Gi, I) and np.array_equal(Gj, J)  # same COO order
    assert np.array_equal(Gv, V, equal_nan=True)
    for name, which in (("r", oracle.R), ("b", oracle.B), ("k", oracle.K)):
        x = scipy.io.mmread(str(tmp_path / f"{name}.mtx")).ravel()
        assert np.array_equal(x, o.vec(which), equal_nan=True), name
    minv = scipy.io.mmread(str(tmp_path / "Minv.mtx")).diagonal()
    expect = np.where(np.repeat(sc.fixed, 3) != 0, 0.0, 1.0 / np.repeat(sc.mass, 3))
    assert np.array_equal(minv, expect)


def test_contact_forces_per_collision(gpu):
    sc = scenes.pile(33, 1500)
    o, g = both(sc, solver=system.BUILD_PLAIN)
    o.step()
    g.step()
    normal, friction = g.contactForcePerCollision()
    _, _, nap, _ = o.contacts()
    gam = o.vec(oracle.GAMMA).reshape(-1, 3)  # no bilateral rows in this scene
    h = sc.params["h"]
    expect_n = gam[:, :1] * nap[:, :3] / h
    assert np.allclose(normal, expect_n, rtol=1e-14, atol=1e-14)
    u = np.array([oracle.frame(n)[0] for n in nap[:, :3]])
    w = np.array([oracle.frame(n)[1] for n in nap[:, :3]])
    assert np.allclose(friction, (gam[:, 1:2] * u + gam[:, 2:3] * w) / h, rtol=1e-12, atol=1e-14)
    # summed per body they give the reference's f_contact of the free bodies
    fc = np.zeros((sc.n, 3))
    a, b, _, _ = g.contacts()
    np.add.at(fc, a, normal + friction)
    np.add.at(fc, b, -(normal + friction))
    free = sc.fixed == 0
    assert np.allclose(fc[free], g.f_contact.reshape(-1, 3)[free], rtol=1e-9, atol=1e-9)


def test_export_import_round_trip(gpu, tmp_path):
    sc = scenes.pile(34, 1200)
    g = system.from_scene(sc, system.BUILD_MIXED)
    g.step()
    f = tmp_path / "sys.csv"
    io.export_system(g, str(f), precision=17)
    h = system.System(system.BUILD_MIXED)
    assert io.import_system(h, str(f)) == sc.n
    h.initializeSystem()
    assert np.array_equal(h.p, g.p) and np.array_equal(h.v, g.v)
    assert np.array_equal(h.geometry(), g.geometry()) and np.array_equal(h.fixed(), g.fixed())
    # the reference's default precision writes 6 significant digits
    io.export_system(g, str(tmp_path / "sys6.csv"))
    p6 = io.read_system(str(tmp_path / "sys6.csv"))[0]
    assert np.allclose(p6.ravel(), g.p, rtol=1e-5, atol=1e-6)
    io.export_collisions(g, str(tmp_path / "contacts.csv"))
    io.output_contact_force_per_collision(g, str(tmp_path / "forces.csv"))
    assert len((tmp_path / "contacts.csv").read_text().splitlines()) == g.numCollisions
    assert len((tmp_path / "forces.csv").read_text().splitlines()) == g.numCollisions + 1


def test_energies(gpu):
    sc = scenes.free_fall(4, n=5)
    g = system.from_scene(sc, system.BUILD_FP64)
    g.step()
    p, v = g.p.reshape(-1, 3), g.v.reshape(-1, 3)
    grav = np.asarray(sc.params["gravity"])
    pe = -np.sum(sc.mass[:, None] * grav * p)
    ke = 0.5 * np.sum(sc.mass[:, None] * v * v)
    assert g.getPotentialEnergy() == pytest.approx(pe, rel=1e-12)
    assert g.getKineticEnergy() == pytest.approx(ke, rel=1e-12)
    assert g.getStrainEnergy() == 0.0
