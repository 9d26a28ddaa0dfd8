"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bit-exact: the contact set, its normals, gaps and frames, k, b and r, the
Schur product N x, the cone projection, and the whole plain FP64 APGD
trajectory (gamma, iterations, residual, v, p) -- the plain path performs
the oracle's operations in the oracle's order.
Tolerance-level: the persistent solvers, which reorder reductions and use
N y = (1+b) N g_k - b N g_{k-1}: FP64 gamma within 1e-6 relative and mixed
FP32/FP64 gamma within 1e-3 relative of the FP64 oracle, both at a residual
<= tol.
"""
import numpy as np
import pytest

import oracle
from dvide_amd import scenes, system
from tests.helpers import both, pair_keys, rel

pytestmark = pytest.mark.gpu


def bits_equal(x, y):
    """Bitwise equality, with NaN matching NaN (payloads may differ)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    if x.shape != y.shape:
        return False
    nx, ny = np.isnan(x), np.isnan(y)
    return np.array_equal(nx, ny) and np.array_equal(x[~nx].view(np.uint64), y[~ny].view(np.uint64))


def _same_contacts(o, g):
    a0, b0, n0, l0 = o.contacts()
    a1, b1, n1, l1 = g.contacts()
    assert a0.shape == a1.shape, (a0.shape, a1.shape)
    assert np.array_equal(pair_keys(a0, b0), pair_keys(a1, b1))
    assert bits_equal(n0, n1), "normals/gaps differ"
    assert bits_equal(l0, l1), "locations differ"
    return a0.size


def test_projection_bit_exact(gpu):
    rng = np.random.default_rng(7)
    g = rng.normal(size=(4096, 3))
    g[:8] = [[0, 0, 0], [1e-15, 1, 1], [-1, 0, 0], [-1, 0.1, 0], [2, 0.1, 0.1], [1, 3, 4], [0, 1, 0], [-5, 1, 1]]
    for mu in (0.0, 0.25, 0.5, 1.3):
        out = system.project_cone(g, mu)
        ref = np.array([oracle.project3(x, mu) for x in g])
        assert bits_equal(out, ref), mu


@pytest.mark.parametrize("mode", [1, 0])
def test_contacts_cloud_bit_exact(gpu, mode):
    sc = scenes.cloud(11, n=3000, edge=1.5)
    o, g = both(sc, geom_mode=mode)
    o.assemble()
    g.assemble()
    nc = _same_contacts(o, g)
    assert nc > 1000


def test_assembly_bit_exact_lattice(gpu):
    sc = scenes.lattice_box(1001)
    o, g = both(sc)
    # drop the lattice onto itself for a few steps to create contacts
    for _ in range(60):
        o.step()
        g.step()
    o.assemble()
    g.assemble()
    _same_contacts(o, g)
    for which_o, which_g in ((oracle.R, system.VEC_R), (oracle.B, system.VEC_B), (oracle.K, system.VEC_K)):
        a = o.vec(which_o)
        b = g.vector(which_g)
        assert bits_equal(a, b), which_g


def test_schur_product_bit_exact(gpu):
    sc = scenes.compaction(2002, n=4000, width=1.5, lid=scenes.LID_BILATERAL)
    o, g = both(sc)
    nc = o.assemble()
    assert g.assemble() == nc and nc > 1000
    rng = np.random.default_rng(3)
    x = rng.normal(size=o.vec(oracle.R).size)
    y0 = o.schur(x)
    y1 = g.schur_product(x)
    assert bits_equal(y0, y1)


def test_plain_trajectory_bit_exact(gpu):
    """200-step lattice drop: the plain FP64 path tracks the oracle bit for bit."""
    sc = scenes.lattice_box(1001)
    o, g = both(sc, solver=system.SOLVER_PLAIN)
    for k in range(200):
        o.step()
        g.step()
        st = g.stats()
        assert st.num_contacts == o.num_contacts, k
        assert st.iterations == o.iterations, k
        assert st.residual == o.residual, k
    for wo, wg in ((oracle.P, system.VEC_P), (oracle.V, system.VEC_V), (oracle.GAMMA, system.VEC_GAMMA),
                   (oracle.F_CONTACT, system.VEC_F_CONTACT)):
        assert bits_equal(o.vec(wo), g.vector(wg)), wg


@pytest.mark.parametrize("solver,tol,tol_gamma", [(system.SOLVER_FUSED, 1e-6, 1e-6),
                                                   (system.SOLVER_MIXED, 1e-4, 1e-3)])
def test_fused_solvers_match_oracle(gpu, solver, tol, tol_gamma):
    """Graph-driven solvers against the oracle's FP64 APGD on the same step."""
    sc = scenes.compaction(2002, n=4000, width=1.5, lid=scenes.LID_BILATERAL)
    o, g = both(sc, solver=solver, tol=tol, max_iterations=20000)
    o.step()
    g.step()
    st = g.stats()
    assert st.num_contacts == o.num_contacts
    assert st.residual < tol
    # the residual the solver reports is the reference's getResidual of the
    # gamma it returns, evaluated by the oracle in FP64
    res_o = o.residual_of(g.gamma)
    assert abs(res_o - st.residual) <= 1e-3 * tol + 1e-9 * abs(st.residual), (res_o, st.residual)
    go, gg = o.vec(oracle.GAMMA), g.gamma
    assert rel(gg, go) < tol_gamma, rel(gg, go)
    assert rel(g.v, o.vec(oracle.V)) < tol_gamma


def test_ball_on_box_gpu(gpu):
    for solver in (system.SOLVER_PLAIN, system.SOLVER_FUSED, system.SOLVER_MIXED):
        g = system.from_scene(scenes.ball_on_box(0.0), solver)
        g.DoTimeStep()
        gam = g.gamma
        assert abs(gam[0] - 0.0981) < 1e-6 * (1e3 if solver == system.SOLVER_MIXED else 1), (solver, gam)
        assert abs(g.f_contact[4] - 9.81) < 1e-4 * (1e2 if solver == system.SOLVER_MIXED else 1)


def test_free_fall_gpu(gpu):
    sc = scenes.free_fall(5, n=8)
    g = system.from_scene(sc, system.SOLVER_FUSED)
    p0, v0 = g.p.copy(), g.v.copy()
    g.DoTimeStep()
    assert g.numCollisions == 0
    h = sc.params["h"]
    v1 = v0.reshape(-1, 3) + h * np.array(sc.params["gravity"])
    assert np.allclose(g.v.reshape(-1, 3), v1, rtol=0, atol=1e-15)
    assert np.allclose(g.p, p0 + h * g.v, rtol=0, atol=1e-15)
