"""Known answers of the whole time step on the GPU solvers.

These are the closed forms the oracle's own tests pin (test_oracle.py),
applied to the product path through the C ABI:

- a ball resting on a box under lateral gravity g_x, mu = 0.25, sticks for
  |g_x| < mu g and slides above; sliding, the cone complementarity relaxation
  of Coulomb friction gives v_x (1 + mu^2) = h (g_x - mu g) and v_n = mu v_x;
- contact forces act in equal and opposite pairs, so they cancel over all
  bodies, and every free body's velocity changes by h (g + f_contact / m)
  (the update, System.cu:569-583);
- a settled pile rests on its container, whose reaction equals its weight.
"""
import numpy as np
import pytest

from dvide_amd import scenes, system

pytestmark = pytest.mark.gpu

SOLVERS = [system.BUILD_PLAIN, system.BUILD_FP64, system.BUILD_MIXED]
NAMES = ["plain", "fused", "mixed"]
# the mixed solver rounds D, M^-1 and the products to binary32 and is run to
# a residual of 1e-6 here; the FP64 solvers to the scene's 1e-8
TOL = {system.BUILD_PLAIN: 1e-8, system.BUILD_FP64: 1e-8, system.BUILD_MIXED: 1e-6}


@pytest.mark.parametrize("solver", SOLVERS, ids=NAMES)
@pytest.mark.parametrize("ratio,sticks", [(0.2, True), (0.24, True), (0.26, False), (0.4, False)])
def test_ball_on_box_stick_slip(gpu, solver, ratio, sticks):
    g = 9.81
    gx = ratio * g
    sc = scenes.ball_on_box(gx, 1e-2)
    mu, h = sc.params["mu"], sc.params["h"]
    assert mu == 0.25
    s = system.from_scene(sc, solver, tol=TOL[solver])
    s.DoTimeStep()
    assert s.numCollisions == 1
    vx, vy, vz = s.v[3:6]
    lim = 1e-7 if solver != system.BUILD_MIXED else 1e-5
    assert abs(vz) < 1e-12
    if sticks:
        assert abs(vx) < lim and abs(vy) < lim, (vx, vy)
    else:
        vx_exact = h * (gx - mu * g) / (1 + mu * mu)
        assert abs(vx - vx_exact) < lim + 1e-5 * abs(vx_exact), (vx, vx_exact)
        assert abs(vy - mu * vx) < lim + 1e-4 * abs(vx), (vy, mu * vx)
    s.close()


@pytest.mark.parametrize("solver", SOLVERS, ids=NAMES)
def test_momentum_identities_in_a_pile(gpu, solver):
    sc = scenes.pile(5, 600)
    s = system.from_scene(sc, solver, tol=1e-6 if solver != system.BUILD_MIXED else 1e-5,
                          max_iterations=20000)
    v0 = s.v.copy()
    s.DoTimeStep()
    assert s.numCollisions > 300
    fc = s.f_contact.reshape(-1, 3)
    assert np.abs(fc.sum(0)).max() < 1e-9 * np.abs(fc).max()
    free = sc.fixed == 0
    m = sc.mass[free][:, None]
    dv = (s.v - v0).reshape(-1, 3)[free]
    h = sc.params["h"]
    rhs = h * (m * np.asarray(sc.params["gravity"]) + fc[free])
    np.testing.assert_allclose(m * dv, rhs, rtol=1e-9, atol=1e-12 * np.abs(rhs).max())
    s.close()


@pytest.mark.parametrize("solver", SOLVERS[1:], ids=NAMES[1:])
def test_settled_pile_carries_its_weight(gpu, solver):
    sc = scenes.pile(6, 300)
    s = system.from_scene(sc, solver, tol=TOL[solver], max_iterations=20000)
    for _ in range(60):
        s.DoTimeStep()
    fc = s.f_contact.reshape(-1, 3)
    weight = 9.81 * sc.mass[sc.fixed == 0].sum()
    assert fc[sc.fixed != 0, 1].sum() == pytest.approx(-weight, rel=2e-3)
    assert np.median(np.abs(s.v)) < 1e-4
    s.close()
