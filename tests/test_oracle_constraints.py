"""Pins the oracle's bilateral DOF rows, getCCPViolation and anti-relaxation
by closed forms (the reference holds no known answers for them).

Bilateral rows (System.cu:622-638, 1526-1544): a row is +1 at DOF a and -1 at
DOF b, and b = (p_a - p_b - v (t - t0)) / h, with v = 0 before t0.  A single
unprojected row is solved exactly by APGD (to its tolerance), so the
velocity after a step is fixed by the row and momentum alone.

getCCPViolation (System.cu:1605-1643): (min(gamma_n, 0), min((D v + b)_n, 0),
gamma_n . (D v + b)_n, |min(mu gamma_n - |gamma_t|, 0)|_2), checked on
hand-set gamma and v whose values are worked out here from the contact's
normal.

Anti-relaxation (APGD.cu:104-114, 335-345) adds mu |(D v)_t| to the normal
row of r, which cancels the separation v_n = mu |v_t| that the cone
complementarity problem gives a sliding contact: a ball sliding on a box
then keeps v_y = 0 and slows by the full Coulomb force, v_x = h (g_x - mu g).
"""
import numpy as np
import pytest

import oracle

from helpers import tiny_scene

FAR = [[0.0, 5.0, 0.0], [10.0, 5.0, 0.0]]


@pytest.fixture(autouse=True)
def one_thread():
    """Scenes of one or two rows: OpenMP only adds overhead."""
    prev = oracle.set_threads(0)
    oracle.set_threads(1)
    yield
    oracle.set_threads(prev)


SPH = [[0.1, 0, 0], [0.1, 0, 0]]


def test_bilateral_rows_of_D():
    sc = tiny_scene(FAR, SPH, bilaterals=[(1, -1, 0.0, 0.0), (3, 4, 0.2, 0.0)])
    o = oracle.OracleSystem(sc)
    o.assemble()
    I, J, V, rows, cols = o.D()
    assert rows == 2 and cols == 6
    assert list(zip(I, J, V)) == [(0, 1, 1.0), (1, 3, 1.0), (1, 4, -1.0)]


@pytest.mark.parametrize("vel,t0", [(0.3, 0.0), (-0.2, 0.0), (0.5, 2.5e-3)])
def test_bilateral_dof_follows_its_velocity(vel, t0):
    """Row p_y - p0 driven at vel from t0 on, against gravity: before t0 the
    DOF is held (v_y = 0); the stabilisation uses the position at the start
    of the step, so the DOF's velocity reaches vel one step after t0 and its
    position lags the drive by one step: p_y(t_n + h) = p0 + vel (t_n - t0)."""
    h = 1e-3
    sc = tiny_scene(FAR[:1], SPH[:1], h=h, tol=1e-10, max_iterations=30000, vel=[[0.7, 0.0, -0.4]],
                    bilaterals=[(1, -1, vel, t0)])
    o = oracle.OracleSystem(sc)
    p0 = 5.0
    for n in range(6):
        t = n * h
        p_old = o.vec(oracle.P)[1]
        o.step()
        v, p = o.vec(oracle.V), o.vec(oracle.P)
        drive = 0.0 if t < t0 else vel
        # one unprojected row: APGD's residual is |N gamma + r| = |v_y + b|,
        # so v_y is -b to within the reported residual
        b = (p_old - p0 - drive * (t - t0)) / h
        assert abs(v[1] + b) <= o.residual * (1 + 1e-9) + 1e-13
        assert o.residual < 1e-8
        # the closed-form trajectory: p_y after the step is p0 + drive (t - t0)
        assert p[1] == pytest.approx(p0 + drive * (t - t0), abs=1e-10)
        if t < t0 or n == 0:
            assert abs(v[1]) < 1e-8  # held against gravity
        if n >= 1 and (t - h) >= t0:
            assert v[1] == pytest.approx(vel, abs=1e-8)
        # the other DOFs are free: x, z keep their velocity
        assert v[0] == 0.7 and v[2] == -0.4
        assert len(o.vec(oracle.GAMMA)) == 1


def test_bilateral_relative_row_conserves_momentum():
    """Row p_a,y - p_b,y of two free bodies (masses 2 and 3) driven at 0.4:
    the row's impulse is internal, so m1 v1 + m2 v2 changes by h (m1 + m2) g
    only, and v1 - v2 = -b = -(p1 - p2 - 0.4 t)/h."""
    h = 1e-3
    m1, m2 = 2.0, 3.0
    sc = tiny_scene(FAR, SPH, mass=[m1, m2], h=h, tol=1e-10, max_iterations=30000,
                    vel=[[0, 1.0, 0], [0, -0.5, 0]], bilaterals=[(1, 4, 0.4, 0.0)])
    o = oracle.OracleSystem(sc)
    for n in range(4):
        p_old, v_old = o.vec(oracle.P), o.vec(oracle.V)
        t = n * h
        o.step()
        v = o.vec(oracle.V)
        b = (p_old[1] - p_old[4] - 0.4 * t) / h
        mom_old = m1 * v_old[1] + m2 * v_old[4]
        # two equations: v1 - v2 = -b; m1 v1 + m2 v2 = mom_old - h (m1 + m2) 9.81
        A = np.array([[1.0, -1.0], [m1, m2]])
        rhs = np.array([-b, mom_old - h * (m1 + m2) * 9.81])
        v1, v2 = np.linalg.solve(A, rhs)
        # the row holds to APGD's reported residual |v1 - v2 + b|; momentum exactly
        assert abs(v[1] - v1) <= o.residual + 1e-12 and abs(v[4] - v2) <= o.residual + 1e-12
        assert m1 * v[1] + m2 * v[4] == pytest.approx(rhs[1], abs=1e-12)
        assert o.residual < 1e-8


def ball_on_box(vel_ball=(0.0, 0.0, 0.0), gx=0.0, h=1e-2, mu=0.25, tol=1e-10, bilaterals=()):
    """Box (fixed, body 0) under a unit-mass ball (body 1) touching its top."""
    return tiny_scene([[0, -0.5, 0], [0, 0.5, 0]], [[5, 0.5, 5], [0.5, 0, 0]], fixed=[1, 0],
                      vel=[[0, 0, 0], list(vel_ball)], gravity=(gx, -9.81, 0.0), h=h, mu=mu, tol=tol,
                      max_iterations=100000, bilaterals=bilaterals)


def test_ccp_violation_of_a_converged_contact_is_zero():
    o = oracle.OracleSystem(ball_on_box())
    o.step()
    viol = o.ccp_violation()
    assert viol[0] == 0.0 and viol[3] == 0.0
    assert abs(viol[1]) < 1e-9 and abs(viol[2]) < 1e-9


def test_ccp_violation_known_values():
    """Hand-set gamma and v on the ball-on-box contact.  Its normal (from the
    ball B to the box A) is (0,-1,0), so (D v)_n = n.v_A - n.v_B = v_ball,y,
    and b_n = gap / h = 0 for a touching ball."""
    o = oracle.OracleSystem(ball_on_box(mu=0.25))
    o.assemble()
    assert o.num_contacts == 1
    _, _, nap, _ = o.contacts()
    np.testing.assert_allclose(nap[0, :3], [0, -1, 0], atol=1e-15)
    assert nap[0, 3] == pytest.approx(0.0, abs=1e-15)
    o.set_vec(oracle.GAMMA, [-0.5, 0.3, 0.4])
    o.set_vec(oracle.V, [0, 0, 0, 0.7, -0.2, 0.1])
    viol = o.ccp_violation()
    # min(gamma_n,0) = -0.5; min(v_n,0) = -0.2; gamma_n v_n = 0.1;
    # mu gamma_n - |gamma_t| = -0.125 - 0.5 = -0.625
    np.testing.assert_allclose(viol, [-0.5, -0.2, 0.1, 0.625], rtol=1e-15, atol=1e-16)
    # inside the cone and separating: every component vanishes
    o.set_vec(oracle.GAMMA, [2.0, 0.3, 0.4])
    o.set_vec(oracle.V, [0, 0, 0, 0.7, 0.2, 0.1])
    viol = o.ccp_violation()
    np.testing.assert_allclose(viol, [0.0, 0.0, 0.4, 0.0], rtol=1e-15, atol=1e-16)


def test_ccp_violation_sums_over_contacts():
    """Two balls on the box: the minimum, the complementarity sum and the
    2-norm of the cone violations over the two contacts."""
    sc = tiny_scene([[0, -0.5, 0], [-1, 0.5, 0], [1, 0.5, 0]], [[5, 0.5, 5], [0.5, 0, 0], [0.5, 0, 0]],
                    fixed=[1, 0, 0], h=1e-2, mu=0.5)
    o = oracle.OracleSystem(sc)
    o.assemble()
    assert o.num_contacts == 2
    o.set_vec(oracle.GAMMA, [1.0, 0.3, 0.4, -2.0, 0.0, 1.2])
    o.set_vec(oracle.V, [0, 0, 0, 0, -0.3, 0, 0, 0.6, 0])
    viol = o.ccp_violation()
    # cone: 0.5*1 - 0.5 = 0 (on the boundary), 0.5*(-2) - 1.2 = -2.2
    np.testing.assert_allclose(viol, [-2.0, -0.3, 1.0 * -0.3 + -2.0 * 0.6, 2.2], rtol=1e-15)


def test_ccp_violation_takes_contact_rows_after_the_bilateral_rows():
    """With one bilateral row first, the contact's gamma and (D v + b) rows
    are rows 1..3 of the flat vectors.  (The reference sizes its D v
    temporary for the contact rows only, System.cu:1616-1628, which does not
    match D's row count once bilateral rows exist; the oracle and the GPU
    path take the contact rows at their offset.)"""
    o = oracle.OracleSystem(ball_on_box(mu=0.25, bilaterals=[(0, -1, 0.0, 0.0)]))
    o.assemble()
    o.set_vec(oracle.GAMMA, [9.0, -0.5, 0.3, 0.4])
    o.set_vec(oracle.V, [0, 0, 0, 0.7, -0.2, 0.1])
    np.testing.assert_allclose(o.ccp_violation(), [-0.5, -0.2, 0.1, 0.625], rtol=1e-15, atol=1e-16)


@pytest.mark.parametrize("gx", [0.4 * 9.81, 0.6 * 9.81])
def test_anti_relaxation_removes_the_sliding_separation(gx):
    h, mu, g = 1e-2, 0.25, 9.81
    plain = oracle.OracleSystem(ball_on_box(gx=gx, h=h, mu=mu, tol=1e-12))
    plain.step()
    vx, vy = plain.vec(oracle.V)[3:5]
    assert vy == pytest.approx(mu * vx, rel=1e-6)  # without: v_n = mu |v_t|
    ar = oracle.OracleSystem(ball_on_box(gx=gx, h=h, mu=mu, tol=1e-12), anti_relax=True)
    ar.step()
    vx, vy, vz = ar.vec(oracle.V)[3:6]
    assert abs(vy) < 1e-9
    assert vx == pytest.approx(h * (gx - mu * g), rel=1e-7)
    assert vz == 0.0
    gam = ar.vec(oracle.GAMMA)
    # gamma_n = m g h, friction on the cone boundary opposing the slide
    assert gam[0] == pytest.approx(g * h, rel=1e-7)
    assert gam[1] == pytest.approx(mu * g * h, rel=1e-6)
    # r keeps the last anti-relaxation term: r_n = b_n + (D M^-1 k)_n + mu |(D v)_t|
    r = ar.vec(oracle.R)
    assert r[0] == pytest.approx(-g * h + mu * vx, rel=1e-6)


def test_anti_relaxation_leaves_a_sticking_contact_alone():
    gx = 0.1 * 9.81
    a = oracle.OracleSystem(ball_on_box(gx=gx, tol=1e-12))
    b = oracle.OracleSystem(ball_on_box(gx=gx, tol=1e-12), anti_relax=True)
    a.step()
    b.step()
    np.testing.assert_allclose(b.vec(oracle.V), a.vec(oracle.V), atol=1e-10)
    assert np.abs(b.vec(oracle.V)).max() < 1e-9
