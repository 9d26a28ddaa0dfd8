"""Pins the CPU oracle (oracle/liboracle.so) without the reference.

The reference has no golden vectors or known-answer tests for this path and
cannot be built here (CUDA + CUSP).  Each check below is chosen so that a
plausible mistake in the restatement fails it: closed-form geometry, brute
force over all pairs, the variational characterisation of the cone
projection, dense products, exact momentum identities and known mechanical
answers.  No check here restates the oracle's own formula.
"""
import itertools

import numpy as np
import pytest

import oracle
from dvide_amd import scenes

from helpers import pair_keys, tiny_scene


# --- narrowphase: closed forms ------------------------------------------------

@pytest.mark.parametrize("geom_mode", [0, 1])
def test_sphere_sphere_closed_form(geom_mode):
    # 3-4-5 triangle: centre distance 0.5 exactly, radii 0.25 + 0.3
    sc = tiny_scene([[0, 0, 0], [0.3, 0.4, 0]], [[0.25, 0, 0], [0.3, 0, 0]])
    o = oracle.OracleSystem(sc, geom_mode=geom_mode)
    kept, n, gap = o.narrow_pair(0, 1)
    assert kept
    assert gap == pytest.approx(0.5 - 0.55, abs=1e-7)
    # normal points from B (body 1) to A (body 0)
    np.testing.assert_allclose(n, [-0.6, -0.8, 0.0], atol=1e-7)


def test_sphere_sphere_envelope_keep_rule():
    env = 1e-3
    for d, keep in ((0.4 + 0.9 * env, True), (0.4 + 1.1 * env, False), (0.35, True)):
        sc = tiny_scene([[0, 0, 0], [d, 0, 0]], [[0.2, 0, 0], [0.2, 0, 0]], envelope=env)
        o = oracle.OracleSystem(sc, geom_mode=0)
        kept, n, gap = o.narrow_pair(0, 1)
        assert kept == keep, d
        if keep:
            assert gap == pytest.approx(d - 0.4, abs=1e-12)
            np.testing.assert_allclose(n, [-1, 0, 0], atol=1e-15)


@pytest.mark.parametrize("centre,expect_n,expect_gap", [
    ([0.0, 1.3, 0.0], [0, 1, 0], 0.3 - 0.3),          # above the top face, touching
    ([0.0, 1.25, 0.0], [0, 1, 0], 0.25 - 0.3),        # penetrating the top face
    ([2.2, 0.5, 0.0], [1, 0, 0], 0.2 - 0.3),          # beside the +x face
    ([2.15, 1.15, 0.0], [np.sqrt(0.5), np.sqrt(0.5), 0], np.hypot(0.15, 0.15) - 0.3),  # edge region
])
def test_sphere_box_closed_form(centre, expect_n, expect_gap):
    # box centred at (0,0.5,0) half extents (2,0.5,1); sphere radius 0.3
    sc = tiny_scene([[0, 0.5, 0], centre], [[2, 0.5, 1], [0.3, 0, 0]], fixed=[1, 0])
    o = oracle.OracleSystem(sc, geom_mode=0)
    kept, n, gap = o.narrow_pair(0, 1)
    assert kept
    assert gap == pytest.approx(expect_gap, abs=1e-12)
    # idA = box (0), idB = sphere (1): normal from the sphere to the box
    np.testing.assert_allclose(n, -np.asarray(expect_n, float), atol=1e-12)


def test_geometry_mode_one_rounds_through_binary32():
    # mode 1 lengths are binary32-rounded: the gap of an irrational distance
    # differs from the FP64 value by about one binary32 ulp
    sc = tiny_scene([[0, 0, 0], [0.1, 0.2, 0.3]], [[0.2, 0, 0], [0.2, 0, 0]])
    g0 = oracle.OracleSystem(sc, geom_mode=0).narrow_pair(0, 1)[2]
    g1 = oracle.OracleSystem(sc, geom_mode=1).narrow_pair(0, 1)[2]
    d = np.sqrt(0.14)
    assert g0 == pytest.approx(d - 0.4, abs=1e-15)
    assert abs(g1 - g0) < 2 * np.spacing(np.float32(d)) and g1 == float(np.float32(d)) - 0.4


# --- broadphase + narrowphase against brute force ----------------------------

def brute_pairs(o, scene):
    """All pairs the narrowphase keeps, minus pairs of one collision family
    other than -1 (CollisionDetector.cu:132,164)."""
    keys = []
    fam = scene.family
    for a, b in itertools.combinations(range(scene.n), 2):
        if fam[a] == fam[b] and fam[a] != -1:
            continue
        kept, _, _ = o.narrow_pair(a, b)
        if kept:
            keys.append((a << 32) | b)
    return np.array(sorted(keys), np.uint64)


@pytest.mark.parametrize("seed,bins", [(1, 20), (2, 7), (3, 1), (4, 33)])
def test_contact_set_equals_brute_force(seed, bins):
    sc = scenes.cloud(seed, n=300, edge=1.0)
    o = oracle.OracleSystem(sc, bins=bins)
    nc = o.collide()
    a, b, nap, _ = o.contacts()
    assert nc == len(a) > 50
    got = pair_keys(a, b)
    assert np.all(np.diff(got.astype(np.int64)) > 0)  # canonical, no duplicates
    np.testing.assert_array_equal(got, brute_pairs(o, sc))
    for i in range(0, nc, 17):  # every record carries its pair's narrowphase result
        _, n, gap = o.narrow_pair(int(a[i]), int(b[i]))
        np.testing.assert_array_equal(nap[i, :3], n)
        assert nap[i, 3] == gap


def test_collision_families_respected():
    sc = scenes.cloud(5, n=300, edge=1.0)
    o = oracle.OracleSystem(sc)
    o.collide()
    a, b, _, _ = o.contacts()
    fam = sc.family
    assert (fam != -1).sum() > 20
    # pairs within one family other than -1 are skipped (CollisionDetector.cu:132,164)
    same = (fam[a] == fam[b]) & (fam[a] != -1)
    assert not np.any(same)
    assert np.any((fam[a] != -1) | (fam[b] != -1))


# --- contact frame -------------------------------------------------------------

@pytest.mark.parametrize("geom_mode", [0, 1])
def test_frame_characterisation(geom_mode):
    oracle.lib().or_set_geom_mode(geom_mode)  # frame() uses the oracle's current mode
    rng = np.random.default_rng(0)
    ns = [rng.normal(size=3) for _ in range(200)]
    ns += [np.array(v, float) for v in ([1, 0, 0], [0, 1, 0], [0, 0, 1], [0, -1, 0], [1, 1, 0], [0, 1, 1])]
    for n in ns:
        n = n / np.linalg.norm(n)
        u, w = oracle.frame(n)
        tol = 1e-6 if geom_mode == 1 else 1e-12
        # orthonormal and right-handed: w = n x u
        assert abs(np.dot(u, n)) < tol and abs(np.linalg.norm(u) - 1) < tol
        np.testing.assert_allclose(w, np.cross(n, u), atol=tol * 10)
        # the plane u lies in and its sign fix u uniquely (System.cu:753-762)
        if n[2] != 0:
            assert u[1] == 0 and u[0] > 0
        elif n[0] != 0:
            assert u[1] == 0 and u[2] > 0
        else:
            assert u[2] == 0 and u[0] > 0


# --- Coulomb cone projection ----------------------------------------------------

def in_cone(g, mu, tol=1e-12):
    return np.hypot(g[1], g[2]) <= mu * g[0] + tol


def test_projection_variational_inequality():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        mu = rng.choice([0.1, 0.25, 0.5, 1.0, 2.0])
        x = rng.normal(size=3) * rng.choice([1e-3, 1.0, 1e3])
        p = oracle.project3(x, mu)
        assert in_cone(p, mu, 1e-9 * max(1, np.abs(x).max()))
        # <x - p, y - p> <= 0 for every y in the cone (p is the nearest point)
        for _ in range(8):
            yn = abs(rng.normal()) * np.abs(x).max()
            t = rng.normal(size=2)
            t *= mu * yn * rng.uniform() / max(np.linalg.norm(t), 1e-300)
            y = np.array([yn, t[0], t[1]])
            assert np.dot(x - p, y - p) <= 1e-9 * max(1.0, np.dot(x, x))
        np.testing.assert_allclose(oracle.project3(p, mu), p, rtol=1e-12, atol=1e-300)  # idempotent


def test_projection_nearest_point_brute_force():
    rng = np.random.default_rng(2)
    mu = 0.3
    for _ in range(20):
        x = rng.normal(size=3)
        p = oracle.project3(x, mu)
        # dense search over the cone boundary and apex
        gn = np.linspace(0, 3, 601)[:, None]
        ang = np.linspace(0, 2 * np.pi, 721)[None, :]
        cand = np.stack([np.broadcast_to(gn, (601, 721)), mu * gn * np.cos(ang), mu * gn * np.sin(ang)], -1)
        best = np.min(np.linalg.norm(cand - x, axis=-1))
        if in_cone(x, mu):
            best = 0.0
        assert np.linalg.norm(p - x) <= best + 1e-6


def test_projection_special_cases():
    # mu = 0: normal component clamped at zero, tangential dropped (APGD.cu:51-57)
    np.testing.assert_array_equal(oracle.project3([0.7, 0.3, -0.2], 0.0), [0.7, 0, 0])
    np.testing.assert_array_equal(oracle.project3([-0.7, 0.3, -0.2], 0.0), [0, 0, 0])
    # inside the cone: unchanged; inside the polar cone: zero
    np.testing.assert_array_equal(oracle.project3([1.0, 0.1, 0.1], 0.5), [1.0, 0.1, 0.1])
    np.testing.assert_array_equal(oracle.project3([-1.0, 0.1, 0.1], 0.5), [0, 0, 0])
    # |gamma_n| < 1e-14 with a tangential part outside the cone maps to zero
    # (the reference's guard, APGD.cu:66), although the exact projection does not
    np.testing.assert_array_equal(oracle.project3([1e-15, 1.0, 0.0], 0.5), [0, 0, 0])
    # by hand: with mu = 1 the cone's boundary in the (n, u) plane is the line
    # u = n; the nearest point of (0.2, 1) on it is (0.6, 0.6)
    np.testing.assert_allclose(oracle.project3([0.2, 1.0, 0.0], 1.0), [0.6, 0.6, 0.0], rtol=1e-15)


# --- Jacobian and Schur product ---------------------------------------------------

def dense_D(o):
    I, J, V, rows, cols = o.D()
    D = np.zeros((rows, cols))
    np.add.at(D, (I, J), V)
    return D


def test_jacobian_blocks_from_contacts():
    sc = scenes.cloud(7, n=60, edge=0.5)
    o = oracle.OracleSystem(sc)
    nc = o.assemble()
    assert nc > 10
    D = dense_D(o)
    a, b, nap, _ = o.contacts()
    for c in range(nc):
        F = D[3 * c:3 * c + 3, 3 * a[c]:3 * a[c] + 3]
        np.testing.assert_array_equal(F[0], nap[c, :3])  # row n is the stored normal
        u, w = oracle.frame(nap[c, :3])
        np.testing.assert_array_equal(F[1], u)
        np.testing.assert_array_equal(F[2], w)
        np.testing.assert_array_equal(D[3 * c:3 * c + 3, 3 * b[c]:3 * b[c] + 3], -F)  # D_B = -D_A
        row = D[3 * c:3 * c + 3].copy()
        row[:, 3 * a[c]:3 * a[c] + 3] = 0
        row[:, 3 * b[c]:3 * b[c] + 3] = 0
        assert not row.any()


def test_schur_product_dense():
    sc = scenes.cloud(8, n=40, edge=0.4)
    o = oracle.OracleSystem(sc)
    nc = o.assemble()
    assert 5 < nc <= 60
    D = dense_D(o)
    minv = np.where(np.repeat(sc.fixed, 3) != 0, 0.0, 1.0 / np.repeat(sc.mass, 3))
    N = D @ np.diag(minv) @ D.T
    rng = np.random.default_rng(3)
    for _ in range(3):
        x = rng.normal(size=3 * nc)
        np.testing.assert_allclose(o.schur(x), N @ x, rtol=1e-12, atol=1e-12 * np.abs(N @ x).max())
    assert np.allclose(N, N.T) and np.linalg.eigvalsh(N).min() > -1e-10


# --- time step: known answers -----------------------------------------------------

def test_free_fall_semi_implicit_euler():
    sc = scenes.free_fall(3, n=6)
    o = oracle.OracleSystem(sc)
    p0, v0 = o.vec(0).copy(), o.vec(1).copy()
    h = sc.params["h"]
    g = np.tile(np.asarray(sc.params["gravity"]), sc.n)
    p, v = p0, v0
    for _ in range(5):
        o.step()
        assert o.num_contacts == 0
        v = v + h * g
        p = p + h * v
    np.testing.assert_allclose(o.vec(1), v, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(o.vec(0), p, rtol=1e-14, atol=1e-14)


def test_ball_on_box_normal_impulse():
    sc = scenes.ball_on_box(0.0, 1e-2)
    o = oracle.OracleSystem(sc)
    o.step()
    assert o.num_contacts == 1 and o.residual < 1e-8
    gam = o.vec(2)
    # m g h on the normal row, nothing tangential; f_contact = m g upward
    np.testing.assert_allclose(gam, [9.81e-2, 0, 0], atol=1e-9)
    np.testing.assert_allclose(o.vec(3)[3:], [0, 9.81, 0], atol=1e-6)
    np.testing.assert_allclose(o.vec(1)[3:], [0, 0, 0], atol=1e-7)  # tolerance-limited


@pytest.mark.parametrize("gx,sticks", [(0.2 * 9.81, True), (0.24 * 9.81, True), (0.26 * 9.81, False),
                                       (0.4 * 9.81, False)])
def test_ball_on_box_stick_slip(gx, sticks):
    sc = scenes.ball_on_box(gx, 1e-2)
    mu = sc.params["mu"]
    assert mu == 0.25
    o = oracle.OracleSystem(sc)
    o.step()
    vx, vy = o.vec(1)[3], o.vec(1)[4]
    h = sc.params["h"]
    if sticks:
        assert abs(vx) < 1e-7 and abs(vy) < 1e-7
    else:
        # the cone complementarity relaxation of Coulomb friction: a sliding
        # contact separates at v_n = mu |v_t|, so gamma_n = m (g h + mu v_x),
        # friction mu gamma_n, and v_x (1 + mu^2) = h (g_x - mu g)
        assert vx == pytest.approx(h * (gx - mu * 9.81) / (1 + mu * mu), rel=1e-5)
        assert vy == pytest.approx(mu * vx, rel=1e-4)


def test_momentum_identities_in_a_pile():
    sc = scenes.pile(seed=5, target_contacts=600)
    o = oracle.OracleSystem(sc, tol=1e-6)
    v0 = o.vec(1).copy()
    o.step()
    fc = o.vec(3).reshape(-1, 3)
    # contact forces are equal and opposite in pairs: they cancel over all bodies
    assert np.abs(fc.sum(0)).max() < 1e-9 * np.abs(fc).max()
    # free bodies: m (v1 - v0) = h (m g + f_contact)
    free = sc.fixed == 0
    m = sc.mass[free][:, None]
    dv = (o.vec(1) - v0).reshape(-1, 3)[free]
    h = sc.params["h"]
    rhs = h * (m * np.asarray(sc.params["gravity"]) + fc[free])
    np.testing.assert_allclose(m * dv, rhs, rtol=1e-9, atol=1e-12 * np.abs(rhs).max())


def test_settled_pile_carries_its_weight():
    sc = scenes.pile(seed=6, target_contacts=300)
    o = oracle.OracleSystem(sc, tol=1e-8, max_iterations=20000)
    for _ in range(60):
        o.step()
    fc = o.vec(3).reshape(-1, 3)
    weight = 9.81 * sc.mass[sc.fixed == 0].sum()
    # the fixed container's reaction equals the weight once the pile is at rest
    assert fc[sc.fixed != 0, 1].sum() == pytest.approx(-weight, rel=2e-3)
    assert np.median(np.abs(o.vec(1))) < 1e-4


def test_apgd_optimality_at_convergence():
    sc = scenes.pile(seed=9, target_contacts=800)
    o = oracle.OracleSystem(sc, tol=1e-7, max_iterations=50000)
    o.step()
    assert o.residual < 1e-7
    g = o.vec(2).reshape(-1, 3)
    mu = sc.params["mu"]
    assert np.all(np.hypot(g[:, 1], g[:, 2]) <= mu * g[:, 0] + 1e-14)  # in the cone
    assert o.objective <= 0.0  # gamma = 0 is feasible with objective 0
    viol = o.ccp_violation()
    assert viol[0] > -1e-12                     # gamma_n >= 0
    assert abs(viol[2]) < 1e-3 * np.abs(g[:, 0]).sum()  # complementarity


def test_residual_zero_at_known_solution():
    sc = scenes.ball_on_box(0.0, 1e-2)
    o = oracle.OracleSystem(sc)
    o.assemble()
    assert o.residual_of(np.array([9.81e-2, 0, 0])) < 1e-9
    assert o.residual_of(np.array([0.2, 0, 0])) > 1e-3


@pytest.fixture
def one_thread():
    prev = oracle.set_threads(0)
    oracle.set_threads(1)
    yield
    oracle.set_threads(prev)


@pytest.mark.parametrize("kind", [oracle.PGS, oracle.PGJ])
def test_projected_solvers_reach_the_apgd_velocities(one_thread, kind):
    """PGS and PGJ (PGS.cu:149-447, PGJ.cu:151-210) solve the same cone
    complementarity problem by another route.  The velocities of its solution
    are unique (gamma is not), so converged runs of the two methods must give
    the same v: an independent check of APGD's solution."""
    sc = scenes.pile(5, 60)
    free = np.repeat(sc.fixed == 0, 3)
    a = oracle.OracleSystem(sc, tol=1e-8, max_iterations=100000)
    p = oracle.OracleSystem(sc, tol=1e-10, max_iterations=4000000, solver=kind)
    a.step()
    p.step()
    assert a.num_contacts == p.num_contacts > 20
    assert a.residual < 1e-8
    assert p.residual < 1e-10 and p.iterations < 4000000
    va, vp = a.vec(oracle.V)[free], p.vec(oracle.V)[free]
    assert np.abs(vp - va).max() < 1e-5 * np.abs(va).max()


def test_projected_solvers_refuse_bilateral_rows(one_thread):
    sc = scenes.compaction(2002, n=300, width=0.6, lid=scenes.LID_BILATERAL)
    p = oracle.OracleSystem(sc, tol=1e-6, max_iterations=10, solver=oracle.PGS)
    p.step()
    assert p.iterations == -1
    # the refused step goes on with no impulses (a gamma of one zero per row)
    g = p.vec(oracle.GAMMA)
    assert not g.any() and np.isfinite(p.vec(oracle.V)).all()
    with pytest.raises(ValueError):
        oracle.OracleSystem(sc, solver=2)
