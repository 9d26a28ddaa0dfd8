"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bit-exact: the contact set, its normals, gaps and frames, k, b and r, the
Schur product N x, the cone projection, and the whole plain FP64 APGD
trajectory (gamma, iterations, residual, v, p) -- the plain path performs
the oracle's operations in the oracle's order.
Tolerance-level: the graph-driven solvers, which reorder reductions and use
N y = (1+b) N g_k - b N g_{k-1}: the FP64 solver's gamma within 1e-6 of the
oracle's at tolerance 1e-6; the mixed solver's early iterates within 2e-5 of
the oracle's, and its solution at tolerance 1e-4 as close to the exact one
as the oracle's at that tolerance (gamma itself is not unique on piles).
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
    o, g = both(sc, solver=system.BUILD_PLAIN)
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


def test_fp64_solver_matches_oracle(gpu):
    """Graph-driven FP64 APGD against the oracle's on the same step: the
    products are summed in another order, so the trajectories agree to
    rounding and the solutions to well below the tolerance."""
    sc = scenes.compaction(2002, n=4000, width=1.5, lid=scenes.LID_BILATERAL)
    o, g = both(sc, solver=system.BUILD_FP64, tol=1e-6, max_iterations=20000)
    o.step()
    g.step()
    st = g.stats()
    assert st.num_contacts == o.num_contacts
    assert st.residual < 1e-6
    # the residual the solver reports is the reference's getResidual of the
    # gamma it returns, evaluated by the oracle in FP64
    res_o = o.residual_of(g.gamma)
    assert abs(res_o - st.residual) <= 1e-3 * 1e-6 + 1e-9 * abs(st.residual), (res_o, st.residual)
    assert rel(g.gamma, o.vec(oracle.GAMMA)) < 1e-6
    assert rel(g.v, o.vec(oracle.V)) < 1e-6


def _free(sc):
    return np.repeat(sc.fixed == 0, 3)


def test_mixed_solver_early_iterations_track_oracle(gpu):
    """The mixed solver performs the reference's APGD steps: after a fixed
    number of iterations its gamma differs from the FP64 oracle's only by
    the binary32 rounding of D, M^-1 and the products."""
    sc = scenes.compaction(2002, n=4000, width=1.5, lid=scenes.LID_BILATERAL)
    for m in (1, 5, 10):
        o, g = both(sc, solver=system.BUILD_MIXED, tol=0.0, max_iterations=m)
        o.step()
        g.step()
        assert g.stats().iterations == o.iterations == m
        assert rel(g.gamma, o.vec(oracle.GAMMA)) < 2e-5, m


def test_mixed_solver_at_matched_residual(gpu):
    """Mixed FP32/FP64 APGD at tolerance 1e-4 against the FP64 oracle.

    gamma is not unique here (N is rank-deficient), D^T gamma is: the test
    compares the residual the oracle evaluates in FP64 on the returned gamma,
    and the velocities and contact forces of the free bodies against the
    oracle's solution at tolerance 1e-8: the mixed solution must be as
    close to it as the oracle's own solution at the same tolerance (within a
    factor 3), and within 2e-2 relative in absolute terms."""
    sc = scenes.compaction(2002, n=4000, width=1.5, lid=scenes.LID_BILATERAL)
    tol = 1e-4
    exact = oracle.OracleSystem(sc, tol=1e-8, max_iterations=100000)
    exact.step()
    assert exact.residual < 1e-7
    o, g = both(sc, solver=system.BUILD_MIXED, tol=tol, max_iterations=20000)
    o.step()
    g.step()
    st = g.stats()
    assert st.num_contacts == o.num_contacts
    assert st.residual < tol
    assert o.residual_of(g.gamma) < 1.1 * tol
    free = _free(sc)
    v_star, f_star = exact.vec(oracle.V)[free], exact.vec(oracle.F_CONTACT)[free]
    for vec_o, vec_g, star in ((oracle.V, g.v, v_star), (oracle.F_CONTACT, g.f_contact, f_star)):
        err_mixed = rel(vec_g[free], star)
        err_oracle = rel(o.vec(vec_o)[free], star)
        assert err_mixed < max(3 * err_oracle, 1e-6), (vec_o, err_mixed, err_oracle)
        assert err_mixed < 2e-2, (vec_o, err_mixed)


def test_ball_on_box_gpu(gpu):
    for solver in (system.BUILD_PLAIN, system.BUILD_FP64, system.BUILD_MIXED):
        g = system.from_scene(scenes.ball_on_box(0.0), solver)
        g.DoTimeStep()
        gam = g.gamma
        assert abs(gam[0] - 0.0981) < 1e-6 * (1e3 if solver == system.BUILD_MIXED else 1), (solver, gam)
        assert abs(g.f_contact[4] - 9.81) < 1e-4 * (1e2 if solver == system.BUILD_MIXED else 1)


def test_free_fall_gpu(gpu):
    sc = scenes.free_fall(5, n=8)
    g = system.from_scene(sc, system.BUILD_FP64)
    p0, v0 = g.p.copy(), g.v.copy()
    g.DoTimeStep()
    assert g.numCollisions == 0
    h = sc.params["h"]
    v1 = v0.reshape(-1, 3) + h * np.array(sc.params["gravity"])
    assert np.allclose(g.v.reshape(-1, 3), v1, rtol=0, atol=1e-15)
    assert np.allclose(g.p, p0 + h * g.v, rtol=0, atol=1e-15)


def test_rows_per_thread_agree(gpu):
    """One or two contacts per thread in the contact pass P1 only
    move rows between lanes: the partial sums keep the one-row-per-thread
    order, so the same ten iterations are bit-identical (both builds)."""
    sc = scenes.pile(31, 6000)
    for solver in (system.BUILD_FP64, system.BUILD_MIXED):
        out = []
        for r in (1, 2):
            g = system.from_scene(sc, solver, tol=0.0, max_iterations=10)
            g.setRowsPerThread(r)
            g.step()
            out.append((g.stats().iterations, g.gamma.copy(), g.v.copy()))
            g.close()
        assert out[0][0] == out[1][0] == 10
        assert bits_equal(out[1][1], out[0][1]) and bits_equal(out[1][2], out[0][2]), solver


def test_profiled_solves_are_unchanged(gpu):
    """Timing P1, P2 or the decide kernel (DV_PARAM_PROFILE) repeats a
    kernel in a burst, which must reproduce its outputs: the same 60
    iterations are bit-identical with and without it (both builds).  A pass
    has two kernels, so kernel 3 is refused."""
    sc = scenes.pile(31, 6000)
    for solver in (system.BUILD_FP64, system.BUILD_MIXED):
        out = {}
        for k in (system.PROFILE_OFF, system.PROFILE_P1, system.PROFILE_P2, system.PROFILE_DECIDE):
            g = system.from_scene(sc, solver, tol=0.0, max_iterations=60)
            g.setProfileKernel(k)
            g.step()
            st = g.stats()
            assert st.iterations == 60
            assert (st.prof_kernel_count > 0) == (k != system.PROFILE_OFF), (solver, k)
            out[k] = (g.gamma.copy(), g.v.copy())
            g.close()
        for k in (system.PROFILE_P1, system.PROFILE_P2, system.PROFILE_DECIDE):
            assert bits_equal(out[k][0], out[0][0]) and bits_equal(out[k][1], out[0][1]), (solver, k)
    g = system.from_scene(sc, system.BUILD_MIXED)
    with pytest.raises(system.DvideError):
        g.setProfileKernel(3)
    g.close()


def test_fp64_solver_velocities_match_pgs(gpu):
    """The graph-driven FP64 APGD against the oracle's projected Gauss-Seidel
    (PGS.cu:149-447), another algorithm for the same cone complementarity
    problem: the velocities of the solution are unique, so they agree."""
    sc = scenes.pile(5, 60)
    prev = oracle.set_threads(0)
    oracle.set_threads(1)
    try:
        p = oracle.OracleSystem(sc, tol=1e-10, max_iterations=4000000, solver=oracle.PGS)
        p.step()
    finally:
        oracle.set_threads(prev)
    assert p.residual < 1e-10
    g = system.from_scene(sc, system.BUILD_FP64, tol=1e-8, max_iterations=100000)
    g.step()
    assert g.stats().residual < 1e-8
    free = np.repeat(sc.fixed == 0, 3)
    vp = p.vec(oracle.V)[free]
    assert np.abs(g.v[free] - vp).max() < 1e-5 * np.abs(vp).max()


@pytest.mark.parametrize("solver,bound", [(system.BUILD_FP64, 1e-13), (system.BUILD_MIXED, 2e-6)])
def test_pass_product_matches_oracle(gpu, solver, bound):
    """The product the graph-driven solvers iterate with -- J = F^T x per
    contact, P2's per-body gather and the contact rows a P1 forms from it -- against the
    oracle's COO product D (M^-1 (D^T x)) on a random x, bilateral rows
    included: FP64 to rounding, the mixed build to binary32 rounding (the
    frames, M^-1, J and tmp are binary32 there)."""
    sc = scenes.compaction(2002, n=4000, width=1.5, lid=scenes.LID_BILATERAL)
    o, g = both(sc)
    nc = o.assemble()
    assert g.assemble() == nc and nc > 1000
    x = np.random.default_rng(11).normal(size=o.vec(oracle.R).size)
    ref = o.schur(x)
    y = g.pass_product(x, solver)
    assert rel(y, ref) < bound, rel(y, ref)
    assert np.abs(y - ref).max() < 10 * bound * np.abs(ref).max()
    # the scene's own state is untouched: the step afterwards is the plain one
    o.step()
    g.step()
    assert bits_equal(g.v, o.vec(oracle.V))


@pytest.mark.parametrize("tol", [1e-4, 1e-6])
def test_mixed_gamma_where_gamma_is_unique(gpu, tol):
    """One layer of spheres on a floor (sliding and sticking), one contact
    each: N is block diagonal and gamma is unique, so the mixed solver's
    gamma must be within 1e-3 relative of the FP64 oracle's at the same
    tolerance, and of the oracle's converged (1e-8) solution."""
    sc = scenes.layer(7)
    exact = oracle.OracleSystem(sc, tol=1e-8, max_iterations=100000)
    exact.step()
    assert exact.residual < 1e-8
    o = oracle.OracleSystem(sc, tol=tol, max_iterations=100000)
    o.step()
    g = system.from_scene(sc, system.BUILD_MIXED, tol=tol, max_iterations=100000)
    g.step()
    st = g.stats()
    assert st.num_contacts == o.num_contacts == 256
    assert st.residual < tol
    assert o.residual_of(g.gamma) < 1.1 * tol
    assert rel(g.gamma, o.vec(oracle.GAMMA)) < 1e-3
    assert rel(g.gamma, exact.vec(oracle.GAMMA)) < 1e-3
    # and the FP64 graph-driven build to the oracle's own accuracy
    f = system.from_scene(sc, system.BUILD_FP64, tol=1e-8, max_iterations=100000)
    f.step()
    assert rel(f.gamma, exact.vec(oracle.GAMMA)) < 1e-6


def test_ccp_violation_matches_oracle(gpu):
    """getCCPViolation (System.cu:1605-1643) of a bilateral-lid pile: on the
    plain build, whose trajectory is the oracle's bit for bit, the four
    values are bit-exact; on the graph-driven FP64 build they are the
    oracle's evaluation of the same gamma and v, bit for bit."""
    sc = scenes.compaction(2002, n=4000, width=1.5, lid=scenes.LID_BILATERAL)
    o, g = both(sc, solver=system.BUILD_PLAIN, tol=1e-5, max_iterations=20000)
    for _ in range(3):
        o.step()
        g.step()
    viol = g.getCCPViolation()
    assert bits_equal(viol, o.ccp_violation())
    assert viol[0] <= 0 and viol[1] <= 0 and viol[3] >= 0
    f = system.from_scene(sc, system.BUILD_FP64, tol=1e-5, max_iterations=20000)
    f.step()
    o1 = oracle.OracleSystem(sc, tol=1e-5, max_iterations=20000)
    o1.step()
    o1.set_vec(oracle.GAMMA, f.gamma)
    o1.set_vec(oracle.V, f.v)
    assert bits_equal(f.getCCPViolation(), o1.ccp_violation())
