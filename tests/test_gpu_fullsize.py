"""The benchmark workload at its full size (112,000 spheres, about 256k
contacts, D of 768k x 336k) against the oracle, on properties that do not
need an oracle solve at that size:

- the contact set, normals, gaps and locations, and r, b, k are bit-exact;
- the Schur product N x is bit-exact;
- the FP64 graph-driven solver's step: the residual the oracle evaluates in
  FP64 on the returned gamma is the one the solver reports, below the
  tolerance; the mixed solver's is within 10% of the tolerance;
- each free body's momentum change is h (m g + f_contact), the contact
  forces cancel between the two bodies of each contact, and gamma lies in
  the friction cone.
"""
import numpy as np
import pytest

import oracle
from dvide_amd import scenes, system
from tests.helpers import both, pair_keys

pytestmark = pytest.mark.gpu

TOL = 1e-4


def bits_equal(x, y):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    if x.shape != y.shape:
        return False
    nx, ny = np.isnan(x), np.isnan(y)
    return np.array_equal(nx, ny) and np.array_equal(x[~nx].view(np.uint64), y[~ny].view(np.uint64))


@pytest.fixture(scope="module")
def workload():
    return scenes.compaction(2002, lid=scenes.LID_NONE)


def test_full_size_contacts_assembly_and_product(gpu, workload):
    o, g = both(workload, tol=TOL)
    nc = o.assemble()
    assert g.assemble() == nc and nc > 250000
    a0, b0, n0, l0 = o.contacts()
    a1, b1, n1, l1 = g.contacts()
    assert np.array_equal(pair_keys(a0, b0), pair_keys(a1, b1))
    assert bits_equal(n0, n1) and bits_equal(l0, l1)
    for wo, wg in ((oracle.R, system.VEC_R), (oracle.B, system.VEC_B), (oracle.K, system.VEC_K)):
        assert bits_equal(o.vec(wo), g.vector(wg)), wg
    x = np.random.default_rng(5).normal(size=3 * nc)
    assert bits_equal(o.schur(x), g.schur_product(x))
    g.close()


@pytest.mark.parametrize("solver,bound", [(system.BUILD_FP64, 1e-13), (system.BUILD_MIXED, 2e-6)])
def test_full_size_pass_product(gpu, workload, solver, bound):
    """The product the solvers iterate with (J = F^T x, P2's gather, the
    contact rows the next P1 forms) against the oracle's N x at the full size: FP64 to rounding, the
    mixed build to binary32 rounding."""
    o, g = both(workload, tol=TOL)
    nc = o.assemble()
    assert g.assemble() == nc and nc > 250000
    x = np.random.default_rng(6).normal(size=3 * nc)
    ref = o.schur(x)
    y = g.pass_product(x, solver)
    err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    print(f"pass product, build {solver}: relative error {err:.2e} at {nc} contacts")
    assert err < bound
    g.close()


@pytest.mark.parametrize("solver", [system.BUILD_FP64, system.BUILD_MIXED])
def test_full_size_step(gpu, workload, solver):
    o, g = both(workload, solver=solver, tol=TOL, max_iterations=20000)
    g.step()
    st = g.stats()
    assert st.num_contacts > 250000 and st.iterations < 20000
    assert st.residual < TOL
    o.assemble()  # the oracle's D, r at the same (initial) state
    gamma = g.gamma
    res_o = o.residual_of(gamma)
    if solver == system.BUILD_FP64:
        assert abs(res_o - st.residual) <= 1e-3 * TOL, (res_o, st.residual)
    else:
        assert res_o < 1.1 * TOL, res_o
    # momentum of each free body over the step
    h = workload.params["h"]
    grav = np.asarray(workload.params["gravity"])
    free = workload.fixed == 0
    m = workload.mass[free][:, None]
    dv = g.v.reshape(-1, 3)[free] - workload.vel[free]
    fc = g.f_contact.reshape(-1, 3)[free]
    lhs = m * dv
    rhs = h * (m * grav + fc)
    assert np.abs(lhs - rhs).max() <= 1e-9 * max(np.abs(rhs).max(), 1.0)
    # gamma in the cone, up to the solver's tolerance
    gc = gamma.reshape(-1, 3)
    mu = workload.params["mu"]
    excess = np.hypot(gc[:, 1], gc[:, 2]) - mu * gc[:, 0]
    assert excess.max() <= 1e-12 * max(np.abs(gc).max(), 1.0)
    assert gc[:, 0].min() >= 0.0
    g.close()
