"""Four times the benchmark workload (450,000 spheres in a 10 x 10 m box,
about a million contacts, some 2,000 row tiles): the sizes past which the
grids, tile tables and partial-sum arrays must still hold.

- the contact set, normals and gaps, and r are bit-exact with the oracle;
- a mixed-precision step, capped at 3,000 iterations: the residual the
  oracle evaluates in FP64 on the returned gamma is the one the solver
  reports (to binary32 rounding), gamma lies in the friction cone, and each
  free body's momentum change is h (m g + f_contact).
"""
import numpy as np
import pytest

import oracle
from dvide_amd import scenes, system
from tests.helpers import both, pair_keys

pytestmark = pytest.mark.gpu


def bits_equal(x, y):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    if x.shape != y.shape:
        return False
    nx, ny = np.isnan(x), np.isnan(y)
    return np.array_equal(nx, ny) and np.array_equal(x[~nx].view(np.uint64), y[~ny].view(np.uint64))


@pytest.fixture(scope="module")
def workload():
    return scenes.compaction(2002, n=450000, width=10.0, lid=scenes.LID_NONE)


def test_max_size_contacts_and_r(gpu, workload):
    o, g = both(workload)
    nc = o.assemble()
    assert g.assemble() == nc and nc > 900000
    a0, b0, n0, _ = o.contacts()
    a1, b1, n1, _ = g.contacts()
    assert np.array_equal(pair_keys(a0, b0), pair_keys(a1, b1))
    assert bits_equal(n0, n1)
    assert bits_equal(o.vec(oracle.R), g.vector(system.VEC_R))
    g.close()


def test_max_size_mixed_step(gpu, workload):
    o, g = both(workload, solver=system.BUILD_MIXED, tol=1e-4, max_iterations=3000)
    g.step()
    st = g.stats()
    assert st.num_contacts > 900000 and 0 < st.iterations <= 3000
    o.assemble()
    gamma = g.gamma
    res_o = o.residual_of(gamma)
    assert abs(res_o - st.residual) <= 0.1 * st.residual, (res_o, st.residual)
    gc = gamma.reshape(-1, 3)
    mu = workload.params["mu"]
    assert gc[:, 0].min() >= 0.0
    assert (np.hypot(gc[:, 1], gc[:, 2]) - mu * gc[:, 0]).max() <= 1e-12 * max(np.abs(gc).max(), 1.0)
    h = workload.params["h"]
    grav = np.asarray(workload.params["gravity"])
    free = workload.fixed == 0
    m = workload.mass[free][:, None]
    dv = g.v.reshape(-1, 3)[free] - workload.vel[free]
    rhs = h * (m * grav + g.f_contact.reshape(-1, 3)[free])
    assert np.abs(m * dv - rhs).max() <= 1e-9 * max(np.abs(rhs).max(), 1.0)
    g.close()
