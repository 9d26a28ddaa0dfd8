"""The batched and served workloads at their full shape, checked member by
member on properties that need no oracle solve at that size, and sampled
members against the oracle's FP64 evaluation of the returned impulses.

- 512 independent piles, contacts log-uniform in [1k, 100k], radii
  U[0.03, 0.06] m, mu U[0.1, 0.5], one mixed-precision batch step (about 8M
  contacts) from the fresh deposits;
- 256 resident piles, contacts log-uniform in [1k, 200k], radii
  U[0.02, 0.03] m, mu 0.4, one request each through the server.

For every member: the step converged to the tolerance, gamma lies in the
friction cone, and each free body's momentum change is h (m g + f_contact).
For sampled members: contacts bit-exact with the oracle and the residual the
oracle evaluates in FP64 on the returned gamma within 10% of the tolerance.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from dvide_amd import scenes, system

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _generate(specs):
    with ThreadPoolExecutor(16) as ex:
        return list(ex.map(lambda a: scenes.pile(*a), specs))


def _check_member(sc, g, tag):
    st = g.stats()
    if st.num_contacts == 0:
        return
    assert st.residual < TOL, (tag, st.residual)
    gc = g.gamma.reshape(-1, 3)
    mu = sc.params["mu"]
    excess = np.hypot(gc[:, 1], gc[:, 2]) - mu * gc[:, 0]
    assert excess.max() <= 1e-12 * max(np.abs(gc).max(), 1.0), tag
    assert gc[:, 0].min() >= 0.0, tag
    h = sc.params["h"]
    grav = np.asarray(sc.params["gravity"])
    free = sc.fixed == 0
    m = sc.mass[free][:, None]
    lhs = m * (g.v.reshape(-1, 3)[free] - sc.vel[free])
    rhs = h * (m * grav + g.f_contact.reshape(-1, 3)[free])
    assert np.abs(lhs - rhs).max() <= 1e-9 * max(np.abs(rhs).max(), 1.0), tag


def _against_oracle(sc, g, tag):
    o = oracle.OracleSystem(sc, tol=TOL)
    nc = o.assemble()
    assert g.stats().num_contacts == nc, tag
    a0, b0, _, _ = o.contacts()
    a1, b1, _, _ = g.contacts()
    assert np.array_equal(a0, a1) and np.array_equal(b0, b1), tag
    assert o.residual_of(g.gamma) < 1.1 * TOL, tag


def test_batch_of_512_piles(gpu):
    rng = np.random.default_rng(4004)
    targets = np.exp(rng.uniform(np.log(1e3), np.log(1e5), 512)).astype(int)
    mus = rng.uniform(0.1, 0.5, 512)
    scs = _generate([(4004 + i, int(targets[i]), 0.03, 0.06, float(mus[i])) for i in range(512)])
    members = [system.from_scene(sc, system.BUILD_MIXED, tol=TOL, max_iterations=100000) for sc in scs]
    batch = system.Batch(system.BUILD_MIXED)
    for g in members:
        batch.add(g)
    batch.step()
    st = batch.stats()
    assert st.num_scenes == 512 and st.total_contacts > 5_000_000
    for i, (sc, g) in enumerate(zip(scs, members)):
        _check_member(sc, g, i)
    # the oracle on a sample: the smallest, the largest and two in between
    order = np.argsort(targets)
    for i in (order[0], order[170], order[340], order[-1]):
        _against_oracle(scs[i], members[i], int(i))
    batch.close()
    for g in members:
        g.close()


def test_served_256_piles(gpu):
    rng = np.random.default_rng(5005)
    targets = np.exp(rng.uniform(np.log(1e3), np.log(2e5), 256)).astype(int)
    scs = _generate([(5005 + i, int(targets[i]), 0.02, 0.03, 0.4) for i in range(256)])
    members = [system.from_scene(sc, system.BUILD_MIXED, tol=TOL, max_iterations=100000) for sc in scs]
    srv = system.Server(system.BUILD_MIXED, max_slots=64)
    tickets = [srv.submit(g) for g in members]
    done = srv.drain()
    assert sorted(done) == sorted(tickets)
    for i, (sc, g) in enumerate(zip(scs, members)):
        _check_member(sc, g, i)
    order = np.argsort(targets)
    for i in (order[0], order[128], order[-1]):
        _against_oracle(scs[i], members[i], int(i))
    srv.close()
    for g in members:
        g.close()
