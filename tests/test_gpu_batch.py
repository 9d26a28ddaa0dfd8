"""Batched and served solves (dv_batch_*, dv_server_*) on the GPU.

A scene's partial sums are formed per tile of its own rows and reduced in a
fixed order, so its whole step -- contacts, iterations, gamma, v, p -- is
bit-identical alone, in a batch and in a server mix, whatever the other
scenes, slots or admission order.  The tests hold the mixed path to that
bar as well as the FP64 path.
"""
import numpy as np
import pytest

import oracle
from dvide_amd import scenes, system
from tests.helpers import rel

pytestmark = pytest.mark.gpu

# (seed, target contacts): uneven sizes, one with a bilateral-driven lid
CASES = [("pile", 21, 3000), ("pile", 22, 800), ("compaction", 2002, 0), ("pile", 24, 6000)]


def make(kind, seed, nc):
    if kind == "pile":
        return scenes.pile(seed, nc, 0.03, 0.06, 0.1 + 0.1 * (seed % 4))
    return scenes.compaction(seed, n=3000, width=1.2, lid=scenes.LID_BILATERAL)


def bits_equal(x, y):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    if x.shape != y.shape:
        return False
    nx, ny = np.isnan(x), np.isnan(y)
    return np.array_equal(nx, ny) and np.array_equal(x[~nx].view(np.uint64), y[~ny].view(np.uint64))


def same_state(a, b, tag):
    sa, sb = a.stats(), b.stats()
    assert sa.num_contacts == sb.num_contacts, tag
    assert sa.iterations == sb.iterations, (tag, sa.iterations, sb.iterations)
    assert sa.residual == sb.residual, tag
    for vec in (system.VEC_GAMMA, system.VEC_V, system.VEC_P):
        assert bits_equal(a.vector(vec), b.vector(vec)), (tag, vec)


def systems(solver, tol=1e-5):
    return [system.from_scene(make(*c), solver, tol=tol, max_iterations=5000) for c in CASES]


@pytest.mark.parametrize("solver", [system.BUILD_FP64, system.BUILD_MIXED])
def test_batch_equals_alone(gpu, solver):
    alone = systems(solver)
    together = systems(solver)
    batch = system.Batch(solver)
    for g in together:
        batch.add(g)
    for k in range(3):
        for g in alone:
            g.step()
        batch.step()
        st = batch.stats()
        assert st.num_scenes == len(CASES)
        assert st.total_contacts == sum(g.stats().num_contacts for g in alone)
        for i, (a, b) in enumerate(zip(alone, together)):
            same_state(a, b, (k, i))
    batch.close()


def test_batch_with_a_scene_without_contacts(gpu):
    """A member with nothing to solve (free fall) is only updated, and the
    others are solved as alone."""
    solver = system.BUILD_MIXED

    def members():
        return [system.from_scene(make("pile", 31, 900), solver, tol=1e-5, max_iterations=3000),
                system.from_scene(scenes.free_fall(7, n=12), solver, tol=1e-5, max_iterations=3000),
                system.from_scene(make("pile", 32, 400), solver, tol=1e-5, max_iterations=3000)]

    alone, together = members(), members()
    batch = system.Batch(solver)
    for g in together:
        batch.add(g)
    for k in range(2):
        for g in alone:
            g.step()
        batch.step()
        assert together[1].stats().num_contacts == 0
        for i, (a, b) in enumerate(zip(alone, together)):
            same_state(a, b, (k, i))
    batch.close()


def test_batch_of_many_equals_alone(gpu):
    """A batch large enough that its scenes' collision detection and
    assembly are queued from several host threads gives the same results."""
    solver = system.BUILD_MIXED
    cases = [("pile", 100 + i, 200 + 37 * i) for i in range(40)]
    alone = [system.from_scene(make(*c), solver, tol=1e-5, max_iterations=3000) for c in cases]
    together = [system.from_scene(make(*c), solver, tol=1e-5, max_iterations=3000) for c in cases]
    batch = system.Batch(solver)
    for g in together:
        batch.add(g)
    for k in range(2):
        for g in alone:
            g.step()
        batch.step()
        for i, (a, b) in enumerate(zip(alone, together)):
            same_state(a, b, (k, i))
    batch.close()


@pytest.mark.parametrize("max_slots,budget", [(8, 0.0), (2, 0.0), (4, 1.0)])
def test_server_mix_equals_alone(gpu, max_slots, budget):
    """Slots fewer than requests and a budget that admits one scene at a time
    change only when each request runs, not its result."""
    solver = system.BUILD_MIXED
    alone = systems(solver)
    served = systems(solver)
    srv = system.Server(solver, max_slots=max_slots, hbm_budget=budget)
    for k in range(2):
        for g in alone:
            g.step()
        tickets = {srv.submit(g): i for i, g in enumerate(served)}
        done = srv.drain()
        assert sorted(done) == sorted(tickets)
        for i, (a, b) in enumerate(zip(alone, served)):
            same_state(a, b, (k, i))
    srv.close()


def test_server_request_without_contacts(gpu):
    """A request with nothing to solve completes at admission, beside
    requests that are solved."""
    solver = system.BUILD_MIXED

    def members():
        return [system.from_scene(make("pile", 41, 700), solver, tol=1e-5, max_iterations=3000),
                system.from_scene(scenes.free_fall(9, n=10), solver, tol=1e-5, max_iterations=3000),
                system.from_scene(make("pile", 42, 500), solver, tol=1e-5, max_iterations=3000)]

    alone, served = members(), members()
    for g in alone:
        g.step()
    srv = system.Server(solver, max_slots=2)
    tickets = [srv.submit(g) for g in served]
    done = srv.drain()
    assert sorted(done) == sorted(tickets)
    assert served[1].stats().num_contacts == 0
    for i, (a, b) in enumerate(zip(alone, served)):
        same_state(a, b, i)
    srv.close()


def test_server_staggered_arrivals(gpu):
    """Requests submitted while others iterate join at chunk boundaries."""
    solver = system.BUILD_FP64
    alone = systems(solver)
    served = systems(solver)
    for g in alone:
        g.step()
    srv = system.Server(solver, max_slots=4)
    done = []
    for g in served:
        srv.submit(g)
        done += srv.poll()
    done += srv.drain()
    assert len(done) == len(served)
    for i, (a, b) in enumerate(zip(alone, served)):
        same_state(a, b, i)
    srv.close()


def test_server_rejects_a_second_request_in_flight(gpu):
    g = system.from_scene(make("pile", 21, 3000), system.BUILD_MIXED)
    srv = system.Server(system.BUILD_MIXED, max_slots=2)
    srv.submit(g)
    with pytest.raises(system.DvideError):
        srv.submit(g)
    srv.drain()
    srv.close()


def test_server_diverged_request_frees_its_slot(gpu):
    """A request whose solve turns non-finite (a sphere centre inside a box
    has no normal, so D has NaN rows) fails with DV_ERR_DIVERGED after its
    slot is freed: its ticket is reported by failed(), the server drains,
    and with two slots the third request, which waits for a free one, is
    still admitted; the others are solved as they are alone."""
    solver = system.BUILD_MIXED
    bad = system.from_scene(scenes.cloud(31, n=400, edge=0.6), solver)

    def good():
        return [system.from_scene(make("pile", 41, 700), solver, tol=1e-5, max_iterations=3000),
                system.from_scene(make("pile", 42, 500), solver, tol=1e-5, max_iterations=3000)]

    alone, served = good(), good()
    for g in alone:
        g.step()
    srv = system.Server(solver, max_slots=2)
    t_bad = srv.submit(bad)
    t_good = [srv.submit(g) for g in served]
    done, errors = [], 0
    for _ in range(100000):  # a chunk per poll; bounded so that a stuck server fails the test
        if not srv.pending():
            break
        try:
            done += srv.poll()
        except system.DvideError as e:
            assert e.code == -6  # DV_ERR_DIVERGED
            errors += 1
            done += e.done
    assert srv.pending() == 0
    assert errors == 1 and srv.failed() == [t_bad]
    assert sorted(done) == sorted(t_good)
    for i, (a, b) in enumerate(zip(alone, served)):
        same_state(a, b, ("after the diverged request", i))
    for g in alone:  # the freed slots take the next requests
        g.step()
    again = [srv.submit(g) for g in served]
    assert sorted(srv.drain()) == sorted(again)
    for i, (a, b) in enumerate(zip(alone, served)):
        same_state(a, b, ("second requests", i))
    srv.close()


def test_batch_rejects_the_plain_solver(gpu):
    with pytest.raises(system.DvideError):
        system.Batch(system.BUILD_PLAIN)


def _check_against_oracle(scene, g, tol, tag):
    """One step of `g` from `scene`'s initial state against the oracle: the
    residual the oracle evaluates in FP64 on the returned gamma, and the
    free bodies' velocities against the oracle's solution at tolerance 1e-7
    (as close as the oracle's own solution at `tol`, within a factor 3, or
    2e-2 relative)."""
    o = oracle.OracleSystem(scene, tol=tol, max_iterations=100000)
    o.step()
    exact = oracle.OracleSystem(scene, tol=1e-7, max_iterations=30000)
    exact.step()
    assert exact.residual < 1e-7, tag
    st = g.stats()
    assert st.num_contacts == o.num_contacts, tag
    assert st.residual < tol, tag
    o_asm = oracle.OracleSystem(scene, tol=tol)
    o_asm.assemble()
    assert o_asm.residual_of(g.gamma) < 1.1 * tol, tag
    free = np.repeat(scene.fixed == 0, 3)
    star = exact.vec(oracle.V)[free]
    err = rel(g.v[free], star)
    assert err < max(3 * rel(o.vec(oracle.V)[free], star), 1e-6) or err < 2e-2, (tag, err)


@pytest.mark.parametrize("solver", [system.BUILD_FP64, system.BUILD_MIXED])
def test_batch_members_match_oracle(gpu, solver):
    """Each member of a batch step against the oracle, not only against its
    own solve alone."""
    tol = 1e-5
    scs = [make(*c) for c in CASES]
    members = [system.from_scene(sc, solver, tol=tol, max_iterations=50000) for sc in scs]
    batch = system.Batch(solver)
    for g in members:
        batch.add(g)
    batch.step()
    for i, (sc, g) in enumerate(zip(scs, members)):
        _check_against_oracle(sc, g, tol, i)
    batch.close()


def test_served_members_match_oracle(gpu):
    tol = 1e-5
    scs = [make(*c) for c in CASES]
    members = [system.from_scene(sc, system.BUILD_MIXED, tol=tol, max_iterations=50000) for sc in scs]
    srv = system.Server(system.BUILD_MIXED, max_slots=2)
    for g in members:
        srv.submit(g)
    assert len(srv.drain()) == len(members)
    for i, (sc, g) in enumerate(zip(scs, members)):
        _check_against_oracle(sc, g, tol, i)
    srv.close()
