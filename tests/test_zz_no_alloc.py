"""dv_schur_product_noalloc: the Schur product N x staged through the
library's fixed device scratch instead of a per-call allocation.  Its result
must equal dv_schur_product's bit for bit, on two calls in a row of which the
second is smaller (the scratch keeps the first call's contents beyond the
second's rows), and it must refuse inputs beyond the scratch's 1,048,576
rows before touching the device."""
import numpy as np
import pytest

from dvide_amd import scenes, system

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 20


def assembled(sc):
    g = system.from_scene(sc, system.BUILD_PLAIN)
    assert g.assemble() > 0
    return g


def bits_equal(x, y):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    return x.shape == y.shape and np.array_equal(x.view(np.uint64), y.view(np.uint64))


def test_noalloc_product_equals_the_allocating_one(gpu):
    big = assembled(scenes.compaction(2002, n=4000, width=1.5, lid=scenes.LID_BILATERAL))
    small = assembled(scenes.pile(23, 900))
    rng = np.random.default_rng(5)
    n_big = big.vector(system.VEC_R).size
    n_small = small.vector(system.VEC_R).size
    assert n_small < n_big
    for g, n in ((big, n_big), (small, n_small)):
        x = rng.normal(size=n)
        ref = g.schur_product(x)
        out = g.schur_product_noalloc(x)
        assert bits_equal(out, ref)
        assert np.abs(ref).max() > 0
    big.close()
    small.close()


def test_noalloc_product_refuses_inputs_beyond_its_scratch(gpu):
    g = assembled(scenes.pile(23, 900))
    x = np.zeros(MAX_ROWS + 1)
    with pytest.raises(system.DvideError) as e:
        g.schur_product_noalloc(x)
    assert e.value.code == -1 and "1,048,576" in str(e.value)
    # a length that fits but is not the scene's row count is refused too
    with pytest.raises(system.DvideError):
        g.schur_product_noalloc(np.zeros(5))
    g.close()
