"""The launch schedules of the graph-driven solvers and their tuner.

No schedule changes a result: contact rows per thread (1 or 2) keep the
one-row-per-thread order of every partial sum, and the XCD-contiguous tile
order and the grid sizes only decide which workgroup takes which tile.  The
tuner therefore measures candidate schedules on live chunks of passes, and a
scene's whole step is bit-identical whatever schedule it ran under, tuned,
fixed or mid-measurement.  A decided schedule is written to the schedule
file and taken from it, unmeasured, by a later process."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from dvide_amd import scenes, system

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bits_equal(x, y):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    return x.shape == y.shape and np.array_equal(x.view(np.uint64), y.view(np.uint64))


@pytest.mark.parametrize("solver", [system.BUILD_FP64, system.BUILD_MIXED])
def test_rows_per_thread_give_identical_results(gpu, solver):
    sc = scenes.pile(31, 6000)
    out = []
    for r in (1, 2):
        g = system.from_scene(sc, solver, tol=1e-5, max_iterations=5000)
        g.setRowsPerThread(r)
        for _ in range(2):
            g.step()
        st = g.stats()
        assert st.rows_per_thread == r and st.schedule_tuning == 0
        out.append((st.iterations, st.residual, g.gamma.copy(), g.v.copy(), g.p.copy()))
        g.close()
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    for k in (2, 3, 4):
        assert bits_equal(out[0][k], out[1][k])


_CHILD = r"""
import json, sys
sys.path.insert(0, %(root)r)
from dvide_amd import scenes, system
mode = sys.argv[1]
out = {}
if mode == "one":
    g = system.from_scene(scenes.pile(33, 8000), system.BUILD_MIXED, tol=1e-6, max_iterations=20000)
    steps = []
    for _ in range(3):
        g.step()
        st = g.stats()
        steps.append([st.schedule_tuning, st.rows_per_thread, st.xcd_order])
    out["steps"] = steps
    out["gamma"] = g.gamma.tolist()
    out["v"] = g.v.tolist()
else:
    members = [system.from_scene(scenes.pile(40 + i, 1500 + 700 * i), system.BUILD_MIXED, tol=1e-6,
                                 max_iterations=20000) for i in range(6)]
    b = system.Batch(system.BUILD_MIXED)
    for m in members:
        b.add(m)
    steps = []
    for _ in range(3):
        b.step()
        st = b.stats()
        steps.append([st.schedule_tuning, st.rows_per_thread, st.xcd_order, st.wg_per_cu_contacts,
                      st.wg_per_cu_bodies])
    out["steps"] = steps
    out["v"] = [x for m in members for x in m.v.tolist()]
print("RESULT " + json.dumps(out))
"""


def _child(mode, env):
    r = subprocess.run([sys.executable, "-c", _CHILD % {"root": ROOT}, mode], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    line = [x for x in r.stdout.splitlines() if x.startswith("RESULT ")][-1]
    return json.loads(line[7:])


@pytest.mark.parametrize("mode", ["one", "many"])
def test_tuned_schedule_is_persisted_and_reused(gpu, mode, tmp_path):
    """First process: the solves measure the candidates, decide, and write
    the class's schedule; second process: the same class is read back and
    runs the same schedule without measuring; a third with tuning off runs
    the default.  All three give the same results bit for bit."""
    path = tmp_path / "schedules.txt"
    env = dict(os.environ, DVIDE_SCHEDULE_FILE=str(path))
    first = _child(mode, env)
    assert first["steps"][0][0] == 1  # the first step measured candidates
    assert first["steps"][-1][0] == 0  # and the class was decided by the last
    lines = path.read_text().split()
    assert len(lines) >= 6  # key rpt xcd wgc wgb score
    second = _child(mode, env)
    assert all(s[0] == 0 for s in second["steps"])
    assert second["steps"][0][1:] == first["steps"][-1][1:]
    assert path.read_text().split() == lines  # nothing measured or added
    off = _child(mode, dict(os.environ, DVIDE_TUNE="0", DVIDE_SCHEDULE_FILE=str(tmp_path / "none.txt")))
    assert all(s[0] == 0 for s in off["steps"])
    for key in ("v", "gamma"):
        if key in first:
            assert first[key] == second[key] == off[key]
