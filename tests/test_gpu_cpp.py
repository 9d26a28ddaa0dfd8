"""The C++ drop-in surface (include/dvide/System.hpp) driven like the
reference's models/test_oneBall.cu, for every solver."""
import os
import re
import subprocess

import pytest

from dvide_amd import system

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "cpp", "bin",
                      "one_ball")


@pytest.mark.parametrize("solver", [system.SOLVER_FUSED, system.SOLVER_MIXED, system.SOLVER_PLAIN])
def test_cpp_one_ball(gpu, solver):
    r = subprocess.run([DRIVER, str(solver)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    vals = dict(re.findall(r"(\w+) ([-0-9.e+]+)", r.stdout))
    assert int(vals["collisions"]) == 1
    tol = 1e-6 if solver != system.SOLVER_MIXED else 1e-5
    assert abs(float(vals["gamma"]) - 9.81e-2) < tol * 9.81e-2 * 10
    assert abs(float(vals["f_contact_y"]) - 9.81) < tol * 9.81 * 10
    assert abs(float(vals["time"]) - 0.03) < 1e-12
