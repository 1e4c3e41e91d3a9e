"""The committed tile-file fixture solved on the GPU: the pressure equals the
oracle's, stored in the file, bit for bit."""
import os

import numpy as np
import pytest

import cirrus_amd as ca

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fixture_solve_matches_the_stored_pressure():
    tree, ch = ca.load_tiles(os.path.join(ROOT, "tests", "golden", "uniform16_solution.ctf"))
    P = ca.Problem(tree)
    P.set(ca.CH_B, ch[1])
    S = ca.GMGSolver()
    S.solve(P, False, 200, 1e-6, 1, 10, 1, False)
    assert S.last["converged"] == 1
    leaf = tree.type == 1
    assert np.array_equal(P.get(ca.CH_X)[leaf], ch[0][leaf])
