"""Manufactured-solution verification of the HIP solver (``tests/test_verification.py:40-51``
of the reference, on the device): the volumetric momentum and energy sources
(``SimpleSolver.set_sources``), observed orders on warped triangles, and the discretisation
errors and iteration counts equal to the oracle's on the same grids."""
import pytest

import oracle.simplec_oracle as O
from gputest_amd import case as C

from mms_cases import orders, run

pytestmark = pytest.mark.gpu


def test_hip_solver_is_second_order_and_equals_the_oracle():
    dev = [run(C, n, energy=True) for n in (10, 20, 40)]
    ref = [run(O, n, energy=True) for n in (10, 20, 40)]
    p = orders([r[0] for r in dev], [r[1] for r in dev])
    print({k: [round(v, 3) for v in vs] for k, vs in p.items()}, [r[2] for r in dev], [r[2] for r in ref], flush=True)
    assert all(r[3].converged for r in dev)
    assert p["u"][-1] > 1.8 and p["v"][-1] > 1.8 and p["T"][-1] > 1.8, p
    assert p["p"][-1] > 1.1, p
    for d, o in zip(dev, ref):
        assert d[2] == o[2]  # the same trajectory
        for k in d[1]:
            assert abs(d[1][k] - o[1][k]) <= 1e-8 * o[1][k], (k, d[1][k], o[1][k])


def test_hip_least_squares_pressure_force_is_inconsistent_on_triangles():
    gg = run(C, 20)[1]
    ls = run(C, 20, pressure_gradient="least-squares")[1]
    assert gg["u"] < 0.3 * ls["u"], (gg, ls)
