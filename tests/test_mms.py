"""Manufactured-solution verification on the oracle (``tests/test_verification.py:11-51`` of
the reference): the manufactured field and its hand-derived sources against finite
differences, observed second order of u, v and T (p above first order) on warped triangles, and
the Green-Gauss pressure force against the least-squares one."""
import numpy as np
import pytest

import oracle.simplec_oracle as O

from mms_cases import exact, orders, run, sources


def test_manufactured_solution_is_consistent():
    """Divergence free, zero on the walls, and the sources equal the steady residuals taken by
    central differences (h = 1e-5) of the exact field."""
    h, nu, kappa = 1e-5, 0.3, 0.1
    x, y = np.random.default_rng(0).uniform(0.1, 0.9, (2, 50))
    e = exact(x, y)

    def d(f, ax):
        return ((exact(x + h, y)[f] - exact(x - h, y)[f]) if ax == 0 else (exact(x, y + h)[f] - exact(x, y - h)[f])) / (2 * h)

    def lap(f):
        return (exact(x + h, y)[f] + exact(x - h, y)[f] + exact(x, y + h)[f] + exact(x, y - h)[f] - 4 * e[f]) / h ** 2

    assert np.abs(d("u", 0) + d("v", 1)).max() < 1e-6
    t = np.linspace(0, 1, 11)
    for xx, yy in ((t, 0 * t), (t, 0 * t + 1), (0 * t, t), (0 * t + 1, t)):
        w = exact(xx, yy)
        assert max(np.abs(w[k]).max() for k in ("u", "v", "T")) < 1e-12
    s = sources(x, y, nu, kappa)
    fu = e["u"] * d("u", 0) + e["v"] * d("u", 1) + d("p", 0) - nu * lap("u")
    fv = e["u"] * d("v", 0) + e["v"] * d("v", 1) + d("p", 1) - nu * lap("v")
    fT = e["u"] * d("T", 0) + e["v"] * d("T", 1) - kappa * lap("T")
    for a, b in ((fu, s["u"]), (fv, s["v"]), (fT, s["T"])):
        assert np.abs(a - b).max() < 1e-3 * max(1.0, np.abs(b).max())


@pytest.mark.slow
def test_oracle_is_second_order_on_warped_triangles():
    """``test_verification.py:40-43`` (u, v > 1.8, p > 1.1), with the energy equation on
    (T > 1.8): grids 10, 20, 40."""
    res = [run(O, n, energy=True) for n in (10, 20, 40)]
    p = orders([r[0] for r in res], [r[1] for r in res])
    print({k: [round(v, 3) for v in vs] for k, vs in p.items()}, [r[2] for r in res])
    assert all(r[3].converged for r in res)
    assert p["u"][-1] > 1.8 and p["v"][-1] > 1.8 and p["T"][-1] > 1.8, p
    assert p["p"][-1] > 1.1, p


@pytest.mark.slow
def test_least_squares_pressure_force_is_inconsistent_on_triangles():
    """``test_verification.py:46-50``: the Green-Gauss face-pressure force is the default
    because the least-squares one is several times less accurate on triangles."""
    gg = run(O, 20)[1]
    ls = run(O, 20, pressure_gradient="least-squares")[1]
    assert gg["u"] < 0.3 * ls["u"], (gg, ls)
