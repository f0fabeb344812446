"""The k-omega SST oracle (``oracle/turbulence_oracle.py``) on the CPU: the reference's own
trajectories (``tests/golden/sst_*.npz``, written by ``make_sst_golden.py``), Spalding's law
in its two limits, the exact wall distance of a straight channel and the viscosity before the
first solve."""
import math

import numpy as np
import pytest

import oracle.simplec_oracle as O
from oracle.turbulence_oracle import B_SPALDING, KAPPA, spalding_uplus
from sst_cases import DNS_YPLUS, GOLDEN, channel_hex, channel_profile, channel_quads, duct_tri, rel, sst_channel_dns


def test_spalding_law_matches_the_reference_and_its_limits():
    z = np.load(f"{GOLDEN}/sst_spalding.npz")
    up = spalding_uplus(z["re_y"])
    np.testing.assert_array_equal(up, z["uplus"])
    re_y = np.array([1e-3, 0.5, 1e5, 1e6])
    up = spalding_uplus(re_y)
    assert np.allclose(up[:2], np.sqrt(re_y[:2]), rtol=2e-3)  # sublayer: u+ = y+
    yp = re_y[2:] / up[2:]
    assert np.allclose(up[2:], np.log(yp) / KAPPA + B_SPALDING, rtol=2e-3)  # log layer
    # the returned u+ solves Spalding's equation with y+ = Re_y / u+
    re_y = np.geomspace(1.0, 1e5, 25)
    up = spalding_uplus(re_y)
    x = KAPPA * up
    rhs = up + math.exp(-KAPPA * B_SPALDING) * (np.exp(x) - 1 - x - x ** 2 / 2 - x ** 3 / 6)
    np.testing.assert_allclose(re_y / up, rhs, rtol=1e-9)
    assert spalding_uplus(np.array([0.0]))[0] == 0.0


def test_wall_distance_is_exact_on_a_straight_channel():
    """As ``tests/test_turbulence.py:24-31``: y = min(y, 1 - y) of the cell centres."""
    mesh, case = channel_quads(O, 1)
    s = O.Simplec(mesh, case)
    y = s.fv.geo.cell_centre[:, 1]
    np.testing.assert_allclose(s.turb.y, np.minimum(y, 1.0 - y), rtol=0, atol=1e-12)
    np.testing.assert_allclose(s.turb.mut, case.density * s.turb.k / s.turb.w, rtol=1e-15)


@pytest.mark.parametrize("build,name,n_iter", [(channel_quads, "sst_channel_quads", 12),
                                               (duct_tri, "sst_duct_tri", 10),
                                               (channel_hex, "sst_channel_hex", 5)])
def test_oracle_follows_the_reference_trajectory(build, name, n_iter):
    g = np.load(f"{GOLDEN}/{name}.npz")
    mesh, case = build(O, n_iter, zones=g["zones"] if "zones" in g.files else None)
    s = O.Simplec(mesh, case)
    t = s.turb
    assert rel(t.y, g["y"]) < 1e-14 and rel(t.yb, g["yb"]) < 1e-14
    np.testing.assert_array_equal(t.wall_cells, g["wall_cells"])
    np.testing.assert_array_equal(t.kb_in, g["kb_in"])
    np.testing.assert_array_equal(t.sb_in, g["sb_in"])
    names = list(g["res_names"])
    for it in range(1, n_iter + 1):
        s.step()
        for f, v in (("U", s.U), ("p", s.p), ("k", t.k), ("w", t.w), ("mut", t.mut)) + \
                ((("T", s.T),) if s.energy else ()):
            assert rel(v, g[f"it{it}_{f}"]) < 1e-10, (it, f, rel(v, g[f"it{it}_{f}"]))
        raw = s.residuals()["raw"]
        np.testing.assert_allclose([raw[n] for n in names], g[f"it{it}_res"], rtol=1e-9)
    assert rel(t.wall_shear_viscosity(s.U), g["mu_w"]) < 1e-12
    assert rel(t.yplus, g["yplus"]) < 1e-12


@pytest.mark.slow
def test_oracle_reproduces_the_reference_channel_validation():
    """The reference's SST pin (``tests/test_turbulence.py:53-58`` on
    ``flowtrace/validation/turbulent.py:62-90``): the Re_tau = 392 half channel converges without
    rollbacks, tau_w within 1 % of 1, and U+ within 5 % of the DNS from y+ = 1."""
    mesh, case = sst_channel_dns(O)
    s = O.Simplec(mesh, case)
    while not s.finished:
        s.step()
    _, err = channel_profile(s.fv.xc, s.U[:, 0])
    tau = s.turb.wall_shear()[s.turb.wall_b]
    assert s.converged and not s.guard_events
    assert abs(float(tau.mean()) - 1.0) < 0.01
    assert np.abs(err[DNS_YPLUS >= 1.0]).max() < 5.0, err


def test_unknown_model_rejected():
    """``tests/test_turbulence.py:70-72``, for the device case and the oracle's."""
    from gputest_amd.case import UnstructuredCase

    with pytest.raises(ValueError):
        UnstructuredCase(turbulence="spalart-allmaras-typo")
    with pytest.raises(ValueError):
        O.Case(turbulence="spalart-allmaras-typo")
