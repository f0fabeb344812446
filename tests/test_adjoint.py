"""The discrete adjoint on the CPU: the torch restatement of one outer iteration has the
oracle's converged state as its fixed point, and the adjoint gradient (SciPy solves) equals
central finite differences (h = 1e-6) of the oracle's converged drag. No reference
counterpart: the reference has no adjoint, so finite differences of the restated forward
are the pin (the forward itself is pinned to the reference's golden vectors)."""
import numpy as np
import pytest
import torch

import oracle.simplec_oracle as O
from oracle.adjoint_oracle import CpuResidualMap, solver_view
from gputest_amd.adjoint import drag_gradient, heat_gradient
from tests.adjoint_cases import case, energy_case, small_mesh, u_ctrl0

NU = 1.0 / 40.0


def _solve(mesh, u, nu):
    c, prof = case(O, u, nu)
    o = O.Simplec(mesh, c)
    while not o.finished:
        o.step()
    assert o.converged
    return o, prof


@pytest.fixture(scope="module")
def base():
    mesh = small_mesh()
    o, prof = _solve(mesh, u_ctrl0(), NU)
    return mesh, o, prof


def test_outer_iteration_restatement_has_the_oracle_fixed_point(base):
    mesh, o, prof = base
    rm = CpuResidualMap(solver_view(o), "cylinder", "inlet", prof)
    w, beta = rm.state(), rm.params()
    out = rm.step(w, beta)
    for name, a, b in zip(("U", "p", "F", "Fb", "grad p", "grad U"), out, w):
        err = float((a - b).detach().abs().max() / max(float(b.detach().abs().max()), 1e-300))
        assert err < 1e-11, (name, err)
    drag = float(rm.objective(w, beta).detach())
    assert abs(drag - o.wall_forces()["cylinder"]["total"][0]) <= 1e-13 * abs(drag)


def test_cpu_adjoint_matches_central_differences(base):
    mesh, o, prof = base
    rm = CpuResidualMap(solver_view(o), "cylinder", "inlet", prof)
    gu, gnu, J, info = drag_gradient(None, prof, rmap=rm, tol=1e-11)
    print("adjoint", info)
    assert info.residual < 1e-10
    u0, h = u_ctrl0(), 1e-6

    def drag(u, nu):
        return _solve(mesh, u, nu)[0].wall_forces()["cylinder"]["total"][0]

    for k in (31,):
        up, um = u0.copy(), u0.copy()
        up[k] += h
        um[k] -= h
        fd = (drag(up, NU) - drag(um, NU)) / (2 * h)
        assert abs(float(gu[k]) - fd) <= 1e-5 * abs(fd), (k, float(gu[k]), fd)
    fd = (drag(u0, NU * (1 + h)) - drag(u0, NU * (1 - h))) / (2 * h * NU)
    assert abs(float(gnu) - fd) <= 1e-5 * abs(fd), (float(gnu), fd)


T_WALL = 350.0


def _solve_energy(mesh, u, T_wall):
    c, prof = energy_case(O, u, NU, T_wall)
    o = O.Simplec(mesh, c)
    while not o.finished:
        o.step()
    assert o.converged
    return o, prof


@pytest.fixture(scope="module")
def heated():
    mesh = small_mesh()
    o, prof = _solve_energy(mesh, u_ctrl0(), T_WALL)
    return mesh, o, prof


def test_restated_iteration_with_energy_has_the_oracle_fixed_point(heated):
    mesh, o, prof = heated
    rm = CpuResidualMap(solver_view(o), "outlet", "inlet", prof, objective="heat", wall_T_zone="cylinder")
    w, beta = rm.state(), rm.params()
    out = rm.step(w, beta)
    for name, a, b in zip(("U", "p", "F", "Fb", "grad p", "grad U", "T", "grad T"), out, w):
        a, b = a.detach(), b.detach()
        err = float((a - b).abs().max() / max(float(b.abs().max()), 1e-300))
        assert err < 1e-11, (name, err)
    J = float(rm.objective(w, beta).detach())
    assert abs(J - o.heat_flow()["outlet"]) <= 1e-12 * abs(J)


def test_cpu_heat_adjoint_matches_central_differences(heated):
    """d(outlet heat flow)/d(inlet control value, cylinder temperature)."""
    mesh, o, prof = heated
    rm = CpuResidualMap(solver_view(o), "outlet", "inlet", prof, objective="heat", wall_T_zone="cylinder")
    gu, gT, J, info = heat_gradient(None, prof, rmap=rm, tol=1e-11)
    print("heat adjoint", info, "J", J, "dJ/dT_wall", float(gT))
    assert info.residual < 1e-10
    u0, h = u_ctrl0(), 1e-6

    def heat(u, T_wall):
        return _solve_energy(mesh, u, T_wall)[0].heat_flow()["outlet"]

    k = 31
    up, um = u0.copy(), u0.copy()
    up[k] += h
    um[k] -= h
    fd = (heat(up, T_WALL) - heat(um, T_WALL)) / (2 * h)
    assert abs(float(gu[k]) - fd) <= 1e-5 * abs(fd), (k, float(gu[k]), fd)
    fd = (heat(u0, T_WALL * (1 + h)) - heat(u0, T_WALL * (1 - h))) / (2 * h * T_WALL)
    assert abs(float(gT) - fd) <= 1e-5 * abs(fd), (float(gT), fd)
