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
from gputest_amd.adjoint import drag_gradient
from tests.adjoint_cases import case, small_mesh, u_ctrl0

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
        err = float((a - b).abs().max() / max(float(b.abs().max()), 1e-300))
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
