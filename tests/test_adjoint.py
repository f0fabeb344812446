"""The discrete adjoint on the CPU. No reference counterpart: the reference has no adjoint, so
the pins are the oracle (itself pinned to the reference's golden vectors):

* the torch restatement of one outer iteration has the oracle's converged state as its fixed
  point;
* the hand-written vector-Jacobian products equal torch autograd's through the restatement;
* the adjoint gradient (SciPy solves in the products) equals the independent tangent reference
  (``oracle/tangent_oracle.py``: central-difference Jacobian-vector products of the oracle's own
  outer iteration, nothing of ``gputest_amd.adjoint``) along a random direction over all 65
  parameters, at the first and last profile points and in nu, and central finite differences
  (h = 1e-6) of the oracle's converged drag along the random direction and at one profile
  point (those re-converge the oracle per perturbed parameter vector: marked slow).

The converged base states come from the cache ``tests/golden/adjoint_*_converged.npz``, checked
against the oracle's own iteration where they are loaded."""
import os

import numpy as np
import pytest
import torch

import oracle.simplec_oracle as O
from oracle.adjoint_oracle import CpuResidualMap, solver_view
from oracle.tangent_oracle import TangentReference
from gputest_amd.adjoint import drag_gradient, heat_gradient
from gputest_amd.adjoint_vjp import Linearisation
from tests.adjoint_cases import HALF, case, energy_case, small_mesh, u_ctrl0

NU = 1.0 / 40.0


GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _solve(mesh, u, nu):
    c, prof = case(O, u, nu)
    o = O.Simplec(mesh, c)
    while not o.finished:
        o.step()
    assert o.converged
    return o, prof


def _converged(mesh, c, name):
    """The oracle's converged state of a case, from its cache
    (``tests/golden/make_adjoint_golden.py``), checked to be a fixed point of the oracle's outer
    iteration: one more iteration moves no field by more than 1e-12 of its size."""
    g = np.load(os.path.join(GOLD, name))

    def load():
        o = O.Simplec(mesh, c)
        st = {k: g[k] for k in ("U", "p", "F", "Fb")}
        if "T" in g.files:
            st["T"] = g["T"]
        o.load_state({**st, "iteration": int(g["iteration"]), "res_continuity_norm": float(g["res_continuity_norm"])})
        o.set_lagged_gradients(g["gp"], list(g["gU"]), g["gT"] if "gT" in g.files else None)
        return o

    o, probe = load(), load()
    probe.case.max_iter = probe.iteration + 1
    probe.step()
    for k in ("U", "p", "F", "Fb") + (("T",) if probe.energy else ()):
        a, b = getattr(probe, k), getattr(o, k)
        assert np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1e-300), k
    return o


@pytest.fixture(scope="module")
def base():
    mesh = small_mesh()
    c, prof = case(O, u_ctrl0(), NU)
    return mesh, _converged(mesh, c, "adjoint_drag_converged.npz"), prof


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


def _direction():
    d = np.random.default_rng(7).standard_normal(65)
    return d / np.linalg.norm(d)


@pytest.mark.parametrize("energy", [False, True])
def test_hand_written_products_equal_autograd(base, heated, energy):
    """(dPhi/dw)^T v and (dPhi/dbeta)^T v by the hand-written transposes against torch
    autograd through the restated iteration, for random v; and dJ/dw, dJ/dbeta."""
    mesh, o, prof = heated if energy else base
    kw = dict(objective="heat", wall_T_zone="cylinder") if energy else {}
    rm = CpuResidualMap(solver_view(o), "outlet" if energy else "cylinder", "inlet", prof, **kw)
    w, beta = rm.state(), rm.params()
    out = rm.step(w, beta)
    rng = np.random.default_rng(1)
    bars = [torch.as_tensor(rng.standard_normal(tuple(t.shape))) for t in out]
    ref = torch.autograd.grad(out, w + beta, grad_outputs=bars, allow_unused=True)
    lin = Linearisation(rm, w, beta)
    hw, hb = lin.vjp(bars)
    for a, b in zip(hw + hb, ref):
        b = torch.zeros_like(a) if b is None else b
        assert float((a - b).abs().max()) <= 1e-12 * max(float(b.abs().max()), 1e-300)
    J = rm.objective(w, beta)
    gref = torch.autograd.grad(J, w + beta, allow_unused=True)
    J2, gw, gb = lin.objective_grad()
    assert float(J2) == float(J)
    for a, b in zip(gw + gb, gref):
        b = torch.zeros_like(a) if b is None else b
        assert float((a - b).abs().max()) <= 1e-13 * max(float(b.abs().max()), 1e-300)


def test_cpu_adjoint_matches_the_independent_tangent_reference(base):
    """The adjoint gradient against the tangent reference built from the oracle alone, along a
    random direction over all 65 parameters, at profile points 0 and 63 (which no inlet face of
    this mesh reaches: both derivatives are zero) and in nu."""
    mesh, o, prof = base
    rm = CpuResidualMap(solver_view(o), "cylinder", "inlet", prof)
    gu, gnu, J, info = drag_gradient(None, prof, rmap=rm, tol=1e-11)
    grad = np.r_[gu.numpy(), float(gnu)]
    c, _ = case(O, u_ctrl0(), NU)
    tr = TangentReference(_converged(mesh, c, "adjoint_drag_converged.npz"), u_ctrl0(), -HALF, HALF)
    assert tr.fixed_point_residual() < 1e-12
    d = _direction()
    for name, v in (("random", d), ("k=0", np.eye(65)[0]), ("k=63", np.eye(65)[63]), ("nu", np.eye(65)[64])):
        t, rec = tr.directional(v)
        a = float(grad @ v)
        print(name, "adjoint", a, "tangent", t, rec)
        assert abs(a - t) <= 1e-5 * max(abs(t), 1e-3 * np.abs(grad).max()), (name, a, t)


@pytest.mark.slow
def test_cpu_adjoint_matches_central_differences_along_a_random_direction(base):
    """Central differences (h = 1e-6) of the oracle's converged drag, re-converged from rest for
    each perturbed parameter vector, along the random direction over all 65 parameters."""
    mesh, o, prof = base
    rm = CpuResidualMap(solver_view(o), "cylinder", "inlet", prof)
    gu, gnu, J, info = drag_gradient(None, prof, rmap=rm, tol=1e-11)
    grad = np.r_[gu.numpy(), float(gnu)]
    d, h = _direction(), 1e-6

    def drag(beta):
        return _solve(mesh, beta[:64], beta[64])[0].wall_forces()["cylinder"]["total"][0]

    b0 = np.r_[u_ctrl0(), NU]
    fd = (drag(b0 + h * d) - drag(b0 - h * d)) / (2 * h)
    print("random direction: adjoint", float(grad @ d), "central differences", fd)
    assert abs(float(grad @ d) - fd) <= 1e-5 * abs(fd)


@pytest.mark.slow
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
    c, prof = energy_case(O, u_ctrl0(), NU, T_WALL)
    return mesh, _converged(mesh, c, "adjoint_heat_converged.npz"), prof


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


@pytest.mark.slow
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
