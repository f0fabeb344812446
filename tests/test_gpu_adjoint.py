"""The discrete adjoint on the GPU (HIP forward solver; hand-written vector-Jacobian products
with the transposed momentum BiCGSTAB and pressure AMG-PCG on the device) against: torch
autograd's products through the restated iteration; the CPU adjoint on the oracle's forward;
the independent tangent reference built from the oracle alone (``oracle/tangent_oracle.py``)
along a random direction over all 65 parameters and at profile points 0 and 63; central finite
differences of the GPU forward along the random direction, at one profile point and in nu;
for the drag and (with the energy equation) the outlet heat flow; the autograd wrapper."""
import numpy as np
import pytest
import torch

import oracle.simplec_oracle as O
from oracle.adjoint_oracle import CpuResidualMap, solver_view
from oracle.tangent_oracle import TangentReference
from gputest_amd import case as C
from gputest_amd.adjoint import AdjointDrag, ResidualMap, drag_gradient, heat_gradient
from gputest_amd.adjoint_vjp import Linearisation
from gputest_amd.simple import SimpleSolver
from tests.adjoint_cases import HALF, case, energy_case, small_mesh, u_ctrl0

pytestmark = pytest.mark.gpu
NU = 1.0 / 40.0
DEV = "cuda"


def _gpu(mesh, u, nu):
    c, prof = case(C, u, nu)
    s = SimpleSolver(mesh, c, DEV)
    while not s.finished:
        s.step()
    assert s.converged
    return s, prof


def test_gpu_adjoint_matches_cpu_adjoint_and_central_differences():
    mesh, u0 = small_mesh(), u_ctrl0()
    s, prof = _gpu(mesh, u0, NU)
    gu, gnu, J, info = drag_gradient(s, prof)
    print("gpu adjoint", info, "J", J)  # noqa: T201
    assert info.residual < 1e-9
    c, prof_o = case(O, u0, NU)
    o = O.Simplec(mesh, c)
    while not o.finished:
        o.step()
    gu_c, gnu_c, J_c, info_c = drag_gradient(None, prof_o, rmap=CpuResidualMap(solver_view(o), "cylinder", "inlet",
                                                                                 prof_o), tol=1e-11)
    print("cpu adjoint", info_c, "J", J_c)
    assert abs(J - J_c) <= 1e-10 * abs(J_c)
    ref = np.r_[gu_c.numpy(), float(gnu_c)]
    got = np.r_[gu.cpu().numpy(), float(gnu)]
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("gpu vs cpu adjoint", err)
    assert err < 1e-7
    h = 1e-6

    def drag(u, nu):
        si, pi = _gpu(mesh, u, nu)
        rm = ResidualMap(si, "cylinder", "inlet", pi)
        return float(rm.objective(rm.state(), rm.params()).detach())

    k = 31
    up, um = u0.copy(), u0.copy()
    up[k] += h
    um[k] -= h
    fd = (drag(up, NU) - drag(um, NU)) / (2 * h)
    print("u", k, float(gu[k]), fd)
    assert abs(float(gu[k]) - fd) <= 1e-5 * abs(fd)
    fd = (drag(u0, NU * (1 + h)) - drag(u0, NU * (1 - h))) / (2 * h * NU)
    print("nu", float(gnu), fd)
    assert abs(float(gnu) - fd) <= 1e-5 * abs(fd)


def test_hand_written_products_equal_autograd_on_the_device():
    mesh, u0 = small_mesh(), u_ctrl0()
    for energy in (False, True):
        s, prof = _gpu_energy(mesh, u0, T_WALL) if energy else _gpu(mesh, u0, NU)
        kw = dict(objective="heat", wall_T_zone="cylinder") if energy else {}
        rm = ResidualMap(s, "outlet" if energy else "cylinder", "inlet", prof, **kw)
        w, beta = rm.state(), rm.params()
        out = rm.step(w, beta)
        gen = torch.Generator(device=DEV).manual_seed(3)
        bars = [torch.randn(t.shape, dtype=torch.float64, device=DEV, generator=gen) for t in out]
        ref = torch.autograd.grad(out, w + beta, grad_outputs=bars, allow_unused=True)
        hw, hb = Linearisation(rm, w, beta).vjp(bars)
        for a, b in zip(hw + hb, ref):
            b = torch.zeros_like(a) if b is None else b
            err = float((a - b).abs().max()) / max(float(b.abs().max()), 1e-300)
            assert err <= 1e-10, (energy, err)  # scatter order on the device (atomic adds)


def test_gpu_adjoint_matches_the_tangent_reference_and_central_differences_along_a_random_direction():
    mesh, u0 = small_mesh(), u_ctrl0()
    s, prof = _gpu(mesh, u0, NU)
    gu, gnu, J, info = drag_gradient(s, prof)
    print("gpu adjoint (hand-written products)", info)
    grad = np.r_[gu.cpu().numpy(), float(gnu)]
    c, _ = case(O, u0, NU)
    o = O.Simplec(mesh, c)
    while not o.finished:
        o.step()
    tr = TangentReference(o, u0, -HALF, HALF)
    d = np.random.default_rng(7).standard_normal(65)
    d /= np.linalg.norm(d)
    for name, v in (("random", d), ("k=0", np.eye(65)[0]), ("k=63", np.eye(65)[63])):
        t, rec = tr.directional(v)
        a = float(grad @ v)
        print(name, "gpu adjoint", a, "tangent", t, rec)
        assert abs(a - t) <= 1e-5 * max(abs(t), 1e-3 * np.abs(grad).max()), (name, a, t)
    h = 1e-6

    def drag(beta):
        si, pi = _gpu(mesh, beta[:64], beta[64])
        rm = ResidualMap(si, "cylinder", "inlet", pi)
        return float(rm.objective(rm.state(), rm.params()).detach())

    b0 = np.r_[u0, NU]
    fd = (drag(b0 + h * d) - drag(b0 - h * d)) / (2 * h)
    print("random direction: gpu adjoint", float(grad @ d), "central differences of the gpu forward", fd)
    assert abs(float(grad @ d) - fd) <= 1e-5 * abs(fd)


def test_autograd_function_returns_the_adjoint_gradient():
    mesh = small_mesh()
    holder = {}

    def make_solver(u, nu):
        c, prof = case(C, u, nu)
        holder["prof"] = prof
        return SimpleSolver(mesh, c, DEV), prof

    u = torch.tensor(u_ctrl0(), dtype=torch.float64, device=DEV, requires_grad=True)
    nu = torch.tensor(NU, dtype=torch.float64, device=DEV, requires_grad=True)
    J = AdjointDrag.apply(u, nu, make_solver)
    (2.0 * J).backward()
    s, prof = _gpu(mesh, u_ctrl0(), NU)
    gu, gnu, J2, _ = drag_gradient(s, prof)
    assert abs(float(J.detach()) - J2) <= 1e-12 * abs(J2)
    assert torch.allclose(u.grad, 2.0 * gu, rtol=1e-8, atol=0)
    assert abs(float(nu.grad) - 2.0 * float(gnu)) <= 1e-8 * abs(float(gnu))


T_WALL = 350.0


def _gpu_energy(mesh, u, T_wall):
    c, prof = energy_case(C, u, NU, T_wall)
    s = SimpleSolver(mesh, c, DEV)
    while not s.finished:
        s.step()
    assert s.converged
    return s, prof


def test_gpu_heat_adjoint_matches_cpu_adjoint_and_central_differences():
    """d(outlet heat flow)/d(inlet control values, cylinder temperature)."""
    mesh, u0 = small_mesh(), u_ctrl0()
    s, prof = _gpu_energy(mesh, u0, T_WALL)
    gu, gT, J, info = heat_gradient(s, prof)
    print("gpu heat adjoint", info, "J", J, "dJ/dT_wall", float(gT))
    assert info.residual < 1e-9
    assert abs(J - s.heat_flow()["outlet"]) <= 1e-12 * abs(J)
    c, prof_o = energy_case(O, u0, NU, T_WALL)
    o = O.Simplec(mesh, c)
    while not o.finished:
        o.step()
    rm = CpuResidualMap(solver_view(o), "outlet", "inlet", prof_o, objective="heat", wall_T_zone="cylinder")
    gu_c, gT_c, J_c, info_c = heat_gradient(None, prof_o, rmap=rm, tol=1e-11)
    print("cpu heat adjoint", info_c, "J", J_c)
    assert abs(J - J_c) <= 1e-10 * abs(J_c)
    ref = np.r_[gu_c.numpy(), float(gT_c)]
    got = np.r_[gu.cpu().numpy(), float(gT)]
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("gpu vs cpu heat adjoint", err)
    assert err < 1e-7
    h = 1e-6

    def heat(u, T_wall):
        return _gpu_energy(mesh, u, T_wall)[0].heat_flow()["outlet"]

    k = 31
    up, um = u0.copy(), u0.copy()
    up[k] += h
    um[k] -= h
    fd = (heat(up, T_WALL) - heat(um, T_WALL)) / (2 * h)
    print("u", k, float(gu[k]), fd)
    assert abs(float(gu[k]) - fd) <= 1e-5 * abs(fd)
    fd = (heat(u0, T_WALL * (1 + h)) - heat(u0, T_WALL * (1 - h))) / (2 * h * T_WALL)
    print("T_wall", float(gT), fd)
    assert abs(float(gT) - fd) <= 1e-5 * abs(fd)
