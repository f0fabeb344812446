"""Discrete adjoint of the converged SIMPLEC solution (no reference counterpart).

The steady solution w* = (U, p, F, Fb, grad p, grad U) is the fixed point of one outer
iteration w -> Phi(w; beta) of ``SimpleSolver.step`` (``flowtrace/unstructured/simple.py:
500-703``) taken with exact inner solves: the iterates of the inexact solves share that
fixed point, since at w* every inner solve starts at its own solution. For an objective
J(w; beta) (the drag on a wall zone, ``simple.py:970-1001``)

    dJ/dbeta = dJ/dbeta|_w + lam^T dPhi/dbeta,   (I - dPhi/dw)^T lam = (dJ/dw)^T.

``ResidualMap`` restates Phi in torch (FP64, on the solver's device) from the same
discretisation as the HIP kernels. The vector-Jacobian products of Phi are hand-written
(``adjoint_vjp.Linearisation``): the transpose of every stage of the outer iteration, applied in
reverse order from the intermediates of one forward pass at w*, with the two transposed
solves on the device: the momentum operator transposed by swapping each face's two
coefficients (Jacobi-BiCGSTAB on the face form) and the symmetric pressure operator
(SA-AMG-PCG). The adjoint system is solved by restarted GMRES on those products (the forward
iterations are not stored: the single step is recomputed for the backward pass). The same
products through torch autograd over ``ResidualMap.step`` (whose two solves are autograd nodes
with the same transposed solves) remain as a cross-check (``vjp="autograd"``). ``drag_gradient`` returns dJ/d(inlet profile
control values) and dJ/d(nu); ``AdjointDrag`` wraps forward solve and adjoint backward as a
``torch.autograd.Function``. With the energy equation the state also holds T and its lagged
gradient, Phi also takes the energy step (``simple.py:778-813``, after the corrections, with
the corrected fluxes), and ``heat_gradient`` returns d(heat flow through a boundary zone,
``simple.py:1003-1031``)/d(inlet profile control values, the temperature of a fixed-temperature
wall zone).

Scope: laminar, no periodic pairs or body force, velocity inlets with a profile given by
``InletProfile``.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional

import numpy as np
import torch

from . import _native as nat
from .amg import SmoothedAggregationAMG
from .krylov import Workspace, bicgstab_ldu, pcg_ldu

WALL, INLET, OUTLET, PRESSURE, SYMMETRY = range(5)
T_ADIABATIC, T_FIXED, T_FLUX = range(3)


class InletProfile:
    """Inlet velocity u_x(y) by linear interpolation of ``n`` control values at equally
    spaced heights over [y0, y1] (u_y = 0): ub = W u_ctrl on the inlet faces."""

    def __init__(self, y0: float, y1: float, values):
        self.y0, self.y1 = float(y0), float(y1)
        self.values = np.asarray(values, float).copy()
        self.y = np.linspace(self.y0, self.y1, self.values.size)

    def weights(self, yf: np.ndarray) -> np.ndarray:
        """(n_faces, n_ctrl) interpolation matrix."""
        n = self.values.size
        t = np.clip((yf - self.y0) / (self.y1 - self.y0) * (n - 1), 0.0, n - 1)
        i = np.minimum(np.floor(t).astype(np.int64), n - 2)
        a = t - i
        W = np.zeros((yf.size, n))
        W[np.arange(yf.size), i] = 1.0 - a
        W[np.arange(yf.size), i + 1] = a
        return W

    def __call__(self, xf: np.ndarray) -> np.ndarray:
        u = self.weights(xf[:, 1]) @ self.values
        return np.stack([u, np.zeros_like(u)] + [np.zeros_like(u)] * (xf.shape[1] - 2), axis=1)


# ------------------------------------------------------------------------ linear-solve nodes
class _MomentumSolve(torch.autograd.Function):
    """x = A^-1 b for A = (diag, aPN, aNP) in face form, valued at the fixed point (x given);
    backward: A^T lam = g, then -lam x^T on the coefficients."""

    @staticmethod
    def forward(ctx, diag, aPN, aNP, b, x_known, rmap):
        ctx.rmap = rmap
        ctx.save_for_backward(diag, aPN, aNP, x_known)
        return x_known.clone()

    @staticmethod
    def backward(ctx, g):
        diag, aPN, aNP, x = ctx.saved_tensors
        rm = ctx.rmap
        lam = rm.solve_momentum_T(diag, aPN, aNP, g)
        P, N = rm.P, rm.N
        return -lam * x, -lam[P] * x[N], -lam[N] * x[P], lam, None, None


class _PressureSolve(torch.autograd.Function):
    """pp = A_p^-1 rhs for A_p = (cf faces, cb boundary faces), valued at the fixed point;
    A_p is symmetric, so the backward solve is the forward solve."""

    @staticmethod
    def forward(ctx, cf, cb, rhs, x_known, rmap):
        ctx.rmap = rmap
        ctx.save_for_backward(cf, cb, x_known)
        return x_known.clone()

    @staticmethod
    def backward(ctx, g):
        cf, cb, x = ctx.saved_tensors
        rm = ctx.rmap
        lam = rm.solve_pressure(g)
        P, N, Pb = rm.P, rm.N, rm.Pb
        return -(lam[P] - lam[N]) * (x[P] - x[N]), -lam[Pb] * x[Pb], lam, None, None


@dataclass
class AdjointInfo:
    gmres_iterations: int
    residual: float
    seconds: float
    vjp_count: int


class ResidualMap:
    """One SIMPLEC outer iteration Phi(w; beta) in torch on the state of ``solver``."""

    def __init__(self, solver, zone: str = "cylinder", inlet: str = "inlet", profile: Optional[InletProfile] = None,
                 solve_tol: float = 1e-13, objective: str = "drag", wall_T_zone: Optional[str] = None):
        c = solver.case
        if getattr(c, "periodic", None) or any(c.body_force) or any(c.periodic_pressure_gradient):
            raise NotImplementedError("the adjoint covers flow without periodic pairs or forcing")
        if getattr(solver, "turb", None) is not None:
            raise NotImplementedError("the adjoint covers laminar flow")
        if objective not in ("drag", "heat"):
            raise ValueError("objective must be 'drag' or 'heat'")
        if objective == "heat" and not solver.energy:
            raise ValueError("the heat-flow objective needs the energy equation")
        self.objective_kind = objective
        self.s, self.case, self.fv = solver, c, solver.fv
        fv, dev = solver.fv, solver.device
        self.device, self.dim, self.nc = dev, solver.dim, solver.nc
        L = lambda t: t.long()  # noqa: E731
        self.P, self.N, self.Pb = L(fv.P), L(fv.N), L(fv.Pb)
        self.cface = L(fv.cface)
        self.w, self.alpha, self.wl, self.wlb = fv.w, fv.alpha, fv.wl, fv.wlb
        self.S, self.d, self.Tv, self.dxm = fv.S, fv.d, fv.Tvec, fv.dxm
        self.dPf, self.dNf = fv.dP_f, fv.dN_f
        self.Sb, self.db, self.dbt, self.nb, self.alpha_b = fv.Sb, fv.db, fv.db_t, fv.nb, fv.alpha_b
        self.vol, self.Ginv = fv.vol, fv.Ginv
        kind = torch.as_tensor(solver.bkind_h, device=dev).long()
        f = lambda m: m.to(torch.float64)  # noqa: E731
        self.m_wall, self.m_inlet = f(kind == WALL), f(kind == INLET)
        self.m_pres = f((kind == OUTLET) | (kind == PRESSURE))
        self.m_sym = f(kind == SYMMETRY)
        self.m_dirvel = f((kind == WALL) | (kind == INLET) | (kind == SYMMETRY))
        self.ub_fixed0 = torch.as_tensor(solver.ub_fixed_h, device=dev)
        self.pb_fixed = torch.as_tensor(solver.pb_fixed_h, device=dev)
        self.singular = solver.singular_p
        zb = fv.bzone
        zid = solver.mesh.zone_by_name(zone).id
        on_zone = (zb == zid) if objective == "heat" else (zb == zid) & (solver.bkind_h == WALL)
        self.zone_mask = torch.as_tensor(on_zone, device=dev).to(torch.float64)
        self.energy = bool(solver.energy)
        if self.energy:
            tk = np.asarray(solver.tkind_h)
            self.m_tfix = torch.as_tensor(tk == T_FIXED, device=dev).to(torch.float64)
            self.m_tflux = torch.as_tensor(tk == T_FLUX, device=dev).to(torch.float64)
            self.m_out = torch.as_tensor(solver.bkind_h == OUTLET, device=dev)
            self.tval0 = torch.as_tensor(np.asarray(solver.tval_h, np.float64), device=dev)
            self.k_over_cp, self.cp, self.k = float(c.conductivity / c.cp), float(c.cp), float(c.conductivity)
            self.relax_T = float(c.relax_T)
            self.Sb_mag = fv.Sb_mag
            self.wall_T_faces = None
            if wall_T_zone is not None:
                wz = solver.mesh.zone_by_name(wall_T_zone).id
                sel = (zb == wz) & (tk == T_FIXED)
                if not sel.any():
                    raise ValueError(f"zone {wall_T_zone!r} has no fixed-temperature faces")
                self.wall_T_faces = torch.as_tensor(sel, device=dev).to(torch.float64)
                self.wall_T0 = float(np.asarray(solver.tval_h)[sel][0])
        iz = solver.mesh.zone_by_name(inlet)
        self.inlet_faces = np.flatnonzero(zb == iz.id)
        self.profile = profile
        if profile is not None:
            W = profile.weights(fv.xfb_h[self.inlet_faces, 1])
            self.W = torch.as_tensor(W, device=dev)
            self.inlet_idx = torch.as_tensor(self.inlet_faces, device=dev)
        self.rho = float(c.density)
        self.relax_u, self.relax_p = float(solver.relax_u), float(solver.relax_p)
        # relaxation of the velocity update inside Phi: any value gives the same fixed point (the
        # converged fields depend on relax_u only through the d coefficient of the flux
        # interpolation, which keeps relax_u), and it sets how fast the adjoint GMRES converges
        self.update_relax_u = self.relax_u
        self.simplec = solver.algo == "SIMPLEC"
        self.second_order = c.scheme == "second-order"
        self.gg = c.pressure_gradient == "green-gauss"
        self.solve_tol = solve_tol
        self.gp_free = self._lag_free_mask()
        self.ws = Workspace(dev)
        self._amg = None
        self.vjp_count = 0

    # ---------------------------------------------------------------- state and parameters
    def state(self) -> List[torch.Tensor]:
        """The converged state as leaf tensors: U (n, dim), p, F, Fb, grad p (n, dim),
        grad U (dim, n, dim), and with energy T and grad T (n, dim)."""
        s = self.s
        U = s.U.t().contiguous().clone()
        fields = (U, s.p, s.F, s.Fb, s.gp, s.gU) + ((s.T, s.gT) if self.energy else ())
        return [t.detach().clone().requires_grad_(True) for t in fields]

    def params(self, u_ctrl=None, nu=None, wall_T=None) -> List[torch.Tensor]:
        """[inlet control values, nu] and, with a fixed-temperature wall zone, its temperature."""
        u = torch.as_tensor(self.profile.values if u_ctrl is None else u_ctrl, device=self.device,
                            dtype=torch.float64)
        v = torch.as_tensor(self.case.viscosity / self.case.density if nu is None else nu, device=self.device,
                            dtype=torch.float64)
        out = [u.detach().clone().requires_grad_(True), v.detach().clone().reshape(()).requires_grad_(True)]
        if self.energy and self.wall_T_faces is not None:
            tw = torch.as_tensor(self.wall_T0 if wall_T is None else wall_T, device=self.device, dtype=torch.float64)
            out.append(tw.detach().clone().reshape(()).requires_grad_(True))
        return out

    def _tval(self, beta):
        if self.wall_T_faces is None or len(beta) < 3:
            return self.tval0
        return self.tval0 * (1.0 - self.wall_T_faces) + beta[2] * self.wall_T_faces

    def _ub_fixed(self, beta):
        if self.profile is None:
            return self.ub_fixed0
        ux = self.W @ beta[0]
        ub = self.ub_fixed0.clone()
        col = torch.zeros(ub.shape[0], self.dim, dtype=torch.float64, device=self.device)
        col = col.index_copy(0, self.inlet_idx, torch.stack([ux] + [torch.zeros_like(ux)] * (self.dim - 1), dim=1))
        keep = torch.ones(ub.shape[0], 1, dtype=torch.float64, device=self.device)
        keep[self.inlet_idx] = 0.0
        return ub * keep + col

    def _lag_free_mask(self) -> torch.Tensor:
        """(n, 1): 0 for the cells whose lagged pressure gradient is not determined by the
        fixed point, else 1. The boundary pressure of a non-outlet face is extrapolated with
        the cell's own lagged gradient, p_P + g . db, and the Green-Gauss sum over the cell's
        faces returns sum_b S_b (db_b . g) / vol for it. When that map has eigenvalue 1 (a
        cell whose other faces all have such boundary faces, e.g. the corner triangle between
        an inlet and a symmetry plane: S_int d_int^T = vol I - sum_b S_b db_b^T), the gradient
        component along its eigenvector is carried unchanged by every outer iteration, so it is
        held fixed here instead of entering the adjoint state."""
        fv, dim = self.fv, self.dim
        Pb = fv.Pb_h
        recon = np.asarray(self.m_pres.cpu().numpy() == 0.0)
        M = np.zeros((self.nc, dim, dim))
        np.add.at(M, Pb[recon], np.einsum("fa,fb->fab", fv.Sb.cpu().numpy()[recon], fv.db.cpu().numpy()[recon]))
        M /= fv.vol_h[:, None, None]
        cnt = np.bincount(Pb[recon], minlength=self.nc)
        free = np.ones((self.nc, 1))
        for c in np.flatnonzero(cnt >= dim):
            if np.any(np.abs(np.linalg.eigvals(M[c]) - 1.0) < 1e-9):
                free[c] = 0.0
        return torch.as_tensor(free, device=self.device)

    # ---------------------------------------------------------------- discrete operators
    def _sc(self, own=None, nei=None, bnd=None):
        out = torch.zeros(self.nc, *(own if own is not None else bnd).shape[1:], dtype=torch.float64,
                          device=self.device)
        if own is not None:
            out = out.index_add(0, self.P, own)
        if nei is not None:
            out = out.index_add(0, self.N, nei)
        if bnd is not None:
            out = out.index_add(0, self.Pb, bnd)
        return out

    def _interp(self, phi):
        w = self.w if phi.dim() == 1 else self.w[:, None]
        return w * phi[self.P] + (1.0 - w) * phi[self.N]

    def _ls(self, phi, phib):
        dphi = self.wl * (phi[self.N] - phi[self.P])
        r = self._sc(dphi[:, None] * self.d, dphi[:, None] * self.d,
                     (self.wlb * (phib - phi[self.Pb]))[:, None] * self.db)
        return torch.einsum("cab,cb->ca", self.Ginv, r)

    def _gg(self, phi, phib, grad):
        pf = self._interp(phi) + torch.sum(self._interp(grad) * self.dxm, dim=1)
        s = self._sc(pf[:, None] * self.S, -(pf[:, None] * self.S), phib[:, None] * self.Sb)
        return s / self.vol[:, None]

    def _ubvals(self, U, gU, ub_fixed):
        UP = U[self.Pb] + torch.stack([torch.sum(gU[k][self.Pb] * self.dbt, dim=1) for k in range(self.dim)], dim=1)
        un = torch.sum(UP * self.nb, dim=1)
        sym = UP - un[:, None] * self.nb
        return (self.m_wall + self.m_inlet)[:, None] * ub_fixed + self.m_sym[:, None] * sym \
            + self.m_pres[:, None] * UP

    def _pbvals(self, p, gp):
        pP = p[self.Pb] + torch.sum(gp[self.Pb] * self.db, dim=1)
        return self.m_pres * self.pb_fixed + (1.0 - self.m_pres) * pP

    # ---------------------------------------------------------------- Phi and J
    def step(self, w: List[torch.Tensor], beta: List[torch.Tensor]) -> List[torch.Tensor]:
        U, p, F, Fb, gp_lag, gU_lag = w[:6]
        gp_lag = self.gp_free * gp_lag + (1.0 - self.gp_free) * gp_lag.detach()
        mu = self.rho * beta[1]
        ub_fixed = self._ub_fixed(beta)
        dim, P, N, Pb = self.dim, self.P, self.N, self.Pb
        ub = self._ubvals(U, gU_lag, ub_fixed)
        pb = self._pbvals(p, gp_lag)
        gp = self._ls(p, pb)
        if self.gg:
            gp = self._gg(p, pb, gp)
        gU = [self._ls(U[:, k], ub[:, k]) for k in range(dim)]
        src = -gp * self.vol[:, None]
        D = mu * self.alpha
        relu = torch.nn.functional.relu
        aPN, aNP = -(D + relu(-F)), -(D + relu(F))
        Db = mu * self.alpha_b * self.m_dirvel
        diag = self._sc(D + relu(F), D + relu(-F), Db + relu(Fb))
        au, au_upd = self.relax_u, self.update_relax_u
        diag_r = diag / au
        diag_u = diag / au_upd
        Ustar = []
        for k in range(dim):
            phi, g, phib = U[:, k], gU[k], ub[:, k]
            rhs = self._sc(bnd=(Db + relu(-Fb)) * phib)
            tn = mu * torch.sum(self._interp(g) * self.Tv, dim=1)
            rhs = rhs + self._sc(tn, -tn)
            if self.second_order:
                phP, phN = phi[P], phi[N]
                hoP = phP + torch.sum(g[P] * self.dPf, dim=1)
                hoN = phN + torch.sum(g[N] * self.dNf, dim=1)
                up = F >= 0
                ho = torch.minimum(torch.maximum(torch.where(up, hoP, hoN), torch.minimum(phP, phN)),
                                   torch.maximum(phP, phN))
                dc = F * (ho - torch.where(up, phP, phN))
                rhs = rhs - self._sc(dc, -dc)
            rhs = rhs + src[:, k]
            b = rhs + ((1.0 - au_upd) / au_upd) * diag * phi
            Ustar.append(_MomentumSolve.apply(diag_u, aPN, aNP, b, phi.detach(), self))
        Ustar = torch.stack(Ustar, dim=1)
        if self.simplec:
            denom = diag_r - self._sc(-aPN, -aNP)
            denom = torch.where(denom > 1e-12 * diag_r, denom, diag_r)
        else:
            denom = diag_r
        dP = self.vol / denom
        dF = self._interp(dP)
        gpf = self._interp(gp)
        rho = self.rho
        Fst = rho * (torch.sum(self._interp(Ustar) * self.S, dim=1)
                     - dF * self.alpha * (p[N] - p[P] - torch.sum(gpf * self.d, dim=1)))
        Fb_in = rho * torch.sum(ub_fixed * self.Sb, dim=1) * self.m_inlet
        Fb_pr = rho * (torch.sum(Ustar[Pb] * self.Sb, dim=1)
                       - dP[Pb] * self.alpha_b * (self.pb_fixed - p[Pb] - torch.sum(gp[Pb] * self.db, dim=1)))
        Fbst = Fb_in + Fb_pr * self.m_pres
        cf = rho * dF * self.alpha
        cb = rho * dP[Pb] * self.alpha_b * self.m_pres
        mass = self._sc(Fst, -Fst, Fbst)
        self._pressure_system(cf.detach(), cb.detach())
        pp = _PressureSolve.apply(cf, cb, -mass, torch.zeros_like(mass), self)
        F_new = Fst + cf * (pp[P] - pp[N])
        Fb_new = Fbst + cb * pp[Pb]
        p_new = p + self.relax_p * pp
        if self.singular:
            p_new = p_new - p_new.mean()
        gpp = self._ls(pp, (1.0 - self.m_pres) * pp[Pb])
        U_new = Ustar - dP[:, None] * gpp
        out = [U_new, p_new, F_new, Fb_new, gp, torch.stack(gU, dim=0)]
        if self.energy:
            out += self._energy(w[6], w[7], F_new, Fb_new, beta)
        return out

    def _energy(self, T, gT_lag, F, Fb, beta):
        """The energy step (``simple.py:778-813``) with the corrected fluxes: boundary
        temperatures (fixed faces, backflow at outlets, else the owner value reconstructed with
        the lagged gradient), the LS gradient, upwind convection with the deferred second-order
        correction, k/cp diffusion with the non-orthogonal term, the heat-flux source, relaxation,
        and the exact solve."""
        P, N, Pb = self.P, self.N, self.Pb
        relu = torch.nn.functional.relu
        tval = self._tval(beta)
        TP = T[Pb] + torch.sum(gT_lag[Pb] * self.dbt, dim=1)
        Tb = self.m_tfix * tval + (1.0 - self.m_tfix) * TP
        back = (self.m_out & (Fb < 0)).to(torch.float64)
        Tb = back * tval + (1.0 - back) * Tb
        gT = self._ls(T, Tb)
        gam = self.k_over_cp
        D = gam * self.alpha
        aPN, aNP = -(D + relu(-F)), -(D + relu(F))
        Db = gam * self.alpha_b * self.m_tfix
        diag = self._sc(D + relu(F), D + relu(-F), Db + relu(Fb))
        rhs = self._sc(bnd=(Db + relu(-Fb)) * Tb)
        tn = gam * torch.sum(self._interp(gT) * self.Tv, dim=1)
        rhs = rhs + self._sc(tn, -tn)
        if self.second_order:
            phP, phN = T[P], T[N]
            hoP = phP + torch.sum(gT[P] * self.dPf, dim=1)
            hoN = phN + torch.sum(gT[N] * self.dNf, dim=1)
            up = F >= 0
            ho = torch.minimum(torch.maximum(torch.where(up, hoP, hoN), torch.minimum(phP, phN)),
                               torch.maximum(phP, phN))
            dc = F * (ho - torch.where(up, phP, phN))
            rhs = rhs - self._sc(dc, -dc)
        rhs = rhs + self._sc(bnd=self.m_tflux * tval * self.Sb_mag / self.cp)
        aT = self.relax_T
        b = rhs + ((1.0 - aT) / aT) * diag * T
        T_new = _MomentumSolve.apply(diag / aT, aPN, aNP, b, T.detach(), self)
        return [T_new, gT]

    def objective(self, w: List[torch.Tensor], beta: List[torch.Tensor]) -> torch.Tensor:
        """Drag: x-component of the pressure and viscous force on the wall zone
        (``simple.py:970-1001``); or the heat flow into the domain through the zone, conduction
        plus enthalpy (``simple.py:1003-1031``)."""
        if self.objective_kind == "heat":
            return self._heat_flow(w, beta)
        U, p, F, Fb, gp, gU = w[:6]
        mu = self.rho * beta[1]
        pb = self._pbvals(p, gp)
        du = U[self.Pb] - self._ub_fixed(beta)
        du_t = du - torch.sum(du * self.nb, dim=1)[:, None] * self.nb
        fx = pb * self.Sb[:, 0] + mu * self.alpha_b * du_t[:, 0]
        return torch.sum(self.zone_mask * fx)

    def _heat_flow(self, w, beta):
        Fb, T = w[3], w[6]
        tval = self._tval(beta)
        TP = T[self.Pb]
        Tb = self.m_tfix * tval + (1.0 - self.m_tfix) * TP
        back = (self.m_out & (Fb < 0)).to(torch.float64)
        Tb = back * tval + (1.0 - back) * Tb
        cond = self.k * self.alpha_b * (Tb - TP) * self.m_tfix + self.m_tflux * tval * self.Sb_mag
        q = cond - self.cp * Fb * Tb
        return torch.sum(self.zone_mask * q)

    # ---------------------------------------------------------------- transposed solves
    def _off(self, a_own, a_nei):
        """Half-edge coefficients of the face form: row P holds a_own, row N holds a_nei."""
        f = self.cface >> 1
        return torch.where((self.cface & 1).bool(), a_nei[f], a_own[f])

    def solve_momentum_T(self, diag, aPN, aNP, g):
        """A^T lam = g: A^T has aNP in the owner rows and aPN in the neighbour rows."""
        diag = diag.detach().contiguous()
        off = self._off(aNP.detach(), aPN.detach()).contiguous()
        inv = (1.0 / diag).contiguous()
        lam = torch.zeros_like(g)
        b = g.detach().contiguous()
        if float(torch.linalg.vector_norm(b)) == 0.0:
            return lam
        info = bicgstab_ldu(self.fv, diag, off, b, inv, lam, max(self.solve_tol, 1e-12), 5000, True, ws=self.ws)
        if not info.converged and not info.residual < 1e-10:
            raise RuntimeError(f"transposed momentum solve did not converge: {info}")
        return lam

    def _pressure_system(self, cf, cb):
        if self._amg is not None:
            return
        diag = self._sc(cf, cf, cb).contiguous()
        off = self._off(-cf, -cf).contiguous()
        self._p_diag, self._p_off = diag, off
        self._amg = SmoothedAggregationAMG.from_ldu(self.fv, diag, off, singular=self.singular)

    def solve_pressure(self, g):
        b = g.detach().contiguous()
        if self.singular:
            b = b - b.mean()
        x = torch.zeros_like(b)
        if float(torch.linalg.vector_norm(b)) == 0.0:
            return x
        info = pcg_ldu(self.fv, self._p_diag, self._p_off, b, x, self.solve_tol, 2000, self._amg, self.singular,
                       ws=self.ws)
        if not info.converged and not info.residual < 1e-10:
            raise RuntimeError(f"pressure solve did not converge: {info}")
        return x


def _flat(ts):
    return torch.cat([t.reshape(-1) for t in ts])


def _unflat(v, like):
    out, o = [], 0
    for t in like:
        n = t.numel()
        out.append(v[o:o + n].reshape(t.shape))
        o += n
    return out


def gmres(apply: Callable[[torch.Tensor], torch.Tensor], b: torch.Tensor, tol: float = 1e-10, restart: int = 60,
          maxiter: int = 600):
    """Restarted GMRES (modified Gram-Schmidt) from x = 0; returns (x, iterations, relative residual)."""
    x = torch.zeros_like(b)
    bn = float(torch.linalg.vector_norm(b))
    if bn == 0.0:
        return x, 0, 0.0
    total, rel = 0, 1.0
    while total < maxiter:
        r = b - apply(x) if total else b.clone()
        beta = float(torch.linalg.vector_norm(r))
        rel = beta / bn
        if rel < tol:
            break
        m = min(restart, maxiter - total)
        V = [r / beta]
        H = torch.zeros(m + 1, m, dtype=torch.float64)
        g = torch.zeros(m + 1, dtype=torch.float64)
        g[0] = beta
        cs, sn = torch.zeros(m, dtype=torch.float64), torch.zeros(m, dtype=torch.float64)
        k_done = 0
        for k in range(m):
            w = apply(V[k])
            for j in range(k + 1):
                H[j, k] = float(torch.dot(w, V[j]))
                w = w - H[j, k] * V[j]
            hk1 = float(torch.linalg.vector_norm(w))
            H[k + 1, k] = hk1
            for j in range(k):
                t = cs[j] * H[j, k] + sn[j] * H[j + 1, k]
                H[j + 1, k] = -sn[j] * H[j, k] + cs[j] * H[j + 1, k]
                H[j, k] = t
            den = math.hypot(float(H[k, k]), float(H[k + 1, k]))
            cs[k], sn[k] = H[k, k] / den, H[k + 1, k] / den
            H[k, k] = den
            H[k + 1, k] = 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            k_done = k + 1
            total += 1
            rel = abs(float(g[k + 1])) / bn
            if rel < tol or hk1 == 0.0:
                break
            V.append(w / hk1)
        y = torch.linalg.solve_triangular(H[:k_done, :k_done], g[:k_done, None], upper=True)[:, 0]
        for j in range(k_done):
            x = x + float(y[j]) * V[j]
        if rel < tol:
            r = b - apply(x)
            rel = float(torch.linalg.vector_norm(r)) / bn
            if rel < tol * 10:
                break
    return x, total, rel


def drag_gradient(solver, profile: InletProfile, zone: str = "cylinder", inlet: str = "inlet", tol: float = 1e-10,
                  restart: int = 200, maxiter: int = 1000, rmap: Optional[ResidualMap] = None, vjp: str = "hand"):
    """dJ/d(u_ctrl) and dJ/d(nu) of the drag J on ``zone`` at the converged state of
    ``solver`` (run with ``profile`` as its inlet and nu = viscosity / density).
    ``vjp``: "hand" (the hand-written transposes of ``adjoint_vjp``) or "autograd" (torch
    autograd through ``ResidualMap.step``, kept as a cross-check).
    Returns (grad_u (n_ctrl,), grad_nu, J, AdjointInfo)."""
    rm = rmap if rmap is not None else ResidualMap(solver, zone, inlet, profile)
    grads, J, info = _GRADIENT[vjp](rm, tol, restart, maxiter)
    return grads[0], grads[1], J, info


def heat_gradient(solver, profile: InletProfile, zone: str = "outlet", wall_T_zone: str = "cylinder",
                  inlet: str = "inlet", tol: float = 1e-10, restart: int = 200, maxiter: int = 1000,
                  rmap: Optional[ResidualMap] = None, vjp: str = "hand"):
    """dJ/d(u_ctrl) and dJ/d(T_wall) of the heat flow J into the domain through ``zone`` at the
    converged state of ``solver`` (energy on; ``wall_T_zone`` a fixed-temperature wall).
    Returns (grad_u (n_ctrl,), grad_T_wall, J, AdjointInfo)."""
    rm = rmap if rmap is not None else ResidualMap(solver, zone, inlet, profile, objective="heat",
                                                   wall_T_zone=wall_T_zone)
    grads, J, info = _GRADIENT[vjp](rm, tol, restart, maxiter)
    return grads[0], grads[2], J, info


def _gradient_hand(rm, tol, restart, maxiter):
    """The adjoint gradient with the hand-written transposes: one linearisation at the fixed
    point, J_w and J_beta by hand, GMRES on lam = J_w^T + Phi_w^T lam, then J_beta + Phi_beta^T lam."""
    from .adjoint_vjp import Linearisation

    t0 = time.perf_counter()
    with torch.no_grad():
        w, beta = rm.state(), rm.params()
        lin = Linearisation(rm, w, beta)
        J, gJw, gJb = lin.objective_grad()
        shapes = lin.w

        def vjp_w(v):
            rm.vjp_count += 1
            return _flat(lin.vjp(_unflat(v, shapes))[0])

        lam, its, rel = gmres(lambda v: v - vjp_w(v), _flat(gJw), tol=tol, restart=restart, maxiter=maxiter)
        gb = lin.vjp(_unflat(lam, shapes))[1]
        grads = [gJb[i] + gb[i] for i in range(len(beta))]
    if rm.device.type == "cuda":
        torch.cuda.synchronize(rm.device)
    info = AdjointInfo(its, rel, time.perf_counter() - t0, rm.vjp_count)
    return [g.detach() for g in grads], float(J), info


def _gradient(rm, tol, restart, maxiter):
    with torch.enable_grad():  # also when called from an autograd backward pass
        return _gradient_autograd(rm, tol, restart, maxiter)


def _gradient_autograd(rm, tol, restart, maxiter):
    t0 = time.perf_counter()
    w = rm.state()
    beta = rm.params()
    # Phi and J once at the fixed point: their graphs are reused by every product
    out = rm.step(w, beta)
    J = rm.objective(w, beta)
    gJ = torch.autograd.grad(J, w + beta, retain_graph=True, allow_unused=True)
    gJw = _flat([g if g is not None else torch.zeros_like(t) for g, t in zip(gJ[:len(w)], w)])
    gJb = [g if g is not None else torch.zeros_like(t) for g, t in zip(gJ[len(w):], beta)]
    out_flat = out

    def vjp_w(v):
        rm.vjp_count += 1
        vs = _unflat(v, w)
        g = torch.autograd.grad(out_flat, w, grad_outputs=vs, retain_graph=True, allow_unused=True)
        return _flat([gi if gi is not None else torch.zeros_like(t) for gi, t in zip(g, w)])

    lam, its, rel = gmres(lambda v: v - vjp_w(v), gJw, tol=tol, restart=restart, maxiter=maxiter)
    gb = torch.autograd.grad(out_flat, beta, grad_outputs=_unflat(lam, w), retain_graph=False, allow_unused=True)
    grads = [gJb[i] + (gb[i] if gb[i] is not None else 0.0) for i in range(len(beta))]
    if rm.device.type == "cuda":
        torch.cuda.synchronize(rm.device)
    info = AdjointInfo(its, rel, time.perf_counter() - t0, rm.vjp_count)
    return [g.detach() if torch.is_tensor(g) else torch.as_tensor(g) for g in grads], float(J.detach()), info


_GRADIENT = {"hand": _gradient_hand, "autograd": _gradient}


class AdjointDrag(torch.autograd.Function):
    """drag = AdjointDrag.apply(u_ctrl, nu, make_solver): forward runs ``make_solver(profile,
    nu)`` to convergence and returns the drag; backward returns the discrete-adjoint
    gradients with respect to u_ctrl and nu."""

    @staticmethod
    def forward(ctx, u_ctrl, nu, make_solver, zone="cylinder"):
        s, profile = make_solver(u_ctrl.detach().cpu().numpy(), float(nu))
        while not s.finished:
            s.step()
        ctx.solver, ctx.profile, ctx.zone = s, profile, zone
        rm = ResidualMap(s, zone, "inlet", profile)
        J = rm.objective(rm.state(), rm.params())
        return J.detach().to(u_ctrl.device)

    @staticmethod
    def backward(ctx, g):
        gu, gnu, _, info = drag_gradient(ctx.solver, ctx.profile, ctx.zone)
        ctx.info = info
        return g * gu.to(g.device), g * gnu.to(g.device), None, None


__all__ = ["InletProfile", "ResidualMap", "drag_gradient", "heat_gradient", "AdjointDrag", "gmres", "AdjointInfo"]
