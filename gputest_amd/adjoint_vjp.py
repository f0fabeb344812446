"""Hand-written vector-Jacobian products of the SIMPLEC outer iteration at its fixed point
(no reference counterpart; the forward it differentiates is ``ResidualMap.step``, which restates
``flowtrace/unstructured/simple.py:500-703`` and, with energy, ``778-813``).

``Linearisation(rm, w, beta)`` runs one forward pass of the outer iteration Phi at the
converged state w* (no autograd tape), keeps the intermediates the transposes need
(checkpointed recomputation: nothing of the forward iterations is stored), and ``vjp(bar)``
returns (dPhi/dw)^T bar and (dPhi/dbeta)^T bar by applying the transpose of every stage in
reverse order:

* the corrections: F = F* + cf (p'_P - p'_N), Fb = Fb* + cb p'_b, p += alpha_p p', U = U* - dP
  grad p' (p' = 0 at the fixed point, so only the p' paths survive);
* the pressure solve: the transposed pressure operator is the operator itself (one AMG-PCG solve
  on the device), then the mass imbalance and the pressure-system coefficients;
* Rhie-Chow (``simple.py:583-613``): the face interpolation of U*, the d coefficient and the
  pressure-gradient correction, the inlet and outlet boundary fluxes;
* the SIMPLEC d coefficient, dP = V / (diag_r - sum a_nb), with its guard;
* the momentum solves: A^T lambda = bar U* with A^T = the face operator with each face's two
  coefficients swapped (Jacobi-BiCGSTAB on the device), then -lambda x^T on the coefficients;
* the coefficients and right-hand side (``simple.py:340-374``): the upwind switches (relu' with
  relu'(0) = 0), diffusion, boundary sources, the non-orthogonal term and the bounded
  second-order deferred correction (the clip's active bound, ties split evenly);
* the least-squares gradients (G^-1 is symmetric, so its transpose is itself) and the
  Green-Gauss gradient with its skew correction;
* the boundary values reconstructed with the lagged gradients (walls, inlets, outlets,
  symmetry projection) and the inlet profile's interpolation weights;
* with the energy equation, the same chain for T, whose coefficients read the corrected fluxes.

Every scatter is an ``index_add`` over the face owners and neighbours. ``objective_grad`` gives
dJ/dw and dJ/dbeta of the drag or the heat flow, also by hand.
"""
from __future__ import annotations

from typing import List

import torch


def _relu_d(x):
    """relu'(x) with relu'(0) = 0."""
    return (x > 0).to(x.dtype)


def _split(a, b):
    """d min(a, b) / d(a, b) and d max(a, b) / d(a, b), ties split evenly."""
    lt, eq = (a < b).to(a.dtype), (a == b).to(a.dtype)
    gt = 1.0 - lt - eq
    return (lt + 0.5 * eq, gt + 0.5 * eq), (gt + 0.5 * eq, lt + 0.5 * eq)


class Linearisation:
    def __init__(self, rm, w: List[torch.Tensor], beta: List[torch.Tensor]):
        self.rm = rm
        with torch.no_grad():
            self.w = [t.detach() for t in w]
            self.beta = [t.detach() for t in beta]
            self._forward()

    # ---------------------------------------------------------------- small operators
    def sc(self, own=None, nei=None, bnd=None):
        rm = self.rm
        ref = own if own is not None else (nei if nei is not None else bnd)
        out = torch.zeros(rm.nc, *ref.shape[1:], dtype=torch.float64, device=rm.device)
        if own is not None:
            out.index_add_(0, rm.P, own)
        if nei is not None:
            out.index_add_(0, rm.N, nei)
        if bnd is not None:
            out.index_add_(0, rm.Pb, bnd)
        return out

    def interp_t(self, gf, out):
        """Transpose of the face interpolation w phi_P + (1 - w) phi_N, added into ``out``."""
        rm = self.rm
        w = rm.w if gf.dim() == 1 else rm.w[:, None]
        out.index_add_(0, rm.P, w * gf)
        out.index_add_(0, rm.N, (1.0 - w) * gf)
        return out

    def ls_t(self, gbar, phi_bar, phib_bar):
        """Transpose of the LS gradient g = G^-1 r(phi, phib); adds into phi_bar and phib_bar."""
        rm = self.rm
        rbar = torch.einsum("cab,cb->ca", rm.Ginv, gbar)  # G^-1 symmetric
        ddphi = torch.sum(rbar[rm.P] * rm.d, dim=1) + torch.sum(rbar[rm.N] * rm.d, dim=1)
        t = rm.wl * ddphi
        phi_bar.index_add_(0, rm.N, t)
        phi_bar.index_add_(0, rm.P, -t)
        e = rm.wlb * torch.sum(rbar[rm.Pb] * rm.db, dim=1)
        phib_bar += e
        phi_bar.index_add_(0, rm.Pb, -e)

    def gg_t(self, gbar, phi_bar, phib_bar, gls_bar):
        """Transpose of the Green-Gauss gradient with skew correction; adds into the three."""
        rm = self.rm
        sbar = gbar / rm.vol[:, None]
        pfbar = torch.sum((sbar[rm.P] - sbar[rm.N]) * rm.S, dim=1)
        phib_bar += torch.sum(sbar[rm.Pb] * rm.Sb, dim=1)
        self.interp_t(pfbar, phi_bar)
        self.interp_t(pfbar[:, None] * rm.dxm, gls_bar)

    # ---------------------------------------------------------------- forward at the fixed point
    def _forward(self):
        rm, w, beta = self.rm, self.w, self.beta
        U, p, F, Fb, gp_lag, gU_lag = w[:6]
        P, N, Pb, dim = rm.P, rm.N, rm.Pb, rm.dim
        relu = torch.relu
        self.mu = rm.rho * beta[1]
        self.ub_fixed = rm._ub_fixed(beta).detach()
        self.UP = U[Pb] + torch.stack([torch.sum(gU_lag[k][Pb] * rm.dbt, dim=1) for k in range(dim)], dim=1)
        self.ub = rm._ubvals(U, gU_lag, self.ub_fixed)
        self.pb = rm._pbvals(p, gp_lag)
        self.gp_ls = rm._ls(p, self.pb)
        self.gp = rm._gg(p, self.pb, self.gp_ls) if rm.gg else self.gp_ls
        self.gU = [rm._ls(U[:, k], self.ub[:, k]) for k in range(dim)]
        D = self.mu * rm.alpha
        self.aPN, self.aNP = -(D + relu(-F)), -(D + relu(F))
        self.Db = self.mu * rm.alpha_b * rm.m_dirvel
        self.diag = self.sc(D + relu(F), D + relu(-F), self.Db + relu(Fb))
        au = rm.relax_u
        self.diag_r = self.diag / au
        if rm.simplec:
            den = self.diag_r - self.sc(-self.aPN, -self.aNP)
            self.keep = (den > 1e-12 * self.diag_r).to(torch.float64)
            self.denom = self.keep * den + (1.0 - self.keep) * self.diag_r
        else:
            self.keep = None
            self.denom = self.diag_r
        self.dP = rm.vol / self.denom
        self.dF = rm._interp(self.dP)
        self.gpf = rm._interp(self.gp)
        self.cf = rm.rho * self.dF * rm.alpha
        self.cb = rm.rho * self.dP[Pb] * rm.alpha_b * rm.m_pres
        rm._pressure_system(self.cf, self.cb)
        self.pgrad_f = p[N] - p[P] - torch.sum(self.gpf * rm.d, dim=1)
        self.pgrad_b = rm.pb_fixed - p[Pb] - torch.sum(self.gp[Pb] * rm.db, dim=1)
        if rm.energy:
            self._forward_energy()

    def _forward_energy(self):
        rm, w = self.rm, self.w
        T, gT_lag = w[6], w[7]
        F, Fb = w[2], w[3]  # the corrected fluxes equal the state's at the fixed point
        Pb = rm.Pb
        relu = torch.relu
        tval = rm._tval(self.beta)
        self.tval = tval
        self.TP = T[Pb] + torch.sum(gT_lag[Pb] * rm.dbt, dim=1)
        Tb0 = rm.m_tfix * tval + (1.0 - rm.m_tfix) * self.TP
        self.back = (rm.m_out & (Fb < 0)).to(torch.float64)
        self.Tb = self.back * tval + (1.0 - self.back) * Tb0
        self.gT = rm._ls(T, self.Tb)
        gam = rm.k_over_cp
        D = gam * rm.alpha
        self.TaPN, self.TaNP = -(D + relu(-F)), -(D + relu(F))
        self.TDb = gam * rm.alpha_b * rm.m_tfix
        self.Tdiag = self.sc(D + relu(F), D + relu(-F), self.TDb + relu(Fb))

    # ---------------------------------------------------------------- transposes
    def _deferred_t(self, phi, g, F, dcbar, phi_bar, g_bar, F_bar):
        """Transpose of dc = F (clip(ho_up, min, max) - phi_up) (``simple.py:366-372``)."""
        rm = self.rm
        P, N = rm.P, rm.N
        phP, phN = phi[P], phi[N]
        hoP = phP + torch.sum(g[P] * rm.dPf, dim=1)
        hoN = phN + torch.sum(g[N] * rm.dNf, dim=1)
        up = (F >= 0).to(torch.float64)
        x = up * hoP + (1.0 - up) * hoN
        lo, hi = torch.minimum(phP, phN), torch.maximum(phP, phN)
        y = torch.maximum(x, lo)
        ho = torch.minimum(y, hi)
        phup = up * phP + (1.0 - up) * phN
        F_bar += dcbar * (ho - phup)
        hobar = F * dcbar
        (dy_min, dhi_min), _ = _split(y, hi)
        ybar, hibar = hobar * dy_min, hobar * dhi_min
        _, (dx_max, dlo_max) = _split(x, lo)
        xbar, lobar = ybar * dx_max, ybar * dlo_max
        (dP_lo, dN_lo), (dP_hi, dN_hi) = _split(phP, phN)
        phPbar = lobar * dP_lo + hibar * dP_hi + up * xbar - up * F * dcbar
        phNbar = lobar * dN_lo + hibar * dN_hi + (1.0 - up) * xbar - (1.0 - up) * F * dcbar
        hoPbar, hoNbar = up * xbar, (1.0 - up) * xbar
        phi_bar.index_add_(0, P, phPbar)
        phi_bar.index_add_(0, N, phNbar)
        g_bar.index_add_(0, P, hoPbar[:, None] * rm.dPf)
        g_bar.index_add_(0, N, hoNbar[:, None] * rm.dNf)

    def vjp(self, bar: List[torch.Tensor]):
        """(dPhi/dw)^T bar and (dPhi/dbeta)^T bar at the fixed point: lists shaped like w and beta."""
        rm, w = self.rm, self.w
        with torch.no_grad():
            return self._vjp(rm, w, [b.detach() for b in bar])

    def _vjp(self, rm, w, bar):
        U, p, F, Fb, gp_lag, gU_lag = w[:6]
        P, N, Pb, dim = rm.P, rm.N, rm.Pb, rm.dim
        z = torch.zeros_like
        U_bar, p_bar, F_bar, Fb_bar, gpl_bar = z(U), z(p), z(F), z(Fb), z(gp_lag)
        gUl_bar = z(gU_lag)
        beta_bar = [z(b) for b in self.beta]
        ubf_bar = z(self.ub_fixed)
        Un_b, pn_b, Fn_b, Fbn_b, gp_out_b, gU_out_b = bar[:6]
        Fn_b, Fbn_b = Fn_b.clone(), Fbn_b.clone()
        if rm.energy:
            self._energy_t(bar[6], bar[7], Fn_b, Fbn_b, U_bar, beta_bar)
        # ---- corrections (p' = 0 at the fixed point, so grad p' = 0)
        if rm.singular:
            pn_b = pn_b - pn_b.mean()
        p_bar += pn_b
        pp_bar = rm.relax_p * pn_b
        Fst_bar, Fbst_bar = Fn_b, Fbn_b
        cfF = self.cf * Fn_b  # F = F* + cf (p'_P - p'_N), Fb = Fb* + cb p'_b
        pp_bar.index_add_(0, P, cfF)
        pp_bar.index_add_(0, N, -cfF)
        pp_bar.index_add_(0, Pb, self.cb * Fbn_b)
        Ustar_bar = Un_b.clone()
        gpp_bar = -self.dP[:, None] * Un_b
        ppb_bar = z(Fb)
        self.ls_t(gpp_bar, pp_bar, ppb_bar)
        pp_bar.index_add_(0, Pb, (1.0 - rm.m_pres) * ppb_bar)
        # ---- pressure solve (symmetric operator); the coefficient terms vanish with p' = 0
        lam_p = rm.solve_pressure(pp_bar)
        mass_bar = -lam_p
        Fst_bar = Fst_bar + mass_bar[P] - mass_bar[N]
        Fbst_bar = Fbst_bar + mass_bar[Pb]
        # ---- Rhie-Chow fluxes and the d coefficient
        rho = rm.rho
        dF_bar = -rho * rm.alpha * Fst_bar * self.pgrad_f
        c = rho * self.dF * rm.alpha * Fst_bar
        p_bar.index_add_(0, N, -c)
        p_bar.index_add_(0, P, c)
        gpf_bar = c[:, None] * rm.d
        Ustar_f = rho * Fst_bar[:, None] * rm.S
        self.interp_t(Ustar_f, Ustar_bar)
        ubf_bar += (rm.m_inlet * rho * Fbst_bar)[:, None] * rm.Sb
        fb = rm.m_pres * rho * Fbst_bar
        Ustar_bar.index_add_(0, Pb, fb[:, None] * rm.Sb)
        dP_bar = z(p)
        dP_bar.index_add_(0, Pb, -fb * rm.alpha_b * self.pgrad_b)
        cbp = fb * self.dP[Pb] * rm.alpha_b
        p_bar.index_add_(0, Pb, cbp)
        gp_bar = gp_out_b.clone()
        gp_bar.index_add_(0, Pb, cbp[:, None] * rm.db)
        self.interp_t(gpf_bar, gp_bar)
        self.interp_t(dF_bar, dP_bar)
        # dP = vol / denom
        den_bar = -(self.dP * self.dP / rm.vol) * dP_bar
        diag_bar = z(p)
        aPN_bar, aNP_bar = z(F), z(F)
        if self.keep is not None:
            kd = self.keep * den_bar
            diag_r_bar = kd + (1.0 - self.keep) * den_bar
            aPN_bar += kd[P]
            aNP_bar += kd[N]
        else:
            diag_r_bar = den_bar
        diag_bar += diag_r_bar / rm.relax_u
        # ---- momentum solves (U* = U at the fixed point) and right-hand sides
        au_upd = rm.update_relax_u
        cu = (1.0 - au_upd) / au_upd
        relu = torch.relu
        cbb = self.Db + relu(-Fb)
        cbb_bar = z(Fb)
        mu_bar = torch.zeros((), dtype=torch.float64, device=rm.device)
        ub_bar = z(self.ub)
        gU_bar = [gU_out_b[k].clone() for k in range(dim)]
        gpr_bar = z(gp_bar)  # gradient of the pressure source, added to gp_bar below
        for k in range(dim):
            lam = rm.solve_momentum_T(self.diag / au_upd, self.aPN, self.aNP, Ustar_bar[:, k])
            xk = U[:, k]
            diag_bar += -(lam * xk) / au_upd
            aPN_bar += -lam[P] * xk[N]
            aNP_bar += -lam[N] * xk[P]
            rhs_bar = lam
            diag_bar += cu * xk * rhs_bar
            U_bar[:, k] += cu * self.diag * rhs_bar
            rb = rhs_bar[Pb]
            cbb_bar += rb * self.ub[:, k]
            ub_bar[:, k] += cbb * rb
            tn_bar = rhs_bar[P] - rhs_bar[N]
            G = torch.sum(rm._interp(self.gU[k]) * rm.Tv, dim=1)
            mu_bar = mu_bar + torch.sum(tn_bar * G)
            self.interp_t((self.mu * tn_bar)[:, None] * rm.Tv, gU_bar[k])
            if rm.second_order:
                dc_bar = -tn_bar
                Ubk = torch.zeros_like(xk)
                self._deferred_t(xk, self.gU[k], F, dc_bar, Ubk, gU_bar[k], F_bar)
                U_bar[:, k] += Ubk
            gpr_bar[:, k] += -rm.vol * rhs_bar
        gp_bar += gpr_bar
        # ---- coefficients
        D_bar = -aPN_bar - aNP_bar
        F_bar += (F < 0).to(F.dtype) * aPN_bar - _relu_d(F) * aNP_bar
        D_bar = D_bar + diag_bar[P] + diag_bar[N]
        F_bar += _relu_d(F) * diag_bar[P] - (F < 0).to(F.dtype) * diag_bar[N]
        Db_bar = diag_bar[Pb] + cbb_bar
        Fb_bar += _relu_d(Fb) * diag_bar[Pb] - (Fb < 0).to(Fb.dtype) * cbb_bar
        mu_bar = mu_bar + torch.sum(D_bar * rm.alpha) + torch.sum(Db_bar * rm.alpha_b * rm.m_dirvel)
        # ---- gradients of this iteration
        for k in range(dim):
            ubk = torch.zeros_like(Fb)
            Ucol = torch.zeros_like(p)
            self.ls_t(gU_bar[k], Ucol, ubk)
            U_bar[:, k] += Ucol
            ub_bar[:, k] += ubk
        pb_bar = z(Fb)
        if rm.gg:
            gls_bar = z(gp_bar)
            self.gg_t(gp_bar, p_bar, pb_bar, gls_bar)
            self.ls_t(gls_bar, p_bar, pb_bar)
        else:
            self.ls_t(gp_bar, p_bar, pb_bar)
        # ---- boundary values with the lagged gradients
        rp = (1.0 - rm.m_pres) * pb_bar
        p_bar.index_add_(0, Pb, rp)
        gpl_bar.index_add_(0, Pb, rp[:, None] * rm.db)
        gpl_bar *= rm.gp_free
        ubf_bar += (rm.m_wall + rm.m_inlet)[:, None] * ub_bar
        us = rm.m_sym[:, None] * ub_bar
        UP_bar = rm.m_pres[:, None] * ub_bar + us - torch.sum(us * rm.nb, dim=1)[:, None] * rm.nb
        U_bar.index_add_(0, Pb, UP_bar)
        for k in range(dim):
            gUl_bar[k].index_add_(0, Pb, UP_bar[:, k:k + 1] * rm.dbt)
        # ---- parameters
        self._params_t(ubf_bar, mu_bar, beta_bar)
        out = [U_bar, p_bar, F_bar, Fb_bar, gpl_bar, gUl_bar]
        if rm.energy:
            out += [self._T_bar, self._gTl_bar]
        return out, beta_bar

    def _params_t(self, ubf_bar, mu_bar, beta_bar):
        rm = self.rm
        if rm.profile is not None:
            beta_bar[0] += rm.W.t() @ ubf_bar[rm.inlet_idx, 0]
        beta_bar[1] += rm.rho * mu_bar

    def _energy_t(self, Tn_bar, gT_out_bar, Fn_bar, Fbn_bar, U_bar, beta_bar):
        """Transpose of the energy step; adds the flux adjoints into Fn_bar / Fbn_bar and keeps
        the temperature and lagged-gradient adjoints for ``_vjp``."""
        rm, w = self.rm, self.w
        T, gT_lag, F, Fb = w[6], w[7], w[2], w[3]
        P, N, Pb = rm.P, rm.N, rm.Pb
        relu = torch.relu
        z = torch.zeros_like
        T_bar, gTl_bar, tval_bar = z(T), z(gT_lag), z(self.tval)
        aT = rm.relax_T
        lam = rm.solve_momentum_T(self.Tdiag / aT, self.TaPN, self.TaNP, Tn_bar)
        diag_bar = -(lam * T) / aT
        aPN_bar, aNP_bar = -lam[P] * T[N], -lam[N] * T[P]
        cT = (1.0 - aT) / aT
        diag_bar = diag_bar + cT * T * lam
        T_bar += cT * self.Tdiag * lam
        rb = lam[Pb]
        cbb = self.TDb + relu(-Fb)
        cbb_bar = rb * self.Tb
        Tb_bar = cbb * rb
        tval_bar += rm.m_tflux * rm.Sb_mag / rm.cp * rb
        tn_bar = lam[P] - lam[N]
        gT_bar = gT_out_bar.clone()
        self.interp_t((rm.k_over_cp * tn_bar)[:, None] * rm.Tv, gT_bar)
        if rm.second_order:
            self._deferred_t(T, self.gT, F, -tn_bar, T_bar, gT_bar, Fn_bar)
        Fn_bar += (F < 0).to(F.dtype) * aPN_bar - _relu_d(F) * aNP_bar
        Fn_bar += _relu_d(F) * diag_bar[P] - (F < 0).to(F.dtype) * diag_bar[N]
        Fbn_bar += _relu_d(Fb) * diag_bar[Pb] - (Fb < 0).to(Fb.dtype) * cbb_bar
        self.ls_t(gT_bar, T_bar, Tb_bar)
        tval_bar += self.back * Tb_bar
        Tb0_bar = (1.0 - self.back) * Tb_bar
        tval_bar += rm.m_tfix * Tb0_bar
        TP_bar = (1.0 - rm.m_tfix) * Tb0_bar
        T_bar.index_add_(0, Pb, TP_bar)
        gTl_bar.index_add_(0, Pb, TP_bar[:, None] * rm.dbt)
        if rm.wall_T_faces is not None and len(beta_bar) > 2:
            beta_bar[2] += torch.sum(tval_bar * rm.wall_T_faces)
        self._T_bar, self._gTl_bar = T_bar, gTl_bar

    # ---------------------------------------------------------------- objective
    def objective_grad(self):
        """J, dJ/dw (shaped like w) and dJ/dbeta of ``rm``'s objective at the state."""
        rm, w, beta = self.rm, self.w, self.beta
        with torch.no_grad():
            z = torch.zeros_like
            gw = [z(t) for t in w]
            gb = [z(b) for b in beta]
            J = rm.objective(w, beta)
            Pb = rm.Pb
            if rm.objective_kind == "heat":
                Fb, T = w[3], w[6]
                tval = rm._tval(beta)
                TP = T[Pb]
                Tb0 = rm.m_tfix * tval + (1.0 - rm.m_tfix) * TP
                back = (rm.m_out & (Fb < 0)).to(torch.float64)
                Tb = back * tval + (1.0 - back) * Tb0
                q = rm.zone_mask
                Tb_bar = q * (rm.k * rm.alpha_b * rm.m_tfix - rm.cp * Fb)
                TP_bar = -q * rm.k * rm.alpha_b * rm.m_tfix
                gw[3] += -rm.cp * Tb * q
                tval_bar = q * rm.m_tflux * rm.Sb_mag + back * Tb_bar
                Tb0_bar = (1.0 - back) * Tb_bar
                tval_bar = tval_bar + rm.m_tfix * Tb0_bar
                TP_bar = TP_bar + (1.0 - rm.m_tfix) * Tb0_bar
                gw[6].index_add_(0, Pb, TP_bar)
                if rm.wall_T_faces is not None and len(gb) > 2:
                    gb[2] += torch.sum(tval_bar * rm.wall_T_faces)
                return J, gw, gb
            mu = rm.rho * beta[1]
            U = w[0]
            ubf = rm._ub_fixed(beta)
            du = U[Pb] - ubf
            du_t = du - torch.sum(du * rm.nb, dim=1)[:, None] * rm.nb
            zm = rm.zone_mask
            pb_bar = zm * rm.Sb[:, 0]
            rp = (1.0 - rm.m_pres) * pb_bar
            gw[1].index_add_(0, Pb, rp)
            gw[4].index_add_(0, Pb, rp[:, None] * rm.db)
            dut_bar = torch.zeros_like(du)
            dut_bar[:, 0] = zm * mu * rm.alpha_b
            du_bar = dut_bar - torch.sum(dut_bar * rm.nb, dim=1)[:, None] * rm.nb
            gw[0].index_add_(0, Pb, du_bar)
            mu_bar = torch.sum(zm * rm.alpha_b * du_t[:, 0])
            self._params_t(-du_bar, mu_bar, gb)
            return J, gw, gb
