"""CPU restatement of the reference's k-omega SST model, for tests only.

Follows ``flowtrace/unstructured/turbulence.py`` (Menter 2003 SST with the automatic wall
treatment) as wired into the outer iteration by ``flowtrace/unstructured/simple.py``:
momentum viscosity and transpose stress (546-548), energy diffusivity (781-782), the k and
omega transports after the energy step (662-669). It is the checker of the device SST path
and is never imported by it.

Pinned by ``tests/golden/sst_*.npz`` (the reference's own FP64 NumPy path on a quad channel,
a heated triangle duct and a 3-D hex channel: wall distance, inlet values, per-iteration
fields, residuals, y+ and the wall-function viscosity; Spalding's law on a Re_y sweep) and
by the closed forms of ``tests/test_sst.py`` (Spalding's sublayer and log-law limits, the
exact wall distance of a straight channel, mu_t = rho k / omega before the first solve).
"""
from __future__ import annotations

import math

import numpy as np

KAPPA = 0.41       # von Karman constant of SST and of Spalding's law (turbulence.py:50)
E_WALL = 9.793     # log-law constant used for the Spalding start value (turbulence.py:55)
B_SPALDING = 5.2   # Spalding's additive constant (turbulence.py:56)
CMU = 0.09
PR_T = 0.85        # turbulent Prandtl number of the energy diffusivity (turbulence.py:59)


def spalding_uplus(re_y, iters: int = 30) -> np.ndarray:
    """u+ from Spalding's law y+ = u+ + e^{-kB}(e^{ku+} - 1 - ku+ - (ku+)^2/2 - (ku+)^3/6) with
    y+ = Re_y / u+, by Newton's method on f(u+) = Re_y/u+ - y+(u+), all faces together
    until every step is below 1e-10 u+ (``turbulence.py:65-87``)."""
    r = np.maximum(np.asarray(re_y, dtype=float), 0.0)
    kap, c = KAPPA, math.exp(-KAPPA * B_SPALDING)
    u = np.where(r < 120.0, np.sqrt(np.maximum(r, 1e-300)),
                 np.log(np.maximum(E_WALL * np.sqrt(np.maximum(r, 1.0)), 1.0)) / kap + 1.0)
    for _ in range(iters):
        x = np.minimum(kap * u, 60.0)
        e = np.exp(x)
        yp = u + c * (e - 1 - x - x ** 2 / 2 - x ** 3 / 6)
        dyp = 1 + c * kap * (e - 1 - x - x ** 2 / 2)
        uu = np.maximum(u, 1e-300)
        f = r / uu - yp
        df = -r / uu ** 2 - dyp
        step = f / df
        u = np.maximum(u - step, 0.5 * u)
        if np.all(np.abs(step) <= 1e-10 * np.maximum(u, 1e-12)):
            break
    return np.where(r > 0, u, 0.0)


def strain_rate(gU, dim):
    """|S| = sqrt(2 S_ij S_ij) of the velocity gradients (``turbulence.py:99-113``)."""
    s2 = 0.0
    for i in range(dim):
        for j in range(dim):
            s2 = s2 + (0.5 * (gU[i][:, j] + gU[j][:, i])) ** 2
    return np.sqrt(2.0 * s2)


class SST:
    """Menter's 2003 k-omega SST (``turbulence.py:116-437``, ``KOmegaSST`` on the shared
    ``TurbulenceModel`` machinery)."""

    sk1, sw1, b1 = 0.85, 0.5, 0.075
    sk2, sw2, b2 = 1.0, 0.856, 0.0828
    bstar, a1 = 0.09, 0.31
    g1 = b1 / bstar - sw1 * KAPPA ** 2 / math.sqrt(bstar)
    g2 = b2 / bstar - sw2 * KAPPA ** 2 / math.sqrt(bstar)
    relax_default = 0.8

    def __init__(self, s):
        from oracle.simplec_oracle import INLET, OUTLET, PRESSURE, WALL

        self.s = s
        c, fv = s.case, s.fv
        self.rho, self.mu = c.density, c.viscosity
        self.nu = self.mu / self.rho
        self.wall_b = s.bkind == WALL
        self.inlet_b = s.bkind == INLET
        self.open_b = (s.bkind == OUTLET) | (s.bkind == PRESSURE)
        self.yb = fv.yb
        self.y = self.wall_distance()
        ywc = np.full(s.nc, np.inf)
        np.minimum.at(ywc, fv.Pb[self.wall_b], self.yb[self.wall_b])
        self.wall_cells = np.unique(fv.Pb[self.wall_b])
        self.y_wall_cell = ywc
        self.y[self.wall_cells] = ywc[self.wall_cells]
        # the value entering at inlets and backflow faces (turbulence.py:153-168)
        speed = np.linalg.norm(s.ub_fixed, axis=1)
        if self.inlet_b.any():
            u_ref = float(speed[self.inlet_b].max())
        else:
            u_ref = float(c.turb_reference_velocity or 1.0)
        self.u_ref = max(u_ref, 1e-6)
        self.kb_in, self.sb_in = np.zeros(fv.n_bf), np.zeros(fv.n_bf)
        for z in s.mesh.boundary_zones():
            spec = s.zone_specs[z.name]
            m = fv.bzone == z.id
            inten = spec.turb_intensity if spec.turb_intensity is not None else c.turb_intensity
            ratio = spec.turb_viscosity_ratio if spec.turb_viscosity_ratio is not None else c.turb_viscosity_ratio
            vel = np.where(speed[m] > 0, speed[m], self.u_ref)
            kz = 1.5 * (inten * vel) ** 2 if spec.k is None else np.full(int(m.sum()), spec.k)
            self.kb_in[m] = np.maximum(kz, 1e-12 * self.u_ref ** 2)
            self.sb_in[m] = self.omega_from(self.kb_in[m], ratio, spec.omega)
        init = c.turb_init or {}
        k0 = init.get("k", 1.5 * (c.turb_intensity * self.u_ref) ** 2)
        w0 = init.get("omega", float(self.omega_from(np.array([k0]), c.turb_viscosity_ratio, None)[0]))
        self.k = np.full(s.nc, float(k0))
        self.w = np.full(s.nc, float(w0))
        self.k_min, self.w_min = 1e-14 * self.u_ref ** 2, 1e-14
        self.relax_k = c.relax_k if c.relax_k is not None else self.relax_default
        self.relax_w = c.relax_w if c.relax_w is not None else self.relax_default
        self.yplus = np.zeros(fv.n_bf)
        self.mut = self.rho * self.k / self.w  # before the first solve (turbulence.py:384-386)

    def omega_from(self, k, ratio, explicit):
        """omega = rho k / (mu r) for a viscosity ratio r, or an explicit value."""
        if explicit is not None:
            return np.full(np.shape(k), float(explicit))
        return self.rho * np.asarray(k) / (self.mu * ratio)

    # --- wall distance (turbulence.py:191-229) ---
    def wall_distance(self) -> np.ndarray:
        s = self.s
        mesh, geo = s.mesh, s.fv.geo
        xc = geo.cell_centre
        walls = [z.id for z in mesh.boundary_zones() if s.zone_specs[z.name].kind == "wall"]
        wf = np.flatnonzero(np.isin(mesh.face_zone, walls))
        if wf.size == 0:
            return np.full(s.nc, 1e30)
        from scipy.spatial import cKDTree

        fc = geo.face_centre[wf]
        kq = min(8, wf.size)
        _, near = cKDTree(fc).query(xc, k=kq)
        near = np.atleast_2d(near).reshape(xc.shape[0], kq)
        best = np.full(xc.shape[0], np.inf)
        if mesh.dim == 2:
            a = mesh.nodes[mesh.face_nodes[mesh.face_ptr[wf]]]
            b = mesh.nodes[mesh.face_nodes[mesh.face_ptr[wf] + 1]]
            for j in range(kq):
                f = near[:, j]
                e = b[f] - a[f]
                t = np.einsum("cd,cd->c", xc - a[f], e) / np.maximum(np.einsum("cd,cd->c", e, e), 1e-300)
                t = np.clip(t, 0, 1)
                best = np.minimum(best, np.linalg.norm(xc - (a[f] + t[:, None] * e), axis=1))
        else:
            nrm = geo.face_area_vec[wf]
            area = np.linalg.norm(nrm, axis=1)
            nrm = nrm / area[:, None]
            rad = 1.2 * np.sqrt(area / np.pi)  # the face as a disc of its area, enlarged 20 %
            for j in range(kq):
                f = near[:, j]
                v = xc - fc[f]
                vn = np.einsum("cd,cd->c", v, nrm[f])
                dt = np.linalg.norm(v - vn[:, None] * nrm[f], axis=1)
                dn = np.abs(vn)
                best = np.minimum(best, np.where(dt <= rad[f], dn, np.sqrt(dn ** 2 + (dt - rad[f]) ** 2)))
        return np.maximum(best, 1e-300)

    # --- coupling to the mean flow ---
    def wall_shear_viscosity(self, U) -> np.ndarray:
        """mu_w on every boundary face: mu Re_y / u+^2 on walls from Spalding's law for the
        wall-parallel velocity relative to the wall (``turbulence.py:399-417``); y+ as a side
        result."""
        s, fv = self.s, self.s.fv
        rel = U[fv.Pb] - s.ub_fixed
        tang = rel - np.einsum("fd,fd->f", rel, fv.nb)[:, None] * fv.nb
        re_y = np.linalg.norm(tang, axis=1) * self.yb / self.nu
        up = spalding_uplus(re_y)
        ok = (re_y > 1e-12) & (up > 0)
        mu_w = np.full(re_y.shape, self.mu)
        mu_w[ok] = self.mu * re_y[ok] / up[ok] ** 2
        yp = np.zeros(re_y.shape)
        yp[ok] = re_y[ok] / up[ok]
        self.yplus = np.where(self.wall_b, yp, 0.0)
        return np.where(self.wall_b, mu_w, self.mu)

    def wall_shear(self) -> np.ndarray:
        """Wall shear stress magnitude on every boundary face, 0 off walls (``turbulence.py:336-350``)."""
        s, fv = self.s, self.s.fv
        UP = s.U[fv.Pb]
        ut = UP - np.einsum("fd,fd->f", UP, fv.nb)[:, None] * fv.nb
        ub = s.ub_fixed
        ut = ut - (ub - np.einsum("fd,fd->f", ub, fv.nb)[:, None] * fv.nb)
        mag = np.linalg.norm(ut, axis=1)
        return np.where(self.wall_b, self.wall_shear_viscosity(s.U) * mag / np.maximum(self.yb, 1e-300), 0.0)

    def momentum_viscosity(self, U):
        """(interior faces: mu + interpolated mu_t, boundary faces: mu + mu_t of the owner, or
        mu_w on walls) (``turbulence.py:235-243``)."""
        fv = self.s.fv
        gb = np.where(self.wall_b, self.wall_shear_viscosity(U), self.mu + self.mut[fv.Pb])
        return self.mu + fv.face_interp(self.mut), gb

    def energy_diffusivity(self, k_over_cp: float):
        """k/cp + mu_t/Pr_t (no turbulent part on wall faces) (``turbulence.py:245-249``)."""
        fv = self.s.fv
        gb = k_over_cp + np.where(self.wall_b, 0.0, self.mut[fv.Pb] / PR_T)
        return k_over_cp + fv.face_interp(self.mut) / PR_T, gb

    def transpose_stress(self, gU):
        """div(mu_t (grad U)^T) per cell, explicit (``turbulence.py:251-261``)."""
        s, fv, dim = self.s, self.s.fv, self.s.dim
        mf = fv.face_interp(self.mut)
        out = np.zeros((s.nc, dim))
        for k in range(dim):
            col = np.stack([gU[j][:, k] for j in range(dim)], axis=1)  # d u_j / d x_k
            out[:, k] = fv.Cdiff @ (mf * np.sum(fv.face_interp_vec(col) * fv.S, axis=1))
        return out

    # --- the two transports (turbulence.py:263-309) ---
    def bvals(self, phi, inflow):
        """Fixed value at inlets and at open faces with backflow, zero gradient elsewhere."""
        fv = self.s.fv
        fixed = self.inlet_b | (self.open_b & (self.s.Fb < 0))
        return np.where(fixed, inflow, phi[fv.Pb]), fixed.astype(np.float64)

    def transport(self, phi, inflow, sigma, Sc, Sp, relax, floor, fixed_cells=None, fixed_vals=None):
        s, fv = self.s, self.s.fv
        phib, fixed = self.bvals(phi, inflow)
        grad = fv.gradient(phi, phib)
        gam = self.mu + sigma * self.mut
        gface = (fv.face_interp(gam), gam[fv.Pb])
        diag, aPN, aNP, Db = s._coefficients(gface, fixed)
        rhs = s._rhs(gface, phi, grad, phib, Db, Sc * fv.vol, scheme=s.case.turb_scheme)
        diag = diag + Sp * fv.vol
        if fixed_cells is not None and fixed_cells.size:
            big = 1e10 * float(np.max(diag))
            diag[fixed_cells] = diag[fixed_cells] + big
            rhs[fixed_cells] = rhs[fixed_cells] + big * fixed_vals
        A = fv.matrix(diag, aPN, aNP)
        keep = phi > floor * (1.0 + 1e-9)  # cells the equation controls (not clipped, not fixed)
        if fixed_cells is not None and fixed_cells.size:
            keep[fixed_cells] = False
        num = np.abs(rhs - A @ phi)[keep]
        den = np.abs(diag * phi)[keep]
        res = float(np.sum(num)) / max(float(np.sum(den)), 1e-300)
        dr = diag / relax
        Ar = fv.matrix(dr, aPN, aNP)
        new, info = s._solve(Ar, rhs + ((1.0 - relax) / relax) * diag * phi, phi, dr, tol=1e-3, maxiter=60)
        return np.maximum(new, floor), res, info

    def blend(self):
        """F1, F2 and grad k . grad omega (``turbulence.py:367-380``)."""
        fv = self.s.fv
        k, w, y = self.k, self.w, self.y
        gk = fv.gradient(k, self.bvals(k, self.kb_in)[0])
        gw = fv.gradient(w, self.bvals(w, self.sb_in)[0])
        kw = np.sum(gk * gw, axis=1)
        cd = np.maximum(2.0 * self.rho * self.sw2 * kw / w, 1e-10)
        sk = np.sqrt(k)
        visc = 500.0 * self.nu / (y * y * w)
        arg1 = np.minimum(np.maximum(sk / (self.bstar * w * y), visc), 4.0 * self.rho * self.sw2 * k / (cd * y * y))
        arg2 = np.maximum(2.0 * sk / (self.bstar * w * y), visc)
        return np.tanh(arg1 ** 4), np.tanh(arg2 ** 2), kw

    def omega_wall(self):
        """sqrt(omega_vis^2 + omega_log^2) in the wall cells (``turbulence.py:390-397``)."""
        yc = self.y_wall_cell[self.wall_cells]
        k = self.k[self.wall_cells]
        vis = 6.0 * self.nu / (self.b1 * yc ** 2)
        log = np.sqrt(np.maximum(k, 0.0)) / (CMU ** 0.25 * KAPPA * yc)
        return np.sqrt(vis ** 2 + log ** 2)

    def solve(self, gU):
        """k then omega, then mu_t from the new fields (``turbulence.py:419-437``)."""
        rho, k, w = self.rho, self.k, self.w
        S = strain_rate(gU, self.s.dim)
        S2 = S * S
        F1, _, kw = self.blend()
        mix = lambda a, b: F1 * a + (1.0 - F1) * b  # noqa: E731
        Pk = np.minimum(self.mut * S2, 10.0 * self.bstar * rho * k * w)
        k_new, rk, ik = self.transport(k, self.kb_in, mix(self.sk1, self.sk2), Pk, self.bstar * rho * w,
                                       self.relax_k, self.k_min)
        cross = (1.0 - F1) * 2.0 * rho * self.sw2 * kw / w
        Sc = mix(self.g1, self.g2) * rho * S2 + np.maximum(cross, 0.0)
        Sp = mix(self.b1, self.b2) * rho * w + np.maximum(-cross, 0.0) / w
        w_new, rw, iw = self.transport(w, self.sb_in, mix(self.sw1, self.sw2), Sc, Sp, self.relax_w, self.w_min,
                                       fixed_cells=self.wall_cells, fixed_vals=self.omega_wall())
        self.k, self.w = k_new, w_new
        _, F2, _ = self.blend()
        self.mut = rho * self.a1 * self.k / np.maximum(self.a1 * self.w, S * F2)
        self.last_info = {"k": ik, "omega": iw}
        return {"k": rk, "omega": rw}

    def state(self):
        return {"k": self.k.copy(), "w": self.w.copy(), "mut": self.mut.copy()}

    def load(self, st):
        if "k" in st:
            self.k = np.array(st["k"], float)
            self.w = np.array(st["w"], float)
            self.mut = np.array(st["mut"], float)
