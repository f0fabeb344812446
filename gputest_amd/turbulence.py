"""k-omega SST (Menter 2003) of the device solver (``flowtrace/unstructured/turbulence.py:116-437``).

``DeviceSST`` holds k, omega and mu_t on the device next to ``SimpleSolver``'s fields and does
the model's three jobs in the outer iteration (``flowtrace/unstructured/simple.py``):

* momentum (546-548): face viscosities mu + mu_t, the Spalding wall-function viscosity on wall
  faces, and the explicit transpose stress div(mu_t (grad U)^T) as a source;
* energy (781-782): the diffusivity k/cp + mu_t / Pr_t;
* after the energy step (662-669): the k and omega transports and the new mu_t.

Set-up happens once per mesh on the host: the wall distance (a nearest-wall-face search, exact
for 2-D segments, a disc model in 3-D, ``turbulence.py:191-229``), the wall cells with their
distance, and the inflow values of k and omega per boundary face (``turbulence.py:153-168``).
Everything per iteration runs in ``ftmi_sst_*`` kernels (``ftmi_turb.hip``), the LS gradient
kernel and the LDU BiCGSTAB.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _native as nat

WALL, INLET, OUTLET, PRESSURE, SYMMETRY = range(5)
KAPPA = 0.41
SK1, SW1, B1 = 0.85, 0.5, 0.075
SK2, SW2, B2 = 1.0, 0.856, 0.0828
BSTAR, A1 = 0.09, 0.31
G1 = B1 / BSTAR - SW1 * KAPPA ** 2 / math.sqrt(BSTAR)
G2 = B2 / BSTAR - SW2 * KAPPA ** 2 / math.sqrt(BSTAR)
RELAX_DEFAULT = 0.8


def wall_distance(mesh, geo, wall_zone_ids) -> np.ndarray:
    """Distance from every cell centre to the nearest wall face: the 8 nearest wall-face
    centres are candidates; in 2-D the distance to each candidate segment is exact, in 3-D a
    face is a disc of its area enlarged by 20 % (``turbulence.py:191-229``)."""
    xc = geo.cell_centre
    wf = np.flatnonzero(np.isin(mesh.face_zone, list(wall_zone_ids)))
    if wf.size == 0:
        return np.full(mesh.n_cells, 1e30)
    from scipy.spatial import cKDTree

    fc = geo.face_centre[wf]
    kq = min(8, wf.size)
    _, cand = cKDTree(fc).query(xc, k=kq)
    cand = np.atleast_2d(cand).reshape(xc.shape[0], kq)
    best = np.full(xc.shape[0], np.inf)
    if mesh.dim == 2:
        p0 = mesh.nodes[mesh.face_nodes[mesh.face_ptr[wf]]]
        p1 = mesh.nodes[mesh.face_nodes[mesh.face_ptr[wf] + 1]]
        for j in range(kq):
            f = cand[:, j]
            seg = p1[f] - p0[f]
            t = np.clip(np.einsum("cd,cd->c", xc - p0[f], seg) / np.maximum(np.einsum("cd,cd->c", seg, seg), 1e-300),
                        0, 1)
            best = np.minimum(best, np.linalg.norm(xc - (p0[f] + t[:, None] * seg), axis=1))
    else:
        nrm = geo.face_area_vec[wf]
        area = np.linalg.norm(nrm, axis=1)
        nrm = nrm / area[:, None]
        radius = np.sqrt(area / np.pi) * 1.2
        for j in range(kq):
            f = cand[:, j]
            v = xc - fc[f]
            vn = np.einsum("cd,cd->c", v, nrm[f])
            tang = np.linalg.norm(v - vn[:, None] * nrm[f], axis=1)
            dn = np.abs(vn)
            best = np.minimum(best, np.where(tang <= radius[f], dn, np.sqrt(dn ** 2 + (tang - radius[f]) ** 2)))
    return np.maximum(best, 1e-300)


class DeviceSST:
    """k, omega and mu_t of one ``SimpleSolver`` and the SST update on the device."""

    name = "k-omega-sst"

    def __init__(self, solver):
        s = self.s = solver
        c, fv, dev = s.case, s.fv, s.device
        if s.mixed:
            raise NotImplementedError("k-omega SST runs in FP64 (precision='fp64')")
        self.rho, self.mu = float(c.density), float(c.viscosity)
        self.nu = self.mu / self.rho
        kind = s.bkind_h
        self.wall_b = kind == WALL
        inlet_b = kind == INLET
        walls = [z.id for z in s.mesh.boundary_zones() if s.zone_specs[z.name].kind == "wall"]
        y = wall_distance(s.mesh, fv.geometry, walls)
        yb = fv.yb_h
        ywc = np.full(s.nc, np.inf)
        np.minimum.at(ywc, fv.Pb_h[self.wall_b], yb[self.wall_b])
        self.wall_cells = np.unique(fv.Pb_h[self.wall_b]).astype(np.int32)
        y[self.wall_cells] = ywc[self.wall_cells]
        self.y_h = y
        # inflow values (turbulence.py:153-168)
        speed = np.linalg.norm(s.ub_fixed_h, axis=1)
        u_ref = float(speed[inlet_b].max()) if inlet_b.any() else float(c.turb_reference_velocity or 1.0)
        self.u_ref = max(u_ref, 1e-6)
        kin, win = np.zeros(fv.n_bf), np.zeros(fv.n_bf)
        for z in s.mesh.boundary_zones():
            spec = s.zone_specs[z.name]
            m = fv.bzone == z.id
            inten = spec.turb_intensity if getattr(spec, "turb_intensity", None) is not None else c.turb_intensity
            ratio = spec.turb_viscosity_ratio if getattr(spec, "turb_viscosity_ratio", None) is not None \
                else c.turb_viscosity_ratio
            vel = np.where(speed[m] > 0, speed[m], self.u_ref)
            kz = 1.5 * (inten * vel) ** 2 if getattr(spec, "k", None) is None else np.full(int(m.sum()), spec.k)
            kin[m] = np.maximum(kz, 1e-12 * self.u_ref ** 2)
            win[m] = self._omega_from(kin[m], ratio, getattr(spec, "omega", None))
        init = c.turb_init or {}
        k0 = init.get("k", 1.5 * (c.turb_intensity * self.u_ref) ** 2)
        w0 = init.get("omega", float(self._omega_from(np.array([k0]), c.turb_viscosity_ratio, None)[0]))
        k_h, w_h = np.full(s.nc, float(k0)), np.full(s.nc, float(w0))
        self.kmin, self.wmin = 1e-14 * self.u_ref ** 2, 1e-14
        self.relax_k = float(c.relax_k if c.relax_k is not None else RELAX_DEFAULT)
        self.relax_w = float(c.relax_w if c.relax_w is not None else RELAX_DEFAULT)
        # device state and work
        f64 = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float64), device=dev)  # noqa: E731
        z = lambda *sh: torch.zeros(*sh, dtype=torch.float64, device=dev)  # noqa: E731
        nc, nf, nb, dim = s.nc, fv.n_if, fv.n_bf, s.dim
        self.k, self.w = f64(k_h), f64(w_h)
        self.mut = f64(self.rho * k_h / w_h)  # before the first solve (turbulence.py:384-386)
        self.y, self.kin, self.win = f64(y), f64(kin), f64(win)
        self.kb, self.wb, self.fixed, self.yplus = z(nb), z(nb), z(nb), z(nb)
        self.gk, self.gw = z(nc, dim), z(nc, dim)
        self.S, self.gam_k, self.sck, self.spk = z(nc), z(nc), z(nc), z(nc)
        self.gam_w, self.scw, self.spw = z(nc), z(nc), z(nc)
        self.diag, self.rhs, self.diag_r, self.b, self.inv_r = z(nc), z(nc), z(nc), z(nc), z(nc)
        self.off = z(2 * nf)
        self.xk, self.xw = z(nc), z(nc)
        self.work = z(int(nat.lib().ftmi_sst_work_doubles()))
        held = np.zeros(nc, np.uint8)
        held[self.wall_cells] = 1
        self.held = torch.as_tensor(held, device=dev)
        self.cells = torch.as_tensor(self.wall_cells, device=dev)
        self.ywc = f64(ywc[self.wall_cells])
        self.gm_f, self.gm_b, self.ge_f, self.ge_b = z(nf), z(nb), z(nf), z(nb)
        self.tstress = z(dim, nc)
        self.gUc, self.ubc, self.pbc = z(dim, nc, dim), z(dim, nb), z(nb)
        self.iters = torch.zeros(1, dtype=torch.int32, device=dev)
        cf = self.coef = nat.FtmiSstCoef()
        cf.rho, cf.mu = self.rho, self.mu
        cf.two_rho_sw2 = 2.0 * self.rho * SW2
        cf.four_rho_sw2 = 4.0 * self.rho * SW2
        cf.nu500 = 500.0 * self.nu
        cf.ten_bstar_rho = 10.0 * BSTAR * self.rho
        cf.bstar_rho = BSTAR * self.rho
        cf.rho_a1 = self.rho * A1
        cf.g1, cf.g2 = G1, G2
        self.last_info = None

    def _omega_from(self, k, ratio, explicit):
        if explicit is not None:
            return np.full(np.shape(k), float(explicit))
        return self.rho * np.asarray(k) / (self.mu * ratio)

    # ------------------------------------------------------------------ coupling
    def momentum_terms(self, L, M, B, s):
        """Face viscosities (mu + mu_t, wall-function mu_w on walls) and the transpose stress of
        the iteration's velocity and gradients (``turbulence.py:235-243, 251-261, 399-417``)."""
        sv = self.s
        nat.check(L.ftmi_sst_face_gamma(M, B, nat.ptr(self.mut), self.mu, 0, nat.ptr(self.gm_f), nat.ptr(self.gm_b),
                                        s), "sst face viscosity")
        nat.check(L.ftmi_sst_wall(M, B, nat.ptr(sv.U), nat.ptr(sv.fv.yb), self.mu, self.nu, nat.ptr(self.iters),
                                  nat.ptr(self.gm_b), nat.ptr(self.yplus), s), "sst wall function")
        nat.check(L.ftmi_sst_tstress(M, nat.ptr(self.mut), nat.ptr(sv.gU), nat.ptr(self.tstress), s),
                  "sst transpose stress")
        return self.gm_f, self.gm_b, [self.tstress[q] for q in range(sv.dim)]

    def energy_terms(self, L, M, B, s, k_over_cp: float):
        """k/cp + mu_t / Pr_t on faces, k/cp on wall faces (``turbulence.py:245-249``)."""
        nat.check(L.ftmi_sst_face_gamma(M, B, nat.ptr(self.mut), float(k_over_cp), 1, nat.ptr(self.ge_f),
                                        nat.ptr(self.ge_b), s), "sst face diffusivity")
        return self.ge_f, self.ge_b

    # ------------------------------------------------------------------ the update
    def solve(self, L, M, B, s, resid_k, resid_w):
        """k, then omega, then mu_t (``turbulence.py:419-437``) from the corrected velocity's
        gradients; resid_k / resid_w: device slots of 3 doubles (sum |r|, sum |diag phi|, max
        diag) of each transport."""
        from .krylov import bicgstab_ldu

        sv, fv, dim = self.s, self.s.fv, self.s.dim
        # gradients of the corrected velocity, boundary values with the iteration's gradients
        nat.check(L.ftmi_simplec_bvals(M, B, nat.ptr(sv.U), nat.ptr(sv.p), nat.ptr(sv.gU), None, nat.ptr(self.ubc),
                                       nat.ptr(self.pbc), s), "sst velocity bvals")
        arr = ctypes.c_void_p * dim
        nat.check(L.ftmi_ls_gradients(M, dim, arr(*[sv.U[q].data_ptr() for q in range(dim)]),
                                      arr(*[self.ubc[q].data_ptr() for q in range(dim)]),
                                      arr(*[self.gUc[q].data_ptr() for q in range(dim)]), None, s), "sst gU")
        nat.check(L.ftmi_sst_bvals(M, B, nat.ptr(sv.Fb), nat.ptr(self.k), nat.ptr(self.w), nat.ptr(self.kin),
                                   nat.ptr(self.win), nat.ptr(self.kb), nat.ptr(self.wb), nat.ptr(self.fixed), s),
                  "sst bvals")
        two = ctypes.c_void_p * 2
        nat.check(L.ftmi_ls_gradients(M, 2, two(self.k.data_ptr(), self.w.data_ptr()),
                                      two(self.kb.data_ptr(), self.wb.data_ptr()),
                                      two(self.gk.data_ptr(), self.gw.data_ptr()), None, s), "sst gradients")
        nat.check(L.ftmi_sst_cell(M, ctypes.byref(self.coef), nat.ptr(self.gUc), nat.ptr(self.gk), nat.ptr(self.gw),
                                  nat.ptr(self.k), nat.ptr(self.w), nat.ptr(self.mut), nat.ptr(self.y),
                                  nat.ptr(self.S), nat.ptr(self.gam_k), nat.ptr(self.sck), nat.ptr(self.spk),
                                  nat.ptr(self.gam_w), nat.ptr(self.scw), nat.ptr(self.spw), s), "sst cell")
        infos = {}
        for name, phi, grad, phib, gam, sc, sp, held, floor, relax, x, out in (
                ("k", self.k, self.gk, self.kb, self.gam_k, self.sck, self.spk, None, self.kmin, self.relax_k,
                 self.xk, resid_k),
                ("omega", self.w, self.gw, self.wb, self.gam_w, self.scw, self.spw, self.held, self.wmin,
                 self.relax_w, self.xw, resid_w)):
            nat.check(L.ftmi_sst_assemble(M, nat.ptr(sv.F), nat.ptr(sv.Fb), nat.ptr(gam), nat.ptr(phi), nat.ptr(grad),
                                          nat.ptr(phib), nat.ptr(self.fixed), nat.ptr(sc), nat.ptr(sp), nat.ptr(held),
                                          float(floor), float(relax), nat.ptr(self.diag), nat.ptr(self.rhs),
                                          nat.ptr(self.diag_r), nat.ptr(self.off), nat.ptr(self.b),
                                          nat.ptr(self.inv_r), nat.ptr(x), nat.ptr(self.work), nat.ptr(out), s),
                      f"sst {name} assembly")
            if held is not None:
                nat.check(L.ftmi_sst_hold(int(self.wall_cells.size), nat.ptr(self.cells), nat.ptr(self.ywc),
                                          nat.ptr(self.k), nat.ptr(phi), nat.ptr(out), self.nu, float(relax),
                                          nat.ptr(self.diag), nat.ptr(self.rhs), nat.ptr(self.diag_r),
                                          nat.ptr(self.b), nat.ptr(self.inv_r), s), "sst omega wall cells")
            infos[name] = bicgstab_ldu(fv, self.diag_r, self.off, self.b, self.inv_r, x, 1e-3, 60, True, ws=sv.ws)
        nat.check(L.ftmi_sst_update(sv.nc, nat.ptr(self.xk), nat.ptr(self.xw), float(self.kmin), float(self.wmin),
                                    nat.ptr(self.y), nat.ptr(self.S), ctypes.byref(self.coef), nat.ptr(self.k),
                                    nat.ptr(self.w), nat.ptr(self.mut), s), "sst update")
        self.last_info = infos
        return infos

    def boundary_viscosity(self) -> np.ndarray:
        """The viscosity the momentum equation uses on every boundary face for the current
        velocity (host array): mu + mu_t, or the wall-function value on walls."""
        sv, fv = self.s, self.s.fv
        L, s = nat.lib(), nat.stream_ptr()
        M, B = ctypes.byref(fv.c_mesh), ctypes.byref(sv.c_bc)
        nat.check(L.ftmi_sst_face_gamma(M, B, nat.ptr(self.mut), self.mu, 0, nat.ptr(self.gm_f),
                                        nat.ptr(self.gm_b), s), "sst face viscosity")
        nat.check(L.ftmi_sst_wall(M, B, nat.ptr(sv.U), nat.ptr(fv.yb), self.mu, self.nu, nat.ptr(self.iters),
                                  nat.ptr(self.gm_b), nat.ptr(self.yplus), s), "sst wall function")
        return self.gm_b.cpu().numpy()

    def wall_shear(self) -> np.ndarray:
        """Wall shear stress magnitude on every boundary face, 0 off walls: the wall-function
        viscosity times the wall-parallel velocity relative to the wall over y_b
        (``turbulence.py:336-350``)."""
        sv, fv = self.s, self.s.fv
        mu_w = self.boundary_viscosity()
        U = sv.U.cpu().numpy().T
        nb, ub = fv.nb_h, sv.ub_fixed_h
        UP = U[fv.Pb_h]
        ut = UP - np.einsum("fd,fd->f", UP, nb)[:, None] * nb
        ut = ut - (ub - np.einsum("fd,fd->f", ub, nb)[:, None] * nb)
        mag = np.linalg.norm(ut, axis=1)
        return np.where(self.wall_b, mu_w * mag / np.maximum(fv.yb_h, 1e-300), 0.0)

    # ------------------------------------------------------------------ state
    def state(self) -> dict:
        return {"k": self.k.cpu().numpy(), "w": self.w.cpu().numpy(), "mut": self.mut.cpu().numpy()}

    def snapshot(self) -> dict:
        return {"k": self.k.clone(), "w": self.w.clone(), "mut": self.mut.clone()}

    def load(self, st: dict):
        if "k" in st:
            for name in ("k", "w", "mut"):
                v = st[name]
                getattr(self, name).copy_(v if torch.is_tensor(v) else torch.as_tensor(np.asarray(v, float),
                                                                                        device=self.s.device))

    def fields(self) -> dict:
        k, w, mut = self.k.cpu().numpy(), self.w.cpu().numpy(), self.mut.cpu().numpy()
        yp = np.zeros(self.s.nc)
        yb = self.yplus.cpu().numpy()
        wf = np.flatnonzero(self.wall_b)
        np.maximum.at(yp, self.s.fv.Pb_h[wf], yb[wf])
        return {"turb-kinetic-energy": k, "specific-diss-rate": w, "viscosity-turb": mut,
                "viscosity-ratio": mut / self.mu, "wall-distance": self.y_h.copy(), "y-plus": yp}
