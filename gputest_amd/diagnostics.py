"""Algebraic, conservation and health diagnostics of one outer iteration, computed on the
device (``flowtrace/unstructured/diagnostics.py``).

The metrics and their definitions are the reference's (its module docstring states them):
componentwise (Oettli-Prager) backward errors of the momentum, energy and pressure-correction
systems before and after their solves, with rows below ``gamma_K s_ref`` excluded; local and
global continuity of the predictor and corrected fluxes; the outer change of the fields; and
non-finite counts. What differs is where they are computed: the solver leaves the systems of
its latest iteration in device memory (the LDU operators, relaxed right-hand sides, the
iteration-start fields ``U_old``/``p_old``/``T_old``), and ``evaluate`` runs the HIP reductions of
``ftmi_diag.hip`` over them, each writing into one device buffer of doubles, which comes to the
host in a single transfer. The unrelaxed right-hand side is recovered on the fly as
``b = b_r - (1 - alpha)/alpha a_P u^n``.

Mixed precision keeps the transport systems in FP32 only; their diagnostics are not
evaluated (``available`` is False with the reason).
"""
from __future__ import annotations

import ctypes
import math
from typing import Dict, List, Sequence, Tuple

import numpy as np
import torch

from . import _native as nat

DIAGNOSTICS_VERSION = 1
COMPONENTS = ("x-velocity", "y-velocity", "z-velocity")
_ROWSET_STATS = ("row_scaled", "worst_row", "active_rows", "excluded_rows", "excluded_abs_max", "nonfinite_rows",
                 "abs_l1", "abs_max", "l2")
_VSTATS = ("diff_norm_max", "norm_max", "min", "max", "nonfinite", "sum", "sum_abs", "sum_sq")


def unit_roundoff(dtype) -> float:
    """u = eps / 2 of the working precision."""
    return float(np.finfo(np.dtype(dtype)).eps) / 2.0


def gamma(k: int, u: float) -> float:
    """Higham's gamma_k = k u / (1 - k u), the relative error bound of a k-term sum."""
    ku = k * u
    return ku / (1.0 - ku) if ku < 1.0 else math.inf


def ratio(num: float, den: float) -> float:
    """num / den with 0/0 = 0 and x/0 = inf (x > 0); non-finite inputs give inf."""
    if not (math.isfinite(num) and math.isfinite(den)):
        return math.inf
    if den > 0.0:
        return num / den
    return 0.0 if num == 0.0 else math.inf


class _Device:
    """Named slots of one device buffer, filled by the diagnostics kernels, read once."""

    def __init__(self, device, n_rows: int, nvec: int):
        L = nat.lib()
        self.L, self.st = L, nat.stream_ptr()
        self.out = torch.zeros(1024, dtype=torch.float64, device=device)
        self.work = torch.empty(int(L.ftmi_diag_work_doubles()), dtype=torch.float64, device=device)
        self.rs = torch.empty((nvec, n_rows), dtype=torch.float64, device=device)
        self.pos = 0
        self.slots: Dict[str, Tuple[int, Sequence[str]]] = {}

    def _slot(self, name: str, keys: Sequence[str]) -> int:
        at = self.pos
        self.slots[name] = (at, tuple(keys))
        self.pos += len(keys)
        if self.pos > self.out.numel():
            raise RuntimeError("diagnostics: output buffer too small")
        return self.out.data_ptr() + 8 * at

    def rows(self, j: int, n: int, dg, ptr, nbr, off, b, x, dgb=None, c=0.0, xb=None):
        """r -> rs[2j], s -> rs[2j + 1]"""
        r, s = self.rs[2 * j], self.rs[2 * j + 1]
        nat.check(self.L.ftmi_diag_rows(n, nat.ptr(dg), nat.ptr(ptr), nat.ptr(nbr), nat.ptr(off), nat.ptr(b),
                                        nat.ptr(dgb), float(c), nat.ptr(xb), nat.ptr(x), nat.ptr(r), nat.ptr(s),
                                        self.st), "diag rows")
        return r, s

    def flux(self, j: int, fv, F, Fb):
        m, t = self.rs[2 * j], self.rs[2 * j + 1]
        nat.check(self.L.ftmi_diag_cell_flux(fv.nc, nat.ptr(fv.cptr), nat.ptr(fv.cface), nat.ptr(fv.bptr),
                                             nat.ptr(fv.bface), nat.ptr(F), nat.ptr(Fb), nat.ptr(m), nat.ptr(t),
                                             self.st), "diag cell flux")
        return m, t

    def rowset(self, prefix: str, names: Sequence[str], pairs, K: int, u: float):
        n = int(pairs[0][0].numel())
        keys = ["s_ref"] + [f"{nm}|{k}" for nm in names for k in _ROWSET_STATS]
        out = self._slot(prefix, keys)
        arr = ctypes.c_void_p * len(pairs)
        nat.check(self.L.ftmi_diag_rowset(n, len(pairs), arr(*[r.data_ptr() for r, _ in pairs]),
                                          arr(*[s.data_ptr() for _, s in pairs]), float(gamma(K, u)), out,
                                          nat.ptr(self.work), self.st), "diag rowset")

    def vstats(self, name: str, x, y=None, ncomp: int = 1):
        x = x.contiguous()
        n = x.numel() // ncomp
        out = self._slot(name, _VSTATS)
        nat.check(self.L.ftmi_diag_vstats(n, ncomp, n, nat.ptr(x), nat.ptr(y), out, nat.ptr(self.work), self.st),
                  "diag vstats")

    def fetch(self) -> Dict[str, float]:
        host = self.out[: self.pos].cpu().numpy()  # the one device-to-host transfer
        vals = {}
        for name, (at, keys) in self.slots.items():
            for i, k in enumerate(keys):
                vals[f"{name}|{k}"] = float(host[at + i])
        return vals


def _row_report(vals: Dict[str, float], prefix: str, name: str, centres: np.ndarray) -> dict:
    key = f"{prefix}|{name}"
    nonfinite = int(vals[key + "|nonfinite_rows"])
    out = {"row_scaled": math.inf if nonfinite else vals[key + "|row_scaled"],
           "abs_l1": vals[key + "|abs_l1"], "abs_max": vals[key + "|abs_max"], "abs_l2": vals[key + "|l2"],
           "active_rows": int(vals[key + "|active_rows"]), "excluded_rows": int(vals[key + "|excluded_rows"]),
           "excluded_abs_max": vals[key + "|excluded_abs_max"], "nonfinite_rows": nonfinite}
    worst = int(vals[key + "|worst_row"])
    if out["active_rows"] and not nonfinite and worst >= 0:
        out["worst_cell"] = worst
        out["worst_cell_centre"] = [float(v) for v in centres[worst]]
    return out


def _linear_info(info):
    if info is None:
        return None
    res = float(info.residual)
    return {"iterations": int(info.iterations), "reported_residual": res if math.isfinite(res) else None,
            "converged": bool(info.converged)}


def _mesh_counts(fv) -> Tuple[int, int]:
    """(largest terms of a matrix row, largest faces of a cell), cached on the mesh."""
    k = getattr(fv, "_diag_counts", None)
    if k is None:
        rows = int(np.max(np.diff(fv.indptr_h))) + 1 if fv.nc else 1
        faces = np.diff(np.asarray(fv.cptr_h)) + np.diff(fv.bptr.cpu().numpy())
        k = (rows, int(faces.max()) if fv.nc else 1)
        fv._diag_counts = k
    return k


def evaluate(solver) -> dict:
    """Diagnostics of the solver's latest outer iteration (``diagnostics.py:214-360``)."""
    cap = getattr(solver, "_capture", None)
    if cap is None:
        return {"available": False, "version": DIAGNOSTICS_VERSION,
                "reason": "no outer iteration has been completed since the solver was created or restored"}
    if solver.mixed:
        return {"available": False, "version": DIAGNOSTICS_VERSION,
                "reason": "mixed precision keeps the transport systems in FP32 only"}
    fv, dim, nc = solver.fv, solver.dim, solver.nc
    u = unit_roundoff(np.float64)
    K_rows, K_cont = _mesh_counts(fv)
    names = COMPONENTS[:dim]
    dev = _Device(solver.device, nc, 2 * max(dim, 2))
    ptr, nbr = fv.cptr, fv.cnbr
    au = cap["alpha_u"]
    c_u = (1.0 - au) / au
    # momentum: assembled system at U^n, then the relaxed system at U*
    pre = [dev.rows(k, nc, solver.diag, ptr, nbr, solver.off, solver.bmom[k], solver.U_old[k], solver.diag, c_u,
                    solver.U_old[k]) for k in range(dim)]
    dev.rowset("mom_pre", names, pre, K_rows, u)
    sol = [dev.rows(k, nc, solver.diag_r, ptr, nbr, solver.off, solver.bmom[k], solver.Ustar[k]) for k in range(dim)]
    dev.rowset("mom_sol", names, sol, K_rows, u)
    for k in range(dim):  # ||b_r - A_r U^n||_2, the start of the linear solve
        r0, _ = dev.rows(0, nc, solver.diag_r, ptr, nbr, solver.off, solver.bmom[k], solver.U_old[k])
        dev.vstats(f"mom_true_r0|{names[k]}", r0)
    energy = solver.energy and cap.get("energy_info") is not None
    if energy:
        aT = cap["alpha_T"]
        pre = [dev.rows(0, nc, solver.diagT, ptr, nbr, solver.offT, solver.bT, solver.T_old, solver.diagT,
                        (1.0 - aT) / aT, solver.T_old)]
        dev.rowset("T_pre", ["energy"], pre, K_rows, u)
        sol = [dev.rows(0, nc, solver.diagT_r, ptr, nbr, solver.offT, solver.bT, solver.Tnew)]
        dev.rowset("T_sol", ["energy"], sol, K_rows, u)
        r0, _ = dev.rows(0, nc, solver.diagT_r, ptr, nbr, solver.offT, solver.bT, solver.T_old)
        dev.vstats("T_true_r0", r0)
    # pressure correction A_p p' = -m*
    pc = [dev.rows(0, nc, solver.diag_p, ptr, nbr, solver.offp, solver.rhs_p, solver.pp)]
    dev.rowset("pcorr", ["p-correction"], pc, K_rows, u)
    # local continuity of the predictor and the corrected fluxes
    for tag, F, Fb in (("pred", solver.Fst, solver.Fbst), ("corr", solver.F, solver.Fb)):
        m, t = dev.flux(0, fv, F, Fb)
        dev.rowset("cont_" + tag, ["cells"], [(m, t)], K_cont, u)
        dev.vstats(f"cont_{tag}|m", m)
        dev.vstats(f"cont_{tag}|F", F)
        dev.vstats(f"cont_{tag}|Fb", Fb)
    # outer change and health
    dev.vstats("U", solver.U, solver.U_old, ncomp=dim)
    dev.vstats("U_old", solver.U_old, ncomp=dim)
    dev.vstats("p", solver.p, solver.p_old)
    if energy:
        dev.vstats("T", solver.T, solver.T_old)
    vals = dev.fetch()

    centres = fv.xc_h
    out: dict = {"available": True, "version": DIAGNOSTICS_VERSION, "iteration": int(cap["iteration"]),
                 "rolled_back": bool(cap.get("rolled_back", False)), "precision": "float64", "unit_roundoff": u,
                 "row_terms_max": K_rows, "cell_faces_max": K_cont}
    mom = {"definition": "pre_solve: b - A U^n (assembled, unrelaxed); solve: b_r - A_r U* (system passed "
                         "to the linear solver)",
           "relax_u": au,
           "row_scale_reference": {"pre_solve": vals["mom_pre|s_ref"], "solve": vals["mom_sol|s_ref"]},
           "excluded_row_threshold": {"pre_solve": gamma(K_rows, u) * vals["mom_pre|s_ref"],
                                      "solve": gamma(K_rows, u) * vals["mom_sol|s_ref"]}}
    for k, nm in enumerate(names):
        p_rep = _row_report(vals, "mom_pre", nm, centres)
        s_rep = _row_report(vals, "mom_sol", nm, centres)
        s_rep["linear_solver"] = _linear_info(cap["mom_info"][k])
        s_rep["true_reduction_l2"] = ratio(s_rep["abs_l2"], math.sqrt(vals[f"mom_true_r0|{nm}|sum_sq"]))
        mom[nm] = {"pre_solve": p_rep, "solve": s_rep}
    out["momentum"] = mom
    if energy:
        p_rep = _row_report(vals, "T_pre", "energy", centres)
        s_rep = _row_report(vals, "T_sol", "energy", centres)
        s_rep["linear_solver"] = _linear_info(cap.get("energy_info"))
        s_rep["true_reduction_l2"] = ratio(s_rep["abs_l2"], math.sqrt(vals["T_true_r0|sum_sq"]))
        out["energy"] = {"pre_solve": p_rep, "solve": s_rep, "relax_T": cap["alpha_T"]}
    pcr = _row_report(vals, "pcorr", "p-correction", centres)
    pcr["linear_solver"] = _linear_info(cap.get("p_info"))
    pcr["relative_l2"] = ratio(pcr["abs_l2"], vals["cont_pred|cells|l2"])
    pcr["definition"] = "A_p p' = -m* (rhs = minus the predictor imbalance); residual = -(corrected imbalance)"
    out["pressure_correction"] = pcr
    cont: dict = {"units": "kg/s (per unit depth in 2-D), positive = net outflow"}
    for tag, label in (("pred", "predictor"), ("corr", "corrected")):
        rep = _row_report(vals, "cont_" + tag, "cells", centres)
        flux_l1 = vals[f"cont_{tag}|F|sum_abs"] + vals[f"cont_{tag}|Fb|sum_abs"]
        cont[label] = {"local_max_relative": rep["row_scaled"], "abs_l1": rep["abs_l1"], "abs_max": rep["abs_max"],
                       "abs_l2": rep["abs_l2"], "l1_over_face_flux": ratio(rep["abs_l1"], flux_l1),
                       "active_cells": rep["active_rows"], "excluded_cells": rep["excluded_rows"],
                       "excluded_abs_max": rep["excluded_abs_max"], "nonfinite_cells": rep["nonfinite_rows"],
                       "sum_of_cells": vals[f"cont_{tag}|m|sum"]}
        if "worst_cell" in rep:
            cont[label]["worst_cell"] = rep["worst_cell"]
            cont[label]["worst_cell_centre"] = rep["worst_cell_centre"]
    g_signed, g_through = vals["cont_corr|Fb|sum"], vals["cont_corr|Fb|sum_abs"]
    cont["global_boundary"] = {"signed": g_signed, "throughflow": g_through,
                               "relative": ratio(abs(g_signed), g_through)}
    out["continuity"] = cont

    rolled = out["rolled_back"]
    ubf = np.asarray(solver.ub_fixed_h, float).reshape(-1, dim)
    U_ref = max(vals["U|norm_max"], vals["U_old|norm_max"],
                float(np.sqrt((ubf * ubf).sum(axis=1)).max()) if ubf.size else 0.0)
    change: dict = {"velocity": None if rolled else ratio(vals["U|diff_norm_max"], U_ref), "velocity_reference": U_ref}
    rho = float(solver.case.density)
    p_ref = max(vals["p|max"] - vals["p|min"], rho * U_ref * U_ref)
    change["pressure"] = None if rolled else ratio(vals["p|diff_norm_max"], p_ref)
    change["pressure_reference"] = p_ref
    health = {"velocity": int(vals["U|nonfinite"]), "pressure": int(vals["p|nonfinite"]),
              "face_mass_flux": int(vals["cont_corr|F|nonfinite"]),
              "boundary_mass_flux": int(vals["cont_corr|Fb|nonfinite"])}
    if energy:
        from .simple import T_FIXED

        fixed = np.asarray(solver.tkind_h) == T_FIXED
        tv = np.asarray(solver.tval_h, float)[fixed]
        hi = max(vals["T|max"], float(tv.max()) if tv.size else -math.inf)
        lo = min(vals["T|min"], float(tv.min()) if tv.size else math.inf)
        change["temperature"] = None if rolled else ratio(vals["T|diff_norm_max"], hi - lo)
        change["temperature_reference"] = hi - lo
        health["temperature"] = int(vals["T|nonfinite"])
    out["field_change"] = change
    out["health"] = {"finite": not any(health.values()), "nonfinite_values": health}
    return out


def summary(d: dict) -> dict:
    """Flat headline numbers of ``evaluate`` (``diagnostics.py:363-382``)."""
    if not d.get("available"):
        return {}
    mom = d["momentum"]
    comps = [k for k in COMPONENTS if k in mom]
    out = {"iteration": d["iteration"],
           "momentum_pre_row_scaled": max(mom[k]["pre_solve"]["row_scaled"] for k in comps),
           "momentum_solve_row_scaled": max(mom[k]["solve"]["row_scaled"] for k in comps),
           "continuity_predictor_local_max": d["continuity"]["predictor"]["local_max_relative"],
           "continuity_corrected_local_max": d["continuity"]["corrected"]["local_max_relative"],
           "continuity_global_relative": d["continuity"]["global_boundary"]["relative"],
           "velocity_change": d["field_change"]["velocity"], "pressure_change": d["field_change"]["pressure"],
           "finite": d["health"]["finite"]}
    if "energy" in d:
        out["energy_pre_row_scaled"] = d["energy"]["pre_solve"]["row_scaled"]
    return out


def rowset_metrics(pairs: List[Tuple[torch.Tensor, torch.Tensor]], K: int, u: float) -> Dict[str, float]:
    """The row-set metrics of given device (r, s) vectors (the reduction ``evaluate`` uses)."""
    dev = _Device(pairs[0][0].device, 1, 2)
    dev.rowset("t", [f"v{i}" for i in range(len(pairs))], pairs, K, u)
    return dev.fetch()


__all__ = ["evaluate", "summary", "rowset_metrics", "unit_roundoff", "gamma", "ratio", "DIAGNOSTICS_VERSION"]
