"""CPU restatement of the outer-iteration diagnostics (test infrastructure: the checker of
``gputest_amd.diagnostics``; nothing in the package imports it).

Follows ``flowtrace/unstructured/diagnostics.py``: ``_rows_of_system`` (lines 103-108: r = b - A x,
s = |b| + |A| |x| with |A| from the assembled CSR entries), ``_row_set`` (110-134: one reference
scale per set, rows with s <= gamma_K s_ref excluded, first-maximum worst row), and
``evaluate`` (214-360) on the oracle's captured systems. Pinned by the identities the
reference's own tests check (``tests/test_diagnostics.py``): the solved-system reduction equals
BiCGSTAB's reported residual, the corrected imbalance equals minus the pressure-correction
residual, the global boundary sum equals the sum of cell imbalances, a zero-scale system
reports zeros, and the synthetic row set of ``test_small_scale_rows_are_excluded``.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp

COMPONENTS = ("x-velocity", "y-velocity", "z-velocity")


def unit_roundoff(dtype=np.float64) -> float:
    return float(np.finfo(np.dtype(dtype)).eps) / 2.0


def gamma(k, u):
    ku = k * u
    return ku / (1.0 - ku) if ku < 1.0 else math.inf


def ratio(num, den):
    if not (math.isfinite(num) and math.isfinite(den)):
        return math.inf
    if den > 0.0:
        return num / den
    return 0.0 if num == 0.0 else math.inf


def rows_of_system(A, b, x):
    A = sp.csr_matrix(A)
    absA = sp.csr_matrix((np.abs(A.data), A.indices, A.indptr), shape=A.shape)
    return b - A @ x, np.abs(b) + absA @ np.abs(x)


def row_set(pairs, K, u):
    """{name: metrics} and s_ref of (name, r, s) triples sharing one reference scale."""
    s_ref = max(float(np.where(np.isfinite(s), s, 0.0).max()) if s.size else 0.0 for _, _, s in pairs)
    thr = gamma(K, u) * s_ref
    out = {}
    for name, r, s in pairs:
        ar = np.abs(r)
        finite = np.isfinite(r) & np.isfinite(s)
        active = finite & (s > thr)
        excluded = finite & ~active
        q = np.where(active, ar / np.where(active, s, 1.0), 0.0)
        fr = np.where(finite, ar, 0.0)
        nonfinite = int((~finite).sum())
        rep = {"row_scaled": math.inf if nonfinite else float(q.max(initial=0.0)),
               "active_rows": int(active.sum()), "excluded_rows": int(excluded.sum()),
               "excluded_abs_max": float(np.where(excluded, ar, 0.0).max(initial=0.0)), "nonfinite_rows": nonfinite,
               "abs_l1": float(fr.sum()), "abs_max": float(fr.max(initial=0.0)), "abs_l2": float(np.sqrt((fr * fr).sum()))}
        if rep["active_rows"] and not nonfinite:
            rep["worst_cell"] = int(np.argmax(q))
        out[name] = rep
    return out, s_ref


def evaluate(o) -> dict:
    """Diagnostics of the oracle's latest iteration (the reference's ``evaluate`` restated)."""
    cap = getattr(o, "_capture", None)
    if cap is None or "U_old" not in cap:
        return {"available": False}
    fv, dim = o.fv, o.dim
    u = unit_roundoff()
    K_rows = int(np.max(np.diff(sp.csr_matrix(cap["A"]).indptr))) + 1
    nf_cell = np.bincount(fv.P, minlength=o.nc) + np.bincount(fv.N, minlength=o.nc) + np.bincount(fv.Pb, minlength=o.nc)
    K_cont = int(nf_cell.max())
    names = COMPONENTS[:dim]
    U_old, U_star = cap["U_old"], cap["U_star"]
    pre, s_pre = row_set([(names[k],) + rows_of_system(cap["A"], cap["b"][k], U_old[:, k]) for k in range(dim)],
                         K_rows, u)
    sol, s_sol = row_set([(names[k],) + rows_of_system(cap["Ar"], cap["br"][k], U_star[:, k]) for k in range(dim)],
                         K_rows, u)
    mom = {"row_scale_reference": {"pre_solve": s_pre, "solve": s_sol}}
    for k, nm in enumerate(names):
        r0 = cap["br"][k] - cap["Ar"] @ U_old[:, k]
        sol[nm]["true_reduction_l2"] = ratio(sol[nm]["abs_l2"], float(np.sqrt((r0 * r0).sum())))
        info = cap["mom_info"][k]
        sol[nm]["linear_solver"] = {"iterations": int(info.iterations), "reported_residual": float(info.residual),
                                    "converged": bool(info.converged)}
        mom[nm] = {"pre_solve": pre[nm], "solve": sol[nm]}
    out = {"available": True, "iteration": int(cap["iteration"]), "rolled_back": bool(cap["rolled_back"]),
           "row_terms_max": K_rows, "cell_faces_max": K_cont, "momentum": mom}
    e = cap.get("energy")
    if e is not None:
        p_rep, _ = row_set([("energy",) + rows_of_system(e["A"], e["b"], e["T_old"])], K_rows, u)
        s_rep, _ = row_set([("energy",) + rows_of_system(e["Ar"], e["br"], e["T_new"])], K_rows, u)
        r0 = e["br"] - e["Ar"] @ e["T_old"]
        s_rep["energy"]["true_reduction_l2"] = ratio(s_rep["energy"]["abs_l2"], float(np.sqrt((r0 * r0).sum())))
        out["energy"] = {"pre_solve": p_rep["energy"], "solve": s_rep["energy"]}
    pc, _ = row_set([("p",) + rows_of_system(cap["Ap"], -cap["mass"], cap["pp"])], K_rows, u)
    pcr = pc["p"]
    cont = {}
    for tag, label, F, Fb in (("pred", "predictor", cap["Fst"], cap["Fbst"]), ("corr", "corrected", cap["F"], cap["Fb"])):
        m = fv.net_outflow(F, Fb)
        t = (np.bincount(fv.P, np.abs(F), o.nc) + np.bincount(fv.N, np.abs(F), o.nc)
             + np.bincount(fv.Pb, np.abs(Fb), o.nc))
        rep, _ = row_set([("cells", m, t)], K_cont, u)
        rep = rep["cells"]
        cont[label] = {"local_max_relative": rep["row_scaled"], "abs_l1": rep["abs_l1"], "abs_max": rep["abs_max"],
                       "abs_l2": rep["abs_l2"], "active_cells": rep["active_rows"],
                       "excluded_cells": rep["excluded_rows"], "sum_of_cells": float(m.sum()),
                       "l1_over_face_flux": ratio(rep["abs_l1"], float(np.abs(F).sum() + np.abs(Fb).sum()))}
        if "worst_cell" in rep:
            cont[label]["worst_cell"] = rep["worst_cell"]
    Fb = cap["Fb"]
    cont["global_boundary"] = {"signed": float(Fb.sum()), "throughflow": float(np.abs(Fb).sum()),
                               "relative": ratio(abs(float(Fb.sum())), float(np.abs(Fb).sum()))}
    pcr["relative_l2"] = ratio(pcr["abs_l2"], cont["predictor"]["abs_l2"])
    pi = cap["p_info"]
    pcr["linear_solver"] = {"iterations": int(pi.iterations), "reported_residual": float(pi.residual),
                            "converged": bool(pi.converged)}
    out["pressure_correction"] = pcr
    out["continuity"] = cont
    U_new, p_old, p_new = cap["U_new"], cap["p_old"], cap["p_new"]
    vmag = max(float(np.sqrt((U_new * U_new).sum(1)).max()), float(np.sqrt((U_old * U_old).sum(1)).max()))
    if fv.Pb.size:
        vmag = max(vmag, float(np.sqrt((o.ub_fixed * o.ub_fixed).sum(1)).max()))
    dU = float(np.sqrt(((U_new - U_old) ** 2).sum(1)).max())
    p_ref = max(float(p_new.max() - p_new.min()), o.case.density * vmag * vmag)
    rolled = out["rolled_back"]
    out["field_change"] = {"velocity": None if rolled else ratio(dU, vmag), "velocity_reference": vmag,
                           "pressure": None if rolled else ratio(float(np.abs(p_new - p_old).max()), p_ref),
                           "pressure_reference": p_ref}
    return out
