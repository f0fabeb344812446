"""Write the golden input/output vectors of the reference's FP64 SIMPLEC outer iteration.

Run once, in a container where the reference checkout is mounted (its path is the
first argument, default ``/root/reference``). The reference is imported as a
library, with its NumPy path forced (``FLOWTRACE_NO_NATIVE=1``, ``backend="numpy"``),
which the reference's CI calls "the reference" (``.github/workflows/ci.yml:62-75``).
Only data is written: ``.npz`` arrays and a ``.json`` of hashes, next to this file.

What is recorded (each pins something that no closed form pins):

* ``cavity64_simplec300.npz``: the 64x64 lid-driven cavity, SIMPLEC, relax 0.7/0.3,
  AMG-PCG to 1e-8, 300 outer iterations with ``tol=0`` (``flowtrace/unstructured/simple.py:500-703``):
  end fields, fields after iterations 1 and 10, and per-iteration inner counts and residuals.
* ``capture_<mesh>.npz``: the state before one outer iteration (fields plus the lagged
  gradients the iteration reads) and every intermediate system of that iteration
  (``simple.py:649-651``), on a warped-triangle cavity, an extruded 3-D hex cavity, a
  periodic channel with a pressure-gradient drive, and a heated duct with energy.
* ``amg_<mesh>.npz``: the smoothed-aggregation hierarchy of one captured pressure matrix
  (``flowtrace/linalg/amg.py:80-117``): aggregates, prolongators, coarse matrices.
* ``short_<case>.npz``: fields after a few iterations of reduced-size versions of the
  heated-duct and 3-D channel workloads.
* ``mesh_hashes.json``: SHA-256 of the generator outputs at full workload sizes
  (``flowtrace/mesh/unstructured.py:430-556, 647-711, 397-423``).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
os.environ["FLOWTRACE_NO_NATIVE"] = "1"
sys.path.insert(0, REF)

from flowtrace.hardware.backend import get_backend  # noqa: E402
from flowtrace.linalg.amg import SmoothedAggregationAMG, _aggregate, _strength  # noqa: E402
from flowtrace.mesh.unstructured import cartesian_mesh_2d, extrude_to_3d, make_periodic  # noqa: E402
from flowtrace.unstructured.case import BoundarySpec, UnstructuredCase  # noqa: E402
from flowtrace.unstructured.simple import SimpleSolver  # noqa: E402

BE = get_backend("numpy")


def _h(a) -> str:
    a = np.ascontiguousarray(a)
    return hashlib.sha256(a.tobytes()).hexdigest()


def _fv_arrays(fv) -> dict:
    """The FV precompute of ``flowtrace/unstructured/fvmesh.py:93-197``."""
    out = {}
    for k in ("P", "N", "Pb", "S", "Sb", "d", "db", "w", "alpha", "alpha_b", "Tvec", "wl", "wlb", "Ginv", "dP_f",
              "dN_f", "dxm", "vol", "nb", "db_t", "Sb_mag", "xc", "xfb"):
        out["fv_" + k] = np.asarray(getattr(fv, k))
    out["fv_indptr"], out["fv_indices"] = np.asarray(fv.indptr), np.asarray(fv.indices)
    out["fv_bzone"] = np.asarray(fv.bzone)
    out["fv_non_orth_deg"] = np.float64(fv.non_orthogonality_deg)
    return out


def _state(s) -> dict:
    st = {"U": s.U.copy(), "p": s.p.copy(), "F": s.F.copy(), "Fb": s.Fb.copy()}
    if s.energy:
        st["T"] = s.T.copy()
    if s._gp is not None:
        st["gp_lag"] = s._gp.copy()
        st["gU_lag"] = np.stack(s._gU, axis=0)
    if s.energy and s._gT is not None:
        st["gT_lag"] = s._gT.copy()
    return st


def _capture(s, it_before: int, tag: str, params: dict):
    """Run ``it_before`` iterations, then record the state and one more iteration's systems."""
    for _ in range(it_before):
        s.step()
    pre = {"pre_" + k: v for k, v in _state(s).items()}
    res_max5 = s._res_max5.get("continuity", 0.0)
    s.step()
    c = s._capture
    out = dict(pre)
    out.update(_fv_arrays(s.fv))
    out["res_max5_before"] = np.float64(res_max5)
    out["relax_p"] = np.float64(s.relax_p)
    out["A_data"] = c["A"].data.copy()
    out["b"] = np.stack(c["b"], axis=0)
    out["Ar_data"] = c["Ar"].data.copy()
    out["br"] = np.stack(c["br"], axis=0)
    out["U_star"] = c["U_star"].copy()
    out["mom_iters"] = np.array([i.iterations for i in c["mom_info"]])
    out["Fst"], out["Fbst"], out["mass"] = c["Fst"].copy(), c["Fbst"].copy(), c["mass"].copy()
    out["Ap_data"] = c["Ap"].data.copy()
    out["pp"] = c["pp"].copy()
    out["p_iters"] = np.int64(c["p_info"].iterations)
    out["F_new"], out["Fb_new"] = c["F"].copy(), c["Fb"].copy()
    out["U_new"], out["p_new"] = c["U_new"].copy(), c["p_new"].copy()
    if s.energy:
        e = c["energy"]
        out["T_A_data"], out["T_b"] = e["A"].data.copy(), e["b"].copy()
        out["T_new"] = e["T_new"].copy()
        out["T_iters"] = np.int64(e["info"].iterations)
    r = s.residuals()
    for k, v in r["scaled"].items():
        out["res_scaled_" + k] = np.float64(v)
    for k, v in r["raw"].items():
        out["res_raw_" + k] = np.float64(v)
    out["params"] = np.array(json.dumps(params))
    np.savez_compressed(os.path.join(HERE, f"capture_{tag}.npz"), **out)
    return s


def _amg(Ap, singular: bool, tag: str):
    """Hierarchy of ``SmoothedAggregationAMG`` (no pyamg here) on one pressure matrix."""
    import scipy.sparse as sp

    A = sp.csr_matrix(Ap, dtype=np.float64)
    out = {"A0_data": A.data, "A0_indices": A.indices, "A0_indptr": A.indptr}
    lvl = 0
    while A.shape[0] > 300:
        S = _strength(A, 0.08)
        agg, na = _aggregate(S.indptr.astype(np.int64), S.indices.astype(np.int64), A.shape[0])
        out[f"L{lvl}_S_indptr"], out[f"L{lvl}_S_indices"] = S.indptr, S.indices
        out[f"L{lvl}_agg"], out[f"L{lvl}_na"] = np.asarray(agg), np.int64(na)
        if na >= A.shape[0] * 0.9:
            break
        A = _next_level(A, agg, na, out, lvl)
        lvl += 1
    M = SmoothedAggregationAMG(Ap, BE, singular=singular)
    out["n_levels"] = np.int64(M.n_levels)
    out["coarse_inv"] = M._coarse_inv
    for k, (Ah, Ph, Rh, Dinv) in enumerate(M.levels):
        out[f"L{k}_P_data"], out[f"L{k}_P_indices"], out[f"L{k}_P_indptr"] = Ph.data, Ph.indices, Ph.indptr
        out[f"L{k}_Dinv"] = Dinv
    rng = np.random.default_rng(7)
    b = rng.standard_normal(Ap.shape[0])
    if singular:
        b -= b.mean()
    out["vcycle_b"], out["vcycle_x"] = b, M(b)
    np.savez_compressed(os.path.join(HERE, f"amg_{tag}.npz"), **out)


def _next_level(A, agg, na, out, lvl):
    """The level update of ``amg.py:95-104``, kept here only to record the coarse matrices."""
    import scipy.sparse as sp

    from flowtrace.linalg.amg import _spectral_radius

    T = sp.csr_matrix((np.ones(A.shape[0]), (np.arange(A.shape[0]), agg)), shape=(A.shape[0], na))
    cnt = np.asarray(T.sum(axis=0)).ravel()
    T = T @ sp.diags(1.0 / np.sqrt(cnt))
    Dinv = 1.0 / A.diagonal()
    DA = sp.diags(Dinv) @ A
    rho = _spectral_radius(DA)
    P = (T - (4.0 / 3.0 / rho) * (DA @ T)).tocsr()
    Ac = (P.T @ A @ P).tocsr()
    out[f"L{lvl}_rho"] = np.float64(rho)
    out[f"L{lvl+1}_A_data"], out[f"L{lvl+1}_A_indices"], out[f"L{lvl+1}_A_indptr"] = Ac.data, Ac.indices, Ac.indptr
    return Ac


def cavity_case(n_iter, relax_p=0.3, pressure_tol=1e-8, dim=2):
    vel = [1.0, 0.0] if dim == 2 else [1.0, 0.0, 0.0]
    return UnstructuredCase(boundaries={"top": BoundarySpec("wall", velocity=vel)}, viscosity=0.01,
                            algorithm="SIMPLEC", relax_u=0.7, relax_p=relax_p, pressure_tol=pressure_tol,
                            max_iter=n_iter, tol=0.0, backend="numpy")


def duct_mesh(nx, ny, seed=0):
    return cartesian_mesh_2d(nx, ny, 4.0, 1.0, names=("inlet", "outlet", "bottom", "top"),
                             types=("velocity-inlet", "pressure-outlet", "wall", "wall"), triangulate=True,
                             jitter=0.2, seed=seed)


def duct_solver(mesh, n_iter, pressure_tol=0.1):
    """Heated duct: parabolic inlet u = 4 y (1 - y) (u_max = 1), Re = u_max H / nu = 500,
    inlet T = 300, wall heat flux 5000, Pr = mu c_p / k = 0.7 with c_p = 1.

    ``BoundarySpec`` has no per-face inlet velocity (``simple.py:211-216``), so the inlet
    faces' fixed velocity is overwritten after construction and the face fluxes rebuilt
    from it, exactly as ``simple.py:110-112`` builds them."""
    mu = 1.0 / 500.0
    bnd = {"inlet": BoundarySpec("inlet", velocity=[1.0, 0.0], value=300.0),
           "outlet": BoundarySpec("outlet", pressure=0.0, value=300.0),
           "bottom": BoundarySpec("wall", thermal="flux", value=5000.0),
           "top": BoundarySpec("wall", thermal="flux", value=5000.0)}
    case = UnstructuredCase(boundaries=bnd, viscosity=mu, energy=True, conductivity=mu / 0.7, cp=1.0,
                            init_T=300.0, max_iter=n_iter, tol=0.0, pressure_tol=pressure_tol, backend="numpy")
    s = SimpleSolver(mesh, case, BE)
    zid = mesh.zone_by_name("inlet").id
    m = s.fv.bzone == zid
    y = s.fv.xfb[m, 1]
    s.ub_fixed[m, 0] = 4.0 * y * (1.0 - y)
    s.ub_fixed[m, 1] = 0.0
    s.F = case.density * np.sum(s.fv.face_interp_vec(s.U) * s.fv.S, axis=1)
    s.Fb = s._boundary_flux_fixed(s._velocity_bvals(s.U))
    s._reset_guard()
    return s


def channel_mesh(nx, ny, nz, stretch=2.0):
    eta = np.linspace(0.0, 1.0, ny + 1)
    ys = 0.5 * (1.0 + np.tanh(stretch * (2.0 * eta - 1.0)) / np.tanh(stretch))
    xs = np.linspace(0.0, 4.0, nx + 1)
    m2 = cartesian_mesh_2d(nx, ny, names=("left", "right", "bottom", "top"), xs=xs, ys=ys)
    return make_periodic(extrude_to_3d(m2, nz, 1.0), "left", "right")


def channel_case(n_iter, pressure_tol=0.1):
    """Periodic channel, laminar, Re_bulk = U_b h / nu = 2000 with U_b = 1 (h = 1):
    plane Poiseuille U_b = G h^2 / (12 mu) gives the mean pressure gradient -G."""
    mu = 1.0 / 2000.0
    G = 12.0 * mu
    return UnstructuredCase(viscosity=mu, periodic_pressure_gradient=[-G, 0.0, 0.0], max_iter=n_iter, tol=0.0,
                            pressure_tol=pressure_tol, backend="numpy")


def main():
    # 1. the 64x64 cavity, 300 outer iterations
    mesh = cartesian_mesh_2d(64, 64)
    s = SimpleSolver(mesh, cavity_case(300), BE)
    hist = {"mom_iters": [], "p_iters": [], "res_u": [], "res_v": [], "res_c": [], "raw_c": []}
    snaps = {}
    while not s.finished:
        s.step()
        ll = s.last_linear
        hist["mom_iters"].append(ll["momentum_iters"])
        hist["p_iters"].append(ll["pressure_iters"])
        r = s.residuals()
        hist["res_u"].append(r["scaled"]["x-velocity"])
        hist["res_v"].append(r["scaled"]["y-velocity"])
        hist["res_c"].append(r["scaled"]["continuity"])
        hist["raw_c"].append(r["raw"]["continuity"])
        if s.iteration in (1, 10, 100):
            for k, v in _state(s).items():
                snaps[f"it{s.iteration}_{k}"] = v
    out = {k: np.asarray(v) for k, v in hist.items()}
    out.update(snaps)
    out.update({"U": s.U, "p": s.p, "F": s.F, "Fb": s.Fb})
    np.savez_compressed(os.path.join(HERE, "cavity64_simplec300.npz"), **out)
    print("cavity64 done", s.iteration)

    # 2. captured iterations (tight pressure tolerance so the correction is solver-independent)
    p_tri = {"mesh": "cartesian_mesh_2d(16,16,triangulate=True,warp=0.05)", "relax_p": 0.8, "pressure_tol": 1e-12}
    s = SimpleSolver(cartesian_mesh_2d(16, 16, triangulate=True, warp=0.05), cavity_case(50, 0.8, 1e-12), BE)
    _capture(s, 3, "tri16", p_tri)
    Ap_tri = s._capture["Ap"]
    p_hex = {"mesh": "extrude_to_3d(cartesian_mesh_2d(6,5,warp=0.05),4,0.5)", "relax_p": 0.8, "pressure_tol": 1e-12}
    s = SimpleSolver(extrude_to_3d(cartesian_mesh_2d(6, 5, warp=0.05), 4, 0.5), cavity_case(50, 0.8, 1e-12, 3), BE)
    _capture(s, 3, "hex3d", p_hex)
    p_per = {"mesh": "channel_mesh(8,8,4)", "pressure_tol": 1e-12}
    s = SimpleSolver(channel_mesh(8, 8, 4), channel_case(50, 1e-12), BE)
    _capture(s, 3, "periodic", p_per)
    p_duct = {"mesh": "duct_mesh(24,8)", "pressure_tol": 1e-12}
    s = duct_solver(duct_mesh(24, 8), 50, 1e-12)
    _capture(s, 3, "duct", p_duct)
    print("captures done")

    # 3. AMG hierarchy of a pressure matrix large enough for two levels
    s = SimpleSolver(cartesian_mesh_2d(40, 40, triangulate=True, jitter=0.2), cavity_case(5, 0.8, 1e-8), BE)
    for _ in range(3):
        s.step()
    _amg(s._capture["Ap"].tocsr() if hasattr(s._capture["Ap"], "tocsr") else s._capture["Ap"], True, "tri40")
    del Ap_tri

    # 4. short runs of reduced-size workloads at the default pressure tolerance (0.1)
    s = duct_solver(duct_mesh(40, 20), 20)
    while not s.finished:
        s.step()
    np.savez_compressed(os.path.join(HERE, "short_duct40x20.npz"), U=s.U, p=s.p, F=s.F, Fb=s.Fb, T=s.T,
                        relax_p=np.float64(s.relax_p), iters=np.int64(s.iteration))
    s = SimpleSolver(channel_mesh(16, 8, 8), channel_case(20), BE)
    while not s.finished:
        s.step()
    np.savez_compressed(os.path.join(HERE, "short_channel16x8x8.npz"), U=s.U, p=s.p, F=s.F, Fb=s.Fb,
                        relax_p=np.float64(s.relax_p), iters=np.int64(s.iteration))
    print("short runs done")

    # 5. generator hashes at the full workload sizes
    hashes = {}
    m = duct_mesh(1000, 500)
    hashes["duct1000x500"] = {k: _h(getattr(m, k)) for k in ("nodes", "face_nodes", "owner", "neighbour",
                                                              "face_zone")}
    m = channel_mesh(128, 64, 64)
    hashes["channel128x64x64"] = {k: _h(getattr(m, k)) for k in ("nodes", "face_ptr", "face_nodes", "owner",
                                                                  "neighbour", "face_zone")}
    hashes["channel128x64x64"]["faces_a"] = _h(m.periodic[0]["faces_a"])
    hashes["channel128x64x64"]["faces_b"] = _h(m.periodic[0]["faces_b"])
    m = cartesian_mesh_2d(64, 64)
    hashes["cavity64"] = {k: _h(getattr(m, k)) for k in ("nodes", "face_nodes", "owner", "neighbour", "face_zone")}
    with open(os.path.join(HERE, "mesh_hashes.json"), "w") as fh:
        json.dump(hashes, fh, indent=1, sort_keys=True)
    print("hashes done")


if __name__ == "__main__":
    main()
