"""Write the golden vectors of the reference's stability guard and boundary heat flows.

Run once, in a container where the reference checkout is mounted (its path is the first
argument, default ``/root/reference``), with the reference's NumPy path forced
(``FLOWTRACE_NO_NATIVE=1``, ``backend="numpy"``). Only data is written, next to this file:

* ``guard_sheared24.npz``: the reference's own guard test case
  (``tests/test_planner.py:93-113``, ``tests/test_diagnostics.py:170-193``): the 24 x 24 cavity
  sheared by 45 degrees, nu = 0.1, automatic pressure relaxation, tol 1e-6, max 3000
  iterations. Recorded: every guard event (iteration detected, iteration restored to, the
  new relax_p, the raw continuity that tripped it), the per-iteration raw continuity,
  relax_p and inner iteration counts, the fields after iterations 40 and 100 and at the end,
  and the same case with a user-set relax_p = 0.8 (no rollbacks; the iteration at which its
  continuity stops being finite).
* ``heat_flow_channel.npz`` / ``heat_flow_duct.npz``: ``SimpleSolver.heat_flow()`` per
  boundary zone (``flowtrace/unstructured/simple.py:1003-1031``) with the fields it was
  computed from, at several iterations of the reference's heated channel
  (``tests/test_unstructured.py:46-59``, a fixed-temperature wall) and of the 24 x 8 heated
  duct (flux walls, parabolic inlet).
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
os.environ["FLOWTRACE_NO_NATIVE"] = "1"
sys.path.insert(0, REF)
sys.path.insert(0, HERE)

from flowtrace.mesh.unstructured import cartesian_mesh_2d  # noqa: E402
from flowtrace.unstructured.case import BoundarySpec, UnstructuredCase  # noqa: E402
from flowtrace.unstructured.simple import SimpleSolver  # noqa: E402

from make_golden import BE, duct_mesh, duct_solver  # noqa: E402


def sheared(relax_p):
    mesh = cartesian_mesh_2d(24, 24)
    mesh.nodes[:, 0] = mesh.nodes[:, 0] + mesh.nodes[:, 1]
    case = UnstructuredCase(boundaries={"top": BoundarySpec("wall", velocity=[1.0, 0.0])}, viscosity=0.1,
                            max_iter=3000, tol=1e-6, backend="numpy", relax_p=relax_p)
    return SimpleSolver(mesh, case, BE)


def guard_case():
    s = sheared(None)
    hist = {"raw_c": [], "relax_p": [], "p_iters": [], "mom_iters": []}
    snaps = {}
    while not s.finished and np.isfinite(s.residuals()["raw"].get("continuity", 0.0)):
        s.step()
        hist["raw_c"].append(s._last_raw["continuity"] if not s._capture.get("rolled_back") else np.nan)
        hist["relax_p"].append(s.relax_p)
        hist["p_iters"].append(s.last_linear["pressure_iters"])
        hist["mom_iters"].append(s.last_linear["momentum_iters"])
        if s.iteration in (40, 100) and f"it{s.iteration}_U" not in snaps:
            for k in ("U", "p", "F", "Fb"):
                snaps[f"it{s.iteration}_{k}"] = getattr(s, k).copy()
    ev = s.guard_events
    out = {k: np.asarray(v, dtype=np.float64) for k, v in hist.items()}
    out.update(snaps)
    out.update({"U": s.U, "p": s.p, "F": s.F, "Fb": s.Fb, "iterations": np.int64(s.iteration),
                "converged": np.bool_(s.converged), "relax_p_final": np.float64(s.relax_p),
                "relax_p_initial": np.float64(hist["relax_p"][0] if not ev or ev[0]["detected_at"] > 1
                                              else np.nan),
                "non_orth_deg": np.float64(s.fv.non_orthogonality_deg),
                "ev_detected_at": np.asarray([e["detected_at"] for e in ev], np.int64),
                "ev_restored_to": np.asarray([e["restored_to"] for e in ev], np.int64),
                "ev_relax_p": np.asarray([e["relax_p"] for e in ev], np.float64),
                "ev_continuity": np.asarray([e["continuity"] for e in ev], np.float64),
                "events_json": np.frombuffer(json.dumps(ev).encode(), np.uint8)})
    print("guard: iterations", s.iteration, "converged", s.converged, "events", ev)

    f = sheared(0.8)
    with np.errstate(all="ignore"):
        while np.isfinite(f.residuals()["raw"].get("continuity", 0.0)) and not f.finished:
            f.step()
    out["fixed_iterations"] = np.int64(f.iteration)
    out["fixed_converged"] = np.bool_(f.converged)
    out["fixed_events"] = np.int64(len(f.guard_events))
    print("fixed relax_p: stopped at", f.iteration, "converged", f.converged)
    np.savez_compressed(os.path.join(HERE, "guard_sheared24.npz"), **out)


def _heat_record(s, at, out):
    hf = s.heat_flow()
    for name, v in hf.items():
        out[f"it{at}_q_{name}"] = np.float64(v)
    out[f"it{at}_T"], out[f"it{at}_Fb"] = s.T.copy(), s.Fb.copy()
    out[f"it{at}_U"], out[f"it{at}_p"], out[f"it{at}_F"] = s.U.copy(), s.p.copy(), s.F.copy()
    if s._gp is not None:
        out[f"it{at}_gp_lag"] = s._gp.copy()


def heat_cases():
    mesh = cartesian_mesh_2d(60, 12, 5.0, 1.0, names=("inlet", "outlet", "bottom", "top"),
                             types=("velocity-inlet", "pressure-outlet", "wall", "wall"))
    bnd = {"inlet": BoundarySpec("inlet", normal_speed=1.0, value=0.0), "outlet": BoundarySpec("outlet"),
           "bottom": BoundarySpec("wall", thermal="temperature", value=1.0), "top": BoundarySpec("wall")}
    case = UnstructuredCase(boundaries=bnd, viscosity=0.02, energy=True, conductivity=0.02, max_iter=1500, tol=1e-6,
                            backend="numpy")
    s = SimpleSolver(mesh, case, BE)
    out = {}
    while not s.finished:
        s.step()
        if s.iteration in (10, 100):
            _heat_record(s, s.iteration, out)
    _heat_record(s, 0, out)
    out["iterations"], out["converged"] = np.int64(s.iteration), np.bool_(s.converged)
    np.savez_compressed(os.path.join(HERE, "heat_flow_channel.npz"), **out)
    print("channel heat flow", s.iteration, s.converged, s.heat_flow())

    s = duct_solver(duct_mesh(24, 8), 400, 1e-8)
    s.case.tol = 1e-8
    out = {}
    while not s.finished:
        s.step()
        if s.iteration in (5, 50):
            _heat_record(s, s.iteration, out)
    _heat_record(s, 0, out)
    out["iterations"], out["converged"] = np.int64(s.iteration), np.bool_(s.converged)
    np.savez_compressed(os.path.join(HERE, "heat_flow_duct.npz"), **out)
    print("duct heat flow", s.iteration, s.converged, s.heat_flow())


if __name__ == "__main__":
    guard_case()
    heat_cases()
