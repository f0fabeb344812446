"""Write the golden vectors of the reference's k-omega SST model.

Run once, in a container where the reference checkout is mounted (its path is the first
argument, default ``/root/reference``), like ``make_golden.py``: the reference is imported as
a library with its NumPy path forced (``FLOWTRACE_NO_NATIVE=1``, ``backend="numpy"``). Only
data is written, as ``.npz`` files next to this script.

What is recorded (``flowtrace/unstructured/turbulence.py:116-440`` wired into
``flowtrace/unstructured/simple.py:546-548, 662-669, 781-782``):

* ``sst_spalding.npz``: u+ of Spalding's law on a sweep of Re_y (``turbulence.py:65-87``).
* ``sst_channel_quads.npz``: 2-D channel [0, 3] x [0, 1], 30 x 12 quads, uniform inlet u = 1,
  outlet, no-slip walls, nu = 1e-3, SST with the case defaults; after each of 12 outer
  iterations the fields U, p, F, Fb, k, omega, mu_t and the residuals; at the end the wall
  distance, the wall-face distance y_b, the inlet values of k and omega, y+ and the
  wall-function viscosity.
* ``sst_duct_tri.npz``: the same on 24 x 10 jittered triangles of [0, 4] x [0, 1] with energy
  (heated walls, Pr_t = 0.85 in the diffusivity), 10 iterations.
* ``sst_channel_hex.npz``: an extruded 3-D channel (10 x 6 x 4 hexes, walls on four sides),
  5 iterations: pins the 3-D branch of the wall distance.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
os.environ["FLOWTRACE_NO_NATIVE"] = "1"
sys.path.insert(0, REF)

from flowtrace.hardware.backend import get_backend  # noqa: E402
from flowtrace.mesh.unstructured import cartesian_mesh_2d, extrude_to_3d  # noqa: E402
from flowtrace.unstructured.case import BoundarySpec, UnstructuredCase  # noqa: E402
from flowtrace.unstructured.simple import SimpleSolver  # noqa: E402
from flowtrace.unstructured.turbulence import spalding_uplus  # noqa: E402

BE = get_backend("numpy")


def _run(s, n_iter, path, extra=None):
    out = {}
    t = s.turb
    out["y"], out["yb"], out["kb_in"], out["sb_in"] = t.y.copy(), t.yb.copy(), t.kb_in.copy(), t.sb_in.copy()
    out["wall_cells"] = np.asarray(t.wall_cells, np.int64)
    out["k0"], out["w0"], out["mut0"] = t.k.copy(), t.w.copy(), t.mut.copy()
    names = None
    for it in range(1, n_iter + 1):
        s.step()
        out[f"it{it}_U"], out[f"it{it}_p"] = s.U.copy(), s.p.copy()
        out[f"it{it}_F"], out[f"it{it}_Fb"] = s.F.copy(), s.Fb.copy()
        out[f"it{it}_k"], out[f"it{it}_w"], out[f"it{it}_mut"] = t.k.copy(), t.w.copy(), t.mut.copy()
        if s.energy:
            out[f"it{it}_T"] = s.T.copy()
        res = s.residuals()
        raw = res.get("raw", res)
        names = sorted(raw)
        out[f"it{it}_res"] = np.array([float(raw[n]) for n in names])
        out[f"it{it}_rolled_back"] = np.int64(bool(getattr(s, "_capture", None) is None))
    out["res_names"] = np.array(names)
    out["mu_w"] = np.asarray(t.wall_shear_viscosity(s.U), np.float64)
    out["yplus"] = np.asarray(t.yplus, np.float64)
    if extra:
        out.update(extra)
    np.savez_compressed(path, **out)
    print("wrote", path, s.iteration, "iterations", s.residuals())


def main():
    re_y = np.concatenate([[0.0, 1e-8, 1e-3, 0.5, 3.0], np.geomspace(5.0, 1e6, 40)])
    np.savez_compressed(os.path.join(HERE, "sst_spalding.npz"), re_y=re_y, uplus=spalding_uplus(re_y))

    mu = 1e-3
    m = cartesian_mesh_2d(30, 12, 3.0, 1.0, names=("inlet", "outlet", "bottom", "top"),
                          types=("velocity-inlet", "pressure-outlet", "wall", "wall"))
    bnd = {"inlet": BoundarySpec("inlet", velocity=[1.0, 0.0]), "outlet": BoundarySpec("outlet"),
           "bottom": BoundarySpec("wall"), "top": BoundarySpec("wall")}
    case = UnstructuredCase(boundaries=bnd, viscosity=mu, turbulence="k-omega-sst", max_iter=12, tol=0.0,
                            backend="numpy")
    _run(SimpleSolver(m, case, BE), 12, os.path.join(HERE, "sst_channel_quads.npz"))

    m = cartesian_mesh_2d(24, 10, 4.0, 1.0, names=("inlet", "outlet", "bottom", "top"),
                          types=("velocity-inlet", "pressure-outlet", "wall", "wall"), triangulate=True,
                          jitter=0.2, seed=3)
    bnd = {"inlet": BoundarySpec("inlet", velocity=[1.0, 0.0], value=300.0, turb_intensity=0.08),
           "outlet": BoundarySpec("outlet", value=300.0),
           "bottom": BoundarySpec("wall", thermal="flux", value=50.0),
           "top": BoundarySpec("wall", thermal="temperature", value=310.0)}
    case = UnstructuredCase(boundaries=bnd, viscosity=2e-3, energy=True, conductivity=2e-3 / 0.7, cp=1.0,
                            init_T=300.0, turbulence="k-omega-sst", max_iter=10, tol=0.0, backend="numpy")
    _run(SimpleSolver(m, case, BE), 10, os.path.join(HERE, "sst_duct_tri.npz"))

    m2 = cartesian_mesh_2d(10, 6, 2.0, 1.0, names=("inlet", "outlet", "bottom", "top"),
                           types=("velocity-inlet", "pressure-outlet", "wall", "wall"))
    m = extrude_to_3d(m2, 4, 0.8)
    zones = [z.name for z in m.boundary_zones()]
    bnd = {z: BoundarySpec("wall") for z in zones}
    bnd["inlet"] = BoundarySpec("inlet", velocity=[1.0, 0.0, 0.0])
    bnd["outlet"] = BoundarySpec("outlet")
    case = UnstructuredCase(boundaries=bnd, viscosity=1e-3, turbulence="k-omega-sst", max_iter=5, tol=0.0,
                            backend="numpy")
    _run(SimpleSolver(m, case, BE), 5, os.path.join(HERE, "sst_channel_hex.npz"),
         extra={"zones": np.array(zones)})


if __name__ == "__main__":
    main()
