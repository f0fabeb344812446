"""The k-omega SST cases of ``tests/golden/make_sst_golden.py``, for the oracle (``mod`` =
``oracle.simplec_oracle``) and for the device solver (``mod`` = ``gputest_amd.case``)."""
import os

import numpy as np

from gputest_amd.mesh import cartesian_mesh_2d, extrude_to_3d, geometric, make_periodic

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _case(mod, boundaries, **kw):
    if hasattr(mod, "Simplec"):
        return mod.Case(boundaries={k: mod.Boundary(**v) for k, v in boundaries.items()}, **kw)
    return mod.UnstructuredCase(boundaries={k: mod.BoundarySpec(**v) for k, v in boundaries.items()}, **kw)


def channel_quads(mod, n_iter, zones=None):
    mesh = cartesian_mesh_2d(30, 12, 3.0, 1.0, names=("inlet", "outlet", "bottom", "top"),
                             types=("velocity-inlet", "pressure-outlet", "wall", "wall"))
    bnd = {"inlet": dict(kind="inlet", velocity=[1.0, 0.0]), "outlet": dict(kind="outlet"),
           "bottom": dict(kind="wall"), "top": dict(kind="wall")}
    return mesh, _case(mod, bnd, viscosity=1e-3, turbulence="k-omega-sst", max_iter=n_iter, tol=0.0)


def duct_tri(mod, n_iter, zones=None):
    mesh = cartesian_mesh_2d(24, 10, 4.0, 1.0, names=("inlet", "outlet", "bottom", "top"),
                             types=("velocity-inlet", "pressure-outlet", "wall", "wall"), triangulate=True,
                             jitter=0.2, seed=3)
    bnd = {"inlet": dict(kind="inlet", velocity=[1.0, 0.0], value=300.0, turb_intensity=0.08),
           "outlet": dict(kind="outlet", value=300.0),
           "bottom": dict(kind="wall", thermal="flux", value=50.0),
           "top": dict(kind="wall", thermal="temperature", value=310.0)}
    return mesh, _case(mod, bnd, viscosity=2e-3, energy=True, conductivity=2e-3 / 0.7, cp=1.0, init_T=300.0,
                       turbulence="k-omega-sst", max_iter=n_iter, tol=0.0)


def channel_hex(mod, n_iter, zones=None):
    m2 = cartesian_mesh_2d(10, 6, 2.0, 1.0, names=("inlet", "outlet", "bottom", "top"),
                           types=("velocity-inlet", "pressure-outlet", "wall", "wall"))
    mesh = extrude_to_3d(m2, 4, 0.8)
    names = list(zones) if zones is not None else [z.name for z in mesh.boundary_zones()]
    bnd = {z: dict(kind="wall") for z in names}
    bnd["inlet"] = dict(kind="inlet", velocity=[1.0, 0.0, 0.0])
    bnd["outlet"] = dict(kind="outlet")
    return mesh, _case(mod, bnd, viscosity=1e-3, turbulence="k-omega-sst", max_iter=n_iter, tol=0.0)


# The fully developed channel at Re_tau = 392.24 against the DNS of Moser, Kim & Mansour, Phys.
# Fluids 11 (1999) 943: the 16 points (y+, U+) the reference samples (flowtrace/validation/
# turbulent.py:20-24), u_tau = 1, delta = 1.
RE_TAU = 392.24
DNS_YPLUS = np.array([1.0628, 1.8888, 4.9810, 9.5307, 15.521, 19.797, 29.858, 50.958, 68.804, 98.400,
                      150.92, 198.91, 251.08, 301.61, 349.01, 392.24])
DNS_UPLUS = np.array([1.0605, 1.8799, 4.8032, 8.1754, 10.836, 11.966, 13.479, 14.902, 15.614, 16.493,
                      17.664, 18.535, 19.256, 19.758, 20.046, 20.133])


def sst_channel_dns(mod, max_iter=20000):
    """The reference's SST validation case (flowtrace/validation/turbulent.py:47-75): a half
    channel, wall at y = 0 and symmetry at y = 1, 4 x 50 cells graded to a first cell of 1.3e-3
    (y+ ~ 0.25 at the centre), periodic in x, driven by a body force 1 (dp/dx = -1)."""
    ys = geometric(50, 1.0, 1.3e-3)
    mesh = cartesian_mesh_2d(4, len(ys) - 1, xs=np.linspace(0.0, 0.2, 5), ys=ys,
                             names=("in", "out", "wall", "centre"), types=("wall", "wall", "wall", "symmetry"))
    mesh = make_periodic(mesh, "in", "out")
    bnd = {"wall": dict(kind="wall"), "centre": dict(kind="symmetry")}
    return mesh, _case(mod, bnd, viscosity=1.0 / RE_TAU, body_force=[1.0, 0.0, 0.0], turbulence="k-omega-sst",
                       turb_reference_velocity=15.0, init_velocity=[15.0, 0.0, 0.0], relax_u=0.9, max_iter=max_iter,
                       tol=1e-6)


def channel_profile(xc, u):
    """U at the DNS heights from the first column of cells, and its error in percent."""
    first = np.isclose(xc[:, 0], xc[:, 0].min())
    order = np.argsort(xc[first, 1])
    U = np.interp(DNS_YPLUS / RE_TAU, xc[first, 1][order], u[first][order])
    return U, 100.0 * (U - DNS_UPLUS) / DNS_UPLUS
