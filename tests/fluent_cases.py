"""The reference's square-cylinder regression (``tests/test_square_cylinder.py:1-53``): its
committed GAMBIT mesh (5 824 quads, D = 4, rear face at x = -18), Re = 40 with nu = 0.1 and a
unit inlet speed, and the drag coefficients and recirculation lengths the reference pins for
side walls and for symmetric sides."""
import os

import numpy as np
from scipy.interpolate import LinearNDInterpolator

MESH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "SquareCylinder-Coarse.msh.gz")
D, X_REAR = 4.0, -18.0
# (sides, Cd, Lr/D) pinned by the reference, tolerances 2e-4 and 5e-3 as there
PINS = [("wall", 2.12276, 2.3485), ("symmetry", 1.90654, 2.8977)]


def boundaries(mod, mesh, sides):
    from gputest_amd.case import default_boundaries

    out = {k: mod.Boundary(v.kind, normal_speed=v.normal_speed) for k, v in default_boundaries(mesh).items()}
    if sides == "symmetry":
        out["SideWalls"] = mod.Boundary("symmetry")
    return out


def case(mod, mesh, sides):
    return mod.Case(boundaries=boundaries(mod, mesh, sides), viscosity=0.1, max_iter=3000, tol=1e-7)


def drag_lift(forces):
    q = 0.5 * 1.0 * 1.0 ** 2 * D
    f = forces["BluffWalls"]["total"]
    return f[0] / q, f[1] / q


def recirculation_length(centres, u):
    """Distance behind the rear face to where u on the axis turns positive, in D."""
    ip = LinearNDInterpolator(centres, u)
    xs = np.linspace(X_REAR + 1e-3, 30.0, 6000)
    v = ip(xs, np.zeros_like(xs))
    k = np.nonzero((v[:-1] < 0) & (v[1:] >= 0))[0][0]
    x0 = xs[k] - v[k] * (xs[k + 1] - xs[k]) / (v[k + 1] - v[k])
    return (x0 - X_REAR) / D


# The reference's example meshes (examples/README.md), gzipped by make_fluent_fixture.py, and the
# cases its command lines run on them:
#   solve heated_duct_tri.msh --energy --mu 0.02 --k 0.02 --bc inlet=inlet:1.0:T0 --bc floor=wall::T1
#   solve cavity_quad.msh --mu 0.01 --bc lid=wall:1,0
# (flowtrace/cli.py:207-238 turns "inlet:1.0:T0" into a normal inflow speed 1 at T = 0 and
# "wall::T1" into a wall at T = 1; the other zones keep their defaults).
_GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EXAMPLES = {"heated_duct_tri": os.path.join(_GOLDEN, "example_heated_duct_tri.msh.gz"),
            "cavity_quad": os.path.join(_GOLDEN, "example_cavity_quad.msh.gz"),
            "elbow_mix": os.path.join(_GOLDEN, "example_elbow_mix.msh.gz")}


def example_case(mod, mesh, name, max_iter, tol=0.0):
    """The README's case on ``mesh``, for the oracle (``mod`` = ``oracle.simplec_oracle``) or
    the device solver (``mod`` = ``gputest_amd.case``)."""
    from gputest_amd.case import default_boundaries

    B = mod.Boundary if hasattr(mod, "Simplec") else mod.BoundarySpec
    Cs = mod.Case if hasattr(mod, "Simplec") else mod.UnstructuredCase
    bnd = {k: B(v.kind, normal_speed=v.normal_speed) for k, v in default_boundaries(mesh).items()}
    if name == "heated_duct_tri":
        bnd["inlet"] = B("inlet", normal_speed=1.0, value=0.0)
        bnd["floor"] = B("wall", thermal="temperature", value=1.0)
        return Cs(boundaries=bnd, viscosity=0.02, energy=True, conductivity=0.02, cp=1.0, max_iter=max_iter, tol=tol)
    if name == "cavity_quad":
        bnd["lid"] = B("wall", velocity=[1.0, 0.0])
        return Cs(boundaries=bnd, viscosity=0.01, max_iter=max_iter, tol=tol)
    raise ValueError(name)
