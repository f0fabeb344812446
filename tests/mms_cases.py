"""Code verification by the method of manufactured solutions, as the reference runs it on the
unstructured SIMPLEC solver (``flowtrace/validation/mms.py:148-196``, tests
``tests/test_verification.py:11-51``).

The manufactured field on the unit square comes from the stream function
psi = sin^2(pi x) sin^2(pi y): u = d psi / dy, v = -d psi / dx (divergence free, zero on every
wall), p = cos(pi x) cos(pi y) (zero normal derivative on the walls), T = sin(pi x) sin(pi y)
(zero on the walls). With rho = cp = 1 the sources are the steady residuals
f = (u . grad) u + grad p - nu lap u and f_T = u . grad T - kappa lap T, whose derivatives are
written out below by hand (and checked against finite differences in ``test_mms.py``). The
discrete solution is compared with the exact one in the volume-weighted L2 norm on a sequence
of grids; the observed order is log(e_coarse / e_fine) / log(h_coarse / h_fine) with
h = sqrt(mean cell area)."""
import math

import numpy as np

from gputest_amd.mesh import cartesian_mesh_2d

PI = math.pi


def exact(x, y):
    sx, sy = np.sin(PI * np.asarray(x, float)), np.sin(PI * np.asarray(y, float))
    cx, cy = np.cos(PI * np.asarray(x, float)), np.cos(PI * np.asarray(y, float))
    return {"u": PI * sx * sx * (2 * sy * cy), "v": -PI * (2 * sx * cx) * sy * sy, "p": cx * cy, "T": sx * sy}


def sources(x, y, nu, kappa):
    x, y = np.asarray(x, float), np.asarray(y, float)
    sx, cx, sy, cy = np.sin(PI * x), np.cos(PI * x), np.sin(PI * y), np.cos(PI * y)
    s2x, c2x, s2y, c2y = np.sin(2 * PI * x), np.cos(2 * PI * x), np.sin(2 * PI * y), np.cos(2 * PI * y)
    u, v = PI * sx ** 2 * s2y, -PI * s2x * sy ** 2
    # first and second derivatives of u = pi sin^2(pi x) sin(2 pi y)
    ux, uy = PI ** 2 * s2x * s2y, 2 * PI ** 2 * sx ** 2 * c2y
    uxx, uyy = 2 * PI ** 3 * c2x * s2y, -4 * PI ** 3 * sx ** 2 * s2y
    # and of v = -pi sin(2 pi x) sin^2(pi y)
    vx, vy = -2 * PI ** 2 * c2x * sy ** 2, -PI ** 2 * s2x * s2y
    vxx, vyy = 4 * PI ** 3 * s2x * sy ** 2, -2 * PI ** 3 * s2x * c2y
    px, py = -PI * sx * cy, -PI * cx * sy
    T, Tx, Ty = sx * sy, PI * cx * sy, PI * sx * cy
    return {"u": u * ux + v * uy + px - nu * (uxx + uyy),
            "v": u * vx + v * vy + py - nu * (vxx + vyy),
            "T": u * Tx + v * Ty + kappa * 2 * PI ** 2 * T}


def l2(e, w):
    return float(np.sqrt(np.sum(w * e ** 2) / np.sum(w)))


def run(mod, n, nu=0.5, kappa=0.5, energy=False, triangulate=True, warp=0.05, pressure_gradient="green-gauss"):
    """One grid of ``mms.py:166-194`` with the oracle (``mod`` = ``oracle.simplec_oracle``) or the
    device solver (``mod`` = ``gputest_amd.case``): all four sides no-slip walls (T = 0 with
    energy), relax_u 0.8, the default second-order scheme, to the scaled residual 1e-8.
    Returns (h, {quantity: L2 error}, iterations, solver)."""
    oracle = hasattr(mod, "Simplec")
    mesh = cartesian_mesh_2d(n, n, triangulate=triangulate, warp=warp)
    B = mod.Boundary if oracle else mod.BoundarySpec
    wall = B("wall", thermal="temperature", value=0.0) if energy else B("wall")
    Cs = mod.Case if oracle else mod.UnstructuredCase
    case = Cs(boundaries={z: wall for z in ("left", "right", "bottom", "top")}, density=1.0, viscosity=nu,
              energy=energy, conductivity=kappa, cp=1.0, max_iter=20000, tol=1e-8, relax_u=0.8,
              pressure_gradient=pressure_gradient)
    if oracle:
        s = mod.Simplec(mesh, case)
        xc, vol = s.fv.xc, s.fv.vol
    else:
        from gputest_amd.simple import SimpleSolver

        s = SimpleSolver(mesh, case, "cuda")
        xc, vol = s.cell_centres(), np.asarray(s.fv.vol_h)
    src = sources(xc[:, 0], xc[:, 1], nu, kappa)
    s.set_sources(np.column_stack([src["u"], src["v"]]), src["T"] if energy else None)
    while not s.finished:
        s.step()
    if oracle:
        U, p, T = s.U, s.p, s.T
    else:
        U, p = s.U.cpu().numpy().T, s.p.cpu().numpy()
        T = s.T.cpu().numpy() if energy else None
    ex = exact(xc[:, 0], xc[:, 1])
    pe = ex["p"] - np.sum(ex["p"] * vol) / vol.sum()
    ph = p - np.sum(p * vol) / vol.sum()
    err = {"u": l2(U[:, 0] - ex["u"], vol), "v": l2(U[:, 1] - ex["v"], vol), "p": l2(ph - pe, vol)}
    if energy:
        err["T"] = l2(T - ex["T"], vol)
    return float(np.sqrt(vol.mean())), err, s.iteration, s


def orders(hs, errs):
    return {k: [math.log(errs[i][k] / errs[i + 1][k]) / math.log(hs[i] / hs[i + 1]) for i in range(len(hs) - 1)]
            for k in errs[0]}
