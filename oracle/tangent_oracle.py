"""ORACLE (test infrastructure only; never part of the product path).

An independent CPU reference for the derivatives of a converged objective: the tangent
(forward-mode) linearisation of the oracle's own outer iteration at its converged state. It
uses nothing of ``gputest_amd.adjoint``.

At the converged state w* the outer iteration ``oracle.simplec_oracle.Simplec.step`` with exact
inner solves is a map w -> Phi(w; beta) with w* = Phi(w*; beta), the state being the fields
(U, p, F, Fb) and the lagged gradients (grad p, grad U) that the next iteration's boundary
values use (``simple.py:500-703``). For a parameter direction d,

    dw = (I - Phi_w)^-1 Phi_beta d,      dJ/dbeta . d = J_beta d + J_w dw,

with J the drag on a wall zone (``simple.py:970-1001``). Every derivative here is a central
difference of the oracle's own functions (``Phi`` with inner solves to 1e-14 and the guard off,
the wall force), and the linear system is solved by SciPy's GMRES on those Jacobian-vector
products. The parameters are the inlet profile's control values, linearly interpolated over
equally spaced heights (the profile parametrisation of the adjoint cases), and the kinematic
viscosity.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

from oracle import simplec_oracle as O


def interp_weights(y: np.ndarray, y0: float, y1: float, n: int) -> np.ndarray:
    """(faces, n) weights of the piecewise-linear profile through n equally spaced values."""
    t = np.clip((y - y0) / (y1 - y0) * (n - 1), 0.0, n - 1)
    i = np.minimum(np.floor(t).astype(np.int64), n - 2)
    a = t - i
    W = np.zeros((y.size, n))
    W[np.arange(y.size), i] = 1.0 - a
    W[np.arange(y.size), i + 1] = a
    return W


class TangentReference:
    """Tangent derivatives of the drag of a converged oracle solver ``o`` (laminar, no energy)
    whose inlet zone ``inlet`` carries the profile of ``u_ctrl`` over heights [y0, y1]."""

    def __init__(self, o: "O.Simplec", u_ctrl, y0: float, y1: float, zone: str = "cylinder", inlet: str = "inlet",
                 eps: float = 1e-7):
        self.o, self.zone, self.eps = o, zone, eps
        self.u_ctrl = np.asarray(u_ctrl, float).copy()
        self.nu0 = float(o.case.viscosity) / float(o.case.density)
        fv = o.fv
        zid = o.mesh.zone_by_name(inlet).id
        self.inlet = np.flatnonzero(fv.bzone == zid)
        self.W = interp_weights(fv.xfb[self.inlet, 1], y0, y1, self.u_ctrl.size)
        o.guard_enabled = False
        o.case.tol = 0.0
        o.case.pressure_tol = 1e-14
        o._solve = self._exact_solve
        self.w0 = self.pack()
        self.beta0 = np.concatenate([self.u_ctrl, [self.nu0]])

    # ---------------------------------------------------------------- state and parameters
    @staticmethod
    def _exact_solve(A, b, x0, diag, tol=0.1, maxiter=30):
        inv = 1.0 / diag
        return O.bicgstab(lambda v: A @ v, b, M=lambda v: inv * v, x0=x0, tol=1e-14, maxiter=5000,
                          relative_to_initial=True)

    def pack(self) -> np.ndarray:
        o = self.o
        return np.concatenate([o.U.ravel(), o.p, o.F, o.Fb, o._gp.ravel(), np.stack(o._gU).ravel()])

    def unpack(self, w: np.ndarray):
        o, n, dim = self.o, self.o.nc, self.o.dim
        nf, nb = o.fv.n_if, o.fv.n_bf
        k = 0

        def take(size, shape):
            nonlocal k
            out = w[k:k + size].reshape(shape).copy()
            k += size
            return out

        o.U = take(n * dim, (n, dim))
        o.p = take(n, (n,))
        o.F = take(nf, (nf,))
        o.Fb = take(nb, (nb,))
        o._gp = take(n * dim, (n, dim))
        gU = take(dim * n * dim, (dim, n, dim))
        o._gU = [gU[i] for i in range(dim)]

    def set_params(self, beta: np.ndarray):
        o = self.o
        ux = self.W @ beta[:-1]
        o.ub_fixed[self.inlet, 0] = ux
        o.ub_fixed[self.inlet, 1:] = 0.0
        o.case.viscosity = float(beta[-1]) * float(o.case.density)

    def phi(self, w: np.ndarray, beta: np.ndarray) -> np.ndarray:
        """One outer iteration from state w with parameters beta."""
        self.unpack(w)
        self.set_params(beta)
        self.o.converged = False
        self.o.iteration = 10  # past the start-up rules; the AMG age only picks a preconditioner
        self.o.step()
        return self.pack()

    def drag(self, w: np.ndarray, beta: np.ndarray) -> float:
        self.unpack(w)
        self.set_params(beta)
        return float(self.o.wall_forces()[self.zone]["total"][0])

    # ---------------------------------------------------------------- derivatives
    def _h(self, x, v):
        return self.eps * max(float(np.linalg.norm(x)), 1.0) / max(float(np.linalg.norm(v)), 1e-300)

    def jvp_w(self, v):
        h = self._h(self.w0, v)
        return (self.phi(self.w0 + h * v, self.beta0) - self.phi(self.w0 - h * v, self.beta0)) / (2 * h)

    def jvp_beta(self, d):
        h = self._h(self.beta0, d)
        return (self.phi(self.w0, self.beta0 + h * d) - self.phi(self.w0, self.beta0 - h * d)) / (2 * h)

    def fixed_point_residual(self) -> float:
        w1 = self.phi(self.w0, self.beta0)
        return float(np.linalg.norm(w1 - self.w0) / np.linalg.norm(self.w0))

    def directional(self, d, tol=1e-9, restart=300, maxiter=4):
        """dJ/dbeta . d and the GMRES record (matrix-vector products, relative residual); GMRES
        restarts every ``restart`` products, at most ``maxiter`` cycles."""
        d = np.asarray(d, float)
        rhs = self.jvp_beta(d)
        n = rhs.size
        count = {"it": 0}

        def mv(v):
            count["it"] += 1
            return v - self.jvp_w(v)

        op = spla.LinearOperator((n, n), matvec=mv, dtype=np.float64)
        dw, info = spla.gmres(op, rhs, rtol=tol, restart=restart, maxiter=maxiter)
        res = float(np.linalg.norm(rhs - mv(dw)) / np.linalg.norm(rhs))
        hw = self._h(self.w0, dw)
        Jw = (self.drag(self.w0 + hw * dw, self.beta0) - self.drag(self.w0 - hw * dw, self.beta0)) / (2 * hw)
        hb = self._h(self.beta0, d)
        Jb = (self.drag(self.w0, self.beta0 + hb * d) - self.drag(self.w0, self.beta0 - hb * d)) / (2 * hb)
        self.unpack(self.w0)
        self.set_params(self.beta0)
        return Jb + Jw, {"gmres_matvecs": count["it"], "residual": res, "info": int(info)}
