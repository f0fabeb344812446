"""ORACLE (test infrastructure only; never part of the product path).

CPU reference for the discrete adjoint: the oracle's converged FP64 SIMPLEC state wrapped
in the attributes ``gputest_amd.adjoint.ResidualMap`` reads, with the two linear solves of
the adjoint products done by SciPy's sparse direct solver on the CPU instead of the HIP
Krylov solvers. Used by ``tests/`` to check the device adjoint against (1) the fixed-point
property of the restated outer iteration at the oracle's converged state, (2) central
finite differences of the oracle's converged drag (and, with the energy equation, outlet heat
flow) and (3) this CPU gradient.
"""
from __future__ import annotations

import types

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from gputest_amd.adjoint import ResidualMap
from gputest_amd.fvmesh import FVMesh


def solver_view(o, device="cpu"):
    """The attributes of a converged oracle ``Simplec`` that ``ResidualMap`` reads."""
    fv = FVMesh(o.mesh, device)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float64), device=device)  # noqa: E731
    v = types.SimpleNamespace()
    v.case, v.fv, v.mesh, v.device = o.case, fv, o.mesh, torch.device(device)
    v.dim, v.nc, v.energy = o.dim, o.nc, o.energy
    v.bkind_h, v.ub_fixed_h, v.pb_fixed_h = o.bkind, o.ub_fixed, o.pb_fixed
    v.singular_p, v.relax_p, v.relax_u, v.algo = o.singular_p, o.relax_p, o.relax_u, o.algo
    v.U, v.p, v.F, v.Fb = t(o.U.T), t(o.p), t(o.F), t(o.Fb)
    v.gp, v.gU = t(o._gp), t(np.stack(o._gU, axis=0))
    if o.energy:
        v.tkind_h, v.tval_h = o.tkind, o.tval
        v.T, v.gT = t(o.T), t(o._gT)
    return v


class CpuResidualMap(ResidualMap):
    """ResidualMap with direct sparse solves on the CPU (the reference adjoint)."""

    def _csr(self, diag, off):
        return self.fv.ldu_to_csr_host(diag.detach().cpu().numpy(), off.detach().cpu().numpy()).tocsc()

    def solve_momentum_T(self, diag, aPN, aNP, g):
        off = self._off(aNP.detach(), aPN.detach())
        x = spla.spsolve(self._csr(diag, off), g.detach().cpu().numpy())
        return torch.as_tensor(x, dtype=torch.float64, device=g.device)

    def _pressure_system(self, cf, cb):
        if self._amg is None:
            self._p_diag = self._sc(cf, cf, cb)
            self._p_off = self._off(-cf, -cf)
            A = self._csr(self._p_diag, self._p_off)
            if self.singular:
                A = A + sp.csc_matrix(np.ones((1, 1)) * 0.0, shape=A.shape)
            self._amg = spla.splu(A.tocsc())

    def solve_pressure(self, g):
        b = g.detach().cpu().numpy()
        if self.singular:
            b = b - b.mean()
        x = self._amg.solve(b)
        if self.singular:
            x = x - x.mean()
        return torch.as_tensor(x, dtype=torch.float64, device=g.device)
