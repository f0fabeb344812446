"""Device-resident Krylov solvers (``flowtrace/linalg/krylov.py``) on HIP kernels.

``bicgstab_jacobi`` is the reference's native entry (``flowtrace/native/__init__.py:220-237``)
on a device CSR matrix; ``bicgstab_ldu`` and ``pcg_ldu`` run on the face (LDU) form the
SIMPLEC stages assemble. All return ``(x, KrylovInfo)`` like the reference.

The mixed-precision forms (no reference counterpart: the reference runs one precision
throughout, ``simple.py:77``) keep FP32 matrices and vectors inside the iterations:
``bicgstab_ldu_mixed`` solves for the FP32 correction of an FP64 iterate, ``pcg_ldu_mixed``
refines an FP64 pressure correction with FP32 AMG-PCG solves until the FP64 residual meets
the tolerance.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _native as nat


@dataclass
class KrylovInfo:
    iterations: int
    residual: float  # final relative residual
    converged: bool
    rounds: int = 0  # mixed precision: FP32 corrections applied


class Workspace:
    """Reusable device work buffers keyed by (kind, n)."""

    def __init__(self, device):
        self.device = torch.device(device)
        self._bufs = {}

    def get(self, kind: str, n: int) -> torch.Tensor:
        key = (kind, n)
        t = self._bufs.get(key)
        if t is None:
            L = nat.lib()
            nbytes = L.ftmi_bicgstab_work_bytes(n) if kind == "bicgstab" else L.ftmi_pcg_work_bytes(n)
            t = torch.zeros(int(nbytes), dtype=torch.uint8, device=self.device)
            self._bufs[key] = t
        return t


_default_ws = {}


def _ws(device) -> Workspace:
    key = str(device)
    if key not in _default_ws:
        _default_ws[key] = Workspace(device)
    return _default_ws[key]


def _run(fn, *args):
    it, res = ctypes.c_int64(0), ctypes.c_double(0.0)
    st = fn(*args, ctypes.byref(it), ctypes.byref(res), nat.stream_ptr())
    nat.check(st, "krylov")
    return KrylovInfo(int(it.value), float(res.value), bool(st == 1))


def bicgstab_jacobi(A, b, x0, inv_diag, tol: float, maxiter: int, relative_to_initial: bool = True, ws=None):
    """Jacobi-right-preconditioned BiCGSTAB on a ``DeviceCSR`` (``krylov.py:68-121`` rules)."""
    x = x0.clone().contiguous()
    n = b.shape[0]
    work = (ws or _ws(b.device)).get("bicgstab", n)
    info = _run(nat.lib().ftmi_bicgstab_jacobi, n, nat.ptr(A.indptr), nat.ptr(A.indices), nat.ptr(A.data),
                nat.ptr(b.contiguous()), nat.ptr(inv_diag.contiguous()), nat.ptr(x), float(tol), int(maxiter),
                int(bool(relative_to_initial)), nat.ptr(work))
    return x, info


def bicgstab_ldu(fv, diag, off, b, inv_diag, x, tol: float, maxiter: int, relative_to_initial=True, ws=None,
                 sliced_off=None):
    """In place on ``x`` (x0 in, solution out); the operator is (diag, off) on ``fv``'s adjacency.
    With ``sliced_off`` (``off`` already in ``fv``'s slice layout, ``fv.slice_values``) the solve
    reads the sliced operator: the same sums, coalesced loads."""
    n = b.shape[0]
    work = (ws or _ws(b.device)).get("bicgstab", n)
    if sliced_off is not None:
        return _run(nat.lib().ftmi_bicgstab_ldu_sliced, n, nat.ptr(diag), nat.ptr(fv.sl_base), nat.ptr(fv.sl_deg),
                    nat.ptr(fv.sl_nbr), nat.ptr(sliced_off), nat.ptr(b), nat.ptr(inv_diag), nat.ptr(x), float(tol),
                    int(maxiter), int(bool(relative_to_initial)), nat.ptr(work))
    return _run(nat.lib().ftmi_bicgstab_ldu, n, nat.ptr(diag), nat.ptr(fv.cptr), nat.ptr(fv.cnbr), nat.ptr(off),
                nat.ptr(b), nat.ptr(inv_diag), nat.ptr(x), float(tol), int(maxiter), int(bool(relative_to_initial)),
                nat.ptr(work))


def pcg_ldu(fv, diag, off, b, x, tol: float, maxiter: int, amg=None, project=False, ws=None, sliced_off=None):
    """PCG from zero into ``x``, preconditioned by one V-cycle of ``amg`` (``krylov.py:25-65``).
    With ``sliced_off`` (``off`` already in ``fv``'s slice layout, ``fv.slice_values``) the
    operator products read the sliced half-edges: the same sums, coalesced loads."""
    n = b.shape[0]
    work = (ws or _ws(b.device)).get("pcg", n)
    if sliced_off is not None:
        return _run(nat.lib().ftmi_pcg_ldu_sliced, n, nat.ptr(diag), nat.ptr(fv.sl_base), nat.ptr(fv.sl_deg),
                    nat.ptr(fv.sl_nbr), nat.ptr(sliced_off), nat.ptr(b), nat.ptr(x), float(tol), int(maxiter),
                    amg.handle if amg is not None else None, int(bool(project)), nat.ptr(work))
    return _run(nat.lib().ftmi_pcg_ldu, n, nat.ptr(diag), nat.ptr(fv.cptr), nat.ptr(fv.cnbr), nat.ptr(off),
                nat.ptr(b), nat.ptr(x), float(tol), int(maxiter), amg.handle if amg is not None else None,
                int(bool(project)), nat.ptr(work))


def pcg_csr(A, b, tol: float, maxiter: int, amg=None, project=False, ws=None):
    n = b.shape[0]
    x = torch.empty_like(b)
    work = (ws or _ws(b.device)).get("pcg", n)
    info = _run(nat.lib().ftmi_pcg_csr, n, nat.ptr(A.indptr), nat.ptr(A.indices), nat.ptr(A.data),
                nat.ptr(b.contiguous()), nat.ptr(x), float(tol), int(maxiter),
                amg.handle if amg is not None else None, int(bool(project)), nat.ptr(work))
    return x, info


def bicgstab_ldu_mixed(fv, ex32, off32, b32, inv32, x, tol: float, maxiter: int, relative_to_initial=True,
                       ws=None, b_is_residual: bool = False, sliced_off=None):
    """FP32 Jacobi-BiCGSTAB for the correction of the FP64 iterate ``x`` (updated in place);
    the operator is (ex32, off32) in difference form, ex32 holding the row sums. ``b32`` is the
    right-hand side, or with ``b_is_residual`` the residual b - A x already formed (the
    mixed-precision transport assembly writes it that way). With ``sliced_off`` (``off32`` in
    ``fv``'s slice layout, ``fv.slice_values``) and ``b_is_residual`` the solve reads the
    sliced operator: the same sums, coalesced loads."""
    n = b32.shape[0]
    work = (ws or _ws(b32.device)).get("bicgstab", n)
    if sliced_off is not None and b_is_residual:
        return _run(nat.lib().ftmi_bicgstab_ldu_mixed_sliced, n, nat.ptr(ex32), nat.ptr(fv.sl_base),
                    nat.ptr(fv.sl_deg), nat.ptr(fv.sl_nbr), nat.ptr(sliced_off), nat.ptr(b32), nat.ptr(inv32),
                    nat.ptr(x), float(tol), int(maxiter), int(bool(relative_to_initial)), nat.ptr(work))
    return _run(nat.lib().ftmi_bicgstab_ldu_mixed, n, nat.ptr(ex32), nat.ptr(fv.cptr), nat.ptr(fv.cnbr),
                nat.ptr(off32), nat.ptr(b32), nat.ptr(inv32), nat.ptr(x), float(tol), int(maxiter),
                int(bool(relative_to_initial)), int(bool(b_is_residual)), nat.ptr(work))


def pcg_ldu_mixed(fv, diag, off, off32, b, x, tol: float, maxiter: int, amg32=None, project=False,
                  max_refine: int = 4, ws=None):
    """FP64 iterative refinement of ``x`` (in place; pass zeros to start from zero) with FP32
    AMG-PCG corrections; ``off32`` is ``off`` rounded to FP32, ``amg32`` an
    ``SmoothedAggregationAMG.to_fp32()`` twin."""
    n = b.shape[0]
    work = (ws or _ws(b.device)).get("pcg", n)
    it, res, rounds = ctypes.c_int64(0), ctypes.c_double(0.0), ctypes.c_int64(0)
    st = nat.lib().ftmi_pcg_ldu_mixed(n, nat.ptr(diag), nat.ptr(fv.cptr), nat.ptr(fv.cnbr), nat.ptr(off),
                                      nat.ptr(off32), nat.ptr(b), nat.ptr(x), float(tol),
                                      int(maxiter), amg32.handle if amg32 is not None else None,
                                      int(bool(project)), int(max_refine), nat.ptr(work), ctypes.byref(it),
                                      ctypes.byref(res), ctypes.byref(rounds), nat.stream_ptr())
    nat.check(st, "pcg mixed")
    return KrylovInfo(int(it.value), float(res.value), bool(st == 1), int(rounds.value))


def pcg_ldu_f32(fv, diag32, off32, b32, x32, tol: float, maxiter: int, amg32=None, project=False, ws=None):
    """Plain FP32 AMG-PCG from zero into ``x32``."""
    n = b32.shape[0]
    work = (ws or _ws(b32.device)).get("pcg", n)
    return _run(nat.lib().ftmi_pcg_ldu_f32, n, nat.ptr(diag32), nat.ptr(fv.cptr), nat.ptr(fv.cnbr), nat.ptr(off32),
                nat.ptr(b32), nat.ptr(x32), float(tol), int(maxiter), amg32.handle if amg32 is not None else None,
                int(bool(project)), nat.ptr(work))


def to_f32(x: torch.Tensor, out: torch.Tensor = None) -> torch.Tensor:
    """FP32 copy of an FP64 device vector by the library's cast kernel."""
    if out is None:
        out = torch.empty(x.shape, dtype=torch.float32, device=x.device)
    nat.check(nat.lib().ftmi_cast_f64_f32(x.numel(), nat.ptr(x), nat.ptr(out), nat.stream_ptr()), "cast")
    return out
