"""ORACLE INFRASTRUCTURE (CPU baseline timing only; never part of the product path).

The reference's own CPU path, for timing it on the host cores beside the GPU: the reference
runs its NumPy SIMPLEC with the hot operators dispatched to its C++/OpenMP kernels
(``flowtrace/native/fvm_kernels.cpp``, bound in ``flowtrace/native/__init__.py:97-113`` and
dispatched from ``flowtrace/unstructured/fvmesh.py:208-348`` and ``simple.py:489-497``). Here
the oracle's restatement of the NumPy host code (``oracle.simplec_oracle``) calls the
reference's kernels compiled from that source file into ``oracle/_ref/libflowtrace_fvm.so``
(``oracle/Makefile``): the LS and Green-Gauss gradients, the face interpolation, the CSR
assembly, every incidence and matrix product (``ft_csr_spmm``) and the momentum/energy
Jacobi-BiCGSTAB. The AMG set-up and the pressure PCG stay in NumPy/SciPy, as in the
reference's CPU path without pyamg. Threads: the kernels' OpenMP team (``OMP_NUM_THREADS``).

Only ``bench.py``'s ``cpu_baseline`` leg and ``tests/`` use this module.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import scipy.sparse as sp

from oracle import simplec_oracle as O

REF_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libflowtrace_fvm.so")
_P, _I64, _I = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
_SIGS = {"ft_max_threads": ([], _I), "ft_csr_spmm": ([_I64, _I64] + [_P] * 5, None),
         "ft_face_interp": ([_I64, _I64] + [_P] * 5, None), "ft_ls_gradient": ([_I64, _I64] + [_P] * 14, None),
         "ft_gg_gradient": ([_I64, _I64] + [_P] * 16, None), "ft_assemble": ([_I64, _I64, _I64] + [_P] * 7, None),
         "ft_bicgstab_jacobi": ([_I64] + [_P] * 6 + [ctypes.c_double, _I64, _I] + [_P] * 3, _I)}
_lib = None


def available() -> bool:
    return os.path.exists(REF_LIB)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(REF_LIB)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes, fn.restype = args, res
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class KernelCSR(sp.csr_matrix):
    """A SciPy CSR matrix whose products with float64 vectors and (n, k) blocks run in the
    reference's ``ft_csr_spmm``; every other operation is SciPy's."""

    def _native(self):
        if getattr(self, "_ip", None) is None or self._ip.size != self.indptr.size:
            self.sort_indices()
            self._ip = np.ascontiguousarray(self.indptr, np.int32)
            self._ix = np.ascontiguousarray(self.indices, np.int32)
        self._dv = _f64(self.data)
        return self._ip, self._ix, self._dv

    def __matmul__(self, x):
        if isinstance(x, np.ndarray) and x.dtype == np.float64 and x.ndim in (1, 2) and x.shape[0] == self.shape[1]:
            ip, ix, dv = self._native()
            x = np.ascontiguousarray(x)
            k = 1 if x.ndim == 1 else x.shape[1]
            y = np.empty((self.shape[0],) if x.ndim == 1 else (self.shape[0], k))
            lib().ft_csr_spmm(self.shape[0], k, _p(ip), _p(ix), _p(dv), _p(x), _p(y))
            return y
        return sp.csr_matrix.__matmul__(sp.csr_matrix(self), x)


def _kernel_csr(m) -> KernelCSR:
    out = KernelCSR(m)
    out.sort_indices()
    return out


class KernelFV(O.FV):
    """``oracle.simplec_oracle.FV`` with its operators in the reference's C++ kernels (the
    dispatch of ``fvmesh.py:268-348`` when the native library is loaded)."""

    def attach_kernels(self):
        self.Cown, self.Cnei = _kernel_csr(self.Cown), _kernel_csr(self.Cnei)
        self.Cb, self.Cdiff = _kernel_csr(self.Cb), _kernel_csr(self.Cdiff)
        i32, i64 = (lambda a: np.ascontiguousarray(a, np.int32)), (lambda a: np.ascontiguousarray(a, np.int64))
        cd, cb = self.Cdiff, self.Cb
        self._k = {"iptr": i32(cd.indptr), "iidx": i32(cd.indices), "isgn": _f64(cd.data), "bptr": i32(cb.indptr),
                   "bidx": i32(cb.indices), "P": i64(self.P), "N": i64(self.N), "w": _f64(self.w),
                   "wl": _f64(self.wl), "d": _f64(self.d), "wlb": _f64(self.wlb), "db": _f64(self.db),
                   "Ginv": _f64(self.Ginv), "S": _f64(self.S), "Sb": _f64(self.Sb), "vol": _f64(self.vol),
                   "dxm": _f64(self.dxm), "slot_diag": i32(self.slot_diag), "slot_PN": i32(self.slot_PN),
                   "slot_NP": i32(self.slot_NP)}

    def gradient(self, phi, phib):
        k = self._k
        phi, phib = _f64(phi), _f64(phib)  # held here: ctypes gets bare addresses
        out = np.empty((self.nc, self.dim))
        lib().ft_ls_gradient(self.nc, self.dim, _p(k["iptr"]), _p(k["iidx"]), _p(k["bptr"]), _p(k["bidx"]),
                             _p(k["P"]), _p(k["N"]), _p(k["wl"]), _p(k["d"]), _p(k["wlb"]), _p(k["db"]),
                             _p(k["Ginv"]), _p(phi), _p(phib), _p(out))
        return out

    def gradient_gg(self, phi, phib, grad=None):
        k = self._k
        phi, phib = _f64(phi), _f64(phib)
        g = _f64(grad) if grad is not None else None
        out = np.empty((self.nc, self.dim))
        lib().ft_gg_gradient(self.nc, self.dim, _p(k["iptr"]), _p(k["iidx"]), _p(k["isgn"]), _p(k["bptr"]),
                             _p(k["bidx"]), _p(k["P"]), _p(k["N"]), _p(k["w"]), _p(k["S"]), _p(k["Sb"]),
                             _p(k["vol"]), _p(k["dxm"]), _p(phi), _p(phib), _p(g) if g is not None else None,
                             _p(out))
        return out

    def _interp(self, a, width):
        a = _f64(a)
        out = np.empty((self.n_if,) if width == 1 else (self.n_if, width))
        lib().ft_face_interp(self.n_if, width, _p(self._k["w"]), _p(self._k["P"]), _p(self._k["N"]), _p(a), _p(out))
        return out

    def face_interp(self, phi):
        return self._interp(phi, 1)

    def face_interp_vec(self, g):
        return self._interp(g, g.shape[1])

    def matrix(self, diag, aPN, aNP):
        k = self._k
        diag, aPN, aNP = _f64(diag), _f64(aPN), _f64(aNP)
        data = np.empty(self.nnz)
        lib().ft_assemble(self.nnz, self.nc, self.n_if, _p(k["slot_diag"]), _p(k["slot_PN"]), _p(k["slot_NP"]),
                          _p(diag), _p(aPN), _p(aNP), _p(data))
        A = KernelCSR((data, self.indices, self.indptr), shape=(self.nc, self.nc))
        return A


def _kernel_bicgstab(A, b, x0, diag, tol=0.1, maxiter=30):
    """``SimpleSolver._solve`` on the native path (``simple.py:489-497``, ``ft_bicgstab_jacobi``)."""
    ip, ix, dv = A._native()
    n = A.shape[0]
    x = np.array(x0, dtype=np.float64, copy=True, order="C")
    inv, b = _f64(1.0 / diag), _f64(b)
    work = np.empty(7 * n)
    it, res = ctypes.c_int64(0), ctypes.c_double(0.0)
    ok = lib().ft_bicgstab_jacobi(n, _p(ip), _p(ix), _p(dv), _p(b), _p(inv), _p(x), float(tol), int(maxiter),
                                  1, _p(work), ctypes.byref(it), ctypes.byref(res))
    return x, O.KrylovInfo(int(it.value), float(res.value), bool(ok))


def with_reference_kernels(s: "O.Simplec") -> "O.Simplec":
    """Switch an oracle solver to the reference's compiled kernels (in place)."""
    s.fv.__class__ = KernelFV
    s.fv.attach_kernels()
    s._solve = _kernel_bicgstab
    return s


def threads() -> int:
    return int(lib().ft_max_threads())
