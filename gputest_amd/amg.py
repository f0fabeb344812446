"""Smoothed-aggregation AMG for the pressure-correction system (``flowtrace/linalg/amg.py``).

Set-up follows ``amg.py:80-117``: symmetric strength ``|a_ij| >= theta sqrt(a_ii a_jj)``,
three-pass greedy aggregation (``ftmi_amg_aggregate_host``), column-normalised tentative
prolongator ``T``, ``rho(D^-1 A)`` by 15 power iterations from ``default_rng(0)``, smoothed
prolongator ``P = T - 4/(3 rho) D^-1 A T``, Galerkin ``A_c = P^T A P``, until at most
``max_coarse`` unknowns remain, then a dense inverse (shifted by ``1e-10 mean|diag|`` when the
system is singular). The hierarchy is built on the host, like the reference's, and uploaded
to HBM; the V(2,2) weighted-Jacobi cycles run as HIP kernels (``ftmi_amg_vcycle``, and inside
``ftmi_pcg_ldu``).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

from . import _native as nat


def strength(A: sp.csr_matrix, theta: float) -> sp.csr_matrix:
    C = A.tocoo()
    dg = np.abs(A.diagonal())
    keep = (C.row != C.col) & (np.abs(C.data) >= theta * np.sqrt(dg[C.row] * dg[C.col]))
    S = sp.csr_matrix((np.ones(int(keep.sum())), (C.row[keep], C.col[keep])), shape=A.shape)
    S.sort_indices()
    return S


def aggregate(S: sp.csr_matrix):
    n = S.shape[0]
    ptr = np.ascontiguousarray(S.indptr, np.int64)
    idx = np.ascontiguousarray(S.indices, np.int64)
    agg = np.empty(n, np.int64)
    na = nat.lib().ftmi_amg_aggregate_host(n, ptr.ctypes.data, idx.ctypes.data, agg.ctypes.data)
    return agg, int(na)


_RNG_CACHE = {}


def _start_vector(n: int) -> np.ndarray:
    v = _RNG_CACHE.get(n)
    if v is None:
        v = np.random.default_rng(0).standard_normal(n)
        if len(_RNG_CACHE) < 64:
            _RNG_CACHE[n] = v
    return v


def spectral_radius(M: sp.csr_matrix, iters: int = 15) -> float:
    x = _start_vector(M.shape[0]).copy()
    lam = 1.0
    for _ in range(iters):
        y = M @ x
        lam = np.linalg.norm(y) / max(np.linalg.norm(x), 1e-300)
        x = y / max(np.linalg.norm(y), 1e-300)
    return lam


class SmoothedAggregationAMG:
    """A device-resident SA-AMG hierarchy; ``amg(b)`` applies one V-cycle on the GPU."""

    def __init__(self, A: sp.csr_matrix, device="cuda", theta=0.08, max_coarse=300, max_levels=20, sweeps=2,
                 omega=0.8, singular=False):
        self.device = torch.device(device)
        self.omega, self.sweeps = omega, sweeps
        A = sp.csr_matrix(A, dtype=np.float64)
        host = []
        while A.shape[0] > max_coarse and len(host) < max_levels - 1:
            S = strength(A, theta)
            agg, na = aggregate(S)
            if na >= A.shape[0] * 0.9:
                break
            n = A.shape[0]
            T = sp.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, na))
            cnt = np.asarray(T.sum(axis=0)).ravel()
            T = T @ sp.diags(1.0 / np.sqrt(cnt))
            Dinv = 1.0 / A.diagonal()
            DA = sp.diags(Dinv) @ A
            rho = spectral_radius(DA)
            P = (T - (4.0 / 3.0 / rho) * (DA @ T)).tocsr()
            host.append((A, P, Dinv, agg))
            A = (P.T @ A @ P).tocsr()
        self.coarse_n = A.shape[0]
        Ad = A.toarray()
        if singular:
            Ad = Ad + 1e-10 * np.abs(np.diag(Ad)).mean() * np.eye(Ad.shape[0])
        try:
            cinv = np.linalg.inv(Ad)
        except np.linalg.LinAlgError:
            cinv = np.linalg.pinv(Ad)
        self.aggregates = [h[3] for h in host]
        self.host_levels = [(h[0], h[1], h[2]) for h in host]
        self.coarse_inv_h = cinv
        self._upload(host, cinv)

    def _upload(self, host, cinv):
        dev = self.device
        f64 = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float64), device=dev)  # noqa: E731
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.int32), device=dev)  # noqa: E731
        L = nat.lib()
        self._keep = []
        self.handle = L.ftmi_amg_create(float(self.omega), int(self.sweeps))
        for A, P, Dinv, _ in host:
            P.sort_indices()
            R = P.T.tocsr()
            R.sort_indices()
            n = A.shape[0]
            t = {"a": (i32(A.indptr), i32(A.indices), f64(A.data)), "p": (i32(P.indptr), i32(P.indices), f64(P.data)),
                 "r": (i32(R.indptr), i32(R.indices), f64(R.data)), "dinv": f64(Dinv),
                 "v": [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(4)]}
            self._keep.append(t)
            nat.check(L.ftmi_amg_add_level(self.handle, n, *[nat.ptr(x) for x in t["a"]], *[nat.ptr(x) for x in t["p"]],
                                           *[nat.ptr(x) for x in t["r"]], nat.ptr(t["dinv"]),
                                           *[nat.ptr(v) for v in t["v"]]), "amg_add_level")
        nc = cinv.shape[0]
        c = {"inv": f64(cinv), "b": torch.zeros(nc, dtype=torch.float64, device=dev),
             "x": torch.zeros(nc, dtype=torch.float64, device=dev)}
        self._keep.append(c)
        nat.check(L.ftmi_amg_set_coarse(self.handle, nc, nat.ptr(c["inv"]), nat.ptr(c["b"]), nat.ptr(c["x"])),
                  "amg_set_coarse")

    @property
    def n_levels(self) -> int:
        return len(self.host_levels) + 1

    def operator_complexity(self) -> float:
        nnz = sum(A.nnz for A, _, _ in self.host_levels) + self.coarse_n ** 2
        return nnz / max(self.host_levels[0][0].nnz if self.host_levels else 1, 1)

    def __call__(self, b: torch.Tensor) -> torch.Tensor:
        b = b.contiguous()
        x = torch.empty_like(b)
        nat.check(nat.lib().ftmi_amg_vcycle(self.handle, nat.ptr(b), nat.ptr(x), nat.stream_ptr()), "amg_vcycle")
        return x

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                nat.lib().ftmi_amg_destroy(h)
            except Exception:
                pass
            self.handle = None
