"""Smoothed-aggregation AMG for the pressure-correction system (``flowtrace/linalg/amg.py``).

The algorithm is the reference's (``amg.py:80-146``): symmetric strength
``|a_ij| >= theta sqrt(|a_ii a_jj|)`` with theta = 0.08, three-pass greedy aggregation,
column-normalised tentative prolongator ``T``, ``rho(D^-1 A)`` from 15 power iterations
started at ``default_rng(0).standard_normal(n)``, smoothed prolongator
``P = T - 4/(3 rho) D^-1 A T``, Galerkin ``A_c = P^T A P``, levels until at most
``max_coarse`` unknowns (or until aggregation stops coarsening), and a dense inverse of the
coarsest matrix (shifted by ``1e-10 mean|diag|`` when the system is singular); V(2,2)
weighted-Jacobi cycles with omega = 0.8.

Set-up runs on the GPU (``ftmi_amgs_*``) except the greedy aggregation, which is
sequential by definition (its result depends on the node order) and runs on the host over
the strength graph copied back per level (the finest level's mask, one byte per entry of a
pattern the host already holds; a coarse level's strong entries as a CSR of their own). The
coarsest dense inverse (at most ``max_coarse`` squared entries) is computed on the host. A
rebuild for new values (``rebuild_ldu``) keeps every level whose strength mask did not change.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

from . import _native as nat

_RNG_CACHE = {}
_DEV_RNG_CACHE = {}


def start_vector(n: int) -> np.ndarray:
    """The power iteration's start vector, ``default_rng(0).standard_normal(n)``."""
    v = _RNG_CACHE.get(n)
    if v is None:
        v = np.random.default_rng(0).standard_normal(n)
        if len(_RNG_CACHE) < 64:
            _RNG_CACHE[n] = v
    return v


def start_vector_device(n: int, device) -> torch.Tensor:
    """A fresh device copy of ``start_vector(n)`` (the power iteration overwrites it); the
    uploaded vector is kept per (size, device), so a rebuild copies it within HBM."""
    key = (n, str(device))
    v = _DEV_RNG_CACHE.get(key)
    if v is None:
        v = torch.as_tensor(start_vector(n), device=device)
        if len(_DEV_RNG_CACHE) < 64:
            _DEV_RNG_CACHE[key] = v
    return v.clone()


def _pinned(n: int, dtype) -> torch.Tensor:
    return torch.empty(n, dtype=dtype, pin_memory=True)


class DevCSR:
    """A CSR matrix in HBM (int32 pattern, float64 values)."""

    def __init__(self, nrows, ncols, ptr, col, val):
        self.n, self.ncols, self.ptr, self.col, self.val = int(nrows), int(ncols), ptr, col, val

    @property
    def nnz(self) -> int:
        return int(self.col.numel())

    @classmethod
    def from_scipy(cls, A, device):
        A = sp.csr_matrix(A, dtype=np.float64)
        A.sort_indices()
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.int32), device=device)  # noqa: E731
        return cls(A.shape[0], A.shape[1], i32(A.indptr), i32(A.indices),
                   torch.as_tensor(np.ascontiguousarray(A.data), device=device))

    def to_scipy(self) -> sp.csr_matrix:
        return sp.csr_matrix((self.val.cpu().numpy(), self.col.cpu().numpy(), self.ptr.cpu().numpy()),
                             shape=(self.n, self.ncols))


def strength(A: sp.csr_matrix, theta: float) -> sp.csr_matrix:
    """Host form of the strength graph (tests and diagnostics)."""
    C = A.tocoo()
    dg = np.abs(A.diagonal())
    keep = (C.row != C.col) & (np.abs(C.data) >= theta * np.sqrt(dg[C.row] * dg[C.col]))
    S = sp.csr_matrix((np.ones(int(keep.sum())), (C.row[keep], C.col[keep])), shape=A.shape)
    S.sort_indices()
    return S


def aggregate(S: sp.csr_matrix):
    """Greedy aggregation of a host strength graph (``ftmi_amg_aggregate_host``)."""
    n = S.shape[0]
    ptr = np.ascontiguousarray(S.indptr, np.int64)
    idx = np.ascontiguousarray(S.indices, np.int64)
    agg = np.empty(n, np.int64)
    na = nat.lib().ftmi_amg_aggregate_host(n, ptr.ctypes.data, idx.ctypes.data, agg.ctypes.data)
    return agg, int(na)


class SmoothedAggregationAMG:
    """A device-resident SA-AMG hierarchy; ``amg(b)`` applies one V-cycle on the GPU.

    ``A`` is a host SciPy matrix or a ``DevCSR``; ``from_ldu`` builds the hierarchy of a
    pressure matrix held in the face (LDU) form of an ``FVMesh``."""

    debug = False

    def __init__(self, A, device="cuda", theta=0.08, max_coarse=300, max_levels=20, sweeps=2, omega=0.8,
                 singular=False, host_pattern=None, refreshable=False, coarse_sweeps=0):
        self.device = torch.device(device)
        self.refreshable = refreshable  # keep what ``refresh`` needs (the transpose map of each P)
        self.theta, self.max_coarse, self.max_levels = theta, max_coarse, max_levels
        # coarse_sweeps > 0: when the level count cuts the hierarchy above max_coarse unknowns,
        # the coarsest system is solved by that many weighted-Jacobi sweeps from zero instead of
        # a dense inverse (the reference always inverts it densely, amg.py:104-112)
        self.coarse_sweeps = int(coarse_sweeps)
        if self.coarse_sweeps == 1:
            raise ValueError("coarse_sweeps must be 0 (dense inverse) or at least 2")
        self.omega, self.sweeps, self.singular = omega, sweeps, singular
        A0 = A if isinstance(A, DevCSR) else DevCSR.from_scipy(A, self.device)
        if host_pattern is None and not isinstance(A, DevCSR):
            Ah = sp.csr_matrix(A)
            Ah.sort_indices()
            host_pattern = (np.ascontiguousarray(Ah.indptr, np.int32), np.ascontiguousarray(Ah.indices, np.int32))
        self._setup(A0, host_pattern)

    @classmethod
    def from_ldu(cls, fv, diag, off, singular=False, **kw):
        L = nat.lib()
        data = torch.empty(fv.nnz, dtype=torch.float64, device=fv.device)
        nat.check(L.ftmi_amgs_ldu_to_csr(fv.nc, nat.ptr(fv.cptr), nat.ptr(fv.he_slot), nat.ptr(fv.slot_diag),
                                         nat.ptr(fv.indptr), nat.ptr(diag), nat.ptr(off), nat.ptr(data),
                                         nat.stream_ptr()), "ldu_to_csr")
        A0 = DevCSR(fv.nc, fv.nc, fv.indptr, fv.indices, data)
        return cls(A0, fv.device, singular=singular, host_pattern=(fv.indptr_h, fv.indices_h), **kw)

    # ------------------------------------------------------------------ set-up
    def _spgemm(self, X: DevCSR, Y: DevCSR, xscale=None, agg=None, tval=None, omega_num=None, omega_scale=0.0,
                keep=None):
        """C = X Y (count pass, scan, fill pass). With ``keep`` (a dict) the count-pass work is kept
        there, so ``_spgemm_fill`` can recompute the values of the same pattern later."""
        L, st, dev = nat.lib(), nat.stream_ptr(), self.device
        n = X.n
        cnt = torch.zeros(n, dtype=torch.int32, device=dev)  # rows over the hash capacity stay 0
        long_rows = torch.empty(n, dtype=torch.int32, device=dev)
        n_long = torch.zeros(1, dtype=torch.int32, device=dev)
        if keep is not None:
            keep.update(cnt=cnt, long_rows=long_rows, n_long=n_long)
        args = (n, Y.ncols, nat.ptr(X.ptr), nat.ptr(X.col), nat.ptr(X.val), nat.ptr(Y.ptr), nat.ptr(Y.col),
                nat.ptr(Y.val),
                nat.ptr(xscale), nat.ptr(agg), nat.ptr(tval), nat.ptr(omega_num), float(omega_scale))
        work = (nat.ptr(self._ovf), nat.ptr(long_rows), nat.ptr(n_long), st)
        nat.check(L.ftmi_amgs_spgemm(0, *args, nat.ptr(cnt), None, None, None, *work), "spgemm")
        ptr = self._scan(cnt)
        nnz = int(ptr[n].item())
        col = torch.empty(nnz, dtype=torch.int32, device=dev)
        val = torch.empty(nnz, dtype=torch.float64, device=dev)
        nat.check(L.ftmi_amgs_spgemm(1, *args, nat.ptr(cnt), nat.ptr(ptr), nat.ptr(col), nat.ptr(val), *work),
                  "spgemm")
        if self.debug:
            o = self._ovf.cpu().numpy()
            print(f"[amg] spgemm rows {n} nnz {nnz} overflow {o.tolist()}", flush=True)
        return DevCSR(n, Y.ncols, ptr, col, val)

    def _spgemm_fill(self, X: DevCSR, Y: DevCSR, C: DevCSR, keep: dict, xscale=None, agg=None, tval=None,
                     omega_num=None, omega_scale=0.0):
        """Values of C = X Y into C's existing pattern (the fill pass alone; no host read)."""
        args = (X.n, Y.ncols, nat.ptr(X.ptr), nat.ptr(X.col), nat.ptr(X.val), nat.ptr(Y.ptr), nat.ptr(Y.col),
                nat.ptr(Y.val), nat.ptr(xscale), nat.ptr(agg), nat.ptr(tval), nat.ptr(omega_num), float(omega_scale))
        nat.check(nat.lib().ftmi_amgs_spgemm(1, *args, nat.ptr(keep["cnt"]), nat.ptr(C.ptr), nat.ptr(C.col),
                                             nat.ptr(C.val), nat.ptr(self._ovf), nat.ptr(keep["long_rows"]),
                                             nat.ptr(keep["n_long"]), nat.stream_ptr()), "spgemm fill")

    def _scan(self, cnt):
        n = cnt.numel()
        out = torch.empty(n + 1, dtype=torch.int32, device=self.device)
        nat.check(nat.lib().ftmi_scan_i32(n, nat.ptr(cnt), nat.ptr(out), nat.ptr(self._bsum), nat.stream_ptr()), "scan")
        return out

    def _transpose(self, P: DevCSR, with_perm: bool = False):
        """R = P^T (columns sorted in each row). With ``with_perm`` also the entry map perm with
        R.val[j] = P.val[perm[j]] (what a numeric refresh re-gathers R through): the transpose
        is then made of entry numbers once and R's values gathered through it."""
        L, st, dev = nat.lib(), nat.stream_ptr(), self.device
        na = P.ncols
        cnt = torch.zeros(na, dtype=torch.int32, device=dev)
        nat.check(L.ftmi_amgs_count_cols(P.nnz, nat.ptr(P.col), nat.ptr(cnt), st), "count_cols")
        rptr = self._scan(cnt)
        fill = torch.zeros(na, dtype=torch.int32, device=dev)
        rcol = torch.empty(P.nnz, dtype=torch.int32, device=dev)
        rval = torch.empty(P.nnz, dtype=torch.float64, device=dev)
        if not with_perm:
            nat.check(L.ftmi_amgs_transpose_fill(P.n, na, nat.ptr(P.ptr), nat.ptr(P.col), nat.ptr(P.val),
                                                 nat.ptr(rptr), nat.ptr(fill), nat.ptr(rcol), nat.ptr(rval), st),
                      "transpose")
            return DevCSR(na, P.n, rptr, rcol, rval)
        idx = torch.arange(P.nnz, dtype=torch.float64, device=dev)
        rnum = torch.empty(P.nnz, dtype=torch.float64, device=dev)
        nat.check(L.ftmi_amgs_transpose_fill(P.n, na, nat.ptr(P.ptr), nat.ptr(P.col), nat.ptr(idx), nat.ptr(rptr),
                                             nat.ptr(fill), nat.ptr(rcol), nat.ptr(rnum), st), "transpose perm")
        perm = rnum.to(torch.int64)
        nat.check(L.ftmi_gather_f64(P.nnz, nat.ptr(perm), nat.ptr(P.val), nat.ptr(rval), st), "gather")
        return DevCSR(na, P.n, rptr, rcol, rval), perm

    def _setup(self, A: DevCSR, host_pattern):
        import time

        dev = self.device
        t_start = time.perf_counter()
        self._reset_timing()
        self._ovf = torch.zeros(4, dtype=torch.int32, device=dev)
        self._bsum = torch.empty(2048, dtype=torch.int64, device=dev)
        self._partials = torch.empty(2048 * 8, dtype=torch.float64, device=dev)
        # rebuild_ldu's per-level "mask changed" flags (device) and their host copy
        # (slot max_levels: matches_ldu's "values differ" flag)
        self._flags = torch.zeros(self.max_levels + 1, dtype=torch.int32, device=dev)
        self._flags_h = _pinned(self.max_levels + 1, torch.int32).zero_()
        self._gen = 0
        self._stage = None
        self.levels, self.aggregates, self.rhos, self._plans, self._masks = [], [], [], [], []
        self._grow(A, host_pattern)
        self._check_overflow()
        self._prepare_coarse()
        self._upload()
        self.timing["total"] = time.perf_counter() - t_start

    def _reset_timing(self):
        # host seconds per phase: waiting for the strength mask, the greedy aggregation, the
        # whole set-up; levels whose aggregates a rebuild kept (equal strength masks)
        self.timing = {"wait_mask": 0.0, "aggregate": 0.0, "total": 0.0, "levels_kept": 0}

    def _strength_mask(self, A: DevCSR, compare_level=None):
        """D, D^-1 and the strength mask (one byte per entry of A) of level matrix A. With
        ``compare_level`` = k the mask is also compared with level k's kept mask on the device:
        ``_flags[k]`` then holds ``_gen`` where they differ."""
        L, st = nat.lib(), nat.stream_ptr()
        n = A.n
        diag = torch.empty(n, dtype=torch.float64, device=self.device)
        dinv = torch.empty(n, dtype=torch.float64, device=self.device)
        nat.check(L.ftmi_amgs_diag(n, A.nnz, nat.ptr(A.ptr), nat.ptr(A.col), nat.ptr(A.val), nat.ptr(diag),
                                   nat.ptr(dinv), st), "diag")
        mask = torch.empty(A.nnz, dtype=torch.uint8, device=self.device)
        if compare_level is None:
            nat.check(L.ftmi_amgs_strength(n, A.nnz, nat.ptr(A.ptr), nat.ptr(A.col), nat.ptr(A.val), nat.ptr(diag),
                                           float(self.theta), nat.ptr(mask), st), "strength")
        else:
            k = compare_level
            nat.check(L.ftmi_amgs_strength_cmp(n, A.nnz, nat.ptr(A.ptr), nat.ptr(A.col), nat.ptr(A.val),
                                               nat.ptr(diag), float(self.theta), nat.ptr(mask),
                                               nat.ptr(self._masks[k]), nat.ptr(self._flags) + 4 * k, self._gen,
                                               st), "strength (compared)")
        return diag, dinv, mask

    def _next_gen(self) -> int:
        self._gen = self._gen % (2 ** 31 - 1) + 1  # a value no flag holds yet (flags start at 0)
        return self._gen

    def _stage_ldu(self, fv, diag, off) -> DevCSR:
        """The CSR values of the face-form matrix (diag, off) into the staging buffer, beside the
        finest matrix's own values (which keep describing the hierarchy). Returns that matrix."""
        A0 = self.levels[0][0] if self.levels else self.coarse
        if self._stage is None or self._stage.numel() != A0.nnz:
            self._stage = torch.empty(A0.nnz, dtype=torch.float64, device=self.device)
        nat.check(nat.lib().ftmi_amgs_ldu_to_csr(fv.nc, nat.ptr(fv.cptr), nat.ptr(fv.he_slot), nat.ptr(fv.slot_diag),
                                                 nat.ptr(fv.indptr), nat.ptr(diag), nat.ptr(off),
                                                 nat.ptr(self._stage), nat.stream_ptr()), "ldu_to_csr")
        return A0

    def matches_ldu(self, fv, diag, off) -> bool:
        """Whether the face-form matrix (diag, off) is, bit for bit, the finest matrix this
        hierarchy was last set up or rebuilt for: then a set-up of it gives this hierarchy
        again, entry for entry, and it can be used as it is. One device pass, one host read."""
        L, st = nat.lib(), nat.stream_ptr()
        A0 = self._stage_ldu(fv, diag, off)
        gen, slot = self._next_gen(), self.max_levels
        nat.check(L.ftmi_differs_u64(A0.nnz, nat.ptr(self._stage), nat.ptr(A0.val), nat.ptr(self._flags) + 4 * slot,
                                     gen, st), "differs")
        self._flags_h[slot:].copy_(self._flags[slot:], non_blocking=True)
        torch.cuda.current_stream(self.device).synchronize()
        return int(self._flags_h[slot]) != gen

    def _changed(self, upto: int):
        """Which of levels 0..upto-1 have a new strength mask that differs from the kept one
        (one device-to-host copy of their flags)."""
        if upto == 0:
            return []
        self._flags_h[:upto].copy_(self._flags[:upto], non_blocking=True)
        torch.cuda.current_stream(self.device).synchronize()
        return [int(v) == self._gen for v in self._flags_h[:upto].tolist()]

    def _grow(self, A: DevCSR, pattern, first=None):
        """Add levels below the current ones, starting from level matrix A (whose host pattern
        is ``pattern`` when known), until the coarsening rule stops; A ends as the coarse
        matrix. ``first`` = (diag, dinv, mask) of A when already computed."""
        import time

        L, st, dev = nat.lib(), nat.stream_ptr(), self.device
        clock = time.perf_counter
        f64 = lambda *s: torch.empty(*s, dtype=torch.float64, device=dev)  # noqa: E731
        partials = self._partials
        while A.n > self.max_coarse and len(self.levels) < self.max_levels - 1:
            n = A.n
            diag, dinv, mask = first if first is not None else self._strength_mask(A)
            first = None
            # The strength graph goes to pinned host memory while the GPU runs the power
            # iteration, which does not depend on the aggregates: the mask when the host holds
            # the level's pattern (the finest level), else the strong entries alone as a CSR
            # of their own (a coarse level keeps a fraction of its entries).
            if pattern is not None:
                mask_p = _pinned(A.nnz, torch.uint8)
                mask_p.copy_(mask, non_blocking=True)
                strong = None
            else:
                cnt = torch.empty(n, dtype=torch.int32, device=dev)
                nat.check(L.ftmi_amgs_strong_count(n, nat.ptr(A.ptr), nat.ptr(mask), nat.ptr(cnt), st), "strong count")
                s_ptr = self._scan(cnt)
                s_nnz = int(s_ptr[n].item())
                s_col = torch.empty(max(s_nnz, 1), dtype=torch.int32, device=dev)
                nat.check(L.ftmi_amgs_strong_fill(n, nat.ptr(A.ptr), nat.ptr(A.col), nat.ptr(mask), nat.ptr(s_ptr),
                                                  nat.ptr(s_col), st), "strong fill")
                sp_p, sc_p = _pinned(n + 1, torch.int32), _pinned(max(s_nnz, 1), torch.int32)
                sp_p.copy_(s_ptr, non_blocking=True)
                sc_p.copy_(s_col, non_blocking=True)
                strong = (sp_p.numpy(), sc_p.numpy())
            copied = torch.cuda.Event()
            copied.record()
            x = start_vector_device(n, dev)
            y, rho = f64(2 * n), f64(2)
            nat.check(L.ftmi_amgs_power(n, A.nnz, nat.ptr(A.ptr), nat.ptr(A.col), nat.ptr(A.val), nat.ptr(dinv),
                                        nat.ptr(x), nat.ptr(y), 15, nat.ptr(partials), nat.ptr(rho), st), "power")
            t0 = clock()
            copied.synchronize()
            t1 = clock()
            agg_p, size_p = _pinned(n, torch.int32), _pinned(n, torch.int32)
            agg_h, size_h = agg_p.numpy(), size_p.numpy()
            if strong is None:
                ptr_h, col_h = pattern
                mask_h = mask_p.numpy()
                na = int(L.ftmi_amg_aggregate_mask_host(n, ptr_h.ctypes.data, col_h.ctypes.data, mask_h.ctypes.data,
                                                        agg_h.ctypes.data, size_h.ctypes.data))
            else:
                na = int(L.ftmi_amg_aggregate_csr_host(n, strong[0].ctypes.data, strong[1].ctypes.data,
                                                       agg_h.ctypes.data, size_h.ctypes.data))
            self.timing["wait_mask"] += t1 - t0
            self.timing["aggregate"] += clock() - t1
            if na >= n * 0.9:
                break
            agg = torch.empty(n, dtype=torch.int32, device=dev)
            agg.copy_(agg_p, non_blocking=True)
            size = torch.empty(na, dtype=torch.int32, device=dev)
            size.copy_(size_p[:na], non_blocking=True)
            t_ptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
            t_val = f64(n)
            nat.check(L.ftmi_amgs_tentative(n, nat.ptr(agg), nat.ptr(size), nat.ptr(t_ptr), nat.ptr(t_val), st),
                      "tentative")
            T = DevCSR(n, na, t_ptr, agg, t_val)
            plan = {"T": T, "diag": diag, "y": y, "rho": rho, "kP": {}, "kAP": {}, "kRAP": {}}
            P = self._spgemm(A, T, xscale=dinv, agg=agg, tval=t_val, omega_num=rho, omega_scale=4.0 / 3.0,
                             keep=plan["kP"])
            if self.refreshable:
                R, plan["perm"] = self._transpose(P, with_perm=True)
            else:
                R, plan["perm"] = self._transpose(P), None
            AP = self._spgemm(A, P, keep=plan["kAP"])
            Ac = self._spgemm(R, AP, keep=plan["kRAP"])
            plan["AP"] = AP
            self._plans.append(plan)
            self._masks.append(mask)
            self.levels.append((A, P, R, dinv))
            self.aggregates.append(agg_h.astype(np.int64))  # a copy: the pinned buffer is recycled
            self.rhos.append(rho)
            A, pattern = Ac, None
        self.coarse = A

    def _check_overflow(self):
        ovf = self._ovf.cpu().numpy()
        if ovf[0]:
            raise RuntimeError(f"AMG set-up: {ovf[0]} sparse-product rows exceeded the LDS hash capacity "
                               f"(max distinct {ovf[1]}, max candidates {ovf[2]}, e.g. row {ovf[3]})")

    def rebuild_ldu(self, fv, diag, off, staged=False, keep=False) -> bool:
        """The hierarchy of a new pressure matrix (face form of ``fv``), equal to a full set-up
        of it (``from_ldu``) value for value: the greedy aggregation is a function of the level's
        pattern and strength mask alone, so wherever a level's new strength mask equals the one
        its aggregates came from, the aggregates (and, below them, every sparsity pattern) are
        those a fresh aggregation would give, and the level's values are recomputed on the kept
        patterns as ``refresh`` does. From the first level whose mask differs, the levels are
        set up anew. Needs a hierarchy set up with ``refreshable=True``.

        ``staged``: the new values are already in the staging buffer (``matches_ldu`` left them
        there). ``keep``: when the finest level's mask changes (every level would be set up
        anew), leave this hierarchy untouched and return False, so that the caller can set up a
        new one and keep this one; True otherwise."""
        import time

        if not self.refreshable:
            raise RuntimeError("this hierarchy was set up without refreshable=True")
        t_start = time.perf_counter()
        self._reset_timing()
        A0 = self.levels[0][0] if self.levels else self.coarse
        if not staged:
            self._stage_ldu(fv, diag, off)
        nlev = len(self.levels)
        # The finest level's mask is checked first (one host read), since refreshing the whole
        # hierarchy for nothing costs most when it changes. The coarser levels are then refreshed
        # on their kept patterns while their new masks are compared with the kept ones on the
        # device (one host read for all of them); a level below the first changed mask has then
        # received values computed on stale aggregates, and is set up anew below.
        self._next_gen()
        firsts, flags = [], []
        for k in range(nlev):
            if k == 0:
                first = self._strength_mask(DevCSR(A0.n, A0.ncols, A0.ptr, A0.col, self._stage), compare_level=0)
                changed = self._changed(1)[0]
                if changed and keep:
                    self.timing["total"] = time.perf_counter() - t_start
                    return False
                A0.val.copy_(self._stage)
                firsts.append(first)
                if changed:
                    flags = [True]
                    break
            else:
                first = self._strength_mask(self.levels[k][0], compare_level=k)
                firsts.append(first)
            self._refresh_level(k, first)
        else:
            if nlev == 0:
                A0.val.copy_(self._stage)
            flags = self._changed(nlev)
        k = flags.index(True) if True in flags else nlev
        first = firsts[k] if k < nlev else None
        if k < nlev or self.coarse.n > self.max_coarse:
            # new aggregates from level k on (or the coarse matrix may now coarsen further)
            A = self.levels[k][0] if k < nlev else self.coarse
            del self.levels[k:], self._plans[k:], self._masks[k:], self.aggregates[k:], self.rhos[k:]
            self._grow(A, (fv.indptr_h, fv.indices_h) if k == 0 else None, first)
            self._check_overflow()
        self.timing["levels_kept"] = k
        if k == nlev and len(self.levels) == nlev and self.coarse.n == self.coarse_n and \
                self._iterative() == self._coarse_iterative:
            self._update_coarse()  # the handle's pointers stay valid
        else:
            self._prepare_coarse()
            self._destroy_handle()
            self._upload()
        self.timing["total"] = time.perf_counter() - t_start
        return True

    # ------------------------------------------------------------------ coarse solve
    def _iterative(self) -> bool:
        """Whether the coarsest system is solved by sweeps: requested, below at least one level,
        and larger than max_coarse (the level count cut the hierarchy)."""
        return self.coarse_sweeps > 0 and bool(self.levels) and self.coarse.n > self.max_coarse

    def _prepare_coarse(self):
        """The coarse solve of the current coarsest matrix, in new buffers: its dense inverse
        (host) or, solved by sweeps, its inverse diagonal (device)."""
        self._coarse_iterative = self._iterative()
        c = self.coarse
        self.coarse_n = c.n
        if self._coarse_iterative:
            self.coarse_inv_h = None
            self._cdiag = torch.empty(c.n, dtype=torch.float64, device=self.device)
            self._cdinv = torch.empty(c.n, dtype=torch.float64, device=self.device)
            self._coarse_diag()
        else:
            self.coarse_inv_h = self._coarse_inverse()
            self.coarse_n = self.coarse_inv_h.shape[0]

    def _coarse_diag(self):
        """D^-1 of the coarsest matrix and the weight of its sweeps, 1 / (1.1 rho(D^-1 A_c)) by
        the set-up's power iteration: the sweeps then contract every mode (omega lambda < 1), which
        the smoother's 0.8 does not guarantee on Galerkin operators with rho above 2.5."""
        L, st, c = nat.lib(), nat.stream_ptr(), self.coarse
        nat.check(L.ftmi_amgs_diag(c.n, c.nnz, nat.ptr(c.ptr), nat.ptr(c.col), nat.ptr(c.val), nat.ptr(self._cdiag),
                                   nat.ptr(self._cdinv), st), "diag")
        x = start_vector_device(c.n, self.device)
        y = torch.empty(2 * c.n, dtype=torch.float64, device=self.device)
        rho = torch.empty(2, dtype=torch.float64, device=self.device)
        nat.check(L.ftmi_amgs_power(c.n, c.nnz, nat.ptr(c.ptr), nat.ptr(c.col), nat.ptr(c.val), nat.ptr(self._cdinv),
                                    nat.ptr(x), nat.ptr(y), 15, nat.ptr(self._partials), nat.ptr(rho), st), "power")
        self._comega = 1.0 / (1.1 * max(float(rho[0].item()), 1e-300))

    def _set_coarse_iter(self):
        c = self.coarse
        nat.check(nat.lib().ftmi_amg_set_coarse_iter(self.handle, self.coarse_n, c.nnz, nat.ptr(c.ptr), nat.ptr(c.col),
                                                     nat.ptr(c.val), nat.ptr(self._cdinv), self.coarse_sweeps,
                                                     float(self._comega), nat.ptr(self._cb), nat.ptr(self._cx),
                                                     nat.ptr(self._cy)), "amg_set_coarse_iter")

    def _update_coarse(self):
        """The coarse solve of new coarse values, in place (same size and kind)."""
        if self._coarse_iterative:
            self._coarse_diag()
            self._set_coarse_iter()  # the new weight
        else:
            self.coarse_inv_h = self._coarse_inverse()
            self._cinv.copy_(torch.as_tensor(np.ascontiguousarray(self.coarse_inv_h), device=self.device))

    def _coarse_inverse(self) -> np.ndarray:
        Ad = self.coarse.to_scipy().toarray()
        if self.singular:
            Ad = Ad + 1e-10 * np.abs(np.diag(Ad)).mean() * np.eye(Ad.shape[0])
        try:
            return np.linalg.inv(Ad)
        except np.linalg.LinAlgError:
            return np.linalg.pinv(Ad)

    # ------------------------------------------------------------------ numeric refresh
    def refresh(self):
        """Recompute every value of the hierarchy for new level-0 values (already in
        ``levels[0][0].val``), keeping the aggregates and all sparsity patterns: D^-1, rho, the
        smoothed prolongator, R, A P and R (A P) by the fill passes alone, and the coarse inverse.
        No aggregation and no host read except the small coarse matrix. The result is a valid
        smoothed-aggregation preconditioner of the new matrix built on the old aggregates (not
        the reference's hierarchy, which re-aggregates)."""
        if not self.refreshable:
            raise RuntimeError("this hierarchy was set up without refreshable=True")
        L, st = nat.lib(), nat.stream_ptr()
        for k, ((A, P, R, dinv), plan) in enumerate(zip(self.levels, self._plans)):
            nat.check(L.ftmi_amgs_diag(A.n, A.nnz, nat.ptr(A.ptr), nat.ptr(A.col), nat.ptr(A.val),
                                       nat.ptr(plan["diag"]), nat.ptr(dinv), st), "diag")
            self._refresh_level(k)
        self._update_coarse()

    def _refresh_level(self, k, first=None):
        """Values of level k's P, R, A P and of the next level's matrix from level k's current
        values, on the kept patterns. D and D^-1 are ``first``[:2] when given (moved into the
        level's buffers), else already current."""
        L, st = nat.lib(), nat.stream_ptr()
        A, P, R, dinv = self.levels[k]
        plan = self._plans[k]
        if first is not None:
            plan["diag"].copy_(first[0])
            dinv.copy_(first[1])
        x = start_vector_device(A.n, self.device)
        nat.check(L.ftmi_amgs_power(A.n, A.nnz, nat.ptr(A.ptr), nat.ptr(A.col), nat.ptr(A.val), nat.ptr(dinv),
                                    nat.ptr(x), nat.ptr(plan["y"]), 15, nat.ptr(self._partials),
                                    nat.ptr(plan["rho"]), st), "power")
        T = plan["T"]
        self._spgemm_fill(A, T, P, plan["kP"], xscale=dinv, agg=T.col, tval=T.val, omega_num=plan["rho"],
                          omega_scale=4.0 / 3.0)
        nat.check(L.ftmi_gather_f64(R.nnz, nat.ptr(plan["perm"]), nat.ptr(P.val), nat.ptr(R.val), st), "gather")
        self._spgemm_fill(A, P, plan["AP"], plan["kAP"])
        nxt = self.levels[k + 1][0] if k + 1 < len(self.levels) else self.coarse
        self._spgemm_fill(R, plan["AP"], nxt, plan["kRAP"])

    def refresh_ldu(self, fv, diag, off):
        """``refresh`` for a new pressure matrix given in the face (LDU) form of ``fv``."""
        A0 = self.levels[0][0] if self.levels else self.coarse
        nat.check(nat.lib().ftmi_amgs_ldu_to_csr(fv.nc, nat.ptr(fv.cptr), nat.ptr(fv.he_slot), nat.ptr(fv.slot_diag),
                                                 nat.ptr(fv.indptr), nat.ptr(diag), nat.ptr(off), nat.ptr(A0.val),
                                                 nat.stream_ptr()), "ldu_to_csr")
        self.refresh()

    def _upload(self):
        dev, L = self.device, nat.lib()
        self.handle = L.ftmi_amg_create(float(self.omega), int(self.sweeps))
        self._vecs = []
        for A, P, R, dinv in self.levels:
            # the V-cycle writes each level vector before reading it
            v = [torch.empty(A.n, dtype=torch.float64, device=dev) for _ in range(4)]
            self._vecs.append(v)
            nat.check(L.ftmi_amg_add_level(self.handle, A.n, A.nnz, P.nnz, nat.ptr(A.ptr), nat.ptr(A.col), nat.ptr(A.val),
                                           nat.ptr(P.ptr), nat.ptr(P.col), nat.ptr(P.val), nat.ptr(R.ptr),
                                           nat.ptr(R.col), nat.ptr(R.val), nat.ptr(dinv), *[nat.ptr(t) for t in v]),
                      "amg_add_level")
        nc = self.coarse_n
        self._cb = torch.zeros(nc, dtype=torch.float64, device=dev)
        self._cx = torch.zeros(nc, dtype=torch.float64, device=dev)
        if self._coarse_iterative:
            self._cinv = None
            self._cy = torch.zeros(nc, dtype=torch.float64, device=dev)
            self._set_coarse_iter()
        else:
            self._cinv = torch.as_tensor(np.ascontiguousarray(self.coarse_inv_h), device=dev)
            nat.check(L.ftmi_amg_set_coarse(self.handle, nc, nat.ptr(self._cinv), nat.ptr(self._cb),
                                            nat.ptr(self._cx)), "amg_set_coarse")

    def to_fp32(self) -> "AmgFP32":
        """The same hierarchy with FP32 values and vectors (the mixed-precision pressure solve)."""
        return AmgFP32(self)

    # ------------------------------------------------------------------ use
    @property
    def n_levels(self) -> int:
        return len(self.levels) + 1

    def prolongator(self, k: int) -> sp.csr_matrix:
        return self.levels[k][1].to_scipy()

    def operator_complexity(self) -> float:
        nnz = sum(A.nnz for A, _, _, _ in self.levels) + self.coarse.nnz
        return nnz / max(self.levels[0][0].nnz if self.levels else self.coarse.nnz, 1)

    def __call__(self, b: torch.Tensor) -> torch.Tensor:
        b = b.contiguous()
        x = torch.empty_like(b)
        nat.check(nat.lib().ftmi_amg_vcycle(self.handle, nat.ptr(b), nat.ptr(x), nat.stream_ptr()), "amg_vcycle")
        return x

    def _destroy_handle(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                nat.lib().ftmi_amg_destroy(h)
            except Exception:
                pass
            self.handle = None

    def __del__(self):
        self._destroy_handle()


class AmgFP32:
    """FP32 twin of a set-up ``SmoothedAggregationAMG``: the patterns are shared, the values
    (A, P, R, D^-1, the coarse inverse) are rounded once to FP32, and ``amg(b)`` applies one
    V-cycle in FP32 arithmetic. Holds a reference to its FP64 source, whose patterns it uses."""

    def __init__(self, src: SmoothedAggregationAMG):
        from .krylov import to_f32

        L, dev = nat.lib(), src.device
        self.src, self.device = src, dev
        self.handle = L.ftmi_amg_create_f32(float(src.omega), int(src.sweeps))
        self._keep, self._vals = [], []
        for A, P, R, dinv in src.levels:
            av, pv, rv, dv = to_f32(A.val), to_f32(P.val), to_f32(R.val), to_f32(dinv)
            self._vals.append((av, pv, rv, dv))
            v = [torch.empty(A.n, dtype=torch.float32, device=dev) for _ in range(4)]
            self._keep += [av, pv, rv, dv] + v
            nat.check(L.ftmi_amg_add_level_f32(self.handle, A.n, A.nnz, P.nnz, nat.ptr(A.ptr), nat.ptr(A.col), nat.ptr(av),
                                               nat.ptr(P.ptr), nat.ptr(P.col), nat.ptr(pv), nat.ptr(R.ptr),
                                               nat.ptr(R.col), nat.ptr(rv), nat.ptr(dv), *[nat.ptr(t) for t in v]),
                      "amg_add_level_f32")
        nc = src.coarse_n
        self._cb = torch.zeros(nc, dtype=torch.float32, device=dev)
        self._cx = torch.zeros(nc, dtype=torch.float32, device=dev)
        if src._coarse_iterative:
            self._cval, self._cdinv = to_f32(src.coarse.val), to_f32(src._cdinv)
            self._cy = torch.zeros(nc, dtype=torch.float32, device=dev)
            self._set_coarse_iter()
        else:
            self._cinv = self._coarse_f32()
            nat.check(L.ftmi_amg_set_coarse_f32(self.handle, nc, nat.ptr(self._cinv), nat.ptr(self._cb),
                                                nat.ptr(self._cx)), "amg_set_coarse_f32")

    def refresh(self):
        """Round the (refreshed) FP64 values of the source into this twin's FP32 buffers."""
        from .krylov import to_f32

        for (A, P, R, dinv), (av, pv, rv, dv) in zip(self.src.levels, self._vals):
            to_f32(A.val, av)
            to_f32(P.val, pv)
            to_f32(R.val, rv)
            to_f32(dinv, dv)
        if self.src._coarse_iterative:
            to_f32(self.src.coarse.val, self._cval)
            to_f32(self.src._cdinv, self._cdinv)
            self._set_coarse_iter()  # the source's new weight
        else:
            self._cinv.copy_(self._coarse_f32())

    def _set_coarse_iter(self):
        src, c = self.src, self.src.coarse
        nat.check(nat.lib().ftmi_amg_set_coarse_iter_f32(self.handle, src.coarse_n, c.nnz, nat.ptr(c.ptr),
                                                         nat.ptr(c.col), nat.ptr(self._cval), nat.ptr(self._cdinv),
                                                         src.coarse_sweeps, float(src._comega), nat.ptr(self._cb),
                                                         nat.ptr(self._cx), nat.ptr(self._cy)),
                  "amg_set_coarse_iter_f32")

    def _coarse_f32(self) -> torch.Tensor:
        """The coarse inverse in FP32. A regular system takes the FP64 inverse rounded. A singular
        one (the projected pressure system) takes the inverse on the range of the coarse matrix:
        the FP64 inverse of the shifted matrix has an eigenvalue of order 1e10 / |diag| along the
        null vector, harmless in FP64, but in FP32 it lifts the rounding error of a projected
        residual (1e-7 |r|) to a correction far above |r|, and the FP32 PCG stalls."""
        from .krylov import to_f32

        src = self.src
        if not src.singular:
            return to_f32(src._cinv)
        Ad = src.coarse.to_scipy().toarray()
        lam, V = np.linalg.eigh(0.5 * (Ad + Ad.T))
        keep = np.abs(lam) > 1e-8 * np.abs(lam).max()
        K = (V[:, keep] / lam[keep]) @ V[:, keep].T
        return torch.as_tensor(np.ascontiguousarray(K, np.float32), device=self.device)

    def __call__(self, b: torch.Tensor) -> torch.Tensor:
        b = b.contiguous()
        x = torch.empty_like(b)
        nat.check(nat.lib().ftmi_amg_vcycle_f32(self.handle, nat.ptr(b), nat.ptr(x), nat.stream_ptr()), "vcycle_f32")
        return x

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                nat.lib().ftmi_amg_destroy(h)
            except Exception:
                pass
            self.handle = None
