"""Device-resident finite-volume mesh and the reference's operator surface on the GPU.

``FVMesh`` pre-computes, once on the host, the face and cell arrays of
``flowtrace/unstructured/fvmesh.py:61-197`` (periodic pairs become interior faces, the
over-relaxed split ``S = alpha d + T``, inverse-distance least-squares matrices, the fixed
cell-cell CSR pattern with face slots), adds the cell-centric face adjacency the fused
SIMPLEC kernels walk, and keeps everything in HBM as torch tensors on ``cuda``.

Its methods are the reference's operator surface (``fvmesh.py:254-356``: ``scatter``,
``net_outflow``, ``gradient``, ``gradient_gg``, ``face_interp``, ``face_interp_vec``,
``matrix``, ``matrix_like``); each runs a libftmi HIP kernel on the current stream.

Layout in HBM (float64 values, int32 indices):

* interior faces (nf): P, N, w, alpha, wl; S, d, Tvec, dP_f, dN_f, dxm as (nf, dim);
* boundary faces (nb): Pb, alpha_b, wlb, Sb_mag; Sb, db, db_t, nb as (nb, dim);
* cells (nc): vol, Ginv (nc, dim, dim);
* adjacency: cptr (nc+1), cface/cnbr (2 nf): a cell's interior faces in face order,
  cface = face << 1 | (cell is the neighbour); bptr (nc+1), bface (nb);
* CSR pattern of cell-cell matrices: indptr (nc+1), indices (nc + 2 nf), slot_diag (nc),
  and the per-row assembly list of ``ftmi_assemble``.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

from . import _native as nat
from .mesh import UnstructuredMesh, compute_geometry


def _inverse_2x2_3x3(G: np.ndarray) -> np.ndarray:
    """Batched inverse, pseudo-inverse where a matrix is singular (``fvmesh.py:25-39``)."""
    out = np.empty_like(G)
    ok = np.abs(np.linalg.det(G)) > 1e-300
    if ok.any():
        out[ok] = np.linalg.inv(G[ok])
        bad = ~ok | ~np.isfinite(out).all(axis=(1, 2))
    else:
        bad = np.ones(G.shape[0], dtype=bool)
    if bad.any():
        out[bad] = np.linalg.pinv(G[bad])
    return out


class DeviceCSR:
    """A CSR matrix in HBM; ``A @ x`` runs ``ftmi_csr_spmm``."""

    def __init__(self, data, indices, indptr, shape):
        self.data, self.indices, self.indptr, self.shape = data, indices, indptr, shape

    def __matmul__(self, x):
        x = x.contiguous()
        k = 1 if x.dim() == 1 else x.shape[1]
        y = torch.empty((self.shape[0],) if x.dim() == 1 else (self.shape[0], k), dtype=torch.float64,
                        device=x.device)
        nat.check(nat.lib().ftmi_csr_spmm(self.shape[0], k, nat.ptr(self.indptr), nat.ptr(self.indices),
                                          nat.ptr(self.data), nat.ptr(x), nat.ptr(y), nat.stream_ptr()), "csr_spmm")
        return y

    def to_scipy(self) -> sp.csr_matrix:
        return sp.csr_matrix((self.data.cpu().numpy(), self.indices.cpu().numpy(), self.indptr.cpu().numpy()),
                             shape=self.shape)


class FVMesh:
    def __init__(self, mesh: UnstructuredMesh, device="cuda", geometry=None):
        self.mesh = mesh
        self.device = torch.device(device)
        dim = self.dim = mesh.dim
        geo = self.geometry = geometry if geometry is not None else compute_geometry(mesh)
        nc = self.nc = mesh.n_cells
        own, nei = mesh.owner, mesh.neighbour
        inner = nei >= 0
        bnd = ~inner
        if mesh.periodic:
            pa = np.concatenate([np.asarray(p["faces_a"], np.int64) for p in mesh.periodic])
            pb = np.concatenate([np.asarray(p["faces_b"], np.int64) for p in mesh.periodic])
        else:
            pa = pb = np.zeros(0, np.int64)
        bnd[pa] = False
        bnd[pb] = False
        n_inner = int(inner.sum())
        self.if_idx = np.concatenate([np.flatnonzero(inner), pa])
        self.bf_idx = np.flatnonzero(bnd)
        P = np.concatenate([own[inner], own[pa]])
        N = np.concatenate([nei[inner], own[pb]])
        Pb = own[bnd]
        if np.any(P == N):
            raise ValueError("a periodic face connects a cell to itself: use at least 2 cells across the period")
        nf, nb = P.size, Pb.size
        self.n_if, self.n_bf = nf, nb
        self.periodic = []
        start = n_inner
        for p in mesh.periodic:
            n = len(p["faces_a"])
            self.periodic.append({"zones": (mesh.zones[p["zone_a"]].name, mesh.zones[p["zone_b"]].name),
                                  "faces": np.arange(start, start + n)})
            start += n
        xc = geo.cell_centre
        S, Sb = geo.face_area_vec[self.if_idx], geo.face_area_vec[bnd]
        xf, xfb = geo.face_centre[self.if_idx], geo.face_centre[bnd]
        xN = xc[N].copy()
        if mesh.periodic:
            xN[n_inner:] -= np.concatenate([np.repeat(np.asarray(p["translation"], float)[None, :dim],
                                                      len(p["faces_a"]), axis=0) for p in mesh.periodic])
        d = xN - xc[P]
        db = xfb - xc[Pb]
        dd = np.einsum("fd,fd->f", d, d)
        w = np.clip(np.einsum("fd,fd->f", xN - xf, d) / dd, 0.0, 1.0)
        Sd = np.einsum("fd,fd->f", S, d)
        SS = np.einsum("fd,fd->f", S, S)
        SSb = np.einsum("fd,fd->f", Sb, Sb)
        alpha = SS / Sd
        alpha_b = SSb / np.einsum("fd,fd->f", Sb, db)
        Tvec = S - alpha[:, None] * d
        self.non_orthogonality_deg = float(np.degrees(np.arccos(np.clip(Sd / np.sqrt(SS * dd), -1, 1))).max()) \
            if nf else 0.0
        wl = 1.0 / dd
        wlb = 1.0 / np.einsum("fd,fd->f", db, db)
        G = np.zeros((nc, dim, dim))
        for a in range(dim):
            for b in range(a, dim):
                o = wl * d[:, a] * d[:, b]
                G[:, a, b] = (np.bincount(P, o, minlength=nc) + np.bincount(N, o, minlength=nc)
                              + np.bincount(Pb, wlb * db[:, a] * db[:, b], minlength=nc))
                G[:, b, a] = G[:, a, b]
        Ginv = _inverse_2x2_3x3(G)
        del G
        Sb_mag = np.sqrt(SSb)
        nbu = Sb / Sb_mag[:, None]
        db_t = db - np.einsum("fd,fd->f", db, nbu)[:, None] * nbu
        dxm = xf - (w[:, None] * xc[P] + (1.0 - w)[:, None] * xN)
        self.bzone = mesh.face_zone[bnd]
        self.yb_h = np.abs(np.einsum("fd,fd->f", db, Sb)) / Sb_mag
        self.P_h, self.N_h, self.Pb_h = P, N, Pb
        self.xc_h, self.xfb_h, self.vol_h = xc, xfb, geo.cell_volume
        self.Sb_h, self.nb_h, self.alpha_b_h, self.Sb_mag_h = Sb, nbu, alpha_b, Sb_mag

        # --- adjacency: a cell's interior faces in face order (= rows of Cdiff), boundary faces
        cells = np.concatenate([P, N])
        faces = np.concatenate([np.arange(nf), np.arange(nf)])
        side = np.concatenate([np.zeros(nf, np.int64), np.ones(nf, np.int64)])
        other = np.concatenate([N, P])
        order = np.lexsort((faces, cells))
        cptr = np.zeros(nc + 1, np.int64)
        np.cumsum(np.bincount(cells, minlength=nc), out=cptr[1:])
        cface = (faces[order] << 1) | side[order]
        cnbr = other[order]
        border = np.argsort(Pb, kind="stable")
        bptr = np.zeros(nc + 1, np.int64)
        np.cumsum(np.bincount(Pb, minlength=nc), out=bptr[1:])
        bface = border
        # --- CSR pattern of cell-cell operators with face slots (fvmesh.py:147-162)
        rows = np.concatenate([np.arange(nc), P, N])
        cols = np.concatenate([np.arange(nc), N, P])
        pat = sp.csr_matrix((np.ones(rows.size, np.int8), (rows, cols)), shape=(nc, nc))
        pat.sum_duplicates()
        pat.sort_indices()
        keys = np.repeat(np.arange(nc, dtype=np.int64), np.diff(pat.indptr)) * nc + pat.indices
        self.nnz = int(pat.nnz)
        if self.nnz >= 2 ** 31 or 2 * nf >= 2 ** 31:
            raise ValueError("mesh too large for int32 indices")
        slot_diag = np.searchsorted(keys, np.arange(nc, dtype=np.int64) * (nc + 1)).astype(np.int32)
        slot_PN = np.searchsorted(keys, P.astype(np.int64) * nc + N).astype(np.int32)
        slot_NP = np.searchsorted(keys, N.astype(np.int64) * nc + P).astype(np.int32)
        indptr, indices = pat.indptr.astype(np.int32), pat.indices.astype(np.int32)
        src_ptr = np.empty(nc + 1, np.int32)
        src = np.empty(4 * nf, np.int32)
        nat.check(nat.lib().ftmi_assemble_plan_host(nc, nf, indptr.ctypes.data, slot_PN.ctypes.data,
                                                    slot_NP.ctypes.data, src_ptr.ctypes.data, src.ctypes.data),
                  "assemble_plan")
        # slot of each half-edge (for rebuilding CSR data from the LDU form on the host)
        he_slot = np.where(cface & 1, slot_NP[cface >> 1], slot_PN[cface >> 1])
        self.indptr_h, self.indices_h, self.slot_diag_h, self.he_slot_h = indptr, indices, slot_diag, he_slot
        self.slot_PN_h, self.slot_NP_h = slot_PN, slot_NP
        self.cptr_h, self.cnbr_h = cptr.astype(np.int32), cnbr.astype(np.int32)
        self.cface_h = cface.astype(np.int32)

        dev = self.device
        f64 = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float64), device=dev)  # noqa: E731
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.int32), device=dev)  # noqa: E731
        self.P, self.N, self.Pb = i32(P), i32(N), i32(Pb)
        self.w, self.alpha, self.wl = f64(w), f64(alpha), f64(wl)
        self.S, self.d, self.Tvec, self.dxm = f64(S), f64(d), f64(Tvec), f64(dxm)
        self.dP_f, self.dN_f = f64(xf - xc[P]), f64(xf - xN)
        self.Sb, self.db, self.db_t, self.nb = f64(Sb), f64(db), f64(db_t), f64(nbu)
        self.alpha_b, self.wlb, self.Sb_mag = f64(alpha_b), f64(wlb), f64(Sb_mag)
        self.vol, self.Ginv = f64(geo.cell_volume), f64(Ginv)
        self.xc, self.xfb, self.yb = f64(xc), f64(xfb), f64(self.yb_h)
        self.cptr, self.cface, self.cnbr = i32(cptr), i32(cface), i32(cnbr)
        self.bptr, self.bface = i32(bptr), i32(bface)
        self.indptr, self.indices, self.slot_diag = i32(indptr), i32(indices), i32(slot_diag)
        self.src_ptr, self.src = i32(src_ptr), i32(src)
        self.he_slot = i32(he_slot)
        self._slices(cptr, cnbr)
        # the reference's incidence adjacency for the drop-in kernels (iidx sorted = cface >> 1)
        self.iidx = i32(cface >> 1)
        self.isgn = f64(np.where(cface & 1, -1.0, 1.0))
        self.c_mesh = self._mesh_struct()
        self._incidence = None

    # ------------------------------------------------------------------ C struct
    def _mesh_struct(self) -> nat.FtmiMesh:
        m = nat.FtmiMesh()
        m.nc, m.nf, m.nb, m.dim = self.nc, self.n_if, self.n_bf, self.dim
        for name in ("P", "N", "w", "alpha", "wl", "S", "d", "Tvec", "dP_f", "dN_f", "dxm", "Pb", "Sb", "db",
                     "db_t", "alpha_b", "wlb", "Sb_mag", "vol", "Ginv", "cptr", "cface", "cnbr", "bptr", "bface"):
            setattr(m, name, nat.ptr(getattr(self, name)))
        m.nb_unit = nat.ptr(self.nb)
        return m

    @property
    def xf(self):
        return self.xc[self.P.long()] + self.dP_f

    # ------------------------------------------------------------------ reference operator surface
    def _inc(self):
        """Incidence matrices Cown, Cnei, Cb, Cdiff (cells x faces) as device CSR (fvmesh.py:138-145)."""
        if self._incidence is None:
            nc, nf, nb = self.nc, self.n_if, self.n_bf
            one = np.ones(nf)
            Cown = sp.csr_matrix((one, (self.P_h, np.arange(nf))), shape=(nc, nf))
            Cnei = sp.csr_matrix((one, (self.N_h, np.arange(nf))), shape=(nc, nf))
            Cb = sp.csr_matrix((np.ones(nb), (self.Pb_h, np.arange(nb))), shape=(nc, nb))
            Cdiff = (Cown - Cnei).tocsr()
            dev = self.device

            def up(M):
                M.sort_indices()
                return DeviceCSR(torch.as_tensor(M.data, dtype=torch.float64, device=dev),
                                 torch.as_tensor(M.indices.astype(np.int32), device=dev),
                                 torch.as_tensor(M.indptr.astype(np.int32), device=dev), M.shape)

            self._incidence = {k: up(v) for k, v in (("own", Cown), ("nei", Cnei), ("b", Cb), ("diff", Cdiff))}
        return self._incidence

    def scatter(self, f_int=None, f_nei=None, f_bnd=None):
        """Sum face values into cells: owner gets f_int, neighbour f_nei, boundary owner f_bnd."""
        C = self._inc()
        out = None
        for M, v in ((C["own"], f_int), (C["nei"], f_nei), (C["b"], f_bnd)):
            if v is None:
                continue
            t = M @ v
            out = t if out is None else out + t
        return out

    def net_outflow(self, F, Fb):
        C = self._inc()
        return C["diff"] @ F + C["b"] @ Fb

    def gradient(self, phi, phi_b):
        """Least-squares cell gradient (nc, dim) of a scalar with boundary face values."""
        phi, phi_b = phi.contiguous(), phi_b.contiguous()
        out = torch.empty((self.nc, self.dim), dtype=torch.float64, device=self.device)
        nat.check(nat.lib().ftmi_ls_gradient(
            self.nc, self.dim, nat.ptr(self.cptr), nat.ptr(self.iidx), nat.ptr(self.bptr), nat.ptr(self.bface),
            nat.ptr(self.P), nat.ptr(self.N), nat.ptr(self.wl), nat.ptr(self.d), nat.ptr(self.wlb), nat.ptr(self.db),
            nat.ptr(self.Ginv), nat.ptr(phi), nat.ptr(phi_b), nat.ptr(out), nat.stream_ptr()), "ls_gradient")
        return out

    def gradient_gg(self, phi, phi_b, grad=None):
        """Green-Gauss gradient with the skewness correction from ``grad`` when given."""
        phi, phi_b = phi.contiguous(), phi_b.contiguous()
        g = grad.contiguous() if grad is not None else None
        out = torch.empty((self.nc, self.dim), dtype=torch.float64, device=self.device)
        nat.check(nat.lib().ftmi_gg_gradient(
            self.nc, self.dim, nat.ptr(self.cptr), nat.ptr(self.iidx), nat.ptr(self.isgn), nat.ptr(self.bptr),
            nat.ptr(self.bface), nat.ptr(self.P), nat.ptr(self.N), nat.ptr(self.w), nat.ptr(self.S),
            nat.ptr(self.Sb), nat.ptr(self.vol), nat.ptr(self.dxm), nat.ptr(phi), nat.ptr(phi_b), nat.ptr(g),
            nat.ptr(out), nat.stream_ptr()), "gg_gradient")
        return out

    def _interp(self, a, k):
        a = a.contiguous()
        out = torch.empty((self.n_if,) if k == 1 else (self.n_if, k), dtype=torch.float64, device=self.device)
        nat.check(nat.lib().ftmi_face_interp(self.n_if, k, nat.ptr(self.w), nat.ptr(self.P), nat.ptr(self.N),
                                             nat.ptr(a), nat.ptr(out), nat.stream_ptr()), "face_interp")
        return out

    def face_interp(self, phi):
        return self._interp(phi, 1)

    def face_interp_vec(self, g):
        return self._interp(g, g.shape[1])

    def matrix(self, diag, a_PN, a_NP) -> DeviceCSR:
        """CSR matrix with the given diagonal and off-diagonal face coefficients."""
        data = torch.empty(self.nnz, dtype=torch.float64, device=self.device)
        nat.check(nat.lib().ftmi_assemble(self.nnz, self.nc, self.n_if, nat.ptr(self.indptr), nat.ptr(self.slot_diag),
                                          nat.ptr(self.src_ptr), nat.ptr(self.src), nat.ptr(diag.contiguous()),
                                          nat.ptr(a_PN.contiguous()), nat.ptr(a_NP.contiguous()), nat.ptr(data),
                                          nat.stream_ptr()), "assemble")
        return DeviceCSR(data, self.indices, self.indptr, (self.nc, self.nc))

    def matrix_like(self, data) -> DeviceCSR:
        return DeviceCSR(data, self.indices, self.indptr, (self.nc, self.nc))

    def _slices(self, cptr, cnbr):
        """The half-edges in 64-row slices (the layout of ``ftmi_bicgstab_ldu_sliced``): slice g
        holds rows 64g..64g+63, the k-th half-edge of row i at sl_base[g] + 64 k + i % 64 with
        slice width max(deg) over its rows; sl_pos maps each half-edge (cptr order) there."""
        nc = self.nc
        deg = np.diff(np.asarray(cptr, np.int64))
        if deg.size and deg.max() > 255:
            self.sl_base = None  # more than 255 faces per cell: the sliced path is not used
            return
        ng = (nc + 63) // 64
        d = np.zeros(ng * 64, np.int64)
        d[:nc] = deg
        width = d.reshape(ng, 64).max(axis=1)
        base = np.concatenate([[0], np.cumsum(64 * width)])
        rows = np.repeat(np.arange(nc), deg)
        k = np.arange(rows.size) - np.asarray(cptr, np.int64)[rows]
        pos = base[rows // 64] + 64 * k + rows % 64
        self.sl_size = int(base[-1])
        snbr = np.zeros(self.sl_size, np.int64)  # padding entries point at row 0 and are never read
        snbr[pos] = cnbr
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.int32), device=self.device)  # noqa: E731
        self.sl_base, self.sl_nbr, self.sl_pos = i32(base[:-1]), i32(snbr), i32(pos)
        self.sl_deg = torch.as_tensor(deg.astype(np.uint8), device=self.device)

    def slice_values(self, off, out=None):
        """Half-edge values (cptr order, FP64 or FP32) in the slice layout (padding entries are
        never read)."""
        if out is None:
            out = torch.zeros(self.sl_size, dtype=off.dtype, device=self.device)
        fn = nat.lib().ftmi_ldu_slice_f32 if off.dtype == torch.float32 else nat.lib().ftmi_ldu_slice
        nat.check(fn(off.numel(), nat.ptr(self.sl_pos), nat.ptr(off), nat.ptr(out), nat.stream_ptr()), "ldu_slice")
        return out

    # ------------------------------------------------------------------ host helpers
    def ldu_to_csr_host(self, diag: np.ndarray, off: np.ndarray) -> sp.csr_matrix:
        """The CSR matrix (fixed pattern) of an LDU-form operator, on the host."""
        data = np.zeros(self.nnz)
        data[self.slot_diag_h] = diag
        np.add.at(data, self.he_slot_h, off)
        return sp.csr_matrix((data, self.indices_h, self.indptr_h), shape=(self.nc, self.nc))

    def ldu_off_from_faces(self, aPN: np.ndarray, aNP: np.ndarray) -> np.ndarray:
        """Half-edge coefficients (the LDU ``off`` array, host) from per-face aPN (row P,
        column N) and aNP (row N, column P)."""
        f = self.cface_h >> 1
        return np.where(self.cface_h & 1, np.asarray(aNP)[f], np.asarray(aPN)[f])

    def memory_bytes(self) -> int:
        return sum(v.numel() * v.element_size() for v in vars(self).values() if isinstance(v, torch.Tensor))
