"""Face-based unstructured meshes, their finite-volume geometry, and the generators
of the SIMPLEC workloads (host-side set-up, run once per mesh).

The data model is the reference's (``flowtrace/mesh/unstructured.py:29-129``): a list
of faces, each with its nodes, an owner cell and a neighbour cell (-1 on boundary
faces), face zones and cell zones, and translational periodic pairs. Area vectors
point from the owner to the neighbour (out of the domain on boundary faces).

The generators reproduce the reference's node coordinates, face order and random
draws exactly (the seeded jitter is drawn X first, then Y, as
``flowtrace/mesh/unstructured.py:455-459``), because the golden vectors of the
reference solver are indexed by that numbering. SHA-256 hashes of the generator
outputs at full workload sizes are checked in the tests.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np


@dataclass
class Zone:
    id: int
    name: str
    type: str  # Fluent zone type: interior, wall, velocity-inlet, pressure-outlet, symmetry, periodic, fluid
    kind: str  # "face" or "cell"
    first: int = 0
    last: int = -1

    @property
    def count(self) -> int:
        return self.last - self.first + 1


@dataclass
class UnstructuredMesh:
    """Fluent-style face-based mesh (``flowtrace/mesh/unstructured.py:43-129``)."""

    dim: int
    nodes: np.ndarray
    face_ptr: np.ndarray
    face_nodes: np.ndarray
    owner: np.ndarray
    neighbour: np.ndarray
    face_zone: np.ndarray
    n_cells: int
    cell_zone: np.ndarray
    zones: Dict[int, Zone] = field(default_factory=dict)
    name: str = "mesh"
    # translational periodic pairs: faces_b[i] = faces_a[i] shifted by translation
    periodic: List[dict] = field(default_factory=list)

    @property
    def n_faces(self) -> int:
        return int(self.owner.shape[0])

    @property
    def n_nodes(self) -> int:
        return int(self.nodes.shape[0])

    def face_zones(self) -> List[Zone]:
        return [z for z in self.zones.values() if z.kind == "face"]

    def cell_zones(self) -> List[Zone]:
        return [z for z in self.zones.values() if z.kind == "cell"]

    def periodic_zone_ids(self) -> set:
        out = set()
        for p in self.periodic:
            out.update((p["zone_a"], p["zone_b"]))
        return out

    def boundary_zones(self) -> List[Zone]:
        """Face zones that take a boundary condition (``unstructured.py:74-77``)."""
        per = self.periodic_zone_ids()
        return [z for z in self.face_zones() if z.type != "interior" and z.id not in per]

    def zone_by_name(self, name: str) -> Optional[Zone]:
        return next((z for z in self.zones.values() if z.name == name), None)

    def faces_of_zone(self, zid: int) -> np.ndarray:
        return np.flatnonzero(self.face_zone == zid)

    def cell_polygons_2d(self):
        """Counter-clockwise node loop of every 2-D cell (``unstructured.py:103-129``).

        Each face runs n0 -> n1 counter-clockwise around its owner and n1 -> n0 around
        its neighbour. A cell's loop starts at the first node of its first edge in a
        stable sort of the edges by cell, then follows the edges."""
        if self.dim != 2:
            raise ValueError("cell loops exist only for 2-D meshes")
        fp = self.face_ptr
        n0, n1 = self.face_nodes[fp[:-1]], self.face_nodes[fp[:-1] + 1]
        inner = self.neighbour >= 0
        cell = np.concatenate([self.owner, self.neighbour[inner]])
        head = np.concatenate([n0, n1[inner]])
        tail = np.concatenate([n1, n0[inner]])
        order = np.argsort(cell, kind="stable")
        cell, head, tail = cell[order], head[order], tail[order]
        ptr = np.zeros(self.n_cells + 1, dtype=np.int64)
        np.cumsum(np.bincount(cell, minlength=self.n_cells), out=ptr[1:])
        loops = np.empty(head.size, dtype=np.int64)
        hl, tl = head.tolist(), tail.tolist()
        for c in range(self.n_cells):
            lo, hi = int(ptr[c]), int(ptr[c + 1])
            succ = dict(zip(hl[lo:hi], tl[lo:hi]))
            start = cur = hl[lo]
            for k in range(lo, hi):
                loops[k] = cur
                cur = succ.get(cur, start)
        return ptr, loops


@dataclass
class Geometry:
    face_centre: np.ndarray
    face_area_vec: np.ndarray
    cell_centre: np.ndarray
    cell_volume: np.ndarray
    flipped_faces: int = 0

    @property
    def face_area(self) -> np.ndarray:
        return np.sqrt(np.einsum("fd,fd->f", self.face_area_vec, self.face_area_vec))


def face_geometry(mesh: UnstructuredMesh):
    """Face centres and area vectors (``unstructured.py:171-212``).

    2-D: a segment n0 -> n1 has centre (n0 + n1)/2 and area vector (n1 - n0) x z.
    3-D: a k-node polygon is fanned into triangles about its node mean; the area
    vector is the sum of the triangle area vectors and the centre is the
    area-weighted mean of the triangle centroids (weights projected on the normal)."""
    X, fp, fn = mesh.nodes, mesh.face_ptr, mesh.face_nodes
    nf = mesh.n_faces
    if mesh.dim == 2:
        if np.any(np.diff(fp) != 2):
            raise ValueError("2-D faces must have exactly 2 nodes")
        a, b = X[fn[fp[:-1]]], X[fn[fp[:-1] + 1]]
        e = b - a
        return 0.5 * (a + b), np.stack([e[:, 1], -e[:, 0]], axis=1)
    fc = np.zeros((nf, 3))
    S = np.zeros((nf, 3))
    counts = np.diff(fp)
    for k in np.unique(counts):
        idx_all = np.flatnonzero(counts == k)
        for lo in range(0, idx_all.size, 1 << 19):
            idx = idx_all[lo:lo + (1 << 19)]
            pts = X[fn[fp[idx][:, None] + np.arange(k)[None, :]]]
            mid = pts.mean(axis=1)
            nxt = np.roll(pts, -1, axis=1)
            tri = 0.5 * np.cross(pts - mid[:, None, :], nxt - mid[:, None, :])
            Sk = tri.sum(axis=1)
            unit = Sk / np.maximum(np.linalg.norm(Sk, axis=1, keepdims=True), 1e-300)
            wt = np.einsum("mkd,md->mk", tri, unit)
            cen = (mid[:, None, :] + pts + nxt) / 3.0
            ws = wt.sum(axis=1, keepdims=True)
            fc[idx] = (wt[:, :, None] * cen).sum(axis=1) / np.where(np.abs(ws) > 0, ws, 1.0)
            S[idx] = Sk
    return fc, S


def compute_geometry(mesh: UnstructuredMesh, repair: bool = True) -> Geometry:
    """Cell volumes and centroids by pyramid decomposition (``unstructured.py:215-261``).

    The apex of every face pyramid is the mean of the cell's face centres; a face whose
    area vector points into its owner is flipped when ``repair`` is set."""
    dim, nc = mesh.dim, mesh.n_cells
    own, nei = mesh.owner, mesh.neighbour
    inner = nei >= 0
    fc, S = face_geometry(mesh)
    nin = nei[inner]
    cnt = np.bincount(own, minlength=nc) + np.bincount(nin, minlength=nc)
    apex = np.empty((nc, dim))
    for a in range(dim):
        apex[:, a] = np.bincount(own, fc[:, a], minlength=nc) + np.bincount(nin, fc[inner, a], minlength=nc)
    apex /= np.maximum(cnt, 1)[:, None]
    flipped = 0
    if repair:
        to = np.where(inner[:, None], apex[np.where(inner, nei, 0)] - apex[own], fc - apex[own])
        bad = np.einsum("fd,fd->f", S, to) < 0
        flipped = int(bad.sum())
        S[bad] *= -1.0
    frac = 2.0 / 3.0 if dim == 2 else 0.75
    vol = np.zeros(nc)
    mom = np.zeros((nc, dim))
    inner_faces = np.flatnonzero(inner)
    for cells, sgn, faces in ((own, 1.0, None), (nin, -1.0, inner_faces)):
        Sf = S if faces is None else S[faces]
        h = (fc if faces is None else fc[faces]) - apex[cells]
        v = np.einsum("fd,fd->f", Sf * sgn, h) / dim
        vc = v[:, None] * (apex[cells] + frac * h)
        vol += np.bincount(cells, v, minlength=nc)
        for a in range(dim):
            mom[:, a] += np.bincount(cells, vc[:, a], minlength=nc)
    if np.any(vol <= 0):
        raise ValueError(f"{int((vol <= 0).sum())} cells have non-positive volume")
    return Geometry(face_centre=fc, face_area_vec=S, cell_centre=mom / vol[:, None], cell_volume=vol,
                    flipped_faces=flipped)


# ----------------------------------------------------------------------------- zone bookkeeping
def _zone_rank_perm(mesh: UnstructuredMesh) -> np.ndarray:
    order = sorted(mesh.face_zones(), key=lambda z: (z.type != "interior", z.id))
    rank = np.zeros(max(mesh.zones) + 1, dtype=np.int64)
    for r, z in enumerate(order):
        rank[z.id] = r
    return np.argsort(rank[mesh.face_zone], kind="stable")


def sort_faces_by_zone(mesh: UnstructuredMesh) -> UnstructuredMesh:
    """Faces reordered so every face zone is contiguous, interior zones first
    (``unstructured.py:790-810``); periodic face lists are renumbered with them."""
    perm = _zone_rank_perm(mesh)
    counts = np.diff(mesh.face_ptr)[perm]
    starts = mesh.face_ptr[:-1][perm]
    ptr = np.zeros(perm.size + 1, dtype=np.int64)
    np.cumsum(counts, out=ptr[1:])
    idx = np.repeat(starts - ptr[:-1], counts) + np.arange(int(ptr[-1]))
    fz = mesh.face_zone[perm]
    zones = dict(mesh.zones)
    for z in mesh.face_zones():
        where = np.flatnonzero(fz == z.id)
        zones[z.id] = Zone(z.id, z.name, z.type, "face", int(where[0]), int(where[-1]))
    inv = np.empty(perm.size, dtype=np.int64)
    inv[perm] = np.arange(perm.size)
    periodic = [dict(p, faces_a=inv[p["faces_a"]], faces_b=inv[p["faces_b"]]) for p in mesh.periodic]
    return UnstructuredMesh(dim=mesh.dim, nodes=mesh.nodes, face_ptr=ptr, face_nodes=mesh.face_nodes[idx],
                            owner=mesh.owner[perm], neighbour=mesh.neighbour[perm], face_zone=fz,
                            n_cells=mesh.n_cells, cell_zone=mesh.cell_zone, zones=zones, name=mesh.name,
                            periodic=periodic)


def non_orthogonality_deg(mesh: UnstructuredMesh) -> float:
    """Largest angle between an interior face's area vector and the line joining its two cell
    centres (``fvmesh.py:120-121``), in degrees; periodic pairs are left out."""
    geo = compute_geometry(mesh)
    inner = mesh.neighbour >= 0
    if not inner.any():
        return 0.0
    S = geo.face_area_vec[inner]
    d = geo.cell_centre[mesh.neighbour[inner]] - geo.cell_centre[mesh.owner[inner]]
    c = np.einsum("fd,fd->f", S, d) / np.sqrt(np.einsum("fd,fd->f", S, S) * np.einsum("fd,fd->f", d, d))
    return float(np.degrees(np.arccos(np.clip(c, -1.0, 1.0))).max())


def _land_on(src: np.ndarray, dst: np.ndarray, reach: float):
    """For every point of ``src``, the nearest point of ``dst`` no farther than ``reach``, found by
    a sort-and-sweep along the coordinate in which ``dst`` spreads most: the candidates of a point
    are the ``dst`` points whose coordinate there lies within ``reach`` of its own. Returns
    (index, distance), with index -1 and distance inf where no candidate is in reach."""
    axis = int(np.argmax(np.ptp(dst, axis=0))) if dst.size else 0
    order = np.argsort(dst[:, axis], kind="stable")
    key = dst[order, axis]
    lo = np.searchsorted(key, src[:, axis] - reach, side="left")
    width = np.searchsorted(key, src[:, axis] + reach, side="right") - lo
    hit = np.full(src.shape[0], -1, dtype=np.int64)
    dist = np.full(src.shape[0], np.inf)
    for k in range(int(width.max(initial=0))):
        rows = np.flatnonzero(width > k)
        cand = order[lo[rows] + k]
        d = np.sqrt(((dst[cand] - src[rows]) ** 2).sum(axis=1))
        closer = d < dist[rows]
        hit[rows[closer]], dist[rows[closer]] = cand[closer], d[closer]
    return hit, dist


def make_periodic(mesh: UnstructuredMesh, zone_a: str, zone_b: str, translation=None,
                  tol: float = 1e-6) -> UnstructuredMesh:
    """Pair two boundary zones as translational periodic, ``zone_b = zone_a + T``
    (``unstructured.py:397-423``). Each face of ``zone_a``, shifted by T, must land within
    ``tol`` x its size (length in 2-D, square root of its area in 3-D) of a distinct face of
    ``zone_b``; T defaults to the difference of the zones' mean face centres. The zones become
    ``periodic`` / ``periodic-shadow`` and the pair is recorded in ``mesh.periodic`` with the
    partner faces in ``zone_a``'s order (replacing any earlier pair of these zones). Raises
    ValueError for unknown zones, unequal face counts or faces that do not match. Modifies and
    returns ``mesh``."""
    za, zb = mesh.zone_by_name(zone_a), mesh.zone_by_name(zone_b)
    if za is None or zb is None:
        raise ValueError(f"no zone named {zone_a!r} or {zone_b!r} in the mesh")
    fa, fb = mesh.faces_of_zone(za.id), mesh.faces_of_zone(zb.id)
    if fa.size != fb.size:
        raise ValueError(f"zones {zone_a!r} and {zone_b!r} cannot be periodic: {fa.size} vs {fb.size} faces")
    centre, area = face_geometry(mesh)
    ca, cb = centre[fa], centre[fb]
    shift = np.asarray(translation, dtype=float) if translation is not None else cb.mean(axis=0) - ca.mean(axis=0)
    size = np.sqrt((area[fa] ** 2).sum(axis=1)) ** (1.0 / max(mesh.dim - 1, 1))
    allowed = tol * np.maximum(size, 1e-300) + 1e-12
    hit, dist = _land_on(ca + shift, cb, float(allowed.max(initial=0.0)))
    if np.any(dist > allowed) or np.unique(hit).size != hit.size:
        raise ValueError(f"zone {zone_b!r} is not zone {zone_a!r} shifted by {shift}: faces do not pair up")
    za.type, zb.type = "periodic", "periodic-shadow"
    pair = {"zone_a": za.id, "zone_b": zb.id, "faces_a": fa, "faces_b": fb[hit], "translation": shift}
    mesh.periodic = [q for q in mesh.periodic if q["zone_a"] not in (za.id, zb.id)] + [pair]
    return mesh


# ----------------------------------------------------------------------------- generators
def cartesian_mesh_2d(nx: int, ny: int, lx: float = 1.0, ly: float = 1.0,
                      names=("left", "right", "bottom", "top"), types=("wall", "wall", "wall", "wall"),
                      triangulate: bool = False, jitter: float = 0.0, seed: int = 0, x0: float = 0.0,
                      y0: float = 0.0, warp: float = 0.0, xs=None, ys=None) -> UnstructuredMesh:
    """Structured quads (or triangles split along the (i,j)-(i+1,j+1) diagonal) of a
    rectangle as an unstructured mesh (``unstructured.py:430-556``).

    Nodes are numbered ``i (ny+1) + j``; quads ``i ny + j``; triangles ``2 (i ny + j) + t``
    (t = 0 below-right, 1 above-left). Interior faces come first (diagonals, then x =
    const, then y = const, each in (i, j) order), then the four sides. ``jitter`` moves
    interior nodes by U(-jitter, jitter) times the mean spacing; ``warp`` adds the
    smooth distortion sin(2 pi sx) sin(2 pi sy)."""
    if xs is not None:
        xs = np.asarray(xs, dtype=float)
        nx, lx = xs.size - 1, float(xs[-1] - xs[0])
    else:
        xs = x0 + np.linspace(0, lx, nx + 1)
    if ys is not None:
        ys = np.asarray(ys, dtype=float)
        ny, ly = ys.size - 1, float(ys[-1] - ys[0])
    else:
        ys = y0 + np.linspace(0, ly, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    if jitter:
        rng = np.random.default_rng(seed)
        X[1:-1, 1:-1] += rng.uniform(-jitter, jitter, (nx - 1, ny - 1)) * (lx / nx)
        Y[1:-1, 1:-1] += rng.uniform(-jitter, jitter, (nx - 1, ny - 1)) * (ly / ny)
    if warp:
        Lx, Ly = xs[-1] - xs[0], ys[-1] - ys[0]
        sx, sy = (X - xs[0]) / Lx, (Y - ys[0]) / Ly
        bump = warp * np.sin(2 * np.pi * sx) * np.sin(2 * np.pi * sy)
        X, Y = X + bump * Lx, Y + bump * Ly
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)

    def node(i, j):
        return i * (ny + 1) + j

    def grid(ilo, ihi, jlo, jhi):
        I, J = np.meshgrid(np.arange(ilo, ihi), np.arange(jlo, jhi), indexing="ij")
        return I.ravel(), J.ravel()

    segs, own, nei = [], [], []
    i_all, j_all = np.arange(nx), np.arange(ny)
    if triangulate:
        def cell(i, j, t):
            return 2 * (i * ny + j) + t

        n_cells = 2 * nx * ny
        I, J = grid(0, nx, 0, ny)
        segs.append((node(I, J), node(I + 1, J + 1)))
        own.append(cell(I, J, 1))
        nei.append(cell(I, J, 0))
        I, J = grid(1, nx, 0, ny)
        segs.append((node(I, J), node(I, J + 1)))
        own.append(cell(I - 1, J, 0))
        nei.append(cell(I, J, 1))
        I, J = grid(0, nx, 1, ny)
        segs.append((node(I + 1, J), node(I, J)))
        own.append(cell(I, J - 1, 1))
        nei.append(cell(I, J, 0))
        sides = [((node(0, j_all + 1), node(0, j_all)), cell(0, j_all, 1)),
                 ((node(nx, j_all), node(nx, j_all + 1)), cell(nx - 1, j_all, 0)),
                 ((node(i_all, 0), node(i_all + 1, 0)), cell(i_all, 0, 0)),
                 ((node(i_all + 1, ny), node(i_all, ny)), cell(i_all, ny - 1, 1))]
    else:
        def cell(i, j):
            return i * ny + j

        n_cells = nx * ny
        I, J = grid(1, nx, 0, ny)
        segs.append((node(I, J), node(I, J + 1)))
        own.append(cell(I - 1, J))
        nei.append(cell(I, J))
        I, J = grid(0, nx, 1, ny)
        segs.append((node(I + 1, J), node(I, J)))
        own.append(cell(I, J - 1))
        nei.append(cell(I, J))
        sides = [((node(0, j_all + 1), node(0, j_all)), cell(0, j_all)),
                 ((node(nx, j_all), node(nx, j_all + 1)), cell(nx - 1, j_all)),
                 ((node(i_all, 0), node(i_all + 1, 0)), cell(i_all, 0)),
                 ((node(i_all + 1, ny), node(i_all, ny)), cell(i_all, ny - 1))]
    n_int = sum(o.size for o in own)
    zone_id = [np.full(n_int, 3)]
    zones = {2: Zone(2, "fluid", "fluid", "cell", 0, n_cells - 1),
             3: Zone(3, "interior-fluid", "interior", "face", 0, n_int - 1)}
    start = n_int
    for k, (seg, o) in enumerate(sides):
        segs.append(seg)
        own.append(o)
        nei.append(np.full(o.size, -1))
        zone_id.append(np.full(o.size, 4 + k))
        zones[4 + k] = Zone(4 + k, names[k], types[k], "face", start, start + o.size - 1)
        start += o.size
    fnodes = np.concatenate([np.stack(s, axis=1) for s in segs]).astype(np.int64)
    return UnstructuredMesh(
        dim=2, nodes=nodes, face_ptr=np.arange(0, 2 * fnodes.shape[0] + 1, 2, dtype=np.int64),
        face_nodes=fnodes.ravel(), owner=np.concatenate(own).astype(np.int64),
        neighbour=np.concatenate(nei).astype(np.int64), face_zone=np.concatenate(zone_id).astype(np.int64),
        n_cells=n_cells, cell_zone=np.full(n_cells, 2, dtype=np.int64), zones=zones,
        name=f"rect{nx}x{ny}{'-tri' if triangulate else ''}")


def extrude_to_3d(mesh2d: UnstructuredMesh, nz: int, lz: float, front=("front", "symmetry"),
                  back=("back", "symmetry")) -> UnstructuredMesh:
    """Prisms/hexes by extruding a 2-D mesh over nz layers of [0, lz] (``unstructured.py:647-711``).

    Side faces of layer k are the 2-D faces lifted as (a,k)(b,k)(b,k+1)(a,k+1); then come
    the caps z_0 (zone ``back``, reversed loops), the interior caps, and z_nz (``front``);
    faces are finally sorted by zone."""
    if mesh2d.dim != 2:
        raise ValueError("extrusion needs a 2-D mesh")
    nn2, nc2 = mesh2d.n_nodes, mesh2d.n_cells
    zs = np.linspace(0.0, lz, nz + 1)
    nodes = np.concatenate([np.column_stack([mesh2d.nodes, np.full(nn2, z)]) for z in zs])
    a = mesh2d.face_nodes[mesh2d.face_ptr[:-1]]
    b = mesh2d.face_nodes[mesh2d.face_ptr[:-1] + 1]
    has_n = mesh2d.neighbour >= 0
    quads, own, nei, fz = [], [], [], []
    for k in range(nz):
        quads.append(np.stack([a + k * nn2, b + k * nn2, b + (k + 1) * nn2, a + (k + 1) * nn2], axis=1))
        own.append(mesh2d.owner + k * nc2)
        nei.append(np.where(has_n, np.maximum(mesh2d.neighbour, 0) + k * nc2, -1))
        fz.append(mesh2d.face_zone)
    quads = np.concatenate(quads)
    counts = [np.full(quads.shape[0], 4)]
    fnodes = [quads.ravel()]
    own, nei, fz = [np.concatenate(own)], [np.concatenate(nei)], [np.concatenate(fz)]
    lptr, loops = mesh2d.cell_polygons_2d()
    cnt2 = np.diff(lptr)
    # reversing every loop in place: position p of cell c maps to lptr[c] + lptr[c+1] - 1 - p
    rev = np.repeat(lptr[:-1] + lptr[1:] - 1, cnt2) - np.arange(loops.size)
    zmax = max(mesh2d.zones) + 1
    int_zone = next(z.id for z in mesh2d.face_zones() if z.type == "interior")
    cells2 = np.arange(nc2)
    for k in range(nz + 1):
        lifted = loops + k * nn2
        if k == 0:
            fnodes.append(lifted[rev])
            own.append(cells2)
            nei.append(np.full(nc2, -1))
            fz.append(np.full(nc2, zmax))
        else:
            fnodes.append(lifted)
            own.append(cells2 + (k - 1) * nc2)
            nei.append(cells2 + k * nc2 if k < nz else np.full(nc2, -1))
            fz.append(np.full(nc2, int_zone if k < nz else zmax + 1))
        counts.append(cnt2)
    zones = dict(mesh2d.zones)
    tiled = np.tile(mesh2d.cell_zone, nz)
    for z in mesh2d.cell_zones():
        idx = np.flatnonzero(tiled == z.id)
        zones[z.id] = Zone(z.id, z.name, z.type, "cell", int(idx[0]), int(idx[-1]))
    zones[zmax] = Zone(zmax, back[0], back[1], "face")
    zones[zmax + 1] = Zone(zmax + 1, front[0], front[1], "face")
    counts = np.concatenate(counts)
    ptr = np.zeros(counts.size + 1, dtype=np.int64)
    np.cumsum(counts, out=ptr[1:])
    m = UnstructuredMesh(dim=3, nodes=nodes, face_ptr=ptr, face_nodes=np.concatenate(fnodes).astype(np.int64),
                         owner=np.concatenate(own).astype(np.int64), neighbour=np.concatenate(nei).astype(np.int64),
                         face_zone=np.concatenate(fz).astype(np.int64), n_cells=nc2 * nz, cell_zone=tiled,
                         zones=zones, name=mesh2d.name + f"-x{nz}")
    return sort_faces_by_zone(m)


# ----------------------------------------------------------------------------- workload meshes
def duct_mesh(nx: int = 1000, ny: int = 500, jitter: float = 0.2, seed: int = 0) -> UnstructuredMesh:
    """4:1 channel [0,4]x[0,1] of 2 nx ny jittered triangles; zones inlet/outlet/bottom/top."""
    return cartesian_mesh_2d(nx, ny, 4.0, 1.0, names=("inlet", "outlet", "bottom", "top"),
                             types=("velocity-inlet", "pressure-outlet", "wall", "wall"), triangulate=True,
                             jitter=jitter, seed=seed)


def geometric(n: int, length: float, first: float, start: float = 0.0) -> np.ndarray:
    """Nodes of n cells on [start, start + length] whose sizes grow by a constant ratio from
    ``first`` (``flowtrace/mesh/grading.py:9-19``): the ratio q solves first (q^n - 1) / (q - 1)
    = length by Brent's method on (1 + 1e-12, 10); the partial sums are then scaled to the exact
    length. Uniform when ``first`` is too large for growth."""
    if n < 1:
        raise ValueError("need at least one cell")
    if first * n >= length:
        return start + np.linspace(0.0, length, n + 1)
    from scipy.optimize import brentq

    q = brentq(lambda r: first * (r ** n - 1) / (r - 1) - length, 1.0 + 1e-12, 10.0)
    x = np.concatenate([[0.0], np.cumsum(first * q ** np.arange(n))])
    x *= length / x[-1]
    return start + x


def tanh_nodes(n: int, stretch: float = 2.0) -> np.ndarray:
    """n+1 nodes on [0,1] clustered at both ends: 0.5 (1 + tanh(s (2 eta - 1)) / tanh s)."""
    eta = np.linspace(0.0, 1.0, n + 1)
    return 0.5 * (1.0 + np.tanh(stretch * (2.0 * eta - 1.0)) / np.tanh(stretch))


def channel_mesh(nx: int = 128, ny: int = 64, nz: int = 64, stretch: float = 2.0) -> UnstructuredMesh:
    """[0,4]x[0,1]x[0,1] hexes, tanh-stretched in y, periodic in x; walls bottom/top,
    symmetry front/back."""
    m2 = cartesian_mesh_2d(nx, ny, names=("left", "right", "bottom", "top"), xs=np.linspace(0.0, 4.0, nx + 1),
                           ys=tanh_nodes(ny, stretch))
    return make_periodic(extrude_to_3d(m2, nz, 1.0), "left", "right")


def extract_cells(mesh: UnstructuredMesh, keep: np.ndarray, name: str = "body",
                  type: str = "wall") -> UnstructuredMesh:
    """The sub-mesh of the cells where ``keep`` is True (``unstructured.py:714-780``, 2-D or
    3-D, no periodic pairs). Faces between a kept and a dropped cell become boundary faces
    of a new zone ``name``, oriented out of the kept cell; faces of dropped cells only go.
    Cells keep their order; faces are regrouped by zone (interior first)."""
    keep = np.asarray(keep, dtype=bool)
    if mesh.periodic:
        raise NotImplementedError("extracting cells from a mesh with periodic pairs is not supported")
    new_id = np.full(mesh.n_cells, -1, dtype=np.int64)
    new_id[keep] = np.arange(int(keep.sum()))
    own, nei = mesh.owner, mesh.neighbour
    ko = keep[own]
    kn = np.where(nei >= 0, keep[np.maximum(nei, 0)], False)
    use = ko | kn
    iface = (ko ^ kn) & (nei >= 0)  # exactly one side kept: becomes a boundary face
    flip = iface & ~ko  # the kept cell is the neighbour: reverse the face
    counts = np.diff(mesh.face_ptr)
    f_idx = np.flatnonzero(use)
    nodes_of = [mesh.face_nodes[mesh.face_ptr[f]:mesh.face_ptr[f + 1]] for f in f_idx]
    fn = [nd[::-1] if flip[f] else nd for f, nd in zip(f_idx, nodes_of)]
    new_own = np.where(flip[f_idx], new_id[np.maximum(nei[f_idx], 0)], new_id[own[f_idx]])
    new_nei = np.where(iface[f_idx], -1, np.where(nei[f_idx] >= 0, new_id[np.maximum(nei[f_idx], 0)], -1))
    zid = max(mesh.zones) + 1
    fz = np.where(iface[f_idx], zid, mesh.face_zone[f_idx])
    zones = {k: z for k, z in mesh.zones.items() if z.kind == "face" and np.any(fz == k)}
    zones[zid] = Zone(zid, name, type, "face")
    kept_cz = mesh.cell_zone[keep]
    for k, z in mesh.zones.items():
        if z.kind == "cell" and np.any(kept_cz == k):
            w = np.flatnonzero(kept_cz == k)
            zones[k] = Zone(k, z.name, z.type, "cell", int(w[0]), int(w[-1]))
    ptr = np.zeros(f_idx.size + 1, dtype=np.int64)
    np.cumsum(counts[f_idx], out=ptr[1:])
    sub = UnstructuredMesh(dim=mesh.dim, nodes=mesh.nodes, face_ptr=ptr,
                           face_nodes=np.concatenate(fn).astype(np.int64), owner=new_own.astype(np.int64),
                           neighbour=new_nei.astype(np.int64), face_zone=fz.astype(np.int64),
                           n_cells=int(keep.sum()), cell_zone=kept_cz.copy(), zones=zones,
                           name=f"{mesh.name}-{name}")
    return sort_faces_by_zone(sub)


def graded_nodes(a: float, b: float, h_near: float, n: int, toward: str) -> np.ndarray:
    """n+1 nodes on [a, b], geometrically graded to spacing h_near at the ``toward`` end
    ("a" or "b"), with the growth ratio chosen so the n cells span b - a exactly."""
    L = b - a
    if n * h_near >= L:
        return np.linspace(a, b, n + 1)
    lo, hi = 1.0, 2.0
    while h_near * (hi ** n - 1.0) / (hi - 1.0) < L:
        hi *= 2.0
    for _ in range(200):  # bisection on the ratio r: h (r^n - 1) / (r - 1) = L
        r = 0.5 * (lo + hi)
        if h_near * (r ** n - 1.0) / (r - 1.0) < L:
            lo = r
        else:
            hi = r
    r = 0.5 * (lo + hi)
    steps = h_near * r ** np.arange(n)
    steps *= L / steps.sum()
    x = a + np.concatenate([[0.0], np.cumsum(steps)])
    x[-1] = b
    return x if toward == "a" else (a + b) - x[::-1]


def square_cylinder_mesh(n_body: int = 40, upstream: float = 5.0, downstream: float = 15.0,
                         half_height: float = 5.0, growth_cells: int = None, jitter: float = 0.15,
                         seed: int = 0) -> UnstructuredMesh:
    """Square cylinder of side D = 1 centred at the origin in a (upstream + downstream) x
    2 half_height channel, meshed with triangles graded toward the body: n_body cells per
    side on the body, geometric growth away from it (``growth_cells`` cells per outer segment,
    default scaled with n_body), interior nodes off the body jittered by U(-jitter, jitter)
    times the LOCAL spacing (seeded). Zones: inlet (velocity-inlet, x = -upstream), outlet
    (pressure-outlet), bottom/top (symmetry), cylinder (wall). 20D x 10D with n_body = 200
    gives about 0.5M triangles."""
    h = 1.0 / n_body
    m = growth_cells if growth_cells is not None else max(8, int(round(n_body * 1.6)))
    xs = np.concatenate([graded_nodes(-upstream, -0.5, h, m, "b")[:-1], np.linspace(-0.5, 0.5, n_body + 1),
                         graded_nodes(0.5, downstream, h, 2 * m, "a")[1:]])
    ys = np.concatenate([graded_nodes(-half_height, -0.5, h, m, "b")[:-1], np.linspace(-0.5, 0.5, n_body + 1),
                         graded_nodes(0.5, half_height, h, m, "a")[1:]])
    nx, ny = xs.size - 1, ys.size - 1
    base = cartesian_mesh_2d(nx, ny, names=("inlet", "outlet", "bottom", "top"),
                             types=("velocity-inlet", "pressure-outlet", "symmetry", "symmetry"),
                             triangulate=True, xs=xs, ys=ys)
    nodes = base.nodes.copy()
    if jitter:
        rng = np.random.default_rng(seed)
        X, Y = nodes[:, 0].reshape(nx + 1, ny + 1), nodes[:, 1].reshape(nx + 1, ny + 1)
        hx = np.minimum(np.diff(xs)[:-1], np.diff(xs)[1:])  # local spacing at interior nodes
        hy = np.minimum(np.diff(ys)[:-1], np.diff(ys)[1:])
        jx = rng.uniform(-jitter, jitter, (nx - 1, ny - 1)) * hx[:, None]
        jy = rng.uniform(-jitter, jitter, (nx - 1, ny - 1)) * hy[None, :]
        xi, yi = X[1:-1, 1:-1], Y[1:-1, 1:-1]
        on_body = (np.abs(xi) <= 0.5 + 1e-12) & (np.abs(yi) <= 0.5 + 1e-12)
        X[1:-1, 1:-1] = np.where(on_body, xi, xi + jx)
        Y[1:-1, 1:-1] = np.where(on_body, yi, yi + jy)
        nodes = np.stack([X.ravel(), Y.ravel()], axis=1)
    base.nodes = nodes
    # a triangle is inside the body iff its three nodes are (body nodes are not jittered)
    ptr, loops = base.cell_polygons_2d()
    pts = nodes[loops].reshape(-1, 3, 2)
    inside = np.all((np.abs(pts[:, :, 0]) <= 0.5 + 1e-12) & (np.abs(pts[:, :, 1]) <= 0.5 + 1e-12), axis=1)
    mesh = extract_cells(base, ~inside, name="cylinder", type="wall")
    mesh.name = f"square-cylinder-{n_body}"
    return mesh
