"""Fluent mesh files (.msh / legacy .cas, ASCII or binary, optionally gzipped): reader and
writer for ``UnstructuredMesh`` (``flowtrace/mesh/fluent_msh.py:173-525``).

The file is a sequence of parenthesised sections ``(index header data)``. The reader scans the
section structure here and hands the bulk of every ASCII grid section to the native parsers
(``ftmi_parse_hex`` for the hexadecimal face and cell records, ``ftmi_parse_f64`` for node
coordinates); binary sections (indices 20xx single, 30xx double precision) are read in place
with NumPy. Sections read:

* 2 dimension; 10/2010/3010 nodes; 12/2012/3012 cells; 13/2013/3013 faces (2-, 3-, 4-node
  and mixed/polygonal records ``[k] n0 .. n(k-1) c0 c1``); 18/2018 periodic shadow face pairs;
  39/45 zone names and types.
* Everything else (comments, headers, case settings of a .cas) is skipped, ASCII by balanced
  parentheses (strings honoured), binary up to its ``End of Binary Section`` marker.

Header numbers of grid sections are hexadecimal, zone ids of 39/45 decimal. Faces keep the
file's order; a boundary face stored with ``c0 = 0`` has its nodes reversed and its cells
swapped so that the owner exists. Fluent's 3-D faces have their right-hand normal toward
``c0``, the opposite of this package's owner-outward convention, so 3-D faces are reversed on
reading and on writing; 2-D faces need nothing.
"""
from __future__ import annotations

import gzip
import io
import os
import re
import zlib
from typing import Dict, List, Tuple

import numpy as np

from .mesh import UnstructuredMesh, Zone, face_geometry, sort_faces_by_zone

# Fluent boundary-condition codes of face zones (``fluent_msh.py`` / the file-format spec)
BC_CODES = {
    2: "interior", 3: "wall", 4: "pressure-inlet", 5: "pressure-outlet", 7: "symmetry",
    8: "periodic-shadow", 9: "pressure-far-field", 10: "velocity-inlet", 12: "periodic",
    14: "fan", 20: "mass-flow-inlet", 24: "interface", 31: "parent", 36: "outflow", 37: "axis",
}
BC_NAMES = {v: k for k, v in BC_CODES.items()}

_BLANK = b" \t\r\n"
_END_BINARY = b"End of Binary Section"


class FluentMeshError(ValueError):
    """A file that is not a readable Fluent mesh (``fluent_msh.py:35-36``)."""


def reverse_face_nodes(face_ptr: np.ndarray, face_nodes: np.ndarray) -> np.ndarray:
    """Every face's node list reversed (``unstructured.py:156-176``)."""
    counts = np.diff(face_ptr)
    face = np.repeat(np.arange(counts.size), counts)
    pos = np.arange(face_nodes.size) - face_ptr[face]
    return face_nodes[face_ptr[face] + counts[face] - 1 - pos]


# ---------------------------------------------------------------- native bulk parsing
def _parse(data: bytes, kind: str, count_hint: int) -> np.ndarray:
    from . import _native as nat

    L = nat.lib()
    cap = max(count_hint, data.count(b"\n") + 1, 16)
    while True:
        out = np.empty(cap, dtype=np.int64 if kind == "hex" else np.float64)
        fn = L.ftmi_parse_hex if kind == "hex" else L.ftmi_parse_f64
        k = int(fn(data, len(data), out.ctypes.data, cap))
        if k >= 0:
            return out[:k]
        if k == -2 - cap:
            cap *= 2
            continue
        bad = -1 - k
        raise FluentMeshError(f"malformed {'hexadecimal' if kind == 'hex' else 'numeric'} token #{bad}: "
                              f"{data.split()[bad][:32]!r}")


# ---------------------------------------------------------------- section scanner
class _Cursor:
    """Byte cursor over the file with the few moves the section grammar needs."""

    def __init__(self, buf: bytes):
        self.b, self.n, self.p = buf, len(buf), 0

    def blanks(self):
        while self.p < self.n and self.b[self.p] in _BLANK:
            self.p += 1

    def at(self, ch: bytes) -> bool:
        self.blanks()
        return self.b[self.p:self.p + 1] == ch

    def take(self, ch: bytes):
        if not self.at(ch):
            raise FluentMeshError(f"expected {ch!r} at byte {self.p}, found {self.b[self.p:self.p + 32]!r}")
        self.p += 1

    def word(self) -> bytes:
        self.blanks()
        s = self.p
        while self.p < self.n and self.b[self.p] not in _BLANK and self.b[self.p] not in b"()":
            self.p += 1
        return self.b[s:self.p]

    def group(self) -> bytes:
        """Text of a '( ... )' group without nested parentheses; the cursor moves past it."""
        self.take(b"(")
        q = self.b.find(b")", self.p)
        if q < 0:
            raise FluentMeshError("unterminated section data")
        text = self.b[self.p:q]
        self.p = q + 1
        return text

    def skip_group(self):
        """Skip a balanced ASCII '( ... )' group; parentheses inside strings do not count."""
        self.take(b"(")
        depth, b, p = 1, self.b, self.p
        while depth and p < self.n:
            c = b[p]
            if c == 0x22:  # a string: jump to its unescaped closing quote
                q = p + 1
                while True:
                    q = b.find(b'"', q)
                    if q < 0:
                        raise FluentMeshError("unterminated string")
                    if b[q - 1] != 0x5C:
                        break
                    q += 1
                p = q + 1
                continue
            depth += c == 0x28
            depth -= c == 0x29
            p += 1
        self.p = p

    def end_binary(self, index: int):
        q = self.b.find(_END_BINARY, self.p)
        if q < 0:
            raise FluentMeshError(f"binary section {index} has no end marker")
        self.p = q + len(_END_BINARY)
        self.word()  # the section index
        self.take(b")")

    def binary(self, dtype, count: int) -> np.ndarray:
        nbytes = count * np.dtype(dtype).itemsize
        if self.p + nbytes > self.n:
            raise FluentMeshError("binary section runs past the end of the file")
        a = np.frombuffer(self.b, dtype=dtype, count=count, offset=self.p)
        self.p += nbytes
        return a


def _face_records(vals: np.ndarray, nf: int, ftype: int):
    """(counts, nodes, c0, c1) of nf face records; fixed size when ftype is 2, 3 or 4."""
    if ftype in (2, 3, 4):
        need = nf * (ftype + 2)
        if vals.size != need:
            raise FluentMeshError(f"face section: expected {need} integers, found {vals.size}")
        rec = vals.reshape(nf, ftype + 2)
        return np.full(nf, ftype, np.int64), rec[:, :ftype].ravel(), rec[:, ftype], rec[:, ftype + 1]
    counts = np.empty(nf, np.int64)
    starts = np.empty(nf, np.int64)
    p = 0
    for i in range(nf):  # mixed records: each starts with its node count
        if p >= vals.size:
            raise FluentMeshError("face section ends inside a record")
        counts[i] = vals[p]
        starts[i] = p + 1
        p += int(vals[p]) + 3
    if p != vals.size:
        raise FluentMeshError(f"face section: expected {p} integers, found {vals.size}")
    ends = starts + counts
    idx = np.repeat(starts, counts) + (np.arange(counts.sum()) - np.repeat(np.cumsum(counts) - counts, counts))
    return counts, vals[idx], vals[ends], vals[ends + 1]


def _binary_face_ints(cur: _Cursor, nf: int, ftype: int) -> np.ndarray:
    if ftype in (2, 3, 4):
        return cur.binary("<i4", nf * (ftype + 2)).astype(np.int64)
    start, p = cur.p, cur.p
    for _ in range(nf):
        k = int(np.frombuffer(cur.b, "<i4", 1, p)[0])
        p += 4 * (k + 3)
    cur.p = start
    return cur.binary("<i4", (p - start) // 4).astype(np.int64)


def _read_bytes(path) -> bytes:
    with open(path, "rb") as fh:
        buf = fh.read()
    return gzip.decompress(buf) if buf[:2] == b"\x1f\x8b" else buf


def read_fluent_mesh(path_or_bytes, name: str = None) -> UnstructuredMesh:
    """Read a Fluent .msh/.cas (ASCII or binary, optionally gzipped) into an UnstructuredMesh
    (``fluent_msh.py:173-186``): an unparsable file raises ``FluentMeshError`` naming it, a
    missing file ``FileNotFoundError``."""
    try:
        return _read(path_or_bytes, name)
    except FluentMeshError:
        raise
    except (ValueError, IndexError, KeyError, EOFError, gzip.BadGzipFile, zlib.error) as exc:
        where = "" if isinstance(path_or_bytes, (bytes, bytearray)) else f"{path_or_bytes}: "
        raise FluentMeshError(f"{where}not a readable Fluent mesh ({type(exc).__name__}: {exc})") from exc


def _read(src, name) -> UnstructuredMesh:
    if isinstance(src, (bytes, bytearray)):
        buf = bytes(src)
        buf = gzip.decompress(buf) if buf[:2] == b"\x1f\x8b" else buf
        name = name or "mesh"
    else:
        buf = _read_bytes(src)
        name = name or os.path.basename(str(src))
    cur = _Cursor(buf)
    dim = None
    nodes_decl = cells_decl = 0
    node_sets: List[Tuple[int, np.ndarray]] = []
    face_sets = []
    cell_sets = []
    zone_names: Dict[int, Tuple[str, str]] = {}
    shadows = []
    while True:
        cur.blanks()
        if cur.p >= cur.n:
            break
        if cur.b[cur.p] != 0x28:  # stray text between sections: skip the line
            q = cur.b.find(b"\n", cur.p)
            cur.p = cur.n if q < 0 else q + 1
            continue
        s0 = cur.p
        cur.p += 1
        tok = cur.word()
        if not tok.isdigit():
            cur.p = s0
            cur.skip_group()
            continue
        index = int(tok)
        if index == 2:
            dim = int(cur.word())
            cur.take(b")")
        elif index in (10, 2010, 3010):
            h = [int(t, 16) for t in cur.group().split()]
            zid, first, last = h[0], h[1], h[2]
            nd = h[4] if len(h) > 4 else (dim or 3)
            dim = dim or nd
            if zid == 0 or cur.at(b")"):
                nodes_decl = max(nodes_decl, last - first + 1) if zid == 0 else nodes_decl
                cur.take(b")")
                continue
            count = last - first + 1
            if index == 10:
                xyz = _parse(cur.group(), "f64", count * nd)
                cur.take(b")")
            else:
                cur.take(b"(")
                xyz = cur.binary("<f4" if index == 2010 else "<f8", count * nd).astype(np.float64)
                cur.end_binary(index)
            if xyz.size != count * nd:
                raise FluentMeshError(f"node zone {zid}: expected {count * nd} coordinates, found {xyz.size}")
            node_sets.append((first, xyz.reshape(count, nd)))
        elif index in (12, 2012, 3012):
            h = [int(t, 16) for t in cur.group().split()]
            zid, first, last, ztype = h[0], h[1], h[2], h[3] if len(h) > 3 else 1
            if zid == 0:
                cells_decl = last - first + 1 if last >= first else 0
                cur.take(b")")
                continue
            if cur.at(b"("):  # per-cell element types: not needed
                if index == 12:
                    cur.group()
                    cur.take(b")")
                else:
                    cur.take(b"(")
                    cur.binary("<i4", last - first + 1)
                    cur.end_binary(index)
            else:
                cur.take(b")")
            cell_sets.append((zid, first, last, ztype))
        elif index in (13, 2013, 3013):
            h = [int(t, 16) for t in cur.group().split()] + [0, 0]
            zid, first, last, bc, ftype = h[:5]
            if zid == 0 or cur.at(b")"):
                cur.take(b")")
                continue
            nf = last - first + 1
            if index == 13:
                vals = _parse(cur.group(), "hex", nf * ((ftype or 4) + 3))
                cur.take(b")")
            else:
                cur.take(b"(")
                vals = _binary_face_ints(cur, nf, ftype)
                cur.end_binary(index)
            face_sets.append((zid, first, bc) + _face_records(vals, nf, ftype))
        elif index in (18, 2018):
            h = [int(t, 16) for t in cur.group().split()]
            if cur.at(b"("):
                if index == 18:
                    vals = _parse(cur.group(), "hex", 2 * (h[1] - h[0] + 1))
                    cur.take(b")")
                else:
                    cur.take(b"(")
                    vals = cur.binary("<i4", 2 * (h[1] - h[0] + 1)).astype(np.int64)
                    cur.end_binary(index)
                shadows.append((h, vals.reshape(-1, 2) - 1))
            else:
                cur.take(b")")
        elif index in (39, 45):
            words = cur.group().split()
            zone_names[int(words[0])] = (words[1].decode(errors="replace"),
                                         words[2].decode(errors="replace") if len(words) > 2 else f"zone-{words[0]}")
            while cur.at(b"("):
                cur.skip_group()
            cur.take(b")")
        elif index >= 2000:  # another binary section
            cur.p = s0
            cur.end_binary(index)
        else:
            cur.p = s0
            cur.skip_group()
    if dim is None or not node_sets or not face_sets:
        raise FluentMeshError("file does not contain a complete Fluent mesh (nodes/faces missing)")
    return _assemble(dim, nodes_decl, cells_decl, node_sets, face_sets, cell_sets, zone_names, shadows, name)


def _assemble(dim, nodes_decl, cells_decl, node_sets, face_sets, cell_sets, zone_names, shadows, name):
    n_nodes = max([nodes_decl] + [first + len(x) - 1 for first, x in node_sets])
    nodes = np.zeros((n_nodes, dim))
    for first, xyz in node_sets:
        nodes[first - 1:first - 1 + len(xyz)] = xyz[:, :dim]
    face_sets.sort(key=lambda fs: fs[1])  # file order of the face indices
    counts = np.concatenate([fs[3] for fs in face_sets])
    fnodes = np.concatenate([fs[4] for fs in face_sets]) - 1
    c0 = np.concatenate([fs[5] for fs in face_sets]).copy()
    c1 = np.concatenate([fs[6] for fs in face_sets]).copy()
    fzone = np.concatenate([np.full(fs[3].size, fs[0], np.int64) for fs in face_sets])
    ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    flip = (c0 == 0) & (c1 != 0)  # the owner must exist: reverse the face, swap its cells
    if flip.any():
        face = np.repeat(np.arange(counts.size), counts)
        rev = reverse_face_nodes(ptr, fnodes)
        sel = flip[face]
        fnodes = np.where(sel, rev, fnodes)
        c0[flip], c1[flip] = c1[flip], 0
    if dim == 3:
        fnodes = reverse_face_nodes(ptr, fnodes)
    if (c0 <= 0).any():
        raise FluentMeshError("a face has no cell on either side")
    owner = c0 - 1
    neighbour = np.where(c1 > 0, c1 - 1, -1)
    n_cells = max(cells_decl, int(owner.max()) + 1, int(neighbour.max()) + 1)
    zones: Dict[int, Zone] = {}
    start = 0
    for zid, _first, bc, cnt, *_ in face_sets:
        ztype, zname = zone_names.get(zid, (BC_CODES.get(bc, "wall"), f"zone-{zid}"))
        zones[zid] = Zone(zid, zname, ztype, "face", start, start + cnt.size - 1)
        start += cnt.size
    cell_zone = np.zeros(n_cells, np.int64)
    for zid, first, last, ztype in cell_sets:
        cell_zone[first - 1:last] = zid
        zt, zn = zone_names.get(zid, ("fluid" if ztype == 1 else "solid", f"zone-{zid}"))
        zones[zid] = Zone(zid, zn, zt, "cell", first - 1, last - 1)
    if not cell_sets:
        zones[max(zones, default=0) + 1] = Zone(max(zones, default=0) + 1, "fluid", "fluid", "cell", 0, n_cells - 1)
        cell_zone[:] = max(zones)
    mesh = UnstructuredMesh(dim=dim, nodes=nodes, face_ptr=ptr, face_nodes=fnodes.astype(np.int64),
                            owner=owner.astype(np.int64), neighbour=neighbour.astype(np.int64),
                            face_zone=fzone, n_cells=n_cells, cell_zone=cell_zone, zones=zones, name=name)
    if shadows:
        fc, _ = face_geometry(mesh)
        for h, pairs in shadows:
            fa, fb = pairs[:, 0], pairs[:, 1]
            if fa.size == 0 or max(fa.max(), fb.max()) >= mesh.n_faces or min(fa.min(), fb.min()) < 0:
                raise FluentMeshError("periodic face pairs refer to faces outside the mesh")
            mesh.periodic.append({"zone_a": h[2], "zone_b": h[3], "faces_a": fa, "faces_b": fb,
                                  "translation": (fc[fb] - fc[fa]).mean(axis=0)})
    return mesh


def read_any_mesh(path: str):
    """(mesh, format) for a Fluent mesh file (``fluent_h5.py``'s dispatcher, .msh/.cas only)."""
    low = str(path).lower()
    base = low[:-3] if low.endswith(".gz") else low
    if base.endswith((".msh", ".cas")):
        return read_fluent_mesh(path), "fluent-msh"
    raise FluentMeshError(f"{path}: unsupported mesh format (Fluent .msh/.cas, optionally .gz)")


# ---------------------------------------------------------------- writer
def _zone_word(name: str) -> str:
    w = re.sub(r"[^A-Za-z0-9_\-.:]", "-", name).strip("-")
    return w or "zone"


def _element_types(mesh: UnstructuredMesh) -> np.ndarray:
    """Fluent element type per cell from its faces (7 = polygon/polyhedron)."""
    nc = mesh.n_cells
    inner = mesh.neighbour >= 0
    fcount = np.diff(mesh.face_ptr)

    def per_cell(w):
        return (np.bincount(mesh.owner, w, minlength=nc) + np.bincount(mesh.neighbour[inner], w[inner], minlength=nc))

    nfaces = per_cell(np.ones(mesh.n_faces)).astype(int)
    if mesh.dim == 2:
        return np.where(nfaces == 3, 1, np.where(nfaces == 4, 3, 7))
    tri, quad = per_cell((fcount == 3).astype(float)).astype(int), per_cell((fcount == 4).astype(float)).astype(int)
    t = np.full(nc, 7)
    t[(nfaces == 4) & (tri == 4)] = 2  # tetrahedron
    t[(nfaces == 6) & (quad == 6)] = 4  # hexahedron
    t[(nfaces == 5) & (tri == 4) & (quad == 1)] = 5  # pyramid
    t[(nfaces == 5) & (tri == 2) & (quad == 3)] = 6  # wedge
    return t


def write_fluent_mesh(mesh: UnstructuredMesh, path: str, binary: bool = False, title: str = "gputest_amd") -> None:
    """Write a Fluent .msh (ASCII, or binary double precision; gzipped when ``path`` ends in
    .gz) that ``read_fluent_mesh`` reads back to the same mesh (faces sorted by zone)
    (``fluent_msh.py:446-525``)."""
    mesh = sort_faces_by_zone(mesh)
    dim, nn, nc, nf = mesh.dim, mesh.n_nodes, mesh.n_cells, mesh.n_faces
    fn = reverse_face_nodes(mesh.face_ptr, mesh.face_nodes) if dim == 3 else mesh.face_nodes
    hx = lambda v: format(int(v), "x")  # noqa: E731
    out = io.BytesIO()
    w = out.write
    w(f'(0 "{title}: {nc} cells")\n(2 {dim})\n(10 (0 1 {hx(nn)} 0 {dim}))\n'.encode())
    if binary:
        w(f"(3010 (1 1 {hx(nn)} 1 {dim})(".encode())
        w(np.ascontiguousarray(mesh.nodes[:, :dim], "<f8").tobytes())
        w(b")\nEnd of Binary Section 3010)\n")
    else:
        w(f"(10 (1 1 {hx(nn)} 1 {dim})(\n".encode())
        np.savetxt(out, mesh.nodes[:, :dim], fmt="%.17g")
        w(b"))\n")
    w(f"(12 (0 1 {hx(nc)} 0))\n".encode())
    etypes = _element_types(mesh)
    for z in mesh.cell_zones():
        cells = np.flatnonzero(mesh.cell_zone == z.id)
        if cells.size == 0:
            continue
        if cells[-1] - cells[0] + 1 != cells.size:
            raise ValueError("cell zones must be contiguous to write a Fluent mesh")
        t = etypes[cells]
        et = int(t[0]) if np.all(t == t[0]) else 0
        w(f"(12 ({hx(z.id)} {hx(cells[0] + 1)} {hx(cells[-1] + 1)} 1 {hx(et)})".encode())
        w(b"(\n" + " ".join(hx(v) for v in t).encode() + b"\n))\n" if et == 0 else b")\n")
    w(f"(13 (0 1 {hx(nf)} 0))\n".encode())
    counts = np.diff(mesh.face_ptr)
    for z in sorted(mesh.face_zones(), key=lambda z: z.first):
        f = np.arange(z.first, z.last + 1)
        k = int(counts[f[0]]) if np.all(counts[f] == counts[f[0]]) and counts[f[0]] in (2, 3, 4) else 0
        if dim == 3 and k == 2:
            k = 0
        c0 = mesh.owner[f] + 1
        c1 = np.where(mesh.neighbour[f] >= 0, mesh.neighbour[f] + 1, 0)
        rows = [([int(counts[i])] if k == 0 else []) + (fn[mesh.face_ptr[i]:mesh.face_ptr[i + 1]] + 1).tolist()
                + [int(a), int(b)] for i, a, b in zip(f, c0, c1)]
        head = f"({hx(z.id)} {hx(f[0] + 1)} {hx(f[-1] + 1)} {hx(BC_NAMES.get(z.type, 3))} {hx(k)})"
        if binary:
            w(f"(2013 {head}(".encode())
            w(np.asarray([v for r in rows for v in r], "<i4").tobytes())
            w(b")\nEnd of Binary Section 2013)\n")
        else:
            w(f"(13 {head}(\n".encode())
            w(("\n".join(" ".join(hx(v) for v in r) for r in rows) + "\n))\n").encode())
    for p in mesh.periodic:
        fa, fb = np.asarray(p["faces_a"]) + 1, np.asarray(p["faces_b"]) + 1
        w(f"(18 (1 {hx(fa.size)} {hx(p['zone_a'])} {hx(p['zone_b'])})(\n".encode())
        w(("\n".join(f"{hx(a)} {hx(b)}" for a, b in zip(fa, fb)) + "\n))\n").encode())
    for z in sorted(mesh.zones.values(), key=lambda z: z.id):
        w(f"(45 ({z.id} {z.type} {_zone_word(z.name)})())\n".encode())
    data = out.getvalue()
    if str(path).endswith(".gz"):
        data = gzip.compress(data)
    with open(path, "wb") as fh:
        fh.write(data)


__all__ = ["FluentMeshError", "read_fluent_mesh", "read_any_mesh", "write_fluent_mesh", "reverse_face_nodes",
           "BC_CODES", "BC_NAMES"]
