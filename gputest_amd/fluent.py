"""Fluent mesh files (.msh / legacy .cas, ASCII or binary, optionally gzipped): reader and
writer for ``UnstructuredMesh`` (``flowtrace/mesh/fluent_msh.py:173-525``).

Reading is three stages:

1. **Index.** ``ftmi_fluent_index`` (host library) scans the whole file once and returns a
   table with one row per top-level section: its index and the byte spans of its header and
   data group. Comments, case settings and unknown sections are skipped there (ASCII ones as
   balanced groups whose strings may hold parentheses, binary ones up to their ``End of Binary
   Section`` marker).
2. **Decode.** Each row goes to the decoder registered for its index in ``_DECODERS``, which
   turns the spans into arrays on a ``_Parts`` accumulator: ASCII numbers through the native
   parsers (``ftmi_parse_hex`` for the hexadecimal face and cell records, ``ftmi_parse_f64`` for
   coordinates), binary payloads (indices 20xx single, 30xx double precision) straight from the
   buffer with NumPy.
3. **Build.** ``_Parts.build`` orders the face zones by their first face index, orients the faces
   and makes the zones.

Sections read: 2 dimension; 10/2010/3010 nodes; 12/2012/3012 cells; 13/2013/3013 faces (2-, 3-,
4-node and mixed/polygonal records ``[k] n0 .. n(k-1) c0 c1``); 18/2018 periodic shadow face
pairs; 39/45 zone names and types. Header numbers of grid sections are hexadecimal, zone ids of
39/45 decimal (the file-format specification).

Orientation. A boundary face stored with ``c0 = 0`` gets its node order reversed and its cells
swapped, so every face has an owner. Fluent's 3-D faces have their right-hand normal toward
``c0``, the opposite of this package's owner-outward convention, so 3-D node orders are reversed
on reading and on writing; 2-D faces need nothing.
"""
from __future__ import annotations

import gzip
import io
import os
import re
import zlib
from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np

from .mesh import UnstructuredMesh, Zone, face_geometry, sort_faces_by_zone

# Fluent boundary-condition codes of face zones (the file-format specification)
BC_CODES = {
    2: "interior", 3: "wall", 4: "pressure-inlet", 5: "pressure-outlet", 7: "symmetry",
    8: "periodic-shadow", 9: "pressure-far-field", 10: "velocity-inlet", 12: "periodic",
    14: "fan", 20: "mass-flow-inlet", 24: "interface", 31: "parent", 36: "outflow", 37: "axis",
}
BC_NAMES = {v: k for k, v in BC_CODES.items()}

_INDEX_ERRORS = {1: "a parenthesis is missing", 2: "a grid data group holds a nested group",
                 3: "a group is not closed", 4: "a string is not closed",
                 5: "a binary section has no 'End of Binary Section' marker"}


class FluentMeshError(ValueError):
    """A file that is not a readable Fluent mesh (``fluent_msh.py:35-36``)."""


def reverse_face_nodes(face_ptr: np.ndarray, face_nodes: np.ndarray) -> np.ndarray:
    """Every face's node list reversed (``unstructured.py:156-176``)."""
    counts = np.diff(face_ptr)
    face = np.repeat(np.arange(counts.size), counts)
    pos = np.arange(face_nodes.size) - face_ptr[face]
    return face_nodes[face_ptr[face] + counts[face] - 1 - pos]


# ---------------------------------------------------------------- native helpers
def _lib():
    from . import _native as nat

    return nat.lib()


def _numbers(text: bytes, kind: str, expect: int) -> np.ndarray:
    """All blank-separated numbers of ``text``: hexadecimal integers or decimal floats."""
    L = _lib()
    fn, dt = (L.ftmi_parse_hex, np.int64) if kind == "hex" else (L.ftmi_parse_f64, np.float64)
    cap = max(expect, 16)
    while True:
        out = np.empty(cap, dtype=dt)
        k = int(fn(text, len(text), out.ctypes.data, cap))
        if k >= 0:
            return out[:k]
        if k == -2 - cap:  # more numbers than expected: make room and parse again
            cap = max(2 * cap, text.count(b" ") + text.count(b"\n") + 1)
            continue
        bad = -1 - k
        raise FluentMeshError(f"number #{bad} of a data group cannot be read: {text.split()[bad][:32]!r}")


def _index(buf: bytes) -> np.ndarray:
    L = _lib()
    cap = max(64, buf.count(b"\n(") + 8)
    err = np.zeros(1, np.int64)
    while True:
        table = np.empty((cap, 6), np.int64)
        k = int(L.ftmi_fluent_index(buf, len(buf), table.ctypes.data, cap, err.ctypes.data))
        if k >= 0:
            return table[:k]
        if k <= -100:
            cap = -100 - k
            continue
        raise FluentMeshError(f"section structure broken at byte {int(err[0])}: {_INDEX_ERRORS.get(-k, k)}")


def _record_starts(vals: np.ndarray, n: int) -> Tuple[np.ndarray, int]:
    vals = np.ascontiguousarray(vals, np.int64)
    starts = np.empty(n, np.int64)
    used = int(_lib().ftmi_fluent_walk(vals.ctypes.data, vals.size, n, starts.ctypes.data))
    if used < 0:
        raise FluentMeshError("a mixed face group ends inside a record")
    return starts, used


# ---------------------------------------------------------------- decoding
@dataclass
class _Face:
    zone: int
    first: int
    bc: int
    counts: np.ndarray
    nodes: np.ndarray  # 1-based
    c0: np.ndarray
    c1: np.ndarray


@dataclass
class _Parts:
    """What the decoders collect; ``build`` makes the mesh."""
    buf: bytes
    dim: int = 0
    n_nodes: int = 0  # declared totals (zone 0 headers)
    n_cells: int = 0
    nodes: List[Tuple[int, np.ndarray]] = field(default_factory=list)  # (first, xyz)
    faces: List[_Face] = field(default_factory=list)
    cells: List[Tuple[int, int, int, int]] = field(default_factory=list)  # (zone, first, last, type)
    names: Dict[int, Tuple[str, str]] = field(default_factory=dict)  # zone id -> (type, name)
    shadows: List[Tuple[int, int, np.ndarray]] = field(default_factory=list)  # (zone a, zone b, pairs)

    def text(self, a, b) -> bytes:
        return self.buf[a:b]

    def binary(self, begin: int, end: int, dtype, count: int) -> np.ndarray:
        """``count`` values of ``dtype`` from ``begin``; the payload must end before ``end``."""
        if begin + count * np.dtype(dtype).itemsize > end:
            raise FluentMeshError("a binary payload is shorter than its header says")
        return np.frombuffer(self.buf, dtype=dtype, count=count, offset=begin)


def _header(parts: _Parts, row) -> List[int]:
    return [int(t, 16) for t in parts.text(row[1], row[2]).split()]


def _dimension(parts: _Parts, idx, row):
    parts.dim = int(parts.text(row[1], row[2]))


def _nodes(parts: _Parts, idx, row):
    h = _header(parts, row)
    zid, first, last = h[0], h[1], h[2]
    nd = h[4] if len(h) > 4 else (parts.dim or 3)
    parts.dim = parts.dim or nd
    if zid == 0:
        parts.n_nodes = max(parts.n_nodes, last - first + 1)
        return
    if not row[5] & 1:
        return
    count = (last - first + 1) * nd
    if idx == 10:
        xyz = _numbers(parts.text(row[3], row[4]), "f64", count)
    else:
        xyz = parts.binary(row[3], row[4], "<f4" if idx == 2010 else "<f8", count).astype(np.float64)
    if xyz.size != count:
        raise FluentMeshError(f"node zone {zid} holds {xyz.size} coordinates, its header says {count}")
    parts.nodes.append((first, xyz.reshape(-1, nd)))


def _cells(parts: _Parts, idx, row):
    h = _header(parts, row)
    zid, first, last = h[0], h[1], h[2]
    if zid == 0:
        parts.n_cells = last - first + 1 if last >= first else 0
        return
    parts.cells.append((zid, first, last, h[3] if len(h) > 3 else 1))  # element types are not needed


def _split_records(vals: np.ndarray, nf: int, ftype: int):
    """(counts, nodes, c0, c1) of ``nf`` records; a fixed record length when ftype is 2, 3 or 4."""
    if ftype in (2, 3, 4):
        if vals.size != nf * (ftype + 2):
            raise FluentMeshError(f"a face group holds {vals.size} integers, {nf} records need {nf * (ftype + 2)}")
        rec = vals.reshape(nf, ftype + 2)
        return np.full(nf, ftype, np.int64), rec[:, :ftype].reshape(-1), rec[:, ftype], rec[:, ftype + 1]
    starts, used = _record_starts(vals, nf)
    if used != vals.size:
        raise FluentMeshError(f"a mixed face group holds {vals.size} integers, its records use {used}")
    counts = vals[starts]
    first_node = starts + 1
    total = int(counts.sum())
    offs = np.arange(total) - np.repeat(np.cumsum(counts) - counts, counts)
    nodes = vals[np.repeat(first_node, counts) + offs]
    after = first_node + counts
    return counts, nodes, vals[after], vals[after + 1]


def _faces(parts: _Parts, idx, row):
    h = _header(parts, row) + [0, 0]
    zid, first, last, bc, ftype = h[:5]
    if zid == 0 or not row[5] & 1:
        return
    nf = last - first + 1
    if idx == 13:
        vals = _numbers(parts.text(row[3], row[4]), "hex", nf * ((ftype or 4) + 3))
    elif ftype in (2, 3, 4):
        vals = parts.binary(row[3], row[4], "<i4", nf * (ftype + 2)).astype(np.int64)
    else:  # mixed binary records: walk them over everything up to the end marker
        room = (row[4] - row[3]) // 4
        everything = parts.binary(row[3], row[4], "<i4", room).astype(np.int64)
        _, used = _record_starts(everything, nf)
        vals = everything[:used]
    parts.faces.append(_Face(zid, first, bc, *_split_records(vals, nf, ftype)))


def _shadow_pairs(parts: _Parts, idx, row):
    if not row[5] & 1:
        return
    h = _header(parts, row)
    count = 2 * (h[1] - h[0] + 1)
    if idx == 18:
        vals = _numbers(parts.text(row[3], row[4]), "hex", count)
    else:
        vals = parts.binary(row[3], row[4], "<i4", count).astype(np.int64)
    parts.shadows.append((h[2], h[3], vals.reshape(-1, 2) - 1))


def _zone_name(parts: _Parts, idx, row):
    words = parts.text(row[1], row[2]).split()
    zid = int(words[0])
    ztype = words[1].decode(errors="replace")
    zname = words[2].decode(errors="replace") if len(words) > 2 else f"zone-{zid}"
    parts.names[zid] = (ztype, zname)


_DECODERS = {2: _dimension, 39: _zone_name, 45: _zone_name}
for _i in (10, 2010, 3010):
    _DECODERS[_i] = _nodes
for _i in (12, 2012, 3012):
    _DECODERS[_i] = _cells
for _i in (13, 2013, 3013):
    _DECODERS[_i] = _faces
for _i in (18, 2018):
    _DECODERS[_i] = _shadow_pairs


def _build(parts: _Parts, name: str) -> UnstructuredMesh:
    if not parts.dim or not parts.nodes or not parts.faces:
        raise FluentMeshError("the file holds no complete mesh (dimension, nodes or faces are missing)")
    dim = parts.dim
    n_nodes = max([parts.n_nodes] + [first + len(x) - 1 for first, x in parts.nodes])
    nodes = np.zeros((n_nodes, dim))
    for first, xyz in parts.nodes:
        nodes[first - 1:first - 1 + len(xyz)] = xyz[:, :dim]
    zones_in_order = sorted(parts.faces, key=lambda f: f.first)
    counts = np.concatenate([f.counts for f in zones_in_order])
    fnodes = np.concatenate([f.nodes for f in zones_in_order]) - 1
    c0 = np.concatenate([f.c0 for f in zones_in_order])
    c1 = np.concatenate([f.c1 for f in zones_in_order])
    fzone = np.repeat(np.array([f.zone for f in zones_in_order], np.int64), [f.counts.size for f in zones_in_order])
    ptr = np.zeros(counts.size + 1, np.int64)
    np.cumsum(counts, out=ptr[1:])
    orphan = (c0 == 0) & (c1 != 0)  # owner missing: the other cell owns the face, reversed
    reverse = orphan if dim == 2 else ~orphan  # 3-D: Fluent's normal points toward c0
    if reverse.any():
        face_of = np.repeat(np.arange(counts.size), counts)
        fnodes = np.where(reverse[face_of], reverse_face_nodes(ptr, fnodes), fnodes)
    owner_1 = np.where(orphan, c1, c0)
    neigh_1 = np.where(orphan, 0, c1)
    if (owner_1 <= 0).any():
        raise FluentMeshError("a face has no cell on either side")
    owner, neighbour = owner_1 - 1, np.where(neigh_1 > 0, neigh_1 - 1, -1)
    n_cells = max(parts.n_cells, int(owner.max()) + 1, int(neighbour.max()) + 1)
    zones: Dict[int, Zone] = {}
    lo = 0
    for f in zones_in_order:
        ztype, zname = parts.names.get(f.zone, (BC_CODES.get(f.bc, "wall"), f"zone-{f.zone}"))
        zones[f.zone] = Zone(f.zone, zname, ztype, "face", lo, lo + f.counts.size - 1)
        lo += f.counts.size
    cell_zone = np.zeros(n_cells, np.int64)
    for zid, first, last, ztype in parts.cells:
        cell_zone[first - 1:last] = zid
        zt, zn = parts.names.get(zid, ("fluid" if ztype == 1 else "solid", f"zone-{zid}"))
        zones[zid] = Zone(zid, zn, zt, "cell", first - 1, last - 1)
    if not parts.cells:  # no cell zone in the file: one fluid zone of every cell
        zid = max(zones, default=0) + 1
        zones[zid] = Zone(zid, "fluid", "fluid", "cell", 0, n_cells - 1)
        cell_zone[:] = zid
    mesh = UnstructuredMesh(dim=dim, nodes=nodes, face_ptr=ptr, face_nodes=fnodes.astype(np.int64),
                            owner=owner.astype(np.int64), neighbour=neighbour.astype(np.int64), face_zone=fzone,
                            n_cells=n_cells, cell_zone=cell_zone, zones=zones, name=name)
    if parts.shadows:
        centre, _ = face_geometry(mesh)
        for za, zb, pairs in parts.shadows:
            if pairs.size == 0 or pairs.min() < 0 or pairs.max() >= mesh.n_faces:
                raise FluentMeshError("periodic shadow pairs name faces the mesh does not have")
            fa, fb = pairs[:, 0], pairs[:, 1]
            mesh.periodic.append({"zone_a": za, "zone_b": zb, "faces_a": fa, "faces_b": fb,
                                  "translation": (centre[fb] - centre[fa]).mean(axis=0)})
    return mesh


def _raw(src) -> Tuple[bytes, str]:
    if isinstance(src, (bytes, bytearray)):
        buf, name = bytes(src), "mesh"
    else:
        with open(src, "rb") as fh:
            buf = fh.read()
        name = os.path.basename(str(src))
    return (gzip.decompress(buf) if buf[:2] == b"\x1f\x8b" else buf), name


def read_fluent_mesh(path_or_bytes, name: str = None) -> UnstructuredMesh:
    """Read a Fluent .msh/.cas (ASCII or binary, optionally gzipped) from a path or from the
    file's bytes (``fluent_msh.py:173-186``). A file that cannot be read as a mesh raises
    ``FluentMeshError`` (a ``ValueError``) naming it and the cause; a missing file raises
    ``FileNotFoundError``."""
    where = "" if isinstance(path_or_bytes, (bytes, bytearray)) else f"{path_or_bytes}: "
    try:
        buf, default_name = _raw(path_or_bytes)
        parts = _Parts(buf)
        for row in _index(buf):
            decode = _DECODERS.get(int(row[0]))
            if decode is not None:
                decode(parts, int(row[0]), row)
        return _build(parts, name or default_name)
    except FluentMeshError as exc:
        raise FluentMeshError(f"{where}{exc}") from None
    except (ValueError, IndexError, KeyError, EOFError, gzip.BadGzipFile, zlib.error) as exc:
        raise FluentMeshError(f"{where}cannot be read as a Fluent mesh: {type(exc).__name__}: {exc}") from exc


def read_any_mesh(path: str):
    """(mesh, format) for a Fluent mesh file (``fluent_h5.py``'s dispatcher, .msh/.cas only)."""
    low = str(path).lower()
    base = low[:-3] if low.endswith(".gz") else low
    if base.endswith((".msh", ".cas")):
        return read_fluent_mesh(path), "fluent-msh"
    raise FluentMeshError(f"{path}: unsupported mesh format (Fluent .msh/.cas, optionally .gz)")


# ---------------------------------------------------------------- writer
_SIGNATURES_3D = {(4, 4, 0): 2, (6, 0, 6): 4, (5, 4, 1): 5, (5, 2, 3): 6}  # (faces, tri, quad) -> type


def _element_types(mesh: UnstructuredMesh) -> np.ndarray:
    """Fluent element type per cell from the shapes of its faces: 1 triangle, 3 quadrilateral,
    2 tetrahedron, 4 hexahedron, 5 pyramid, 6 wedge, 7 polygon/polyhedron."""
    nc = mesh.n_cells
    both = np.concatenate([mesh.owner, mesh.neighbour[mesh.neighbour >= 0]])
    shape = np.diff(mesh.face_ptr)
    shape = np.concatenate([shape, shape[mesh.neighbour >= 0]])
    nfaces = np.bincount(both, minlength=nc)
    if mesh.dim == 2:
        return np.select([nfaces == 3, nfaces == 4], [1, 3], 7)
    tri = np.bincount(both, shape == 3, minlength=nc).astype(np.int64)
    quad = np.bincount(both, shape == 4, minlength=nc).astype(np.int64)
    out = np.full(nc, 7)
    for (f, t, q), code in _SIGNATURES_3D.items():
        out[(nfaces == f) & (tri == t) & (quad == q)] = code
    return out


def _hex_lines(rows: List[np.ndarray]) -> bytes:
    """Lines of blank-separated hexadecimal integers, one per array."""
    return b"".join((" ".join(format(int(v), "x") for v in r) + "\n").encode() for r in rows)


def _hex_table(a: np.ndarray) -> bytes:
    out = io.StringIO()
    np.savetxt(out, a, fmt="%x")
    return out.getvalue().encode()


def _section(index: int, header: str, body: bytes, binary: bool) -> bytes:
    if binary:
        return f"({index} ({header})(".encode() + body + f")\nEnd of Binary Section {index})\n".encode()
    return f"({index} ({header})(\n".encode() + body + b"))\n"


def _node_section(mesh, binary: bool) -> bytes:
    nn, dim = mesh.n_nodes, mesh.dim
    head = f"1 1 {nn:x} 1 {dim}"
    if binary:
        return _section(3010, head, np.ascontiguousarray(mesh.nodes[:, :dim], "<f8").tobytes(), True)
    txt = io.StringIO()
    np.savetxt(txt, mesh.nodes[:, :dim], fmt="%.17g")
    return _section(10, head, txt.getvalue().encode(), False)


def _cell_sections(mesh) -> bytes:
    out = [f"(12 (0 1 {mesh.n_cells:x} 0))\n".encode()]
    kinds = _element_types(mesh)
    for z in mesh.cell_zones():
        cells = np.flatnonzero(mesh.cell_zone == z.id)
        if cells.size == 0:
            continue
        if cells[-1] - cells[0] + 1 != cells.size:
            raise ValueError(f"cell zone {z.name!r} is not a contiguous cell range; Fluent zones must be")
        k = kinds[cells]
        uniform = bool(np.all(k == k[0]))
        head = f"{z.id:x} {cells[0] + 1:x} {cells[-1] + 1:x} 1 {int(k[0]) if uniform else 0:x}"
        if uniform:
            out.append(f"(12 ({head}))\n".encode())
        else:  # a mixed zone lists every cell's type
            out.append(f"(12 ({head})(\n".encode() + _hex_lines([k]) + b"))\n")
    return b"".join(out)


def _face_sections(mesh, nodes_out: np.ndarray, binary: bool) -> bytes:
    out = [f"(13 (0 1 {mesh.n_faces:x} 0))\n".encode()]
    shapes = np.diff(mesh.face_ptr)
    fixed_ok = (2,) if mesh.dim == 2 else (3, 4)
    for z in sorted(mesh.face_zones(), key=lambda z: z.first):
        f = np.arange(z.first, z.last + 1)
        k = int(shapes[f[0]])
        fixed = k in fixed_ok and bool(np.all(shapes[f] == k))
        c0 = mesh.owner[f] + 1
        c1 = np.where(mesh.neighbour[f] >= 0, mesh.neighbour[f] + 1, 0)
        head = f"{z.id:x} {f[0] + 1:x} {f[-1] + 1:x} {BC_NAMES.get(z.type, 3):x} {k if fixed else 0:x}"
        if fixed:
            rec = np.column_stack([nodes_out[mesh.face_ptr[f][:, None] + np.arange(k)] + 1, c0, c1])
        else:
            rows = [np.concatenate([[shapes[i]], nodes_out[mesh.face_ptr[i]:mesh.face_ptr[i + 1]] + 1, [a, b]])
                    for i, a, b in zip(f, c0, c1)]
        if binary:
            flat = rec.reshape(-1) if fixed else np.concatenate(rows)
            out.append(_section(2013, head, np.asarray(flat, "<i4").tobytes(), True))
        else:
            out.append(_section(13, head, _hex_table(rec) if fixed else _hex_lines(rows), False))
    return b"".join(out)


def _periodic_sections(mesh) -> bytes:
    out = []
    for p in mesh.periodic:
        pairs = np.column_stack([np.asarray(p["faces_a"]) + 1, np.asarray(p["faces_b"]) + 1])
        out.append(_section(18, f"1 {len(pairs):x} {p['zone_a']:x} {p['zone_b']:x}", _hex_table(pairs), False))
    return b"".join(out)


def _zone_sections(mesh) -> bytes:
    word = lambda s: re.sub(r"[^A-Za-z0-9_\-.:]", "-", s).strip("-") or "zone"  # noqa: E731
    return b"".join(f"(45 ({z.id} {z.type} {word(z.name)})())\n".encode()
                    for z in sorted(mesh.zones.values(), key=lambda z: z.id))


def write_fluent_mesh(mesh: UnstructuredMesh, path: str, binary: bool = False, title: str = "gputest_amd") -> None:
    """Write a Fluent .msh (ASCII, or binary double precision; gzipped when ``path`` ends in
    .gz) that ``read_fluent_mesh`` reads back to the same mesh with its faces sorted by zone
    (``fluent_msh.py:446-525``). Cell zones must be contiguous cell ranges (``ValueError``)."""
    mesh = sort_faces_by_zone(mesh)
    nodes_out = reverse_face_nodes(mesh.face_ptr, mesh.face_nodes) if mesh.dim == 3 else mesh.face_nodes
    data = b"".join([f'(0 "{title}: {mesh.n_cells} cells")\n(2 {mesh.dim})\n'.encode(),
                     f"(10 (0 1 {mesh.n_nodes:x} 0 {mesh.dim}))\n".encode(),
                     _node_section(mesh, binary), _cell_sections(mesh), _face_sections(mesh, nodes_out, binary),
                     _periodic_sections(mesh), _zone_sections(mesh)])
    if str(path).endswith(".gz"):
        data = gzip.compress(data)
    with open(path, "wb") as fh:
        fh.write(data)


__all__ = ["FluentMeshError", "read_fluent_mesh", "read_any_mesh", "write_fluent_mesh", "reverse_face_nodes",
           "BC_CODES", "BC_NAMES"]
