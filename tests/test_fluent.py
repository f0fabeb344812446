"""Fluent mesh reader/writer (``flowtrace/mesh/fluent_msh.py``) on the CPU: the reference's own
regression mesh, closed-form geometry of that mesh, write/read round trips (ASCII, binary,
gzip; 2-D, 3-D with mixed face records, periodic shadow pairs), the native number parsers
against Python's, malformed input, and the reference's square-cylinder drag and
recirculation pins reproduced by the oracle on the mesh this reader returns."""
import gzip

import numpy as np
import pytest

import oracle.simplec_oracle as O
from gputest_amd import _native as nat
from gputest_amd.case import default_boundaries
from gputest_amd.fluent import FluentMeshError, read_any_mesh, read_fluent_mesh, write_fluent_mesh
from gputest_amd.mesh import cartesian_mesh_2d, compute_geometry, extrude_to_3d, make_periodic, sort_faces_by_zone

from fluent_cases import MESH, PINS, case, drag_lift, recirculation_length


@pytest.fixture(scope="module")
def sqcyl():
    return read_any_mesh(MESH)[0]


def test_zones_are_detected_from_the_fluent_file(sqcyl):
    """As ``tests/test_square_cylinder.py:26-31``."""
    mesh, fmt = read_any_mesh(MESH)
    assert fmt == "fluent-msh" and mesh.dim == 2 and mesh.n_cells == 5824
    b = default_boundaries(mesh)
    assert b["inlet"].kind == "inlet" and b["outlet"].kind == "outlet"
    assert b["BluffWalls"].kind == "wall" and b["SideWalls"].kind == "wall"
    assert {z.type for z in mesh.face_zones()} == {"interior", "wall", "velocity-inlet", "pressure-outlet"}


def test_geometry_closed_forms(sqcyl):
    """The channel [-40, 40] x [-24, 24] minus the 4 x 4 square: area, boundary lengths, closed
    cells and outward boundary normals."""
    g = compute_geometry(sqcyl)
    lo, hi = sqcyl.nodes.min(0), sqcyl.nodes.max(0)
    assert np.allclose(lo, [-40, -24]) and np.allclose(hi, [40, 24])
    assert g.cell_volume.sum() == pytest.approx(80 * 48 - 16, rel=1e-12)
    assert (g.cell_volume > 0).all()
    length = {z.name: g.face_area[z.first:z.last + 1].sum() for z in sqcyl.face_zones()}
    assert length["BluffWalls"] == pytest.approx(16.0, rel=1e-12)
    assert length["inlet"] == pytest.approx(48.0, rel=1e-12) and length["outlet"] == pytest.approx(48.0, rel=1e-12)
    assert length["SideWalls"] == pytest.approx(160.0, rel=1e-12)
    # every cell is closed: its outward area vectors sum to zero
    s = np.zeros((sqcyl.n_cells, 2))
    inner = sqcyl.neighbour >= 0
    np.add.at(s, sqcyl.owner, g.face_area_vec)
    np.add.at(s, sqcyl.neighbour[inner], -g.face_area_vec[inner])
    assert np.abs(s).max() < 1e-12 * np.abs(g.face_area_vec).max()
    z = {zz.name: zz for zz in sqcyl.face_zones()}
    inlet = g.face_area_vec[z["inlet"].first:z["inlet"].last + 1]
    assert (inlet[:, 0] < 0).all()  # out of the domain at x = -40


def _same(a, b):
    assert a.dim == b.dim and a.n_cells == b.n_cells
    np.testing.assert_array_equal(a.face_ptr, b.face_ptr)
    np.testing.assert_array_equal(a.face_nodes, b.face_nodes)
    np.testing.assert_array_equal(a.owner, b.owner)
    np.testing.assert_array_equal(a.neighbour, b.neighbour)
    np.testing.assert_array_equal(a.nodes, b.nodes)
    za = sorted((z.name, z.type, z.kind, z.first, z.last) for z in a.zones.values())
    zb = sorted((z.name, z.type, z.kind, z.first, z.last) for z in b.zones.values())
    assert za == zb


@pytest.mark.parametrize("binary,suffix", [(False, ".msh"), (True, ".msh"), (False, ".msh.gz"), (True, ".cas.gz")])
def test_round_trip_of_the_regression_mesh(sqcyl, tmp_path, binary, suffix):
    path = tmp_path / ("m" + suffix)
    write_fluent_mesh(sqcyl, str(path), binary=binary)
    back, fmt = read_any_mesh(str(path))
    _same(sort_faces_by_zone(sqcyl), back)  # the writer puts interior zones first


def test_round_trip_3d_mixed_faces_and_periodic_pairs(tmp_path):
    """Extruded triangles: tetra-free prisms whose interior zone mixes 3- and 4-node faces
    (mixed records), 3-D face orientation, and a periodic pair written as shadow records."""
    m = extrude_to_3d(cartesian_mesh_2d(5, 4, triangulate=True, warp=0.05), 3, 0.6)
    m = make_periodic(m, "left", "right")
    counts = np.diff(m.face_ptr)
    assert set(np.unique(counts)) == {3, 4}
    for binary in (False, True):
        path = tmp_path / f"prism{int(binary)}.msh"
        write_fluent_mesh(m, str(path), binary=binary)
        back = read_fluent_mesh(str(path))
        ref = sort_faces_by_zone(m)
        _same(ref, back)
        g0, g1 = compute_geometry(ref), compute_geometry(back)
        np.testing.assert_allclose(g1.face_area_vec, g0.face_area_vec, rtol=0, atol=1e-15)
        assert len(back.periodic) == 1
        p0, p1 = ref.periodic[0], back.periodic[0]
        np.testing.assert_array_equal(p1["faces_a"], p0["faces_a"])
        np.testing.assert_array_equal(p1["faces_b"], p0["faces_b"])
        np.testing.assert_allclose(p1["translation"], p0["translation"], atol=1e-12)


def test_boundary_face_stored_with_c0_zero_is_flipped():
    """A boundary face written as (n0 n1 0 c) becomes (n1 n0 c 0): same area vector direction
    out of the domain."""
    text = b"""(2 2)
(10 (1 1 4 1 2)(
0 0
1 0
1 1
0 1
))
(12 (2 1 1 1 3))
(13 (3 1 1 3 2)(
2 1 0 1
))
(13 (4 2 4 3 2)(
2 3 1 0
3 4 1 0
4 1 1 0
))
(45 (2 fluid fluid)())
(45 (3 wall bottom)())
(45 (4 wall rest)())
"""
    m = read_fluent_mesh(text)
    assert m.n_cells == 1 and (m.owner == 0).all() and (m.neighbour == -1).all()
    np.testing.assert_array_equal(m.face_nodes[:2], [0, 1])
    g = compute_geometry(m)
    assert g.face_area_vec[0][1] < 0  # bottom face: outward is -y
    assert g.cell_volume[0] == pytest.approx(1.0)


def test_native_parsers_match_python():
    rng = np.random.default_rng(0)
    ints = rng.integers(0, 2 ** 40, 5000)
    text = " ".join(format(int(v), "x" if i % 2 else "X") for i, v in enumerate(ints)).encode()
    text = text.replace(b" ", b"\n", 100).replace(b" ", b"\t", 50)
    out = np.empty(6000, np.int64)
    k = nat.lib().ftmi_parse_hex(text, len(text), out.ctypes.data, out.size)
    assert k == ints.size
    np.testing.assert_array_equal(out[:k], ints)
    vals = rng.standard_normal(3000) * 10.0 ** rng.integers(-30, 30, 3000)
    text = "  ".join(repr(float(v)) for v in vals).encode() + b"\r\n 1e+001 -4.0000000000e+001"
    out = np.empty(4000)
    k = nat.lib().ftmi_parse_f64(text, len(text), out.ctypes.data, out.size)
    assert k == vals.size + 2
    np.testing.assert_array_equal(out[:vals.size], vals)  # round-trip exact
    assert out[vals.size] == 10.0 and out[vals.size + 1] == -40.0
    # malformed token, too many tokens, empty input
    assert nat.lib().ftmi_parse_hex(b"1 2 zz 4", 8, out.ctypes.data, 10) == -1 - 2
    assert nat.lib().ftmi_parse_hex(b"1 2 3", 5, out.ctypes.data, 2) == -2 - 2
    assert nat.lib().ftmi_parse_f64(b"1.0 x", 5, out.ctypes.data, 10) == -1 - 1
    assert nat.lib().ftmi_parse_f64(b" \n ", 3, out.ctypes.data, 10) == 0


def test_section_index_skips_strings_settings_and_binary_sections():
    """The native section table: a comment whose string holds parentheses, case settings with
    nested groups and a foreign binary section are skipped; grid sections report their header
    and data spans; broken structure is reported with its position."""
    from gputest_amd.fluent import _index

    buf = (b'(0 "a ) tricky ( comment")\n(2 2)\nstray line\n(37 (settings (a (b c)) "x)"))\n'
           b"(2099 (1 2)(\x01)\x00)(\nEnd of Binary Section 2099)\n"
           b"(10 (0 1 3 0 2))\n(10 (1 1 3 1 2)(\n0 0\n1 0\n0 1\n))\n(45 (2 wall w)())\n")
    t = _index(buf)
    assert [int(r[0]) for r in t] == [2, 10, 10, 45]
    assert buf[t[0][1]:t[0][2]] == b"2"
    assert buf[t[2][1]:t[2][2]] == b"1 1 3 1 2" and buf[t[2][3]:t[2][4]].split() == [b"0", b"0", b"1", b"0", b"0", b"1"]
    assert t[1][5] == 0 and t[2][5] == 1
    for bad, what in ((b'(0 "never closed)\n', "string"), (b"(10 (1 1 3 1 2)(\n0 0\n", "not closed"),
                      (b"(3010 (1 1 1 1 2)(\x00\x00)", "End of Binary"), (b"(13 (1 1 2 3 2)(1 (2) 3))", "nested")):
        with pytest.raises(FluentMeshError, match=what):
            read_fluent_mesh(bad)


def test_malformed_files_raise(tmp_path):
    with pytest.raises(FluentMeshError):
        read_fluent_mesh(b"(2 2)\n(10 (1 1 2 1 2)(\n0 0\n1 zz\n))\n")
    with pytest.raises(FluentMeshError):
        read_fluent_mesh(b"(0 \"only a comment\")\n")
    with pytest.raises(FileNotFoundError):
        read_fluent_mesh(str(tmp_path / "missing.msh"))
    with open(MESH, "rb") as fh:
        data = gzip.decompress(fh.read())
    with pytest.raises(FluentMeshError):
        read_fluent_mesh(data[: len(data) // 2])  # cut inside the face sections


@pytest.mark.slow
@pytest.mark.parametrize("sides,cd,lr", PINS)
def test_square_cylinder_re40_oracle_reproduces_the_reference_pins(sqcyl, sides, cd, lr):
    """The oracle on the mesh this reader returns gives the drag coefficient and recirculation
    length the reference pins (``tests/test_square_cylinder.py:34-53``)."""
    s = O.Simplec(sqcyl, case(O, sqcyl, sides))
    while not s.finished:
        s.step()
    assert s.converged
    cdx, cl = drag_lift(s.wall_forces())
    assert abs(cdx - cd) < 2e-4
    assert abs(cl) < 1e-6
    xc = compute_geometry(sqcyl).cell_centre
    assert abs(recirculation_length(xc, s.U[:, 0]) - lr) < 5e-3
