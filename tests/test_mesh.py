"""Generators reproduce the reference's meshes bit for bit; geometry matches its closed forms."""
import hashlib
import json
import os

import numpy as np
import pytest

from gputest_amd.mesh import cartesian_mesh_2d, channel_mesh, compute_geometry, duct_mesh, extrude_to_3d

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _h(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def hashes():
    with open(os.path.join(GOLD, "mesh_hashes.json")) as fh:
        return json.load(fh)


def test_duct_million_triangles_bit_identical(hashes):
    m = duct_mesh(1000, 500)
    assert m.n_cells == 1_000_000
    for k, v in hashes["duct1000x500"].items():
        assert _h(getattr(m, k)) == v, k


def test_cavity_bit_identical(hashes):
    m = cartesian_mesh_2d(64, 64)
    for k, v in hashes["cavity64"].items():
        assert _h(getattr(m, k)) == v, k


def test_channel_hexes_bit_identical(hashes):
    m = channel_mesh(128, 64, 64)
    assert m.n_cells == 524_288
    ref = hashes["channel128x64x64"]
    for k in ("nodes", "face_ptr", "face_nodes", "owner", "neighbour", "face_zone"):
        assert _h(getattr(m, k)) == ref[k], k
    assert _h(m.periodic[0]["faces_a"]) == ref["faces_a"]
    assert _h(m.periodic[0]["faces_b"]) == ref["faces_b"]


def test_geometry_closed_forms():
    # uniform quads: unit volume sum, centroids at cell centres, no flipped faces
    g = compute_geometry(cartesian_mesh_2d(8, 5, 2.0, 1.0))
    assert g.flipped_faces == 0
    assert abs(g.cell_volume.sum() - 2.0) < 1e-14
    assert np.allclose(g.cell_volume, 2.0 / 40)
    # warped triangles and extruded prisms still tile the box exactly
    assert abs(compute_geometry(cartesian_mesh_2d(9, 7, triangulate=True, warp=0.05)).cell_volume.sum() - 1) < 1e-13
    m3 = extrude_to_3d(cartesian_mesh_2d(5, 4, triangulate=True, jitter=0.2), 3, 0.6)
    g3 = compute_geometry(m3)
    assert abs(g3.cell_volume.sum() - 0.6) < 1e-13
    # every closed cell: the outward area vectors sum to zero
    S = g3.face_area_vec
    tot = np.zeros((m3.n_cells, 3))
    for a in range(3):
        tot[:, a] = np.bincount(m3.owner, S[:, a], minlength=m3.n_cells)
        inner = m3.neighbour >= 0
        tot[:, a] -= np.bincount(m3.neighbour[inner], S[inner, a], minlength=m3.n_cells)
    assert np.abs(tot).max() < 1e-14
