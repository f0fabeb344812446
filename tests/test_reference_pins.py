"""The reference's physics tests of the SIMPLE/SIMPLEC path (``tests/test_unstructured.py:27-71``)
on the oracle, and the reference's three example meshes (``examples/``, gzipped fixtures):
read with their zones and closed-form geometry, and the README's cases run by the oracle."""
import numpy as np
import pytest

import oracle.simplec_oracle as O
from gputest_amd.fluent import read_any_mesh
from gputest_amd.mesh import cartesian_mesh_2d, compute_geometry

from fluent_cases import EXAMPLES, example_case
from reference_pins import PINS, ghia_errors, mass_balance, run


@pytest.mark.parametrize("name", list(PINS))
def test_oracle_meets_the_reference_pin(name):
    build, check = PINS[name]
    mesh, case = build(O)
    s = run(O.Simplec(mesh, case))
    print(name, "iterations", s.iteration, check(s))


# name: (cells, {zone: (Fluent type, faces, total length)}, total area)
EXAMPLE_GEOMETRY = {
    "heated_duct_tri": (2400, {"inlet": ("velocity-inlet", 20, 2.0), "outlet": ("pressure-outlet", 20, 2.0),
                               "floor": ("wall", 60, 6.0), "ceiling": ("wall", 60, 6.0)}, 12.0),
    "cavity_quad": (1600, {"left": ("wall", 40, 1.0), "right": ("wall", 40, 1.0), "bottom": ("wall", 40, 1.0),
                           "lid": ("wall", 40, 1.0)}, 1.0),
    "elbow_mix": (648, {"wall.4": ("wall", 130, None), "out": ("pressure-outlet", 10, 100.0),
                        "inf1": ("velocity-inlet", 10, 100.0), "inf2": ("velocity-inlet", 6, 25.0)}, None),
}


@pytest.mark.parametrize("name", list(EXAMPLES))
def test_reference_example_meshes_read_closed(name):
    """ASCII (heated duct; the GAMBIT elbow) and binary double precision (cavity) files: cells,
    boundary zones and their lengths, positive areas, closed cells, no inward boundary face."""
    mesh, fmt = read_any_mesh(EXAMPLES[name])
    cells, zones, area = EXAMPLE_GEOMETRY[name]
    assert fmt == "fluent-msh" and mesh.dim == 2 and mesh.n_cells == cells
    g = compute_geometry(mesh)
    assert g.flipped_faces == 0 and (g.cell_volume > 0).all()
    if area is not None:
        assert g.cell_volume.sum() == pytest.approx(area, rel=1e-12)
    got = {z.name: (z.type, z.count, g.face_area[z.first:z.last + 1].sum()) for z in mesh.boundary_zones()}
    assert sorted(got) == sorted(zones)
    for z, (typ, count, length) in zones.items():
        assert got[z][:2] == (typ, count), (z, got[z])
        if length is not None:
            assert got[z][2] == pytest.approx(length, rel=1e-12)
    s = np.zeros((mesh.n_cells, 2))
    inner = mesh.neighbour >= 0
    np.add.at(s, mesh.owner, g.face_area_vec)
    np.add.at(s, mesh.neighbour[inner], -g.face_area_vec[inner])
    assert np.abs(s).max() < 1e-12 * np.abs(g.face_area_vec).max()


def test_binary_cavity_is_the_uniform_40_by_40_grid():
    mesh = read_any_mesh(EXAMPLES["cavity_quad"])[0]
    ref = cartesian_mesh_2d(40, 40)
    a, b = compute_geometry(mesh), compute_geometry(ref)
    key = lambda c: np.lexsort((c[:, 0], c[:, 1]))  # noqa: E731
    np.testing.assert_allclose(a.cell_centre[key(a.cell_centre)], b.cell_centre[key(b.cell_centre)], atol=1e-15)
    np.testing.assert_allclose(np.sort(a.cell_volume), np.sort(b.cell_volume), rtol=1e-12)


@pytest.mark.slow
def test_oracle_runs_the_example_cases():
    """The README's two command lines: the cavity at Re = 100 against Ghia's profiles (the bar
    of ``test_unstructured.py:34``), and the heated duct's mass balance and temperature bounds
    (the bars of ``test_unstructured.py:56-59``)."""
    mesh = read_any_mesh(EXAMPLES["cavity_quad"])[0]
    s = run(O.Simplec(mesh, example_case(O, mesh, "cavity_quad", 3000, 1e-6)))
    e = ghia_errors(s)
    print("cavity_quad", s.iteration, e)
    assert s.converged and e["u_max_err"] < 0.03 and e["v_max_err"] < 0.03
    mesh = read_any_mesh(EXAMPLES["heated_duct_tri"])[0]
    s = run(O.Simplec(mesh, example_case(O, mesh, "heated_duct_tri", 3000, 1e-6)))
    mb = mass_balance(s)
    print("heated_duct_tri", s.iteration, mb, s.heat_flow())
    assert s.converged
    assert abs(mb["inlet"] + 2.0) < 1e-9 and abs(sum(mb.values())) < 1e-8
    assert -1e-8 <= s.T.min() and s.T.max() <= 1.0 + 1e-6
