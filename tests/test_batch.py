"""Batch construction on the CPU: the concatenated mesh is the cases side by side (same
geometry cell by cell, zones renamed), and the case sampler follows its distribution."""
import numpy as np

from gputest_amd.batch import concatenate_meshes, sample_cavities
from gputest_amd.mesh import cartesian_mesh_2d, compute_geometry


def test_concatenated_mesh_is_the_cases_side_by_side():
    ms = [cartesian_mesh_2d(5, 4), cartesian_mesh_2d(3, 3, triangulate=True, warp=0.05), cartesian_mesh_2d(2, 6)]
    u = concatenate_meshes(ms, ["a/", "b/", "c/"])
    assert u.n_cells == sum(m.n_cells for m in ms) and u.n_faces == sum(m.n_faces for m in ms)
    gu = compute_geometry(u)
    off = 0
    for m, pre in zip(ms, "abc"):
        g = compute_geometry(m)
        assert np.allclose(gu.cell_volume[off:off + m.n_cells], g.cell_volume, rtol=0, atol=1e-15)
        assert np.allclose(gu.cell_centre[off:off + m.n_cells], g.cell_centre, rtol=0, atol=1e-15)
        for z in m.face_zones():
            uz = u.zone_by_name(f"{pre}/{z.name}")
            assert uz is not None and uz.count == z.count and uz.type == z.type
        off += m.n_cells
    # interior faces first, and no face joins two cases
    inner = u.neighbour >= 0
    assert not np.any(inner[np.argmin(inner):])
    case_of = np.repeat(np.arange(3), [m.n_cells for m in ms])
    assert np.all(case_of[u.owner[inner]] == case_of[u.neighbour[inner]])


def test_case_sampler():
    specs = sample_cavities(256, seed=0)
    cells = np.array([s.cells for s in specs])
    assert cells.min() >= 15_000 and cells.max() <= 205_000
    assert sum(s.triangles for s in specs) == 128
    re = np.array([s.reynolds for s in specs])
    assert re.min() >= 100 and re.max() <= 1000
    # log-uniform: the median cell count sits near the geometric mean of the range
    assert 40_000 < np.median(cells) < 80_000
    assert [s.n for s in sample_cavities(8, seed=0)] == [s.n for s in specs[:8]]
    print("total cells", int(cells.sum()))
