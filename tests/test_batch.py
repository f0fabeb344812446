"""Batch construction on the CPU: the concatenated mesh is the cases side by side (same
geometry cell by cell, zones renamed), the case samplers follow their distributions, and the
backward-facing step has the reference step's geometry and conserves mass in the oracle to the
reference's own bar (net imbalance / inflow < 1e-8, ``validation/benchmarks.py:463-487``)."""
import dataclasses

import numpy as np
import pytest

import oracle.simplec_oracle as O
from gputest_amd.batch import CavitySpec, StepSpec, concatenate_meshes, sample_cases, sample_cavities
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


def test_step_mesh_geometry():
    for tri in (False, True):
        spec = StepSpec(4, tri, 100.0, length=6.0)
        m = spec.mesh()
        g = compute_geometry(m)
        assert m.n_cells == spec.cells
        assert abs(g.cell_volume.sum() - (2 * 6.0 - 1.0)) < 1e-12
        assert sorted(z.name for z in m.face_zones() if z.type != "interior") == \
            ["bottom", "inlet", "outlet", "step", "top"]
        fc = g.face_centre
        for name, lo, hi in (("inlet", (0, 1), (0, 2)), ("outlet", (6, 0), (6, 2)), ("step", (0, 0), (1, 1))):
            f = m.faces_of_zone(m.zone_by_name(name).id)
            assert np.all(fc[f] >= np.array(lo) - 1e-12) and np.all(fc[f] <= np.array(hi) + 1e-12), name
        # the step's two faces of length 1 (riser x = 1, tread y = 1), each face out of the fluid
        f = m.faces_of_zone(m.zone_by_name("step").id)
        area = np.linalg.norm(g.face_area_vec[f], axis=1).sum()
        assert abs(area - 2.0) < 1e-12
        out = np.einsum("fd,fd->f", g.face_area_vec[f], fc[f] - g.cell_centre[m.owner[f]])
        assert np.all(out > 0)


def test_case_sampler_mixes_cavities_and_steps():
    specs = sample_cases(256, seed=0)
    kinds = [(type(s).__name__, s.triangles) for s in specs]
    for key in [("CavitySpec", False), ("CavitySpec", True), ("StepSpec", False), ("StepSpec", True)]:
        assert kinds.count(key) == 64
    cells = np.array([s.cells for s in specs])
    assert cells.min() >= 14_000 and cells.max() <= 215_000
    assert 40_000 < np.median(cells) < 80_000
    assert all(isinstance(s, (CavitySpec, StepSpec)) for s in specs)


def _oracle_case(spec, nu, **kw):
    bnd = {k: O.Boundary(**{f.name: getattr(b, f.name) for f in dataclasses.fields(O.Boundary)})
           for k, b in spec.boundaries("", nu).items()}
    return O.Case(boundaries=bnd, viscosity=nu, **kw)


@pytest.mark.slow
def test_step_case_conserves_mass_in_the_oracle():
    """The reference's step check: Re = 100, 6 cells per step height, converged to 1e-6; the
    net boundary mass flux over the inflow below 1e-8; a recirculation behind the step."""
    spec = StepSpec(6, False, 100.0, length=10.0)
    nu = 0.01
    s = O.Simplec(spec.mesh(), _oracle_case(spec, nu, max_iter=3000, tol=1e-6))
    while not s.finished:
        s.step()
    assert s.converged, s.iteration
    Fb = s.Fb
    mb = {z.name: float(Fb[s.fv.bzone == z.id].sum()) for z in s.mesh.face_zones() if (s.fv.bzone == z.id).any()}
    inflow = -mb["inlet"]
    # mean velocity 1 over the unit inlet, by the midpoint rule on n faces: 1 + 1 / (2 n^2)
    assert abs(inflow - (1.0 + 1.0 / (2 * spec.n ** 2))) < 1e-12
    assert abs(sum(mb.values())) / inflow < 1e-8, mb
    xc = s.fv.xc
    near = (xc[:, 0] > 1.0) & (xc[:, 0] < 2.0) & (xc[:, 1] < 0.2)
    assert s.U[near, 0].min() < 0.0
    print("step", s.iteration, "iterations", mb)
