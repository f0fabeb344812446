"""The AMG rebuild (``SmoothedAggregationAMG.rebuild_ldu``) against a full set-up of the same
matrix (``from_ldu``): the reference re-aggregates at every rebuild (``amg.py:80-146``); the
rebuild keeps a level's aggregates only where its new strength mask equals the old one, which
is exactly where a fresh greedy aggregation (a function of the pattern and the mask) gives
them again. So every level must come out bit-identical to the full set-up, whichever level
the masks first differ at."""
import numpy as np
import pytest
import torch

from gputest_amd.amg import SmoothedAggregationAMG
from gputest_amd.fvmesh import FVMesh
from gputest_amd.mesh import cartesian_mesh_2d

pytestmark = pytest.mark.gpu
DEV = "cuda"


def t(a):
    return torch.as_tensor(np.ascontiguousarray(a, np.float64), device=DEV)


def _laplacian(fv, a):
    """Symmetric face-operator system with face weights a (a pinned Laplacian), LDU form."""
    diag = np.zeros(fv.nc)
    np.add.at(diag, fv.P_h, a)
    np.add.at(diag, fv.N_h, a)
    diag[0] += 1.0
    return t(diag), t(fv.ldu_off_from_faces(-a, -a))


def _same(x, y):
    return bool(torch.equal(x, y))


def _assert_identical(amg, ref):
    assert len(amg.levels) == len(ref.levels)
    for k, ((A, P, R, dinv), (Ar, Pr, Rr, dr)) in enumerate(zip(amg.levels, ref.levels)):
        np.testing.assert_array_equal(amg.aggregates[k], ref.aggregates[k])
        for x, y in ((A, Ar), (P, Pr), (R, Rr)):
            assert x.n == y.n and x.ncols == y.ncols
            assert _same(x.ptr, y.ptr) and _same(x.col, y.col), k
            assert _same(x.val, y.val), (k, float((x.val - y.val).abs().max()))
        assert _same(dinv, dr), k
    assert _same(amg.coarse.val, ref.coarse.val)
    np.testing.assert_array_equal(amg.coarse_inv_h, ref.coarse_inv_h)
    b = t(np.random.default_rng(3).standard_normal(amg.levels[0][0].n))
    assert _same(amg(b), ref(b))


@pytest.mark.parametrize("how", ["all kept", "coarse levels new", "all new"])
def test_rebuild_equals_a_full_set_up(how):
    fv = FVMesh(cartesian_mesh_2d(64, 48, triangulate=True, warp=0.05), DEV)
    rng = np.random.default_rng(11)
    nf = fv.n_if
    if how == "all kept":  # values move by 1e-9: no strength test changes its outcome
        a1 = rng.uniform(0.5, 2.0, nf)
        a2 = a1 * (1.0 + 1e-9 * rng.uniform(-1.0, 1.0, nf))
    elif how == "coarse levels new":  # every fine entry strong either way; coarse masks move
        a1, a2 = rng.uniform(0.5, 2.0, nf), rng.uniform(0.5, 2.0, nf)
    else:  # weak fine entries come and go
        a1, a2 = rng.uniform(0.001, 2.0, nf), rng.uniform(0.001, 2.0, nf)
    amg = SmoothedAggregationAMG.from_ldu(fv, *_laplacian(fv, a1), refreshable=True)
    nlev = len(amg.levels)
    amg.rebuild_ldu(fv, *_laplacian(fv, a2))
    ref = SmoothedAggregationAMG.from_ldu(fv, *_laplacian(fv, a2))
    kept = amg.timing["levels_kept"]
    if how == "all kept":
        assert kept == nlev
    elif how == "coarse levels new":
        assert 1 <= kept < nlev
    else:
        assert kept == 0
    _assert_identical(amg, ref)
    # a second rebuild on top of the first (its levels mix kept and new ones)
    amg.rebuild_ldu(fv, *_laplacian(fv, a1))
    _assert_identical(amg, SmoothedAggregationAMG.from_ldu(fv, *_laplacian(fv, a1)))


def test_rebuild_of_the_solvers_pressure_system_equals_a_full_set_up():
    """Inside the solver: after each rebuild of a lid-driven cavity run, the hierarchy equals a
    full set-up of the pressure matrix it was rebuilt for."""
    from gputest_amd import case as C
    from gputest_amd.simple import SimpleSolver

    mesh = cartesian_mesh_2d(40, 40, triangulate=True, warp=0.05)
    s = SimpleSolver(mesh, C.UnstructuredCase(boundaries={"top": C.BoundarySpec("wall", velocity=[1.0, 0.0])},
                                              viscosity=0.01, max_iter=31, tol=0.0), DEV)
    checked = 0
    while not s.finished:
        s.step()
        if s._amg_age == 1 and s.amg_rebuilds > 1:  # this iteration rebuilt the hierarchy
            diag, off = s.diag_p.clone(), s.offp.clone()
            ref = SmoothedAggregationAMG.from_ldu(s.fv, diag, off, singular=s.singular_p)
            _assert_identical(s._amg, ref)
            checked += 1
    assert checked >= 2
