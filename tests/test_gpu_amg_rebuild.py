"""The AMG rebuild (``SmoothedAggregationAMG.rebuild_ldu``) against a full set-up of the same
matrix (``from_ldu``): the reference re-aggregates at every rebuild (``amg.py:80-146``); the
rebuild keeps a level's aggregates only where its new strength mask equals the old one, which
is exactly where a fresh greedy aggregation (a function of the pattern and the mask) gives
them again. So every level must come out bit-identical to the full set-up, whichever level
the masks first differ at. Also the hierarchy cut at a level count, whose coarsest system is
solved by sweeps."""
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


def test_solver_reuses_a_hierarchy_only_for_its_exact_matrix():
    """The solver's hierarchy bookkeeping (``SimpleSolver._reuse_hierarchy``, then
    ``_next_hierarchy``) over a fixed sequence of pressure matrices: a hierarchy is used as it is
    only for the matrix it was set up for, bit for bit; the current one is rebuilt when the finest
    strength mask stays, and a new one is set up (keeping the current one) when it changes.
    After every step the hierarchy in
    use equals a full set-up of the matrix."""
    from gputest_amd import case as C
    from gputest_amd.simple import SimpleSolver

    mesh = cartesian_mesh_2d(64, 48, triangulate=True, warp=0.05)
    s = SimpleSolver(mesh, C.UnstructuredCase(boundaries={"top": C.BoundarySpec("wall", velocity=[1.0, 0.0])},
                                              viscosity=0.01, max_iter=5, tol=0.0), DEV)
    fv = s.fv
    rng = np.random.default_rng(5)
    a1, a2 = rng.uniform(0.001, 2.0, fv.n_if), rng.uniform(0.001, 2.0, fv.n_if)  # weak entries differ
    a1n = a1 * (1.0 + 1e-9 * rng.uniform(-1.0, 1.0, fv.n_if))  # every strength test as for a1
    a2n = a2 * (1.0 + 1e-9 * rng.uniform(-1.0, 1.0, fv.n_if))
    a1u = a1.copy()
    a1u[7] = np.nextafter(a1u[7], 3.0)  # one ulp on one face
    # (current, kept) after each: H1 -; H2 H1; H1 H2; H1 H2; H1u H2; H2 H1u; H1n H2; H2 H1n; H2n H1n; H1n H2n
    seq = [(a1, "new"), (a2, "new"), (a1, "kept"), (a1, "current"), (a1u, "rebuilt"), (a2, "kept"),
           (a1n, "new"), (a2, "kept"), (a2n, "rebuilt"), (a1n, "kept")]
    got = []
    for a, want in seq:
        diag, off = _laplacian(fv, a)
        s.diag_p.copy_(diag)
        s.offp.copy_(off)
        how = (s._reuse_hierarchy(fv) if s._amg is not None else None) or s._next_hierarchy(fv)
        got.append(how)
        _assert_identical(s._amg, SmoothedAggregationAMG.from_ldu(fv, diag, off, singular=s.singular_p))
        assert s._amg.matches_ldu(fv, s.diag_p, s.offp)
    print("hierarchy outcomes", got)
    assert got == [w for _, w in seq]


def _host_vcycle(amg, b):
    """The V(2,2) cycle of amg.py:130-146 in NumPy over the device hierarchy's own matrices,
    with the coarsest system solved as the hierarchy says: a dense inverse, or `coarse_sweeps`
    weighted-Jacobi sweeps from zero with the hierarchy's coarse weight."""
    levels = [(A.to_scipy(), P.to_scipy(), R.to_scipy(), dinv.cpu().numpy()) for A, P, R, dinv in amg.levels]
    w, nu = amg.omega, amg.sweeps
    Ac = amg.coarse.to_scipy()

    def coarse(bc):
        if amg.coarse_inv_h is not None:
            return amg.coarse_inv_h @ bc
        dc, wc = 1.0 / Ac.diagonal(), amg._comega
        x = wc * dc * bc
        for _ in range(amg.coarse_sweeps - 1):
            x = x + wc * dc * (bc - Ac @ x)
        return x

    def cycle(k, bk):
        if k == len(levels):
            return coarse(bk)
        A, P, R, D = levels[k]
        x = w * D * bk
        for _ in range(nu - 1):
            x = x + w * D * (bk - A @ x)
        x = x + P @ cycle(k + 1, R @ (bk - A @ x))
        for _ in range(nu):
            x = x + w * D * (bk - A @ x)
        return x

    return cycle(0, b)


def test_hierarchy_cut_at_three_levels_solves_its_coarsest_system_by_sweeps():
    """max_levels = 3 on a 200 x 200 Laplacian, whose levels are 40,000 and 6,700 unknowns and
    whose third (coarsest) one has 835 (the reference would go on to 124): the coarsest system is
    solved by weighted-Jacobi sweeps instead of a dense inverse. The device V-cycle equals the
    NumPy cycle over the same matrices, its FP32 twin follows it, and the cycle is a symmetric
    positive definite preconditioner (PCG converges)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    from gputest_amd.fvmesh import DeviceCSR
    from gputest_amd.krylov import pcg_csr, to_f32

    m = 200
    T1 = sp.diags([-np.ones(m - 1), 2 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
    A = (sp.kron(T1, sp.identity(m)) + sp.kron(sp.identity(m), T1)).tocsr()
    A.sort_indices()
    amg = SmoothedAggregationAMG(A, DEV, max_levels=3, coarse_sweeps=20)
    assert len(amg.levels) == 2 and amg.coarse.n > amg.max_coarse and amg.coarse_inv_h is None
    # the weight 1 / (1.1 rho) of the set-up's power estimate (a lower bound of rho): the sweeps
    # contract every mode (omega lambda_max < 2) and the weight is not below 1 / (1.1 lambda_max)
    Ac = amg.coarse.to_scipy()
    dh = sp.diags(1.0 / np.sqrt(Ac.diagonal()))
    lam = spla.eigsh(dh @ Ac @ dh, k=1, which="LA", return_eigenvectors=False)[0]
    assert 0.9999 / (1.1 * lam) <= amg._comega < 2.0 / lam, (amg._comega, lam)
    rng = np.random.default_rng(1)
    b = rng.standard_normal(A.shape[0])
    z = amg(t(b)).cpu().numpy()
    zh = _host_vcycle(amg, b)
    assert np.linalg.norm(z - zh) <= 1e-12 * np.linalg.norm(zh)
    z32 = amg.to_fp32()(to_f32(t(b))).double().cpu().numpy()
    assert np.linalg.norm(z32 - zh) <= 1e-5 * np.linalg.norm(zh)
    y = rng.standard_normal(A.shape[0])
    zy = amg(t(y)).cpu().numpy()
    assert abs(y @ z - b @ zy) <= 1e-10 * abs(y @ z)  # symmetric
    Ad = DeviceCSR(t(A.data), torch.as_tensor(A.indices.astype(np.int32), device=DEV),
                   torch.as_tensor(A.indptr.astype(np.int32), device=DEV), A.shape)
    x, info = pcg_csr(Ad, t(b), 1e-10, 200, amg=amg)
    print("pcg with the three-level cycle", info)
    assert info.converged
    assert np.linalg.norm(A @ x.cpu().numpy() - b) <= 1e-9 * np.linalg.norm(b)
