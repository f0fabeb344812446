"""The oracle is pinned: against the reference's own golden vectors, against the reference's
compiled C++ kernels (oracle/_ref, when built), and against closed forms."""
import ctypes
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle.simplec_oracle as O
from gputest_amd.mesh import cartesian_mesh_2d, extrude_to_3d
from tests.golden_cases import capture_case, cavity_case, load_pre_state

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REF_LIB = os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle", "_ref", "libflowtrace_fvm.so")


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("tag", ["tri16", "hex3d", "periodic", "duct"])
def test_oracle_reproduces_captured_iteration(tag):
    g = np.load(os.path.join(GOLD, f"capture_{tag}.npz"))
    mesh, case = capture_case(tag, O)
    s = O.Simplec(mesh, case)
    for k in ("P", "N", "Pb", "w", "alpha", "alpha_b", "Ginv", "dxm", "Tvec", "vol", "db_t", "dP_f", "dN_f"):
        assert rel(getattr(s.fv, k), g["fv_" + k]) < 1e-13, k
    assert s.relax_p == float(g["relax_p"])
    load_pre_state(s, g)
    s.step()
    c = s._capture
    # everything up to the pressure solve is the reference's arithmetic, bit for bit
    assert rel(c["A"].data, g["A_data"]) == 0.0
    assert rel(np.stack(c["b"]), g["b"]) == 0.0
    assert rel(c["Ar"].data, g["Ar_data"]) == 0.0
    assert rel(c["U_star"], g["U_star"]) == 0.0
    assert c["mom_iters"] == g["mom_iters"].tolist()
    for k in ("Fst", "Fbst", "mass"):
        assert rel(c[k], g[k]) == 0.0, k
    assert rel(c["Ap"].data, g["Ap_data"]) == 0.0
    # the correction is solved to 1e-12 with a (differently aged) AMG preconditioner
    assert rel(c["pp"], g["pp"]) < 1e-11
    assert rel(s.F, g["F_new"]) < 1e-11
    assert rel(s.U, g["U_new"]) < 1e-11
    assert rel(s.p, g["p_new"]) < 1e-11
    if case.energy:
        assert c["T_iters"] == int(g["T_iters"])
        assert rel(s.T, g["T_new"]) < 1e-8  # ill-conditioned energy system (60 BiCGSTAB steps)
    for k, v in s.residuals()["scaled"].items():
        assert abs(v - float(g["res_scaled_" + k])) <= 1e-9 * abs(float(g["res_scaled_" + k])) + 1e-15, k


def test_oracle_cavity_300_iterations_bit_identical():
    g = np.load(os.path.join(GOLD, "cavity64_simplec300.npz"))
    s = O.Simplec(cartesian_mesh_2d(64, 64), cavity_case(O))
    its, pits = [], []
    while not s.finished:
        s.step()
        its.append(s.last_linear["momentum_iters"])
        pits.append(s.last_linear["pressure_iters"])
        if s.iteration == 10:
            assert rel(s.U, g["it10_U"]) == 0.0
    assert np.array_equal(its, g["mom_iters"]) and np.array_equal(pits, g["p_iters"])
    assert rel(s.U, g["U"]) == 0.0 and rel(s.p, g["p"]) == 0.0 and rel(s.F, g["F"]) == 0.0


def test_oracle_amg_hierarchy_matches_reference():
    g = np.load(os.path.join(GOLD, "amg_tri40.npz"))
    A = sp.csr_matrix((g["A0_data"], g["A0_indices"], g["A0_indptr"]))
    M = O.SAAMG(A, singular=True)
    assert M.n_levels == int(g["n_levels"])
    for k, (Ak, P, R, Dinv) in enumerate(M.levels):
        assert np.array_equal(M.aggregates[k], g[f"L{k}_agg"])
        assert np.array_equal(P.indptr, g[f"L{k}_P_indptr"]) and np.array_equal(P.indices, g[f"L{k}_P_indices"])
        assert rel(P.data, g[f"L{k}_P_data"]) < 1e-14
    assert rel(M.coarse_inv, g["coarse_inv"]) < 1e-9
    assert rel(M(g["vcycle_b"]), g["vcycle_x"]) < 1e-10


def _mesh_fv(tri=True):
    m = cartesian_mesh_2d(20, 16, triangulate=tri, warp=0.05) if tri else \
        extrude_to_3d(cartesian_mesh_2d(6, 5, warp=0.05), 4, 0.5)
    return O.FV(m)


@pytest.mark.parametrize("tri", [True, False])
def test_gradients_exact_for_linear_fields(tri):
    fv = _mesh_fv(tri)
    a = np.array([0.3, -1.7, 0.9])[: fv.dim]
    phi = fv.xc @ a + 2.0
    phib = fv.xfb @ a + 2.0
    assert np.abs(fv.gradient(phi, phib) - a).max() < 1e-11
    # Green-Gauss: exact for a linear field once the skew correction sees the exact gradient
    g = np.tile(a, (fv.nc, 1))
    assert np.abs(fv.gradient_gg(phi, phib, g) - a).max() < 1e-10
    # and zero for a constant field (closed cells)
    assert np.abs(fv.gradient_gg(np.ones(fv.nc), np.ones(fv.n_bf))).max() < 1e-12
    # a dropped skew term would not be exact on this skewed mesh
    assert np.abs(fv.gradient_gg(phi, phib) - a).max() > 1e-6


def test_matrix_and_pressure_system_properties():
    g = np.load(os.path.join(GOLD, "capture_tri16.npz"))
    mesh, case = capture_case("tri16", O)
    s = O.Simplec(mesh, case)
    load_pre_state(s, g)
    s.step()
    Ap = s._capture["Ap"]
    assert abs(Ap - Ap.T).max() == 0.0  # symmetric
    assert np.abs(Ap @ np.ones(s.nc)).max() < 1e-12 * abs(Ap.diagonal()).max()  # closed cavity: singular
    # a dense solve of the reference's projected system reproduces the correction
    dense = Ap.toarray() + np.ones((s.nc, s.nc))
    pp = np.linalg.solve(dense, -s._capture["mass"])
    assert rel(pp - pp.mean(), g["pp"]) < 1e-9


def test_krylov_against_dense_solves_and_breakdown():
    rng = np.random.default_rng(3)
    n = 60
    A = sp.diags([-np.ones(n - 1), 2.5 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tocsr()
    N = A + 0.4 * sp.eye(n, k=1)
    b = rng.standard_normal(n)
    x, info = O.pcg(lambda v: A @ v, b, tol=1e-12, maxiter=200)
    assert info.converged and rel(x, np.linalg.solve(A.toarray(), b)) < 1e-10
    x, info = O.bicgstab(lambda v: N @ v, b, M=lambda v: v / 2.5, tol=1e-12, maxiter=200)
    assert info.converged and rel(x, np.linalg.solve(N.toarray(), b)) < 1e-10
    x, info = O.pcg(lambda v: np.diag([1.0, 2.0]) @ v, np.array([1.0, 0.0]), M=lambda r: np.array([-r[1], r[0]]))
    assert not info.converged and np.all(np.isfinite(x))


def test_amg_pcg_on_laplacian():
    n = 40
    T = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])
    A = (sp.kron(T, sp.identity(n)) + sp.kron(sp.identity(n), T)).tocsr()
    b = np.random.default_rng(0).standard_normal(n * n)
    M = O.SAAMG(A)
    assert M.n_levels >= 2
    x, info = O.pcg(lambda v: A @ v, b, M=M, tol=1e-10, maxiter=100)
    assert info.converged and info.iterations < 30


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("tri", [True, False])
def test_oracle_operators_match_reference_cpp_kernels(tri):
    lib = ctypes.CDLL(REF_LIB)
    P_ = ctypes.c_void_p
    fv = _mesh_fv(tri)
    rng = np.random.default_rng(0)
    phi, phib = rng.standard_normal(fv.nc), rng.standard_normal(fv.n_bf)
    g = rng.standard_normal((fv.nc, fv.dim))

    def p(a):
        return a.ctypes.data_as(P_)

    i32 = lambda a: np.ascontiguousarray(a, np.int32)  # noqa: E731
    i64 = lambda a: np.ascontiguousarray(a, np.int64)  # noqa: E731
    cd, cb = fv.Cdiff, fv.Cb
    iptr, iidx, isgn = i32(cd.indptr), i32(cd.indices), np.ascontiguousarray(cd.data, float)
    bptr, bidx = i32(cb.indptr), i32(cb.indices)
    P, N = i64(fv.P), i64(fv.N)
    out = np.empty((fv.n_if,))
    lib.ft_face_interp(ctypes.c_int64(fv.n_if), ctypes.c_int64(1), p(fv.w), p(P), p(N), p(phi), p(out))
    assert rel(out, fv.face_interp(phi)) == 0.0
    out = np.empty((fv.nc, fv.dim))
    lib.ft_ls_gradient(ctypes.c_int64(fv.nc), ctypes.c_int64(fv.dim), p(iptr), p(iidx), p(bptr), p(bidx), p(P), p(N),
                       p(fv.wl), p(np.ascontiguousarray(fv.d)), p(fv.wlb), p(np.ascontiguousarray(fv.db)),
                       p(np.ascontiguousarray(fv.Ginv)), p(phi), p(phib), p(out))
    assert rel(out, fv.gradient(phi, phib)) < 1e-13
    gg = np.empty((fv.nc, fv.dim))
    lib.ft_gg_gradient(ctypes.c_int64(fv.nc), ctypes.c_int64(fv.dim), p(iptr), p(iidx), p(isgn), p(bptr), p(bidx),
                       p(P), p(N), p(fv.w), p(np.ascontiguousarray(fv.S)), p(np.ascontiguousarray(fv.Sb)), p(fv.vol),
                       p(np.ascontiguousarray(fv.dxm)), p(phi), p(phib), p(g), p(gg))
    assert rel(gg, fv.gradient_gg(phi, phib, g)) < 1e-13
    d = 4 + rng.random(fv.nc)
    a1, a2 = -rng.random(fv.n_if), -rng.random(fv.n_if)
    data = np.empty(fv.nnz)
    lib.ft_assemble(ctypes.c_int64(fv.nnz), ctypes.c_int64(fv.nc), ctypes.c_int64(fv.n_if), p(i32(fv.slot_diag)),
                    p(i32(fv.slot_PN)), p(i32(fv.slot_NP)), p(d), p(a1), p(a2), p(data))
    assert rel(data, fv.matrix(d, a1, a2).data) == 0.0


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built (reference sources absent)")
def test_oracle_on_the_reference_kernels_follows_the_plain_oracle():
    """The CPU baseline's solver (the oracle with the reference's compiled C++/OpenMP kernels for
    the gradients, face interpolation, assembly, incidence products and BiCGSTAB,
    ``oracle/native_cpu.py``) follows the plain NumPy oracle on the heated duct: same inner
    iteration counts, fields within the kernels' summation-order round-off."""
    from gputest_amd.mesh import duct_mesh
    from oracle import native_cpu as N
    from tests.golden_cases import duct_case

    m = duct_mesh(40, 20)
    a = O.Simplec(m, duct_case(O, max_iter=10))
    b = N.with_reference_kernels(O.Simplec(m, duct_case(O, max_iter=10)))
    for _ in range(10):
        a.step()
        b.step()
        assert b.last_linear["momentum_iters"] == a.last_linear["momentum_iters"]
        assert b.last_linear["pressure_iters"] == a.last_linear["pressure_iters"]
    assert rel(b.U, a.U) < 1e-10 and rel(b.p, a.p) < 1e-10
    assert rel(b.T - 300.0, a.T - 300.0) < 1e-5  # the unconverged 60-step energy BiCGSTAB amplifies round-off


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.slow
def test_reference_energy_iterate_is_not_reproducible_across_its_own_paths():
    """The heated-duct temperature mid-run is not an iterate-level quantity even for the
    reference: its 60-step energy BiCGSTAB never meets its 1e-3 tolerance there, and the
    reference's own two CPU paths (NumPy, and its C++/OpenMP kernels with their own summation
    order) part by several percent in T - 300 after five outer iterations of the 250 x 125 duct
    while U and p agree to 1e-9. This bounds what any port can be held to on T mid-run; T is
    pinned by the captured iteration and at convergence instead."""
    from gputest_amd.mesh import duct_mesh
    from oracle import native_cpu as N
    from tests.golden_cases import duct_case

    m = duct_mesh(250, 125)
    a = O.Simplec(m, duct_case(O, max_iter=5))
    b = N.with_reference_kernels(O.Simplec(m, duct_case(O, max_iter=5)))
    for _ in range(5):
        a.step()
        b.step()
    assert rel(b.U, a.U) < 1e-9 and rel(b.p, a.p) < 1e-9
    dT = float(np.linalg.norm(b.T - a.T) / np.linalg.norm(a.T - 300.0))
    print("T - 300 relative difference between the reference's two CPU paths", dT)
    assert dT > 1e-2
