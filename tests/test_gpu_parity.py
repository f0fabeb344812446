"""Parity of the HIP path (through the C ABI) with the oracle and with the reference's own
golden vectors. Tolerances: bit-exact where the reference's tests demand it (face
interpolation, CSR assembly), 1e-13 for gradients, identical Krylov iteration counts,
1e-10 relative on one captured outer iteration, <= 1e-8 relative L2 after 300 iterations."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle.simplec_oracle as O
from gputest_amd import case as C
from gputest_amd.amg import SmoothedAggregationAMG
from gputest_amd.fvmesh import FVMesh
from gputest_amd.krylov import bicgstab_jacobi, bicgstab_ldu, pcg_csr
from gputest_amd.mesh import cartesian_mesh_2d, extrude_to_3d
from gputest_amd.simple import SimpleSolver
from tests.golden_cases import capture_case, cavity_case, channel_case, duct_case, load_pre_state

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
DEV = "cuda"


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def rel_l2(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def close(a, ref, rtol, scale=0.0):
    """max |a - ref| <= rtol * max(max |ref|, scale): relative, unless ``ref`` is round-off noise
    on a problem whose natural magnitude is ``scale``."""
    a, ref = np.asarray(a, float), np.asarray(ref, float)
    return float(np.abs(a - ref).max()) <= rtol * max(float(np.abs(ref).max()), scale)


def t(a):
    return torch.as_tensor(np.ascontiguousarray(a, np.float64), device=DEV)


def h(x):
    torch.cuda.synchronize()
    return x.cpu().numpy()


def _meshes():
    return {"tri2d": cartesian_mesh_2d(20, 16, triangulate=True, warp=0.05),
            "hex3d": extrude_to_3d(cartesian_mesh_2d(6, 5, warp=0.05), 4, 0.5)}


@pytest.mark.parametrize("name", ["tri2d", "hex3d"])
def test_dropin_operators_match_oracle(name):
    mesh = _meshes()[name]
    ref = O.FV(mesh)
    fv = FVMesh(mesh, DEV)
    rng = np.random.default_rng(0)
    phi, phib = rng.standard_normal(ref.nc), rng.standard_normal(ref.n_bf)
    g = rng.standard_normal((ref.nc, ref.dim))
    F, G = rng.standard_normal(ref.n_if), rng.standard_normal((ref.n_if, ref.dim))
    assert rel(h(fv.face_interp(t(phi))), ref.face_interp(phi)) == 0.0
    assert rel(h(fv.face_interp_vec(t(g))), ref.face_interp_vec(g)) == 0.0
    assert rel(h(fv.gradient(t(phi), t(phib))), ref.gradient(phi, phib)) < 1e-13
    assert rel(h(fv.gradient_gg(t(phi), t(phib))), ref.gradient_gg(phi, phib)) < 1e-13
    assert rel(h(fv.gradient_gg(t(phi), t(phib), t(g))), ref.gradient_gg(phi, phib, g)) < 1e-13
    assert rel(h(fv.net_outflow(t(F), t(phib))), ref.net_outflow(F, phib)) < 1e-15
    assert rel(h(fv.scatter(t(F), t(-F), t(phib))), ref.scatter(F, -F, phib)) < 1e-15
    assert rel(h(fv._inc()["own"] @ t(G)), ref.Cown @ G) == 0.0
    d = 4 + rng.random(ref.nc)
    a1, a2 = -rng.random(ref.n_if), -rng.random(ref.n_if)
    A = fv.matrix(t(d), t(a1), t(a2))
    assert rel(h(A.data), ref.matrix(d, a1, a2).data) == 0.0
    assert rel(h(A @ t(phi)), ref.matrix(d, a1, a2) @ phi) < 1e-15


@pytest.mark.parametrize("tol,maxiter", [(0.1, 30), (1e-10, 500), (1e-10, 3)])
def test_bicgstab_jacobi_follows_the_reference_algorithm(tol, maxiter):
    mesh = cartesian_mesh_2d(24, 24, triangulate=True, warp=0.05)
    ref = O.FV(mesh)
    fv = FVMesh(mesh, DEV)
    rng = np.random.default_rng(2)
    d = 4 + rng.random(ref.nc)
    a1, a2 = -rng.random(ref.n_if), -rng.random(ref.n_if)
    Ah = ref.matrix(d, a1, a2)
    b, x0, inv = rng.standard_normal(ref.nc), rng.standard_normal(ref.nc), 1.0 / d
    xr, ir = O.bicgstab(lambda v: Ah @ v, b, M=lambda v: inv * v, x0=x0, tol=tol, maxiter=maxiter,
                        relative_to_initial=True)
    A = fv.matrix(t(d), t(a1), t(a2))
    xn, inn = bicgstab_jacobi(A, t(b), t(x0), t(inv), tol, maxiter)
    assert (inn.iterations, inn.converged) == (ir.iterations, ir.converged)
    assert abs(inn.residual - ir.residual) <= 1e-10 * max(ir.residual, 1e-300)
    assert rel(h(xn), xr) < 1e-12
    # the same solve on the face (LDU) layout
    off = np.where(h(fv.cface) & 1, a2[h(fv.cface) >> 1], a1[h(fv.cface) >> 1])
    x = t(x0)
    il = bicgstab_ldu(fv, t(d), t(off), t(b), t(inv), x, tol, maxiter, True)
    assert (il.iterations, il.converged) == (ir.iterations, ir.converged)
    assert rel(h(x), xr) < 1e-12


def test_amg_hierarchy_vcycle_and_pcg():
    g = np.load(os.path.join(GOLD, "amg_tri40.npz"))
    A = sp.csr_matrix((g["A0_data"], g["A0_indices"], g["A0_indptr"]))
    M = SmoothedAggregationAMG(A, DEV, singular=True)
    R = O.SAAMG(A, singular=True)
    assert M.n_levels == R.n_levels == int(g["n_levels"])
    for k in range(len(M.aggregates)):
        assert np.array_equal(M.aggregates[k], g[f"L{k}_agg"])
        # the smoothed prolongator built on the device equals the reference's
        P = M.prolongator(k)
        Pr = sp.csr_matrix((g[f"L{k}_P_data"], g[f"L{k}_P_indices"], g[f"L{k}_P_indptr"]), shape=P.shape)
        Pr.sort_indices()  # the reference's P is not in canonical (sorted) order
        assert np.array_equal(P.indptr, Pr.indptr) and np.array_equal(P.indices, Pr.indices)
        assert rel(P.data, Pr.data) < 1e-13
        if f"L{k}_rho" in g.files:
            assert abs(float(h(M.rhos[k])[0]) - float(g[f"L{k}_rho"])) <= 1e-12 * float(g[f"L{k}_rho"])
        if f"L{k + 1}_A_data" in g.files and k + 1 < len(M.levels):
            Ac = M.levels[k + 1][0].to_scipy()
            Ar = sp.csr_matrix((g[f"L{k + 1}_A_data"], g[f"L{k + 1}_A_indices"], g[f"L{k + 1}_A_indptr"]),
                               shape=Ac.shape)
            Ar.sort_indices()
            assert np.array_equal(Ac.indptr, Ar.indptr) and np.array_equal(Ac.indices, Ar.indices)
            assert rel(Ac.data, Ar.data) < 1e-12
    assert rel(M.coarse_inv_h, g["coarse_inv"]) < 1e-8
    x = h(M(t(g["vcycle_b"])))
    assert rel(x, g["vcycle_x"]) < 1e-10
    assert rel(x, R(g["vcycle_b"])) < 1e-12
    # PCG with the device V-cycle reproduces the oracle's iteration count and solution
    from gputest_amd.fvmesh import DeviceCSR

    b = g["vcycle_b"]
    Ad = DeviceCSR(t(A.data), torch.as_tensor(A.indices.astype(np.int32), device=DEV),
                   torch.as_tensor(A.indptr.astype(np.int32), device=DEV), A.shape)
    xo, io = O.pcg(lambda v: A @ v, b, M=R, tol=1e-10, maxiter=200, project=lambda v: v - v.mean())
    xd, idd = pcg_csr(Ad, t(b), 1e-10, 200, amg=M, project=True)
    assert idd.iterations == io.iterations and idd.converged
    assert rel(h(xd), xo) < 1e-9


def test_pcg_amg_on_laplacian():
    n = 40
    T = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])
    A = (sp.kron(T, sp.identity(n)) + sp.kron(sp.identity(n), T)).tocsr()
    from gputest_amd.fvmesh import DeviceCSR

    Ad = DeviceCSR(t(A.data), torch.as_tensor(A.indices.astype(np.int32), device=DEV),
                   torch.as_tensor(A.indptr.astype(np.int32), device=DEV), A.shape)
    b = np.random.default_rng(0).standard_normal(n * n)
    M = SmoothedAggregationAMG(A, DEV)
    x, info = pcg_csr(Ad, t(b), 1e-10, 100, amg=M)
    assert info.converged and info.iterations < 30
    assert np.linalg.norm(A @ h(x) - b) / np.linalg.norm(b) < 1e-9


@pytest.mark.parametrize("tag", ["tri16", "hex3d", "periodic", "duct"])
def test_captured_iteration_matches_reference(tag):
    g = np.load(os.path.join(GOLD, f"capture_{tag}.npz"))
    mesh, case = capture_case(tag, C)
    s = SimpleSolver(mesh, case, DEV)
    assert s.relax_p == float(g["relax_p"])
    load_pre_state(s, g)
    s.step()
    fv = s.fv
    A = fv.ldu_to_csr_host(h(s.diag), h(s.off))
    Ar = fv.ldu_to_csr_host(h(s.diag_r), h(s.off))
    Ap = fv.ldu_to_csr_host(h(s.diag_p), h(s.offp))
    if not case.energy:  # with energy the off-diagonal buffer now holds the energy matrix
        assert rel(A.data, g["A_data"]) < 1e-14
        assert rel(Ar.data, g["Ar_data"]) < 1e-14
    assert rel(h(s.bmom), g["br"]) < 1e-13
    assert s.last_linear["momentum_iters"] == g["mom_iters"].tolist()
    assert rel(h(s.Ustar).T, g["U_star"]) < 1e-11
    assert rel(h(s.Fst), g["Fst"]) < 1e-11
    fscale = float(np.abs(g["Fst"]).max())  # the periodic channel's imbalance is round-off noise
    assert close(h(s.Fbst), g["Fbst"], 1e-11, fscale)
    assert close(h(s.mass), g["mass"], 1e-10, fscale)
    assert rel(Ap.data, g["Ap_data"]) < 1e-12
    pscale = case.density * float(np.abs(g["U_new"]).max()) ** 2
    assert close(h(s.pp), g["pp"], 1e-10, pscale)
    assert rel(h(s.F), g["F_new"]) < 1e-10
    assert rel(h(s.U).T, g["U_new"]) < 1e-10
    assert close(h(s.p), g["p_new"], 1e-10, pscale)
    if case.energy:
        assert s.last_linear["energy_iters"] == int(g["T_iters"])
        assert rel(h(s.T), g["T_new"]) < 1e-7
    for k, v in s.residuals()["scaled"].items():
        ref = float(g["res_scaled_" + k])
        assert abs(v - ref) <= 1e-8 * abs(ref) + 1e-11, k


def test_cavity_300_iterations_within_1e8_of_reference():
    g = np.load(os.path.join(GOLD, "cavity64_simplec300.npz"))
    s = SimpleSolver(cartesian_mesh_2d(64, 64), cavity_case(C), DEV)
    its, pits = [], []
    while not s.finished:
        s.step()
        its.append(s.last_linear["momentum_iters"])
        pits.append(s.last_linear["pressure_iters"])
    U = h(s.U).T
    p = h(s.p)
    err_u = rel_l2(U[:, 0], g["U"][:, 0])
    err_v = rel_l2(U[:, 1], g["U"][:, 1])
    err_p = rel_l2(p - p.mean(), g["p"] - g["p"].mean())
    print(f"cavity64 300 it: rel L2 u {err_u:.2e} v {err_v:.2e} p {err_p:.2e}")
    assert err_u <= 1e-8 and err_v <= 1e-8 and err_p <= 1e-8
    assert np.array_equal(its, g["mom_iters"])
    assert np.array_equal(pits, g["p_iters"])


def _errors(U, p, F, T, g, singular):
    """Errors against the reference trajectory: relative L2 per velocity component (absolute,
    scaled by max |U|, for a component that is round-off noise), pressure as max |dp| over
    max(max |p|, max |U|^2), flux and temperature as relative L2."""
    gU, gp = g["U"], g["p"]
    if singular:
        p, gp = p - p.mean(), gp - gp.mean()
    umax = float(np.abs(gU).max())
    out = {}
    for k in range(U.shape[1]):
        nrm = np.linalg.norm(gU[:, k])
        out[f"U{k}"] = rel_l2(U[:, k], gU[:, k]) if nrm > 1e-6 * np.linalg.norm(gU) else \
            float(np.abs(U[:, k] - gU[:, k]).max()) / umax
    out["p"] = float(np.abs(p - gp).max()) / max(float(np.abs(gp).max()), umax ** 2)
    out["F"] = rel_l2(F, g["F"])
    if T is not None:
        out["T"] = rel_l2(T, g["T"])
    return out


def _oracle_ulp_sensitivity(mesh, case, g):
    """The oracle's own drift from the reference trajectory when every face flux moves by one
    ulp after the first iteration: the round-off amplification of the algorithm on this case."""
    s = O.Simplec(mesh, case)
    s.step()
    s.F = np.nextafter(s.F, np.inf)
    while not s.finished:
        s.step()
    return _errors(s.U, s.p, s.F, s.T if s.energy else None, g, s.singular_p)


@pytest.mark.parametrize("which", ["duct", "channel"])
def test_short_runs_match_reference(which):
    """20 iterations of reduced-size workloads at the default pressure tolerance (0.1) against
    the reference's trajectory: 1e-8, or 10x the oracle's own one-ulp drift where the algorithm
    amplifies round-off beyond that (the noise-driven pressure coupling of the periodic
    channel). Temperature is not iterate-level stable: its 60-step Jacobi-BiCGSTAB does not
    converge and amplifies round-off by orders of magnitude (the reference measured 1e-15 ->
    1e-2 for its own native port); it is pinned by the captured iteration (1e-7, same count)
    and by test_energy_converges_to_the_oracle_solution (1e-6), and bounded here by 1e-3."""
    if which == "duct":
        from gputest_amd.mesh import duct_mesh

        g = np.load(os.path.join(GOLD, "short_duct40x20.npz"))
        mesh, case_c, case_o = duct_mesh(40, 20), duct_case(C, max_iter=20), duct_case(O, max_iter=20)
    else:
        from gputest_amd.mesh import channel_mesh

        g = np.load(os.path.join(GOLD, "short_channel16x8x8.npz"))
        mesh, case_c, case_o = channel_mesh(16, 8, 8), channel_case(C, max_iter=20), channel_case(O, max_iter=20)
    s = SimpleSolver(mesh, case_c, DEV)
    assert s.relax_p == float(g["relax_p"])
    while not s.finished:
        s.step()
    errs = _errors(h(s.U).T, h(s.p), h(s.F), h(s.T) if s.energy else None, g, s.singular_p)
    sens = _oracle_ulp_sensitivity(mesh, case_o, g)
    print(which, "errors", errs, "oracle one-ulp drift", sens)
    for k, e in errs.items():
        bound = 1e-3 if k == "T" else max(1e-8, 10.0 * sens[k])
        assert e <= bound, (k, e, bound)


def test_energy_converges_to_the_oracle_solution():
    """Converged fields with energy (the reference's own comparison for its native port,
    tests/test_native.py:90-103): 16x16 skewed triangles, heated lid, 1e-6."""
    mesh = cartesian_mesh_2d(16, 16, triangulate=True, warp=0.05)

    def mk(mod):
        B = mod.Boundary
        return mod.Case(boundaries={"top": B("wall", velocity=[1.0, 0.0], thermal="temperature", value=1.0),
                                    "bottom": B("wall", thermal="temperature", value=0.0)},
                        viscosity=0.05, energy=True, conductivity=0.05, max_iter=4000, tol=1e-9)

    ref = O.Simplec(mesh, mk(O))
    while not ref.finished:
        ref.step()
    s = SimpleSolver(mesh, mk(C), DEV)
    while not s.finished:
        s.step()
    assert ref.converged and s.converged
    f = s.fields()
    for k, r in (("x-velocity", ref.U[:, 0]), ("y-velocity", ref.U[:, 1]), ("pressure", ref.p),
                 ("temperature", ref.T)):
        e = rel(f[k], r)
        print(k, e, ref.iteration, s.iteration)
        assert e < 1e-6, (k, e)


def test_krylov_at_sizes_that_fill_the_reduction_grid():
    """360,000 unknowns: reductions run on the capped workgroup grid and every workgroup of a
    consuming kernel re-adds the partials (the configuration of the million-cell workload)."""
    n1 = 600
    T1 = sp.diags([-np.ones(n1 - 1), 2 * np.ones(n1), -np.ones(n1 - 1)], [-1, 0, 1])
    A = (sp.kron(T1, sp.identity(n1)) + sp.kron(sp.identity(n1), T1)).tocsr()
    from gputest_amd.fvmesh import DeviceCSR

    def dev(M):
        M = sp.csr_matrix(M)
        M.sort_indices()
        return DeviceCSR(t(M.data), torch.as_tensor(M.indices.astype(np.int32), device=DEV),
                         torch.as_tensor(M.indptr.astype(np.int32), device=DEV), M.shape)

    rng = np.random.default_rng(4)
    b = rng.standard_normal(A.shape[0])
    R = O.SAAMG(A)
    xo, io = O.pcg(lambda v: A @ v, b, M=R, tol=1e-8, maxiter=200)
    M = SmoothedAggregationAMG(A, DEV)
    xd, idd = pcg_csr(dev(A), t(b), 1e-8, 200, amg=M)
    print("pcg iterations", io.iterations, idd.iterations)
    assert idd.converged and idd.iterations == io.iterations
    assert rel(h(xd), xo) < 1e-6
    N = (A + sp.diags(rng.uniform(0, 0.5, A.shape[0])) + 0.3 * sp.eye(A.shape[0], k=1)).tocsr()
    inv = 1.0 / N.diagonal()
    x0 = rng.standard_normal(A.shape[0])
    xr, ir = O.bicgstab(lambda v: N @ v, b, M=lambda v: inv * v, x0=x0, tol=1e-6, maxiter=100,
                        relative_to_initial=True)
    xn, inn = bicgstab_jacobi(dev(N), t(b), t(x0), t(inv), 1e-6, 100)
    print("bicgstab iterations", ir.iterations, inn.iterations)
    assert (inn.iterations, inn.converged) == (ir.iterations, ir.converged)
    assert rel(h(xn), xr) < 1e-8
