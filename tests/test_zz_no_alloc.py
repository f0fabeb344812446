"""The allocation-free BiCGSTAB entry point against the one it mirrors, through the C ABI.

``ftmi_bicgstab_jacobi_noalloc`` must give what ``ftmi_bicgstab_jacobi`` gives with a caller
workspace: the same x bit for bit, the same iteration count, residual and status. Two calls in
a row share the library's device array, the second on a smaller system, so whatever the first
left in the array must not leak into the second. A system larger than the array is refused on
the host before any device work.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from gputest_amd import _native as nat

pytestmark = pytest.mark.gpu


def convection_diffusion(m, peclet, seed):
    """Upwind convection-diffusion on an m x m grid: nonsymmetric, diagonally dominant."""
    rng = np.random.default_rng(seed)
    n = m * m
    i = np.arange(n)
    x, y = i % m, i // m
    rows, cols, vals = [i], [i], [4.5 + peclet * 2.0 + 0.01 * rng.random(n)]
    for dx, dy, c in ((1, 0, -1.0), (-1, 0, -1.0 - peclet), (0, 1, -1.0), (0, -1, -1.0 - peclet)):
        ok = (x + dx >= 0) & (x + dx < m) & (y + dy >= 0) & (y + dy < m)
        rows.append(i[ok])
        cols.append((i + dx + dy * m)[ok])
        vals.append(np.full(ok.sum(), c))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    A.sort_indices()
    return A, rng.standard_normal(n), 0.1 * rng.standard_normal(n)


def device_system(A, b, x0):
    dev = "cuda"
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)  # noqa: E731
    return (t(A.indptr, torch.int32), t(A.indices, torch.int32), t(A.data, torch.float64), t(b, torch.float64),
            t(1.0 / A.diagonal(), torch.float64), t(x0, torch.float64))


def solve(A, b, x0, tol, maxiter, noalloc):
    L = nat.lib()
    ip, ix, dv, bd, inv, x = device_system(A, b, x0)
    n = A.shape[0]
    it, res = ctypes.c_int64(-7), ctypes.c_double(-7.0)
    ptrs = [nat.ptr(ip), nat.ptr(ix), nat.ptr(dv), nat.ptr(bd), nat.ptr(inv), nat.ptr(x)]
    if noalloc:
        st = L.ftmi_bicgstab_jacobi_noalloc(n, *ptrs, float(tol), int(maxiter), 1, ctypes.byref(it),
                                            ctypes.byref(res), nat.stream_ptr())
    else:
        work = torch.zeros(int(L.ftmi_bicgstab_work_bytes(n)), dtype=torch.uint8, device="cuda")
        st = L.ftmi_bicgstab_jacobi(n, *ptrs, float(tol), int(maxiter), 1, nat.ptr(work), ctypes.byref(it),
                                    ctypes.byref(res), nat.stream_ptr())
    torch.cuda.synchronize()
    assert st >= 0, nat.lib().ftmi_last_error()
    return st, int(it.value), float(res.value), x.cpu().numpy()


@pytest.mark.parametrize("tol,maxiter", [(1e-8, 300), (0.1, 30)])
def test_noalloc_equals_the_workspace_entry_point_on_two_calls(tol, maxiter):
    L = nat.lib()
    big = int(L.ftmi_bicgstab_noalloc_max_rows())
    assert big == 1 << 18
    m1 = int(np.sqrt(big))  # 512 x 512 = the array's full size
    cases = [convection_diffusion(m1, 0.5, 1), convection_diffusion(40, 2.0, 2)]
    assert cases[0][0].shape[0] == big and cases[1][0].shape[0] < big
    for A, b, x0 in cases:  # two calls in a row, the second smaller
        print("noalloc case n =", A.shape[0], "tol", tol, flush=True)
        ref = solve(A, b, x0, tol, maxiter, noalloc=False)
        got = solve(A, b, x0, tol, maxiter, noalloc=True)
        assert got[0] == ref[0] and got[1] == ref[1], (got[:3], ref[:3])
        assert got[2] == ref[2] or (np.isnan(got[2]) and np.isnan(ref[2]))
        np.testing.assert_array_equal(got[3], ref[3])
        if tol < 1e-3:
            r = b - A @ got[3]
            assert ref[0] == 1 and np.linalg.norm(r) <= 1e-7 * np.linalg.norm(b - A @ x0)


def test_noalloc_refuses_a_system_larger_than_its_array():
    L = nat.lib()
    n = int(L.ftmi_bicgstab_noalloc_max_rows()) + 1
    it, res = ctypes.c_int64(0), ctypes.c_double(0.0)
    st = L.ftmi_bicgstab_jacobi_noalloc(n, None, None, None, None, None, None, 1e-8, 10, 1, ctypes.byref(it),
                                        ctypes.byref(res), nat.stream_ptr())
    assert st == -1 and b"max_rows" in L.ftmi_last_error()
