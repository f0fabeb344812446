"""CPU-side checks of the C ABI: the library loads and exports every function that
include/ftmi.h declares (no device work is launched)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ftmi.h")
LIB = os.path.join(ROOT, "gputest_amd", "libftmi.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ftmi_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="libftmi.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def declared_prototypes():
    """{name: [parameter types]} from include/ftmi.h."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"\b(ftmi_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        params = [p.strip() for p in m.group(2).split(",") if p.strip() and p.strip() != "void"]
        out[m.group(1)] = params
    return out


def _ctype_of(param):
    import ctypes as ct

    if "*" in param:
        return ct.c_void_p
    base = param.rsplit(" ", 1)[0].replace("const", "").strip()
    return {"int": ct.c_int, "int64_t": ct.c_int64, "double": ct.c_double}[base]


def test_binding_matches_header_prototypes():
    """Every ctypes signature has the header's parameter count and kinds (pointer / int /
    int64 / double), so a binding mistake fails here instead of on the GPU."""
    from gputest_amd import _native as nat

    protos = declared_prototypes()
    assert set(protos) == set(nat._SIGS)
    for name, params in protos.items():
        args, _ = nat._SIGS[name]
        assert len(args) == len(params), (name, len(args), len(params))
        for a, p in zip(args, params):
            assert a == _ctype_of(p), (name, p, a)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libftmi.so not built")
def test_abi_version_and_binding():
    from gputest_amd import _native as nat

    assert nat.lib().ftmi_abi_version() == nat.ABI_VERSION
    assert set(nat._SIGS) == set(declared_functions())


def test_host_aggregation_matches_oracle():
    """The product's host aggregation (C++) equals the oracle's restatement on a real strength graph."""
    if not os.path.exists(LIB):
        pytest.skip("libftmi.so not built")
    import numpy as np
    import scipy.sparse as sp

    import oracle.simplec_oracle as O
    from gputest_amd import amg

    g = np.load(os.path.join(ROOT, "tests", "golden", "amg_tri40.npz"))
    A = sp.csr_matrix((g["A0_data"], g["A0_indices"], g["A0_indptr"]))
    S = amg.strength(A, 0.08)
    agg, na = amg.aggregate(S)
    assert np.array_equal(agg, g["L0_agg"]) and na == int(g["L0_na"])
    agg2, na2 = O.aggregate(S.indptr, S.indices, S.shape[0])
    assert np.array_equal(agg, agg2) and na == na2
    # the forms the device set-up hands over: the pattern with a strength mask per entry (the
    # finest level) and the strong entries as a CSR of their own (coarse levels)
    from gputest_amd import _native as nat

    A.sort_indices()
    n = A.shape[0]
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    dg = np.abs(A.diagonal())
    mask = ((A.indices != rows) & (np.abs(A.data) >= 0.08 * np.sqrt(dg[rows] * dg[A.indices]))).astype(np.uint8)
    ptr, col = np.ascontiguousarray(A.indptr, np.int32), np.ascontiguousarray(A.indices, np.int32)
    sptr, scol = np.ascontiguousarray(S.indptr, np.int32), np.ascontiguousarray(S.indices, np.int32)
    L = nat.lib()
    for call in (lambda a, z: L.ftmi_amg_aggregate_mask_host(n, ptr.ctypes.data, col.ctypes.data, mask.ctypes.data,
                                                            a.ctypes.data, z.ctypes.data),
                 lambda a, z: L.ftmi_amg_aggregate_csr_host(n, sptr.ctypes.data, scol.ctypes.data, a.ctypes.data,
                                                           z.ctypes.data)):
        a32, size = np.empty(n, np.int32), np.empty(n, np.int32)
        assert call(a32, size) == na
        assert np.array_equal(a32, agg)
        assert np.array_equal(size[:na], np.bincount(agg, minlength=na))


def test_assemble_plan_covers_every_slot():
    if not os.path.exists(LIB):
        pytest.skip("libftmi.so not built")
    import numpy as np

    from gputest_amd import _native as nat
    import oracle.simplec_oracle as O
    from gputest_amd.mesh import cartesian_mesh_2d

    fv = O.FV(cartesian_mesh_2d(7, 5, triangulate=True))
    nc, nf = fv.nc, fv.n_if
    ip = np.ascontiguousarray(fv.indptr, np.int32)
    sPN, sNP = np.ascontiguousarray(fv.slot_PN, np.int32), np.ascontiguousarray(fv.slot_NP, np.int32)
    src_ptr, src = np.empty(nc + 1, np.int32), np.empty(4 * nf, np.int32)
    assert nat.lib().ftmi_assemble_plan_host(nc, nf, ip.ctypes.data, sPN.ctypes.data, sNP.ctypes.data,
                                             src_ptr.ctypes.data, src.ctypes.data) == 0
    # host replay of the plan reproduces the reference assembly exactly
    rng = np.random.default_rng(0)
    d, a1, a2 = 4 + rng.random(nc), -rng.random(nf), -rng.random(nf)
    data = np.zeros(fv.nnz)
    data[fv.slot_diag] = d
    for c in range(nc):
        for e in range(src_ptr[c], src_ptr[c + 1], 2):
            code, slot = src[e], src[e + 1]
            assert ip[c] <= slot < ip[c + 1]
            data[slot] += a2[code >> 1] if code & 1 else a1[code >> 1]
    assert np.array_equal(data, fv.matrix(d, a1, a2).data)
