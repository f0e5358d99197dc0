"""The header-only C++ API (RXMeshStatic, Query<N>::dispatch<Op>, Iterator,
VertexAttribute, for_each_*) on the GPU, through tests/cpp/api_test.hip.

Query rows written by user lambdas through the API must equal the C-ABI's
CSR (itself pinned against the oracle in test_gpu_queries.py) bit for bit;
the reference-style vertex-normal kernel must match the FP64 oracle."""
import ctypes
import os

import numpy as np
import pytest

import rxmesh_amd as rx
from conftest import ROOT, obj
from oracle.oracle import OracleMesh
from rxmesh_amd import meshgen as mg

pytestmark = pytest.mark.gpu

API_LIB = os.path.join(ROOT, "tests", "cpp", "_build", "libapi_test.so")
OPS = {"VV": 3, "VE": 4, "VF": 5, "FV": 6, "FE": 7, "FF": 8, "EV": 9, "EF": 11, "EVD": 12}

_api = None


def api():
    global _api
    if _api is None:
        assert os.path.exists(API_LIB), "C++ API test library missing: run python tests/cpp/build.py"
        rx.lib()
        L = ctypes.CDLL(API_LIB)
        vp, u32, i, i64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_int64
        L.api_query.argtypes = [vp, u32, u32, i, i, i, vp, vp, i64]
        L.api_vertex_normals.argtypes = [vp, u32, vp, u32, u32, vp]
        L.api_misc.argtypes = [vp, u32, u32, ctypes.c_char_p, i]
        L.api_arap_weights.argtypes = [vp, u32, vp, u32, u32, vp]
        L.api_arap_weights_mat.argtypes = [vp, u32, vp, u32, u32, vp]
        _api = L
    return _api


def _case(name):
    if name.endswith(".obj"):
        return obj(name)
    if name == "grid":
        return mg.grid(24, seed=3)
    if name == "disk":
        return mg.disk(20, seed=4)
    if name == "ico":
        return mg.icosphere(3, seed=5, sigma=0.01)
    if name == "torus":
        return mg.torus(24, 16)
    raise KeyError(name)


CASES = [("plane.obj", 512), ("cube.obj", 512), ("sphere3.obj", 128), ("bunnyhead.obj", 256), ("grid", 64),
         ("disk", 96), ("ico", 128), ("torus", 64)]


def api_query(F, ps, op, oriented=False, active_mod=1, m=None):
    m = m or rx.Mesh(F, patch_size=ps)
    nsrc = m.num_src(op)
    row = np.zeros(nsrc + 1, np.int32)
    cap = max(m.query_nnz(op), 1)
    col = np.zeros(cap, np.int32)
    F = np.ascontiguousarray(F, dtype=np.uint32)
    rc = api().api_query(F.ctypes.data, len(F), ps, OPS[op], int(oriented), active_mod, row.ctypes.data,
                         col.ctypes.data, cap)
    assert rc == 0, rc
    return row, col[:row[-1]]


@pytest.mark.parametrize("op", list(OPS))
@pytest.mark.parametrize("case,ps", CASES)
def test_api_query_equals_capi(case, ps, op):
    V, F = _case(case)
    m = rx.Mesh(F, patch_size=ps)
    rp, col = m.query(op)
    row, c = api_query(F, ps, op, m=m)
    assert np.array_equal(row, rp.cpu().numpy())
    assert np.array_equal(c, col.cpu().numpy()[:len(c)])


@pytest.mark.parametrize("op", ["VV", "VE"])
@pytest.mark.parametrize("case,ps", [c for c in CASES if c[0] != "bunnyhead.obj"])
def test_api_oriented_query_equals_capi(case, ps, op):
    V, F = _case(case)
    m = rx.Mesh(F, patch_size=ps)
    rp, col = m.query(op, oriented=True)
    row, c = api_query(F, ps, op, oriented=True, m=m)
    assert np.array_equal(row, rp.cpu().numpy())
    assert np.array_equal(c, col.cpu().numpy()[:len(c)])


@pytest.mark.parametrize("op", ["VV", "VF", "FF", "EF", "EV"])
def test_api_active_set(op):
    """dispatch(..., active_set): only sources whose predicate holds get a
    row; their rows are unchanged."""
    V, F = _case("ico")
    m = rx.Mesh(F, patch_size=128)
    rp, col = (t.cpu().numpy() for t in m.query(op))
    row, c = api_query(F, 128, op, active_mod=3, m=m)
    n = len(rp) - 1
    sizes = np.diff(row)
    for s in range(n):
        if s % 3 == 0:
            assert sizes[s] == rp[s + 1] - rp[s]
            assert np.array_equal(c[row[s]:row[s + 1]], col[rp[s]:rp[s + 1]])
        else:
            assert sizes[s] == 0


@pytest.mark.parametrize("case,ps", [("sphere3.obj", 128), ("bunnyhead.obj", 256), ("grid", 64), ("torus", 64)])
def test_api_vertex_normal_kernel(case, ps):
    V, F = _case(case)
    V = np.ascontiguousarray(V, dtype=np.float32)
    F = np.ascontiguousarray(F, dtype=np.uint32)
    out = np.zeros_like(V)
    rc = api().api_vertex_normals(F.ctypes.data, len(F), V.ctypes.data, len(V), ps, out.ctypes.data)
    assert rc == 0
    ref = OracleMesh(F).vertex_normals(V.astype(np.float64), "max")
    # fp32 with atomics in any order vs FP64: 1e-5 of the largest normal
    err = np.abs(out - ref).max() / np.abs(ref).max()
    assert err < 1e-5, err


@pytest.mark.parametrize("case,ps", [("sphere3.obj", 128), ("disk", 96), ("torus", 64), ("ico", 128)])
def test_api_arap_weights_kernel(case, ps):
    """The DiffARAP edge-weight kernel, spelled as the reference app writes
    it over Query::dispatch<Op::EVDiamond>, equals rxm_arap_weights per edge
    (same fp32 formula; 1e-6 relative for contraction differences)."""
    import torch
    V, F = _case(case)
    V = np.ascontiguousarray(V, dtype=np.float32)
    F = np.ascontiguousarray(F, dtype=np.uint32)
    m = rx.Mesh(F, patch_size=ps)
    out = np.zeros(m.nE, np.float32)
    rc = api().api_arap_weights(F.ctypes.data, len(F), V.ctypes.data, len(V), ps, out.ctypes.data)
    assert rc == 0
    l2g = m.array("v_l2g")
    w = m.arap_weights(torch.from_numpy(np.ascontiguousarray(V[l2g].T)).cuda()).cpu().numpy()
    assert np.all(np.abs(out - w) <= 1e-6 * np.maximum(1.0, np.abs(w)))
    assert (w > 0).any()


@pytest.mark.parametrize("case,ps", [("sphere3.obj", 128), ("disk", 96), ("ico", 128)])
def test_api_sparse_matrix_arap_weights(case, ps):
    """The reference's calc_edge_weights_mat, unchanged, writing into the
    SparseMatrix mirror: both symmetric entries of every edge, read back
    through SparseMatrix::operator(), equal rxm_arap_weights."""
    import torch
    V, F = _case(case)
    V = np.ascontiguousarray(V, dtype=np.float32)
    F = np.ascontiguousarray(F, dtype=np.uint32)
    m = rx.Mesh(F, patch_size=ps)
    out = np.zeros((m.nE, 2), np.float32)
    rc = api().api_arap_weights_mat(F.ctypes.data, len(F), V.ctypes.data, len(V), ps, out.ctypes.data)
    assert rc == 0
    l2g = m.array("v_l2g")
    w = m.arap_weights(torch.from_numpy(np.ascontiguousarray(V[l2g].T)).cuda()).cpu().numpy()
    for k in range(2):
        assert np.all(np.abs(out[:, k] - w) <= 1e-6 * np.maximum(1.0, np.abs(w)))
    assert (w > 0).any()


@pytest.mark.parametrize("case,ps", [("sphere3.obj", 128), ("disk", 96), ("torus", 64)])
def test_api_host_side(case, ps):
    V, F = _case(case)
    F = np.ascontiguousarray(F, dtype=np.uint32)
    buf = ctypes.create_string_buffer(4096)
    fails = api().api_misc(F.ctypes.data, len(F), ps, buf, 4096)
    assert fails == 0, buf.value.decode()
