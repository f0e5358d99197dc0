"""The reference's own app kernels, compiled unchanged from the reference tree
against this library's C++ API mirror (oracle/_ref/libref_apps.so, built by
`make -C oracle ref` where the reference tree is present), run on the GPU and
compared with this library's C-ABI kernels and the FP64 oracle:

  - compute_vertex_normal (apps/VertexNormal/vertex_normal_kernel.cuh:10-42)
    against rxm_vertex_normals (Max weights): within 1e-5 per vertex,
    relative to the largest normal (the reference kernel adds with float
    atomics, so its sums come in any order);
  - mcf_A_setup / mcf_B_setup (apps/MCF/mcf_kernels.cuh:302-426) into the
    SparseMatrix / DenseMatrix mirrors: the cotangent system's pattern equals
    rxm_cot_csr_assemble's entry for entry and its values, and B = M x,
    agree within 1e-5 of the largest entry; the uniform-weight system agrees
    with the FP64 oracle's within 1e-5;
  - matvec (:114-213), the matrix-free product rxm_mcf_matfree_cg iterates
    with: A x within 1e-5 of the FP64 oracle's product, cotangent and
    uniform weights.
Skipped when the library was not built (no reference tree where the build
ran)."""
import ctypes
import os

import numpy as np
import pytest

from oracle.oracle import OracleMesh
from rxmesh_amd import meshgen as mg

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libref_apps.so")
PATCH = 256


def _lib():
    if not os.path.exists(LIB):
        pytest.skip("oracle/_ref/libref_apps.so not built (needs the reference tree at build time)")
    import rxmesh_amd as rx
    rx.lib()  # torch's HIP runtime first, then the library the reference kernels link
    L = ctypes.CDLL(LIB)
    u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
    i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
    u, i, f = ctypes.c_uint32, ctypes.c_int, ctypes.c_float
    L.ref_vertex_normals.argtypes = [u32p, u, f32p, u, u, f32p]
    L.ref_mcf_system.argtypes = [u32p, u, f32p, u, u, f, i, i32p, i32p, f32p, ctypes.c_int64,
                                 ctypes.POINTER(ctypes.c_int64), f32p]
    L.ref_mcf_matvec.argtypes = [u32p, u, f32p, u, u, f, i, f32p, f32p]
    return L


def _mesh(name):
    if name == "ico4":
        return mg.icosphere(4, seed=1, sigma=0.01)
    if name == "ico_perm":
        return mg.permute_mesh(*mg.icosphere(4, seed=2, sigma=0.005), seed=2)
    if name == "torus":
        return mg.torus(48, 24)
    return mg.disk(40, seed=3)


@pytest.mark.parametrize("name", ["ico4", "ico_perm", "torus", "disk"])
def test_reference_vertex_normal_kernel_matches(name):
    import torch
    import rxmesh_amd as rx
    L = _lib()
    V, F = _mesh(name)
    V32 = np.ascontiguousarray(V, np.float32)
    Fu = np.ascontiguousarray(F, np.uint32)
    ref = np.zeros_like(V32)
    assert L.ref_vertex_normals(Fu, len(Fu), V32, len(V32), PATCH, ref) == 0
    m = rx.Mesh(Fu, patch_size=PATCH)
    l2g = m.array("v_l2g")
    x = torch.from_numpy(np.ascontiguousarray(V32[l2g].T)).cuda()
    n = m.vertex_normals(x, "max")
    torch.cuda.synchronize()
    got = np.empty_like(V32)
    got[l2g] = n.cpu().numpy().T
    scale = np.linalg.norm(ref, axis=1).max()
    assert np.linalg.norm(got - ref, axis=1).max() <= 1e-5 * scale


@pytest.mark.parametrize("name", ["ico4", "ico_perm", "torus", "disk"])
@pytest.mark.parametrize("dt", [1e-2, 10.0])
def test_reference_mcf_setup_matches_cot_assembly(name, dt):
    import torch
    import rxmesh_amd as rx
    L = _lib()
    V, F = _mesh(name)
    V32 = np.ascontiguousarray(V, np.float32)
    Fu = np.ascontiguousarray(F, np.uint32)
    m = rx.Mesh(Fu, patch_size=PATCH)
    cap = m.cot_nnz()
    rows, cols = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    vals, B = np.zeros(cap, np.float32), np.zeros_like(V32)
    nnz = ctypes.c_int64()
    assert L.ref_mcf_system(Fu, len(Fu), V32, len(V32), PATCH, dt, 0, rows, cols, vals, cap, ctypes.byref(nnz), B) == 0
    assert nnz.value == cap
    l2g = m.array("v_l2g")
    x = torch.from_numpy(np.ascontiguousarray(V32[l2g].T)).cuda()
    rp, col, val, mass = (t.cpu().numpy() for t in m.cot_assemble(x, dt))
    # the library's rows in global ids, entry for entry in the reference's row layout
    grow = l2g[np.repeat(np.arange(len(rp) - 1), np.diff(rp))]
    gcol = l2g[col]
    order = np.lexsort((np.arange(len(grow)), grow))
    rorder = np.lexsort((np.arange(len(rows)), rows))
    assert np.array_equal(grow[order], rows[rorder]) and np.array_equal(gcol[order], cols[rorder])
    assert np.abs(val[order] - vals[rorder]).max() <= 1e-5 * np.abs(vals).max()
    gmass = np.empty(len(V32), np.float32)
    gmass[l2g] = mass
    Mx = gmass[:, None].astype(np.float64) * V32
    assert np.abs(B - Mx).max() <= 1e-5 * np.abs(Mx).max()


@pytest.mark.parametrize("name", ["ico4", "disk"])
def test_reference_mcf_setup_uniform_matches_oracle(name):
    L = _lib()
    V, F = _mesh(name)
    V32 = np.ascontiguousarray(V, np.float32)
    Fu = np.ascontiguousarray(F, np.uint32)
    o = OracleMesh(Fu).set_patches(np.zeros(len(Fu), np.uint32))
    orow, ocol, oval, omass = o.uniform_assemble(0.5)
    cap = len(oval)
    rows, cols = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    vals, B = np.zeros(cap, np.float32), np.zeros_like(V32)
    nnz = ctypes.c_int64()
    assert L.ref_mcf_system(Fu, len(Fu), V32, len(V32), PATCH, 0.5, 1, rows, cols, vals, cap, ctypes.byref(nnz), B) == 0
    import scipy.sparse as sp
    n = len(V32)
    A = sp.coo_matrix((vals.astype(np.float64), (rows, cols)), shape=(n, n)).toarray()
    Ao = sp.coo_matrix((oval, (orow, ocol)), shape=(n, n)).toarray()
    assert np.abs(A - Ao).max() <= 1e-5 * np.abs(Ao).max()
    Mx = omass[:, None] * V32.astype(np.float64)
    assert np.abs(B - Mx).max() <= 1e-5 * np.abs(Mx).max()


@pytest.mark.parametrize("name", ["ico4", "torus", "disk"])
@pytest.mark.parametrize("uniform", [0, 1])
def test_reference_matvec_matches_oracle_product(name, uniform):
    L = _lib()
    V, F = _mesh(name)
    V32 = np.ascontiguousarray(V, np.float32)
    Fu = np.ascontiguousarray(F, np.uint32)
    rng = np.random.default_rng(4)
    xin = rng.standard_normal(V32.shape).astype(np.float32)
    out = np.zeros_like(V32)
    dt = 0.25
    assert L.ref_mcf_matvec(Fu, len(Fu), V32, len(V32), PATCH, dt, uniform, xin, out) == 0
    o = OracleMesh(Fu).set_patches(np.zeros(len(Fu), np.uint32))
    orow, ocol, oval, omass = (o.uniform_assemble(dt) if uniform else o.cot_assemble(V32.astype(np.float64), dt))
    import scipy.sparse as sp
    n = len(V32)
    A = sp.coo_matrix((oval, (orow, ocol)), shape=(n, n)).tocsr()
    ref = A @ xin.astype(np.float64)
    scale = (abs(A) @ np.abs(xin.astype(np.float64))).max()
    assert np.abs(out - ref).max() <= 1e-5 * scale
