"""The Python host side refuses tensors the C-ABI cannot use -- wrong dtype,
wrong shape (a (V, 3) array where (3, V) SoA is expected), a transposed
(non-contiguous) view, an undersized output buffer, host memory -- with a
ValueError raised before anything reaches a kernel. Runs on the CPU: the
mesh is built host-only and every call here is refused before a launch
(arguments are checked in order, so each case puts the bad tensor first)."""
import numpy as np
import pytest
import torch

import rxmesh_amd as rx
from rxmesh_amd import meshgen as mg


@pytest.fixture(scope="module")
def mesh():
    pos, fv = mg.icosphere(2, seed=0)
    return rx.Mesh(fv, patch_size=64, host_only=True)


def _v3(m, dtype=torch.float32):
    return torch.zeros(3, m.nV, dtype=dtype)


def test_checker_rules():
    t = torch.zeros(3, 10)
    with pytest.raises(ValueError, match="dtype"):
        rx._dev(t.half(), "x", (3, 10))
    with pytest.raises(ValueError, match="shape"):
        rx._dev(torch.zeros(10, 3), "x", (3, 10))
    with pytest.raises(ValueError, match="contiguous"):
        rx._dev(torch.zeros(10, 3).T, "x", (3, 10))
    with pytest.raises(ValueError, match="at least 11"):
        rx._dev(torch.zeros(10, dtype=torch.int32), "col", None, "int32", 11)
    with pytest.raises(ValueError, match="1-D"):
        rx._dev(torch.zeros(2, 6, dtype=torch.int32), "col", None, "int32", 11)
    with pytest.raises(ValueError, match="GPU"):
        rx._dev(t, "x", (3, 10))
    with pytest.raises(ValueError, match="torch tensor"):
        rx._dev(np.zeros((3, 10), np.float32), "x", (3, 10))
    assert rx._opt(None, "x", (3, 10)) is None


def test_wrong_dtype_refused(mesh):
    m = mesh
    with pytest.raises(ValueError, match="dtype"):
        m.vertex_normals(_v3(m, torch.float16))
    with pytest.raises(ValueError, match="dtype"):
        m.smooth(_v3(m, torch.float64))
    with pytest.raises(ValueError, match="dtype"):
        m.query("VV", row_ptr=torch.zeros(m.nV + 1, dtype=torch.int64), col=torch.zeros(1, dtype=torch.int32))


def test_aos_positions_refused(mesh):
    m = mesh
    with pytest.raises(ValueError, match="shape"):
        m.vertex_normals(torch.zeros(m.nV, 3))
    with pytest.raises(ValueError, match="shape"):
        m.gaussian_curvature(torch.zeros(m.nV, 3))
    with pytest.raises(ValueError, match="shape"):
        m.arap_weights(torch.zeros(m.nV, 3), out=torch.zeros(m.nE))


def test_transposed_view_refused(mesh):
    m = mesh
    with pytest.raises(ValueError, match="contiguous"):
        m.vertex_normals(torch.zeros(m.nV, 3).T)  # (3, V) shape over (V, 3) memory
    with pytest.raises(ValueError, match="contiguous"):
        m.cot_assemble(torch.zeros(m.nV, 3).T, 1e-3, out=(None, None, None, None))


def test_undersized_buffers_refused(mesh):
    m = mesh
    with pytest.raises(ValueError, match="shape"):
        m.query("VV", row_ptr=torch.zeros(m.nV, dtype=torch.int32), col=torch.zeros(1, dtype=torch.int32))
    with pytest.raises(ValueError, match="shape"):
        m.symdir3_energy(torch.zeros(4, m.nF - 1), _v3(m), out=torch.zeros(m.nF))
    with pytest.raises(ValueError, match="shape"):
        m.symdir_rest(torch.zeros(3, m.nV - 1))


def test_cotangent_system_checked_against_mesh(mesh):
    m = mesh
    nnz = m.cot_nnz()
    A = (torch.zeros(m.nV + 1, dtype=torch.int32), torch.zeros(nnz, dtype=torch.int32), torch.zeros(nnz - 7),
         torch.zeros(m.nV))
    with pytest.raises(ValueError, match="val"):
        m.mcf_explicit(A, _v3(m), 1, out=_v3(m))
    with pytest.raises(ValueError, match="val"):
        m.mcf_implicit_cg(A, _v3(m), out=_v3(m))


def test_host_memory_refused(mesh):
    m = mesh
    with pytest.raises(ValueError, match="GPU"):
        m.vertex_normals(_v3(m), out=_v3(m))
    with pytest.raises(ValueError, match="GPU"):
        m.symdir3_grad(torch.zeros(4, m.nF), _v3(m), out=_v3(m))
    with pytest.raises(ValueError, match="GPU"):
        rx.build_edges_device(torch.zeros(5, 3, dtype=torch.int32), 3)
    with pytest.raises(ValueError, match=r"\(F, 3\)"):
        rx.build_edges_device(torch.zeros(15, dtype=torch.int32), 3)
