"""rxmesh_amd: MI355X-native patch-based static mesh query engine.

Python host side over the C-ABI (include/rxmesh_amd.h). PyTorch provides
device memory and streams; all compute runs in the HIP kernels of
librxmesh_amd.so. There is no CPU fallback: if the library is missing the
import of the native handle fails loudly.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "librxmesh_amd.so")

OP = {"VV": 3, "VE": 4, "VF": 5, "FV": 6, "FE": 7, "FF": 8, "EV": 9, "EF": 11, "EVD": 12}
STATUS = {0: "ok", 1: "invalid argument", 2: "non-manifold mesh", 3: "patch too big for LDS/16-bit ids",
          4: "I/O error", 5: "device error", 6: "unsupported", 7: "out of memory",
          8: "beyond the fixed scratch of an allocation-free entry point"}
ARR = {"ev": 0, "face_patch": 1, "vertex_patch": 2, "edge_patch": 3, "patches_val": 4,
       "patches_offset": 5, "ribbon_val": 6, "ribbon_offset": 7, "v_l2g": 8, "e_l2g": 9,
       "f_l2g": 10, "v_g2l": 11, "e_g2l": 12, "f_g2l": 13, "patch_counts": 14,
       "patch_offsets": 15, "ltog_v": 16, "ltog_e": 17, "ltog_f": 18, "ev_local": 19,
       "fe_local": 20, "vv2_roff": 21, "vv2_nbr": 22, "vv2_xlin": 23, "vv2_offsets": 24}
RXM_HOST_ONLY = 1
RXM_PATCHER_KMEANS = 2
RXM_BUILD_SERIAL = 4


class RXMError(RuntimeError):
    def __init__(self, code, what):
        super().__init__("%s failed: %s (status %d)" % (what, STATUS.get(code, "?"), code))
        self.code = code


class MeshInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in [
        "num_vertices", "num_edges", "num_faces", "num_patches", "patch_size",
        "max_vertices_per_patch", "max_edges_per_patch", "max_faces_per_patch",
        "max_owned_vertices", "max_owned_edges", "max_owned_faces", "is_closed",
        "is_edge_manifold", "max_valence", "max_edge_incident_faces",
        "max_face_adjacent_faces", "num_ribbon_faces", "on_device"]] + [("device_bytes", ctypes.c_uint64),
                                                                          ("is_consistently_oriented", ctypes.c_uint32)]


_lib = None

# exported symbols of include/rxmesh_amd.h with their argtypes
_vp, _i, _u32, _f = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_float
_SIGS = {
    "rxm_mesh_create": [_vp, _u32, _u32, _u32, ctypes.c_char_p, _u32, ctypes.POINTER(_vp)],
    "rxm_mesh_create_with_patches": [_vp, _u32, _u32, _vp, _u32, _u32, ctypes.POINTER(_vp)],
    "rxm_mesh_destroy": [_vp],
    "rxm_gaussian_curvature": [_vp, _vp, _vp, _vp, _vp],
    "rxm_arap_weights": [_vp, _vp, _vp, _vp],
    "rxm_build_edges_device": [_vp, _u32, _u32, _vp, _vp, _vp, _vp],
    "rxm_arap_rotations": [_vp, _vp, _vp, _vp, _vp, _vp],
    "rxm_arap_energy_grad": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "rxm_mcf_implicit_cg": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _f, _f, _vp, _vp, _vp, _vp],
    "rxm_mcf_matfree_cg": [_vp, _vp, _vp, _vp, _f, _i, _i, _i, _f, _f, _vp, _vp, _vp, _vp],
    "rxm_mesh_info": [_vp, ctypes.POINTER(MeshInfo)],
    "rxm_mesh_array_size": [_vp, _i, ctypes.POINTER(ctypes.c_uint64)],
    "rxm_mesh_get_array": [_vp, _i, _vp, ctypes.c_uint64],
    "rxm_mesh_save_patches": [_vp, ctypes.c_char_p],
    "rxm_query_nnz": [_vp, _i, ctypes.POINTER(ctypes.c_int64)],
    "rxm_query_csr": [_vp, _i, _i, _vp, _vp, _vp],
    "rxm_query_vv_vf": [_vp, _vp, _vp, _vp, _vp, _vp],
    "rxm_vertex_normals": [_vp, _i, _vp, _vp, _vp],
    "rxm_laplacian_smooth": [_vp, _vp, _vp, _vp, _f, _i, _vp],
    "rxm_to_linear": [_vp, _i, _i, _vp, _vp, _vp],
    "rxm_laplacian_smooth_csr": [_vp, _vp, _vp, _vp, _vp, _vp, _f, _i, _vp],
    "rxm_mesh_device_view": [_vp, _vp, ctypes.c_uint64],
    "rxm_vv_plan_build": [_vp, _vp],
    "rxm_set_launch_threads": [_vp, _i, _i],
    "rxm_mesh_create_batch": [_vp, _vp, _vp, _u32, _u32, _u32, ctypes.POINTER(_vp)],
    "rxm_batch_count": [_vp, ctypes.POINTER(_u32)],
    "rxm_mesh_batch_from": [_vp, _u32, _u32, ctypes.POINTER(_vp)],
    "rxm_batch_offsets": [_vp, _i, _vp, _u32],
    "rxm_symdir3_energy": [_vp, _vp, _vp, _vp, _vp],
    "rxm_symdir3_grad_step": [_vp, _vp, _vp, _f, _vp, _vp],
    "rxm_symdir3_grad": [_vp, _vp, _vp, _vp, _vp],
    "rxm_smooth_symdir_steps": [_vp, _vp, _vp, _vp, _f, _f, _i, _vp],
    "rxm_smooth_symdir_steps_each": [_vp, _vp, _vp, _vp, _u32, _f, _f, _i, _vp],
    "rxm_query_smooth_symdir_steps": [_vp, _vp, _vp, _vp, _f, _f, _i, _vp, _vp, _vp, _vp, _vp],
    "rxm_query_smooth_symdir_steps_each": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _f, _f, _i, _vp],
    "rxm_graph_capture_smooth_symdir": [_vp, _vp, _vp, _vp, _f, _f, _i, ctypes.POINTER(_vp)],
    "rxm_graph_launch": [_vp, _vp],
    "rxm_graph_destroy": [_vp],
    "rxm_cot_nnz": [_vp, ctypes.POINTER(ctypes.c_int64)],
    "rxm_cot_csr_assemble": [_vp, _vp, _f, _vp, _vp, _vp, _vp, _vp],
    "rxm_spmv3": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "rxm_mcf_explicit": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp],
    "rxm_symdir_rest": [_vp, _vp, _vp, _vp],
    "rxm_symdir_energy": [_vp, _vp, _vp, _vp, _vp, _vp],
    "rxm_symdir_energy_noalloc": [_vp, _vp, _vp, _vp, _vp, _vp],
    "rxm_symdir_grad": [_vp, _vp, _vp, _vp, _vp, _vp],
    "rxm_to_global": [_vp, _i, _i, _vp, _vp, _vp],
}


def _init_torch_runtime_first():
    """torch carries its own copy of the HIP runtime next to the system one
    this library links. With both in one process, torch finds no device if
    the system runtime initialised first; torch first, both work. So when a
    device is present, torch's runtime is initialised before the library is
    loaded (device_count() itself initialises nothing)."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.device_count() > 0:
        torch.cuda.init()


def lib():
    """Load librxmesh_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("librxmesh_amd.so not built: run `python -m rxmesh_amd.build` "
                              "(no CPU fallback exists)")
        _init_torch_runtime_first()
        L = ctypes.CDLL(LIB_PATH)
        for name, argt in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = argt
            fn.restype = ctypes.c_int
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise RXMError(rc, what)


def _stream(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if hasattr(stream, "cuda_stream"):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(stream)


def _dev(t, name, shape=None, dtype="float32", min_numel=None):
    """ctypes pointer to a tensor the C-ABI reads or writes, after checking
    what the C-ABI cannot see through a raw pointer: the dtype, the exact
    shape (or, for buffers sized by a count, a minimum element count),
    contiguity and that the tensor lives on the GPU. Raises ValueError before
    anything is launched, so a transposed view, a (V, 3) array where (3, V)
    SoA is expected, an undersized buffer or a float16 tensor never reaches a
    kernel."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError("%s: expected a torch tensor, got %s" % (name, type(t).__name__))
    want = {"float32": torch.float32, "int32": torch.int32}[dtype]
    if t.dtype != want:
        raise ValueError("%s: expected dtype %s, got %s" % (name, want, t.dtype))
    if shape is not None and tuple(t.shape) != tuple(int(n) for n in shape):
        raise ValueError("%s: expected shape %s, got %s" % (name, tuple(int(n) for n in shape), tuple(t.shape)))
    if min_numel is not None and (t.dim() != 1 or t.numel() < min_numel):
        raise ValueError("%s: expected a 1-D tensor of at least %d elements, got shape %s"
                         % (name, min_numel, tuple(t.shape)))
    if not t.is_contiguous():
        raise ValueError("%s: tensor must be contiguous (call .contiguous())" % name)
    if not t.is_cuda:
        raise ValueError("%s: tensor must be on the GPU" % name)
    return ctypes.c_void_p(t.data_ptr())


def _opt(t, *args, **kw):
    """_dev for an optional tensor (None passes a null pointer)."""
    return None if t is None else _dev(t, *args, **kw)


def build_edges_device(fv, num_vertices, stream=None):
    """Edges of a triangle list on the device, numbered as the host build
    numbers them (first appearance over the half-edges). fv: (F, 3) int32
    device tensor. Returns (ev (E, 2) int32, fe (3F,) int32)."""
    import torch
    if not isinstance(fv, torch.Tensor) or fv.dim() != 2 or fv.shape[1] != 3:
        raise ValueError("fv: expected an (F, 3) tensor")
    nF = fv.shape[0]
    pf = _dev(fv, "fv", (nF, 3), "int32")
    ev = torch.empty(2 * 3 * nF, dtype=torch.int32, device=fv.device)
    fe = torch.empty(3 * nF, dtype=torch.int32, device=fv.device)
    ne = ctypes.c_uint32()
    _check(lib().rxm_build_edges_device(pf, nF, int(num_vertices), ctypes.c_void_p(ev.data_ptr()),
                                        ctypes.c_void_p(fe.data_ptr()), ctypes.byref(ne), _stream(stream)),
           "rxm_build_edges_device")
    return ev[: 2 * ne.value].view(-1, 2), fe


class Mesh:
    """One static patch mesh (host build + device slabs).

    Parameters mirror RXMeshStatic(fv, patcher_file, patch_size)
    (include/rxmesh/rxmesh_static.h:61-100)."""

    def __init__(self, fv, num_vertices=0, patch_size=512, patch_file=None, face_patch=None,
                 host_only=False, batch=None, parts=None, patcher="bisect", serial_build=False):
        """fv: F x 3 triangles. Alternatives: batch = list of fv arrays (their
        disjoint union, built here), parts = list of built Mesh objects
        (their union, reusing the host build)."""
        h = ctypes.c_void_p()
        if patcher not in ("kmeans", "bisect"):
            raise ValueError("patcher must be 'kmeans' or 'bisect'")
        flags = ((RXM_HOST_ONLY if host_only else 0) | (RXM_PATCHER_KMEANS if patcher == "kmeans" else 0)
                 | (RXM_BUILD_SERIAL if serial_build else 0))
        self._fv = None
        if parts is not None:
            self._parts = list(parts)
            arr = (ctypes.c_void_p * len(parts))(*[q._h.value for q in parts])
            rc = lib().rxm_mesh_batch_from(arr, len(parts), flags, ctypes.byref(h))
        elif batch is not None:
            fvs = [np.ascontiguousarray(f, dtype=np.uint32).reshape(-1, 3) for f in batch]
            self._fv = fvs
            ptrs = (ctypes.c_void_p * len(fvs))(*[f.ctypes.data for f in fvs])
            nfs = np.array([f.shape[0] for f in fvs], np.uint32)
            nvs = np.array([int(f.max()) + 1 for f in fvs], np.uint32)
            rc = lib().rxm_mesh_create_batch(ptrs, nfs.ctypes.data, nvs.ctypes.data, len(fvs), patch_size,
                                             flags, ctypes.byref(h))
        else:
            fv = np.ascontiguousarray(fv, dtype=np.uint32).reshape(-1, 3)
            self._fv = fv
            if face_patch is not None:
                fp = np.ascontiguousarray(face_patch, dtype=np.uint32)
                rc = lib().rxm_mesh_create_with_patches(fv.ctypes.data, fv.shape[0], num_vertices, fp.ctypes.data,
                                                        int(fp.max()) + 1, flags, ctypes.byref(h))
            else:
                pf = patch_file.encode() if patch_file else None
                rc = lib().rxm_mesh_create(fv.ctypes.data, fv.shape[0], num_vertices, patch_size, pf, flags,
                                           ctypes.byref(h))
        _check(rc, "rxm_mesh_create")
        self._h = h
        self.info = self._info()
        self.nV = self.info["num_vertices"]
        self.nE = self.info["num_edges"]
        self.nF = self.info["num_faces"]
        self.P = self.info["num_patches"]

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _check(lib().rxm_mesh_destroy(self._h), "rxm_mesh_destroy")
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _info(self):
        i = MeshInfo()
        _check(lib().rxm_mesh_info(self._h, ctypes.byref(i)), "rxm_mesh_info")
        return {n: getattr(i, n) for n, _ in MeshInfo._fields_}

    def array(self, name):
        which = ARR[name]
        n = ctypes.c_uint64()
        _check(lib().rxm_mesh_array_size(self._h, which, ctypes.byref(n)), "rxm_mesh_array_size")
        dt = np.uint16 if name in ("fe_local", "vv2_roff", "vv2_nbr") else np.uint32
        out = np.zeros(max(n.value, 1), dt)
        _check(lib().rxm_mesh_get_array(self._h, which, out.ctypes.data, out.nbytes), "rxm_mesh_get_array")
        return out[: n.value]

    def save_patches(self, path):
        _check(lib().rxm_mesh_save_patches(self._h, path.encode()), "rxm_mesh_save_patches")

    # -- argument shapes ------------------------------------------------
    def _v3(self, t, name):
        """(3, V) float32 SoA positions / vertex vectors over linear ids."""
        return _dev(t, name, (3, self.nV))

    def _mr(self, Mr):
        """(4, F) float32 rest shapes (rxm_symdir_rest)."""
        return _dev(Mr, "Mr", (4, self.nF))

    # -- queries --------------------------------------------------------
    def query_nnz(self, op):
        n = ctypes.c_int64()
        _check(lib().rxm_query_nnz(self._h, OP[op], ctypes.byref(n)), "rxm_query_nnz")
        return n.value

    def query_vv_vf(self, vv=None, vf=None, stream=None):
        """VV and VF CSRs (same as query("VV") and query("VF")) from one launch."""
        import torch
        nvv, nvf = self.query_nnz("VV"), self.query_nnz("VF")
        if vv is None:
            vv = (torch.empty(self.nV + 1, dtype=torch.int32, device="cuda"),
                  torch.empty(max(nvv, 1), dtype=torch.int32, device="cuda"))
        if vf is None:
            vf = (torch.empty(self.nV + 1, dtype=torch.int32, device="cuda"),
                  torch.empty(max(nvf, 1), dtype=torch.int32, device="cuda"))
        args = (_dev(vv[0], "vv row_ptr", (self.nV + 1,), "int32"), _dev(vv[1], "vv col", None, "int32", nvv),
                _dev(vf[0], "vf row_ptr", (self.nV + 1,), "int32"), _dev(vf[1], "vf col", None, "int32", nvf))
        _check(lib().rxm_query_vv_vf(self._h, *args, _stream(stream)), "rxm_query_vv_vf")
        return vv, vf

    def num_src(self, op):
        return {"V": self.nV, "E": self.nE, "F": self.nF}[op[0]]

    def query(self, op, oriented=False, row_ptr=None, col=None, stream=None):
        """CSR (row_ptr, col) of `op` over linear ids, as int32 device tensors."""
        import torch
        nnz = self.query_nnz(op)
        if row_ptr is None:
            row_ptr = torch.empty(self.num_src(op) + 1, dtype=torch.int32, device="cuda")
        if col is None:
            col = torch.empty(max(nnz, 1), dtype=torch.int32, device="cuda")
        prp = _dev(row_ptr, "row_ptr", (self.num_src(op) + 1,), "int32")
        pcol = _dev(col, "col", None, "int32", nnz)
        _check(lib().rxm_query_csr(self._h, OP[op], int(oriented), prp, pcol, _stream(stream)), "rxm_query_csr")
        return row_ptr, col

    # -- geometry -------------------------------------------------------
    def vertex_normals(self, pos, mode="max", out=None, stream=None):
        import torch
        if out is None:
            out = torch.empty_like(pos)
        _check(lib().rxm_vertex_normals(self._h, 0 if mode == "max" else 1, self._v3(pos, "pos"),
                                        self._v3(out, "out"), _stream(stream)), "rxm_vertex_normals")
        return out

    def gaussian_curvature(self, pos, out=None, stream=None):
        """(gc, mixed area) per vertex in linear order, from (3, V) positions."""
        import torch
        if out is None:
            out = (torch.empty(self.nV, dtype=torch.float32, device=pos.device),
                   torch.empty(self.nV, dtype=torch.float32, device=pos.device))
        gc, am = out
        _check(lib().rxm_gaussian_curvature(self._h, self._v3(pos, "pos"), _dev(gc, "gc", (self.nV,)),
                                            _opt(am, "area_mixed", (self.nV,)), _stream(stream)),
               "rxm_gaussian_curvature")
        return gc, am

    # -- ARAP-style deformation ---------------------------------------------
    def arap_weights(self, pos, out=None, stream=None):
        """Cotangent edge weights (E,) in linear edge order, 0 on boundary edges."""
        import torch
        if out is None:
            out = torch.empty(self.nE, dtype=torch.float32, device=pos.device)
        _check(lib().rxm_arap_weights(self._h, self._v3(pos, "pos"), _dev(out, "w", (self.nE,)), _stream(stream)),
               "rxm_arap_weights")
        return out

    def arap_rotations(self, rest, x, w, out=None, stream=None):
        """Best rotation per vertex as (V, 9): row v is R_v, row-major."""
        import torch
        if out is None:
            out = torch.empty(self.nV, 9, dtype=torch.float32, device=x.device)
        _check(lib().rxm_arap_rotations(self._h, self._v3(rest, "rest"), self._v3(x, "x"),
                                        _dev(w, "w", (self.nE,)), _dev(out, "R", (self.nV, 9)), _stream(stream)),
               "rxm_arap_rotations")
        return out

    def arap_energy_grad(self, rest, x, w, R, total=None, grad=None, stream=None):
        """(energy as a 1-element device tensor, dE/dX as (3, V)), rotations fixed."""
        import torch
        if total is None:
            total = torch.empty(1, dtype=torch.float32, device=x.device)
        if grad is None:
            grad = torch.empty_like(x)
        _check(lib().rxm_arap_energy_grad(self._h, self._v3(rest, "rest"), self._v3(x, "x"),
                                          _dev(w, "w", (self.nE,)), _dev(R, "R", (self.nV, 9)),
                                          _dev(total.view(-1), "total", None, "float32", 1), self._v3(grad, "grad"),
                                          _stream(stream)), "rxm_arap_energy_grad")
        return total, grad

    def smooth(self, pos, lam=0.5, iters=1, out=None, tmp=None, stream=None):
        import torch
        if out is None:
            out = torch.empty_like(pos)
        if tmp is None and iters > 1:
            tmp = torch.empty_like(pos)
        _check(lib().rxm_laplacian_smooth(self._h, self._v3(pos, "pos"), self._v3(out, "out"),
                                          None if tmp is None else self._v3(tmp, "tmp"), float(lam), int(iters),
                                          _stream(stream)), "rxm_laplacian_smooth")
        return out

    def build_vv_plan(self, stream=None):
        _check(lib().rxm_vv_plan_build(self._h, _stream(stream)), "rxm_vv_plan_build")

    def set_launch_threads(self, kernel, threads):
        _check(lib().rxm_set_launch_threads(self._h, kernel, threads), "rxm_set_launch_threads")

    def smooth_csr(self, vv, pos, lam=0.5, iters=1, out=None, tmp=None, stream=None):
        """Smoothing on a materialised VV query vv = (row_ptr, col)."""
        import torch
        if out is None:
            out = torch.empty_like(pos)
        if tmp is None and iters > 1:
            tmp = torch.empty_like(pos)
        _check(lib().rxm_laplacian_smooth_csr(self._h, _dev(vv[0], "vv row_ptr", (self.nV + 1,), "int32"),
                                              _dev(vv[1], "vv col", None, "int32", self.query_nnz("VV")),
                                              self._v3(pos, "pos"), self._v3(out, "out"),
                                              None if tmp is None else self._v3(tmp, "tmp"), float(lam), int(iters),
                                              _stream(stream)), "rxm_laplacian_smooth_csr")
        return out

    # -- batches ----------------------------------------------------------
    def batch_count(self):
        n = ctypes.c_uint32()
        _check(lib().rxm_batch_count(self._h, ctypes.byref(n)), "rxm_batch_count")
        return n.value

    def batch_offsets(self, kind=0):
        k = self.batch_count()
        out = np.zeros(k + 1, np.uint32)
        _check(lib().rxm_batch_offsets(self._h, {"V": 0, "E": 1, "F": 2, "P": 3}.get(kind, kind),
                                       out.ctypes.data, k + 1), "rxm_batch_offsets")
        return out

    def symdir3_energy(self, Mr, x, out=None, stream=None):
        import torch
        if out is None:
            out = torch.empty(self.nF, dtype=torch.float32, device="cuda")
        _check(lib().rxm_symdir3_energy(self._h, self._mr(Mr), self._v3(x, "x"), _dev(out, "energy", (self.nF,)),
                                        _stream(stream)), "rxm_symdir3_energy")
        return out

    def symdir3_grad_step(self, Mr, x, eta, out=None, stream=None):
        import torch
        if out is None:
            out = torch.empty_like(x)
        _check(lib().rxm_symdir3_grad_step(self._h, self._mr(Mr), self._v3(x, "x"), float(eta),
                                           self._v3(out, "out"), _stream(stream)), "rxm_symdir3_grad_step")
        return out

    def symdir3_grad(self, Mr, x, out=None, stream=None):
        """dE/dx (3, V) of the 3-D symmetric Dirichlet energy (rxm_symdir3_grad)."""
        import torch
        if out is None:
            out = torch.empty_like(x)
        _check(lib().rxm_symdir3_grad(self._h, self._mr(Mr), self._v3(x, "x"), self._v3(out, "grad"),
                                      _stream(stream)), "rxm_symdir3_grad")
        return out

    def smooth_symdir_steps(self, Mr, x, tmp, lam, eta, steps, stream=None):
        _check(lib().rxm_smooth_symdir_steps(self._h, self._mr(Mr), self._v3(x, "x"), self._v3(tmp, "tmp"),
                                             float(lam), float(eta), int(steps), _stream(stream)),
               "rxm_smooth_symdir_steps")
        return x

    def vv_vf_buffers(self):
        """Device buffers ((vv_row_ptr, vv_col), (vf_row_ptr, vf_col)) sized for query_vv_vf."""
        import torch
        i32 = dict(dtype=torch.int32, device="cuda")
        return ((torch.empty(self.nV + 1, **i32), torch.empty(max(self.query_nnz("VV"), 1), **i32)),
                (torch.empty(self.nV + 1, **i32), torch.empty(max(self.query_nnz("VF"), 1), **i32)))

    def _vv_vf_ptrs(self, vv, vf):
        return (_dev(vv[0], "vv_row_ptr", (self.nV + 1,), "int32"),
                _dev(vv[1], "vv_col", None, "int32", self.query_nnz("VV")),
                _dev(vf[0], "vf_row_ptr", (self.nV + 1,), "int32"),
                _dev(vf[1], "vf_col", None, "int32", self.query_nnz("VF")))

    def query_smooth_symdir_steps(self, Mr, x, tmp, lam, eta, steps, vv, vf, stream=None):
        """smooth_symdir_steps with the VV and VF CSRs (into vv, vf) queried at every step."""
        _check(lib().rxm_query_smooth_symdir_steps(self._h, self._mr(Mr), self._v3(x, "x"), self._v3(tmp, "tmp"),
                                                   float(lam), float(eta), int(steps), *self._vv_vf_ptrs(vv, vf),
                                                   _stream(stream)), "rxm_query_smooth_symdir_steps")
        return x

    def capture_smooth_symdir(self, Mr, x, tmp, lam, eta, steps):
        return Graph(self, Mr, x, tmp, lam, eta, steps)

    # -- cotangent system / explicit MCF ------------------------------
    def cot_nnz(self):
        n = ctypes.c_int64()
        _check(lib().rxm_cot_nnz(self._h, ctypes.byref(n)), "rxm_cot_nnz")
        return n.value

    def _system(self, A):
        """Pointers of a cotangent system (row_ptr, col, val, mass) as
        rxm_cot_csr_assemble lays it out, every array checked against the
        mesh's sizes (the patch-local kernels read val by the mesh's own row
        layout)."""
        rp, col, val, mass = A[:4]
        nnz = self.cot_nnz()
        pv, pm = _dev(val, "val", (nnz,)), _dev(mass, "mass", (self.nV,))  # the arrays the kernels read
        return _dev(rp, "row_ptr", (self.nV + 1,), "int32"), _dev(col, "col", (nnz,), "int32"), pv, pm

    def cot_assemble(self, pos, dt, out=None, stream=None):
        """(row_ptr, col, val, mass) of A = M + dt L on device; `out` reuses
        the four tensors of an earlier call."""
        import torch
        if out is None:
            nnz = self.cot_nnz()
            out = (torch.empty(self.nV + 1, dtype=torch.int32, device="cuda"),
                   torch.empty(nnz, dtype=torch.int32, device="cuda"),
                   torch.empty(nnz, dtype=torch.float32, device="cuda"),
                   torch.empty(self.nV, dtype=torch.float32, device="cuda"))
        _check(lib().rxm_cot_csr_assemble(self._h, self._v3(pos, "pos"), float(dt), *self._system(out),
                                          _stream(stream)), "rxm_cot_csr_assemble")
        return tuple(out)

    def spmv3(self, A, x, out=None, stream=None):
        """y = A x over the mesh's V rows for a CSR A = (row_ptr, col, val, ...)."""
        import torch
        rp, col, val = A[:3]
        if out is None:
            out = torch.empty_like(x)
        pc = _dev(col, "col", (col.numel(),), "int32")
        pv = _dev(val, "val", (col.numel(),))
        _check(lib().rxm_spmv3(self._h, _dev(rp, "row_ptr", (self.nV + 1,), "int32"), pc, pv, self._v3(x, "x"),
                               self._v3(out, "out"), _stream(stream)), "rxm_spmv3")
        return out

    def mcf_implicit_cg(self, A, x, max_iter=100, abs_tol=1e-6, rel_tol=0.0, out=None, stats=True, stream=None):
        """Implicit MCF step: CG on A X = M x from X = x. Returns (X, iterations,
        start |r|^2, final |r|^2) -- the last three None when stats is False
        (no host synchronisation)."""
        import torch
        if out is None:
            out = torch.empty_like(x)
        it, r0, r1 = ctypes.c_int(), ctypes.c_float(), ctypes.c_float()
        refs = (ctypes.byref(it), ctypes.byref(r0), ctypes.byref(r1)) if stats else (None, None, None)
        _check(lib().rxm_mcf_implicit_cg(self._h, *self._system(A), self._v3(x, "x"), self._v3(out, "out"),
                                         int(max_iter), float(abs_tol), float(rel_tol), *refs, _stream(stream)),
               "rxm_mcf_implicit_cg")
        if not stats:
            return out, None, None, None
        return out, it.value, r0.value, r1.value

    def mcf_matfree_cg(self, coords, x, dt, uniform=False, precond=False, max_iter=100, abs_tol=1e-6, rel_tol=0.0,
                       out=None, stats=True, stream=None):
        """Implicit MCF step with the system never assembled (weights recomputed
        from `coords` in every product): CG, or Jacobi-preconditioned CG, on
        A X = M x from X = x. Returns (X, iterations, start residual, final
        residual) -- the last three None when stats is False."""
        import torch
        if out is None:
            out = torch.empty_like(x)
        it, r0, r1 = ctypes.c_int(), ctypes.c_float(), ctypes.c_float()
        refs = (ctypes.byref(it), ctypes.byref(r0), ctypes.byref(r1)) if stats else (None, None, None)
        _check(lib().rxm_mcf_matfree_cg(self._h, self._v3(coords, "coords"), self._v3(x, "x"), self._v3(out, "out"),
                                        float(dt), int(bool(uniform)), int(bool(precond)), int(max_iter),
                                        float(abs_tol), float(rel_tol), *refs, _stream(stream)),
               "rxm_mcf_matfree_cg")
        if not stats:
            return out, None, None, None
        return out, it.value, r0.value, r1.value

    def mcf_explicit(self, A, x, iters, out=None, tmp=None, stream=None):
        import torch
        if out is None:
            out = torch.empty_like(x)
        if tmp is None and iters > 1:
            tmp = torch.empty_like(x)
        _check(lib().rxm_mcf_explicit(self._h, *self._system(A), self._v3(x, "x"), self._v3(out, "out"),
                                      None if tmp is None else self._v3(tmp, "tmp"), int(iters), _stream(stream)),
               "rxm_mcf_explicit")
        return out

    # -- symmetric Dirichlet --------------------------------------------
    def symdir_rest(self, rest_pos, stream=None):
        import torch
        pr = self._v3(rest_pos, "rest_pos")
        Mr = torch.empty(4, self.nF, dtype=torch.float32, device="cuda")
        _check(lib().rxm_symdir_rest(self._h, pr, self._mr(Mr), _stream(stream)), "rxm_symdir_rest")
        return Mr

    def symdir_energy(self, Mr, uv, energy=None, total=None, stream=None, noalloc=False):
        """(total, per-face energies). noalloc=True calls the allocation-free
        entry point (rxm_symdir_energy_noalloc): same kernels and results, its
        partial sums in the library's fixed per-device scratch."""
        import torch
        if energy is None:
            energy = torch.empty(self.nF, dtype=torch.float32, device="cuda")
        if total is None:
            total = torch.empty((), dtype=torch.float32, device="cuda")
        name = "rxm_symdir_energy_noalloc" if noalloc else "rxm_symdir_energy"
        _check(getattr(lib(), name)(self._h, self._mr(Mr), _dev(uv, "uv", (2, self.nV)),
                                    _dev(energy, "energy", (self.nF,)),
                                    _dev(total.view(-1), "total", None, "float32", 1), _stream(stream)), name)
        return total, energy

    def symdir_grad(self, Mr, uv, scale=None, out=None, stream=None):
        import torch
        if out is None:
            out = torch.empty_like(uv)
        ps = None if scale is None else _dev(scale.view(-1), "scale", None, "float32", 1)
        _check(lib().rxm_symdir_grad(self._h, self._mr(Mr), _dev(uv, "uv", (2, self.nV)), ps,
                                     _dev(out, "grad", (2, self.nV)), _stream(stream)), "rxm_symdir_grad")
        return out

    def _perm_args(self, x, out, kind):
        n = (self.nV, self.nE, self.nF)[kind]
        shape = (n,) if x.dim() == 1 else (x.shape[0], n)
        return x.shape[0] if x.dim() == 2 else 1, _dev(x, "x", shape), _dev(out, "out", shape)

    def to_linear(self, x, kind=0, out=None, stream=None):
        """x: (ncomp, n) or (n,) float32 device tensor in global order -> linear order."""
        import torch
        if out is None:
            out = torch.empty_like(x)
        ncomp, px, po = self._perm_args(x, out, kind)
        _check(lib().rxm_to_linear(self._h, kind, ncomp, px, po, _stream(stream)), "rxm_to_linear")
        return out

    def to_global(self, x, kind=0, out=None, stream=None):
        import torch
        if out is None:
            out = torch.empty_like(x)
        ncomp, px, po = self._perm_args(x, out, kind)
        _check(lib().rxm_to_global(self._h, kind, ncomp, px, po, _stream(stream)), "rxm_to_global")
        return out


class Graph:
    """A hipGraph of the smoothing + symmetric-Dirichlet step loop
    (rxm_graph_capture_smooth_symdir); the tensors are baked in."""

    def __init__(self, mesh, Mr, x, tmp, lam, eta, steps):
        self._keep = (mesh, Mr, x, tmp)
        g = ctypes.c_void_p()
        _check(lib().rxm_graph_capture_smooth_symdir(mesh._h, mesh._mr(Mr), mesh._v3(x, "x"), mesh._v3(tmp, "tmp"),
                                                     float(lam), float(eta), int(steps), ctypes.byref(g)),
               "rxm_graph_capture_smooth_symdir")
        self._g = g

    def launch(self, stream=None):
        _check(lib().rxm_graph_launch(self._g, _stream(stream)), "rxm_graph_launch")

    def close(self):
        if getattr(self, "_g", None) is not None and self._g.value:
            _check(lib().rxm_graph_destroy(self._g), "rxm_graph_destroy")
        self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def smooth_symdir_steps_each(meshes, Mrs, xs, tmps, lam, eta, steps, stream=None):
    """Plain per-mesh launches of the step loop over separate meshes."""
    n = len(meshes)
    if not (len(Mrs) == len(xs) == len(tmps) == n):
        raise ValueError("smooth_symdir_steps_each: one Mr, x and tmp per mesh")
    arr = lambda items: (ctypes.c_void_p * n)(*[p.value for p in items])  # noqa: E731
    _check(lib().rxm_smooth_symdir_steps_each((ctypes.c_void_p * n)(*[m._h.value for m in meshes]),
                                              arr([m._mr(t) for m, t in zip(meshes, Mrs)]),
                                              arr([m._v3(t, "x") for m, t in zip(meshes, xs)]),
                                              arr([m._v3(t, "tmp") for m, t in zip(meshes, tmps)]),
                                              n, float(lam), float(eta), int(steps), _stream(stream)),
           "rxm_smooth_symdir_steps_each")


def query_smooth_symdir_steps_each(meshes, Mrs, xs, tmps, bufs, lam, eta, steps, stream=None):
    """Plain per-mesh launches of the step loop with the per-step VV + VF
    query; bufs[k] = (vv, vf) CSR buffers of mesh k (Mesh.vv_vf_buffers)."""
    n = len(meshes)
    if not (len(Mrs) == len(xs) == len(tmps) == len(bufs) == n):
        raise ValueError("query_smooth_symdir_steps_each: one Mr, x, tmp and buffer set per mesh")
    arr = lambda items: (ctypes.c_void_p * n)(*[p.value for p in items])  # noqa: E731
    ptrs = [m._vv_vf_ptrs(*b) for m, b in zip(meshes, bufs)]
    _check(lib().rxm_query_smooth_symdir_steps_each((ctypes.c_void_p * n)(*[m._h.value for m in meshes]),
                                                    arr([m._mr(t) for m, t in zip(meshes, Mrs)]),
                                                    arr([m._v3(t, "x") for m, t in zip(meshes, xs)]),
                                                    arr([m._v3(t, "tmp") for m, t in zip(meshes, tmps)]),
                                                    *[arr([p[j] for p in ptrs]) for j in range(4)],
                                                    n, float(lam), float(eta), int(steps), _stream(stream)),
           "rxm_query_smooth_symdir_steps_each")


class SymmetricDirichlet:
    """torch.autograd.Function wrapper: energy = sum_f E_f(uv) over a mesh.

    forward runs rxm_symdir_energy; backward runs the hand-written adjoint
    rxm_symdir_grad (per-face Jacobians are recomputed, not stored).
    uv: (2, V) float32 SoA over linear vertex ids."""

    _fn = None

    @classmethod
    def apply(cls, mesh, Mr, uv):
        if cls._fn is None:
            import torch

            class _SymDir(torch.autograd.Function):
                @staticmethod
                def forward(ctx, uv_, mesh_, Mr_):
                    uv_ = uv_.detach().contiguous()
                    Mr_ = Mr_.contiguous()
                    total, _ = mesh_.symdir_energy(Mr_, uv_)
                    ctx.mesh = mesh_
                    ctx.Mr = Mr_
                    ctx.save_for_backward(uv_)
                    return total

                @staticmethod
                def backward(ctx, g):
                    (uv_,) = ctx.saved_tensors
                    g = g.detach().to(torch.float32).contiguous()
                    return ctx.mesh.symdir_grad(ctx.Mr, uv_, scale=g), None, None

            cls._fn = _SymDir
        return cls._fn.apply(uv, mesh, Mr)


def symmetric_dirichlet(mesh, Mr, uv):
    """Differentiable sum of symmetric Dirichlet energies (see SymmetricDirichlet)."""
    return SymmetricDirichlet.apply(mesh, Mr, uv)
