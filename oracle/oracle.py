"""ctypes wrapper of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the cpu_baseline leg of ``bench.py`` -- never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

OP = {"V": 0, "E": 1, "F": 2, "VV": 3, "VE": 4, "VF": 5, "FV": 6, "FE": 7,
      "FF": 8, "EV": 9, "EE": 10, "EF": 11, "EVD": 12}

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp = ctypes.c_void_p
        L.ora_create.restype = vp
        L.ora_create.argtypes = [_u32p, ctypes.c_uint32]
        L.ora_destroy.argtypes = [vp]
        L.ora_set_patches.restype = ctypes.c_int
        L.ora_set_patches.argtypes = [vp, _u32p, _u32p, _u32p, ctypes.c_uint32]
        L.ora_counts.argtypes = [vp, _u64p]
        L.ora_stats.argtypes = [vp, _u32p]
        L.ora_get_ev.argtypes = [vp, _u32p]
        L.ora_get_ff.argtypes = [vp, _u32p, _u32p]
        L.ora_get_ownership.argtypes = [vp, _u32p, _u32p]
        L.ora_get_ribbon.argtypes = [vp, _u32p, _u32p]
        L.ora_patch_counts.argtypes = [vp, ctypes.c_uint32, _u32p]
        L.ora_patch_ltog.argtypes = [vp, ctypes.c_uint32, ctypes.c_int, _u32p]
        L.ora_patch_ev.argtypes = [vp, ctypes.c_uint32, _u16p]
        L.ora_patch_fe.argtypes = [vp, ctypes.c_uint32, _u16p]
        L.ora_linear_to_global.argtypes = [vp, ctypes.c_int, _u32p]
        L.ora_query_nnz.restype = ctypes.c_int64
        L.ora_query_nnz.argtypes = [vp, ctypes.c_int]
        L.ora_query.argtypes = [vp, ctypes.c_int, ctypes.c_int, _i32p, _i32p]
        L.ora_global_relation.restype = ctypes.c_int64
        L.ora_global_relation.argtypes = [vp, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.ora_vertex_normals.argtypes = [vp, _f64p, ctypes.c_int, _f64p]
        L.ora_gaussian_curvature.argtypes = [vp, _f64p, _f64p, _f64p]
        L.ora_mcf_cg.restype = ctypes.c_int
        L.ora_mcf_cg.argtypes = [ctypes.c_int64, _i32p, _i32p, _f64p, _f64p, ctypes.c_uint32, _f64p,
                                 ctypes.c_int, ctypes.c_double, ctypes.c_double, _f64p, _f64p]
        L.ora_smooth_uniform.argtypes = [vp, _f64p, ctypes.c_double, ctypes.c_int, _f64p]
        L.ora_mcf_pcg.restype = ctypes.c_int
        L.ora_mcf_pcg.argtypes = L.ora_mcf_cg.argtypes
        L.ora_uniform_assemble.restype = ctypes.c_int64
        L.ora_uniform_assemble.argtypes = [vp, ctypes.c_double, _i32p, _i32p, _f64p, _f64p]
        L.ora_cot_assemble.restype = ctypes.c_int64
        L.ora_cot_assemble.argtypes = [vp, _f64p, ctypes.c_double, _i32p, _i32p, _f64p, _f64p]
        L.ora_mcf_explicit.argtypes = [ctypes.c_int64, _i32p, _i32p, _f64p, _f64p,
                                       ctypes.c_uint32, _f64p, ctypes.c_int, _f64p]
        L.ora_rest_shapes.argtypes = [vp, _f64p, _f64p]
        L.ora_symdir_energy.argtypes = [vp, _f64p, _f64p, _f64p]
        L.ora_symdir_grad_dual.argtypes = [vp, _f64p, _f64p, _f64p]
        L.ora_symdir_grad_closed.argtypes = [vp, _f64p, _f64p, _f64p]
        L.ora_symdir3_energy.argtypes = [vp, _f64p, _f64p, _f64p]
        L.ora_symdir3_grad.argtypes = [vp, _f64p, _f64p, _f64p]
        L.ora_smooth_symdir_step.argtypes = [vp, _f64p, _f64p, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_int, _f64p]
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class OracleMesh:
    """FP64 CPU restatement of one RXMeshStatic build (global-id view)."""

    def __init__(self, fv):
        self.fv = _c(fv, np.uint32).reshape(-1, 3)
        self._h = lib().ora_create(self.fv.ravel(), self.fv.shape[0])
        self._patched = False

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.ora_destroy(self._h)
            self._h = None

    # -- a1: supporting structures -------------------------------------
    def counts(self):
        out = np.zeros(6, np.uint64)
        lib().ora_counts(self._h, out)
        return dict(nV=int(out[0]), nE=int(out[1]), nF=int(out[2]), P=int(out[3]),
                    ribbon=int(out[4]), ff=int(out[5]))

    def stats(self):
        out = np.zeros(5, np.uint32)
        lib().ora_stats(self._h, out)
        return dict(closed=bool(out[0]), manifold=bool(out[1]), max_valence=int(out[2]),
                    max_edge_faces=int(out[3]), max_face_adj=int(out[4]))

    @property
    def nV(self):
        return self.counts()["nV"]

    @property
    def nE(self):
        return self.counts()["nE"]

    @property
    def nF(self):
        return self.counts()["nF"]

    def ev(self):
        out = np.zeros(2 * self.nE, np.uint32)
        lib().ora_get_ev(self._h, out)
        return out.reshape(-1, 2)

    def ff(self):
        c = self.counts()
        off = np.zeros(c["nF"] + 1, np.uint32)
        val = np.zeros(max(c["ff"], 1), np.uint32)
        lib().ora_get_ff(self._h, off, val)
        return off, val[: c["ff"]]

    # -- a2..a5: patches -------------------------------------------------
    def set_patches(self, face_patch, patches_val=None, patches_offset=None):
        face_patch = _c(face_patch, np.uint32)
        P = int(face_patch.max()) + 1 if face_patch.size else 0
        if patches_val is None:
            # faces grouped by patch in ascending face order (patcher.cu:631-637)
            patches_val = np.argsort(face_patch, kind="stable").astype(np.uint32)
            patches_offset = np.cumsum(np.bincount(face_patch, minlength=P)).astype(np.uint32)
        patches_val = _c(patches_val, np.uint32)
        patches_offset = _c(patches_offset, np.uint32)[:P]
        rc = lib().ora_set_patches(self._h, face_patch, patches_val, patches_offset, P)
        if rc != 0:
            raise ValueError("patch exceeds 16-bit local ids")
        self._patched = True
        return self

    @property
    def P(self):
        return self.counts()["P"]

    def ownership(self):
        c = self.counts()
        vp = np.zeros(c["nV"], np.uint32)
        ep = np.zeros(c["nE"], np.uint32)
        lib().ora_get_ownership(self._h, vp, ep)
        return vp, ep

    def ribbon(self):
        c = self.counts()
        val = np.zeros(max(c["ribbon"], 1), np.uint32)
        off = np.zeros(max(c["P"], 1), np.uint32)
        lib().ora_get_ribbon(self._h, val, off)
        return val[: c["ribbon"]], off[: c["P"]]

    def patch_counts(self, p):
        out = np.zeros(6, np.uint32)
        lib().ora_patch_counts(self._h, p, out)
        return out

    def patch_ltog(self, p, kind):
        n = int(self.patch_counts(p)[kind])
        out = np.zeros(max(n, 1), np.uint32)
        lib().ora_patch_ltog(self._h, p, kind, out)
        return out[:n]

    def patch_ev(self, p):
        n = int(self.patch_counts(p)[1])
        out = np.zeros(max(2 * n, 1), np.uint16)
        lib().ora_patch_ev(self._h, p, out)
        return out[: 2 * n]

    def patch_fe(self, p):
        n = int(self.patch_counts(p)[2])
        out = np.zeros(max(3 * n, 1), np.uint16)
        lib().ora_patch_fe(self._h, p, out)
        return out[: 3 * n]

    def linear_to_global(self, kind):
        c = self.counts()
        n = [c["nV"], c["nE"], c["nF"]][kind]
        out = np.zeros(max(n, 1), np.uint32)
        lib().ora_linear_to_global(self._h, kind, out)
        return out[:n]

    # -- a8..a11: queries ------------------------------------------------
    def query(self, op, oriented=False):
        """CSR in linear-id space, deterministic order (row_ptr, col)."""
        o = OP[op]
        c = self.counts()
        nsrc = {"V": c["nV"], "E": c["nE"], "F": c["nF"]}[op[0]]
        nnz = lib().ora_query_nnz(self._h, o)
        rp = np.zeros(nsrc + 1, np.int32)
        col = np.zeros(max(nnz, 1), np.int32)
        lib().ora_query(self._h, o, int(oriented), rp, col)
        return rp, col[:nnz]

    def global_relation(self, op):
        """Brute-force ground truth in global ids (ascending rows)."""
        o = OP[op]
        c = self.counts()
        nsrc = {"V": c["nV"], "E": c["nE"], "F": c["nF"]}[op[0]]
        n = lib().ora_global_relation(self._h, o, None, None)
        rp = np.zeros(nsrc + 1, np.int32)
        col = np.zeros(max(n, 1), np.int32)
        lib().ora_global_relation(self._h, o, rp.ctypes.data, col.ctypes.data)
        return rp, col[:n]

    # -- geometry (global ids) -------------------------------------------
    def vertex_normals(self, pos, mode="max"):
        pos = _c(pos, np.float64).reshape(-1, 3)
        out = np.zeros_like(pos)
        lib().ora_vertex_normals(self._h, pos.ravel(), 0 if mode == "max" else 1, out.reshape(-1))
        return out

    def gaussian_curvature(self, pos):
        """(gc, mixed area) per vertex, global ids."""
        pos = _c(pos, np.float64).reshape(-1, 3)
        gc = np.zeros(len(pos), np.float64)
        am = np.zeros(len(pos), np.float64)
        lib().ora_gaussian_curvature(self._h, pos.ravel(), gc, am)
        return gc, am

    def smooth_uniform(self, pos, lam, iters):
        pos = _c(pos, np.float64).reshape(-1, 3)
        out = np.zeros_like(pos)
        lib().ora_smooth_uniform(self._h, pos.ravel(), lam, iters, out.reshape(-1))
        return out

    def cot_assemble(self, pos, dt):
        pos = _c(pos, np.float64).reshape(-1, 3)
        c = self.counts()
        nnz = c["nV"] + 2 * c["nE"]
        rows = np.zeros(nnz, np.int32)
        cols = np.zeros(nnz, np.int32)
        vals = np.zeros(nnz, np.float64)
        mass = np.zeros(c["nV"], np.float64)
        n = lib().ora_cot_assemble(self._h, pos.ravel(), dt, rows, cols, vals, mass)
        return rows[:n], cols[:n], vals[:n], mass

    def mcf_explicit(self, rows, cols, vals, mass, x0, iters):
        x0 = _c(x0, np.float64).reshape(-1, 3)
        out = np.zeros_like(x0)
        lib().ora_mcf_explicit(len(rows), _c(rows, np.int32), _c(cols, np.int32),
                               _c(vals, np.float64), _c(mass, np.float64), x0.shape[0],
                               x0.ravel(), iters, out.reshape(-1))
        return out

    def mcf_cg(self, rows, cols, vals, mass, x0, max_iter=100, abs_tol=1e-6, rel_tol=0.0, precond=False):
        """(x, iterations, start residual, final residual) of the reference's CG
        (|r|^2) or, with precond, its Jacobi-preconditioned CG (<r, r / diag A>)
        on A x = M x0."""
        x0 = _c(x0, np.float64).reshape(-1, 3)
        out = np.zeros_like(x0)
        res = np.zeros(2, np.float64)
        fn = lib().ora_mcf_pcg if precond else lib().ora_mcf_cg
        it = fn(len(rows), _c(rows, np.int32), _c(cols, np.int32), _c(vals, np.float64),
                _c(mass, np.float64), x0.shape[0], x0.ravel(), int(max_iter), float(abs_tol),
                float(rel_tol), out.reshape(-1), res)
        return out, it, res[0], res[1]

    def uniform_assemble(self, dt):
        """COO (rows, cols, vals) and right-hand-side scale (valence) of the MCF
        app's uniform-weight system, laid out as cot_assemble's."""
        c = self.counts()
        nnz = c["nV"] + 2 * c["nE"]
        rows = np.zeros(nnz, np.int32)
        cols = np.zeros(nnz, np.int32)
        vals = np.zeros(nnz, np.float64)
        mass = np.zeros(c["nV"], np.float64)
        n = lib().ora_uniform_assemble(self._h, float(dt), rows, cols, vals, mass)
        return rows[:n], cols[:n], vals[:n], mass

    def rest_shapes(self, pos):
        pos = _c(pos, np.float64).reshape(-1, 3)
        out = np.zeros((self.nF, 4), np.float64)
        lib().ora_rest_shapes(self._h, pos.ravel(), out.reshape(-1))
        return out

    def symdir_energy(self, Mr, uv):
        out = np.zeros(self.nF, np.float64)
        lib().ora_symdir_energy(self._h, _c(Mr, np.float64).ravel(), _c(uv, np.float64).ravel(), out)
        return out

    def symdir_grad(self, Mr, uv, method="closed"):
        out = np.zeros((self.nV, 2), np.float64)
        fn = lib().ora_symdir_grad_closed if method == "closed" else lib().ora_symdir_grad_dual
        fn(self._h, _c(Mr, np.float64).ravel(), _c(uv, np.float64).ravel(), out.reshape(-1))
        return out


def _symdir3(self, Mr, x):
    out = np.zeros(self.nF, np.float64)
    lib().ora_symdir3_energy(self._h, _c(Mr, np.float64).ravel(), _c(x, np.float64).ravel(), out)
    return out


def _symdir3_grad(self, Mr, x):
    out = np.zeros((self.nV, 3), np.float64)
    lib().ora_symdir3_grad(self._h, _c(Mr, np.float64).ravel(), _c(x, np.float64).ravel(), out.reshape(-1))
    return out


def _smooth_symdir_steps(self, Mr, x, lam, eta, steps):
    out = np.zeros((self.nV, 3), np.float64)
    lib().ora_smooth_symdir_step(self._h, _c(Mr, np.float64).ravel(), _c(x, np.float64).ravel(), lam, eta,
                                 steps, out.reshape(-1))
    return out


OracleMesh.symdir3_energy = _symdir3
OracleMesh.symdir3_grad = _symdir3_grad
OracleMesh.smooth_symdir_steps = _smooth_symdir_steps


# -- ARAP-style deformation (reference apps/DiffARAP), FP64, global ids -------
def _arap_weights(self, pos):
    """Per global edge: max(0, (cot_q + cot_s) / 2) over the EVDiamond
    [p, q, r, s] (arap_kernels.h:8-54), cot at v = dot(p - v, r - v) /
    |cross(p - v, r - v)|; 0 for edges lacking a side, which the reference
    skips and leaves at the matrix's reset value."""
    P = np.asarray(pos, np.float64).reshape(-1, 3)
    _, col = self.global_relation("EVD")
    d = col.reshape(-1, 4).astype(np.int64)
    ok = (d[:, 1] >= 0) & (d[:, 3] >= 0)
    dd = np.where(d < 0, 0, d)
    p, q, r, s = P[dd[:, 0]], P[dd[:, 1]], P[dd[:, 2]], P[dd[:, 3]]

    def cot(v):
        u0, u1 = p - v, r - v
        return (u0 * u1).sum(1) / np.linalg.norm(np.cross(u0, u1), axis=1)

    with np.errstate(divide="ignore", invalid="ignore"):
        w = (cot(q) + cot(s)) / 2
    return np.where(ok, np.maximum(0.0, np.nan_to_num(w, nan=0.0)), 0.0)


def _arap_covariances(self, rest, x, w):
    """S_v = sum over edges (v, j) of w (P_v - P_j)(X_v - X_j)^T."""
    P = np.asarray(rest, np.float64).reshape(-1, 3)
    X = np.asarray(x, np.float64).reshape(-1, 3)
    ev = self.ev().astype(np.int64)
    e, ep = P[ev[:, 0]] - P[ev[:, 1]], X[ev[:, 0]] - X[ev[:, 1]]
    outer = np.asarray(w, np.float64)[:, None, None] * e[:, :, None] * ep[:, None, :]
    S = np.zeros((len(P), 3, 3))
    np.add.at(S, ev[:, 0], outer)
    np.add.at(S, ev[:, 1], outer)  # (P_j - P_v)(X_j - X_v)^T is the same product
    return S


def _arap_rotations(self, rest, x, w):
    """Per vertex R = V U^T from S = U Sigma V^T (LAPACK SVD), with the sign
    of U's column of the smallest singular value flipped when det < 0
    (arap_kernels.h:60-117): the rotation maximising tr(R S). (V, 3, 3)."""
    S = _arap_covariances(self, rest, x, w)
    U, sig, Vt = np.linalg.svd(S)
    R = np.einsum("vji,vkj->vik", Vt, U)  # V U^T
    neg = np.linalg.det(R) < 0
    if neg.any():
        U2 = U[neg].copy()
        U2[:, :, 2] *= -1.0  # numpy orders singular values descending
        R[neg] = np.einsum("vji,vkj->vik", Vt[neg], U2)
    return R


def _arap_energy_grad(self, rest, x, w, R):
    """E = sum over edges (i, j) = EV order (i the larger global id) of
    w |(X_i - X_j) - R_i (P_i - P_j)|^2 (diff_arap.cu:87-118) and dE/dX
    with the rotations held fixed."""
    P = np.asarray(rest, np.float64).reshape(-1, 3)
    X = np.asarray(x, np.float64).reshape(-1, 3)
    R = np.asarray(R, np.float64).reshape(-1, 3, 3)
    w = np.asarray(w, np.float64)
    ev = self.ev().astype(np.int64)  # first-seen orientation; the EV query puts the larger id first
    i, j = ev.max(1), ev.min(1)
    d = (X[i] - X[j]) - np.einsum("eab,eb->ea", R[i], P[i] - P[j])
    E = float((w * (d * d).sum(1)).sum())
    g = np.zeros_like(X)
    np.add.at(g, i, 2 * w[:, None] * d)
    np.add.at(g, j, -2 * w[:, None] * d)
    return E, g


OracleMesh.arap_weights = _arap_weights
OracleMesh.arap_covariances = _arap_covariances
OracleMesh.arap_rotations = _arap_rotations
OracleMesh.arap_energy_grad = _arap_energy_grad


def read_patch_archive(path):
    """Decode a reference patch file (cereal PortableBinary of
    include/rxmesh/patcher/patcher.h:162-182): u8 endianness flag, nine u32
    scalars, seven (u64 length + u32[]) vectors, one f32."""
    import struct
    d = open(path, "rb").read()
    if d[0] != 1:
        raise ValueError("big-endian patch archives are not supported")
    off = 1
    names = ["patch_size", "num_patches", "num_vertices", "num_edges", "num_faces",
             "num_seeds", "max_num_patches", "num_components", "num_lloyd_run"]
    out = dict(zip(names, struct.unpack_from("<9I", d, off)))
    off += 36
    for name in ["face_patch", "vertex_patch", "edge_patch", "patches_val",
                 "patches_offset", "ribbon_ext_val", "ribbon_ext_offset"]:
        (n,) = struct.unpack_from("<Q", d, off)
        off += 8
        out[name] = np.frombuffer(d, np.uint32, n, off).copy()
        off += 4 * n
    (out["patching_time_ms"],) = struct.unpack_from("<f", d, off)
    return out
