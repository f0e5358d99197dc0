"""The product's host connectivity build (C++, behind the C-ABI, no GPU
needed with RXM_HOST_ONLY) against the oracle, plus patcher invariants and
error handling. CPU only."""
import ctypes
import os
import re

import numpy as np
import pytest

import rxmesh_amd as rx
from conftest import ROOT, fixture_path, obj
from oracle.oracle import OracleMesh, read_patch_archive
from rxmesh_amd import meshgen as mg


def compare_with_oracle(m, o, F):
    assert np.array_equal(m.array("ev"), o.ev().ravel())
    vp, ep = o.ownership()
    assert np.array_equal(m.array("vertex_patch"), vp)
    assert np.array_equal(m.array("edge_patch"), ep)
    rv, ro = o.ribbon()
    assert np.array_equal(m.array("ribbon_val"), rv)
    assert np.array_equal(m.array("ribbon_offset"), ro)
    pc = m.array("patch_counts").reshape(-1, 6)
    po = m.array("patch_offsets").reshape(3, -1)
    ltog = [m.array("ltog_v"), m.array("ltog_e"), m.array("ltog_f")]
    evl, fel = m.array("ev_local"), m.array("fe_local")
    for p in range(m.P):
        assert np.array_equal(pc[p], o.patch_counts(p))
        for k in range(3):
            assert np.array_equal(ltog[k][po[k][p]:po[k][p] + pc[p][k]], o.patch_ltog(p, k))
        e = evl[po[1][p]:po[1][p] + pc[p][1]]
        oe = o.patch_ev(p).reshape(-1, 2)
        assert np.array_equal(e & 0xFFFF, oe[:, 0]) and np.array_equal(e >> 16, oe[:, 1])
        assert np.array_equal(fel[3 * po[2][p]:3 * po[2][p] + 3 * pc[p][2]], o.patch_fe(p))
    for k, nm in enumerate(["v_l2g", "e_l2g", "f_l2g"]):
        assert np.array_equal(m.array(nm), o.linear_to_global(k))


@pytest.mark.parametrize("name", ["sphere3", "torus"])
def test_reference_patch_archive_replay(name):
    V, F = obj(name + ".obj")
    path = fixture_path(name + "_patches")
    m = rx.Mesh(F, patch_file=path, host_only=True)
    a = read_patch_archive(path)
    o = OracleMesh(F).set_patches(a["face_patch"], a["patches_val"], a["patches_offset"])
    assert m.P == a["num_patches"]
    compare_with_oracle(m, o, F)


@pytest.mark.parametrize("case", ["plane", "cube", "bunnyhead", "c0", "c1", "disk"])
@pytest.mark.parametrize("patch_size", [64, 512])
@pytest.mark.parametrize("patcher", ["bisect", "kmeans"])
def test_patcher_then_build_matches_oracle(case, patch_size, patcher):
    if case in ("plane", "cube", "bunnyhead"):
        V, F = obj(case + ".obj")
    elif case == "disk":
        V, F = mg.disk(48)
    else:
        V, F = mg.config_mesh(case)
    m = rx.Mesh(F, patch_size=patch_size, host_only=True, patcher=patcher)
    fp = m.array("face_patch")
    # patcher invariants: every face in exactly one patch, no patch above the cap
    cnt = np.bincount(fp, minlength=m.P)
    assert cnt.sum() == len(F) and cnt.min() >= 1 and cnt.max() <= max(patch_size, 1)
    o = OracleMesh(F).set_patches(fp, m.array("patches_val"), m.array("patches_offset"))
    compare_with_oracle(m, o, F)


@pytest.mark.parametrize("patcher", ["bisect", "kmeans"])
def test_patcher_is_deterministic(patcher):
    V, F = mg.config_mesh("c1")
    a = rx.Mesh(F, host_only=True, patcher=patcher).array("face_patch")
    b = rx.Mesh(F, host_only=True, patcher=patcher).array("face_patch")
    assert np.array_equal(a, b)


def test_patch_file_round_trip(tmp_path):
    V, F = mg.icosphere(4, seed=2, sigma=0.01)
    m = rx.Mesh(F, patch_size=128, host_only=True)
    path = str(tmp_path / "p.bin")
    m.save_patches(path)
    a = read_patch_archive(path)
    assert np.array_equal(a["face_patch"], m.array("face_patch"))
    assert np.array_equal(a["vertex_patch"], m.array("vertex_patch"))
    assert np.array_equal(a["ribbon_ext_val"], m.array("ribbon_val"))
    m2 = rx.Mesh(F, patch_file=path, host_only=True)
    assert np.array_equal(m2.array("ltog_v"), m.array("ltog_v"))


def test_explicit_face_patch():
    V, F = obj("sphere3.obj")
    fp = (np.arange(len(F)) * 5 // len(F)).astype(np.uint32)
    m = rx.Mesh(F, face_patch=fp, host_only=True)
    assert m.P == 5
    o = OracleMesh(F).set_patches(fp)
    compare_with_oracle(m, o, F)


def test_mesh_info_statistics():
    V, F = obj("plane.obj")
    i = rx.Mesh(F, host_only=True).info
    assert (i["num_vertices"], i["num_edges"], i["num_faces"]) == (9, 16, 8)
    assert i["is_closed"] == 0 and i["is_edge_manifold"] == 1 and i["max_valence"] == 8
    V, F = mg.config_mesh("c1")
    i = rx.Mesh(F, host_only=True).info
    assert (i["num_vertices"], i["num_edges"], i["num_faces"]) == (40962, 122880, 81920)
    assert i["is_closed"] == 1 and i["is_consistently_oriented"] == 1


@pytest.mark.parametrize("flip", [0, 5, 81919])
def test_consistent_orientation_flag(flip):
    """One face traversed backwards makes its edges run the same way in both
    faces: the flag drops (oriented queries then need a fixed mesh); the
    plane (boundary edges, single faces) and the icosphere keep it."""
    V, F = obj("plane.obj")
    assert rx.Mesh(F, host_only=True).info["is_consistently_oriented"] == 1
    V, F = mg.config_mesh("c1")
    F = F.copy()
    F[flip] = F[flip][::-1]
    i = rx.Mesh(F, host_only=True).info
    assert i["is_edge_manifold"] == 1 and i["is_consistently_oriented"] == 0


# ---- error handling: status codes, never exit (util/macros.h:77-89 exits) --
def test_errors_are_status_codes(tmp_path):
    with pytest.raises(rx.RXMError) as e:
        rx.Mesh(np.array([[0, 0, 1]], np.uint32), host_only=True)  # degenerate face
    assert e.value.code == 1
    with pytest.raises(rx.RXMError) as e:  # no faces
        rx.Mesh(np.zeros((0, 3), np.uint32), host_only=True)
    assert e.value.code == 1
    Vg, Fg = mg.grid(8)
    with pytest.raises(rx.RXMError) as e:  # vertices referenced by no face have no owner patch
        rx.Mesh(Fg, num_vertices=len(Vg) + 3, host_only=True)
    assert e.value.code == 1
    bad = tmp_path / "bad_patches"
    bad.write_bytes(b"\x01" + b"\x00" * 10)
    V, F = obj("sphere3.obj")
    with pytest.raises(rx.RXMError) as e:
        rx.Mesh(F, patch_file=str(bad), host_only=True)
    assert e.value.code == 4
    with pytest.raises(rx.RXMError) as e:  # archive of another mesh
        rx.Mesh(F, patch_file=fixture_path("torus_patches"), host_only=True)
    assert e.value.code == 1
    with pytest.raises(rx.RXMError):
        rx.Mesh(F, patch_size=4, host_only=True)
    with pytest.raises(rx.RXMError) as e:  # one patch of 81,920 faces: beyond 16-bit local ids
        rx.Mesh(mg.config_mesh("c1")[1], patch_size=1 << 20, host_only=True)
    assert e.value.code == 3
    m = rx.Mesh(F, host_only=True)
    # device entry points on a host-only mesh report "unsupported"
    rc = rx.lib().rxm_query_csr(m._h, 3, 0, ctypes.c_void_p(16), ctypes.c_void_p(16), None)
    assert rc == 6
    assert rx.lib().rxm_query_csr(m._h, 10, 0, ctypes.c_void_p(16), ctypes.c_void_p(16), None) == 1


def test_query_nnz_counts():
    V, F = mg.config_mesh("c1")
    m = rx.Mesh(F, host_only=True)
    E, Fn = m.nE, m.nF
    for op, n in {"VV": 2 * E, "VE": 2 * E, "VF": 3 * Fn, "FV": 3 * Fn, "FE": 3 * Fn,
                  "FF": 3 * Fn, "EV": 2 * E, "EF": 2 * E}.items():
        assert m.query_nnz(op) == n


# ---- the C-ABI library loads and exports every declared symbol -----------
def test_capi_exports_every_declared_symbol():
    hdrs = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
            if f.endswith(".h")]
    names = set()
    for h in hdrs:
        txt = open(h).read()
        names |= set(re.findall(r"^\s*rxm_status\s+(rxm_\w+)\s*\(", txt, re.M))
    assert len(names) >= 12
    L = ctypes.CDLL(rx.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


# ---- batches: the disjoint union of meshes ---------------------------------
def _batch_meshes():
    return [mg.icosphere(3, seed=1, sigma=0.01), mg.grid(20, seed=2), mg.icosphere(2, seed=3),
            mg.disk(24, seed=4)]


def test_batch_union_matches_separate_meshes():
    ms = _batch_meshes()
    b = rx.Mesh(None, batch=[f for _, f in ms], patch_size=128, host_only=True)
    sep = [rx.Mesh(f, patch_size=128, host_only=True) for _, f in ms]
    assert b.batch_count() == len(ms)
    ov, oe, of, op = (b.batch_offsets(k) for k in "VEFP")
    assert ov[-1] == b.nV == sum(m.nV for m in sep) and of[-1] == b.nF and op[-1] == b.P
    pc = b.array("patch_counts").reshape(-1, 6)
    for k, m in enumerate(sep):
        assert np.array_equal(b.array("v_l2g")[ov[k]:ov[k + 1]] - ov[k], m.array("v_l2g"))
        assert np.array_equal(b.array("f_l2g")[of[k]:of[k + 1]] - of[k], m.array("f_l2g"))
        assert np.array_equal(b.array("face_patch")[of[k]:of[k + 1]] - op[k], m.array("face_patch"))
        assert np.array_equal(pc[op[k]:op[k + 1]], m.array("patch_counts").reshape(-1, 6))
    for q in ("VV", "VE", "VF", "FV", "FE", "FF", "EV", "EF"):
        assert b.query_nnz(q) == sum(m.query_nnz(q) for m in sep)
    # the union is itself a mesh the oracle accepts, with identical layout
    fv_all = np.concatenate([f + ov[k] for k, (_, f) in enumerate(ms)])
    o = OracleMesh(fv_all).set_patches(b.array("face_patch"), b.array("patches_val"), b.array("patches_offset"))
    compare_with_oracle(b, o, fv_all)


# ---- the recursive-bisection patcher ------------------------------------------
@pytest.mark.parametrize("case", ["c0", "c1", "disk", "torus"])
def test_bisect_patcher_invariants_and_layout(case):
    if case == "disk":
        V, F = mg.disk(48)
    elif case == "torus":
        V, F = mg.torus(48, 24)
    else:
        V, F = mg.config_mesh(case)
    ps = 256
    m = rx.Mesh(F, patch_size=ps, host_only=True, patcher="bisect")
    fp = m.array("face_patch")
    cnt = np.bincount(fp, minlength=m.P)
    # none above the cap; close to the fewest the cap allows (pieces moved to
    # keep both halves of a cut connected can add a few patches)
    pmin = -(-len(F) // ps)
    assert pmin <= m.P <= 1.1 * pmin and cnt.min() >= 1 and cnt.max() <= ps
    # every patch is edge-connected
    import scipy.sparse as sps
    from scipy.sparse.csgraph import connected_components
    o = OracleMesh(F)
    off, cols = o.ff()
    rows = np.repeat(np.arange(len(F)), np.diff(off.astype(np.int64)))
    keep = fp[rows] == fp[cols]
    A = sps.coo_matrix((np.ones(keep.sum()), (rows[keep], cols[keep])), shape=(len(F), len(F)))
    ncomp, lab = connected_components(A, directed=False)
    assert ncomp == m.P
    assert np.array_equal(rx.Mesh(F, patch_size=ps, host_only=True, patcher="bisect").array("face_patch"), fp)
    o.set_patches(fp, m.array("patches_val"), m.array("patches_offset"))
    compare_with_oracle(m, o, F)


def test_bisect_patcher_components_get_their_own_patches():
    (V1, F1), (V2, F2) = mg.icosphere(2), mg.grid(12)
    F = np.concatenate([F1, F2 + len(V1)]).astype(np.uint32)
    m = rx.Mesh(F, patch_size=100, host_only=True, patcher="bisect")
    fp = m.array("face_patch")
    assert m.P <= 1.1 * (-(-len(F1) // 100) + -(-len(F2) // 100))
    assert not set(fp[:len(F1)]) & set(fp[len(F1):])


# ---- the two-step smoothing plan ------------------------------------------------
@pytest.mark.parametrize("case", ["c1", "disk", "bunnyhead"])
def test_vv2_plan_rows_are_owner_rows(case):
    """Every local vertex's row in the two-step plan lists its full one-ring
    in its owner's canonical VV order (the oracle's VV row), and every
    neighbour index stays inside the patch's extended local space."""
    if case == "disk":
        V, F = mg.disk(40, seed=2)
    elif case == "bunnyhead":
        V, F = obj("bunnyhead.obj")
    else:
        V, F = mg.config_mesh(case)
    m = rx.Mesh(F, patch_size=256, host_only=True)
    o = OracleMesh(F).set_patches(m.array("face_patch"), m.array("patches_val"), m.array("patches_offset"))
    orp, ocol = o.query("VV")                    # linear-id CSR
    l2g = m.array("v_l2g")
    P = m.P
    offs = m.array("vv2_offsets")
    rbase, nbase, xbase, nx = offs[:P + 1], offs[P + 1:2 * P + 2], offs[2 * P + 2:3 * P + 3], offs[3 * P + 3:]
    roff, nbr, xlin = m.array("vv2_roff"), m.array("vv2_nbr"), m.array("vv2_xlin")
    pc = m.array("patch_counts").reshape(-1, 6)
    po = m.array("patch_offsets").reshape(3, -1)
    ltog = m.array("ltog_v")
    g2l = m.array("v_g2l")
    for p in range(P):
        nv = pc[p, 0]
        ext = np.concatenate([ltog[po[0][p]:po[0][p] + nv], l2g[xlin[xbase[p]:xbase[p] + nx[p]]]])
        assert len(np.unique(ext)) == len(ext)
        ro = roff[rbase[p]:rbase[p] + nv + 1]
        assert ro[0] == 0 and ro[-1] <= nbase[p + 1] - nbase[p]
        for l in range(nv):
            row = nbr[nbase[p] + ro[l]:nbase[p] + ro[l + 1]]
            assert np.all(row < nv + nx[p])
            lin = g2l[ext[l]]
            want = l2g[ocol[orp[lin]:orp[lin + 1]]]
            assert np.array_equal(ext[row], want)


def test_concurrent_builds_equal_serial_builds():
    """Meshes built from several threads at once (the C-ABI releases the GIL),
    each with its host build on its own thread only (serial_build), export the
    same arrays as the same meshes built one by one with every host thread,
    and a batch built with its meshes in parallel (one thread each) equals the
    union of the serial ones."""
    from concurrent.futures import ThreadPoolExecutor
    meshes = [mg.icosphere(4, seed=k, sigma=0.01)[1] for k in range(3)] + [mg.grid(40 + 7 * k, seed=k)[1]
                                                                          for k in range(3)]
    meshes += [mg.permute_mesh(*mg.config_mesh("c1"), seed=5)[1]]
    serial = [rx.Mesh(f, patch_size=256, host_only=True) for f in meshes]
    with ThreadPoolExecutor(max_workers=4) as pool:
        conc = list(pool.map(lambda f: rx.Mesh(f, patch_size=256, host_only=True, serial_build=True), meshes))
    names = [a for a in rx.ARR if not a.startswith("vv2")]
    for a, b in zip(serial, conc):
        for n in names:
            assert np.array_equal(a.array(n), b.array(n)), n
    batch = rx.Mesh(None, batch=meshes, patch_size=256, host_only=True)
    union = rx.Mesh(None, parts=serial, host_only=True)
    for n in names:
        assert np.array_equal(batch.array(n), union.array(n)), n
