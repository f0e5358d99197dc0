"""GPU parity: the HIP query engine (through the C-ABI) against the oracle.

Connectivity is compared bit for bit: same row pointers, same entries in the
same order (ascending source-local order, or the canonical CCW ring when
oriented). Geometry is compared to the FP64 oracle evaluated on the same
fp32 inputs, norm-wise per vertex: max_v |x - x64| / |x64| <= tolerance."""
import numpy as np
import pytest

from conftest import fixture_path, obj
from oracle.oracle import OracleMesh, read_patch_archive
from rxmesh_amd import meshgen as mg

pytestmark = pytest.mark.gpu

OPS = ["VV", "VE", "VF", "FV", "FE", "FF", "EV", "EF"]
_CACHE = {}


def case(name):
    """(positions, fv, product mesh, oracle mesh) with identical patches."""
    if name in _CACHE:
        return _CACHE[name]
    import rxmesh_amd as rx
    print("case", name, flush=True)
    if name in ("sphere3", "torus"):
        V, F = obj(name + ".obj")
        path = fixture_path(name + "_patches")
        m = rx.Mesh(F, patch_file=path)
        a = read_patch_archive(path)
        o = OracleMesh(F).set_patches(a["face_patch"], a["patches_val"], a["patches_offset"])
    else:
        if name in ("plane", "cube", "bunnyhead"):
            V, F = obj(name + ".obj")
            ps = 512
        elif name == "c0":
            V, F = mg.config_mesh("c0")
            ps = 512
        elif name == "c1":
            V, F = mg.config_mesh("c1")
            ps = 512
        elif name == "c1_p128":
            V, F = mg.config_mesh("c1")
            ps = 128
        elif name == "disk":
            V, F = mg.disk(64, seed=1)
            ps = 256
        elif name == "c1_kmeans":
            V, F = mg.config_mesh("c1")
            ps = 512
        m = rx.Mesh(F, patch_size=ps, patcher="kmeans" if name.endswith("_kmeans") else "bisect")
        o = OracleMesh(F).set_patches(m.array("face_patch"), m.array("patches_val"),
                                      m.array("patches_offset"))
    _CACHE[name] = (V, F, m, o)
    return _CACHE[name]


def lin_pos(m, V):
    """Positions as a (3, V) fp32 SoA device tensor in linear order."""
    import torch
    l2g = m.array("v_l2g")
    return torch.from_numpy(np.ascontiguousarray(V[l2g].T.astype(np.float32))).cuda()


def to_global_np(m, x_lin):
    l2g = m.array("v_l2g")
    out = np.empty_like(x_lin)
    out[:, l2g] = x_lin
    return out


NAMES = ["sphere3", "torus", "plane", "cube", "bunnyhead", "c0", "c1", "c1_p128", "disk", "c1_kmeans"]


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("op", OPS + ["EVD"])
def test_query_bit_exact(name, op):
    import torch
    V, F, m, o = case(name)
    rp, col = m.query(op)
    torch.cuda.synchronize()
    orp, ocol = o.query(op)
    assert np.array_equal(rp.cpu().numpy(), orp)
    assert np.array_equal(col.cpu().numpy()[: len(ocol)], ocol)


@pytest.mark.parametrize("name", ["sphere3", "torus", "plane", "cube", "bunnyhead", "c1", "disk", "c1_kmeans"])
@pytest.mark.parametrize("op", ["VV", "VE"])
@pytest.mark.parametrize("threads", [256, 512])
def test_oriented_query_bit_exact(name, op, threads):
    import torch
    V, F, m, o = case(name)
    m.set_launch_threads(6, threads)
    rp, col = m.query(op, oriented=True)
    torch.cuda.synchronize()
    m.set_launch_threads(6, 256)
    orp, ocol = o.query(op, oriented=True)
    assert np.array_equal(rp.cpu().numpy(), orp)
    assert np.array_equal(col.cpu().numpy(), ocol)


def test_oriented_paths_refuse_inconsistent_orientation():
    """A face traversed backwards leaves half-edges with two faces: the
    oriented query and the cotangent assembly return RXM_ERR_NON_MANIFOLD
    while the unoriented queries still run."""
    import rxmesh_amd as rx
    import torch
    V, F = mg.config_mesh("c1")
    F = F.copy()
    F[7] = F[7][::-1]
    m = rx.Mesh(F)
    for op in ("VV", "VE"):
        with pytest.raises(rx.RXMError) as e:
            m.query(op, oriented=True)
        assert e.value.code == 2
    with pytest.raises(rx.RXMError) as e:
        m.cot_assemble(torch.zeros(3, m.nV, device="cuda"), 1e-3)
    assert e.value.code == 2
    rp, col = m.query("VV")
    torch.cuda.synchronize()
    assert int(rp[-1]) == m.query_nnz("VV")


@pytest.mark.parametrize("name", ["sphere3", "torus", "plane", "cube", "bunnyhead", "c1", "c1_p128", "disk",
                                  "c1_kmeans"])
def test_query_vv_vf_one_launch(name):
    """The fused VV + VF launch writes both CSRs exactly as the oracle (and
    the separate queries) do."""
    import torch
    V, F, m, o = case(name)
    (vrp, vcol), (frp, fcol) = m.query_vv_vf()
    torch.cuda.synchronize()
    for op, rp, col in (("VV", vrp, vcol), ("VF", frp, fcol)):
        orp, ocol = o.query(op)
        assert np.array_equal(rp.cpu().numpy(), orp)
        assert np.array_equal(col.cpu().numpy()[: len(ocol)], ocol)


@pytest.mark.parametrize("seed", range(8))
def test_random_meshes_and_patch_sizes(seed):
    """Seeded sweep: a permuted jittered sphere, disk or torus with a random
    patch size (64..512) and patcher; every query, oriented VV / VE and the
    fused VV + VF launch equal the oracle bit for bit."""
    import torch
    import rxmesh_amd as rx
    rng = np.random.default_rng(1000 + seed)
    kind = ("sphere", "disk", "torus")[seed % 3]
    if kind == "sphere":
        V, F = mg.icosphere(int(rng.integers(3, 5)), seed=seed, sigma=0.01)
    elif kind == "disk":
        V, F = mg.disk(int(rng.integers(24, 60)), seed=seed)
    else:
        V, F = mg.torus(int(rng.integers(20, 60)), int(rng.integers(12, 30)))
    V, F = mg.permute_mesh(V, F, seed=seed)
    ps = int(rng.integers(64, 513))
    patcher = "kmeans" if seed % 2 else "bisect"
    print("case", kind, len(F), ps, patcher, flush=True)
    m = rx.Mesh(F, patch_size=ps, patcher=patcher)
    o = OracleMesh(F).set_patches(m.array("face_patch"), m.array("patches_val"), m.array("patches_offset"))
    for op in OPS + ["EVD"]:
        rp, col = m.query(op)
        torch.cuda.synchronize()
        orp, ocol = o.query(op)
        assert np.array_equal(rp.cpu().numpy(), orp), op
        assert np.array_equal(col.cpu().numpy()[: len(ocol)], ocol), op
    for op in ("VV", "VE"):
        rp, col = m.query(op, oriented=True)
        torch.cuda.synchronize()
        orp, ocol = o.query(op, oriented=True)
        assert np.array_equal(rp.cpu().numpy(), orp) and np.array_equal(col.cpu().numpy(), ocol), op
    (vrp, vcol), (frp, fcol) = m.query_vv_vf()
    torch.cuda.synchronize()
    for op, rp, col in (("VV", vrp, vcol), ("VF", frp, fcol)):
        orp, ocol = o.query(op)
        assert np.array_equal(rp.cpu().numpy(), orp) and np.array_equal(col.cpu().numpy()[: len(ocol)], ocol)


def test_largest_patch_that_fits():
    """One patch holding a whole 5,120-face sphere (7,680 edges: the VV
    transpose needs about 111 KB of LDS, beyond the 64 KB default): every
    query is bit-exact and the normals and smoothing match FP64, or an entry
    point reports RXM_ERR_PATCH_TOO_BIG cleanly (FF only, at this size)."""
    import torch
    import rxmesh_amd as rx
    V, F = mg.icosphere(4, seed=1, sigma=0.01)
    m = rx.Mesh(F, patch_size=8192)
    assert m.P == 1
    o = OracleMesh(F).set_patches(m.array("face_patch"), m.array("patches_val"), m.array("patches_offset"))
    refused = []
    for op in OPS + ["EVD"]:
        print("op", op, flush=True)
        try:
            rp, col = m.query(op)
        except rx.RXMError as e:
            assert e.code == 3
            refused.append(op)
            continue
        torch.cuda.synchronize()
        orp, ocol = o.query(op)
        assert np.array_equal(rp.cpu().numpy(), orp)
        assert np.array_equal(col.cpu().numpy()[: len(ocol)], ocol)
    # FF stages FE, the edge offsets, per-face counts and 3F slots: just over 160 KB here
    assert set(refused) <= {"FF"}
    V32 = V.astype(np.float32)
    x = lin_pos(m, V32)
    n = m.vertex_normals(x, "max")
    s = m.smooth(x, 0.5, 4)
    torch.cuda.synchronize()
    got = to_global_np(m, n.cpu().numpy()).T
    ref = o.vertex_normals(V32.astype(np.float64), "max")
    assert (np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)).max() <= 1e-5
    got = to_global_np(m, s.cpu().numpy()).T
    ref = o.smooth_uniform(V32.astype(np.float64), 0.5, 4)
    assert (np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)).max() <= 1e-5


def test_evdiamond_inconsistent_orientation():
    """With one face traversed backwards, two faces run some edges the same
    way: the EVDiamond slot takes the later face (serial order of the
    reference's face pass), on the device as in the oracle."""
    import rxmesh_amd as rx
    import torch
    V, F = mg.config_mesh("c1")
    F = F.copy()
    F[7] = F[7][::-1]
    m = rx.Mesh(F, patch_size=512)
    o = OracleMesh(F).set_patches(m.array("face_patch"), m.array("patches_val"), m.array("patches_offset"))
    rp, col = m.query("EVD")
    torch.cuda.synchronize()
    orp, ocol = o.query("EVD")
    assert np.array_equal(rp.cpu().numpy(), orp)
    assert np.array_equal(col.cpu().numpy(), ocol)


def test_query_set_semantics_vs_bruteforce():
    # the reference verifier (tests/RXMesh_test/rxmesh_test.h:343-440): per
    # source element, outputs are owned (linear ids) and equal the ground truth
    # as a set with the same count
    V, F, m, o = case("c1")
    for op in OPS:
        rp, col = (t.cpu().numpy() for t in m.query(op))
        src = m.array({"V": "v_l2g", "E": "e_l2g", "F": "f_l2g"}[op[0]])
        dst = m.array({"V": "v_l2g", "E": "e_l2g", "F": "f_l2g"}[op[1]])
        grp, gcol = o.global_relation(op)
        for l in range(0, len(rp) - 1, 97):
            s = src[l]
            got = sorted(dst[col[rp[l]:rp[l + 1]]].tolist())
            assert got == sorted(gcol[grp[s]:grp[s + 1]].tolist())


def test_queries_are_deterministic():
    import torch
    V, F, m, o = case("c1_p128")
    for op in OPS:
        a = [t.clone() for t in m.query(op)]
        b = m.query(op)
        torch.cuda.synchronize()
        assert torch.equal(a[0], b[0]) and torch.equal(a[1][: m.query_nnz(op)], b[1][: m.query_nnz(op)])


@pytest.mark.parametrize("name", ["c0", "c1", "sphere3", "bunnyhead", "disk"])
@pytest.mark.parametrize("mode", ["max", "area"])
def test_vertex_normals_vs_fp64(name, mode):
    import torch
    V, F, m, o = case(name)
    V32 = V.astype(np.float32)
    n = m.vertex_normals(lin_pos(m, V32), mode)
    torch.cuda.synchronize()
    got = to_global_np(m, n.cpu().numpy()).T
    ref = o.vertex_normals(V32.astype(np.float64), mode)
    err = np.linalg.norm(got - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-30)
    # per-result tolerance: 1e-5 relative, norm-wise per vertex
    assert err.max() <= 1e-5, err.max()


@pytest.mark.parametrize("name", ["c1", "c0", "torus"])
def test_smoothing_vs_fp64(name):
    import torch
    V, F, m, o = case(name)
    V32 = V.astype(np.float32)
    x = lin_pos(m, V32)
    out = m.smooth(x, 0.5, 10)
    torch.cuda.synchronize()
    got = to_global_np(m, out.cpu().numpy()).T
    ref = o.smooth_uniform(V32.astype(np.float64), 0.5, 10)
    err = np.linalg.norm(got - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-30)
    assert err.max() <= 1e-5, err.max()
    # deterministic: a second run is bitwise identical
    out2 = m.smooth(x, 0.5, 10)
    assert torch.equal(out, out2)


def test_smoothing_launch_configs_agree():
    """Every launch configuration the autotuner may pick (workgroup size, LDS
    layout, one or two Jacobi steps per launch) gives bit-identical results:
    same neighbours, same summation order."""
    import torch
    V, F, m, o = case("c1")
    x = lin_pos(m, V.astype(np.float32))
    from rxmesh_amd import autotune
    for iters in (2, 3, 4, 7):
        outs = []
        for cfg in autotune.smooth_configs():
            autotune.apply(m, cfg)
            outs.append(m.smooth(x, 0.5, iters))
        autotune.apply(m, {"smooth_threads": 128, "smooth_layout": 0, "smooth_fused": 0})
        torch.cuda.synchronize()
        for y in outs[1:]:
            assert torch.equal(outs[0], y)


def test_vv_plan_matches_vv_query():
    # the patch-local plan decodes (through the owner table) to the VV CSR
    import ctypes
    import rxmesh_amd as rx
    V, F, m, o = case("c1_p128")
    m.build_vv_plan()
    orp, ocol = o.query("VV")
    # smoothing one step from the plan equals a host step over the oracle CSR
    x = lin_pos(m, V.astype(np.float32))
    out = m.smooth(x, 0.5, 1).cpu().numpy().astype(np.float64)
    xl = x.cpu().numpy().astype(np.float64)
    ref = np.empty_like(xl)
    for r in range(len(orp) - 1):
        nb = ocol[orp[r]:orp[r + 1]]
        ref[:, r] = xl[:, r] + 0.5 * (xl[:, nb].mean(1) - xl[:, r])
    assert np.abs(out - ref).max() <= 1e-6


@pytest.mark.parametrize("name", ["c1", "c0", "bunnyhead"])
def test_smoothing_on_vv_csr_vs_fp64(name):
    import torch
    V, F, m, o = case(name)
    V32 = V.astype(np.float32)
    x = lin_pos(m, V32)
    vv = m.query("VV")
    out = m.smooth_csr(vv, x, 0.5, 10)
    torch.cuda.synchronize()
    got = to_global_np(m, out.cpu().numpy()).T
    ref = o.smooth_uniform(V32.astype(np.float64), 0.5, 10)
    err = np.linalg.norm(got - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-30)
    assert err.max() <= 1e-5, err.max()
    assert torch.equal(out, m.smooth_csr(vv, x, 0.5, 10))
    # one lane per row sums each row in CSR order, the patch-local smoothing's
    # neighbour order: the two paths agree bit for bit
    assert torch.equal(out, m.smooth(x, 0.5, 10))


def test_permutation_round_trip():
    import torch
    V, F, m, o = case("c1")
    xg = torch.from_numpy(np.ascontiguousarray(V.T.astype(np.float32))).cuda()
    xl = m.to_linear(xg)
    back = m.to_global(xl)
    torch.cuda.synchronize()
    assert torch.equal(back, xg)
    assert torch.equal(xl, lin_pos(m, V.astype(np.float32)))


@pytest.mark.parametrize("name", ["sphere3", "bunnyhead", "c1", "c1_p128", "disk"])
def test_stable_and_atomic_transposes_agree(name):
    """Knob 10 switches the transposed queries between the LDS atomic
    histogram + row sort (0), the atomic-free stable counting sort (1:
    ballot / mbcnt ranks, rows in entry order) and the one-pass form (2, the
    default: per-edge / per-face keys held in registers, rows sorted in
    registers; VV, VE, VF, EF and FF by their own launch, VV + VF in one
    counting sort by the fused launch): identical CSRs, bit for bit, for every
    transposed op and for the fused launch."""
    import torch
    V, F, m, o = case(name)
    got = {}
    for knob in (0, 1, 2):
        m.set_launch_threads(10, knob)
        got[knob] = {op: tuple(t.cpu() for t in m.query(op)) for op in ("VV", "VE", "VF", "EF", "FF")}
        (a, b), (c, d) = m.query_vv_vf()
        got[knob]["VVF"] = (a.cpu(), b.cpu(), c.cpu(), d.cpu())
    m.set_launch_threads(10, 2)
    for knob in (1, 2):
        for op in got[0]:
            for u, v in zip(got[0][op], got[knob][op]):
                assert torch.equal(u, v), (knob, op)
