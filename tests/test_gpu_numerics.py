"""GPU parity of the numerics on top of the query engine: cotangent CSR
assembly, SpMV-3 / explicit MCF, symmetric Dirichlet energy and its
hand-written adjoint (also through torch.autograd)."""
import numpy as np
import pytest

from oracle.oracle import OracleMesh
from rxmesh_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def _mesh(name):
    import rxmesh_amd as rx
    if name == "ico4":
        V, F = mg.icosphere(4, seed=1, sigma=0.01)
        ps = 256
    elif name == "ico5_perm":
        V, F = mg.permute_mesh(*mg.icosphere(5, seed=2, sigma=0.005), seed=2)
        ps = 512
    elif name == "torus":
        V, F = mg.torus(96, 48)
        ps = 512
    elif name == "disk":
        V, F = mg.disk(64, seed=0)
        ps = 256
    elif name == "c3":
        V, F = mg.config_mesh("c3")
        ps = 512
    m = rx.Mesh(F, patch_size=ps)
    o = OracleMesh(F).set_patches(m.array("face_patch"), m.array("patches_val"), m.array("patches_offset"))
    return V, F, m, o


def _lin(m, X, kind="v_l2g"):
    import torch
    l2g = m.array(kind)
    return torch.from_numpy(np.ascontiguousarray(X[l2g].T.astype(np.float32))).cuda()


def _glob(m, Y, kind="v_l2g"):
    l2g = m.array(kind)
    out = np.empty_like(Y)
    out[..., l2g] = Y
    return out


@pytest.mark.parametrize("name", ["ico4", "ico5_perm", "torus", "disk"])
def test_cot_assembly_vs_fp64(name):
    import torch
    V, F, m, o = _mesh(name)
    V32 = V.astype(np.float32)
    dt = 1e-3
    rp, col, val, mass = m.cot_assemble(_lin(m, V32), dt)
    torch.cuda.synchronize()
    rp, col, val, mass = (t.cpu().numpy() for t in (rp, col, val, mass))
    orows, ocols, ovals, omass = o.cot_assemble(V32.astype(np.float64), dt)
    l2g = m.array("v_l2g")
    # pattern: row = [diag, VV query row] in the same order as the oracle
    assert rp[-1] == len(ocols) == m.cot_nnz()
    g2l = m.array("v_g2l")
    # compare entry by entry in row order (oracle rows are in linear order too)
    order = np.argsort(g2l[orows], kind="stable")
    assert np.array_equal(col, g2l[ocols[order]])
    ref = ovals[order]
    scale = np.abs(ref).max()
    assert np.abs(val - ref).max() <= 1e-5 * scale
    gm = _glob(m, mass)
    assert np.abs(gm - omass).max() <= 1e-5 * np.abs(omass).max()
    # row sums of L vanish: A 1 = M (up to fp32 rounding)
    rs = np.add.reduceat(val.astype(np.float64), rp[:-1])
    assert np.abs(rs - mass).max() <= 1e-5 * np.abs(val).max()


@pytest.mark.parametrize("name", ["ico5_perm", "disk"])
def test_cot_assembly_workgroup_sizes_agree(name):
    """Knob 7 (256 / 512 threads per patch) changes the launch only: the
    systems are bit-identical, and so are two calls in a row."""
    import torch
    V, F, m, o = _mesh(name)
    x = _lin(m, V.astype(np.float32))
    m.set_launch_threads(7, 256)
    a = [t.clone() for t in m.cot_assemble(x, 1e-3)]
    m.set_launch_threads(7, 512)
    b = m.cot_assemble(x, 1e-3)
    c = m.cot_assemble(x, 1e-3, out=tuple(torch.empty_like(t) for t in b))
    torch.cuda.synchronize()
    for u, v, w in zip(a, b, c):
        assert torch.equal(u, v) and torch.equal(v, w)


@pytest.mark.parametrize("name", ["ico4", "torus"])
@pytest.mark.parametrize("lanes", [1, 2, 4, 8])
def test_spmv_and_mcf_vs_fp64(name, lanes):
    import torch
    V, F, m, o = _mesh(name)
    m.set_launch_threads(5, lanes)
    V32 = V.astype(np.float32)
    dt = 1e-5
    x = _lin(m, V32)
    A = m.cot_assemble(x, dt)
    # all-ones SpMV equals the mass (SparseMatrixEdgeLen-style closed form)
    ones = torch.ones_like(x)
    y = m.spmv3(A, ones)
    torch.cuda.synchronize()
    assert torch.allclose(y[0], A[3], rtol=1e-5, atol=1e-6 * float(A[3].abs().max()))
    # 50 explicit MCF steps vs FP64 on the oracle's system
    out = m.mcf_explicit(A, x, 50)
    torch.cuda.synchronize()
    got = _glob(m, out.cpu().numpy()).T
    orows, ocols, ovals, omass = o.cot_assemble(V32.astype(np.float64), dt)
    ref = o.mcf_explicit(orows, ocols, ovals, omass, V32.astype(np.float64), 50)
    err = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert err.max() <= 1e-5, err.max()
    # determinism
    out2 = m.mcf_explicit(A, x, 50)
    assert torch.equal(out, out2)
    m.set_launch_threads(5, 8)


@pytest.mark.parametrize("name", ["disk", "c3"])
def test_symdir_energy_and_grad_vs_fp64(name):
    import torch
    V, F, m, o = _mesh(name)
    V32 = V.astype(np.float32)
    uv = mg.uv_init(V, F, seed=0).astype(np.float32)
    Mr = m.symdir_rest(_lin(m, V32))
    uvl = _lin(m, uv)
    total, energy = m.symdir_energy(Mr, uvl)
    grad = m.symdir_grad(Mr, uvl)
    torch.cuda.synchronize()
    # rest shapes vs FP64
    oMr = o.rest_shapes(V32.astype(np.float64))
    gMr = _glob(m, Mr.cpu().numpy(), "f_l2g").T
    assert np.abs(gMr - oMr).max() <= 1e-5 * np.abs(oMr).max()
    # energies use the oracle's FP64 evaluation on the GPU's fp32 rest shapes
    # (isolates the energy kernel from the rest-shape rounding)
    oE = o.symdir_energy(gMr.astype(np.float64), uv.astype(np.float64))
    gE = _glob(m, energy.cpu().numpy(), "f_l2g")
    assert np.all(np.isfinite(oE))
    rel = np.abs(gE - oE) / oE
    assert rel.max() <= 1e-5, rel.max()
    assert abs(float(total) - oE.sum()) <= 1e-5 * oE.sum()
    # gradient vs FP64 closed form and the forward-mode dual restatement
    g64 = o.symdir_grad(gMr.astype(np.float64), uv.astype(np.float64), "closed")
    d64 = o.symdir_grad(gMr.astype(np.float64), uv.astype(np.float64), "dual")
    gg = _glob(m, grad.cpu().numpy()).T
    gmax = np.abs(g64).max()
    # per-result tolerance: max |g - g64| <= 1e-5 * max |g64|
    assert np.abs(gg - g64).max() <= 1e-5 * gmax
    assert np.abs(gg - d64).max() <= 1e-5 * gmax


def test_symdir_grad_vs_central_differences():
    import torch
    V, F, m, o = _mesh("disk")
    V32 = V.astype(np.float32)
    uv = mg.uv_init(V, F, seed=0).astype(np.float32)
    Mr = m.symdir_rest(_lin(m, V32))
    grad = m.symdir_grad(Mr, _lin(m, uv))
    torch.cuda.synchronize()
    gMr = _glob(m, Mr.cpu().numpy(), "f_l2g").T.astype(np.float64)
    gg = _glob(m, grad.cpu().numpy()).T
    uv64 = uv.astype(np.float64)
    h = 1e-6
    gmax = np.abs(gg).max()
    for v in np.random.default_rng(1).choice(len(uv), 20, replace=False):
        for c in range(2):
            up, dn = uv64.copy(), uv64.copy()
            up[v, c] += h
            dn[v, c] -= h
            fd = (o.symdir_energy(gMr, up).sum() - o.symdir_energy(gMr, dn).sum()) / (2 * h)
            assert abs(fd - gg[v, c]) <= 1e-5 * gmax + 1e-6 * abs(fd)


def test_symdir_autograd_function():
    import torch
    import rxmesh_amd as rx
    V, F, m, o = _mesh("disk")
    V32 = V.astype(np.float32)
    uv = mg.uv_init(V, F, seed=0).astype(np.float32)
    Mr = m.symdir_rest(_lin(m, V32))
    uvl = _lin(m, uv).requires_grad_(True)
    e = rx.symmetric_dirichlet(m, Mr, uvl)
    (3.0 * e).backward()
    direct = m.symdir_grad(Mr, uvl.detach())
    torch.cuda.synchronize()
    assert torch.allclose(uvl.grad, 3.0 * direct, rtol=1e-6, atol=0)
    # a backtracking gradient step decreases the energy (the energy is a
    # barrier: too long a step flips a face and returns FLT_MAX)
    x = uvl.detach().clone()
    e0 = float(m.symdir_energy(Mr, x)[0])
    g = m.symdir_grad(Mr, x)
    step, e1 = 1e-3, float("inf")
    while step > 1e-12:
        e1 = float(m.symdir_energy(Mr, x - step * g)[0])
        if e1 < e0:
            break
        step *= 0.5
    assert e1 < e0


def test_symdir_flipped_face():
    import torch
    import rxmesh_amd as rx
    F = np.array([[0, 1, 2], [1, 3, 2]], np.uint32)
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float64)
    m = rx.Mesh(F, patch_size=64)
    Mr = m.symdir_rest(_lin(m, V.astype(np.float32)))
    uv = V[:, :2].copy()
    uv[3] = [-1.0, -1.0]  # flips face 1
    total, energy = m.symdir_energy(Mr, _lin(m, uv))
    torch.cuda.synchronize()
    gE = _glob(m, energy.cpu().numpy(), "f_l2g")
    assert gE[1] == np.finfo(np.float32).max and np.isfinite(gE[0])


@pytest.mark.parametrize("name", ["ico4", "ico5_perm", "torus", "disk"])
def test_gaussian_curvature_vs_fp64(name):
    """Gaussian curvature against the FP64 restatement of the reference's
    serial version. The angle defect 2 pi - sum(angles) cancels in fp32, so
    it is compared absolutely (1e-5 rad: a handful of fp32 angle roundings);
    the mixed area relatively (1e-5); gc = defect / area follows. Two calls
    are bitwise identical (no atomics)."""
    import torch
    V, F, m, o = _mesh(name)
    V32 = V.astype(np.float32)
    gc, am = m.gaussian_curvature(_lin(m, V32))
    gc2, am2 = m.gaussian_curvature(_lin(m, V32))
    torch.cuda.synchronize()
    assert torch.equal(gc, gc2) and torch.equal(am, am2)
    g = _glob(m, gc.cpu().numpy().astype(np.float64))
    a = _glob(m, am.cpu().numpy().astype(np.float64))
    rg, ra = o.gaussian_curvature(V32.astype(np.float64))
    assert np.abs(a - ra).max() <= 1e-5 * np.abs(ra).max()
    assert np.abs(g * a - rg * ra).max() <= 1e-5
    # Gauss-Bonnet on the closed meshes: the oracle's defects sum to 2 pi chi,
    # the fp32 ones within the per-vertex bound times the vertex count
    if name != "disk":
        chi = 0 if name == "torus" else 2
        assert abs((rg * ra).sum() - 2 * np.pi * chi) < 1e-9
        assert abs((g * a).sum() - 2 * np.pi * chi) <= 1e-5 * len(g)


@pytest.mark.parametrize("name,dt,fwd_tol", [("ico4", 1e-2, 1e-5), ("torus", 1e-2, 1e-5), ("ico4", 10.0, 2e-4)])
def test_mcf_implicit_cg_vs_direct_solve(name, dt, fwd_tol):
    """The device CG on A X = M x (reference CGSolver semantics) solves the
    fp32 system to fp32 accuracy: the FP64 residual of its solution is below
    1e-5 of || |A| |x| || (backward error; CG's recursively updated residual,
    the one the stopping test reads, runs below the true one in fp32 -- the
    reference's solver behaves the same), and it is within fwd_tol (relative
    to the largest coordinate) of the FP64 direct solution: 1e-5 for
    dt = 1e-2; 2e-4 for the reference's default dt = 10, where A = M + 10 L
    shrinks the sphere twentyfold and its conditioning amplifies fp32
    rounding. It needs about as many iterations as the FP64 restatement of
    the same CG, and two solves are bit-identical."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    import torch
    V, F, m, o = _mesh(name)
    x = _lin(m, V.astype(np.float32))
    A = m.cot_assemble(x, dt)
    X, it, r0, r1 = m.mcf_implicit_cg(A, x, max_iter=500, abs_tol=1e-12)
    X2, it2, _, _ = m.mcf_implicit_cg(A, x, max_iter=500, abs_tol=1e-12)
    torch.cuda.synchronize()
    assert torch.equal(X, X2) and it == it2
    rp, col, val, mass = (t.cpu().numpy() for t in A)
    n = len(rp) - 1
    # scipy keeps the index arrays it is given and sorts them in place on use;
    # the rows here are in [diag, VV] order, so it gets copies
    As = sp.csr_matrix((val.astype(np.float64), col.copy(), rp.copy()), shape=(n, n))
    b = mass.astype(np.float64)[:, None] * x.cpu().numpy().T.astype(np.float64)
    xs = sla.spsolve(As.tocsc(), b)
    got = X.cpu().numpy().T.astype(np.float64)
    assert r1 < 1e-12 and 0 < it < 500
    assert np.linalg.norm(As @ got - b) <= 1e-5 * np.linalg.norm(abs(As) @ np.abs(got))
    assert np.abs(got - xs).max() <= fwd_tol * np.abs(xs).max()
    # iteration count of the FP64 restatement on the same (global-order) system
    rows = np.repeat(np.arange(n), np.diff(rp))
    l2g = m.array("v_l2g")
    _, it64, s0, _ = o.mcf_cg(l2g[rows], l2g[col], val.astype(np.float64), _glob(m, mass.astype(np.float64)),
                              V.astype(np.float32).astype(np.float64), 500, 1e-12, 0.0)
    assert abs(it - it64) <= max(3, it64 // 5)
    assert abs(r0 - s0) <= 1e-4 * s0


def test_mcf_implicit_cg_stopping_rules():
    """The reference solver's exits, on the device: dt = 0 makes A = M, so x
    solves the system and no iteration runs; max_iter = 0 returns x; a
    relative tolerance stops on |r|^2 / |r0|^2 after as many iterations as
    the FP64 restatement takes (within 20%); a solve without statistics (no
    host synchronisation) writes the same bits as one with them."""
    import torch
    V, F, m, o = _mesh("torus")
    x = _lin(m, V.astype(np.float32))
    X, it, r0, r1 = m.mcf_implicit_cg(m.cot_assemble(x, 0.0), x, max_iter=50, abs_tol=1e-20)
    assert it == 0 and r0 == 0.0 and torch.equal(X, x)
    A = m.cot_assemble(x, 1e-2)
    X, it, r0, r1 = m.mcf_implicit_cg(A, x, max_iter=0, abs_tol=0.0)
    assert it == 0 and torch.equal(X, x) and r1 == r0 > 0
    X, it, r0, r1 = m.mcf_implicit_cg(A, x, max_iter=500, abs_tol=0.0, rel_tol=1e-6)
    assert 0 < it < 500 and r1 / r0 < 1e-6
    Y, _, _, _ = m.mcf_implicit_cg(A, x, max_iter=500, abs_tol=0.0, rel_tol=1e-6, stats=False)
    torch.cuda.synchronize()
    assert torch.equal(X, Y)
    rp, col, val, mass = (t.cpu().numpy() for t in A)
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    l2g = m.array("v_l2g")
    _, it64, _, _ = o.mcf_cg(l2g[rows], l2g[col], val.astype(np.float64), _glob(m, mass.astype(np.float64)),
                             V.astype(np.float32).astype(np.float64), 500, 0.0, 1e-6)
    assert abs(it - it64) <= max(3, it64 // 5)


@pytest.mark.parametrize("name,dt,uniform,precond,fwd_tol", [
    ("ico4", 1e-2, False, False, 1e-5), ("ico4", 1e-2, False, True, 1e-5), ("torus", 1e-2, False, True, 1e-5),
    ("ico4", 10.0, False, False, 2e-4), ("ico5_perm", 1e-2, True, False, 1e-5), ("disk", 1.0, True, True, 1e-5)])
def test_mcf_matfree_cg_vs_direct_solve(name, dt, uniform, precond, fwd_tol):
    """The matrix-free implicit step (weights recomputed from the coordinates
    in every product; the reference's matvec / precond_matvec under its CG /
    PCG) solves the FP64 system of the same weights -- cotangent or uniform
    -- to fp32 accuracy: backward error below 1e-5, within fwd_tol of the
    FP64 direct solution (the tolerances of the assembled CG), about as many
    iterations as the FP64 restatement of the same (P)CG, two solves
    bit-identical; for cotangent weights it agrees with the assembled-system
    CG to the same tolerance."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    import torch
    V, F, m, o = _mesh(name)
    x = _lin(m, V.astype(np.float32))
    kw = dict(uniform=uniform, precond=precond, max_iter=500, abs_tol=1e-12)
    X, it, r0, r1 = m.mcf_matfree_cg(x, x, dt, **kw)
    X2, it2, _, r2 = m.mcf_matfree_cg(x, x, dt, **kw)
    torch.cuda.synchronize()
    assert torch.equal(X, X2) and it == it2 and r1 == r2
    V64 = V.astype(np.float32).astype(np.float64)
    rows, cols, vals, mass = o.uniform_assemble(dt) if uniform else o.cot_assemble(V64, dt)
    n = len(V)
    As = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    b = mass[:, None] * V64
    xs = sla.spsolve(As.tocsc(), b)
    got = _glob(m, X.cpu().numpy()).T.astype(np.float64)
    assert r1 < 1e-12 and 0 < it < 500
    assert np.linalg.norm(As @ got - b) <= 1e-5 * np.linalg.norm(abs(As) @ np.abs(got))
    assert np.abs(got - xs).max() <= fwd_tol * np.abs(xs).max()
    _, it64, s0, _ = o.mcf_cg(rows, cols, vals, mass, V64, 500, 1e-12, 0.0, precond=precond)
    assert abs(it - it64) <= max(3, it64 // 5)
    assert abs(r0 - s0) <= 1e-4 * s0
    if not uniform and not precond:
        Xa, ita, _, _ = m.mcf_implicit_cg(m.cot_assemble(x, dt), x, max_iter=500, abs_tol=1e-12)
        torch.cuda.synchronize()
        assert abs(ita - it) <= max(2, it // 10)
        assert float((Xa - X).abs().max()) <= fwd_tol * float(Xa.abs().max())


def test_mcf_matfree_cg_stopping_rules_and_refusals():
    """dt = 0 with cotangent weights gives A = M, so x solves A X = M x
    exactly and no iteration runs; max_iter = 0 returns x; a relative
    tolerance stops on the residual ratio; cotangent weights on a mesh with
    boundary rings are refused (unsupported) before any launch."""
    import torch
    import rxmesh_amd as rx
    V, F, m, o = _mesh("torus")
    x = _lin(m, V.astype(np.float32))
    X, it, r0, r1 = m.mcf_matfree_cg(x, x, 0.0, max_iter=50, abs_tol=1e-20)
    assert it == 0 and r0 == 0.0 and torch.equal(X, x)
    X, it, r0, r1 = m.mcf_matfree_cg(x, x, 1e-2, precond=True, max_iter=0, abs_tol=0.0)
    assert it == 0 and torch.equal(X, x) and r1 == r0 > 0
    X, it, r0, r1 = m.mcf_matfree_cg(x, x, 1e-2, precond=True, max_iter=500, abs_tol=0.0, rel_tol=1e-6)
    assert 0 < it < 500 and r1 / r0 < 1e-6
    V, F, m, o = _mesh("disk")
    x = _lin(m, V.astype(np.float32))
    with pytest.raises(rx.RXMError) as e:
        m.mcf_matfree_cg(x, x, 1e-2)
    assert e.value.code == 6


def _rot(axis, angle):
    import math
    a = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


@pytest.mark.parametrize("name", ["ico4", "torus", "disk"])
def test_arap_weights_vs_fp64(name):
    """ARAP edge weights per edge within 1e-5 x max(1, |w64|) of the FP64
    oracle (fp32 cotangents of the same fp32 positions); boundary edges 0."""
    import torch
    V, F, m, o = _mesh(name)
    V32 = V.astype(np.float32)
    w = m.arap_weights(_lin(m, V32))
    torch.cuda.synchronize()
    got = _glob(m, w.cpu().numpy().astype(np.float64), "e_l2g")
    ref = o.arap_weights(V32.astype(np.float64))
    assert np.all(np.abs(got - ref) <= 1e-5 * np.maximum(1.0, np.abs(ref)))
    assert np.array_equal(got == 0.0, ref == 0.0) or name == "disk"
    assert torch.equal(w, m.arap_weights(_lin(m, V32)))


@pytest.mark.parametrize("name", ["ico4", "torus", "disk"])
def test_arap_rotations_energy_grad_vs_fp64(name):
    """Rotations (Horn's quaternion form on the device, LAPACK SVD in the
    oracle) agree to 1e-4 per entry wherever the covariance fixes R well
    (singular gap sigma2 + sign(det) sigma3 > 1e-2 sigma1: 95%+ of vertices
    here); the energy (relative 1e-5) and gradient (1e-4 of its largest
    entry) agree with FP64 on the device's rotations; reruns bit-identical."""
    import torch
    V, F, m, o = _mesh(name)
    rng = np.random.default_rng(7)
    V32 = V.astype(np.float32)
    X32 = (V + 0.3 * np.median(np.linalg.norm(V[F[:, 0]] - V[F[:, 1]], axis=1))
           * rng.standard_normal(V.shape)).astype(np.float32)
    P, X = _lin(m, V32), _lin(m, X32)
    w = m.arap_weights(P)
    R = m.arap_rotations(P, X, w)
    E, g = m.arap_energy_grad(P, X, w, R)
    torch.cuda.synchronize()
    w64 = o.arap_weights(V32.astype(np.float64))
    R64 = o.arap_rotations(V32.astype(np.float64), X32.astype(np.float64), w64)
    Rg = _glob(m, R.cpu().numpy().astype(np.float64).T).T.reshape(-1, 3, 3)
    S = o.arap_covariances(V32.astype(np.float64), X32.astype(np.float64), w64)
    sig = np.linalg.svd(S, compute_uv=False)
    gap = sig[:, 1] + np.sign(np.linalg.det(S)) * sig[:, 2]
    ok = gap > 1e-2 * sig[:, 0]
    assert ok.mean() > 0.95
    assert np.abs(Rg - R64)[ok].max() <= 1e-4
    assert np.abs(np.einsum("vij,vkj->vik", Rg, Rg) - np.eye(3)).max() <= 1e-5
    wg = _glob(m, w.cpu().numpy().astype(np.float64), "e_l2g")
    E64, g64 = o.arap_energy_grad(V32.astype(np.float64), X32.astype(np.float64), wg, Rg)
    assert abs(float(E.cpu()[0]) - E64) <= 1e-5 * E64
    gg = _glob(m, g.cpu().numpy().astype(np.float64)).T
    assert np.abs(gg - g64).max() <= 1e-4 * np.abs(g64).max()
    R2 = m.arap_rotations(P, X, w)
    E2, g2 = m.arap_energy_grad(P, X, w, R2)
    assert torch.equal(R, R2) and torch.equal(E, E2) and torch.equal(g, g2)


def test_arap_rigid_motion():
    """X = Q P + t: every rotation is Q to 1e-5, the energy and gradient
    vanish up to fp32 rounding of the edge vectors."""
    import torch
    V, F, m, o = _mesh("ico4")
    Q = _rot([0.3, -1.0, 0.5], 0.9)
    X = V @ Q.T + np.array([0.2, -0.1, 0.4])
    P, Xd = _lin(m, V.astype(np.float32)), _lin(m, X.astype(np.float32))
    w = m.arap_weights(P)
    R = m.arap_rotations(P, Xd, w)
    E, g = m.arap_energy_grad(P, Xd, w, R)
    torch.cuda.synchronize()
    Rg = R.cpu().numpy().reshape(-1, 3, 3)
    assert np.abs(Rg - Q).max() <= 1e-5
    _, g0 = m.arap_energy_grad(P, P, w, torch.zeros_like(R))  # R = 0: the gradient of sum w |X_i - X_j|^2
    scale = float(torch.abs(g0).max())
    assert float(torch.abs(g).max()) <= 1e-4 * scale
    assert float(E.cpu()[0]) <= 1e-9 * float(m.arap_energy_grad(P, P, w, torch.zeros_like(R))[0].cpu()[0])


@pytest.mark.parametrize("seed", range(4))
def test_random_meshes_numerics_vs_fp64(seed):
    """Seeded sweep over permuted meshes with random patch sizes: cotangent
    system (pattern exact, values 1e-5 of the largest), vertex normals and
    four smoothing iterations (1e-5 per vertex, norm-wise), Gaussian
    curvature (defect 1e-5 rad, mixed area 1e-5 relative) and ARAP weights
    (1e-5 x max(1, |w|)) against the FP64 oracle."""
    import torch
    import rxmesh_amd as rx
    rng = np.random.default_rng(2000 + seed)
    if seed % 2 == 0:
        V, F = mg.icosphere(int(rng.integers(3, 5)), seed=seed, sigma=0.01)
    else:
        V, F = mg.disk(int(rng.integers(24, 48)), seed=seed)
    V, F = mg.permute_mesh(V, F, seed=seed)
    ps = int(rng.integers(96, 513))
    print("case", seed, len(F), ps, flush=True)
    m = rx.Mesh(F, patch_size=ps)
    o = OracleMesh(F).set_patches(m.array("face_patch"), m.array("patches_val"), m.array("patches_offset"))
    V32 = V.astype(np.float32)
    V64 = V32.astype(np.float64)
    x = _lin(m, V32)
    rp, col, val, mass = (t.cpu().numpy() for t in m.cot_assemble(x, 1e-3))
    orows, ocols, ovals, omass = o.cot_assemble(V64, 1e-3)
    g2l = m.array("v_g2l")
    order = np.argsort(g2l[orows], kind="stable")
    assert np.array_equal(col, g2l[ocols[order]])
    assert np.abs(val - ovals[order]).max() <= 1e-5 * np.abs(ovals).max()
    assert np.abs(_glob(m, mass) - omass).max() <= 1e-5 * np.abs(omass).max()

    def rel(got, ref):
        return (np.linalg.norm(got - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-30)).max()

    n = _glob(m, m.vertex_normals(x, "max").cpu().numpy()).T
    assert rel(n, o.vertex_normals(V64, "max")) <= 1e-5
    s = _glob(m, m.smooth(x, 0.5, 4).cpu().numpy()).T
    assert rel(s, o.smooth_uniform(V64, 0.5, 4)) <= 1e-5
    gc, am = (t.cpu().numpy() for t in m.gaussian_curvature(x))
    rg, ra = o.gaussian_curvature(V64)
    assert np.abs(_glob(m, gc * am) - rg * ra).max() <= 1e-5
    assert np.abs(_glob(m, am) - ra).max() <= 1e-5 * np.abs(ra).max()
    w = _glob(m, m.arap_weights(x).cpu().numpy().astype(np.float64), "e_l2g")
    wr = o.arap_weights(V64)
    assert np.all(np.abs(w - wr) <= 1e-5 * np.maximum(1.0, np.abs(wr)))
