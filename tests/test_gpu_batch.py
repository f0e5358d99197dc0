"""GPU parity of batched execution: a batch is the disjoint union of its
meshes, run in one launch per kernel (or a persistent work-queue launch), its
step loop captured in a hipGraph; results must equal per-mesh runs bit for
bit and the FP64 oracle within tolerance."""
import numpy as np
import pytest

from oracle.oracle import OracleMesh
from rxmesh_amd import meshgen as mg

pytestmark = pytest.mark.gpu

LAM, ETA = 0.5, 1e-3


def _meshes():
    return [mg.icosphere(3, seed=1, sigma=0.01), mg.grid(20, seed=2), mg.icosphere(4, seed=3, sigma=0.005),
            mg.disk(24, seed=4), mg.grid(33, seed=5)]


_C = {}


def setup():
    if "b" in _C:
        return _C["b"]
    import torch
    import rxmesh_amd as rx
    ms = _meshes()
    b = rx.Mesh(None, batch=[f for _, f in ms], patch_size=128)
    sep = [rx.Mesh(f, patch_size=128) for _, f in ms]
    ov = b.batch_offsets("V")
    X = np.concatenate([p for p, _ in ms]).astype(np.float32)
    l2g = b.array("v_l2g")
    x = torch.from_numpy(np.ascontiguousarray(X[l2g].T)).cuda()
    _C["b"] = (ms, b, sep, ov, X, x)
    return _C["b"]


def lin(m, X):
    import torch
    return torch.from_numpy(np.ascontiguousarray(X[m.array("v_l2g")].T.astype(np.float32))).cuda()


def glob(m, Y):
    out = np.empty_like(Y)
    out[:, m.array("v_l2g")] = Y
    return out


def test_union_queries_equal_per_mesh_queries():
    import torch
    ms, b, sep, ov, X, x = setup()
    ofs = {"V": ov, "E": b.batch_offsets("E"), "F": b.batch_offsets("F")}
    for op in ("VV", "VF", "FF", "EV"):
        rp, col = (t.cpu().numpy() for t in b.query(op))
        src, dst = ofs[op[0]], ofs[op[1]]
        for k, m in enumerate(sep):
            mrp, mcol = (t.cpu().numpy() for t in m.query(op))
            r = rp[src[k]:src[k + 1] + 1]
            assert np.array_equal(r - r[0], mrp)
            assert np.array_equal(col[r[0]:r[-1]] - dst[k], mcol[: len(mcol)])


def test_symdir3_step_vs_fp64():
    import torch
    ms, b, sep, ov, X, x = setup()
    Mr = b.symdir_rest(x)
    # deform, then one gradient step
    Xd = (X.astype(np.float64) * 1.05 + 0.01 * np.sin(5 * X)).astype(np.float32)
    xd = lin(b, Xd)
    out = b.symdir3_grad_step(Mr, xd, ETA)
    e = b.symdir3_energy(Mr, xd)
    torch.cuda.synchronize()
    fv_all = np.concatenate([f + ov[k] for k, (_, f) in enumerate(ms)])
    o = OracleMesh(fv_all).set_patches(b.array("face_patch"), b.array("patches_val"), b.array("patches_offset"))
    gMr = np.empty((4, b.nF), np.float32)
    gMr[:, b.array("f_l2g")] = Mr.cpu().numpy()
    g64 = o.symdir3_grad(gMr.T.astype(np.float64), Xd.astype(np.float64))
    got = glob(b, out.cpu().numpy()).T
    step = (Xd.astype(np.float64) - got) / ETA
    assert np.abs(step - g64).max() <= 1e-3 * np.abs(g64).max() + 2e-7 / ETA
    E64 = o.symdir3_energy(gMr.T.astype(np.float64), Xd.astype(np.float64))
    gE = np.empty(b.nF, np.float32)
    gE[b.array("f_l2g")] = e.cpu().numpy()
    assert np.abs(gE - E64).max() <= 1e-5 * np.abs(E64).max()


def test_symdir3_grad_vs_fp64_every_mesh():
    """The gradient itself (rxm_symdir3_grad) on the deformed batch: within
    1e-5 of each mesh's largest FP64 gradient entry, on every mesh, and equal
    to what the gradient step applies."""
    import torch
    ms, b, sep, ov, X, x = setup()
    Mr = b.symdir_rest(x)
    Xd = (X.astype(np.float64) * 1.05 + 0.01 * np.sin(5 * X)).astype(np.float32)
    xd = lin(b, Xd)
    g = b.symdir3_grad(Mr, xd)
    step = b.symdir3_grad_step(Mr, xd, ETA)
    torch.cuda.synchronize()
    # the step kernel may contract x - eta g into one fma: equal to an ulp or two
    assert torch.allclose(step, xd - ETA * g, rtol=0, atol=2.5e-7 * float(xd.abs().max()))
    fv_all = np.concatenate([f + ov[k] for k, (_, f) in enumerate(ms)])
    o = OracleMesh(fv_all).set_patches(b.array("face_patch"), b.array("patches_val"), b.array("patches_offset"))
    gMr = np.empty((4, b.nF), np.float32)
    gMr[:, b.array("f_l2g")] = Mr.cpu().numpy()
    g64 = o.symdir3_grad(gMr.T.astype(np.float64), Xd.astype(np.float64))
    got = glob(b, g.cpu().numpy()).T
    for k in range(len(ms)):
        sl = slice(ov[k], ov[k + 1])
        assert np.abs(got[sl] - g64[sl]).max() <= 1e-5 * np.abs(g64[sl]).max(), k


def test_batched_steps_vs_fp64_oracle():
    import torch
    ms, b, sep, ov, X, x = setup()
    Mr = b.symdir_rest(x)
    xs = x.clone()
    tmp = torch.empty_like(xs)
    b.smooth_symdir_steps(Mr, xs, tmp, LAM, ETA, 5)
    torch.cuda.synchronize()
    got = glob(b, xs.cpu().numpy()).T
    fv_all = np.concatenate([f + ov[k] for k, (_, f) in enumerate(ms)])
    o = OracleMesh(fv_all).set_patches(b.array("face_patch"), b.array("patches_val"), b.array("patches_offset"))
    gMr = np.empty((4, b.nF), np.float32)
    gMr[:, b.array("f_l2g")] = Mr.cpu().numpy()
    ref = o.smooth_symdir_steps(gMr.T.astype(np.float64), X.astype(np.float64), LAM, ETA, 5)
    err = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert err.max() <= 1e-5, err.max()


def test_batched_equals_plain_per_mesh_bitwise():
    import torch
    import rxmesh_amd as rx
    ms, b, sep, ov, X, x = setup()
    Mr = b.symdir_rest(x)
    xb = x.clone()
    tb = torch.empty_like(xb)
    b.smooth_symdir_steps(Mr, xb, tb, LAM, ETA, 4)
    xs, ts, Mrs = [], [], []
    for k, m in enumerate(sep):
        xk = lin(m, X[ov[k]:ov[k + 1]])
        Mrs.append(m.symdir_rest(xk))
        xs.append(xk.clone())
        ts.append(torch.empty_like(xk))
    rx.smooth_symdir_steps_each(sep, Mrs, xs, ts, LAM, ETA, 4)
    torch.cuda.synchronize()
    gb = glob(b, xb.cpu().numpy())
    for k, m in enumerate(sep):
        assert np.array_equal(gb[:, ov[k]:ov[k + 1]], glob(m, xs[k].cpu().numpy()))


def test_graph_replays_match_eager():
    # every replay checked against the eager loop, with eager work and
    # synchronisations between replays
    import torch
    ms, b, sep, ov, X, x = setup()
    Mr = b.symdir_rest(x)
    xe = x.clone()
    te = torch.empty_like(xe)
    xg = x.clone()
    tg = torch.empty_like(xg)
    g = b.capture_smooth_symdir(Mr, xg, tg, LAM, ETA, 3)
    for rep in range(4):
        b.smooth_symdir_steps(Mr, xe, te, LAM, ETA, 3)
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(xg, xe), rep
        _ = b.query("VV")  # unrelated eager work between replays
        torch.cuda.synchronize()
    g.close()


def test_work_queue_mode_equals_grid_mode():
    import torch
    ms, b, sep, ov, X, x = setup()
    Mr = b.symdir_rest(x)
    outs = []
    for mode, th in ((0, 256), (1, 256), (1, 128), (0, 128)):
        b.set_launch_threads(2, mode)
        b.set_launch_threads(1, th)
        xs = x.clone()
        tmp = torch.empty_like(xs)
        b.smooth_symdir_steps(Mr, xs, tmp, LAM, ETA, 3)
        outs.append(xs)
    b.set_launch_threads(2, 0)
    b.set_launch_threads(1, 256)
    torch.cuda.synchronize()
    for o in outs[1:]:
        assert torch.equal(o, outs[0])
    # graph capture of the work-queue launch (counter reset by a kernel node)
    b.set_launch_threads(2, 1)
    xg = x.clone()
    tg = torch.empty_like(xg)
    g = b.capture_smooth_symdir(Mr, xg, tg, LAM, ETA, 3)
    for rep in range(3):
        xg.copy_(x)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(xg, outs[0]), rep
    g.close()
    b.set_launch_threads(2, 0)


def test_query_step_loop_equals_plan_loop_and_queries():
    """The step loop with the per-step VV + VF query leaves the same positions
    as the loop without it, its CSRs equal rxm_query_vv_vf's, and the plain
    per-mesh form equals the batched one bit for bit."""
    import torch
    import rxmesh_amd as rx
    ms, b, sep, ov, X, x = setup()
    Mr = b.symdir_rest(x)
    xq, tq = x.clone(), torch.empty_like(x)
    vv, vf = b.vv_vf_buffers()
    b.query_smooth_symdir_steps(Mr, xq, tq, LAM, ETA, 3, vv, vf)
    xp, tp = x.clone(), torch.empty_like(x)
    b.smooth_symdir_steps(Mr, xp, tp, LAM, ETA, 3)
    rvv, rvf = b.vv_vf_buffers()
    b.query_vv_vf(vv=rvv, vf=rvf)
    torch.cuda.synchronize()
    assert torch.equal(xq, xp)
    for got, ref in ((vv, rvv), (vf, rvf)):
        assert torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1])
    xs, ts, Mrs, bufs = [], [], [], []
    for k, m in enumerate(sep):
        xk = lin(m, X[ov[k]:ov[k + 1]])
        Mrs.append(m.symdir_rest(xk))
        xs.append(xk.clone())
        ts.append(torch.empty_like(xk))
        bufs.append(m.vv_vf_buffers())
    rx.query_smooth_symdir_steps_each(sep, Mrs, xs, ts, bufs, LAM, ETA, 3)
    torch.cuda.synchronize()
    gb = glob(b, xq.cpu().numpy())
    for k, m in enumerate(sep):
        assert np.array_equal(gb[:, ov[k]:ov[k + 1]], glob(m, xs[k].cpu().numpy()))
        mvv, mvf = m.query_vv_vf()
        for got, ref in zip(bufs[k], (mvv, mvf)):
            assert torch.equal(got[0], ref[0]) and torch.equal(got[1][: ref[1].numel()], ref[1])


def test_query_step_loop_graph_replays_match_eager():
    """The per-step-query loop captured in a HIP graph (torch's capture of the
    C-ABI launches): every replay equals the eager loop, with unrelated eager
    work and synchronisations between replays."""
    import torch
    ms, b, sep, ov, X, x = setup()
    Mr = b.symdir_rest(x)
    xe, te = x.clone(), torch.empty_like(x)
    xg, tg = x.clone(), torch.empty_like(x)
    qe, qg = b.vv_vf_buffers(), b.vv_vf_buffers()
    run = lambda: b.query_smooth_symdir_steps(Mr, xg, tg, LAM, ETA, 2, *qg)  # noqa: E731
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        run()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    xg.copy_(x)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, capture_error_mode="thread_local"):
        run()
    torch.cuda.synchronize()
    for rep in range(4):
        b.query_smooth_symdir_steps(Mr, xe, te, LAM, ETA, 2, *qe)
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(xg, xe), rep
        for got, ref in zip(qg, qe):
            assert torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1]), rep
        _ = b.query("VF")  # unrelated eager work between replays
        torch.cuda.synchronize()
