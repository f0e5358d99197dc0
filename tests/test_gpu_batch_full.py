"""The batched workload at its full size: 256 meshes of 2k to 1.3M faces
(icospheres and jittered grids, about 48M faces in all) as one union, 20
steps of {VV + VF query, smoothing, symmetric-Dirichlet gradient step}
captured in a HIP graph. The FP64 oracle cannot run every member at this
size in seconds, so the checks are size-independent properties plus the
oracle on a sample of members:
  - the batched loop equals plain per-mesh launches bit for bit on a sample
    (the smallest, the largest and evenly spaced members);
  - every position is finite after the 20 steps;
  - gradient descent alone (lambda 0) lowers the total energy;
  - the batched 3-D gradient is within 1e-5 of each sampled member's largest
    FP64 gradient entry."""
import numpy as np
import pytest

from oracle.oracle import OracleMesh
from rxmesh_amd import meshgen as mg

pytestmark = pytest.mark.gpu

LAM, ETA, STEPS = 0.5, 1e-3, 20
_C = {}


def setup():
    if "s" in _C:
        return _C["s"]
    import torch
    import rxmesh_amd as rx
    specs = mg.batch_specs(count=256)
    geo = [mg.make_batch_mesh(s, seed=k) for k, s in enumerate(specs)]
    sizes = np.array([len(f) for _, f in geo])
    print("batch: %d meshes, %d faces" % (len(geo), sizes.sum()), flush=True)
    u = rx.Mesh(None, batch=[f for _, f in geo], patch_size=512)
    ov = u.batch_offsets("V")
    X = np.concatenate([p for p, _ in geo]).astype(np.float32)
    x = torch.from_numpy(np.ascontiguousarray(X[u.array("v_l2g")].T)).cuda()
    sample = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), *range(0, 256, 37)})
    _C["s"] = (geo, u, ov, X, x, sample)
    return _C["s"]


def _lin(m, X):
    import torch
    return torch.from_numpy(np.ascontiguousarray(X[m.array("v_l2g")].T.astype(np.float32))).cuda()


def _glob(m, Y):
    out = np.empty_like(Y)
    out[:, m.array("v_l2g")] = Y
    return out


def test_full_batch_graph_equals_plain_on_sample_and_stays_finite():
    import torch
    import rxmesh_amd as rx
    geo, u, ov, X, x, sample = setup()
    Mr = u.symdir_rest(x)
    xb, tb = x.clone(), torch.empty_like(x)
    qb = u.vv_vf_buffers()
    run = lambda: u.query_smooth_symdir_steps(Mr, xb, tb, LAM, ETA, STEPS, *qb)  # noqa: E731
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        run()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    xb.copy_(x)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, capture_error_mode="thread_local"):
        run()
    g.replay()
    torch.cuda.synchronize()
    assert bool(torch.isfinite(xb).all())
    gb = _glob(u, xb.cpu().numpy())
    parts = [rx.Mesh(geo[k][1], patch_size=512) for k in sample]
    xs = [_lin(m, geo[k][0]) for m, k in zip(parts, sample)]
    Mrs = [m.symdir_rest(xk) for m, xk in zip(parts, xs)]
    ts = [torch.empty_like(xk) for xk in xs]
    bufs = [m.vv_vf_buffers() for m in parts]
    rx.query_smooth_symdir_steps_each(parts, Mrs, xs, ts, bufs, LAM, ETA, STEPS)
    torch.cuda.synchronize()
    for m, k, xk in zip(parts, sample, xs):
        assert np.array_equal(gb[:, ov[k]:ov[k + 1]], _glob(m, xk.cpu().numpy())), k
    for m in parts:
        m.close()


def test_full_batch_descent_lowers_energy():
    import torch
    geo, u, ov, X, x, sample = setup()
    Mr = u.symdir_rest(x)
    gen = torch.Generator(device="cuda").manual_seed(7)
    x0 = x + 1e-4 * torch.randn(x.shape, generator=gen, device="cuda")  # below every member's edge length
    e0 = float(u.symdir3_energy(Mr, x0).double().sum())
    xd, td = x0.clone(), torch.empty_like(x0)
    u.smooth_symdir_steps(Mr, xd, td, 0.0, ETA, 5)  # lambda 0: gradient descent only
    e1 = float(u.symdir3_energy(Mr, xd).double().sum())
    assert np.isfinite(e0) and np.isfinite(e1)
    assert e1 < e0, (e0, e1)


def test_full_batch_gradient_vs_fp64_on_sample():
    import torch
    geo, u, ov, X, x, sample = setup()
    Mr = u.symdir_rest(x)
    Xd = (X.astype(np.float64) * 1.05 + 0.01 * np.sin(5 * X)).astype(np.float32)
    g = u.symdir3_grad(Mr, _lin(u, Xd))
    torch.cuda.synchronize()
    got = _glob(u, g.cpu().numpy()).T
    gMr = np.empty((4, u.nF), np.float32)
    gMr[:, u.array("f_l2g")] = Mr.cpu().numpy()
    of = u.batch_offsets("F")
    for k in sample:
        _, f = geo[k]
        o = OracleMesh(f)
        sl, fs = slice(ov[k], ov[k + 1]), slice(of[k], of[k + 1])
        g64 = o.symdir3_grad(gMr[:, fs].T.astype(np.float64), Xd[sl].astype(np.float64))
        assert np.abs(got[sl] - g64).max() <= 1e-5 * np.abs(g64).max(), k
