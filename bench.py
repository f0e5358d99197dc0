"""Benchmark of the patch query engine on one MI355X.

Default workload (one "step"): on the icosphere subdivided 9 times
(2,621,442 V / 5,242,880 F, radial noise N(0, 0.005), vertex and face ids
randomly permuted, patch size 512), the VV and VF queries materialised as
int32 CSRs over linear ids by one launch (rxm_query_vv_vf).
value = VV + VF entries written per second (Gentries/s) by that launch,
inputs resident in HBM. Beside it: the query + smoothing step (the same
query, then 10 Jacobi iterations of uniform Laplacian smoothing, lambda 0.5,
fp32, over the engine's patch-local VV plan), the eight queries one by one,
per-kernel timings with their roofline, the CPU port of the same work on the
host, the other workloads and the 256-mesh batch.

Prints ONE JSON line. --gpus N > 1 is rejected (single-GPU build).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2s_perm")
    ap.add_argument("--patch-size", type=int, default=512)
    ap.add_argument("--patcher", default="bisect", choices=("kmeans", "bisect"))
    ap.add_argument("--smooth-iters", type=int, default=10)
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the c0 / c1 / c2s / c2t / c3 workloads")
    ap.add_argument("--cpu-reps", type=int, default=5, help="CPU baseline runs per thread count (median)")
    ap.add_argument("--profile", default=None,
                    help="run only this kernel --reps times (for rocprofv3): query_vv, query_vf, "
                         "smooth_uniform, smooth_2step (with --smooth-fused 1), smooth_ksteps (with --smooth-steps k), "
                         "vertex_normals, gaussian_curvature, cot_assembly, mcf_step, mcf_cg (20 CG iterations, dt 10), "
                         "query_evd, query_vv_vf, arap_weights, arap_rotations, arap_energy_grad, copy")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--query-sort", type=int, default=None,
                    help="--profile: counting-sort form of the queries (0 LDS atomics, 1 atomic-free, 2 one pass for VV + VF; default: library's)")
    ap.add_argument("--smooth-threads", type=int, default=None, help="smoothing workgroup size (default: autotuned)")
    ap.add_argument("--smooth-layout", type=int, default=None, help="smoothing LDS layout 0/1 (default: autotuned)")
    ap.add_argument("--smooth-fused", type=int, default=None,
                    help="1: two Jacobi steps per smoothing launch (default: autotuned)")
    ap.add_argument("--smooth-steps", type=int, default=None,
                    help="1-6: Jacobi steps per launch over the K-level plan, 0: off (default: autotuned)")
    ap.add_argument("--batch-count", type=int, default=256, help="meshes in the batched workload (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--batch-patch-size", type=int, default=None, help="batch patch size (default: autotuned)")
    return ap.parse_args()


def measured_traffic(workload, patch_size, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE + WRITE_SIZE passes, each scaled by
    a plain copy's reading), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "pmc_%s.md" % workload)
    if patch_size != 512 or not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return int(f[2])
    return None


def event_time_ms(fn, reps, groups=5):
    """Duration of one fn() call on the device: the reps calls are queued in
    `groups` runs with a HIP event (on the current stream, the stream every
    rxmesh_amd launch uses here) between runs; the figure is the median over
    runs of a run's time / its calls. The host queues the calls ahead of the
    device, so a run's interval is device time; the median keeps a host stall
    (the Python collector, a descheduled thread) that leaves the device idle
    inside one run out of the figure, and an event per run rather than per
    call keeps the events' own cost out of it. The collector is off while
    the calls are queued."""
    import gc
    import numpy as np
    import torch
    s = torch.cuda.current_stream()
    groups = max(1, min(groups, reps))
    per = max(1, reps // groups)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(groups + 1)]
    fn()
    torch.cuda.synchronize()
    was = gc.isenabled()
    gc.disable()
    try:
        ev[0].record(s)
        for g in range(groups):
            for _ in range(per):
                fn()
            ev[g + 1].record(s)
        ev[-1].synchronize()
    finally:
        if was:
            gc.enable()
    return float(np.median([ev[g].elapsed_time(ev[g + 1]) / per for g in range(groups)]))


def cpu_baseline(reps, m, fv, pos, iters, gpu_entries, build_s):
    """The CPU port (oracle/cpu_port.cpp: OpenMP, fp32, int32 CSRs over
    global ids; kind "port") of the same work on the bench mesh itself: VV +
    VF CSRs from the edge list and fv, then `iters` Jacobi smoothing
    iterations over the VV CSR, at one thread and at every core this process
    may use (OMP_NUM_THREADS), each the median of `reps` runs after a warm-up.
    It is compiled for this host (-O3 -march=native -fopenmp) when a compiler
    is available. The host connectivity build (the reference's CPU build
    restated, on all host threads) is reported beside it. A baseline, not a
    target."""
    import numpy as np
    from oracle import cpu_port as cp
    L, flags = cp.load(native=True)
    cores = L.cpu_max_threads()
    ev = m.array("ev")
    l2g = m.array("v_l2g")
    x = np.ascontiguousarray(pos.T.astype(np.float32))  # global (input) vertex order
    res = {}
    for th in sorted({1, cores}):
        r = reps if th > 1 else 1  # one timed run after the warm-up: ~10 s on one core
        tq, vv, vf = cp.query_vv_vf(L, m.nV, ev, fv, th, r)
        ts, _ = cp.smooth(L, vv, x, 0.5, iters, th, r)
        res[th] = (tq, ts)
    n = len(vv[1]) + len(vf[1])
    assert n == gpu_entries, (n, gpu_entries)
    del l2g
    tq, ts = res[cores]
    return {"value": round(n / tq / 1e9, 4), "unit": "Gentries/s", "cores": cores, "kind": "port",
            "sample": "CPU port, VV + VF CSRs (%d entries) of the bench mesh after one warm-up run: median "
                      "of %d runs at %d threads, one run at 1 thread; %s; %s" % (n, reps, cores, flags, cp.cpu_model()),
            "cpu_model": cp.cpu_model(), "nproc": os.cpu_count(), "build": flags,
            "query_ms": {str(k): round(v[0] * 1e3, 2) for k, v in res.items()},
            "value_1thread": round(n / res[1][0] / 1e9, 4),
            "smooth_ms": {str(k): round(v[1] * 1e3, 2) for k, v in res.items()},
            "step_query_smooth_ms": {str(k): round((v[0] + v[1]) * 1e3, 2) for k, v in res.items()},
            "host_build_s": round(build_s, 3)}


def batch_bench(count, steps, reps, patch_size=None):
    """Batched workload: `count` meshes (face counts log-uniform in
    [2e3, 1.3e6]; icospheres L3-L8 and jittered grids 32..800 alternating),
    `steps` iterations of {VV + VF query CSRs, uniform smoothing step
    (lambda 0.5), symmetric-Dirichlet gradient step (eta 1e-3)}.
    Batched = the union of all meshes, three launches per step, the whole
    loop captured in one hipGraph; plain = separate per-mesh launches from
    one C loop (also timed as a captured graph). The same loop without the
    per-step query (the kernels then read the relations the engine builds
    once per mesh) is timed beside it. Patch size comes from the autotuner
    (timed on every 8th mesh of the batch) unless `patch_size` is given."""
    import numpy as np
    import torch
    import rxmesh_amd as rx
    from rxmesh_amd import autotune
    from rxmesh_amd import meshgen as mg
    t0 = time.perf_counter()
    log = lambda msg: print("[batch %.1fs] %s" % (time.perf_counter() - t0, msg), file=sys.stderr, flush=True)  # noqa: E731
    specs = mg.batch_specs(count=count)
    geo = [mg.make_batch_mesh(s, seed=k) for k, s in enumerate(specs)]
    log("generated %d meshes, %d faces" % (count, sum(len(f) for _, f in geo)))
    lam, eta = 0.5, 1e-3
    ps_times = {}
    if patch_size is None:
        sample = geo[::8]
        patch_size, ps_times = autotune.tune_batch_patch_size([f for _, f in sample], [p for p, _ in sample],
                                                              lam, eta)
        log("autotuned patch size %d: %s" % (patch_size, ps_times))
    t_setup = time.perf_counter()
    # the plain meshes: those above an even share of the faces first, one at a time with every
    # host thread, then the rest concurrently (the C-ABI releases the GIL), largest first, one
    # host thread each (serial_build: no nested thread pools)
    from concurrent.futures import ThreadPoolExecutor
    workers = min(16, os.cpu_count() or 1)
    order = sorted(range(count), key=lambda k: -len(geo[k][1]))
    total = sum(len(f) for _, f in geo)
    nbig = 0
    while nbig < count and workers > 1 and len(geo[order[nbig]][1]) * workers > total:
        nbig += 1
    built = {k: rx.Mesh(geo[k][1], patch_size=patch_size) for k in order[:nbig]}
    with ThreadPoolExecutor(max_workers=workers) as pool:
        built.update(zip(order[nbig:], pool.map(lambda k: rx.Mesh(geo[k][1], patch_size=patch_size,
                                                                  serial_build=True), order[nbig:])))
    parts = [built[k] for k in range(count)]
    log("built %d meshes" % count)
    union = rx.Mesh(None, parts=parts)
    log("union built: %d patches" % union.P)
    xs = [torch.from_numpy(np.ascontiguousarray(p[m.array("v_l2g")].T.astype(np.float32))).cuda()
          for m, (p, _) in zip(parts, geo)]
    x_u = torch.cat(xs, dim=1).contiguous()
    Mr_u = union.symdir_rest(x_u)
    Mrs = [m.symdir_rest(x) for m, x in zip(parts, xs)]
    for m in parts:
        m.build_vv_plan()
    union.build_vv_plan()
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t_setup
    # autotuned launch configuration of the union
    x_t = x_u.clone()
    t_t = torch.empty_like(x_t)
    cfg, _ = autotune.tune_smooth_symdir(union, Mr_u, x_t, t_t, lam, eta, use_cache=False)
    log("autotuned: %s" % cfg)
    del x_t, t_t

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - a) / reps

    def captured(fn):
        """fn() captured once into a torch CUDA graph (HIP graph) on a side stream."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            fn()  # everything built lazily is built before capture
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, capture_error_mode="thread_local"):
            fn()
        torch.cuda.synchronize()
        return g

    xb = x_u.clone()
    tb = torch.empty_like(xb)
    qb = union.vv_vf_buffers()
    xp = [x.clone() for x in xs]
    tp = [torch.empty_like(x) for x in xs]
    qp = [m.vv_vf_buffers() for m in parts]
    # with the per-step query
    run_u = lambda: union.query_smooth_symdir_steps(Mr_u, xb, tb, lam, eta, steps, *qb)  # noqa: E731
    run_p = lambda: rx.query_smooth_symdir_steps_each(parts, Mrs, xp, tp, qp, lam, eta, steps)  # noqa: E731
    g_u = captured(run_u)
    t_graph_q = timed(g_u.replay)
    t_eager_q = timed(run_u)
    log("with query: batched graph %.2f ms, eager %.2f ms" % (t_graph_q * 1e3, t_eager_q * 1e3))
    t_plain_q = timed(run_p)
    g_p = captured(run_p)
    t_plain_graph_q = timed(g_p.replay)
    log("with query: plain per-mesh %.2f ms, as a graph %.2f ms" % (t_plain_q * 1e3, t_plain_graph_q * 1e3))
    del g_u, g_p
    # without it: the engine's own graph capture of the step loop
    graph = union.capture_smooth_symdir(Mr_u, xb, tb, lam, eta, steps)
    t_graph = timed(graph.launch)
    t_eager = timed(lambda: union.smooth_symdir_steps(Mr_u, xb, tb, lam, eta, steps))
    t_plain = timed(lambda: rx.smooth_symdir_steps_each(parts, Mrs, xp, tp, lam, eta, steps))
    log("plan only: batched graph %.2f ms, eager %.2f ms, plain %.2f ms" % (t_graph * 1e3, t_eager * 1e3,
                                                                             t_plain * 1e3))
    graph.close()
    V, E, F = union.nV, union.nE, union.nF
    nnz_vv, nnz_vf = union.query_nnz("VV"), union.query_nnz("VF")
    # algorithmic bytes per step: query (EV u32, FE 3 x u16, both int32 CSRs), smoothing (positions read
    # and written, u16 VV plan), gradient step (rest shapes 16 B per face, positions read and written,
    # u16 face plan: FV and VF rows)
    b_query = 4 * E + 6 * F + 4 * nnz_vv + 4 * nnz_vf + 8 * (V + 1)
    b_smooth = 24 * V + 4 * E
    b_grad = 16 * F + 24 * V + 12 * F
    b_step = b_query + b_smooth + b_grad
    gbs = b_step * steps / t_graph_q / 1e9
    return {"meshes": count, "faces": F, "vertices": V, "patches": union.P, "steps": steps,
            "patch_size": patch_size, "patch_size_autotune_ms": ps_times,
            "setup_s": round(setup_s, 1), "autotune": cfg,
            "ms_batched_graph": round(t_graph_q * 1e3, 3), "ms_batched_eager": round(t_eager_q * 1e3, 3),
            "ms_plain_per_mesh": round(t_plain_q * 1e3, 3),
            "ms_plain_per_mesh_graph": round(t_plain_graph_q * 1e3, 3),
            "meshes_per_s_batched": round(count / t_graph_q, 2), "meshes_per_s_plain": round(count / t_plain_q, 2),
            "speedup_vs_plain": round(t_plain_q / t_graph_q, 3),
            "speedup_vs_plain_graph": round(t_plain_graph_q / t_graph_q, 3),
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_step": b_step},
            "plan_only": {"ms_batched_graph": round(t_graph * 1e3, 3), "ms_batched_eager": round(t_eager * 1e3, 3),
                          "ms_plain_per_mesh": round(t_plain * 1e3, 3),
                          "speedup_vs_plain": round(t_plain / t_graph, 3)}}


def _entry(t_ms, nbytes, unit_count=None, unit=None):
    r = {"us": round(t_ms * 1e3, 2), "algorithmic_bytes": int(nbytes),
         "GB_s": round(nbytes / (t_ms * 1e-3) / 1e9, 1),
         "frac": round(nbytes / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if unit_count is not None:
        r[unit] = round(unit_count / (t_ms * 1e-3) / 1e9, 4)
    return r


def configs_bench(reps, c2s_mesh=None, c2s_pos=None):
    """The other workloads of the suite, each measured with HIP events on
    the launch stream against the kernel's algorithmic bytes (fp32 = 4 B,
    int32 index = 4 B, patch-local u16 = 2 B):
      c0: area-weighted vertex normals + VV query on the jittered 128 x 128 grid;
      c1: the full query suite (VV, VE, VF, EV, EF, FV, FE, FF) and 10
          uniform smoothing iterations on the permuted icosphere L6;
      c2s / c2t: cotangent CSR assembly (A = M + dt L, oriented VV rows) and
          explicit mean-curvature flow x <- x + (M x - A x) / M, 50 SpMV-3
          iterations, dt = 1e-5, on the icosphere L9 and the 2048 x 1024 torus;
      c3: symmetric-Dirichlet energy and its adjoint gradient on the
          256 x 256 grid clipped to a disk, UV = rest x, y + N(0, 0.002)."""
    import numpy as np
    import torch
    import rxmesh_amd as rx
    from rxmesh_amd import meshgen as mg
    out = {}
    t0 = time.perf_counter()
    log = lambda msg: print("[configs %.1fs] %s" % (time.perf_counter() - t0, msg), file=sys.stderr, flush=True)  # noqa: E731

    def lin(m, pos):
        return torch.from_numpy(np.ascontiguousarray(pos[m.array("v_l2g")].T.astype(np.float32))).cuda()

    # c0
    pos, fv = mg.config_mesh("c0")
    m = rx.Mesh(fv)
    V, E, F = m.nV, m.nE, m.nF
    x = lin(m, pos)
    n = torch.empty_like(x)
    rp = torch.empty(V + 1, dtype=torch.int32, device="cuda")
    col = torch.empty(m.query_nnz("VV"), dtype=torch.int32, device="cuda")
    m.vertex_normals(x, "area", out=n)
    t_n = event_time_ms(lambda: m.vertex_normals(x, "area", out=n), reps)
    t_q = event_time_ms(lambda: m.query("VV", row_ptr=rp, col=col), reps)
    out["c0_grid128"] = {"V": V, "F": F, "normals_area": _entry(t_n, 24 * V + 4 * E + 6 * F, F, "Gfaces_s"),
                         "query_vv": _entry(t_q, 4 * E + 8 * E + 4 * (V + 1), 2 * E, "Gentries_s")}
    m.close()
    log("c0 done")

    # c1: every query (topology read + int32 CSR written, as in the main table) and the smoothing
    pos, fv = mg.config_mesh("c1")
    m = rx.Mesh(fv, patch_size=512)
    V, E, F = m.nV, m.nE, m.nF
    x = lin(m, pos)
    topo = {"VV": 4 * E, "VE": 4 * E, "VF": 4 * E + 6 * F, "FV": 4 * E + 6 * F, "FE": 6 * F, "FF": 6 * F,
            "EV": 4 * E, "EF": 6 * F}
    c1 = {"V": V, "F": F, "patches": m.P}
    for op in ("VV", "VE", "VF", "EV", "EF", "FV", "FE", "FF"):
        nnz, nsrc = m.query_nnz(op), m.num_src(op)
        rp_q = torch.empty(nsrc + 1, dtype=torch.int32, device="cuda")
        col_q = torch.empty(max(nnz, 1), dtype=torch.int32, device="cuda")
        t = event_time_ms(lambda: m.query(op, row_ptr=rp_q, col=col_q), reps)
        c1["query_" + op.lower()] = _entry(t, topo[op] + 4 * nnz + 4 * (nsrc + 1), nnz, "Gentries_s")
    y, tmp = torch.empty_like(x), torch.empty_like(x)
    from rxmesh_amd import autotune
    cfg, _ = autotune.tune_smooth(m, x, y, use_cache=False, iters=10)  # launch form for 10 iterations
    m.smooth(x, 0.5, 10, out=y, tmp=tmp)
    t_s = event_time_ms(lambda: m.smooth(x, 0.5, 10, out=y, tmp=tmp), reps)
    c1["smooth_10_iterations"] = dict(_entry(t_s, 10 * (24 * V + 4 * E), 10 * 2 * E, "Gentries_s"), config=cfg)
    out["c1_icosphere_L6"] = c1
    m.close()
    log("c1 done")

    def mcf(m, x, name):
        V, E, F = m.nV, m.nE, m.nF
        dt = 1e-5
        A = m.cot_assemble(x, dt)
        nnz = int(A[1].numel())
        t_a = event_time_ms(lambda: m.cot_assemble(x, dt, out=A), max(10, reps))
        y = torch.empty_like(x)
        tmp = torch.empty_like(x)
        sweep = {}
        for lanes in (1, 2, 4, 8):  # lanes per CSR row of the general SpMV-3: autotuned here
            m.set_launch_threads(5, lanes)
            m.spmv3(A, x, out=y)
            sweep[lanes] = event_time_ms(lambda: m.spmv3(A, x, out=y), reps)
        best = min(sweep, key=sweep.get)
        m.set_launch_threads(5, best)
        m.mcf_explicit(A, x, 50, out=y, tmp=tmp)  # patch-local steps over the VV plan
        t_m = event_time_ms(lambda: m.mcf_explicit(A, x, 50, out=y, tmp=tmp), max(3, reps // 10))
        # implicit step with the reference MCF app's CG defaults (dt 10, |r|^2 < 1e-6, <= 100 iterations)
        A10 = m.cot_assemble(x, 10.0)
        _, cg_it, cg_r0, cg_r1 = m.mcf_implicit_cg(A10, x, 100, 1e-6, 0.0, out=y)
        t_cg = event_time_ms(lambda: m.mcf_implicit_cg(A10, x, 100, 1e-6, 0.0, out=y, stats=False), 3)
        cg_bytes = 8 * nnz + 4 * (V + 1) + 124 * V  # one CG iteration: A p with the p update, then x / r update
        del A10
        # the EVDiamond query and the ARAP-style terms built on it (DiffARAP pieces)
        rp_d = torch.empty(E + 1, dtype=torch.int32, device="cuda")
        col_d = torch.empty(4 * E, dtype=torch.int32, device="cuda")
        m.query("EVD", row_ptr=rp_d, col=col_d)
        t_evd = event_time_ms(lambda: m.query("EVD", row_ptr=rp_d, col=col_d), reps)
        gen = torch.Generator(device="cuda").manual_seed(0)
        xd = x + 1e-3 * torch.randn(x.shape, generator=gen, device="cuda")
        w_a = m.arap_weights(x)
        R_a = m.arap_rotations(x, xd, w_a)
        tot_a, g_a = m.arap_energy_grad(x, xd, w_a, R_a)
        t_aw = event_time_ms(lambda: m.arap_weights(x, out=w_a), reps)
        t_ar = event_time_ms(lambda: m.arap_rotations(x, xd, w_a, out=R_a), reps)
        t_ae = event_time_ms(lambda: m.arap_energy_grad(x, xd, w_a, R_a, total=tot_a, grad=g_a), reps)
        del rp_d, col_d, xd, w_a, R_a, tot_a, g_a
        out[name] = {"V": V, "F": F, "nnz": nnz,
                     "cot_csr_assembly": _entry(t_a, 12 * V + 4 * E + 6 * F + 8 * nnz + 4 * (V + 1) + 4 * V,
                                                nnz, "Gnnz_s"),
                     "mcf_spmv3_iteration": _entry(t_m / 50, 8 * nnz + 4 * (V + 1) + 12 * V + 4 * V + 12 * V,
                                                   nnz, "Gnnz_s"),
                     "mcf_50_iterations_ms": round(t_m, 4),
                     "mcf_implicit_cg": {"ms": round(t_cg, 4), "iterations": cg_it, "start_residual": cg_r0,
                                         "final_residual": cg_r1,
                                         "per_iteration": _entry(t_cg / max(cg_it + 1, 1), cg_bytes, nnz, "Gnnz_s")},
                     "spmv3_csr": _entry(sweep[best], 8 * nnz + 4 * (V + 1) + 12 * V + 12 * V, nnz, "Gnnz_s"),
                     # EV 4E + FE 6F read; row_ptr 4(E+1) + 4 ids per edge written
                     "evd_query": _entry(t_evd, 4 * E + 6 * F + 4 * (E + 1) + 16 * E, 4 * E, "Gentries_s"),
                     # positions 12V + EV 4E + FE 6F read, one weight per edge written
                     "arap_weights": _entry(t_aw, 12 * V + 4 * E + 6 * F + 4 * E, E, "Gedges_s"),
                     # rest + deformed 24V, weights 4E, EV 4E, edge pattern 8E + 4V read; 9 floats per vertex written
                     "arap_rotations": _entry(t_ar, 24 * V + 16 * E + 4 * V + 36 * V, V, "Gverts_s"),
                     # as above plus rotations 36V read; gradient 12V written
                     "arap_energy_grad": _entry(t_ae, 24 * V + 16 * E + 4 * V + 36 * V + 12 * V, E, "Gedges_s"),
                     "spmv_lanes": best,
                     "spmv_lanes_sweep_us": {str(k): round(v * 1e3, 2) for k, v in sweep.items()}}
        del A, y, tmp

    if c2s_mesh is not None:
        mcf(c2s_mesh, c2s_pos, "c2s_icosphere_L9")
        log("c2s done")
    pos, fv = mg.config_mesh("c2t")
    m = rx.Mesh(fv)
    x = lin(m, pos)
    mcf(m, x, "c2t_torus_2048x1024")
    # connectivity build on the device (edge ids as the host build numbers them)
    fv_t = torch.from_numpy(np.ascontiguousarray(fv, dtype=np.int32)).cuda()
    rx.build_edges_device(fv_t, len(pos))
    t_eb = event_time_ms(lambda: rx.build_edges_device(fv_t, len(pos)), 3)
    out["c2t_torus_2048x1024"]["edge_build_device"] = {"ms": round(t_eb, 3), "Gfaces_s": round(len(fv) / (t_eb * 1e-3) / 1e9, 4),
                                                       "note": "whole call: work buffers, sort, scans, sync"}
    del fv_t
    del x
    m.close()
    log("c2t done")

    # c3
    pos, fv = mg.config_mesh("c3")
    m = rx.Mesh(fv)
    V, E, F = m.nV, m.nE, m.nF
    uv = mg.uv_init(pos, fv, seed=0)
    rest = lin(m, pos)
    u = torch.from_numpy(np.ascontiguousarray(uv[m.array("v_l2g")].T.astype(np.float32))).cuda()
    Mr = m.symdir_rest(rest)
    total, energy = m.symdir_energy(Mr, u)
    g = m.symdir_grad(Mr, u)
    t_e = event_time_ms(lambda: m.symdir_energy(Mr, u, energy=energy, total=total), reps)
    t_g = event_time_ms(lambda: m.symdir_grad(Mr, u, out=g), reps)
    topo = 6 * F + 4 * E
    out["c3_disk256"] = {"V": V, "F": F,
                         "symdir_energy": _entry(t_e, 16 * F + 8 * V + topo + 4 * F, F, "Gfaces_s"),
                         "symdir_grad": _entry(t_g, 16 * F + 8 * V + topo + 8 * V, F, "Gfaces_s")}
    m.close()
    log("c3 done")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import rxmesh_amd as rx
    from rxmesh_amd import meshgen as mg

    torch.cuda.set_device(0)
    pos, fv = mg.config_mesh(args.workload)
    # one-time costs of the process (library load, first launches of its
    # device code), timed once: a small mesh first, then the bench mesh
    t_first = time.perf_counter()
    rx.Mesh(mg.config_mesh("c1")[1], patch_size=args.patch_size, patcher=args.patcher).close()
    t_first = time.perf_counter() - t_first
    t_build = time.perf_counter()
    m = rx.Mesh(fv, patch_size=args.patch_size, patcher=args.patcher)  # host build + upload
    t_build = time.perf_counter() - t_build
    V, E, F, P = m.nV, m.nE, m.nF, m.P
    l2g = m.array("v_l2g")
    x = torch.from_numpy(np.ascontiguousarray(pos[l2g].T.astype(np.float32))).cuda()
    out = torch.empty_like(x)
    tmp = torch.empty_like(x)
    nnz_vv, nnz_vf = m.query_nnz("VV"), m.query_nnz("VF")
    vv_rp = torch.empty(V + 1, dtype=torch.int32, device="cuda")
    vv_col = torch.empty(nnz_vv, dtype=torch.int32, device="cuda")
    vf_rp = torch.empty(V + 1, dtype=torch.int32, device="cuda")
    vf_col = torch.empty(nnz_vf, dtype=torch.int32, device="cuda")
    it = args.smooth_iters
    m.build_vv_plan()  # static patch-local VV plan, built once per mesh
    torch.cuda.synchronize()
    from rxmesh_amd import autotune
    if (args.smooth_threads or args.smooth_layout is not None or args.smooth_fused is not None
            or args.smooth_steps is not None):
        smooth_cfg = {"smooth_threads": args.smooth_threads or 128,
                      "smooth_layout": args.smooth_layout if args.smooth_layout is not None else 0,
                      "smooth_fused": args.smooth_fused if args.smooth_fused is not None else 0,
                      "smooth_steps": args.smooth_steps if args.smooth_steps is not None else 0}
        autotune.apply(m, smooth_cfg)
        smooth_sweep = {}
    elif args.profile:
        smooth_cfg, smooth_sweep = {}, {}  # library defaults; --smooth-threads / --smooth-layout choose
    else:
        smooth_cfg, smooth_sweep = autotune.tune_smooth(m, x, out, use_cache=False, iters=it)
    query_sort, query_sort_us = autotune.tune_query_sort(m) if not args.profile else (None, {})
    torch.cuda.synchronize()

    if args.profile:
        if args.query_sort is not None:
            m.set_launch_threads(10, args.query_sort)
        A_cot = m.cot_assemble(x, 1e-5) if args.profile in ("cot_assembly", "mcf_step") else None
        A_cg = m.cot_assemble(x, 10.0) if args.profile == "mcf_cg" else None
        if args.profile.startswith("arap") or args.profile == "query_evd":
            evd_rp = torch.empty(E + 1, dtype=torch.int32, device="cuda")
            evd_col = torch.empty(4 * E, dtype=torch.int32, device="cuda")
            gen = torch.Generator(device="cuda").manual_seed(0)
            xd = x + 1e-3 * torch.randn(x.shape, generator=gen, device="cuda")
            w_a = m.arap_weights(x)
            R_a = m.arap_rotations(x, xd, w_a)
            tot_a, g_a = m.arap_energy_grad(x, xd, w_a, R_a)
        fns = {"query_vv": lambda: m.query("VV", row_ptr=vv_rp, col=vv_col),
               "cot_assembly": lambda: m.cot_assemble(x, 1e-5, out=A_cot),
               "mcf_step": lambda: m.mcf_explicit(A_cot, x, 1, out=out),
               "mcf_cg": lambda: m.mcf_implicit_cg(A_cg, x, 20, 1e-6, 0.0, out=out, stats=False),
               "query_evd": lambda: m.query("EVD", row_ptr=evd_rp, col=evd_col),
               "query_vv_vf": lambda: m.query_vv_vf(vv=(vv_rp, vv_col), vf=(vf_rp, vf_col)),
               "arap_weights": lambda: m.arap_weights(x, out=w_a),
               "arap_rotations": lambda: m.arap_rotations(x, xd, w_a, out=R_a),
               "arap_energy_grad": lambda: m.arap_energy_grad(x, xd, w_a, R_a, total=tot_a, grad=g_a),
               "query_vf": lambda: m.query("VF", row_ptr=vf_rp, col=vf_col),
               "smooth_uniform": lambda: m.smooth(x, 0.5, 1, out=out),
               "smooth_2step": lambda: m.smooth(x, 0.5, 2, out=out, tmp=tmp),
               "smooth_ksteps": lambda: m.smooth(x, 0.5, max(1, args.smooth_steps or 1), out=out, tmp=tmp),
               "smooth_csr": lambda: m.smooth_csr((vv_rp, vv_col), x, 0.5, 1, out=out),
               "vertex_normals": lambda: m.vertex_normals(x, "max", out=out),
               "gaussian_curvature": lambda: m.gaussian_curvature(x, out=(out[0], out[1])),
               "copy": lambda: out.copy_(x)}  # plain copy of 12 V bytes (counter calibration)
        fn = fns[args.profile]
        for _ in range(args.reps):
            fn()
        torch.cuda.synchronize()
        print(json.dumps({"profile": args.profile, "reps": args.reps, "V": V, "E": E, "F": F,
                          "copy_bytes": 12 * V}), flush=True)
        return

    def step():  # one pass of the hot path: the VV and VF CSRs from one launch (rxm_query_vv_vf)
        m.query_vv_vf(vv=(vv_rp, vv_col), vf=(vf_rp, vf_col))

    def step_smooth():  # the query step, then the smoothing iterations over the engine's VV plan
        m.query_vv_vf(vv=(vv_rp, vv_col), vf=(vf_rp, vf_col))
        m.smooth(x, 0.5, it, out=out, tmp=tmp)

    def step_separate():  # the same query + smoothing work with the two queries launched separately
        m.query("VV", row_ptr=vv_rp, col=vv_col)
        m.query("VF", row_ptr=vf_rp, col=vf_col)
        m.smooth(x, 0.5, it, out=out, tmp=tmp)

    def wall_ms(fn, steps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    ms = wall_ms(step, args.steps)
    entries = nnz_vv + nnz_vf
    value = entries / (ms * 1e-3) / 1e9
    ms_smooth = wall_ms(step_smooth, args.steps)
    ms_separate = wall_ms(step_separate, args.steps)

    # per-kernel live timing (HIP events on the launch stream) and algorithmic bytes per launch
    reps = args.kernel_reps
    kern = {}
    t_vvf = event_time_ms(step, reps)
    kern["query_vv_vf"] = (t_vvf, 4 * E + 6 * F + 4 * nnz_vv + 4 * nnz_vf + 8 * (V + 1))
    # the eight queries one by one: topology read (EV u32 per edge, FE 3 x u16 per face) + int32 CSR written
    topo = {"VV": 4 * E, "VE": 4 * E, "VF": 4 * E + 6 * F, "FV": 4 * E + 6 * F, "FE": 6 * F, "FF": 6 * F,
            "EV": 4 * E, "EF": 6 * F}
    queries = {}
    for op in ("VV", "VE", "VF", "EV", "EF", "FV", "FE", "FF"):
        nnz = m.query_nnz(op)
        nsrc = m.num_src(op)
        rp_q = torch.empty(nsrc + 1, dtype=torch.int32, device="cuda")
        col_q = torch.empty(max(nnz, 1), dtype=torch.int32, device="cuda")
        t_q = event_time_ms(lambda: m.query(op, row_ptr=rp_q, col=col_q), reps)
        nb = topo[op] + 4 * nnz + 4 * (nsrc + 1)
        queries[op] = dict(_entry(t_q, nb, nnz, "Gentries_s"), entries=nnz)
        kern["query_%s" % op.lower()] = (t_q, nb)
        del rp_q, col_q
    t_sm = event_time_ms(lambda: m.smooth(x, 0.5, 1, out=out), reps)
    kern["smooth_uniform"] = (t_sm, 24 * V + 4 * E)
    ksteps = int(smooth_cfg.get("smooth_steps", 0))
    fused = bool(smooth_cfg.get("smooth_fused", 0)) and not ksteps
    if fused:  # one launch = two Jacobi steps: twice the entries of one step
        t_s2 = event_time_ms(lambda: m.smooth(x, 0.5, 2, out=out, tmp=tmp), reps)
        kern["smooth_uniform_2step"] = (t_s2, 2 * (24 * V + 4 * E))
    if ksteps:  # one launch = ksteps Jacobi steps over the K-level plan
        t_sk = event_time_ms(lambda: m.smooth(x, 0.5, ksteps, out=out, tmp=tmp), reps)
        kern["smooth_uniform_%dstep" % ksteps] = (t_sk, ksteps * (24 * V + 4 * E))
    t_sc = event_time_ms(lambda: m.smooth_csr((vv_rp, vv_col), x, 0.5, 1, out=out), reps)
    kern["smooth_csr"] = (t_sc, 24 * V + 4 * E)
    t_vn = event_time_ms(lambda: m.vertex_normals(x, "max", out=out), reps)
    kern["vertex_normals"] = (t_vn, 24 * V + 4 * E + 6 * F)
    gcs = (torch.empty(V, dtype=torch.float32, device="cuda"), torch.empty(V, dtype=torch.float32, device="cuda"))
    m.gaussian_curvature(x, out=gcs)
    t_gc = event_time_ms(lambda: m.gaussian_curvature(x, out=gcs), reps)
    kern["gaussian_curvature"] = (t_gc, 12 * V + 4 * E + 6 * F + 8 * V)  # positions, topology, gc + mixed area
    kernels = {k: {"us": round(t * 1e3, 2), "bytes": b, "GB_s": round(b / (t * 1e-3) / 1e9, 1),
                   "frac": round(b / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)} for k, (t, b) in kern.items()}
    # the step is one launch, so its kernel is the roofline kernel
    dt, db = kern["query_vv_vf"]
    ach = db / (dt * 1e-3) / 1e9
    smooth_bytes = it * (24 * V + 4 * E)
    step_bytes = db + smooth_bytes
    line = {
        "metric": "VV-query Gentries/s + %HBM peak; batched grad meshes/s vs plain launches",
        "value": round(value, 4), "unit": "Gentries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "VV + VF query CSRs (one launch), icosphere L9 permuted (%d V / %d F), "
                               "patch size %d" % (V, F, args.patch_size),
                   "mesh": args.workload, "num_patches": P, "patcher": args.patcher,
                   "entries_per_step": entries, "mesh_build_s": round(t_build, 3),
                   "first_small_mesh_build_s": round(t_first, 3)},
        "roofline": {"bound": "hbm", "kernel": "query_vv_vf", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": (measured_traffic(args.workload, args.patch_size, "query_vv_vf")
                                 if args.patcher == "bisect" else None),
                     "algorithmic_bytes": db},
        # the query + smoothing step of the north-star target: the VV + VF query launch, then `it`
        # fp32 Jacobi iterations (lambda 0.5) that read the engine's static patch-local VV plan
        # (ms_separate_queries: VV and VF queried by separate launches, then the smoothing)
        "step_query_smooth": {"smooth_iters": it, "ms": round(ms_smooth, 4),
                              "ms_separate_queries": round(ms_separate, 4),
                              "algorithmic_bytes": step_bytes,
                              "GB_s": round(step_bytes / (ms_smooth * 1e-3) / 1e9, 1),
                              "frac": round(step_bytes / (ms_smooth * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "queries": queries,
        "kernels": kernels,
        "smooth_autotune": {"chosen": smooth_cfg, "us": smooth_sweep},
        # counting-sort form of the transposed queries: 0 LDS-atomic histogram + row sort,
        # 1 atomic-free ballot/mbcnt sort (identical rows); us of one VV + VF launch
        "query_sort_autotune": {"chosen": query_sort, "us": query_sort_us},
    }
    cpu = None
    if not args.no_cpu:  # before the other workloads free the mesh: the CPU port runs on the same mesh
        cpu = cpu_baseline(args.cpu_reps, m, fv, pos, it, entries, t_build)
    if not args.no_configs:
        del out, tmp, vv_rp, vv_col, vf_rp, vf_col
        torch.cuda.empty_cache()
        line["configs"] = configs_bench(reps, m, x)
    # free the single-mesh workload before the batch
    del x
    m.close()
    torch.cuda.empty_cache()
    if args.batch_count > 0:
        line["batch"] = batch_bench(args.batch_count, args.batch_steps, args.batch_reps, args.batch_patch_size)
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
