"""Launch-configuration autotuner.

Times the candidate launch configurations of a mesh (workgroup sizes of the
smoothing and symmetric-Dirichlet kernels, grid vs persistent work-queue
launch of the gradient step, the counting-sort form of the queries) on the
live GPU with HIP events and keeps the fastest. Results are cached in a JSON file keyed by the mesh's shape (patch
count and largest patch), so a later run on a mesh of the same shape reuses
them without timing. Patch size is a build parameter: tune_patch_size()
rebuilds the mesh for each candidate and times the same workload.
"""
import itertools
import json
import os

CACHE_PATH = os.environ.get("RXM_AUTOTUNE_CACHE",
                            os.path.join(os.path.expanduser("~"), ".cache", "rxmesh_amd_autotune.json"))

SMOOTH_THREADS = (64, 128, 256, 512)
GRAD_THREADS = (128, 256)
GRAD_QUEUE = (0, 1)


def _key(mesh, what):
    i = mesh.info
    return "%s:P%d:V%d:nv%d:ne%d:nf%d" % (what, i["num_patches"], i["num_vertices"], i["max_vertices_per_patch"],
                                          i["max_edges_per_patch"], i["max_faces_per_patch"])


def _load():
    try:
        with open(CACHE_PATH) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store(cache):
    try:
        os.makedirs(os.path.dirname(CACHE_PATH), exist_ok=True)
        with open(CACHE_PATH, "w") as fh:
            json.dump(cache, fh, indent=1, sort_keys=True)
    except OSError:
        pass


def _time_ms(fn, reps):
    import torch
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()
    a.record(s)
    for _ in range(reps):
        fn()
    b.record(s)
    b.synchronize()
    return a.elapsed_time(b) / reps


SMOOTH_LAYOUTS = (0, 1)  # LDS staging of positions: 0 three float arrays, 1 float4 per vertex


SMOOTH_FUSED = (0, 1)  # 1: two Jacobi steps per launch over the two-step plan
SMOOTH_STEPS = (2, 3, 4, 5, 6)  # Jacobi steps per launch over the K-level plan
SMOOTH_STEPS_THREADS = (128, 256, 512)


def smooth_configs():
    for fused in SMOOTH_FUSED:
        for st in SMOOTH_THREADS:
            for lay in SMOOTH_LAYOUTS:
                for g4 in ((0, 1) if fused and lay == 0 else (0,)):
                    yield {"smooth_threads": st, "smooth_layout": lay, "smooth_fused": fused, "smooth_steps": 0,
                           "smooth_g4": g4}
    for steps in SMOOTH_STEPS:
        for st in SMOOTH_STEPS_THREADS:
            yield {"smooth_threads": st, "smooth_layout": 0, "smooth_fused": 0, "smooth_steps": steps}


def config_name(cfg):
    if cfg.get("smooth_steps", 0):
        return "t%d/k%d" % (cfg["smooth_threads"], cfg["smooth_steps"])
    return "t%d/l%d/f%d%s" % (cfg["smooth_threads"], cfg["smooth_layout"], cfg["smooth_fused"],
                              "/g4" if cfg.get("smooth_g4", 0) else "")


def apply(mesh, cfg):
    if "smooth_threads" in cfg:
        mesh.set_launch_threads(0, cfg["smooth_threads"])
    if "grad_threads" in cfg:
        mesh.set_launch_threads(1, cfg["grad_threads"])
    if "grad_queue" in cfg:
        mesh.set_launch_threads(2, cfg["grad_queue"])
    if "smooth_layout" in cfg:
        mesh.set_launch_threads(3, cfg["smooth_layout"])
    if "smooth_fused" in cfg:
        mesh.set_launch_threads(4, cfg["smooth_fused"])
    if "smooth_threads" in cfg:
        mesh.set_launch_threads(8, cfg.get("smooth_steps", 0))
        mesh.set_launch_threads(9, cfg.get("smooth_g4", 0))


def tune_smooth(mesh, x, out, lam=0.5, reps=10, use_cache=True, iters=10):
    """Pick workgroup size, LDS layout and steps per launch of uniform
    smoothing from the time of `iters` Jacobi iterations (the workload's
    count: a launch of k steps only pays off over k iterations); returns
    (config, timings in us per iteration)."""
    key = _key(mesh, "smooth%d" % iters)
    cache = _load() if use_cache else {}
    if key in cache:
        apply(mesh, cache[key])
        return cache[key], {}
    import torch
    mesh.build_vv_plan()
    tmp = torch.empty_like(x)
    timings = {}
    best, best_t = None, float("inf")
    for cfg in smooth_configs():
        apply(mesh, cfg)
        t = _time_ms(lambda: mesh.smooth(x, lam, iters, out=out, tmp=tmp), reps) / iters
        timings[config_name(cfg)] = round(t * 1e3, 2)
        if t < best_t:
            best, best_t = cfg, t
    apply(mesh, best)
    cache[key] = best
    if use_cache:
        _store(cache)
    return best, timings


def tune_smooth_symdir(mesh, Mr, x, tmp, lam, eta, steps=2, reps=3, use_cache=True):
    """Pick the launch configuration of the smoothing + symmetric-Dirichlet
    step loop for `mesh`; returns (config, timings)."""
    key = _key(mesh, "smooth_symdir")
    cache = _load() if use_cache else {}
    if key in cache:
        apply(mesh, cache[key])
        return cache[key], {}
    x0 = x.clone()
    timings = {}
    best, best_t = None, float("inf")
    for st, gt, q in itertools.product(SMOOTH_THREADS, GRAD_THREADS, GRAD_QUEUE):
        cfg = {"smooth_threads": st, "grad_threads": gt, "grad_queue": q}
        apply(mesh, cfg)
        t = _time_ms(lambda: mesh.smooth_symdir_steps(Mr, x, tmp, lam, eta, steps), reps)
        timings["%d/%d/%d" % (st, gt, q)] = t
        if t < best_t:
            best, best_t = cfg, t
    x.copy_(x0)
    apply(mesh, best)
    cache[key] = best
    if use_cache:
        _store(cache)
    return best, timings


# transposes: 0 LDS-atomic histogram + row sort, 1 atomic-free ballot/mbcnt sort, 2 the one-pass form
# (VV / VE / VF / EF / FF each in one pass, the fused VV + VF launch in one counting sort; the default)
QUERY_SORTS = (0, 1, 2)


def tune_query_sort(mesh, reps=20):
    """Pick the counting-sort form of the transposed queries (both give the
    same rows) from the time of one VV + VF launch; returns (form, {form:
    us})."""
    vv, vf = mesh.vv_vf_buffers()
    times = {}
    for form in QUERY_SORTS:
        mesh.set_launch_threads(10, form)
        times[form] = round(_time_ms(lambda: mesh.query_vv_vf(vv=vv, vf=vf), reps) * 1e3, 2)
    best = min(times, key=times.get)
    mesh.set_launch_threads(10, best)
    return best, times


PATCH_SIZES = (256, 512, 1024)


def tune_patch_size(fv, make_inputs, run, sizes=PATCH_SIZES, batch=False):
    """Rebuild the mesh for each candidate patch size and time run(mesh,
    *make_inputs(mesh)); returns (best size, {size: ms}). With batch=True, fv
    is a list of face arrays built as one batch. A size whose build or whose
    run is refused (a patch beyond what the workload's kernels hold in LDS or
    in 16-bit local ids) is skipped."""
    import rxmesh_amd as rx
    out = {}
    for ps in sizes:
        try:
            m = rx.Mesh(None, batch=fv, patch_size=ps) if batch else rx.Mesh(fv, patch_size=ps)
            args = make_inputs(m)
            out[ps] = _time_ms(lambda: run(m, *args), 3)
        except rx.RXMError:
            continue
        finally:
            m = None
    if not out:
        raise RuntimeError("tune_patch_size: no candidate patch size runs this workload")
    best = min(out, key=out.get)
    return best, out


def tune_batch_patch_size(fvs, positions, lam, eta, steps=2, sizes=PATCH_SIZES):
    """Patch size for the batched step loop (per-step VV + VF query,
    smoothing, symmetric-Dirichlet gradient step), timed on the meshes
    `fvs` (with vertex positions `positions`, (V, 3) arrays) built as one
    batch; returns (best size, {size: ms per step})."""
    import numpy as np
    import torch

    def make_inputs(m):
        X = np.concatenate(positions).astype(np.float32)
        x = torch.from_numpy(np.ascontiguousarray(X[m.array("v_l2g")].T)).cuda()
        Mr = m.symdir_rest(x)
        vv, vf = m.vv_vf_buffers()
        return Mr, x, torch.empty_like(x), vv, vf

    def run(m, Mr, x, tmp, vv, vf):
        m.query_smooth_symdir_steps(Mr, x, tmp, lam, eta, steps, vv, vf)

    best, times = tune_patch_size(fvs, make_inputs, run, sizes, batch=True)
    return best, {k: round(v / steps, 4) for k, v in times.items()}
